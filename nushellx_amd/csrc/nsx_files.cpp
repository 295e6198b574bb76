// Lanczos-vector files in the reference's layout, for its downstream programs.
//
// NuVec (src/NuVec/NuVec.f90:113-224) reads <prefix>.abz and either <prefix>.lnz
// or <prefix>.trl, where prefix = Nucleus // file_code(J, nA, nB).  The .lnz
// variant (written by Lnzs, src/NuTRL/NuCoreX.f90:738-743) is:
//   .abz  sequential unformatted, 4-byte record markers:
//         rec 1: isave (i4), alpha(1:isave) (r4), beta(1:isave) (r4)
//         rec 2: imax (i4), no_level (i4), tt (r4), no_state (i4), basd (i4)
//         rec 3: bas(1:basd), character(8) each = char(2Ja) char(2Jb) + six NULs
//   .lnz  direct access, record r = vector r, no_state REAL*4, no markers
//         (recl = no_state * rkw / 4 four-byte units, :553-555)
// NuVec diagonalises the tridiagonal (alpha, beta) and forms each wanted state
// as sum_r eigv(r, i) * record r.  Writing the converged Ritz vectors as the
// records with alpha = Ritz values and beta = 0 makes that matrix diagonal, so
// NuVec reproduces exactly these states (in REAL*4) without a Lanczos history.
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iterator>
#include <string>
#include <vector>

#include "../../include/nsx_hv.h"
#include "nsx_api_internal.hpp"

namespace nsx {

// one Fortran sequential record: length marker, payload, length marker
bool write_record(FILE* f, const std::vector<unsigned char>& payload) {
  const int32_t n = (int32_t)payload.size();
  return std::fwrite(&n, 4, 1, f) == 1 &&
         (payload.empty() || std::fwrite(payload.data(), 1, payload.size(), f) == payload.size()) &&
         std::fwrite(&n, 4, 1, f) == 1;
}

// Fortran sequential unformatted records of a whole file
bool read_records(const std::string& path, std::vector<std::vector<unsigned char>>& out) {
  std::ifstream f(path, std::ios::binary);
  if (!f) return false;
  std::vector<unsigned char> d((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
  size_t pos = 0;
  while (pos + 8 <= d.size()) {
    int32_t n, n2;
    std::memcpy(&n, d.data() + pos, 4);
    if (n < 0 || pos + 8 + size_t(n) > d.size()) return false;
    std::memcpy(&n2, d.data() + pos + 4 + n, 4);
    if (n2 != n) return false;
    out.emplace_back(d.begin() + pos + 4, d.begin() + pos + 4 + n);
    pos += 8 + size_t(n);
  }
  return pos == d.size();
}

}  // namespace nsx

namespace {

using nsx::read_records;

struct File {
  FILE* f = nullptr;
  explicit File(const std::string& p) : f(std::fopen(p.c_str(), "wb")) {}
  ~File() {
    if (f) std::fclose(f);
  }
};

bool record(FILE* f, const std::vector<unsigned char>& payload) { return nsx::write_record(f, payload); }

template <typename T>
void put(std::vector<unsigned char>& b, T v) {
  const unsigned char* p = reinterpret_cast<const unsigned char*>(&v);
  b.insert(b.end(), p, p + sizeof(T));
}

}  // namespace

extern "C" int nsx_hv_write_vectors(const nsx_hv* h, const char* prefix, int32_t nvec, const double* evals,
                                    const double* vecs, int32_t no_level, float tt) {
  using nsx::set_error;
  if (!h || !prefix || !evals || !vecs || nvec < 1) { set_error("bad arguments"); return NSX_EARG; }
  const nsx::Model& m = h->model;
  const std::string base(prefix);
  {
    File f(base + ".lnz");
    if (!f.f) { set_error("cannot open " + base + ".lnz for writing"); return NSX_EIO; }
    std::vector<float> rec(size_t(m.N));
    for (int q = 0; q < nvec; ++q) {
      const double* v = vecs + size_t(q) * m.N;
      for (int64_t i = 0; i < m.N; ++i) rec[i] = (float)v[i];
      if (std::fwrite(rec.data(), sizeof(float), rec.size(), f.f) != rec.size()) {
        set_error("write failed: " + base + ".lnz");
        return NSX_EIO;
      }
    }
  }
  File f(base + ".abz");
  if (!f.f) { set_error("cannot open " + base + ".abz for writing"); return NSX_EIO; }
  std::vector<unsigned char> r1, r2, r3;
  put<int32_t>(r1, nvec);
  for (int q = 0; q < nvec; ++q) put<float>(r1, (float)evals[q]);
  for (int q = 0; q < nvec; ++q) put<float>(r1, 0.0f);
  put<int32_t>(r2, nvec);
  put<int32_t>(r2, no_level);
  put<float>(r2, tt);
  put<int32_t>(r2, (int32_t)m.N);
  put<int32_t>(r2, m.basd);
  for (int b = 0; b < m.basd; ++b) {
    unsigned char key[8] = {(unsigned char)m.bas_a[b], (unsigned char)m.bas_b[b], 0, 0, 0, 0, 0, 0};
    r3.insert(r3.end(), key, key + 8);
  }
  if (!record(f.f, r1) || !record(f.f, r2) || !record(f.f, r3)) {
    set_error("write failed: " + base + ".abz");
    return NSX_EIO;
  }
  return NSX_OK;
}

// The thick-restart layout NuTRL leaves behind (src/NuTRL/NuCoreX.f90:526-534,
// records of .trl as :553-555), read by NuVec (src/NuVec/NuVec.f90:124-130,
// 156-157, 174-190):
//   .abz  rec 1: isave, alpha(1:isave), betat(1:isave), matrix(1:isave,1:isave)
//         rec 2: imax, no_level, tt, no_state, basd, count, no_keep, no_conv,
//                first_iter, second_iter, noiter
//         rec 3: bas(1:basd)
//         rec 4: alpha_copy(1:noiter), betat(isave+1), indx(1:noiter)
//   .trl  direct access, record r = vector r (no_state REAL*4, no markers)
// The records are the converged Ritz vectors and the projected matrix is
// diag(Ritz values) with count = 2, the state after a thick restart whose
// arrow couplings have converged: NuVec's tred2/tqli path then returns exactly
// these values and vectors.
extern "C" int nsx_hv_write_trl(const nsx_hv* h, const char* prefix, int32_t nvec, const double* evals,
                                const double* vecs, int32_t no_level, float tt) {
  using nsx::set_error;
  constexpr int32_t kNoiter = 300, kFirstIterT = 80, kIncIter = 3;  // src/NuCommon/NuCommon.f90:111-124
  if (!h || !prefix || !evals || !vecs || nvec < 1 || nvec > kNoiter) { set_error("bad arguments"); return NSX_EARG; }
  const nsx::Model& m = h->model;
  const std::string base(prefix);
  {
    File f(base + ".trl");
    if (!f.f) { set_error("cannot open " + base + ".trl for writing"); return NSX_EIO; }
    std::vector<float> rec(size_t(m.N));
    for (int q = 0; q < nvec; ++q) {
      const double* v = vecs + size_t(q) * m.N;
      for (int64_t i = 0; i < m.N; ++i) rec[i] = (float)v[i];
      if (std::fwrite(rec.data(), sizeof(float), rec.size(), f.f) != rec.size()) {
        set_error("write failed: " + base + ".trl");
        return NSX_EIO;
      }
    }
  }
  File f(base + ".abz");
  if (!f.f) { set_error("cannot open " + base + ".abz for writing"); return NSX_EIO; }
  std::vector<unsigned char> r1, r2, r3, r4;
  put<int32_t>(r1, nvec);
  for (int q = 0; q < nvec; ++q) put<float>(r1, (float)evals[q]);
  for (int q = 0; q < nvec; ++q) put<float>(r1, 0.0f);
  for (int c = 0; c < nvec; ++c)
    for (int r = 0; r < nvec; ++r) put<float>(r1, r == c ? (float)evals[c] : 0.0f);
  for (int32_t x : {nvec, no_level}) put<int32_t>(r2, x);
  put<float>(r2, tt);
  for (int32_t x : {(int32_t)m.N, (int32_t)m.basd, 2, nvec, nvec, kFirstIterT, kIncIter, kNoiter}) put<int32_t>(r2, x);
  for (int b = 0; b < m.basd; ++b) {
    unsigned char key[8] = {(unsigned char)m.bas_a[b], (unsigned char)m.bas_b[b], 0, 0, 0, 0, 0, 0};
    r3.insert(r3.end(), key, key + 8);
  }
  for (int q = 0; q < kNoiter; ++q) put<float>(r4, q < nvec ? (float)evals[q] : 0.0f);
  put<float>(r4, 0.0f);
  for (int q = 0; q < kNoiter; ++q) put<int32_t>(r4, q + 1);
  if (!record(f.f, r1) || !record(f.f, r2) || !record(f.f, r3) || !record(f.f, r4)) {
    set_error("write failed: " + base + ".abz");
    return NSX_EIO;
  }
  return NSX_OK;
}

// NuVec's state file (src/NuVec/NuVec.f90:166,200,291): record no_state, nlev; then
// per level its energy and the state, REAL*4 (rl = 4, src/NuCommon/NuCommon.f90:101).
extern "C" int nsx_hv_write_xvc(const nsx_hv* h, const char* prefix, int32_t nvec, const double* evals,
                                const double* vecs) {
  using nsx::set_error;
  if (!h || !prefix || !evals || !vecs || nvec < 1) { set_error("bad arguments"); return NSX_EARG; }
  const int64_t N = h->model.N;
  const std::string path = std::string(prefix) + ".xvc";
  File f(path);
  if (!f.f) { set_error("cannot open " + path + " for writing"); return NSX_EIO; }
  std::vector<unsigned char> r;
  put<int32_t>(r, (int32_t)N);
  put<int32_t>(r, nvec);
  if (!record(f.f, r)) { set_error("write failed: " + path); return NSX_EIO; }
  for (int q = 0; q < nvec; ++q) {
    r.clear();
    put<float>(r, (float)evals[q]);
    for (int64_t i = 0; i < N; ++i) put<float>(r, (float)vecs[size_t(q) * N + i]);
    if (!record(f.f, r)) { set_error("write failed: " + path); return NSX_EIO; }
  }
  return NSX_OK;
}

// NuXpct's file step (src/NuXpct/NuCoreXE.f90:1431-1470): the states of
// <prefix>.xvc, M(i,j) = (H v_i) . v_j, written to <prefix><operat>.ovl (text,
// one column of 8f10.4 lines per final level) and <prefix><operat>.ovx
// (record nlev, then one REAL*4 record per column).  A file with fewer state
// records than its header announces gives the matrix of the states present.
extern "C" int nsx_hv_xpct_files(nsx_hv* h, const char* prefix, const char* operat, int32_t* nlev_out) {
  using nsx::set_error;
  if (!h || !prefix || !operat) { set_error("bad arguments"); return NSX_EARG; }
  const int64_t N = h->model.N;
  const std::string base(prefix), op(operat);
  std::vector<std::vector<unsigned char>> recs;
  if (!read_records(base + ".xvc", recs) || recs.empty() || recs[0].size() < 8) {
    set_error("cannot read " + base + ".xvc");
    return NSX_EIO;
  }
  int32_t no_st, nlev;
  std::memcpy(&no_st, recs[0].data(), 4);
  std::memcpy(&nlev, recs[0].data() + 4, 4);
  if (no_st != N) { set_error(base + ".xvc: no_state differs from the handle's basis"); return NSX_EFORMAT; }
  if (nlev < 0) { set_error(base + ".xvc: bad level count"); return NSX_EFORMAT; }
  nlev = std::min<int32_t>(nlev, (int32_t)recs.size() - 1);
  if (nlev < 1) { set_error(base + ".xvc: no states"); return NSX_EFORMAT; }
  std::vector<double> V(size_t(N) * nlev);
  for (int q = 0; q < nlev; ++q) {
    const std::vector<unsigned char>& r = recs[q + 1];
    if (r.size() == 4 + 4 * size_t(N)) {
      for (int64_t i = 0; i < N; ++i) {
        float x;
        std::memcpy(&x, r.data() + 4 + 4 * i, 4);
        V[size_t(q) * N + i] = x;
      }
    } else if (r.size() == 8 + 8 * size_t(N)) {
      std::memcpy(&V[size_t(q) * N], r.data() + 8, 8 * size_t(N));
    } else {
      set_error(base + ".xvc: state record of unexpected length");
      return NSX_EFORMAT;
    }
  }
  std::vector<double> M(size_t(nlev) * nlev);
  const int rc = nsx_hv_expect(h, nlev, V.data(), M.data());
  if (rc) return rc;
  {
    File t(base + op + ".ovl");
    if (!t.f) { set_error("cannot open " + base + op + ".ovl for writing"); return NSX_EIO; }
    std::fprintf(t.f, " nushellx_amd (MI355X) expectation values\n\n");
    std::fprintf(t.f, "  Expectation Values and Off Diagonal Matrix Elements\n\n");
    std::fprintf(t.f, "  Interaction/Operator file %s\n\n", op.c_str());
    std::fprintf(t.f, "  Initial levels from 1 ... No_Levels left to right\n");
    std::fprintf(t.f, "  Final   levels from 1 ... No_Levels top  to bottom\n\n");
    for (int j = 0; j < nlev; ++j) {
      for (int i = 0; i < nlev; ++i) {
        std::fprintf(t.f, "%10.4f", M[size_t(j) * nlev + i]);
        if (i % 8 == 7 || i == nlev - 1) std::fprintf(t.f, "\n");
      }
      std::fprintf(t.f, "\n");
    }
    if (std::ferror(t.f)) { set_error("write failed: " + base + op + ".ovl"); return NSX_EIO; }
  }
  File x(base + op + ".ovx");
  if (!x.f) { set_error("cannot open " + base + op + ".ovx for writing"); return NSX_EIO; }
  std::vector<unsigned char> r;
  put<int32_t>(r, nlev);
  bool ok = record(x.f, r);
  for (int j = 0; j < nlev && ok; ++j) {
    r.clear();
    for (int i = 0; i < nlev; ++i) put<float>(r, (float)M[size_t(j) * nlev + i]);
    ok = record(x.f, r);
  }
  if (!ok) { set_error("write failed: " + base + op + ".ovx"); return NSX_EIO; }
  if (nlev_out) *nlev_out = nlev;
  return NSX_OK;
}
