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
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/nsx_hv.h"
#include "nsx_api_internal.hpp"

namespace {

struct File {
  FILE* f = nullptr;
  explicit File(const std::string& p) : f(std::fopen(p.c_str(), "wb")) {}
  ~File() {
    if (f) std::fclose(f);
  }
};

// one Fortran sequential record: length marker, payload, length marker
bool record(FILE* f, const std::vector<unsigned char>& payload) {
  const int32_t n = (int32_t)payload.size();
  return std::fwrite(&n, 4, 1, f) == 1 &&
         (payload.empty() || std::fwrite(payload.data(), 1, payload.size(), f) == payload.size()) &&
         std::fwrite(&n, 4, 1, f) == 1;
}

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
