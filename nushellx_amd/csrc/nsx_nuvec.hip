// NuVec's program step (Vec_OP, src/NuVec/NuVec.f90:99-296) for one nucleus and J:
// read <prefix>.abz and the Lanczos vectors of <prefix>.lnz (Lnzs layout) or
// <prefix>.trl (thick-restart layout), diagonalise the projected matrix, form
// the lowest min(no_level, no_state, isave) states, and write
//   .xvc  the states (record no_state, nlev; per state energy and vector, REAL*4)
//   .xva  per state: energy, average a- and b-shell occupations, and the
//         percentage of the state in each (a partition, b partition, 2Ja, 2Jb)
//   .lev  the energies (f10.3) with a parity mark, after a header
//   .lp   the partition analysis of each state, as text
// The states are formed on the device: one GEMM of the N x isave records by
// the isave x nlev eigenvector block.  The projected eigenproblem is solved on
// the host (nsx::sym_eig): the tridiagonal (alpha, beta) of the .lnz layout and
// of a TRL run without restart (count = 1), the full projected matrix
// otherwise (NuVec's tred2 + tqli, :173-198).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/nsx_hv.h"
#include "nsx_api_internal.hpp"
#include "nsx_eig.hpp"

namespace nsx {

// S[r + N c] = sum_j R[r + N j] E[j + isave c]: the states from the records
__global__ __launch_bounds__(256) void k_nuvec_states(const double* __restrict__ R, const double* __restrict__ E,
                                                     int64_t N, int isave, int nlev, double* __restrict__ S) {
  const int64_t r = blockIdx.x * (int64_t)256 + threadIdx.x;
  if (r >= N) return;
  for (int c = 0; c < nlev; ++c) {
    double s = 0.0;
    for (int j = 0; j < isave; ++j) s = fma(R[r + N * j], E[j + (int64_t)isave * c], s);
    S[r + N * c] = s;
  }
}

}  // namespace nsx

using namespace nsx;

namespace {

template <typename T>
T get(const std::vector<unsigned char>& r, size_t off) {
  T v;
  std::memcpy(&v, r.data() + off, sizeof(T));
  return v;
}

template <typename T>
void put(std::vector<unsigned char>& b, T v) {
  const unsigned char* p = reinterpret_cast<const unsigned char*>(&v);
  b.insert(b.end(), p, p + sizeof(T));
}

bool exists(const std::string& p) {
  std::ifstream f(p, std::ios::binary);
  return (bool)f;
}

// one number in a printf format (list-directed items of the text files)
std::string fmt_f(const char* f, double x) {
  char b[64];
  std::snprintf(b, sizeof b, f, x);
  return b;
}

struct DevMem {
  double* p = nullptr;
  ~DevMem() { (void)hipFree(p); }
};

}  // namespace

extern "C" int nsx_nuvec(nsx_hv* h, const char* prefix, int32_t* nlev_out) {
  if (!h || !prefix) { set_error("bad arguments"); return NSX_EARG; }
  if (h->host_only) { set_error("host-only handle: NuVec's state assembly needs the device"); return NSX_EARG; }
  const Model& m = h->model;
  const int64_t N = m.N;
  const std::string base(prefix);
  const bool lnz = exists(base + ".lnz"), trl = exists(base + ".trl");
  if (!lnz && !trl) { set_error(base + ": Lnz or Trl file not found"); return NSX_EIO; }
  std::vector<std::vector<unsigned char>> rec;
  if (!read_records(base + ".abz", rec) || rec.size() < 3) { set_error("cannot read " + base + ".abz"); return NSX_EIO; }
  // record 1: isave, alpha(isave), beta(isave) [, matrix(isave, isave)]
  const std::vector<unsigned char>& r1 = rec[0];
  if (r1.size() < 4) { set_error(base + ".abz: empty first record"); return NSX_EFORMAT; }
  const int isave = get<int32_t>(r1, 0);
  const size_t need1 = 4 + 8 * size_t(std::max(isave, 0)) + (lnz ? 0 : 4 * size_t(isave) * isave);
  if (isave < 1 || isave > 300 || r1.size() < need1) { set_error(base + ".abz: bad first record"); return NSX_EFORMAT; }
  std::vector<double> alpha(isave), beta(isave);
  for (int i = 0; i < isave; ++i) {
    alpha[i] = get<float>(r1, 4 + 4 * i);
    beta[i] = get<float>(r1, 4 + 4 * (isave + i));
  }
  // record 2: imax, no_level, tt, no_state, basd [, count, ...]
  const std::vector<unsigned char>& r2 = rec[1];
  if (r2.size() < 20 || (!lnz && r2.size() < 24)) { set_error(base + ".abz: bad second record"); return NSX_EFORMAT; }
  int no_level = get<int32_t>(r2, 4);
  const int no_state = get<int32_t>(r2, 12), basd = get<int32_t>(r2, 16);
  const int count = lnz ? 1 : get<int32_t>(r2, 20);
  if (no_state != N || basd != m.basd) { set_error(base + ".abz: no_state or basd differs from the handle's basis"); return NSX_EFORMAT; }
  if (no_level > no_state) no_level = no_state;
  const int nlev = std::max(0, std::min(no_level, isave));
  // the projected matrix (row-major isave x isave)
  std::vector<double> T(size_t(isave) * isave, 0.0);
  if (lnz) {  // tqli on (alpha, beta): beta(i) couples i-1 and i (beta(1) unused)
    for (int i = 0; i < isave; ++i) T[size_t(i) * isave + i] = alpha[i];
    for (int i = 1; i < isave; ++i) T[size_t(i - 1) * isave + i] = T[size_t(i) * isave + i - 1] = beta[i];
  } else if (count == 1) {  // betat(i) couples i and i+1 (:176-180 shift it for tqli)
    for (int i = 0; i < isave; ++i) T[size_t(i) * isave + i] = alpha[i];
    for (int i = 0; i + 1 < isave; ++i) T[size_t(i) * isave + i + 1] = T[size_t(i + 1) * isave + i] = beta[i];
  } else {  // matrix(1:isave, 1:isave), column-major
    for (int c = 0; c < isave; ++c)
      for (int r = 0; r < isave; ++r) T[size_t(r) * isave + c] = get<float>(r1, 4 + 8 * size_t(isave) + 4 * (size_t(c) * isave + r));
  }
  std::vector<double> w(isave), Z(size_t(isave) * isave);
  if (!sym_eig(T.data(), isave, true, w.data(), Z.data())) { set_error("projected eigenproblem did not converge"); return NSX_EGPU; }
  // the Lanczos vectors: isave records of no_state REAL*4
  std::vector<double> Rh(size_t(N) * isave);
  {
    std::ifstream f(base + (lnz ? ".lnz" : ".trl"), std::ios::binary);
    std::vector<float> buf((size_t)N);
    for (int j = 0; j < isave; ++j) {
      if (!f.read(reinterpret_cast<char*>(buf.data()), sizeof(float) * N)) {
        set_error(base + (lnz ? ".lnz" : ".trl") + ": fewer records than isave");
        return NSX_EFORMAT;
      }
      for (int64_t i = 0; i < N; ++i) Rh[size_t(j) * N + i] = buf[i];
    }
  }
  // states on the device: S = R E, E = the nlev lowest eigenvectors
  std::vector<double> S(size_t(N) * std::max(nlev, 1));
  if (nlev > 0) {
    DevMem dR, dE, dS;
    hipError_t e;
    auto fail = [&](hipError_t x) { set_error(std::string("HIP error: ") + hipGetErrorString(x)); return NSX_EGPU; };
    if ((e = hipMalloc(&dR.p, sizeof(double) * Rh.size())) != hipSuccess) return fail(e);
    if ((e = hipMalloc(&dE.p, sizeof(double) * size_t(isave) * nlev)) != hipSuccess) return fail(e);
    if ((e = hipMalloc(&dS.p, sizeof(double) * size_t(N) * nlev)) != hipSuccess) return fail(e);
    if ((e = hipMemcpyAsync(dR.p, Rh.data(), sizeof(double) * Rh.size(), hipMemcpyHostToDevice, h->stream)) != hipSuccess)
      return fail(e);
    if ((e = hipMemcpyAsync(dE.p, Z.data(), sizeof(double) * size_t(isave) * nlev, hipMemcpyHostToDevice, h->stream)) !=
        hipSuccess)
      return fail(e);
    hipLaunchKernelGGL(k_nuvec_states, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, h->stream, dR.p, dE.p, N,
                       isave, nlev, dS.p);
    if ((e = hipGetLastError()) != hipSuccess) return fail(e);
    if ((e = hipMemcpyAsync(S.data(), dS.p, sizeof(double) * size_t(N) * nlev, hipMemcpyDeviceToHost, h->stream)) !=
        hipSuccess)
      return fail(e);
    if ((e = hipStreamSynchronize(h->stream)) != hipSuccess) return fail(e);
  }
  // NuVec keeps coefT in REAL*4: round the states, then analyse them
  for (double& x : S) x = (float)x;
  const SideFiles &A = m.A, &B = m.B;
  const int ns = m.nus.no_shells;
  const int nJa = A.maxJ2 / 2 - A.minJ2 / 2 + 1, nJb = B.maxJ2 / 2 - B.minJ2 / 2 + 1;
  const size_t npc = size_t(A.nsp) * B.nsp * nJa * nJb;
  auto pc_idx = [&](int ap, int bp, int ja2, int jb2) {  // spartpc(anpl, bnpl, ajl/2, bjl/2), column-major
    return size_t(ap) + size_t(A.nsp) * (bp + size_t(B.nsp) * ((ja2 / 2 - A.minJ2 / 2) + size_t(nJa) * (jb2 / 2 - B.minJ2 / 2)));
  };
  std::vector<char> chpar(nlev, '?');
  std::vector<std::vector<double>> ava(nlev, std::vector<double>(ns)), avb(nlev, std::vector<double>(ns)),
      pc(nlev, std::vector<double>(npc));
  std::vector<double> tampl(nlev), para(nlev), npara(nlev);
  for (int i = 0; i < nlev; ++i) {
    const double* v = &S[size_t(i) * N];
    double t = 0.0, pp = 0.0, np = 0.0;
    // basis order (NuVec.f90:136-152): per block, b partition, b state, a partition, a state
    for (int b = 0; b < m.basd; ++b) {
      const int ja = m.bas_a[b], jb = m.bas_b[b];
      int64_t k = m.posnv[b];
      for (int bp = 0; bp < B.nsp; ++bp)
        for (int rs = 0; rs < B.n(jb, bp); ++rs)
          for (int ap = 0; ap < A.nsp; ++ap)
            for (int ls = 0; ls < A.n(ja, ap); ++ls, ++k) {
              const double a2 = v[k] * v[k];
              t += a2;
              if (A.parity[ap] * B.parity[bp] >= 0) pp += a2; else np += a2;
              pc[i][pc_idx(ap, bp, ja, jb)] += a2;
              for (int q = 0; q < ns; ++q) {
                ava[i][q] += A.occ[size_t(ap) * ns + q] * a2;
                avb[i][q] += B.occ[size_t(bp) * ns + q] * a2;
              }
            }
    }
    tampl[i] = t;
    para[i] = pp;
    npara[i] = np;
    if (pp / t > 0.95) chpar[i] = '+';
    if (np / t > 0.95) chpar[i] = '-';
    for (double& x : pc[i]) x = (float)((float)x * 100.0f / (float)t);
  }
  int rc = nsx_hv_write_xvc(h, prefix, std::max(nlev, 1), w.data(), S.data());
  if (nlev == 0) rc = NSX_OK;
  if (rc) return rc;
  // .xva: energy, ava, avb, spartpc (REAL*4)
  {
    FILE* f = std::fopen((base + ".xva").c_str(), "wb");
    if (!f) { set_error("cannot open " + base + ".xva for writing"); return NSX_EIO; }
    bool ok = true;
    for (int i = 0; i < nlev && ok; ++i) {
      std::vector<unsigned char> r;
      put<float>(r, (float)w[i]);
      for (double x : ava[i]) put<float>(r, (float)x);
      for (double x : avb[i]) put<float>(r, (float)x);
      for (double x : pc[i]) put<float>(r, (float)x);
      ok = write_record(f, r);
    }
    ok = (std::fclose(f) == 0) && ok;
    if (!ok) { set_error("write failed: " + base + ".xva"); return NSX_EIO; }
  }
  // .lev and .lp: output_results' header (:298-320), then the energies '(5(f10.3,a1))'
  int nA = 0, nB = 0;  // valence particles of each side: the occupations of its first partition (2T = |nA - nB|)
  for (int q = 0; q < ns; ++q) {
    if (A.nsp > 0) nA += A.occ[q];
    if (B.nsp > 0) nB += B.occ[q];
  }
  auto header = [&](FILE* f) {
    std::fprintf(f, " nushellx_amd (MI355X): NuVec's analysis of the Lanczos states\n");
    std::fprintf(f, " %s\n\n", m.intfile.c_str());
    for (int k = 0; k < 3; ++k) std::fprintf(f, " %s\n", m.nus.desc[k].c_str());
    std::fprintf(f, "\n  2*J,   2*T %12d%12d\n\n  Lowest Energy Levels\n\n", m.J2, std::abs(nA - nB));
  };
  auto energies = [&](FILE* f) {
    for (int i = 0; i < nlev; ++i) {
      std::fprintf(f, "%10.3f%c", w[i], chpar[i]);
      if (i % 5 == 4 || i == nlev - 1) std::fprintf(f, "\n");
    }
  };
  {
    FILE* f = std::fopen((base + ".lev").c_str(), "w");
    if (!f) { set_error("cannot open " + base + ".lev for writing"); return NSX_EIO; }
    header(f);
    energies(f);
    if (std::fclose(f) != 0) { set_error("write failed: " + base + ".lev"); return NSX_EIO; }
  }
  {
    FILE* f = std::fopen((base + ".lp").c_str(), "w");
    if (!f) { set_error("cannot open " + base + ".lp for writing"); return NSX_EIO; }
    header(f);
    std::fprintf(f, "\n  Type a = %c  Type b = %c\n", A.proton ? 'p' : 'n', B.proton ? 'p' : 'n');
    auto parts = [&](const SideFiles& s, const char* title) {
      std::fprintf(f, " %s\n", title);
      for (int p = 0; p < s.nsp; ++p) {
        std::fprintf(f, "%8d", -(p + 1));
        for (int q = 0; q < ns; ++q) {
          std::fprintf(f, "%4d", (int)s.occ[size_t(p) * ns + q]);
          if (q % 18 == 17 && q + 1 < ns) std::fprintf(f, "\n");
        }
        std::fprintf(f, "\n");
      }
    };
    parts(A, "A partitions");
    parts(B, "B partitions");
    auto row11 = [&](const std::vector<double>& x, double t) {
      for (int q = 0; q < ns; ++q) {
        std::fprintf(f, "%7.2f", x[q] / t);
        if (q % 11 == 10 || q == ns - 1) std::fprintf(f, "\n");
      }
    };
    for (int i = 0; i < nlev; ++i) {
      std::fprintf(f, "\n State %12d%s %c\n\n", i + 1, fmt_f("%16.7f", w[i]).c_str(), chpar[i]);
      std::fprintf(f, " Average A type nucleons  \n");
      row11(ava[i], tampl[i]);
      std::fprintf(f, "\n Average B type nucleons  \n");
      row11(avb[i], tampl[i]);
      std::fprintf(f, "  Positive and Negative Parity \n");
      std::fprintf(f, "  +ve %8.2f%% -ve %8.2f%%\n", para[i] * 100.0 / tampl[i], npara[i] * 100.0 / tampl[i]);
      for (int jb = B.minJ2; jb <= B.maxJ2; jb += 2)
        for (int ja = A.minJ2; ja <= A.maxJ2; ja += 2) {
          double amp = 0.0;
          for (int bp = 0; bp < B.nsp; ++bp)
            for (int ap = 0; ap < A.nsp; ++ap) amp += pc[i][pc_idx(ap, bp, ja, jb)];
          if (amp < 1.0) continue;
          std::fprintf(f, "\n 2*JA, 2*JB, amplitude %12d%12d%s\n", ja, jb, fmt_f("%16.7f", amp).c_str());
          for (int bp = 0; bp < B.nsp; ++bp) {
            double ab = 0.0;
            for (int ap = 0; ap < A.nsp; ++ap) ab += pc[i][pc_idx(ap, bp, ja, jb)];
            if (ab < 1.0) continue;
            std::fprintf(f, " B partition, amplitude %12d%s\n", bp + 1, fmt_f("%16.7f", ab).c_str());
            std::fprintf(f, " A partitions, amplitudes\n");
            for (int ap = 0; ap < A.nsp; ++ap) {
              std::fprintf(f, "(%3d,%7.2f)", ap + 1, pc[i][pc_idx(ap, bp, ja, jb)]);
              if (ap % 6 == 5 || ap == A.nsp - 1) std::fprintf(f, "\n");
            }
          }
        }
    }
    energies(f);
    if (std::fclose(f) != 0) { set_error("write failed: " + base + ".lp"); return NSX_EIO; }
  }
  if (nlev_out) *nlev_out = nlev;
  return NSX_OK;
}
