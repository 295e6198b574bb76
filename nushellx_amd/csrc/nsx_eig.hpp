// Small dense symmetric eigensolver for the Lanczos projected matrix (host, FP64).
//
// The projected matrix of a thick-restart cycle is an arrowhead block (Ritz
// values on the diagonal, coupling column k) followed by a tridiagonal tail.
// Householder reflections taken from the bottom row upwards bring it to
// tridiagonal form; rows that are already tridiagonal are skipped, so only the
// arrowhead block costs O(k^3).  The reflectors never touch the last index, so
// the last row of the transformation stays e_{n-1}^T.  The implicit symmetric
// QR algorithm with Wilkinson shifts (Golub & Van Loan, Algorithms 8.3.2-8.3.3)
// then diagonalises the tridiagonal matrix, applying each plane rotation either
// to the full transformation or to its last row only.  The last
// row is all the Lanczos convergence test needs (residual |beta_m s_{m,i}|), so a
// test costs O(k^3 + n^2) instead of a full O(n^3) decomposition.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <numeric>
#include <vector>

namespace nsx {

// Implicit symmetric QR on a tridiagonal matrix (diagonal d[0..n), coupling
// e[k] between k and k+1), after Golub & Van Loan, Matrix Computations (4th
// ed.), Algorithms 8.3.2 and 8.3.3: deflate negligible couplings from the
// bottom, take the Wilkinson shift of the trailing 2x2 of the lowest unreduced
// window, and chase the bulge of one plane rotation down that window.  Each
// rotation R acts on coordinates (k, k+1) as [c s; -s c]; T <- R T R^T, and the
// tracked rows of the transformation take Z <- Z R^T.  z holds nz rows of
// length n (row-major).  On return d holds the eigenvalues (unsorted).
inline bool tridiag_qr(double* d, double* e, int n, double* z, int nz) {
  const double eps = std::numeric_limits<double>::epsilon();
  auto negligible = [&](int k) { return std::fabs(e[k]) <= eps * (std::fabs(d[k]) + std::fabs(d[k + 1])); };
  auto rotate_rows = [&](int k, double c, double s) {
    for (int r = 0; r < nz; ++r) {
      double* zr = z + size_t(r) * n;
      const double u = zr[k], w = zr[k + 1];
      zr[k] = c * u + s * w;
      zr[k + 1] = c * w - s * u;
    }
  };
  int hi = n - 1;
  long budget = 40L * n + 40;
  while (hi > 0) {
    if (negligible(hi - 1)) {  // bottom coupling gone: d[hi] is an eigenvalue
      e[hi - 1] = 0.0;
      --hi;
      continue;
    }
    int lo = hi - 1;  // grow the unreduced window [lo, hi] upwards
    while (lo > 0 && !negligible(lo - 1)) --lo;
    if (lo > 0) e[lo - 1] = 0.0;
    if (--budget < 0) return false;
    // Wilkinson shift: the eigenvalue of the trailing 2x2 nearer to d[hi]
    const double b = e[hi - 1];
    const double half = 0.5 * (d[hi - 1] - d[hi]);
    const double root = std::hypot(half, b);
    const double mu = d[hi] - b * b / (half + (half >= 0.0 ? root : -root));
    double x = d[lo] - mu, y = e[lo];  // first column of T - mu I, rows lo, lo+1
    for (int k = lo; k < hi; ++k) {
      const double r = std::hypot(x, y);
      const double c = r == 0.0 ? 1.0 : x / r;
      const double s = r == 0.0 ? 0.0 : y / r;
      if (k > lo) e[k - 1] = r;  // the bulge at (k+1, k-1) is folded into e[k-1]
      const double p = d[k], t = d[k + 1], off = e[k];
      const double cc = c * c, ss = s * s, cs = c * s;
      d[k] = cc * p + 2.0 * cs * off + ss * t;
      d[k + 1] = ss * p - 2.0 * cs * off + cc * t;
      e[k] = cs * (t - p) + (cc - ss) * off;
      rotate_rows(k, c, s);
      if (k + 1 < hi) {  // row k picks up s * e[k+1] at (k, k+2): the new bulge
        x = e[k];
        y = s * e[k + 1];
        e[k + 1] *= c;
      }
    }
  }
  return true;
}

// A: row-major n x n symmetric.  w: n eigenvalues ascending.
// full: Z is n x n column-major (column c = eigenvector c);
// otherwise Z holds the n last components (Z[c] = s_{n-1, c}).
// Returns false if the QR iteration exceeds its budget (never seen for symmetric input).
inline bool sym_eig(const double* A, int n, bool full, double* w, double* Z) {
  if (n <= 0) return true;
  std::vector<double> a(A, A + size_t(n) * n);
  auto at = [&](int r, int c) -> double& { return a[size_t(r) * n + c]; };
  // transformation kept row-major as q[r * n + c] when full
  std::vector<double> q;
  if (full) {
    q.assign(size_t(n) * n, 0.0);
    for (int i = 0; i < n; ++i) q[size_t(i) * n + i] = 1.0;
  }
  std::vector<double> v(n), p(n);
  for (int i = n - 1; i >= 2; --i) {
    double off = 0.0;
    for (int k = 0; k < i - 1; ++k) off += std::fabs(at(i, k));
    if (off == 0.0) continue;  // row already tridiagonal
    double nrm2 = 0.0;
    for (int k = 0; k < i; ++k) nrm2 += at(i, k) * at(i, k);
    const double xl = at(i, i - 1);
    const double alpha = -std::copysign(std::sqrt(nrm2), xl);
    for (int k = 0; k < i; ++k) v[k] = at(i, k);
    v[i - 1] -= alpha;
    const double h = nrm2 - alpha * xl;  // v^T v / 2; P = I - v v^T / h
    double K = 0.0;
    for (int k = 0; k < i; ++k) {
      double s = 0.0;
      for (int l = 0; l < i; ++l) s += at(k, l) * v[l];
      p[k] = s / h;
      K += v[k] * p[k];
    }
    K /= 2.0 * h;
    for (int k = 0; k < i; ++k) p[k] -= K * v[k];  // p is now q = p - K v
    for (int k = 0; k < i; ++k)
      for (int l = 0; l < i; ++l) at(k, l) -= v[k] * p[l] + p[k] * v[l];
    for (int k = 0; k < i; ++k) at(i, k) = at(k, i) = 0.0;
    at(i, i - 1) = at(i - 1, i) = alpha;
    if (full) {
      for (int r = 0; r < n; ++r) {
        double s = 0.0;
        for (int l = 0; l < i; ++l) s += q[size_t(r) * n + l] * v[l];
        s /= h;
        for (int k = 0; k < i; ++k) q[size_t(r) * n + k] -= s * v[k];
      }
    }
  }
  std::vector<double> d(n), e(n, 0.0);
  for (int k = 0; k < n; ++k) d[k] = at(k, k);
  for (int k = 0; k + 1 < n; ++k) e[k] = at(k + 1, k);  // e[k] couples k and k+1
  // Rows of the accumulated transformation that the rotations update: all n
  // rows, or only the last one.
  const int nz = full ? n : 1;
  std::vector<double> z(size_t(nz) * n, 0.0);
  if (full) z = q;
  else z[n - 1] = 1.0;
  if (!tridiag_qr(d.data(), e.data(), n, z.data(), nz)) return false;
  std::vector<int> idx(n);
  std::iota(idx.begin(), idx.end(), 0);
  std::sort(idx.begin(), idx.end(), [&](int x, int y) { return d[x] < d[y]; });
  for (int c = 0; c < n; ++c) {
    w[c] = d[idx[c]];
    if (full)
      for (int r = 0; r < n; ++r) Z[size_t(c) * n + r] = z[size_t(r) * n + idx[c]];
    else
      Z[c] = z[idx[c]];
  }
  return true;
}

}  // namespace nsx
