// Small dense symmetric eigensolver for the Lanczos projected matrix (host, FP64).
//
// The projected matrix of a thick-restart cycle is an arrowhead block (Ritz
// values on the diagonal, coupling column k) followed by a tridiagonal tail.
// Householder reflections taken from the bottom row upwards bring it to
// tridiagonal form; rows that are already tridiagonal are skipped, so only the
// arrowhead block costs O(k^3).  The reflectors never touch the last index, so
// the last row of the transformation stays e_{n-1}^T.  Implicit QL with a
// Wilkinson-type shift then diagonalises the tridiagonal matrix (the method of
// EISPACK tql2; Golub & Van Loan, Matrix Computations, 8.3), applying each plane
// rotation either to the full transformation or to its last row only.  The last
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

// A: row-major n x n symmetric.  w: n eigenvalues ascending.
// full: Z is n x n column-major (column c = eigenvector c);
// otherwise Z holds the n last components (Z[c] = s_{n-1, c}).
// Returns false if QL fails to converge (never seen for symmetric input).
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
  for (int k = 0; k + 1 < n; ++k) e[k] = at(k + 1, k);
  // z rows tracked by the QL rotations: all n rows, or the last one
  const int nz = full ? n : 1;
  std::vector<double> z(size_t(nz) * n, 0.0);
  if (full) z = q;
  else z[n - 1] = 1.0;
  const double eps = std::numeric_limits<double>::epsilon();
  for (int l = 0; l < n; ++l) {
    int iter = 0, m;
    do {
      for (m = l; m < n - 1; ++m) {
        const double dd = std::fabs(d[m]) + std::fabs(d[m + 1]);
        if (std::fabs(e[m]) <= eps * dd) break;
      }
      if (m == l) break;
      if (++iter > 60) return false;
      // shift from the leading 2x2 block at l
      double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
      double r = std::hypot(g, 1.0);
      g = d[m] - d[l] + e[l] / (g + std::copysign(r, g));
      double s = 1.0, c = 1.0, pp = 0.0;
      int i;
      bool deflated = false;
      for (i = m - 1; i >= l; --i) {
        const double f = s * e[i], b = c * e[i];
        r = std::hypot(f, g);
        e[i + 1] = r;
        if (r == 0.0) {  // underflow: split the matrix here and restart the sweep
          d[i + 1] -= pp;
          e[m] = 0.0;
          deflated = true;
          break;
        }
        s = f / r;
        c = g / r;
        g = d[i + 1] - pp;
        r = (d[i] - g) * s + 2.0 * c * b;
        pp = s * r;
        d[i + 1] = g + pp;
        g = c * r - b;
        for (int k = 0; k < nz; ++k) {
          double* zk = &z[size_t(k) * n];
          const double t = zk[i + 1];
          zk[i + 1] = s * zk[i] + c * t;
          zk[i] = c * zk[i] - s * t;
        }
      }
      if (deflated) continue;
      d[l] -= pp;
      e[l] = g;
      e[m] = 0.0;
    } while (true);
  }
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
