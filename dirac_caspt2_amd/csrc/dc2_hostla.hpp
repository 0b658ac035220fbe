// Host-side dense linear algebra for the small matrices of the Davidson solver.
#pragma once

#include <algorithm>
#include <cmath>
#include <complex>
#include <numeric>
#include <vector>

namespace dc2 {

typedef std::complex<double> zc;

// Cyclic Jacobi eigendecomposition of a Hermitian n x n matrix A (column-major, full
// storage, overwritten).  w: ascending eigenvalues, V: orthonormal eigenvectors
// (column k belongs to w[k]).  Each rotation diagonalises a 2x2 block
// [[a_pp, b], [conj(b), a_qq]], b = |b| e^{i phi}, by U = P R with
// P = diag(1, e^{-i phi}) and the real rotation R of the symmetric block
// [[a_pp, |b|], [|b|, a_qq]]; sweeps run until every off-diagonal element is below
// eps * sqrt(|a_pp a_qq|) or 1e-300 in magnitude.
inline void herm_jacobi(int n, zc* A, double* w, zc* V) {
  auto a = [&](int i, int j) -> zc& { return A[i + (size_t)n * j]; };
  auto v = [&](int i, int j) -> zc& { return V[i + (size_t)n * j]; };
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i) v(i, j) = (i == j) ? 1.0 : 0.0;
  for (int i = 0; i < n; ++i) a(i, i) = a(i, i).real();
  const double eps = 2.2e-16;
  for (int sweep = 0; sweep < 100; ++sweep) {
    bool rotated = false;
    for (int p = 0; p < n - 1; ++p)
      for (int q = p + 1; q < n; ++q) {
        const zc b = a(p, q);
        const double ab = std::abs(b);
        const double app = a(p, p).real(), aqq = a(q, q).real();
        if (ab <= 1e-300 || ab <= eps * std::sqrt(std::fabs(app * aqq)) * 0.5) continue;
        rotated = true;
        const zc ph = b / ab;                       // e^{i phi}
        const double theta = (aqq - app) / (2.0 * ab);
        const double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
        const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
        // columns: A <- A U, U = [[c, s], [-s e^{-i phi}, c e^{-i phi}]]
        const zc eq = std::conj(ph);
        for (int r = 0; r < n; ++r) {
          const zc xp = a(r, p), xq = a(r, q);
          a(r, p) = c * xp - s * eq * xq;
          a(r, q) = s * xp + c * eq * xq;
          const zc vp = v(r, p), vq = v(r, q);
          v(r, p) = c * vp - s * eq * vq;
          v(r, q) = s * vp + c * eq * vq;
        }
        // rows: A <- U^H A
        for (int r = 0; r < n; ++r) {
          const zc xp = a(p, r), xq = a(q, r);
          a(p, r) = c * xp - s * ph * xq;
          a(q, r) = s * xp + c * ph * xq;
        }
        a(p, q) = 0.0;
        a(q, p) = 0.0;
        a(p, p) = a(p, p).real();
        a(q, q) = a(q, q).real();
      }
    if (!rotated) break;
  }
  std::vector<int> idx(n);
  std::iota(idx.begin(), idx.end(), 0);
  std::vector<double> d(n);
  for (int i = 0; i < n; ++i) d[i] = a(i, i).real();
  std::stable_sort(idx.begin(), idx.end(), [&](int x, int y) { return d[x] < d[y]; });
  std::vector<zc> Vs((size_t)n * n);
  for (int k = 0; k < n; ++k) {
    w[k] = d[idx[k]];
    for (int i = 0; i < n; ++i) Vs[i + (size_t)n * k] = v(i, idx[k]);
  }
  std::copy(Vs.begin(), Vs.end(), V);
}

}  // namespace dc2
