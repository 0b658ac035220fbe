// Host-side dense linear algebra for the small matrices of the Davidson solver.
#pragma once

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdint>
#include <limits>
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
// eps * sqrt(|a_pp a_qq|) or 1e-300 in magnitude.  A <- A U is applied to the columns;
// rows p and q outside the 2x2 block then follow from hermiticity, the block itself is
// rotated explicitly.  Complex products are written out in real arithmetic (the library
// complex multiply carries NaN/Inf recovery that costs more than the rotation itself).
inline void herm_jacobi(int n, zc* A, double* w, zc* V) {
  double* a = reinterpret_cast<double*>(A);   // a[2*(i + n*j)] = Re A_ij, [.. + 1] = Im
  double* v = reinterpret_cast<double*>(V);
  const size_t ld = 2 * (size_t)n;
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i) {
      v[2 * i + ld * j] = (i == j) ? 1.0 : 0.0;
      v[2 * i + ld * j + 1] = 0.0;
    }
  for (int i = 0; i < n; ++i) a[2 * i + ld * i + 1] = 0.0;
  const double eps = 2.2e-16;
  // columns p and q of M (n rows): (x_p, x_q) <- (c x_p - s e x_q, s x_p + c e x_q), e = er + i ei
  auto rotate_cols = [n](double* xp, double* xq, double c, double s, double er, double ei) {
    for (int r = 0; r < n; ++r) {
      const double pr = xp[2 * r], pi = xp[2 * r + 1], qr = xq[2 * r], qi = xq[2 * r + 1];
      const double tr = er * qr - ei * qi, ti = er * qi + ei * qr;
      xp[2 * r] = c * pr - s * tr;
      xp[2 * r + 1] = c * pi - s * ti;
      xq[2 * r] = s * pr + c * tr;
      xq[2 * r + 1] = s * pi + c * ti;
    }
  };
  for (int sweep = 0; sweep < 100; ++sweep) {
    bool rotated = false;
    for (int p = 0; p < n - 1; ++p)
      for (int q = p + 1; q < n; ++q) {
        double* cp = a + ld * p;
        double* cq = a + ld * q;
        const double br = cq[2 * p], bi = cq[2 * p + 1];   // b = A_pq
        const double ab = std::hypot(br, bi);
        const double app = cp[2 * p], aqq = cq[2 * q];
        if (ab <= 1e-300 || ab <= eps * std::sqrt(std::fabs(app * aqq)) * 0.5) continue;
        rotated = true;
        const double phr = br / ab, phi = bi / ab;           // e^{i phi}
        const double theta = (aqq - app) / (2.0 * ab);
        const double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
        const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
        // A <- A U and V <- V U with U = [[c, s], [-s e^{-i phi}, c e^{-i phi}]]
        rotate_cols(cp, cq, c, s, phr, -phi);
        rotate_cols(v + ld * p, v + ld * q, c, s, phr, -phi);
        // rows p, q outside the block: A_pr = conj(A_rp), A_qr = conj(A_rq)
        for (int r = 0; r < n; ++r) {
          if (r == p || r == q) continue;
          double* cr = a + ld * r;
          cr[2 * p] = cp[2 * r];
          cr[2 * p + 1] = -cp[2 * r + 1];
          cr[2 * q] = cq[2 * r];
          cr[2 * q + 1] = -cq[2 * r + 1];
        }
        // the block: the diagonalised values of the real rotation of [[app, |b|], [|b|, aqq]]
        cp[2 * p] = app - t * ab;
        cp[2 * p + 1] = 0.0;
        cq[2 * q] = aqq + t * ab;
        cq[2 * q + 1] = 0.0;
        cp[2 * q] = cp[2 * q + 1] = 0.0;
        cq[2 * p] = cq[2 * p + 1] = 0.0;
      }
    if (!rotated) break;
  }
  std::vector<int> idx(n);
  std::iota(idx.begin(), idx.end(), 0);
  std::vector<double> d(n);
  for (int i = 0; i < n; ++i) d[i] = a[2 * i + ld * i];
  std::stable_sort(idx.begin(), idx.end(), [&](int x, int y) { return d[x] < d[y]; });
  std::vector<zc> Vs((size_t)n * n);
  for (int k = 0; k < n; ++k) {
    w[k] = d[idx[k]];
    for (int i = 0; i < n; ++i) Vs[i + (size_t)n * k] = V[i + (size_t)n * idx[k]];
  }
  std::copy(Vs.begin(), Vs.end(), V);
}

// The implicit QL iteration with Wilkinson shifts (the tql2 scheme) on the real symmetric
// tridiagonal (d, e), e[i] coupling i and i+1, without eigenvectors: d becomes the
// eigenvalues (unsorted, e is destroyed) and every sweep's Givens rotations are recorded,
// sweeps[3k .. 3k+2] = (m, istop, first): rotation i = m-1 .. istop of sweep k acts on
// (i, i+1) with cs/sn[first + m-1-i], so that the eigenvectors are accumulated elsewhere
// (on the host by herm_tql, on the device by the eigensolver).  e_m is negligible below eps
// times the running max of |d_l| + |e_l|.  False if an eigenvalue needs more than 90 sweeps
// or the rotations exceed max_rot.
inline bool tql_rotations(int n, std::vector<double>& d, std::vector<double>& e, std::vector<double>& cs,
                          std::vector<double>& sn, std::vector<int>& sweeps, size_t max_rot = SIZE_MAX) {
  // sqrt(a^2 + b^2), the scaled std::hypot only where the squares could leave the range
  auto hyp = [](double a, double b) {
    const double m = std::max(std::fabs(a), std::fabs(b));
    return (m < 1e150 && m > 1e-150) || m == 0.0 ? std::sqrt(a * a + b * b) : std::hypot(a, b);
  };
  cs.clear();
  sn.clear();
  sweeps.clear();
  const double eps = 2.2e-16;
  double tst1 = 0.0;
  for (int l = 0; l < n; ++l) {
    int iter = 0, m;
    tst1 = std::max(tst1, std::fabs(d[l]) + std::fabs(e[l]));
    do {
      for (m = l; m + 1 < n; ++m)
        if (std::fabs(e[m]) <= eps * tst1) break;
      if (m != l) {
        if (++iter > 90) return false;
        double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
        double r = hyp(g, 1.0);
        g = d[m] - d[l] + e[l] / (g + (g >= 0 ? r : -r));
        double s = 1.0, c = 1.0, pp = 0.0;
        const int first = (int)cs.size();
        int i;
        bool underflow = false;
        for (i = m - 1; i >= l; --i) {
          double f = s * e[i], b = c * e[i];
          r = hyp(f, g);
          e[i + 1] = r;
          if (r == 0.0) {
            d[i + 1] -= pp;
            e[m] = 0.0;
            underflow = true;
            break;
          }
          s = f / r;
          c = g / r;
          g = d[i + 1] - pp;
          r = (d[i] - g) * s + 2.0 * c * b;
          pp = s * r;
          d[i + 1] = g + pp;
          g = c * r - b;
          cs.push_back(c);
          sn.push_back(s);
        }
        const int istop = underflow ? i + 1 : l;
        if (istop < m) {
          sweeps.push_back(m);
          sweeps.push_back(istop);
          sweeps.push_back(first);
        }
        if (cs.size() > max_rot) return false;
        if (underflow) continue;
        d[l] -= pp;
        e[l] = g;
        e[m] = 0.0;
      }
    } while (m != l);
  }
  return true;
}

// Householder reduction of a Hermitian matrix to a real symmetric tridiagonal one.  For
// column k the reflector H_k = I - tau v v^H (tau = 2 / v^H v, real) maps x = A[k+1:, k] to
// alpha e_1 with alpha = -e^{i arg x_0} |x|; the trailing block becomes B - v q^H - q v^H with
// p = tau B v, q = p - (tau/2)(v^H p) v.  The complex subdiagonal e_k = alpha is made real by
// D = diag(delta), delta_{k+1} = delta_k e_k / |e_k|, so that A = Q D T D^H Q^H with T real;
// eigenvectors are Q D Z for the eigenvectors Z of T.  A (full column-major storage) is
// overwritten, v of column k kept in A[k+1:, k] with tau[k]; d, e: T's diagonal and
// subdiagonal (e[k] = |alpha_k|).
inline void herm_tridiag(int n, zc* A, std::vector<double>& d, std::vector<double>& e, std::vector<double>& tau,
                         std::vector<zc>& delta) {
  auto at = [&](int i, int j) -> zc& { return A[i + (size_t)n * j]; };
  auto mul = [](zc x, zc y) { return zc(x.real() * y.real() - x.imag() * y.imag(), x.real() * y.imag() + x.imag() * y.real()); };
  auto cmulc = [](zc x, zc y) {   // conj(x) * y
    return zc(x.real() * y.real() + x.imag() * y.imag(), x.real() * y.imag() - x.imag() * y.real());
  };
  d.assign(n, 0.0);
  e.assign(n, 0.0);
  tau.assign(n, 0.0);
  std::vector<zc> sub(n, zc(0.0, 0.0)), p(n), q(n);
  for (int k = 0; k + 1 < n; ++k) {
    const int m = n - k - 1;
    zc* x = &at(k + 1, k);
    double xn2 = 0.0, tail2 = 0.0;
    for (int i = 0; i < m; ++i) xn2 += std::norm(x[i]);
    tail2 = xn2 - std::norm(x[0]);
    if (m == 1 || tail2 <= 0.0) {   // already tridiagonal in this column
      sub[k] = x[0];
      tau[k] = 0.0;
      continue;
    }
    const double xn = std::sqrt(xn2), ax0 = std::abs(x[0]);
    const zc ph = ax0 > 0.0 ? x[0] / ax0 : zc(1.0, 0.0);
    const zc alpha = -xn * ph;
    x[0] -= alpha;                    // v = x - alpha e_1, kept in place of x
    const double t = 2.0 / (2.0 * xn2 + 2.0 * ax0 * xn);
    tau[k] = t;
    sub[k] = alpha;
    // p = tau B v over the trailing block B = A[k+1:, k+1:] (full Hermitian storage)
    for (int i = 0; i < m; ++i) p[i] = zc(0.0, 0.0);
    for (int j = 0; j < m; ++j) {
      const zc vj = x[j];
      const zc* col = &at(k + 1, k + 1 + j);
      for (int i = 0; i < m; ++i) p[i] += mul(col[i], vj);
    }
    zc vhp(0.0, 0.0);
    for (int i = 0; i < m; ++i) {
      p[i] *= t;
      vhp += cmulc(x[i], p[i]);
    }
    const double K = 0.5 * t * vhp.real();
    for (int i = 0; i < m; ++i) q[i] = p[i] - K * x[i];
    for (int j = 0; j < m; ++j) {
      zc* col = &at(k + 1, k + 1 + j);
      const zc cqj = std::conj(q[j]), cvj = std::conj(x[j]);
      for (int i = 0; i < m; ++i) col[i] -= mul(x[i], cqj) + mul(q[i], cvj);
    }
  }
  for (int i = 0; i < n; ++i) d[i] = at(i, i).real();
  delta.assign(n, zc(1.0, 0.0));
  for (int k = 0; k + 1 < n; ++k) {
    const double ae = std::abs(sub[k]);
    e[k] = ae;
    delta[k + 1] = ae > 0.0 ? mul(delta[k], sub[k] / ae) : delta[k];
  }
}

// herm_tridiag on the lower triangle alone (the upper one is neither read nor written): the
// same reflectors, half the flops; the rounding differs from herm_tridiag's.
inline void herm_tridiag_lower(int n, zc* A, std::vector<double>& d, std::vector<double>& e, std::vector<double>& tau,
                         std::vector<zc>& delta) {
  auto at = [&](int i, int j) -> zc& { return A[i + (size_t)n * j]; };
  auto mul = [](zc x, zc y) { return zc(x.real() * y.real() - x.imag() * y.imag(), x.real() * y.imag() + x.imag() * y.real()); };
  auto cmulc = [](zc x, zc y) {   // conj(x) * y
    return zc(x.real() * y.real() + x.imag() * y.imag(), x.real() * y.imag() - x.imag() * y.real());
  };
  d.assign(n, 0.0);
  e.assign(n, 0.0);
  tau.assign(n, 0.0);
  std::vector<zc> sub(n, zc(0.0, 0.0)), p(n), q(n);
  for (int k = 0; k + 1 < n; ++k) {
    const int m = n - k - 1;
    zc* x = &at(k + 1, k);
    double xn2 = 0.0, tail2 = 0.0;
    for (int i = 0; i < m; ++i) xn2 += std::norm(x[i]);
    tail2 = xn2 - std::norm(x[0]);
    if (m == 1 || tail2 <= 0.0) {   // already tridiagonal in this column
      sub[k] = x[0];
      tau[k] = 0.0;
      continue;
    }
    const double xn = std::sqrt(xn2), ax0 = std::abs(x[0]);
    const zc ph = ax0 > 0.0 ? x[0] / ax0 : zc(1.0, 0.0);
    const zc alpha = -xn * ph;
    x[0] -= alpha;                    // v = x - alpha e_1, kept in place of x
    const double t = 2.0 / (2.0 * xn2 + 2.0 * ax0 * xn);
    tau[k] = t;
    sub[k] = alpha;
    // p = tau B v from the lower triangle of B = A[k+1:, k+1:]: column j adds B_ij v_j to
    // p_i (i > j) and conj(B_ij) v_i = B_ji v_i to p_j
    for (int i = 0; i < m; ++i) p[i] = zc(0.0, 0.0);
    for (int j = 0; j < m; ++j) {
      const zc vj = x[j];
      const zc* col = &at(k + 1, k + 1 + j);
      zc acc = col[j].real() * vj;
      for (int i = j + 1; i < m; ++i) {
        p[i] += mul(col[i], vj);
        acc += cmulc(col[i], x[i]);
      }
      p[j] += acc;
    }
    zc vhp(0.0, 0.0);
    for (int i = 0; i < m; ++i) {
      p[i] *= t;
      vhp += cmulc(x[i], p[i]);
    }
    const double K = 0.5 * t * vhp.real();
    for (int i = 0; i < m; ++i) q[i] = p[i] - K * x[i];
    for (int j = 0; j < m; ++j) {
      zc* col = &at(k + 1, k + 1 + j);
      const zc cqj = std::conj(q[j]), cvj = std::conj(x[j]);
      for (int i = j; i < m; ++i) col[i] -= mul(x[i], cqj) + mul(q[i], cvj);
    }
  }
  for (int i = 0; i < n; ++i) d[i] = at(i, i).real();
  delta.assign(n, zc(1.0, 0.0));
  for (int k = 0; k + 1 < n; ++k) {
    const double ae = std::abs(sub[k]);
    e[k] = ae;
    delta[k + 1] = ae > 0.0 ? mul(delta[k], sub[k] / ae) : delta[k];
  }
}

// Hermitian eigendecomposition by the reduction above and the implicit QL iteration with
// Wilkinson shifts (the tql2 scheme), for the small projected problems of the Davidson
// solver, where it needs about a fifth of the flops of Jacobi.  Same contract as
// herm_jacobi: A (column-major, overwritten), w ascending, V orthonormal eigenvectors.
// Returns false if an eigenvalue needs more than 90 QL steps.
inline bool herm_tql(int n, zc* A, double* w, zc* V) {
  auto at = [&](int i, int j) -> zc& { return A[i + (size_t)n * j]; };
  auto mul = [](zc x, zc y) { return zc(x.real() * y.real() - x.imag() * y.imag(), x.real() * y.imag() + x.imag() * y.real()); };
  auto cmulc = [](zc x, zc y) {   // conj(x) * y
    return zc(x.real() * y.real() + x.imag() * y.imag(), x.real() * y.imag() - x.imag() * y.real());
  };
  std::vector<double> d, e, tau;
  std::vector<zc> delta;
  herm_tridiag(n, A, d, e, tau, delta);
  // implicit QL on (d, e), its rotations accumulated into the real Z
  std::vector<double> cs, sn;
  std::vector<int> sweeps;
  if (!tql_rotations(n, d, e, cs, sn, sweeps)) return false;
  std::vector<double> z((size_t)n * n, 0.0);
  for (int i = 0; i < n; ++i) z[i + (size_t)n * i] = 1.0;
  for (size_t k = 0; k < sweeps.size(); k += 3) {
    const int m = sweeps[k], istop = sweeps[k + 1], first = sweeps[k + 2];
    for (int i = m - 1; i >= istop; --i) {
      const double c = cs[first + m - 1 - i], s = sn[first + m - 1 - i];
      double* z0 = &z[(size_t)n * i];
      double* z1 = &z[(size_t)n * (i + 1)];
      for (int r = 0; r < n; ++r) {
        const double zf = z1[r];
        z1[r] = s * z0[r] + c * zf;
        z0[r] = c * z0[r] - s * zf;
      }
    }
  }
  // eigenvectors: X = Q D Z, Q = H_0 H_1 ... applied from the last reflector
  std::vector<int> idx(n);
  std::iota(idx.begin(), idx.end(), 0);
  std::stable_sort(idx.begin(), idx.end(), [&](int x, int y) { return d[x] < d[y]; });
  for (int c2 = 0; c2 < n; ++c2) {
    const int src = idx[c2];
    w[c2] = d[src];
    zc* col = V + (size_t)n * c2;
    for (int i = 0; i < n; ++i) col[i] = delta[i] * z[i + (size_t)n * src];
    for (int k = n - 2; k >= 0; --k) {
      if (tau[k] == 0.0) continue;
      const int m = n - k - 1;
      const zc* v = &at(k + 1, k);
      zc s(0.0, 0.0);
      for (int i = 0; i < m; ++i) s += cmulc(v[i], col[k + 1 + i]);
      s *= tau[k];
      for (int i = 0; i < m; ++i) col[k + 1 + i] -= mul(v[i], s);
    }
  }
  return true;
}

// herm_tql, or herm_jacobi on a copy of A if QL does not converge
inline void herm_eig(int n, zc* A, double* w, zc* V) {
  std::vector<zc> keep(A, A + (size_t)n * n);
  if (!herm_tql(n, A, w, V)) herm_jacobi(n, keep.data(), w, V);
}

// All n eigenvalues (w, ascending) and the eigenvectors of the nvec lowest only (V: n x nvec),
// for the Rayleigh matrices of the Davidson solver, which uses the lowest block alone.  The
// reduction of herm_tridiag_lower (A's lower triangle) and the QL of herm_tql; with Z = R_1 R_2 ... R_K the product of QL's plane rotations in
// the order they were made, column j of Z is R_1 (R_2 (... (R_K e_j))), so the rotations are
// applied to the nvec unit vectors from the last to the first (O(K nvec) instead of the O(K n)
// of accumulating Z), then D and the reflectors (O(n^2 nvec) instead of O(n^3)).  On a QL
// failure, herm_jacobi on the saved matrix and its first nvec eigenvectors.
inline void herm_eig_low(int n, zc* A, double* w, zc* V, int nvec) {
  nvec = std::max(0, std::min(nvec, n));
  std::vector<zc> keep(A, A + (size_t)n * n);
  std::vector<double> d, e, tau, cs, sn;
  std::vector<zc> delta;
  std::vector<int> sweeps;
  herm_tridiag_lower(n, A, d, e, tau, delta);
  cs.reserve(2 * (size_t)n * n);
  sn.reserve(2 * (size_t)n * n);
  if (!tql_rotations(n, d, e, cs, sn, sweeps)) {
    std::vector<zc> Vall((size_t)n * n);
    herm_jacobi(n, keep.data(), w, Vall.data());
    std::copy(Vall.begin(), Vall.begin() + (size_t)n * nvec, V);
    return;
  }
  auto at = [&](int i, int j) -> zc& { return A[i + (size_t)n * j]; };
  auto mul = [](zc x, zc y) { return zc(x.real() * y.real() - x.imag() * y.imag(), x.real() * y.imag() + x.imag() * y.real()); };
  auto cmulc = [](zc x, zc y) {   // conj(x) * y
    return zc(x.real() * y.real() + x.imag() * y.imag(), x.real() * y.imag() - x.imag() * y.real());
  };
  std::vector<int> idx(n);
  std::iota(idx.begin(), idx.end(), 0);
  std::stable_sort(idx.begin(), idx.end(), [&](int x, int y) { return d[x] < d[y]; });
  for (int c = 0; c < n; ++c) w[c] = d[idx[c]];
  // z (n x nvec, row-major by row r: z[r * nvec + c]) = Z e_{idx[c]}
  std::vector<double> z((size_t)n * nvec, 0.0);
  for (int c = 0; c < nvec; ++c) z[(size_t)idx[c] * nvec + c] = 1.0;
  for (size_t k = sweeps.size(); k >= 3; k -= 3) {
    const int m = sweeps[k - 3], istop = sweeps[k - 2], first = sweeps[k - 1];
    for (int i = istop; i <= m - 1; ++i) {   // the sweep's rotations in reverse (i ascending)
      const double c = cs[first + m - 1 - i], s = sn[first + m - 1 - i];
      double* z0 = &z[(size_t)nvec * i];
      double* z1 = &z[(size_t)nvec * (i + 1)];
      for (int col = 0; col < nvec; ++col) {
        const double a = z0[col], b = z1[col];
        z0[col] = c * a + s * b;
        z1[col] = c * b - s * a;
      }
    }
  }
  for (int c2 = 0; c2 < nvec; ++c2) {
    zc* col = V + (size_t)n * c2;
    for (int i = 0; i < n; ++i) col[i] = delta[i] * z[(size_t)nvec * i + c2];
    for (int k = n - 2; k >= 0; --k) {
      if (tau[k] == 0.0) continue;
      const int m = n - k - 1;
      const zc* v = &at(k + 1, k);
      zc s(0.0, 0.0);
      for (int i = 0; i < m; ++i) s += cmulc(v[i], col[k + 1 + i]);
      s *= tau[k];
      for (int i = 0; i < m; ++i) col[k + 1 + i] -= mul(v[i], s);
    }
  }
}


// Eigendecomposition of a Hermitian n x n matrix by the sequence of LAPACK's ZHEEV('V', 'U')
// (the reference's cdiag, diag.F90:66-128): Householder reduction of the upper triangle from
// the last column (zhetd2), the unitary Q from the reflectors (zungtr / zung2l), implicit QL or
// QR iterations with Wilkinson shifts on the real tridiagonal (zsteqr: the end with the smaller
// diagonal element starts the sweeps, rotations from dlartg in the form that keeps r > 0 unless
// |f| > |g| with f < 0, 2 x 2 blocks by dlaev2), then a selection sort of the eigenpairs.  The
// eigenvectors come out with ZHEEV's signs and phases, which the IVO coefficients depend on.
// A (column-major, only the upper triangle read) is not modified; w ascending, V column k for w[k].
// Returns false when the iteration does not converge in 30 n sweeps.
inline bool herm_eig_zheev(int n, const zc* A, double* w, zc* V) {
  auto at = [n](std::vector<zc>& M, int i, int j) -> zc& { return M[(size_t)i + (size_t)n * j]; };
  if (n <= 0) return true;
  if (n == 1) {
    w[0] = A[0].real();
    V[0] = zc(1.0, 0.0);
    return true;
  }
  std::vector<zc> T(A, A + (size_t)n * n), tau(n - 1), x(n), v(n), wv(n);
  std::vector<double> d(n), e(n - 1);
  // zhetd2, UPLO = 'U' (0-based: reflector of column i+1 annihilates rows 0 .. i-2)
  at(T, n - 1, n - 1) = at(T, n - 1, n - 1).real();
  for (int i = n - 1; i >= 1; --i) {
    zc alpha = at(T, i - 1, i);
    double xnorm = 0.0;
    for (int r = 0; r < i - 1; ++r) xnorm = std::hypot(xnorm, std::abs(at(T, r, i)));
    zc taui(0.0, 0.0);
    if (!(xnorm == 0.0 && alpha.imag() == 0.0)) {   // zlarfg
      const double beta = -std::copysign(std::sqrt(std::norm(alpha) + xnorm * xnorm), alpha.real());
      taui = zc((beta - alpha.real()) / beta, -alpha.imag() / beta);
      const zc sc = 1.0 / (alpha - beta);
      for (int r = 0; r < i - 1; ++r) at(T, r, i) *= sc;
      alpha = zc(beta, 0.0);
    }
    e[i - 1] = alpha.real();
    if (taui != zc(0.0, 0.0)) {
      at(T, i - 1, i) = 1.0;
      for (int r = 0; r < i; ++r) v[r] = at(T, r, i);
      // x = taui * H v with H the Hermitian matrix of the upper triangle of T[0:i, 0:i]
      for (int r = 0; r < i; ++r) {
        zc acc(0.0, 0.0);
        for (int c = 0; c < i; ++c) {
          const zc h = r < c ? at(T, r, c) : (r == c ? zc(at(T, r, r).real(), 0.0) : std::conj(at(T, c, r)));
          acc += h * v[c];
        }
        x[r] = taui * acc;
      }
      zc dot(0.0, 0.0);
      for (int r = 0; r < i; ++r) dot += std::conj(x[r]) * v[r];
      const zc al = -0.5 * taui * dot;
      for (int r = 0; r < i; ++r) wv[r] = x[r] + al * v[r];
      for (int c = 0; c < i; ++c)
        for (int r = 0; r <= c; ++r) at(T, r, c) -= v[r] * std::conj(wv[c]) + wv[r] * std::conj(v[c]);
      for (int r = 0; r < i; ++r) at(T, r, r) = at(T, r, r).real();
    } else {
      at(T, i - 1, i - 1) = at(T, i - 1, i - 1).real();
    }
    at(T, i - 1, i) = e[i - 1];
    d[i] = at(T, i, i).real();
    tau[i - 1] = taui;
  }
  d[0] = at(T, 0, 0).real();
  // zungtr (UPLO = 'U'): the reflectors shifted one column left, then zung2l on the
  // leading (n-1) x (n-1) block; the last row and column are those of the unit matrix
  std::vector<zc> Q((size_t)n * n, zc(0.0, 0.0));
  for (int j = 0; j < n - 1; ++j)
    for (int r = 0; r < j; ++r) at(Q, r, j) = at(T, r, j + 1);
  at(Q, n - 1, n - 1) = 1.0;
  const int m = n - 1;
  for (int i = 1; i <= m; ++i) {   // H(i) applied to Q[0:i, 0:i-1] from the left
    const int ii = i - 1, rows = i;
    at(Q, rows - 1, ii) = 1.0;
    for (int c = 0; c < ii; ++c) {
      zc acc(0.0, 0.0);
      for (int r = 0; r < rows; ++r) acc += std::conj(at(Q, r, c)) * at(Q, r, ii);
      const zc f = tau[i - 1] * std::conj(acc);
      for (int r = 0; r < rows; ++r) at(Q, r, c) -= at(Q, r, ii) * f;
    }
    for (int r = 0; r < rows - 1; ++r) at(Q, r, ii) *= -tau[i - 1];
    at(Q, rows - 1, ii) = 1.0 - tau[i - 1];
    for (int r = rows; r < m; ++r) at(Q, r, ii) = 0.0;
  }
  // zsteqr (1-based indices as in LAPACK; d1/e1 address them)
  const double eps = 0.5 * std::numeric_limits<double>::epsilon(), eps2 = eps * eps;
  const double safmin = std::numeric_limits<double>::min();
  auto d1 = [&](int i) -> double& { return d[i - 1]; };
  auto e1 = [&](int i) -> double& { return e[i - 1]; };
  auto lartg = [](double f, double g, double& c, double& s, double& r) {
    if (g == 0.0) {
      c = 1.0; s = 0.0; r = f;
    } else if (f == 0.0) {
      c = 0.0; s = 1.0; r = g;
    } else {
      r = std::sqrt(f * f + g * g);
      c = f / r; s = g / r;
      if (std::fabs(f) > std::fabs(g) && c < 0.0) { c = -c; s = -s; r = -r; }
    }
  };
  auto laev2 = [](double a, double b, double c, double& rt1, double& rt2, double& cs1, double& sn1) {
    const double sm = a + c, df = a - c, adf = std::fabs(df), tb = b + b, ab = std::fabs(tb);
    const double acmx = std::fabs(a) > std::fabs(c) ? a : c, acmn = std::fabs(a) > std::fabs(c) ? c : a;
    double rt;
    if (adf > ab) rt = adf * std::sqrt(1.0 + (ab / adf) * (ab / adf));
    else if (adf < ab) rt = ab * std::sqrt(1.0 + (adf / ab) * (adf / ab));
    else rt = ab * std::sqrt(2.0);
    int sgn1, sgn2;
    if (sm < 0.0) { rt1 = 0.5 * (sm - rt); sgn1 = -1; rt2 = (acmx / rt1) * acmn - (b / rt1) * b; }
    else if (sm > 0.0) { rt1 = 0.5 * (sm + rt); sgn1 = 1; rt2 = (acmx / rt1) * acmn - (b / rt1) * b; }
    else { rt1 = 0.5 * rt; rt2 = -0.5 * rt; sgn1 = 1; }
    double cs;
    if (df >= 0.0) { cs = df + rt; sgn2 = 1; } else { cs = df - rt; sgn2 = -1; }
    if (std::fabs(cs) > ab) {
      const double ct = -tb / cs;
      sn1 = 1.0 / std::sqrt(1.0 + ct * ct); cs1 = ct * sn1;
    } else if (ab == 0.0) {
      cs1 = 1.0; sn1 = 0.0;
    } else {
      const double tn = -cs / tb;
      cs1 = 1.0 / std::sqrt(1.0 + tn * tn); sn1 = tn * cs1;
    }
    if (sgn1 == sgn2) { const double tn = cs1; cs1 = -sn1; sn1 = tn; }
  };
  // columns j0 .. j0+k-1 (1-based) of Q rotated by (c_j, s_j), j = 1 .. k-1, backward or forward
  auto lasr = [&](bool backward, int j0, int k, const double* cw, const double* sw) {
    for (int t = 0; t < k - 1; ++t) {
      const int j = backward ? k - 2 - t : t;
      const double ct = cw[j], st = sw[j];
      zc* a = &at(Q, 0, j0 - 1 + j);
      zc* b = &at(Q, 0, j0 + j);
      for (int r = 0; r < n; ++r) {
        const zc tmp = b[r];
        b[r] = ct * tmp - st * a[r];
        a[r] = st * tmp + ct * a[r];
      }
    }
  };
  std::vector<double> cw(n), sw(n);
  const int nmaxit = 30 * n;
  int jtot = 0, l1 = 1;
  while (l1 <= n) {
    if (l1 > 1) e1(l1 - 1) = 0.0;
    int mm = n;
    for (int q = l1; q <= n - 1; ++q) {
      const double tst = std::fabs(e1(q));
      if (tst == 0.0) { mm = q; break; }
      if (tst <= (std::sqrt(std::fabs(d1(q))) * std::sqrt(std::fabs(d1(q + 1)))) * eps) { e1(q) = 0.0; mm = q; break; }
    }
    int l = l1, lsv = l, lend = mm, lendsv = lend;
    l1 = mm + 1;
    if (lend == l) continue;
    if (std::fabs(d1(lend)) < std::fabs(d1(l))) { lend = lsv; l = lendsv; }
    if (lend > l) {   // QL iteration
      for (;;) {
        int m2 = lend;
        if (l != lend)
          for (int q = l; q <= lend - 1; ++q) {
            const double tst = e1(q) * e1(q);
            if (tst <= (eps2 * std::fabs(d1(q))) * std::fabs(d1(q + 1)) + safmin) { m2 = q; break; }
          }
        if (m2 < lend) e1(m2) = 0.0;
        double p = d1(l);
        if (m2 == l) {
          d1(l) = p; ++l;
          if (l <= lend) continue;
          break;
        }
        if (m2 == l + 1) {
          double rt1, rt2, c, s;
          laev2(d1(l), e1(l), d1(l + 1), rt1, rt2, c, s);
          lasr(true, l, 2, &c, &s);
          d1(l) = rt1; d1(l + 1) = rt2; e1(l) = 0.0; l += 2;
          if (l <= lend) continue;
          break;
        }
        if (jtot == nmaxit) break;
        ++jtot;
        double g = (d1(l + 1) - p) / (2.0 * e1(l));
        double r = std::hypot(g, 1.0);
        g = d1(m2) - p + (e1(l) / (g + std::copysign(r, g)));
        double s = 1.0, c = 1.0;
        p = 0.0;
        for (int i = m2 - 1; i >= l; --i) {
          const double f = s * e1(i), b = c * e1(i);
          lartg(g, f, c, s, r);
          if (i != m2 - 1) e1(i + 1) = r;
          g = d1(i + 1) - p;
          r = (d1(i) - g) * s + 2.0 * c * b;
          p = s * r;
          d1(i + 1) = g + p;
          g = c * r - b;
          cw[i - 1] = c; sw[i - 1] = -s;
        }
        lasr(true, l, m2 - l + 1, cw.data() + l - 1, sw.data() + l - 1);
        d1(l) -= p; e1(l) = g;
      }
    } else {          // QR iteration
      for (;;) {
        int m2 = lend;
        if (l != lend)
          for (int q = l; q >= lend + 1; --q) {
            const double tst = e1(q - 1) * e1(q - 1);
            if (tst <= (eps2 * std::fabs(d1(q))) * std::fabs(d1(q - 1)) + safmin) { m2 = q; break; }
          }
        if (m2 > lend) e1(m2 - 1) = 0.0;
        double p = d1(l);
        if (m2 == l) {
          d1(l) = p; --l;
          if (l >= lend) continue;
          break;
        }
        if (m2 == l - 1) {
          double rt1, rt2, c, s;
          laev2(d1(l - 1), e1(l - 1), d1(l), rt1, rt2, c, s);
          lasr(false, l - 1, 2, &c, &s);
          d1(l - 1) = rt1; d1(l) = rt2; e1(l - 1) = 0.0; l -= 2;
          if (l >= lend) continue;
          break;
        }
        if (jtot == nmaxit) break;
        ++jtot;
        double g = (d1(l - 1) - p) / (2.0 * e1(l - 1));
        double r = std::hypot(g, 1.0);
        g = d1(m2) - p + (e1(l - 1) / (g + std::copysign(r, g)));
        double s = 1.0, c = 1.0;
        p = 0.0;
        for (int i = m2; i <= l - 1; ++i) {
          const double f = s * e1(i), b = c * e1(i);
          lartg(g, f, c, s, r);
          if (i != m2) e1(i - 1) = r;
          g = d1(i) - p;
          r = (d1(i + 1) - g) * s + 2.0 * c * b;
          p = s * r;
          d1(i) = g + p;
          g = c * r - b;
          cw[i - 1] = c; sw[i - 1] = s;
        }
        lasr(false, m2, l - m2 + 1, cw.data() + m2 - 1, sw.data() + m2 - 1);
        d1(l) -= p; e1(l - 1) = g;
      }
    }
    if (jtot >= nmaxit) return false;
  }
  for (int ii = 2; ii <= n; ++ii) {   // selection sort, ascending
    const int i = ii - 1;
    int k = i;
    double p = d1(i);
    for (int j = ii; j <= n; ++j)
      if (d1(j) < p) { k = j; p = d1(j); }
    if (k != i) {
      d1(k) = d1(i); d1(i) = p;
      for (int r = 0; r < n; ++r) std::swap(at(Q, r, i - 1), at(Q, r, k - 1));
    }
  }
  std::copy(d.begin(), d.end(), w);
  std::copy(Q.begin(), Q.end(), V);
  return true;
}

}  // namespace dc2
