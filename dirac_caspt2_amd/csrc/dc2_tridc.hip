// Eigenpairs of a real symmetric tridiagonal matrix by divide and conquer, every step on the
// device (the tridiagonal stage of this library's Hermitian eigensolver, dc2_heev.hip).
//
// Tree.  Node j of depth l covers rows [lo(l, j), lo(l, j+1)) with lo(l, j) = (j n) >> l; its
// children are nodes 2j and 2j+1 of depth l+1, split at m = lo(l+1, 2j+1).  The leaves sit at
// the smallest depth whose nodes hold at most DC_LEAF rows.  Every leaf boundary m is the
// split point of exactly one node, where T = diag(T1, T2) + |beta| u u^T with beta = e[m-1],
// u = e_{m-1} + sign(beta) e_m, and T1, T2 carry d[m-1] - |beta|, d[m] - |beta| (Cuppen's
// tearing); all tears are applied to the leaves' diagonals at once.
//
// Leaves: the implicit QL iteration with Wilkinson shifts (the tql2 scheme) in one wave, the
// eigenvectors accumulated in LDS, sorted ascending.
//
// Merge of a node (eigenpairs (D1, Q1), (D2, Q2) of its children, each ascending):
//   z = Q^T u / sqrt(2), rho = 2 |beta|;  the problem is diag(D) + rho z z^T.
//   deflation (the scheme of LAPACK's dlaed2): in ascending order of D, an entry with
//     rho |z_j| <= tol is an eigenpair as it stands; two entries whose Givens rotation
//     zeroes one z component with |(d_j - d_p) c s| <= tol are rotated (the rotation applied
//     to the columns of Q) and the first of them is an eigenpair;  tol = 8 eps max(|D|, |z|).
//   secular equation 1 + rho sum_i z_i^2 / (d_i - lambda) = 0 for the K remaining entries:
//     root j lies in (d_j, d_j+1) (the last in (d_K-1, d_K-1 + rho |z|^2)); it is held as
//     tau relative to the nearer pole d_o, and found by Newton's method on
//     F(tau) = tau f(tau) (the pole at the origin removed), bisection where a step leaves the
//     bracket, until |F| is below the rounding error of its evaluation; all differences
//     d_i - lambda are formed as (d_i - d_o) - tau.
//   eigenvectors by the Loewner formula (Gu and Eisenstat): zhat_i^2 = prod_j (lambda_j - d_i)
//     / (rho prod_{j != i} (d_j - d_i)), v_j = (zhat_i / (d_i - lambda_j))_i normalised, so
//     that they are orthogonal to working accuracy whatever the accuracy of the roots.
//   new eigenvectors: (rotated Q)[:, non-deflated] V by an FP64 MFMA product, deflated
//     columns copied, all in ascending order of the new eigenvalues.
// The nodes of one depth are merged together (blockIdx.y / z = node), one launch per step;
// sizes stay on the device, so the whole solve is a fixed sequence of launches.
#include "dc2_kernels.hpp"

namespace dc2 {

typedef double d4 __attribute__((ext_vector_type(4)));

namespace {

constexpr int DC_LEAF = 16;
constexpr int DC_NMAX = 2048;
constexpr int DC_ITMAX = 200;   // secular Newton/bisection steps per root
constexpr double DC_EPS = 1.1102230246251565e-16;   // 2^-53

__host__ __device__ inline int node_lo(int n, int depth, int j) { return (int)(((int64_t)j * n) >> depth); }

// sqrt(a^2 + b^2), scaled only where the squares could leave the range
__device__ inline double dc_hyp(double a, double b) {
  const double m = fmax(fabs(a), fabs(b));
  if ((m < 1e150 && m > 1e-150) || m == 0.0) return sqrt(a * a + b * b);
  const double x = a / m, y = b / m;
  return m * sqrt(x * x + y * y);
}

// sum over the wave, the same value in every lane (lane 0's, so that every lane takes the
// same branches)
__device__ inline double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return __shfl(v, 0, 64);
}
__device__ inline double wave_prod(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v *= __shfl_xor(v, o, 64);
  return __shfl(v, 0, 64);
}

struct DcWork {
  double* ndd;   // non-deflated d, z (ascending), at a + i
  double* ndz;
  double* dfd;   // deflated eigenvalues, at a + f
  double* rc;    // rotations (c, s), at a + t
  double* rs;
  double* tau;   // root j: lambda = ndd[a + org] + tau
  double* zh;    // zhat
  int* ndc;      // node-local column of Q of non-deflated entry i
  int* dfc;      // node-local column of Q of deflated entry f
  int* rp;       // rotation t acts on node-local columns (rp, rq)
  int* rq;
  int* org;      // root j: index of its origin pole among the non-deflated entries
  int* pos;      // node-local output column of root j (at a + j) and deflated entry f (at a + K + f)
  int* cnt;      // node j: {K, deflated, rotations}
};

// ------------------------------------------------------------------ leaves
// One wave per leaf: lane r < N owns row r of the leaf's eigenvectors (LDS, column-major).
// The scalar QL recurrence is executed by every lane on the shared d/e (identical values),
// so that no lane waits for another.
__global__ __launch_bounds__(64) void k_dc_leaf(int n, int depth, const double* __restrict__ din,
                                                const double* __restrict__ e, double* __restrict__ Dout,
                                                double* __restrict__ Zout, int* __restrict__ info) {
  __shared__ double sd[DC_LEAF], se[DC_LEAF], Zl[DC_LEAF * DC_LEAF];
  const int j = blockIdx.x, lane = threadIdx.x;
  const int a = node_lo(n, depth, j), c = node_lo(n, depth, j + 1), N = c - a;
  if (N <= 0) return;
  if (lane < N) {
    double d = din[a + lane];
    if (lane == 0 && a > 0) d -= fabs(e[a - 1]);        // tears at the leaf's two ends
    if (lane == N - 1 && c < n) d -= fabs(e[c - 1]);
    sd[lane] = d;
    se[lane] = lane < N - 1 ? e[a + lane] : 0.0;
    for (int col = 0; col < N; ++col) Zl[col * DC_LEAF + lane] = col == lane ? 1.0 : 0.0;
  }
  __syncthreads();
  const bool own = lane < N;
  bool ok = true;
  double tst1 = 0.0;
  for (int l = 0; l < N && ok; ++l) {
    int iter = 0, m;
    tst1 = fmax(tst1, fabs(sd[l]) + fabs(se[l]));
    do {
      for (m = l; m + 1 < N; ++m)
        if (fabs(se[m]) <= 2.0 * DC_EPS * tst1) break;
      if (m != l) {
        if (++iter > 90) {
          ok = false;
          break;
        }
        double g = (sd[l + 1] - sd[l]) / (2.0 * se[l]);
        double r = dc_hyp(g, 1.0);
        g = sd[m] - sd[l] + se[l] / (g + (g >= 0.0 ? r : -r));
        double s = 1.0, cc = 1.0, p = 0.0;
        bool under = false;
        for (int i = m - 1; i >= l; --i) {
          const double f = s * se[i], b = cc * se[i];
          r = dc_hyp(f, g);
          se[i + 1] = r;
          if (r == 0.0) {
            sd[i + 1] -= p;
            se[m] = 0.0;
            under = true;
            break;
          }
          s = f / r;
          cc = g / r;
          g = sd[i + 1] - p;
          r = (sd[i] - g) * s + 2.0 * cc * b;
          p = s * r;
          sd[i + 1] = g + p;
          g = cc * r - b;
          if (own) {
            const double h = Zl[(i + 1) * DC_LEAF + lane], x = Zl[i * DC_LEAF + lane];
            Zl[(i + 1) * DC_LEAF + lane] = s * x + cc * h;
            Zl[i * DC_LEAF + lane] = cc * x - s * h;
          }
        }
        if (under) continue;
        sd[l] -= p;
        se[l] = g;
        se[m] = 0.0;
      }
    } while (m != l);
  }
  if (!ok) {
    if (lane == 0) info[0] = 1;
    return;
  }
  // ascending order (selection sort, columns swapped with their eigenvalues)
  for (int i = 0; i + 1 < N; ++i) {
    int k = i;
    double p = sd[i];
    for (int q = i + 1; q < N; ++q)
      if (sd[q] < p) {
        k = q;
        p = sd[q];
      }
    if (k != i) {
      const double di = sd[i];
      sd[k] = di;
      sd[i] = p;
      if (own) {
        const double t = Zl[i * DC_LEAF + lane];
        Zl[i * DC_LEAF + lane] = Zl[k * DC_LEAF + lane];
        Zl[k * DC_LEAF + lane] = t;
      }
    }
  }
  if (own) {
    Dout[a + lane] = sd[lane];
    for (int col = 0; col < N; ++col) Zout[(a + lane) + (int64_t)n * (a + col)] = Zl[col * DC_LEAF + lane];
  }
}

// ------------------------------------------------------------------ deflation
// One workgroup per node.  A node whose split leaves a child empty passes its child through
// (every entry "deflated", in order).
__global__ __launch_bounds__(256) void k_dc_deflate(int n, int depth, const double* __restrict__ e,
                                                    const double* __restrict__ Din, const double* __restrict__ Q,
                                                    DcWork w) {
  __shared__ double sd[DC_NMAX], sz[DC_NMAX];
  __shared__ int srt[DC_NMAX];
  __shared__ double red[2][256];
  const int j = blockIdx.x, tid = threadIdx.x;
  const int a = node_lo(n, depth, j), m = node_lo(n, depth + 1, 2 * j + 1), c = node_lo(n, depth, j + 1);
  const int N = c - a, n1 = m - a, n2 = c - m;
  if (!(a < m && m < c)) {
    for (int i = tid; i < N; i += 256) {
      w.dfd[a + i] = Din[a + i];
      w.dfc[a + i] = i;
    }
    if (tid == 0) {
      w.cnt[3 * j] = 0;
      w.cnt[3 * j + 1] = N;
      w.cnt[3 * j + 2] = 0;
    }
    return;
  }
  const double beta = e[m - 1];
  const double sgn = beta >= 0.0 ? 1.0 : -1.0, rho = 2.0 * fabs(beta);
  double mx_d = 0.0, mx_z = 0.0;
  for (int i = tid; i < N; i += 256) {
    const double d = Din[a + i];
    const double z = (i < n1 ? Q[(m - 1) + (int64_t)n * (a + i)] : sgn * Q[m + (int64_t)n * (a + i)]) *
                     0.70710678118654752440;
    sd[i] = d;
    sz[i] = z;
    mx_d = fmax(mx_d, fabs(d));
    mx_z = fmax(mx_z, fabs(z));
  }
  red[0][tid] = mx_d;
  red[1][tid] = mx_z;
  __syncthreads();
  // merged ascending order of the two sorted halves (ties: the first child first)
  int rk[DC_NMAX / 256];
#pragma unroll
  for (int u = 0; u < DC_NMAX / 256; ++u) {
    const int i = tid + 256 * u;
    rk[u] = -1;
    if (i < N) {
      const double v = sd[i];
      int lo = 0, hi;
      if (i < n1) {
        hi = n2;   // first index in the second half with value >= v
        while (lo < hi) {
          const int md = (lo + hi) >> 1;
          if (sd[n1 + md] < v) lo = md + 1;
          else hi = md;
        }
        rk[u] = i + lo;
      } else {
        hi = n1;   // first index in the first half with value > v
        while (lo < hi) {
          const int md = (lo + hi) >> 1;
          if (sd[md] <= v) lo = md + 1;
          else hi = md;
        }
        rk[u] = (i - n1) + lo;
      }
    }
  }
  for (int s = 128; s > 0; s >>= 1) {
    if (tid < s) {
      red[0][tid] = fmax(red[0][tid], red[0][tid + s]);
      red[1][tid] = fmax(red[1][tid], red[1][tid + s]);
    }
    __syncthreads();
  }
  const double tol = 8.0 * DC_EPS * fmax(red[0][0], red[1][0]);
  // sorted copies in place of the originals
  double vd[DC_NMAX / 256], vz[DC_NMAX / 256];
#pragma unroll
  for (int u = 0; u < DC_NMAX / 256; ++u) {
    const int i = tid + 256 * u;
    vd[u] = i < N ? sd[i] : 0.0;
    vz[u] = i < N ? sz[i] : 0.0;
  }
  __syncthreads();
#pragma unroll
  for (int u = 0; u < DC_NMAX / 256; ++u) {
    const int i = tid + 256 * u;
    if (i < N) {
      sd[rk[u]] = vd[u];
      sz[rk[u]] = vz[u];
      srt[rk[u]] = i;
    }
  }
  __syncthreads();
  if (tid != 0) return;
  // the serial scan; the next entry's loads are issued one iteration ahead, and the test for
  // a close pair, |(d_j - d_p) c s| <= tol with c s = -z_j z_p / (z_j^2 + z_p^2), is made
  // without a square root or division (those are needed only for a rotation)
  int K = 0, nf = 0, nr = 0, pj = -1;
  double pd = 0.0, pz = 0.0;
  int jn = srt[0];
  double dn = sd[0], zn = sz[0];
  for (int k = 0; k < N; ++k) {
    const int jj = jn;
    const double dj = dn, zj = zn;
    if (k + 1 < N) {
      jn = srt[k + 1];
      dn = sd[k + 1];
      zn = sz[k + 1];
    }
    if (rho * fabs(zj) <= tol) {
      w.dfd[a + nf] = dj;
      w.dfc[a + nf] = jj;
      ++nf;
      continue;
    }
    if (pj < 0) {
      pj = jj;
      pd = dj;
      pz = zj;
      continue;
    }
    const double z2 = zj * zj + pz * pz;
    const bool scaled_ok = z2 > 1e-280 && z2 < 1e280;   // else the exact form below
    bool close;
    if (scaled_ok) {
      close = fabs((dj - pd) * zj * pz) <= tol * z2;
    } else {
      const double t0 = dc_hyp(zj, pz);
      close = fabs((dj - pd) * (zj / t0) * (pz / t0)) <= tol;
    }
    if (close) {
      const double t = dc_hyp(zj, pz);
      const double cs = zj / t, sn = -pz / t;
      w.rp[a + nr] = pj;
      w.rq[a + nr] = jj;
      w.rc[a + nr] = cs;
      w.rs[a + nr] = sn;
      ++nr;
      w.dfd[a + nf] = pd * cs * cs + dj * sn * sn;
      w.dfc[a + nf] = pj;
      ++nf;
      pd = pd * sn * sn + dj * cs * cs;
      pz = t;
      pj = jj;
    } else {
      w.ndd[a + K] = pd;
      w.ndz[a + K] = pz;
      w.ndc[a + K] = pj;
      ++K;
      pj = jj;
      pd = dj;
      pz = zj;
    }
  }
  if (pj >= 0) {
    w.ndd[a + K] = pd;
    w.ndz[a + K] = pz;
    w.ndc[a + K] = pj;
    ++K;
  }
  w.cnt[3 * j] = K;
  w.cnt[3 * j + 1] = nf;
  w.cnt[3 * j + 2] = nr;
}

// ------------------------------------------------------------------ rotations
// Thread per row of the node: clears the row's entries outside its child's block (the two
// off-diagonal blocks of Q), then applies the deflation rotations in order.  A rotation's
// first column is the previous rotation's second one or a column not touched before, so the
// chain is carried in a register.
__global__ __launch_bounds__(256) void k_dc_rotate(int n, int depth, double* __restrict__ Q, DcWork w) {
  const int j = blockIdx.y;
  const int a = node_lo(n, depth, j), m = node_lo(n, depth + 1, 2 * j + 1), c = node_lo(n, depth, j + 1);
  if (!(a < m && m < c)) return;
  const int r = a + blockIdx.x * 256 + threadIdx.x;
  if (r >= c) return;
  double* row = Q + r;
  const int z0 = r < m ? m : a, z1 = r < m ? c : m;
  for (int col = z0; col < z1; ++col) row[(int64_t)n * col] = 0.0;
  const int nr = w.cnt[3 * j + 2];
  int cur = -1;
  double cv = 0.0;
  for (int t = 0; t < nr; ++t) {
    const int p = w.rp[a + t], q = w.rq[a + t];
    const double cs = w.rc[a + t], sn = w.rs[a + t];
    const double x = p == cur ? cv : row[(int64_t)n * (a + p)];
    const double y = row[(int64_t)n * (a + q)];
    row[(int64_t)n * (a + p)] = cs * x + sn * y;
    cv = cs * y - sn * x;
    cur = q;
    row[(int64_t)n * (a + q)] = cv;
  }
}

// ------------------------------------------------------------------ secular equation
// One wave per root; the node's non-deflated (d, z) in LDS.
__global__ __launch_bounds__(256) void k_dc_secular(int n, int depth, const double* __restrict__ e, DcWork w,
                                                    int* __restrict__ info, int report) {
  __shared__ double ld[DC_NMAX], lz[DC_NMAX];
  const int j = blockIdx.y;
  const int K = w.cnt[3 * j];
  if (blockIdx.x * 4 >= K) return;
  const int a = node_lo(n, depth, j), m = node_lo(n, depth + 1, 2 * j + 1);
  const double rho = 2.0 * fabs(e[m - 1]);
  for (int i = threadIdx.x; i < K; i += 256) {
    ld[i] = w.ndd[a + i];
    lz[i] = w.ndz[a + i];
  }
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const int jr = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (jr >= K) return;
  int o;
  double lo, hi;
  if (jr < K - 1) {
    const double mid = 0.5 * (ld[jr + 1] - ld[jr]);
    double s = 0.0;
    for (int i = lane; i < K; i += 64) s += lz[i] * lz[i] / ((ld[i] - ld[jr]) - mid);
    s = wave_sum(s);
    if (1.0 + rho * s >= 0.0) {
      o = jr;
      lo = 0.0;
      hi = mid;
    } else {
      o = jr + 1;
      lo = -mid;
      hi = 0.0;
    }
  } else {
    double s = 0.0;
    for (int i = lane; i < K; i += 64) s += lz[i] * lz[i];
    o = jr;
    lo = 0.0;
    hi = rho * wave_sum(s);
  }
  const double dorg = ld[o], zo2 = lz[o] * lz[o];
  double t = 0.5 * (lo + hi);
  bool conv = false;
  for (int it = 0; it < DC_ITMAX; ++it) {
    double s1 = 0.0, s2 = 0.0, sa = 0.0;
    for (int i = lane; i < K; i += 64) {
      if (i == o) continue;
      const double dl = ld[i] - dorg;
      const double r = 1.0 / (dl - t);
      const double q = lz[i] * lz[i] * r;
      s1 += q;
      s2 += q * dl * r;
      sa += fabs(q);
    }
    s1 = wave_sum(s1);
    s2 = wave_sum(s2);
    sa = wave_sum(sa);
    const double F = t * (1.0 + rho * s1) - rho * zo2;
    // converged once F is within the rounding error of its own evaluation (the stopping rule
    // of LAPACK's dlaed4): close to a pole, 1 + rho s1 can cancel to a few digits, and no
    // step-size test would then be met
    if (fabs(F) <= 8.0 * DC_EPS * (fabs(t) * (1.0 + rho * sa) + rho * zo2)) {
      conv = true;
      break;
    }
    const double Fp = 1.0 + rho * s2;
    // the sign of f is the sign of F times that of tau
    if ((t > 0.0) == (F > 0.0)) hi = t;
    else lo = t;
    const double step = Fp != 0.0 ? F / Fp : 0.0;
    if (Fp != 0.0 && fabs(step) <= 4.0 * DC_EPS * fabs(t)) {
      t -= step;
      conv = true;
      break;
    }
    double tn = t - step;
    if (Fp == 0.0 || !(lo < tn && tn < hi)) {
      tn = 0.5 * (lo + hi);
      if (tn == lo || tn == hi) {
        conv = true;
        break;
      }
    }
    t = tn;
  }
  if (lane == 0) {
    w.org[a + jr] = o;
    w.tau[a + jr] = t;
    if (!conv) {
      info[0] = 2;
      if (report && atomicAdd(info + 1, 1) < 6)
        printf("dc2 secular: depth %d node %d K %d root %d origin %d lo %.17g hi %.17g tau %.17g d[o] %.17g "
               "d[j] %.17g d[j+1] %.17g z[o] %.17g rho %.17g\n",
               depth, j, K, jr, o, lo, hi, t, dorg, ld[jr], jr + 1 < K ? ld[jr + 1] : 0.0, lz[o], rho);
    }
  }
}

// ------------------------------------------------------------------ zhat
// One wave per entry i: zhat_i^2 = (lambda_i - d_i) / rho * prod_{j != i} (lambda_j - d_i) / (d_j - d_i).
__global__ __launch_bounds__(256) void k_dc_zhat(int n, int depth, const double* __restrict__ e, DcWork w) {
  __shared__ double ld[DC_NMAX], lt[DC_NMAX];
  __shared__ int lo_[DC_NMAX];
  const int j = blockIdx.y;
  const int K = w.cnt[3 * j];
  if (blockIdx.x * 4 >= K) return;
  const int a = node_lo(n, depth, j), m = node_lo(n, depth + 1, 2 * j + 1);
  const double rho = 2.0 * fabs(e[m - 1]);
  for (int i = threadIdx.x; i < K; i += 256) {
    ld[i] = w.ndd[a + i];
    lt[i] = w.tau[a + i];
    lo_[i] = w.org[a + i];
  }
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const int i = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (i >= K) return;
  const double di = ld[i];
  double p = 1.0;
  for (int q = lane; q < K; q += 64) {
    if (q == i) continue;
    p *= ((ld[lo_[q]] - di) + lt[q]) / (ld[q] - di);
  }
  p = wave_prod(p) * (((ld[lo_[i]] - di) + lt[i]) / rho);
  if (lane == 0) w.zh[a + i] = copysign(sqrt(fmax(p, 0.0)), w.ndz[a + i]);
}

// ------------------------------------------------------------------ order
// Thread per entry of the node: its rank among the node's new eigenvalues (ties by entry
// index), the eigenvalue stored at that rank.
__global__ __launch_bounds__(256) void k_dc_order(int n, int depth, double* __restrict__ Dout, DcWork w) {
  __shared__ double v[DC_NMAX];
  const int j = blockIdx.y;
  const int a = node_lo(n, depth, j), c = node_lo(n, depth, j + 1), N = c - a;
  if (blockIdx.x * 256 >= N) return;
  const int K = w.cnt[3 * j];
  for (int q = threadIdx.x; q < N; q += 256)
    v[q] = q < K ? w.ndd[a + w.org[a + q]] + w.tau[a + q] : w.dfd[a + q - K];
  __syncthreads();
  const int q = blockIdx.x * 256 + threadIdx.x;
  if (q >= N) return;
  const double x = v[q];
  int r = 0;
  for (int f = 0; f < N; ++f) {
    const double y = v[f];
    r += (y < x) || (y == x && f < q);
  }
  Dout[a + r] = x;
  w.pos[a + q] = r;
}

// ------------------------------------------------------------------ eigenvectors of the update
// One wave per root j: column j of V (K x K, at the node's diagonal block of W).
__global__ __launch_bounds__(256) void k_dc_vec(int n, int depth, double* __restrict__ W, DcWork w) {
  __shared__ double ld[DC_NMAX], lzh[DC_NMAX];
  const int j = blockIdx.y;
  const int K = w.cnt[3 * j];
  if (blockIdx.x * 4 >= K) return;
  const int a = node_lo(n, depth, j);
  for (int i = threadIdx.x; i < K; i += 256) {
    ld[i] = w.ndd[a + i];
    lzh[i] = w.zh[a + i];
  }
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const int jr = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (jr >= K) return;
  const double dorg = ld[w.org[a + jr]], t = w.tau[a + jr];
  double s = 0.0;
  for (int i = lane; i < K; i += 64) {
    const double v = lzh[i] / ((ld[i] - dorg) - t);
    s += v * v;
  }
  const double inv = 1.0 / sqrt(wave_sum(s));
  double* col = W + a + (int64_t)n * (a + jr);
  for (int i = lane; i < K; i += 64) col[i] = lzh[i] / ((ld[i] - dorg) - t) * inv;
}

// ------------------------------------------------------------------ new eigenvectors
// Zout[a + r, a + pos[j]] = sum_i Q[a + r, a + ndc[i]] V[i, j]: 64 x 64 output tiles, four
// waves of 2 x 2 16x16 FP64 MFMA tiles (v_mfma_f64_16x16x4f64), K in steps of 16 through LDS.
constexpr int DG_LDS = 80;
__global__ __launch_bounds__(256) void k_dc_gemm(int n, int depth, const double* __restrict__ Q,
                                                 const double* __restrict__ W, double* __restrict__ Zout, DcWork w) {
  __shared__ double As[16][DG_LDS], Bs[16][DG_LDS];
  __shared__ int cols[16];
  const int j = blockIdx.z;
  const int a = node_lo(n, depth, j), c = node_lo(n, depth, j + 1), N = c - a;
  const int K = w.cnt[3 * j];
  const int m0 = blockIdx.x * 64, n0 = blockIdx.y * 64;
  if (m0 >= N || n0 >= K) return;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = (wave & 1) * 32, wn = (wave >> 1) * 32;
  d4 acc[2][2];
#pragma unroll
  for (int x = 0; x < 2; ++x)
#pragma unroll
    for (int y = 0; y < 2; ++y) acc[x][y] = d4{0.0, 0.0, 0.0, 0.0};
  const double* Wn = W + a + (int64_t)n * a;
  for (int k0 = 0; k0 < K; k0 += 16) {
    if (tid < 16) cols[tid] = k0 + tid < K ? w.ndc[a + k0 + tid] : 0;
    __syncthreads();
#pragma unroll
    for (int e4 = tid; e4 < 64 * 16; e4 += 256) {
      const int mm = e4 & 63, kk = e4 >> 6;
      const int r = m0 + mm;
      As[kk][mm] = (r < N && k0 + kk < K) ? Q[(a + r) + (int64_t)n * (a + cols[kk])] : 0.0;
    }
#pragma unroll
    for (int e4 = tid; e4 < 64 * 16; e4 += 256) {
      const int kk = e4 & 15, nn = e4 >> 4;
      Bs[kk][nn] = (k0 + kk < K && n0 + nn < K) ? Wn[(k0 + kk) + (int64_t)n * (n0 + nn)] : 0.0;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < 16; kk += 4) {
      const int kl = kk + (lane >> 4);
      double av[2], bv[2];
#pragma unroll
      for (int x = 0; x < 2; ++x) {
        av[x] = As[kl][wm + 16 * x + (lane & 15)];
        bv[x] = Bs[kl][wn + 16 * x + (lane & 15)];
      }
#pragma unroll
      for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y) acc[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[x], bv[y], acc[x][y], 0, 0, 0);
    }
    __syncthreads();
  }
  // element r of lane L of a 16x16 tile holds row (L >> 4) + 4 r, column L & 15
#pragma unroll
  for (int x = 0; x < 2; ++x)
#pragma unroll
    for (int y = 0; y < 2; ++y) {
      const int col = n0 + wn + 16 * y + (lane & 15);
      if (col >= K) continue;
      double* out = Zout + a + (int64_t)n * (a + w.pos[a + col]);
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = m0 + wm + 16 * x + (lane >> 4) + 4 * r;
        if (row < N) out[row] = acc[x][y][r];
      }
    }
}

// deflated columns of the rotated Q to their places: one workgroup per deflated entry
__global__ __launch_bounds__(256) void k_dc_copydef(int n, int depth, const double* __restrict__ Q,
                                                    double* __restrict__ Zout, DcWork w) {
  const int j = blockIdx.y, f = blockIdx.x;
  if (f >= w.cnt[3 * j + 1]) return;
  const int a = node_lo(n, depth, j), c = node_lo(n, depth, j + 1), N = c - a;
  const int K = w.cnt[3 * j];
  const double* src = Q + a + (int64_t)n * (a + w.dfc[a + f]);
  double* dst = Zout + a + (int64_t)n * (a + w.pos[a + K + f]);
  for (int r = threadIdx.x; r < N; r += 256) dst[r] = src[r];
}

int leaf_depth(int n) {
  int d = 0;
  while (((n + (1 << d) - 1) >> d) > DC_LEAF) ++d;
  return d;
}

}  // namespace

int64_t tri_dc_work_bytes(int n) {
  const int64_t nn = (int64_t)n * n;
  return 16 * nn + 8 * (int64_t)n + 7 * 8 * (int64_t)n + 6 * 4 * (int64_t)n + 3 * 4 * DC_NMAX / 2 + 32 * 256;
}

void tri_dc(hipStream_t s, int n, double* w, const double* e, double* Z, void* work, int* info_dev, bool report) {
  DC2_CHECK(n >= 1 && n <= DC_NMAX, 1, "tri_dc: n must be in 1..2048");
  const int64_t nn = (int64_t)n * n;
  char* p = reinterpret_cast<char*>(work);
  auto take = [&](int64_t bytes) {
    char* q = p;
    p += (bytes + 255) / 256 * 256;
    return q;
  };
  double* Zb = reinterpret_cast<double*>(take(8 * nn));
  double* W = reinterpret_cast<double*>(take(8 * nn));
  double* Db = reinterpret_cast<double*>(take(8 * n));
  DcWork dw;
  dw.ndd = reinterpret_cast<double*>(take(8 * n));
  dw.ndz = reinterpret_cast<double*>(take(8 * n));
  dw.dfd = reinterpret_cast<double*>(take(8 * n));
  dw.rc = reinterpret_cast<double*>(take(8 * n));
  dw.rs = reinterpret_cast<double*>(take(8 * n));
  dw.tau = reinterpret_cast<double*>(take(8 * n));
  dw.zh = reinterpret_cast<double*>(take(8 * n));
  dw.ndc = reinterpret_cast<int*>(take(4 * n));
  dw.dfc = reinterpret_cast<int*>(take(4 * n));
  dw.rp = reinterpret_cast<int*>(take(4 * n));
  dw.rq = reinterpret_cast<int*>(take(4 * n));
  dw.org = reinterpret_cast<int*>(take(4 * n));
  dw.pos = reinterpret_cast<int*>(take(4 * n));
  dw.cnt = reinterpret_cast<int*>(take(4 * 3 * DC_NMAX / 2));
  DC2_HIP(hipMemsetAsync(info_dev, 0, 2 * sizeof(int), s));
  const int Ld = leaf_depth(n);
  double* Dbuf[2] = {w, Db};
  double* Zbuf[2] = {Z, Zb};
  int q = Ld & 1;   // the last merge writes w and Z
  hipLaunchKernelGGL(k_dc_leaf, dim3(1 << Ld), dim3(64), 0, s, n, Ld, w, e, Dbuf[q], Zbuf[q], info_dev);
  for (int depth = Ld - 1; depth >= 0; --depth) {
    const int nodes = 1 << depth;
    const int nmax = (n + nodes - 1) / nodes;
    double* Din = Dbuf[q];
    double* Q = Zbuf[q];
    double* Dout = Dbuf[q ^ 1];
    double* Zout = Zbuf[q ^ 1];
    hipLaunchKernelGGL(k_dc_deflate, dim3(nodes), dim3(256), 0, s, n, depth, e, Din, Q, dw);
    hipLaunchKernelGGL(k_dc_rotate, dim3(cdiv(nmax, 256), nodes), dim3(256), 0, s, n, depth, Q, dw);
    hipLaunchKernelGGL(k_dc_secular, dim3(cdiv(nmax, 4), nodes), dim3(256), 0, s, n, depth, e, dw, info_dev,
                       (int)report);
    hipLaunchKernelGGL(k_dc_zhat, dim3(cdiv(nmax, 4), nodes), dim3(256), 0, s, n, depth, e, dw);
    hipLaunchKernelGGL(k_dc_order, dim3(cdiv(nmax, 256), nodes), dim3(256), 0, s, n, depth, Dout, dw);
    hipLaunchKernelGGL(k_dc_vec, dim3(cdiv(nmax, 4), nodes), dim3(256), 0, s, n, depth, W, dw);
    hipLaunchKernelGGL(k_dc_gemm, dim3(cdiv(nmax, 64), cdiv(nmax, 64), nodes), dim3(256), 0, s, n, depth, Q, W,
                       Zout, dw);
    hipLaunchKernelGGL(k_dc_copydef, dim3(nmax, nodes), dim3(256), 0, s, n, depth, Q, Zout, dw);
    q ^= 1;
  }
  DC2_HIP(hipGetLastError());
}

}  // namespace dc2
