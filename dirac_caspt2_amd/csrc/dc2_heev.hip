// Hermitian eigensolver for 3 <= n <= 2048 (the matrix its upper triangle defines, as
// rocSOLVER's fill_upper): Householder tridiagonalisation in one persistent
// kernel, divide and conquer on the real tridiagonal (this library's, dc2_tridc.hip, or
// rocSOLVER dstedc), back-transformation with one wave per eigenvector.
//
// Tridiagonalisation (k_hetrd).  Step k (k = 0 .. n-3) applies the reflector
// H_k = I - tau_k v_k v_k^H, v_k supported on rows k+1 .. n-1, to the trailing block
// B = A[k+1:, k+1:]:  p = tau_k B v_k,  K = (tau_k / 2) Re(v_k^H p),  q = p - K v_k,
// B <- B - v_k q^H - q v_k^H.  The reflector of column k maps x = A[k+1:, k] to
// alpha_k e_1 with alpha_k = -e^{i arg x_0} |x| (tau_k = 2 / v^H v is real); alpha_k is the
// complex subdiagonal.  A = Q T' Q^H with T' Hermitian tridiagonal, Q = H_0 H_1 ... H_{n-3};
// D = diag(delta), delta_0 = 1, delta_{k+1} = delta_k alpha_k / |alpha_k|, makes
// T = D^H T' D real symmetric with subdiagonal |alpha_k|, so the eigenvectors are
// X = Q D Z for the eigenvectors Z of T.
//
// Layout: workgroup w of nwg owns the columns j = w + nwg * l (l = wave + 8 c, c < CPW), held
// in registers for the whole reduction; lane L of a wave holds rows L + 64 s.  Per step:
//   (a) every workgroup forms p_j for its columns (a wave reduction per column) and a partial
//       of Re(v^H p); the owner of column k+1 publishes that column as it stands;
//   -- one grid barrier --
//   (b) every workgroup reads p, the partials and column k+1, updates its columns, applies
//       the same update to its copy of column k+1 and derives reflector k+1 from it.
// Every workgroup computes reflector k+1 from identical data with identical code, so all
// hold the same v; workgroup 0 writes d, alpha, tau, delta and the reflectors.  Exchanged
// data (p, partials, column k+1) are double-buffered by the parity of k and read with
// device-coherent loads after the barrier.  All workgroups must be resident together (see
// CoopSlot); each barrier wait has a spin limit after which the kernel raises a flag and
// every workgroup leaves, and the host then uses rocSOLVER instead.
#include "dc2_kernels.hpp"

#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <condition_variable>
#include <map>
#include <mutex>

namespace dc2 {
namespace {

constexpr int HT_WAVES = 8;
constexpr int HT_THREADS = 64 * HT_WAVES;
constexpr int HT_NMAX = 2048;
// columns per wave: 1 up to n = 512 (twice the workgroups of 2, half the work of a wave per
// column step: 3-9% less time at n = 100-512), 2 up to n = 1024 (4 spills registers at
// n > 768), 1 above, where the 2048 / 8 = 256 workgroups of the largest matrix take every
// compute unit.  narrow: 2 up to n = 1280, half the workgroups (80 instead of 160 at
// n = 1274) at 19% more time per reduction, for reductions that run beside other work (the
// subspace lanes of a CASPT2 energy: 88.0 against 90.4 ms per 324-spinor energy, whose
// transforms get the freed units)
inline int hetrd_cpw(int n, bool narrow = false) {
  if (narrow) return n <= 1280 ? 2 : 1;
  return (n > 512 && n <= 1024) ? 2 : 1;
}
inline int hetrd_nwg(int n, bool narrow = false) {
  return (n + HT_WAVES * hetrd_cpw(n, narrow) - 1) / (HT_WAVES * hetrd_cpw(n, narrow));
}

// barrier counters: the top one at bar[0], group g's at bar[32 (g + 1)] (lines of their own)
constexpr int HT_GROUPS = 16;
constexpr int HT_BAR_WORDS = 32 * (HT_GROUPS + 1);

struct HetrdWork {
  cplx* P;          // [2][n] p of the current step
  cplx* NC;         // [2][n] column k+1 as it stands before step k's update
  double* S;        // [2][nwg] partial sums of Re(v^H p)
  unsigned* bar;    // HT_BAR_WORDS barrier counters, zero on entry
  int* err;         // 1: a barrier wait hit its spin limit
};

// exchanged data are written and read with device-scope (coherent) accesses, which do not
// go through the non-coherent part of the per-XCD L2
__device__ inline double ld_coherent(const double* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline cplx ld_coherent(const cplx* p) {
  const double* q = reinterpret_cast<const double*>(p);
  return mk(ld_coherent(q), ld_coherent(q + 1));
}
__device__ inline void st_coherent(double* p, double v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline void st_coherent(cplx* p, cplx v) {
  double* q = reinterpret_cast<double*>(p);
  st_coherent(q, v.x);
  st_coherent(q + 1, v.y);
}

// Grid barrier of step k over all workgroups (all resident); false if the wait was
// abandoned.  Everything exchanged between workgroups during the reduction is written and read
// with device-coherent accesses (write-through stores, loads that bypass the per-XCD L2), so
// the barrier orders completions only and needs no L2 writeback or invalidation: every wave
// waits for its own stores to complete (s_waitcnt 0) before the workgroup barrier, thread 0
// then stores the workgroup's partial (the sum of red[0..7]) to part_dst when that is not
// null, waits for it, and arrives: above 64 workgroups on the counter of its group
// (workgroups wg = g mod HT_GROUPS), the group's last arrival on the top counter, so that no
// counter takes more than max(16, nwg / 16) increments per step; up to 64 on the top counter
// directly (one round trip fewer).  The waiting loop polls the top counter with
// relaxed loads, after which the exchanged data are read with coherent loads.  While thread
// 0 waits, workgroup 0 stores reflector k (held in sv) to Vh, off the path to its arrival.
__device__ bool grid_sync(unsigned* bar, int k, int* err, double* part_dst, const double* red, int n,
                          const cplx* sv, cplx* Vh) {
  __shared__ int ok;
  const int nwg = gridDim.x, g = blockIdx.x % HT_GROUPS;
  const unsigned step = (unsigned)(k + 1);
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  if (threadIdx.x == 0) {
    ok = 1;
    if (part_dst) {
      double ps = 0.0;
#pragma unroll
      for (int ww = 0; ww < HT_WAVES; ++ww) ps += red[ww];
      st_coherent(part_dst, ps);
      __builtin_amdgcn_s_waitcnt(0);
    }
    if (nwg > 64) {
      const unsigned gsize = (unsigned)((nwg - g + HT_GROUPS - 1) / HT_GROUPS);
      const unsigned old =
          __hip_atomic_fetch_add(bar + 32 * (g + 1), 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if (old + 1 == step * gsize) __hip_atomic_fetch_add(bar, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    } else {
      __hip_atomic_fetch_add(bar, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
  }
  if (blockIdx.x == 0)
    for (int i = threadIdx.x; i < n; i += HT_THREADS) Vh[i + (int64_t)n * k] = sv[i];
  if (threadIdx.x == 0) {
    const unsigned target = step * (unsigned)(nwg > 64 ? HT_GROUPS : nwg);
    long spins = 0;
    while (__hip_atomic_load(bar, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
      __builtin_amdgcn_s_sleep(1);
      if (__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0 || ++spins > (1l << 22)) {
        __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        ok = 0;
        break;
      }
    }
  }
  __syncthreads();
  return ok != 0;
}

__device__ inline cplx wave_sum(cplx x) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    x.x += __shfl_xor(x.x, o, 64);
    x.y += __shfl_xor(x.y, o, 64);
  }
  return x;
}
__device__ inline double wave_sum(double x) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
  return x;
}

// sum over the workgroup of one double per thread (same order in every workgroup), with one
// workgroup barrier, which also publishes the LDS stores made before the call.  red must not
// be read by another call between this one and the next workgroup barrier: the two sums of
// a step use red slots of their own.
__device__ double block_sum(double x, double* red) {
  x = wave_sum(x);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) red[wave] = x;
  __syncthreads();
  double s = 0.0;
#pragma unroll
  for (int w = 0; w < HT_WAVES; ++w) s += red[w];
  return s;
}

// Reflector of column k from sc (the column as updated, rows k .. n-1) and x2, this thread's
// part of |sc[k+1 ..]|^2 over rows it wrote itself: d_k, alpha_k, tau_k, v_k into sv (zero at
// rows <= k).  Workgroup 0 writes d, alpha, tau and delta here, v_k at the next barrier.
__device__ void reflector(int n, int k, double x2, const cplx* sc, cplx* sv, double* red, double* s_tau,
                          cplx* s_delta, double* d, cplx* sub, double* tau, cplx* delta, cplx* Vh) {
  const double xn2 = block_sum(x2, red);   // its barrier publishes sc
  const cplx x0 = sc[k + 1];
  const double ax0 = sqrt(cabs2(x0));
  const double tail = xn2 - ax0 * ax0;
  double t = 0.0;
  cplx alpha = x0;
  cplx ph = mk(1.0, 0.0);
  if (n - k - 1 > 1 && tail > 0.0) {
    const double xn = sqrt(xn2);
    if (ax0 > 0.0) ph = mk(x0.x / ax0, x0.y / ax0);
    alpha = mk(-xn * ph.x, -xn * ph.y);
    t = 2.0 / (2.0 * xn2 + 2.0 * ax0 * xn);
  }
  for (int i = threadIdx.x; i < n; i += HT_THREADS) {
    cplx v = mk(0.0, 0.0);
    if (t != 0.0 && i > k) v = (i == k + 1) ? csub(sc[i], alpha) : sc[i];
    sv[i] = v;   // workgroup 0 stores it to Vh at the next grid barrier (after the loop for the last)
  }
  if (threadIdx.x == 0) {
    *s_tau = t;
    if (blockIdx.x == 0) {
      d[k] = sc[k].x;
      sub[k] = alpha;
      tau[k] = t;
      const double aa = sqrt(cabs2(alpha));
      const cplx dk = *s_delta;
      const cplx dn = aa > 0.0 ? cmul(dk, mk(alpha.x / aa, alpha.y / aa)) : dk;
      delta[k + 1] = dn;
      *s_delta = dn;
    }
  }
  __syncthreads();
}

template <int S, int CPW>
__global__ __launch_bounds__(HT_THREADS) void k_hetrd(int n, const cplx* __restrict__ A, int64_t lda, HetrdWork w,
                                                      double* __restrict__ d, cplx* __restrict__ sub,
                                                      double* __restrict__ tau, cplx* __restrict__ delta,
                                                      cplx* __restrict__ Vh) {
  __shared__ cplx sv[64 * S], sq[64 * S], sc[64 * S];
  __shared__ double red[HT_WAVES], redn[HT_WAVES];
  __shared__ double s_tau;
  __shared__ cplx s_delta;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int nwg = gridDim.x, wg = blockIdx.x;
  int col[CPW];
  cplx a[CPW][S];
#pragma unroll
  for (int c = 0; c < CPW; ++c) {
    col[c] = wg + nwg * (wave + HT_WAVES * c);
#pragma unroll
    for (int s = 0; s < S; ++s) {
      const int i = lane + 64 * s;
      a[c][s] = (col[c] < n && i < n) ? A[i + lda * col[c]] : mk(0.0, 0.0);
    }
  }
  if (threadIdx.x == 0) {
    s_delta = mk(1.0, 0.0);
    if (wg == 0) delta[0] = s_delta;
  }
  double x2 = 0.0;
  for (int i = threadIdx.x; i < n; i += HT_THREADS) {   // column 0
    const cplx x = A[i];
    sc[i] = x;
    if (i > 0) x2 += cabs2(x);
  }
  // the loads of the owned columns complete here, once, rather than being waited for again
  // before every store that might use them
  __builtin_amdgcn_s_waitcnt(0);
  reflector(n, 0, x2, sc, sv, redn, &s_tau, &s_delta, d, sub, tau, delta, Vh);
  for (int k = 0; k <= n - 3; ++k) {
    const int par = k & 1;
    const double t = s_tau;
    x2 = 0.0;
    cplx* P = w.P + (int64_t)par * n;
    cplx* NC = w.NC + (int64_t)par * n;
    double* Sp = w.S + (int64_t)par * nwg;
    // (a) p_j = t sum_i conj(A_ij) v_i for the owned columns j >= k+1
    double part = 0.0;
    if (t != 0.0) {
#pragma unroll
      for (int c = 0; c < CPW; ++c) {
        const int j = col[c];
        if (j >= n || j <= k) continue;   // uniform over the wave
        cplx acc = mk(0.0, 0.0);
#pragma unroll
        for (int s = 0; s < S; ++s) {
          const int i = lane + 64 * s;
          if (i > k && i < n) cfma(acc, cconj(a[c][s]), sv[i]);
        }
        acc = wave_sum(acc);
        const cplx pj = mk(t * acc.x, t * acc.y);
        if (lane == 0) st_coherent(P + j, pj);
        const cplx vj = sv[j];
        part += vj.x * pj.x + vj.y * pj.y;   // Re(conj(v_j) p_j)
      }
      if (lane == 0) red[wave] = part;
    }
    // the owner publishes column k+1
#pragma unroll
    for (int c = 0; c < CPW; ++c)
      if (col[c] == k + 1) {
#pragma unroll
        for (int s = 0; s < S; ++s) {
          const int i = lane + 64 * s;
          if (i > k && i < n) st_coherent(NC + i, a[c][s]);
        }
      }
    if (!grid_sync(w.bar, k, w.err, t != 0.0 ? Sp + wg : nullptr, red, n, sv, Vh)) return;
    // (b) one round of coherent loads: column k+1, p and the partials (nwg <= 256, one per
    // thread, summed in the same order in every workgroup); q = p - K v is formed where used.
    // Every load of the round is issued before the first is used (NL per thread), so their
    // latencies overlap instead of adding up over the rows
    // (in batches of two per thread for the largest tiles, whose columns fill the registers)
    constexpr int NL = (64 * S + HT_THREADS - 1) / HT_THREADS;
    constexpr int NB = S >= 30 ? 2 : NL;
    double kp = 0.0;
    if (t != 0.0 && threadIdx.x < nwg) kp = ld_coherent(Sp + threadIdx.x);
#pragma unroll
    for (int u0 = 0; u0 < NL; u0 += NB) {
      cplx lc[NB], lq[NB];
#pragma unroll
      for (int u = 0; u < NB; ++u) {
        const int i = k + 1 + threadIdx.x + (u0 + u) * HT_THREADS;
        lc[u] = mk(0.0, 0.0);
        lq[u] = mk(0.0, 0.0);
        if (u0 + u < NL && i < n) {
          lc[u] = ld_coherent(NC + i);
          if (t != 0.0) lq[u] = ld_coherent(P + i);
        }
      }
#pragma unroll
      for (int u = 0; u < NB; ++u) {
        const int i = k + 1 + threadIdx.x + (u0 + u) * HT_THREADS;
        if (u0 + u < NL && i < n) {
          sc[i] = lc[u];
          if (t != 0.0) sq[i] = lq[u];
        }
      }
    }
    if (t != 0.0) {
      const double K = 0.5 * t * block_sum(kp, red);   // its barrier also publishes sc and sq
      auto qat = [&](int i) {
        const cplx p = sq[i], v = sv[i];
        return mk(p.x - K * v.x, p.y - K * v.y);
      };
#pragma unroll
      for (int c = 0; c < CPW; ++c) {
        const int j = col[c];
        if (j >= n || j <= k) continue;
        const cplx cqj = cconj(qat(j)), cvj = cconj(sv[j]);
#pragma unroll
        for (int s = 0; s < S; ++s) {
          const int i = lane + 64 * s;
          if (i > k && i < n) {
            cplx x = a[c][s];
            x = csub(x, cmul(sv[i], cqj));
            x = csub(x, cmul(qat(i), cvj));
            a[c][s] = x;
          }
        }
      }
      // column k+1 updated by the rows this thread loaded, and its part of the next norm
      const cplx cq1 = cconj(qat(k + 1)), cv1 = cconj(sv[k + 1]);
      for (int i = k + 1 + threadIdx.x; i < n; i += HT_THREADS) {
        cplx x = sc[i];
        x = csub(x, cmul(sv[i], cq1));
        x = csub(x, cmul(qat(i), cv1));
        sc[i] = x;
        if (i > k + 1) x2 += cabs2(x);
      }
    } else {
      for (int i = k + 1 + threadIdx.x; i < n; i += HT_THREADS)
        if (i > k + 1) x2 += cabs2(sc[i]);
    }
    // every thread has finished reading sv and sq of step k before the barrier in reflector,
    // after which v_{k+1} replaces v_k in sv
    reflector(n, k + 1, x2, sc, sv, redn, &s_tau, &s_delta, d, sub, tau, delta, Vh);
  }
  if (wg == 0)   // the last reflector, which no barrier follows
    for (int i = threadIdx.x; i < n; i += HT_THREADS) Vh[i + (int64_t)n * (n - 2)] = sv[i];
  // the last diagonal element, as updated by the last step
#pragma unroll
  for (int c = 0; c < CPW; ++c)
    if (col[c] == n - 1) {
#pragma unroll
      for (int s = 0; s < S; ++s)
        if (lane + 64 * s == n - 1) d[n - 1] = a[c][s].x;
    }
}

// the Hermitian matrix the upper triangle defines (what rocSOLVER's fill_upper reads):
// A_ij = conj(A_ji) below the diagonal, real diagonal
__global__ void k_herm_from_upper(int n, cplx* __restrict__ A, int64_t lda) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)n * n) return;
  const int i = (int)(t % n), j = (int)(t / n);
  if (i > j) A[i + lda * j] = cconj(A[j + lda * i]);
  else if (i == j) A[i + lda * i].y = 0.0;
}

// e_k = |alpha_k| (the real subdiagonal); Z <- identity if Z is given (dstedc reads it)
__global__ void k_tri_prep(int n, const cplx* __restrict__ sub, double* __restrict__ e, double* __restrict__ Z) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n - 1) e[t] = sqrt(cabs2(sub[t]));
  if (Z && t < (int64_t)n * n) Z[t] = (t % (n + 1) == 0) ? 1.0 : 0.0;
}

// X[:, c] = H_0 H_1 ... H_{n-3} (delta .* Z[:, c]); one wave per column, 8 columns per
// workgroup; the reflectors are staged through LDS in chunks of BT_CH (coalesced loads)
constexpr int BT_CH = 8;
template <int S>
__global__ __launch_bounds__(512) void k_hetrd_back(int n, const double* __restrict__ Z, const cplx* __restrict__ delta,
                                                   const double* __restrict__ tau, const cplx* __restrict__ Vh,
                                                   cplx* __restrict__ X, int64_t ldx) {
  __shared__ cplx vs[BT_CH][64 * S];
  __shared__ double ts[BT_CH];
  const int lane = threadIdx.x & 63;
  const int c = blockIdx.x * 8 + (threadIdx.x >> 6);
  const bool active = c < n;   // every wave takes part in the chunk loads
  cplx x[S];
#pragma unroll
  for (int s = 0; s < S; ++s) {
    const int i = lane + 64 * s;
    x[s] = (active && i < n) ? cscale(delta[i], Z[i + (int64_t)n * c]) : mk(0.0, 0.0);
  }
  for (int k0 = n - 3; k0 >= 0; k0 -= BT_CH) {
    const int nk = min(BT_CH, k0 + 1);   // reflectors k0, k0-1, ..., k0-nk+1
    __syncthreads();
    for (int idx = threadIdx.x; idx < nk * n; idx += 512) {
      const int kk = idx / n, i = idx - kk * n, k = k0 - kk;
      vs[kk][i] = i > k ? Vh[i + (int64_t)n * k] : mk(0.0, 0.0);
    }
    if (threadIdx.x < nk) ts[threadIdx.x] = tau[k0 - threadIdx.x];
    __syncthreads();
    if (!active) continue;
    for (int kk = 0; kk < nk; ++kk) {
      const double t = ts[kk];
      if (t == 0.0) continue;
      cplx dot = mk(0.0, 0.0);
#pragma unroll
      for (int s = 0; s < S; ++s) {
        const int i = lane + 64 * s;
        if (i < n) cfma(dot, cconj(vs[kk][i]), x[s]);
      }
      dot = wave_sum(dot);
      const cplx f = mk(t * dot.x, t * dot.y);
#pragma unroll
      for (int s = 0; s < S; ++s) {
        const int i = lane + 64 * s;
        if (i < n) x[s] = csub(x[s], cmul(vs[kk][i], f));
      }
    }
  }
  if (!active) return;
#pragma unroll
  for (int s = 0; s < S; ++s) {
    const int i = lane + 64 * s;
    if (i < n) X[i + ldx * c] = x[s];
  }
}

// Blocked back-transformation (n >= BT_BLOCKED_MIN): the reflectors k = a .. a+nb-1 of panel
// p (a = p nb) multiply as H_a ... H_{a+nb-1} = I - V T V^H (compact WY, T upper triangular:
// T_jj = tau_{a+j}, T[0:j, j] = -tau_{a+j} T[0:j, 0:j] V[:, 0:j]^H v_j), so X = Q D Z is
// D Z with the panels applied from the last to the first as X <- X - (V T) (V^H X), three
// matrix products per panel on the MFMA kernel.
constexpr int BT_NB = 64, BT_BLOCKED_MIN = 256;
constexpr int BT_SB = 4 * BT_NB;   // reflectors per application: four panels merged

// X = D Z (X complex, lda; Z real n x n)
__global__ void k_dz(int n, const double* __restrict__ Z, const cplx* __restrict__ delta, cplx* __restrict__ X,
                     int64_t ldx) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)n * n) return;
  const int i = (int)(t % n), c = (int)(t / n);
  X[i + ldx * c] = cscale(delta[i], Z[t]);
}

// T of every panel from its Gram matrix G = V^H V (nb x nb per panel, column-major) and tau;
// one workgroup of BT_NB threads per panel with G and T in LDS, thread i owns row i of T.
// Reflectors past nrefl (and those with tau = 0) give zero columns.
__global__ __launch_bounds__(BT_NB) void k_wy_t(int nrefl, const double* __restrict__ tau,
                                                const cplx* __restrict__ G, cplx* __restrict__ T) {
  __shared__ cplx gs[BT_NB][BT_NB + 1], ts[BT_NB][BT_NB + 1];   // [column][row]
  const int p = blockIdx.x, i = threadIdx.x, a = p * BT_NB;
  const cplx* g = G + (int64_t)BT_NB * BT_NB * p;
  for (int j = 0; j < BT_NB; ++j) gs[j][i] = g[i + BT_NB * j];
  __syncthreads();
  for (int j = 0; j < BT_NB; ++j) {
    const double tj = a + j < nrefl ? tau[a + j] : 0.0;
    // T[i][j] = -tau_j sum_{l=i}^{j-1} T[i][l] G[l][j] (i < j), tau_j (i = j), 0 (i > j);
    // column j reads only columns l < j, written in earlier steps
    cplx v = mk(0.0, 0.0);
    if (i < j && tj != 0.0) {
      cplx acc = mk(0.0, 0.0);
      for (int l = i; l < j; ++l) cfma(acc, ts[l][i], gs[j][l]);
      v = mk(-tj * acc.x, -tj * acc.y);
    } else if (i == j) {
      v = mk(tj, 0.0);
    }
    ts[j][i] = v;
    __syncthreads();
  }
  cplx* t = T + (int64_t)BT_NB * BT_NB * p;
  for (int j = 0; j < BT_NB; ++j) t[i + BT_NB * j] = ts[j][i];
}

// T of a merged panel: its four panels' T on the diagonal blocks (the rest zero)
__global__ void k_ts_diag(const cplx* __restrict__ T64, cplx* __restrict__ TS) {
  const int p = blockIdx.y, b = blockIdx.x >> 4;   // merged panel, its panel 0..3
  const int e = (blockIdx.x & 15) * 256 + threadIdx.x, i = e % BT_NB, j = e / BT_NB;
  cplx* ts = TS + (int64_t)BT_SB * BT_SB * p;
  const cplx* t = T64 + (int64_t)BT_NB * BT_NB * (4 * p + b);
  for (int bb = 0; bb < 4; ++bb)   // the whole column strip of block b: zeros off its diagonal block
    ts[(BT_NB * bb + i) + (int64_t)BT_SB * (BT_NB * b + j)] = bb == b ? t[i + BT_NB * j] : mk(0.0, 0.0);
}

// X = Q D Z with Q = H_0 H_1 ... H_{n-3}, four panels of 64 reflectors at a time: each
// panel's compact-WY T from its Gram matrix (k_wy_t), the merged T of four panels from
// (I - V1 T1 V1^H)(I - V2 T2 V2^H) = I - [V1 V2] [[T1, -T1 (V1^H V2) T2], [0, T2]] [V1 V2]^H
// (pairs, then the two pairs), then X <- X - (V T)(V^H X) per merged panel, last first: half
// as many products as panels of 64, each four times taller (the products of the 64-reflector
// panels ran at a few TF/s beside the transforms).  Work: G64, T64 (64^2 per panel), GS, TS,
// TM (256^2 per merged panel), VT (n x 256 per merged panel), W (256 x n).
void back_blocked(hipStream_t s, int n, const double* Z, const cplx* delta, const double* tau, const cplx* Vh,
                  cplx* X, int64_t ldx, cplx* G, cplx* T, cplx* GS, cplx* TS, cplx* TM, cplx* VT, cplx* W) {
  const int nrefl = n - 2, nps = (int)cdiv(nrefl, BT_SB), np = 4 * nps;
  const cplx one = mk(1.0, 0.0), zero = mk(0.0, 0.0), mone = mk(-1.0, 0.0);
  const int64_t nn = (int64_t)n * n, sq = (int64_t)BT_SB * BT_SB, ld = BT_SB;
  hipLaunchKernelGGL(k_dz, dim3(cdiv(nn, 256)), dim3(256), 0, s, n, Z, delta, X, ldx);
  // every panel's G = V^H V and T; every merged panel's Gram (the columns of Vh past the last
  // reflector are zero, so are their T columns)
  zgemm(s, 'C', 'N', BT_NB, BT_NB, n, one, Vh, n, (int64_t)n * BT_NB, Vh, n, (int64_t)n * BT_NB, zero, G, BT_NB,
        (int64_t)BT_NB * BT_NB, np);
  hipLaunchKernelGGL(k_wy_t, dim3(np), dim3(BT_NB), 0, s, nrefl, tau, G, T);
  zgemm(s, 'C', 'N', BT_SB, BT_SB, n, one, Vh, n, (int64_t)n * BT_SB, Vh, n, (int64_t)n * BT_SB, zero, GS, ld, sq,
        nps);
  hipLaunchKernelGGL(k_ts_diag, dim3(4 * 16, nps), dim3(256), 0, s, T, TS);
  // pairs (0,1) and (2,3): TS_ab = -T_a (G_ab T_b); then the pairs: TS_[01][23]
  for (int a = 0; a < 4; a += 2) {
    const int b = a + 1;
    const int64_t oab = BT_NB * a + ld * BT_NB * b, oaa = BT_NB * a + ld * BT_NB * a, obb = BT_NB * b + ld * BT_NB * b;
    zgemm(s, 'N', 'N', BT_NB, BT_NB, BT_NB, one, GS + oab, ld, sq, TS + obb, ld, sq, zero, TM, BT_NB,
          (int64_t)BT_NB * BT_NB, nps);
    zgemm(s, 'N', 'N', BT_NB, BT_NB, BT_NB, mone, TS + oaa, ld, sq, TM, BT_NB, (int64_t)BT_NB * BT_NB, zero, TS + oab,
          ld, sq, nps);
  }
  {
    const int h = 2 * BT_NB;
    const int64_t o12 = ld * h, o22 = h + ld * h;
    zgemm(s, 'N', 'N', h, h, h, one, GS + o12, ld, sq, TS + o22, ld, sq, zero, TM, h, (int64_t)h * h, nps);
    zgemm(s, 'N', 'N', h, h, h, mone, TS, ld, sq, TM, h, (int64_t)h * h, zero, TS + o12, ld, sq, nps);
  }
  zgemm(s, 'N', 'N', n, BT_SB, BT_SB, one, Vh, n, (int64_t)n * BT_SB, TS, ld, sq, zero, VT, n, (int64_t)n * BT_SB,
        nps);
  for (int p = nps - 1; p >= 0; --p) {
    const int a = p * BT_SB, nb = std::min(BT_SB, nrefl - a), rows = n - a - 1;
    // rows a+1 .. n-1 are the only ones V touches
    zgemm(s, 'C', 'N', nb, n, rows, one, Vh + (a + 1) + (int64_t)n * a, n, 0, X + (a + 1), ldx, 0, zero, W, nb, 0,
          1);
    zgemm(s, 'N', 'N', rows, n, nb, mone, VT + (a + 1) + (int64_t)n * a, n, 0, W, nb, 0, one, X + (a + 1), ldx, 0,
          1);
  }
}

// The workgroups of one k_hetrd wait for each other, so all of them must be resident at
// once.  Each one fits a compute unit on its own (one workgroup of 8 waves at <= 256 VGPRs
// and 49 KB LDS), and the reductions in flight in this process hold at most one workgroup
// per compute unit between them; other kernels sharing the device always finish, so every
// workgroup of every reduction in flight is eventually scheduled.
std::mutex g_coop_mu;
std::condition_variable g_coop_cv;
int g_coop_units = -1;   // workgroups still available (the device's compute units)

struct CoopSlot {
  int units;
  CoopSlot(int u, int cus) : units(u) {
    std::unique_lock<std::mutex> lk(g_coop_mu);
    if (g_coop_units < 0) g_coop_units = cus;
    g_coop_cv.wait(lk, [&] { return g_coop_units >= units; });
    g_coop_units -= units;
  }
  ~CoopSlot() {
    {
      std::lock_guard<std::mutex> lk(g_coop_mu);
      g_coop_units += units;
    }
    g_coop_cv.notify_all();
  }
};

template <int S, int CPW = 2>
hipError_t launch_hetrd(hipStream_t s, int nwg, int n, const cplx* A, int64_t lda, HetrdWork w, double* d, cplx* sub,
                        double* tau, cplx* delta, cplx* Vh) {
  int per_cu = 0;
  hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(
      &per_cu, reinterpret_cast<const void*>(&k_hetrd<S, CPW>), HT_THREADS, 0);
  if (e != hipSuccess || per_cu < 1) return e != hipSuccess ? e : hipErrorLaunchOutOfResources;
  hipLaunchKernelGGL((k_hetrd<S, CPW>), dim3(nwg), dim3(HT_THREADS), 0, s, n, A, lda, w, d, sub, tau, delta, Vh);
  return hipGetLastError();
}

template <int S>
void launch_back(hipStream_t s, int n, const double* Z, const cplx* delta, const double* tau, const cplx* Vh, cplx* X,
                 int64_t ldx) {
  hipLaunchKernelGGL((k_hetrd_back<S>), dim3(cdiv(n, 8)), dim3(512), 0, s, n, Z, delta, tau, Vh, X, ldx);
}

// device time of every tridiagonalisation, by matrix size (HIP events around the launch)
std::mutex g_hetrd_mu;
std::map<int, std::pair<int64_t, double>> g_hetrd_time;

}  // namespace

void hetrd_stats(int n, double* out) {
  std::lock_guard<std::mutex> lk(g_hetrd_mu);
  auto it = g_hetrd_time.find(n);
  out[0] = it == g_hetrd_time.end() ? 0.0 : (double)it->second.first;
  out[1] = it == g_hetrd_time.end() ? 0.0 : it->second.second;
}

void hetrd_stats_reset() {
  std::lock_guard<std::mutex> lk(g_hetrd_mu);
  g_hetrd_time.clear();
}

// the blocked back-transformation's G, T, GS, TS, TM, VT and W (each rounded to 256 bytes)
int64_t back_blocked_bytes(int n) {
  const int64_t nps = cdiv(n - 2, BT_SB), np = 4 * nps;
  return 16 * (2 * np * BT_NB * BT_NB + 3 * nps * BT_SB * BT_SB + nps * BT_SB * (int64_t)n + (int64_t)BT_SB * n) +
         7 * 256;
}

// columns of the reflector array: whole merged panels of the blocked back-transformation
int64_t vh_cols(int n) { return n >= BT_BLOCKED_MIN ? std::max<int64_t>(n, cdiv(n - 2, BT_SB) * BT_SB) : n; }

int64_t heev_native_work_bytes(int n) {
  const int64_t nn = (int64_t)n * n;
  const int nwg = hetrd_nwg(n);
  int64_t b = (int64_t)n * vh_cols(n) * 16 + nn * 8 + (int64_t)n * (16 * 6 + 8 * 3) + 16 * nwg +
              4 * (HT_BAR_WORDS + 32) + 10 * 256;
  // the tridiagonal divide and conquer's buffers, later the blocked back-transformation's
  b += std::max<int64_t>(tri_dc_work_bytes(n), n >= BT_BLOCKED_MIN ? back_blocked_bytes(n) : 0);
  return b;
}

bool heev_native(hipStream_t s, rocblas_handle h, int n, cplx* A, int64_t lda, double* w, void* work, int* info_dev,
                 int* flags_host, bool own_tridiag, bool report, bool narrow) {
  DC2_CHECK(n >= 3 && n <= HT_NMAX, 1, "heev_native: n must be in 3..2048");
  const int nwg = hetrd_nwg(n, narrow);   // at most hetrd_nwg(n), which sizes the work area
  const int64_t nn = (int64_t)n * n;
  char* p = reinterpret_cast<char*>(work);
  auto take = [&](int64_t bytes) {
    char* q = p;
    p += (bytes + 255) / 256 * 256;
    return q;
  };
  cplx* Vh = reinterpret_cast<cplx*>(take((int64_t)n * vh_cols(n) * 16));
  double* Z = reinterpret_cast<double*>(take(nn * 8));
  cplx* sub = reinterpret_cast<cplx*>(take(n * 16));
  cplx* delta = reinterpret_cast<cplx*>(take(n * 16));
  cplx* P = reinterpret_cast<cplx*>(take(2 * n * 16));
  cplx* NC = reinterpret_cast<cplx*>(take(2 * n * 16));
  double* tau = reinterpret_cast<double*>(take(n * 8));
  double* e = reinterpret_cast<double*>(take(n * 8));
  double* S = reinterpret_cast<double*>(take(2 * nwg * 8));
  unsigned* flags = reinterpret_cast<unsigned*>(take(4 * (HT_BAR_WORDS + 32)));
  HetrdWork hw{P, NC, S, flags, reinterpret_cast<int*>(flags + 1)};   // err beside the top counter
  DC2_HIP(hipMemsetAsync(flags, 0, 4 * (HT_BAR_WORDS + 32), s));
  // reflectors 0 .. n-2 are written; the columns after them take part in the blocked
  // back-transformation's batched products (whole panels), as zeros
  DC2_HIP(hipMemsetAsync(Vh + (int64_t)n * (n - 1), 0, (vh_cols(n) - n + 1) * n * sizeof(cplx), s));
  hipLaunchKernelGGL(k_herm_from_upper, dim3(cdiv(nn, 256)), dim3(256), 0, s, n, A, lda);
  DC2_HIP(hipGetLastError());
  int dev = 0, cus = 0;
  DC2_HIP(hipGetDevice(&dev));
  DC2_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  if (nwg > cus) return false;
  hipError_t st;
  // timing events of this thread, destroyed when it exits (lane threads live for one call)
  struct ThreadEvents {
    hipEvent_t e[2] = {nullptr, nullptr};
    ~ThreadEvents() {
      for (auto x : e)
        if (x) (void)hipEventDestroy(x);
    }
  };
  thread_local ThreadEvents tev;
  hipEvent_t* ev = tev.e;
  for (int q = 0; q < 2; ++q)
    if (!ev[q]) DC2_HIP(hipEventCreate(&ev[q]));
  {
    CoopSlot slot(nwg, cus);
    DC2_HIP(hipEventRecord(ev[0], s));
    // S = rows per lane (n <= 64 S)
    const bool two = hetrd_cpw(n, narrow) == 2;
    if (n <= 128)
      st = two ? launch_hetrd<2, 2>(s, nwg, n, A, lda, hw, w, sub, tau, delta, Vh)
               : launch_hetrd<2, 1>(s, nwg, n, A, lda, hw, w, sub, tau, delta, Vh);
    else if (n <= 256)
      st = two ? launch_hetrd<4, 2>(s, nwg, n, A, lda, hw, w, sub, tau, delta, Vh)
               : launch_hetrd<4, 1>(s, nwg, n, A, lda, hw, w, sub, tau, delta, Vh);
    else if (n <= 512)
      st = two ? launch_hetrd<8, 2>(s, nwg, n, A, lda, hw, w, sub, tau, delta, Vh)
               : launch_hetrd<8, 1>(s, nwg, n, A, lda, hw, w, sub, tau, delta, Vh);
    else if (n <= 768) st = launch_hetrd<12>(s, nwg, n, A, lda, hw, w, sub, tau, delta, Vh);
    else if (n <= 832) st = launch_hetrd<13>(s, nwg, n, A, lda, hw, w, sub, tau, delta, Vh);
    else if (n <= 1024) st = launch_hetrd<16>(s, nwg, n, A, lda, hw, w, sub, tau, delta, Vh);
    else if (n <= 1280)
      st = two ? launch_hetrd<20, 2>(s, nwg, n, A, lda, hw, w, sub, tau, delta, Vh)
               : launch_hetrd<20, 1>(s, nwg, n, A, lda, hw, w, sub, tau, delta, Vh);
    else if (n <= 1536) st = launch_hetrd<24, 1>(s, nwg, n, A, lda, hw, w, sub, tau, delta, Vh);
    else if (n <= 1792) st = launch_hetrd<28, 1>(s, nwg, n, A, lda, hw, w, sub, tau, delta, Vh);
    else st = launch_hetrd<32, 1>(s, nwg, n, A, lda, hw, w, sub, tau, delta, Vh);
    if (st != hipSuccess) {
      (void)hipGetLastError();
      return false;   // e.g. the workgroup does not fit a compute unit: the caller uses rocSOLVER
    }
    DC2_HIP(hipEventRecord(ev[1], s));
    DC2_HIP(hipMemcpyAsync(flags_host, flags, 2 * sizeof(int), hipMemcpyDeviceToHost, s));
    DC2_HIP(hipStreamSynchronize(s));
    if (flags_host[1] != 0) return false;
    float ms = 0.0f;
    DC2_HIP(hipEventElapsedTime(&ms, ev[0], ev[1]));
    std::lock_guard<std::mutex> lk(g_hetrd_mu);
    auto& e = g_hetrd_time[n];
    e.first += 1;
    e.second += 1e-3 * ms;
  }
  if (own_tridiag) {
    hipLaunchKernelGGL(k_tri_prep, dim3(cdiv(n, 256)), dim3(256), 0, s, n, sub, e, nullptr);
    DC2_HIP(hipGetLastError());
    tri_dc(s, n, w, e, Z, p, info_dev, report);   // its buffers end where the back-transformation's begin
  } else {
    hipLaunchKernelGGL(k_tri_prep, dim3(cdiv(nn, 256)), dim3(256), 0, s, n, sub, e, Z);
    DC2_HIP(hipGetLastError());
    const rocblas_status rs = rocsolver_dstedc(h, rocblas_evect_tridiagonal, n, w, e, Z, n, info_dev);
    if (rs != rocblas_status_success) return false;
  }
  if (n >= BT_BLOCKED_MIN) {
    const int64_t nps = cdiv(n - 2, BT_SB), np = 4 * nps;
    cplx* G = reinterpret_cast<cplx*>(take(16 * np * BT_NB * BT_NB));
    cplx* T = reinterpret_cast<cplx*>(take(16 * np * BT_NB * BT_NB));
    cplx* GS = reinterpret_cast<cplx*>(take(16 * nps * BT_SB * BT_SB));
    cplx* TS = reinterpret_cast<cplx*>(take(16 * nps * BT_SB * BT_SB));
    cplx* TM = reinterpret_cast<cplx*>(take(16 * nps * BT_SB * BT_SB));
    cplx* VT = reinterpret_cast<cplx*>(take(16 * nps * BT_SB * (int64_t)n));
    cplx* W = reinterpret_cast<cplx*>(take(16 * (int64_t)BT_SB * n));
    back_blocked(s, n, Z, delta, tau, Vh, A, lda, G, T, GS, TS, TM, VT, W);
    DC2_HIP(hipGetLastError());
    return true;
  }
  if (n <= 128) launch_back<2>(s, n, Z, delta, tau, Vh, A, lda);
  else if (n <= 256) launch_back<4>(s, n, Z, delta, tau, Vh, A, lda);
  else if (n <= 512) launch_back<8>(s, n, Z, delta, tau, Vh, A, lda);
  else if (n <= 768) launch_back<12>(s, n, Z, delta, tau, Vh, A, lda);
  else launch_back<16>(s, n, Z, delta, tau, Vh, A, lda);
  DC2_HIP(hipGetLastError());
  return true;
}

}  // namespace dc2
