// Complex FP64 GEMM on CDNA4 matrix cores (v_mfma_f64_16x16x4f64).
//
// One workgroup = 4 wave64s computing a 64x64 complex tile of C; every wave
// owns a 32x32 sub-tile = 2x2 MFMA tiles of 16x16.  A complex multiply-add is
// four real MFMAs on separate real/imaginary planes staged in LDS.
// The LDS row stride (80 doubles) puts the four k-rows read by one wave in
// disjoint bank halves.
//
// Used for: the four one-index integral transforms (module_intra.F90), the
// RDM contractions (density_matrix.F90), the V contractions and the
// (U L^-1/2)^H B (U L^-1/2) products of every solve_X_subspace.  Single products
// whose output has too few 64x64 tiles (the block Davidson of the CI matrix) go to
// plain FP64 kernels instead: one wave per output row over a long K, one thread per
// output over a short K, one thread per row for at most 8 columns over a long K.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdlib>
#include <map>
#include <mutex>
#include <tuple>

#include "dc2_common.hpp"

namespace dc2 {

typedef double d4 __attribute__((ext_vector_type(4)));

namespace {
constexpr int BM = 64, BN = 64, BK = 16, LDS_STRIDE = 80;

template <char OPA>
__device__ inline cplx loadA(const cplx* __restrict__ A, int64_t lda, int64_t m, int64_t k) {
  if (OPA == 'N') return A[m + lda * k];
  cplx v = A[k + lda * m];
  return OPA == 'C' ? cconj(v) : v;
}

template <char OPB>
__device__ inline cplx loadB(const cplx* __restrict__ B, int64_t ldb, int64_t k, int64_t n) {
  if (OPB == 'N') return B[k + ldb * n];
  if (OPB == 'R') return cconj(B[k + ldb * n]);
  cplx v = B[n + ldb * k];
  return OPB == 'C' ? cconj(v) : v;
}

// blockIdx.z = batch index * sk + K-split index; split kz covers K in [kz*kchunk, (kz+1)*kchunk)
// and writes its partial product to C + kz * sC (the split-K workspace) when sk > 1.
template <char OPA, char OPB>
__global__ __launch_bounds__(256) void k_zgemm(int64_t M, int64_t N, int64_t Kfull, cplx alpha,
                                               const cplx* __restrict__ A, int64_t lda, int64_t sA,
                                               const cplx* __restrict__ B, int64_t ldb, int64_t sB,
                                               cplx beta, cplx* __restrict__ C, int64_t ldc, int64_t sC, int sk,
                                               int64_t kchunk) {
  __shared__ double Ar[BK][LDS_STRIDE], Ai[BK][LDS_STRIDE], Br[BK][LDS_STRIDE], Bi[BK][LDS_STRIDE];
  const int64_t bz = blockIdx.z / sk, kz = blockIdx.z % sk;
  A += bz * sA;
  B += bz * sB;
  C += (sk > 1 ? bz * sk + kz : bz) * sC;   // split K: partial kz of product bz
  const int64_t kbeg = kz * kchunk;
  const int64_t K = (kbeg + kchunk < Kfull) ? kbeg + kchunk : Kfull;   // loop end
  const int64_t m0 = (int64_t)blockIdx.x * BM, n0 = (int64_t)blockIdx.y * BN;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = (wave & 1) * 32, wn = (wave >> 1) * 32;
  d4 accr[2][2], acci[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      accr[i][j] = d4{0.0, 0.0, 0.0, 0.0};
      acci[i][j] = d4{0.0, 0.0, 0.0, 0.0};
    }

  for (int64_t k0 = kbeg; k0 < K; k0 += BK) {
    // stage A (BM x BK) and B (BK x BN) tiles; the index order follows the
    // contiguous direction of the operand in memory
#pragma unroll
    for (int e = tid; e < BM * BK; e += 256) {
      int mm, kk;
      if (OPA == 'N') { mm = e % BM; kk = e / BM; }
      else { kk = e % BK; mm = e / BK; }
      const int64_t gm = m0 + mm, gk = k0 + kk;
      cplx v = (gm < M && gk < K) ? loadA<OPA>(A, lda, gm, gk) : mk(0.0, 0.0);
      Ar[kk][mm] = v.x;
      Ai[kk][mm] = v.y;
    }
#pragma unroll
    for (int e = tid; e < BN * BK; e += 256) {
      int nn, kk;
      if (OPB == 'N' || OPB == 'R') { kk = e % BK; nn = e / BK; }
      else { nn = e % BN; kk = e / BN; }
      const int64_t gn = n0 + nn, gk = k0 + kk;
      cplx v = (gn < N && gk < K) ? loadB<OPB>(B, ldb, gk, gn) : mk(0.0, 0.0);
      Br[kk][nn] = v.x;
      Bi[kk][nn] = v.y;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < BK; kk += 4) {
      const int kl = kk + (lane >> 4);
      double ar[2], ai[2], br[2], bi[2];
#pragma unroll
      for (int i = 0; i < 2; ++i) {
        ar[i] = Ar[kl][wm + 16 * i + (lane & 15)];
        ai[i] = Ai[kl][wm + 16 * i + (lane & 15)];
        br[i] = Br[kl][wn + 16 * i + (lane & 15)];
        bi[i] = Bi[kl][wn + 16 * i + (lane & 15)];
      }
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
          accr[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[i], br[j], accr[i][j], 0, 0, 0);
          accr[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(-ai[i], bi[j], accr[i][j], 0, 0, 0);
          acci[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[i], bi[j], acci[i][j], 0, 0, 0);
          acci[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(ai[i], br[j], acci[i][j], 0, 0, 0);
        }
    }
    __syncthreads();
  }
  // D[row][col] of a 16x16 MFMA tile sits in element (row / 4) of lane col + 16 * (row % 4)
  // (measured on gfx950 with a one-hot probe of __builtin_amdgcn_mfma_f64_16x16x4f64)
  const bool beta_zero = (beta.x == 0.0 && beta.y == 0.0);
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t row = m0 + wm + 16 * i + (lane >> 4) + 4 * r;
        const int64_t col = n0 + wn + 16 * j + (lane & 15);
        if (row < M && col < N) {
          cplx acc = mk(accr[i][j][r], acci[i][j][r]);
          cplx out = cmul(alpha, acc);
          if (!beta_zero) cfma(out, beta, C[row + ldc * col]);
          C[row + ldc * col] = out;
        }
      }
}

// Software-pipelined variant: the next K tile is fetched from global memory into registers
// while the current one is multiplied out of LDS, so load latency hides behind the MFMAs.
// KB: K per tile (16 or 32); PAD: LDS row padding in doubles (bank spread of the
// transposed-operand stores and of the MFMA operand reads).
// UPPER: only the tiles that reach the upper triangle (a Hermitian product; the tiles
// entirely below the diagonal leave at once)
template <char OPA, char OPB, int KB, int PAD, bool UPPER = false>
__global__ __launch_bounds__(256) void k_zgemm2(int64_t M, int64_t N, int64_t Kfull, cplx alpha,
                                                const cplx* __restrict__ A, int64_t lda, int64_t sA,
                                                const cplx* __restrict__ B, int64_t ldb, int64_t sB,
                                                cplx beta, cplx* __restrict__ C, int64_t ldc, int64_t sC, int sk,
                                                int64_t kchunk) {
  int64_t mt, nt;
  xcd_tile(gridDim.x, gridDim.y, blockIdx.x, blockIdx.y, mt, nt);
  if (UPPER && mt * BM > nt * BN + BN - 1) return;   // uniform
  constexpr int LS = BM + PAD;               // BM == BN
  constexpr int NE = BM * KB / 256;          // elements of each operand tile per thread
  __shared__ double Ar[KB][LS], Ai[KB][LS], Br[KB][LS], Bi[KB][LS];
  const int64_t bz = blockIdx.z / sk, kz = blockIdx.z % sk;
  A += bz * sA;
  B += bz * sB;
  C += (sk > 1 ? bz * sk + kz : bz) * sC;   // split K: partial kz of product bz
  const int64_t kbeg = kz * kchunk;
  const int64_t kend = (kbeg + kchunk < Kfull) ? kbeg + kchunk : Kfull;
  const int64_t m0 = mt * BM, n0 = nt * BN;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = (wave & 1) * 32, wn = (wave >> 1) * 32;
  d4 accr[2][2], acci[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      accr[i][j] = d4{0.0, 0.0, 0.0, 0.0};
      acci[i][j] = d4{0.0, 0.0, 0.0, 0.0};
    }
  cplx ra[NE], rb[NE];
  auto gload = [&](int64_t k0) {
#pragma unroll
    for (int r = 0; r < NE; ++r) {
      const int e = tid + 256 * r;
      int mm, kk, nn, kb;
      if (OPA == 'N') { mm = e % BM; kk = e / BM; } else { kk = e % KB; mm = e / KB; }
      if (OPB == 'N' || OPB == 'R') { kb = e % KB; nn = e / KB; } else { nn = e % BN; kb = e / BN; }
      const int64_t gm = m0 + mm, gk = k0 + kk, gn = n0 + nn, gkb = k0 + kb;
      ra[r] = (gm < M && gk < kend) ? loadA<OPA>(A, lda, gm, gk) : mk(0.0, 0.0);
      rb[r] = (gn < N && gkb < kend) ? loadB<OPB>(B, ldb, gkb, gn) : mk(0.0, 0.0);
    }
  };
  auto sstore = [&]() {
#pragma unroll
    for (int r = 0; r < NE; ++r) {
      const int e = tid + 256 * r;
      int mm, kk, nn, kb;
      if (OPA == 'N') { mm = e % BM; kk = e / BM; } else { kk = e % KB; mm = e / KB; }
      if (OPB == 'N' || OPB == 'R') { kb = e % KB; nn = e / KB; } else { nn = e % BN; kb = e / BN; }
      Ar[kk][mm] = ra[r].x;
      Ai[kk][mm] = ra[r].y;
      Br[kb][nn] = rb[r].x;
      Bi[kb][nn] = rb[r].y;
    }
  };
  if (kbeg < kend) {
    gload(kbeg);
    sstore();
  }
  __syncthreads();
  for (int64_t k0 = kbeg; k0 < kend; k0 += KB) {
    const bool more = k0 + KB < kend;
    if (more) gload(k0 + KB);
#pragma unroll
    for (int kk = 0; kk < KB; kk += 4) {
      const int kl = kk + (lane >> 4);
      double ar[2], ai[2], br[2], bi[2];
#pragma unroll
      for (int i = 0; i < 2; ++i) {
        ar[i] = Ar[kl][wm + 16 * i + (lane & 15)];
        ai[i] = Ai[kl][wm + 16 * i + (lane & 15)];
        br[i] = Br[kl][wn + 16 * i + (lane & 15)];
        bi[i] = Bi[kl][wn + 16 * i + (lane & 15)];
      }
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
          accr[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[i], br[j], accr[i][j], 0, 0, 0);
          accr[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(-ai[i], bi[j], accr[i][j], 0, 0, 0);
          acci[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[i], bi[j], acci[i][j], 0, 0, 0);
          acci[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(ai[i], br[j], acci[i][j], 0, 0, 0);
        }
    }
    __syncthreads();
    if (more) {
      sstore();
      __syncthreads();
    }
  }
  const bool beta_zero = (beta.x == 0.0 && beta.y == 0.0);
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t row = m0 + wm + 16 * i + (lane >> 4) + 4 * r;
        const int64_t col = n0 + wn + 16 * j + (lane & 15);
        if (row < M && col < N) {
          cplx acc = mk(accr[i][j][r], acci[i][j][r]);
          cplx out = cmul(alpha, acc);
          if (!beta_zero) cfma(out, beta, C[row + ldc * col]);
          C[row + ldc * col] = out;
        }
      }
}

// ---- shapes with too few output tiles for the MFMA kernel (the block Davidson of the CI
// matrix: Rayleigh-matrix columns, Gram-Schmidt coefficients, subspace updates, H x)

__device__ inline cplx out_value(cplx acc, cplx alpha, cplx beta, const cplx* c) {
  cplx out = cmul(alpha, acc);
  if (beta.x != 0.0 || beta.y != 0.0) cfma(out, beta, *c);
  return out;
}

// op(A) = A^H or A^T, op(B) = B or conj(B) with a small output (M x N): both operands are read
// along K, contiguous.  One wave per row m and up to 4 columns; the lanes stride over K and
// their partial sums meet in a fixed butterfly, so the result does not depend on timing.
constexpr int DOT_NB = 4, DOT_U = 4;
template <char OPA, char OPB>
__global__ __launch_bounds__(256) void k_zgemm_dot(int64_t M, int64_t N, int64_t K, cplx alpha,
                                                   const cplx* __restrict__ A, int64_t lda,
                                                   const cplx* __restrict__ B, int64_t ldb, cplx beta,
                                                   cplx* __restrict__ C, int64_t ldc) {
  const int lane = threadIdx.x & 63;
  const int64_t w = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);   // wave-uniform
  const int64_t nb = (N + DOT_NB - 1) / DOT_NB;
  if (w >= M * nb) return;
  const int64_t m = w % M, n0 = (w / M) * DOT_NB;
  const int nn = (int)(N - n0 < DOT_NB ? N - n0 : DOT_NB);
  cplx acc[DOT_NB];
#pragma unroll
  for (int j = 0; j < DOT_NB; ++j) acc[j] = mk(0.0, 0.0);
  const cplx* a = A + lda * m;
  const cplx* b = B + ldb * n0;
  // DOT_U strides of K per round: all their loads are issued before the first multiply
  for (int64_t k0 = lane; k0 < K; k0 += 64 * DOT_U) {
    cplx x[DOT_U], y[DOT_U][DOT_NB];
#pragma unroll
    for (int u = 0; u < DOT_U; ++u) {
      const int64_t k = k0 + 64 * u;
      x[u] = k < K ? a[k] : mk(0.0, 0.0);
#pragma unroll
      for (int j = 0; j < DOT_NB; ++j) y[u][j] = (k < K && j < nn) ? b[k + ldb * j] : mk(0.0, 0.0);
    }
#pragma unroll
    for (int u = 0; u < DOT_U; ++u) {
      if (k0 + 64 * u >= K) break;
      const cplx xv = OPA == 'C' ? cconj(x[u]) : x[u];
#pragma unroll
      for (int j = 0; j < DOT_NB; ++j)
        if (j < nn) cfma(acc[j], xv, OPB == 'R' ? cconj(y[u][j]) : y[u][j]);
    }
  }
#pragma unroll
  for (int j = 0; j < DOT_NB; ++j) {
    double re = acc[j].x, im = acc[j].y;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      re += __shfl_xor(re, off);
      im += __shfl_xor(im, off);
    }
    acc[j] = mk(re, im);
  }
  if (lane < nn) {
    cplx s = acc[0];
#pragma unroll
    for (int j = 1; j < DOT_NB; ++j)
      if (lane == j) s = acc[j];
    cplx* c = C + m + ldc * (n0 + lane);
    *c = out_value(s, alpha, beta, c);
  }
}

// The same products with very few outputs (the block Davidson's Rayleigh columns and
// Gram-Schmidt coefficients: at most a few hundred rows over K = ndet): one workgroup per row m
// and up to 4 columns, its 256 threads striding over K; the partial sums meet in a butterfly
// per wave and then in wave order through LDS, so the result does not depend on timing.
template <char OPA, char OPB>
__global__ __launch_bounds__(256) void k_zgemm_dotwg(int64_t M, int64_t N, int64_t K, cplx alpha,
                                                     const cplx* __restrict__ A, int64_t lda,
                                                     const cplx* __restrict__ B, int64_t ldb, cplx beta,
                                                     cplx* __restrict__ C, int64_t ldc) {
  __shared__ cplx part[4][DOT_NB];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int64_t w = blockIdx.x;
  const int64_t m = w % M, n0 = (w / M) * DOT_NB;
  const int nn = (int)(N - n0 < DOT_NB ? N - n0 : DOT_NB);
  cplx acc[DOT_NB];
#pragma unroll
  for (int j = 0; j < DOT_NB; ++j) acc[j] = mk(0.0, 0.0);
  const cplx* a = A + lda * m;
  const cplx* b = B + ldb * n0;
  for (int64_t k0 = tid; k0 < K; k0 += 256 * DOT_U) {
    cplx x[DOT_U], y[DOT_U][DOT_NB];
#pragma unroll
    for (int u = 0; u < DOT_U; ++u) {
      const int64_t k = k0 + 256 * u;
      x[u] = k < K ? a[k] : mk(0.0, 0.0);
#pragma unroll
      for (int j = 0; j < DOT_NB; ++j) y[u][j] = (k < K && j < nn) ? b[k + ldb * j] : mk(0.0, 0.0);
    }
#pragma unroll
    for (int u = 0; u < DOT_U; ++u) {
      if (k0 + 256 * u >= K) break;
      const cplx xv = OPA == 'C' ? cconj(x[u]) : x[u];
#pragma unroll
      for (int j = 0; j < DOT_NB; ++j)
        if (j < nn) cfma(acc[j], xv, OPB == 'R' ? cconj(y[u][j]) : y[u][j]);
    }
  }
#pragma unroll
  for (int j = 0; j < DOT_NB; ++j) {
    double re = acc[j].x, im = acc[j].y;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      re += __shfl_xor(re, off);
      im += __shfl_xor(im, off);
    }
    if (lane == 0) part[wave][j] = mk(re, im);
  }
  __syncthreads();
  if (tid < nn) {
    cplx s = part[0][tid];
#pragma unroll
    for (int q = 1; q < 4; ++q) s = cadd(s, part[q][tid]);
    cplx* c = C + m + ldc * (n0 + tid);
    *c = out_value(s, alpha, beta, c);
  }
}

// op(A) = A, op(B) = B or conj(B) with a short K: one thread per output element
template <char OPB>
__global__ __launch_bounds__(256) void k_zgemm_smallk(int64_t M, int64_t N, int64_t K, cplx alpha,
                                                      const cplx* __restrict__ A, int64_t lda,
                                                      const cplx* __restrict__ B, int64_t ldb, cplx beta,
                                                      cplx* __restrict__ C, int64_t ldc) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= M * N) return;
  const int64_t m = i % M, n = i / M;
  cplx acc = mk(0.0, 0.0);
  for (int64_t k = 0; k < K; ++k) {
    cplx y = B[k + ldb * n];
    if (OPB == 'R') y = cconj(y);
    cfma(acc, A[m + lda * k], y);
  }
  cplx* c = C + m + ldc * n;
  *c = out_value(acc, alpha, beta, c);
}

// op(A) = A, op(B) = B or conj(B) with at most TS_NMAX columns and a long K: thread = row m,
// all columns; blockIdx.y = K-split index, partial products to W[kz][n][m] (summed by
// k_splitk_reduce in a fixed order)
constexpr int TS_NMAX = 8, TS_U = 8, TS_KMAX = 64;
template <char OPB>
__global__ __launch_bounds__(256) void k_zgemm_tall(int64_t M, int64_t N, int64_t K, const cplx* __restrict__ A,
                                                    int64_t lda, const cplx* __restrict__ B, int64_t ldb,
                                                    cplx* __restrict__ W, int64_t kchunk) {
  __shared__ cplx Bs[TS_KMAX][TS_NMAX];   // this split's rows of op(B), read as broadcasts
  const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t kz = blockIdx.y, k0 = kz * kchunk;
  const int kc = (int)((k0 + kchunk < K ? k0 + kchunk : K) - k0);   // <= TS_KMAX (host-checked)
  for (int e = threadIdx.x; e < kc * N; e += blockDim.x) {
    const int kk = e % kc, j = e / kc;
    const cplx y = B[k0 + kk + ldb * j];
    Bs[kk][j] = OPB == 'R' ? cconj(y) : y;
  }
  __syncthreads();
  if (m >= M) return;
  cplx acc[TS_NMAX];
#pragma unroll
  for (int j = 0; j < TS_NMAX; ++j) acc[j] = mk(0.0, 0.0);
  const cplx* a = A + m + lda * k0;
  // TS_U rows of K per round: their loads of A are issued before the first multiply
  for (int kb = 0; kb < kc; kb += TS_U) {
    cplx x[TS_U];
#pragma unroll
    for (int u = 0; u < TS_U; ++u) x[u] = kb + u < kc ? a[lda * (kb + u)] : mk(0.0, 0.0);
#pragma unroll
    for (int u = 0; u < TS_U; ++u) {
      if (kb + u >= kc) break;
#pragma unroll
      for (int j = 0; j < TS_NMAX; ++j)
        if (j < N) cfma(acc[j], x[u], Bs[kb + u][j]);
    }
  }
  cplx* w = W + M * N * kz + m;
#pragma unroll
  for (int j = 0; j < TS_NMAX; ++j)   // static indices: acc stays in registers
    if (j < N) w[M * j] = acc[j];
}

// the strict lower triangle as the conjugate of the upper one (the diagonal is kept)
__global__ void k_lower_from_upper(int64_t m, cplx* __restrict__ C, int64_t ldc) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= m * m) return;
  const int64_t i = t % m, j = t / m;
  if (i > j) C[i + ldc * j] = cconj(C[j + ldc * i]);
}

// C = sum_kz W[kz] on and above the diagonal (the split-K partials of a Hermitian product,
// summed in a fixed order), the strict lower triangle as the conjugate of the upper one
__global__ void k_herk_reduce(int64_t m, int sk, const cplx* __restrict__ W, cplx* __restrict__ C, int64_t ldc) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= m * m) return;
  const int64_t i = t % m, j = t / m;
  if (i > j) return;
  cplx acc = mk(0.0, 0.0);
  int z = 0;
  for (; z + 4 <= sk; z += 4) {
    cplx v[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) v[u] = W[t + m * m * (z + u)];
#pragma unroll
    for (int u = 0; u < 4; ++u) acc = cadd(acc, v[u]);
  }
  for (; z < sk; ++z) acc = cadd(acc, W[t + m * m * z]);
  C[i + ldc * j] = acc;
  if (i < j) C[j + ldc * i] = cconj(acc);
}

// C = alpha * sum_kz W[kz] + beta * C (the split-K partial products, fixed summation order);
// blockIdx.y: the product of a batch (partials W + (bz sk + kz) M N, output C + bz sC)
__global__ void k_splitk_reduce(int64_t M, int64_t N, int sk, const cplx* __restrict__ W, cplx alpha, cplx beta,
                                cplx* __restrict__ C, int64_t ldc, int64_t sC) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= M * N) return;
  W += (int64_t)blockIdx.y * sk * M * N;
  C += (int64_t)blockIdx.y * sC;
  const int64_t m = i % M, n = i / M;
  cplx acc = mk(0.0, 0.0);
  int z = 0;
  for (; z + 8 <= sk; z += 8) {   // eight loads in flight, summed in order
    cplx v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) v[u] = W[i + M * N * (z + u)];
#pragma unroll
    for (int u = 0; u < 8; ++u) acc = cadd(acc, v[u]);
  }
  for (; z < sk; ++z) acc = cadd(acc, W[i + M * N * z]);
  cplx out = cmul(alpha, acc);
  if (beta.x != 0.0 || beta.y != 0.0) cfma(out, beta, C[m + ldc * n]);
  C[m + ldc * n] = out;
}

// grow-only split-K workspaces, one per stream (kernels on one stream run in order).  The
// owner of a stream releases its workspace before destroying the stream; the map itself is
// never destroyed, so no device memory is freed by static destructors at process exit,
// when the HIP runtime may already be gone.
std::mutex g_ws_mu;
std::map<hipStream_t, DevBuf<cplx>>* g_ws = new std::map<hipStream_t, DevBuf<cplx>>;
cplx* splitk_workspace(hipStream_t s, size_t n) {
  std::lock_guard<std::mutex> lk(g_ws_mu);
  DevBuf<cplx>& b = (*g_ws)[s];
  b.ensure(n);
  return b.get();
}

GemmStats* g_stats = nullptr;
// kernel variant: 0 k_zgemm, 1..4 k_zgemm2 with (KB, PAD) = (16,1), (16,16), (32,1), (32,16).
// Default 1: measured fastest on every shape of the path (e.g. 61 vs 46 TF/s on 2048^3).
int g_variant = [] {
  const char* v = std::getenv("DC2_GEMM_VARIANT");
  return v ? std::atoi(v) : 1;
}();
std::mutex g_stats_mu;
thread_local int g_tag = 0;
thread_local bool g_no_ws = false;

// tuned choices by exact shape (heap, never destroyed; looked up only when not empty)
struct TuneKey {
  char oa, ob;
  int64_t M, N, K, batch;
  bool operator<(const TuneKey& o) const {
    return std::tie(oa, ob, M, N, K, batch) < std::tie(o.oa, o.ob, o.M, o.N, o.K, o.batch);
  }
};
struct TuneChoice {
  int variant, split;
};
std::mutex g_tune_mu;
std::map<TuneKey, TuneChoice>* g_tune = new std::map<TuneKey, TuneChoice>;
std::atomic<bool> g_tune_any{false};

bool tuned_choice(char oa, char ob, int64_t M, int64_t N, int64_t K, int64_t batch, TuneChoice& c) {
  if (!g_tune_any.load(std::memory_order_relaxed)) return false;
  std::lock_guard<std::mutex> lk(g_tune_mu);
  auto it = g_tune->find(TuneKey{oa, ob, M, N, K, batch});
  if (it == g_tune->end()) return false;
  c = it->second;
  return true;
}

// Kernels for the shapes the 64x64 MFMA tiles serve badly (single products only): small
// outputs with A read along K (k_zgemm_dot), a short K (k_zgemm_smallk), at most 8 columns
// and a long K (k_zgemm_tall).  Returns false when the MFMA kernel is the one to use.
template <char OPA, char OPB>
bool launch_skinny(hipStream_t s, int64_t M, int64_t N, int64_t K, cplx alpha, const cplx* A, int64_t lda,
                   const cplx* B, int64_t ldb, cplx beta, cplx* C, int64_t ldc) {
  constexpr bool b_along_k = OPB == 'N' || OPB == 'R';
  if constexpr ((OPA == 'C' || OPA == 'T') && b_along_k) {
    // a few outputs, or a matrix-vector product over a long K (one wave per output row
    // reads its column of A once; on 64 x 64 MFMA tiles N = 1 would use 1/64 of each tile)
    if (M * N > 2048 && !(N <= DOT_NB && K >= 256)) return false;
    const int64_t waves = M * ((N + DOT_NB - 1) / DOT_NB);
    if (waves < 512 && K >= 1024) {   // too few rows to keep the device busy: a workgroup per row
      hipLaunchKernelGGL((k_zgemm_dotwg<OPA, OPB>), dim3((unsigned)waves), dim3(256), 0, s, M, N, K, alpha, A,
                         lda, B, ldb, beta, C, ldc);
      return true;
    }
    hipLaunchKernelGGL((k_zgemm_dot<OPA, OPB>), dim3((unsigned)cdiv(waves, 4)), dim3(256), 0, s, M, N, K, alpha, A,
                       lda, B, ldb, beta, C, ldc);
    return true;
  } else if constexpr (OPA == 'N' && b_along_k) {
    if (K <= 64 && M * N <= 65536) {
      hipLaunchKernelGGL((k_zgemm_smallk<OPB>), dim3((unsigned)cdiv(M * N, 256)), dim3(256), 0, s, M, N, K, alpha, A,
                         lda, B, ldb, beta, C, ldc);
      return true;
    }
    if (N > TS_NMAX || K <= 64 || M > 65536 || g_no_ws) return false;
    const int64_t rb = cdiv(M, 256);
    const int64_t want = std::max<int64_t>(cdiv(K, TS_KMAX), std::min<int64_t>({cdiv(256, rb), 32, cdiv(K, 16)}));
    const int64_t kchunk = cdiv(K, want), sk = cdiv(K, kchunk);
    if (kchunk > TS_KMAX || sk >= 65536) return false;
    cplx* W = splitk_workspace(s, (size_t)sk * M * N);
    hipLaunchKernelGGL((k_zgemm_tall<OPB>), dim3((unsigned)rb, (unsigned)sk), dim3(256), 0, s, M, N, K, A, lda, B,
                       ldb, W, kchunk);
    hipLaunchKernelGGL(k_splitk_reduce, dim3(cdiv(M * N, 256)), dim3(256), 0, s, M, N, (int)sk, W, alpha, beta, C,
                       ldc, (int64_t)0);
    return true;
  }
  return false;
}

// force_variant > 0: that MFMA variant and split-K factor force_split (the tuner's trials; no
// skinny kernels, no table lookup)
template <char OPA, char OPB>
void launch(hipStream_t s, int64_t M, int64_t N, int64_t K, cplx alpha, const cplx* A, int64_t lda, int64_t sA,
            const cplx* B, int64_t ldb, int64_t sB, cplx beta, cplx* C, int64_t ldc, int64_t sC, int64_t batch,
            int force_variant = 0, int force_split = 1) {
  GemmRecord r{};
  const bool acct = g_stats && g_stats->on;
  if (acct) {
    DC2_HIP(hipEventCreate(&r.a));
    DC2_HIP(hipEventCreate(&r.b));
    DC2_HIP(hipEventRecord(r.a, s));
  }
  auto record = [&](bool mfma) {
    if (!acct) return;
    r.mfma = mfma;
    DC2_HIP(hipEventRecord(r.b, s));
    r.flops = 8.0 * (double)M * (double)N * (double)K * (double)batch;
    r.tag = g_tag;
    r.opA = OPA;
    r.opB = OPB;
    r.M = M;
    r.N = N;
    r.K = K;
    r.batch = batch;
    std::lock_guard<std::mutex> lk(g_stats_mu);
    g_stats->rec.push_back(r);
  };
  if (force_variant == 0 && batch == 1 && g_variant != 0 &&
      launch_skinny<OPA, OPB>(s, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc)) {
    DC2_HIP(hipGetLastError());
    record(false);
    return;
  }
  // split K over workgroups when the output has few tiles and K is long
  const int64_t tiles = (int64_t)cdiv(M, BM) * cdiv(N, BN) * batch;
  int sk = 1;
  int64_t kchunk = K;
  if (tiles < 128 && K >= 256) {
    int64_t want = std::min<int64_t>(K / 128, (256 + tiles - 1) / tiles);
    want = std::max<int64_t>(2, std::min<int64_t>(want, 64));
    kchunk = ((K + want - 1) / want + BK - 1) / BK * BK;
    sk = (int)((K + kchunk - 1) / kchunk);
  }
  int variant = g_variant;
  TuneChoice tc{0, 0};
  int want_split = 0;   // 0: the rule above
  if (force_variant > 0) {
    variant = force_variant;
    want_split = force_split;
  } else if (g_variant != 0 && tuned_choice(OPA, OPB, M, N, K, batch, tc)) {
    variant = tc.variant;
    want_split = tc.split;
  }
  if (want_split > 0) {
    sk = 1;
    kchunk = K;
    if (want_split > 1) {
      kchunk = ((K + want_split - 1) / want_split + BK - 1) / BK * BK;
      sk = (int)((K + kchunk - 1) / kchunk);
    }
  }
  if (batch * sk >= 65536 || g_no_ws) {   // one grid dimension holds batch x split
    sk = 1;
    kchunk = K;
  }
  dim3 grid(cdiv(M, BM), cdiv(N, BN), (unsigned)(batch * sk));
  cplx* Cw = C;
  int64_t ldw = ldc, sw = sC;
  cplx a1 = alpha, b1 = beta;
  if (sk > 1) {
    Cw = splitk_workspace(s, (size_t)sk * M * N * batch);
    ldw = M;
    sw = M * N;
    a1 = mk(1.0, 0.0);
    b1 = mk(0.0, 0.0);
  }
  switch (variant) {
    case 1:
      hipLaunchKernelGGL((k_zgemm2<OPA, OPB, 16, 1>), grid, dim3(256), 0, s, M, N, K, a1, A, lda, sA, B, ldb, sB, b1,
                         Cw, ldw, sw, sk, kchunk);
      break;
    case 2:
      hipLaunchKernelGGL((k_zgemm2<OPA, OPB, 16, 16>), grid, dim3(256), 0, s, M, N, K, a1, A, lda, sA, B, ldb, sB,
                         b1, Cw, ldw, sw, sk, kchunk);
      break;
    case 3:
      hipLaunchKernelGGL((k_zgemm2<OPA, OPB, 32, 1>), grid, dim3(256), 0, s, M, N, K, a1, A, lda, sA, B, ldb, sB, b1,
                         Cw, ldw, sw, sk, kchunk);
      break;
    case 4:
      hipLaunchKernelGGL((k_zgemm2<OPA, OPB, 32, 16>), grid, dim3(256), 0, s, M, N, K, a1, A, lda, sA, B, ldb, sB,
                         b1, Cw, ldw, sw, sk, kchunk);
      break;
    default:
      hipLaunchKernelGGL((k_zgemm<OPA, OPB>), grid, dim3(256), 0, s, M, N, K, a1, A, lda, sA, B, ldb, sB, b1, Cw,
                         ldw, sw, sk, kchunk);
  }
  if (sk > 1)
    hipLaunchKernelGGL(k_splitk_reduce, dim3(cdiv(M * N, 256), (unsigned)batch), dim3(256), 0, s, M, N, sk, Cw, alpha,
                       beta, C, ldc, sC);
  DC2_HIP(hipGetLastError());
  record(true);
}

template <char OPA>
void dispatchB(char opB, hipStream_t s, int64_t M, int64_t N, int64_t K, cplx alpha, const cplx* A, int64_t lda,
               int64_t sA, const cplx* B, int64_t ldb, int64_t sB, cplx beta, cplx* C, int64_t ldc, int64_t sC,
               int64_t batch, int fv = 0, int fs = 1) {
  switch (opB) {
    case 'N': launch<OPA, 'N'>(s, M, N, K, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc, sC, batch, fv, fs); break;
    case 'R': launch<OPA, 'R'>(s, M, N, K, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc, sC, batch, fv, fs); break;
    case 'T': launch<OPA, 'T'>(s, M, N, K, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc, sC, batch, fv, fs); break;
    case 'C': launch<OPA, 'C'>(s, M, N, K, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc, sC, batch, fv, fs); break;
    default: throw Error(1, "zgemm: bad opB");
  }
}

void dispatch(hipStream_t s, char opA, char opB, int64_t M, int64_t N, int64_t K, cplx alpha, const cplx* A,
              int64_t lda, int64_t sA, const cplx* B, int64_t ldb, int64_t sB, cplx beta, cplx* C, int64_t ldc,
              int64_t sC, int64_t batch, int fv = 0, int fs = 1) {
  switch (opA) {
    case 'N': dispatchB<'N'>(opB, s, M, N, K, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc, sC, batch, fv, fs); break;
    case 'T': dispatchB<'T'>(opB, s, M, N, K, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc, sC, batch, fv, fs); break;
    case 'C': dispatchB<'C'>(opB, s, M, N, K, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc, sC, batch, fv, fs); break;
    default: throw Error(1, "zgemm: bad opA");
  }
}
}  // namespace

void set_gemm_stats(GemmStats* st) {
  std::lock_guard<std::mutex> lk(g_stats_mu);
  g_stats = st;
}
void set_gemm_tag(int tag) { g_tag = tag; }
void set_gemm_no_workspace(bool on) { g_no_ws = on; }

void release_stream_workspace(hipStream_t s) {
  std::lock_guard<std::mutex> lk(g_ws_mu);
  g_ws->erase(s);
}

void zgemm(hipStream_t s, char opA, char opB, int64_t M, int64_t N, int64_t K, cplx alpha, const cplx* A,
           int64_t lda, int64_t sA, const cplx* B, int64_t ldb, int64_t sB, cplx beta, cplx* C, int64_t ldc,
           int64_t sC, int64_t batch) {
  if (M <= 0 || N <= 0 || batch <= 0) return;
  DC2_CHECK(batch < 65536, 1, "zgemm: batch too large for one grid dimension");
  DC2_CHECK(cdiv(M, BM) < (1u << 31) && cdiv(N, BN) < 65536, 1, "zgemm: matrix too large for the grid");
  dispatch(s, opA, opB, M, N, K, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc, sC, batch);
}

// C (m x m, ldc) = A A^H for A (m x k, lda): the tiles on and above the diagonal on the MFMA
// kernel, then the lower triangle as the conjugate of the upper one.  Few tiles over a long
// k (the determinant-space Gram matrices: 91 upper tiles over k = 1728) are split over k
// into about 1024 workgroups, at most 12 chunks of at least 128, summed in a fixed order by
// k_herk_reduce (m = 792, k = 1728: 421 us unsplit, 207 / 162 / 164 / 162 / 145 us in 3 / 4
// / 8 / 12 / 16 chunks).  split: 0 this rule, >= 1 that many k chunks (1: unsplit).
void zherk_upper(hipStream_t s, int64_t m, int64_t k, const cplx* A, int64_t lda, cplx* C, int64_t ldc, int split) {
  zgemm_herm(s, m, k, A, lda, A, lda, C, ldc, split);
}

// C (m x m, ldc) = A B^H for A, B (m x k) whose product is Hermitian (B = D A with D real
// diagonal over k, or B = A): the same upper-tile product and mirror as zherk_upper
void zgemm_herm(hipStream_t s, int64_t m, int64_t k, const cplx* A, int64_t lda, const cplx* B, int64_t ldb, cplx* C,
                int64_t ldc, int split) {
  if (m <= 0) return;
  DC2_CHECK(cdiv(m, BM) < 65536, 1, "zgemm_herm: matrix too large for the grid");
  DC2_CHECK(split >= 0 && split <= 64, 1, "zgemm_herm: split 0..64");
  const cplx one = mk(1.0, 0.0), zero = mk(0.0, 0.0);
  const int64_t T = cdiv(m, BM), tiles = T * (T + 1) / 2;
  int64_t want = split;
  if (want == 0 && tiles < 256 && k >= 256) want = std::min<int64_t>(std::min<int64_t>(k / 128, cdiv(1024, tiles)), 12);
  int64_t kchunk = k;
  int sk = 1;
  if (want > 1 && k > 0) {
    kchunk = (cdiv(k, want) + BK - 1) / BK * BK;
    sk = (int)cdiv(k, kchunk);
  }
  if (sk > 1) {
    cplx* W = splitk_workspace(s, (size_t)sk * m * m);
    dim3 grid(T, T, (unsigned)sk);
    hipLaunchKernelGGL((k_zgemm2<'N', 'C', 16, 1, true>), grid, dim3(256), 0, s, m, m, k, one, A, lda, (int64_t)0, B,
                       ldb, (int64_t)0, zero, W, m, m * m, sk, kchunk);
    hipLaunchKernelGGL(k_herk_reduce, dim3(cdiv(m * m, 256)), dim3(256), 0, s, m, sk, W, C, ldc);
    DC2_HIP(hipGetLastError());
    return;
  }
  dim3 grid(T, T, 1);
  hipLaunchKernelGGL((k_zgemm2<'N', 'C', 16, 1, true>), grid, dim3(256), 0, s, m, m, k, one, A, lda, (int64_t)0, B,
                     ldb, (int64_t)0, zero, C, ldc, (int64_t)0, 1, k);
  hipLaunchKernelGGL(k_lower_from_upper, dim3(cdiv(m * m, 256)), dim3(256), 0, s, m, C, ldc);
  DC2_HIP(hipGetLastError());
}

void gemm_set_choice(char opA, char opB, int64_t M, int64_t N, int64_t K, int64_t batch, int variant, int split) {
  DC2_CHECK(variant >= 1 && variant <= 4 && split >= 1 && split <= 64, 1, "gemm choice: variant 1..4, split 1..64");
  std::lock_guard<std::mutex> lk(g_tune_mu);
  (*g_tune)[TuneKey{opA, opB, M, N, K, batch}] = TuneChoice{variant, split};
  g_tune_any.store(true);
}

void gemm_clear_choices() {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  g_tune->clear();
  g_tune_any.store(false);
}

int64_t gemm_choice_count() {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  return (int64_t)g_tune->size();
}

namespace {
// p[i] = seeded values in (-0.5, 0.5) + i (-0.5, 0.5) from a 64-bit integer hash of (seed, i)
// (splitmix64's finaliser), a grid-stride loop
__global__ __launch_bounds__(256) void k_fill_hash(cplx* __restrict__ p, int64_t n, uint64_t seed) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    uint64_t z = seed + 0x9E3779B97F4A7C15ull * (uint64_t)(i + 1);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    const double re = (double)(z >> 40) * (1.0 / 16777216.0) - 0.5;
    const double im = (double)((z >> 16) & 0xFFFFFF) * (1.0 / 16777216.0) - 0.5;
    p[i] = mk(re, im);
  }
}

void fill_hash(hipStream_t s, cplx* p, int64_t n, uint64_t seed) {
  if (n <= 0) return;
  const int64_t blocks = std::min<int64_t>(cdiv(n, 256), 65535);
  hipLaunchKernelGGL(k_fill_hash, dim3((unsigned)blocks), dim3(256), 0, s, p, n, seed);
  DC2_HIP(hipGetLastError());
}
}  // namespace

void gemm_tune(hipStream_t s, char opA, char opB, int64_t M, int64_t N, int64_t K, int64_t batch, int reps,
               double* out) {
  DC2_CHECK(M > 0 && N > 0 && K > 0 && batch > 0 && batch < 65536 && reps >= 1, 1, "gemm_tune: invalid shape");
  DC2_CHECK(std::string("NTC").find(opA) != std::string::npos && std::string("NRTC").find(opB) != std::string::npos, 1,
            "gemm_tune: bad op");
  // scratch operands with seeded nonzero values (an all-zero product draws less power than
  // the real one, so its clocks and timings could differ)
  const int64_t lda = opA == 'N' ? M : K, ldb = (opB == 'N' || opB == 'R') ? K : N;
  const int64_t sA = M * K, sB = K * N, sC = M * N;
  DevBuf<cplx> A((size_t)(sA * batch)), B((size_t)(sB * batch)), C((size_t)(sC * batch));
  fill_hash(s, A.get(), sA * batch, 0x5eed0001ull);
  fill_hash(s, B.get(), sB * batch, 0x5eed0002ull);
  DC2_HIP(hipMemsetAsync(C.get(), 0, sC * batch * sizeof(cplx), s));
  hipEvent_t e0, e1;
  DC2_HIP(hipEventCreate(&e0));
  DC2_HIP(hipEventCreate(&e1));
  const cplx one = mk(1.0, 0.0), zero = mk(0.0, 0.0);
  auto time = [&](int fv, int fs) {   // fv = 0: the built-in choice (the table is bypassed below)
    dispatch(s, opA, opB, M, N, K, one, A.get(), lda, sA, B.get(), ldb, sB, zero, C.get(), M, sC, batch, fv, fs);
    DC2_HIP(hipEventRecord(e0, s));
    for (int r = 0; r < reps; ++r)
      dispatch(s, opA, opB, M, N, K, one, A.get(), lda, sA, B.get(), ldb, sB, zero, C.get(), M, sC, batch, fv, fs);
    DC2_HIP(hipEventRecord(e1, s));
    DC2_HIP(hipEventSynchronize(e1));
    float ms = 0.0f;
    DC2_HIP(hipEventElapsedTime(&ms, e0, e1));
    return 1e-3 * ms / reps;
  };
  // the built-in choice: variant g_variant and the split-K rule, without a table entry
  bool had = false;
  TuneChoice prev{0, 0};
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto it = g_tune->find(TuneKey{opA, opB, M, N, K, batch});
    if (it != g_tune->end()) { had = true; prev = it->second; g_tune->erase(it); }
    g_tune_any.store(!g_tune->empty());
  }
  const double t_builtin = time(0, 1);
  if (had) gemm_set_choice(opA, opB, M, N, K, batch, prev.variant, prev.split);
  const int64_t tiles = (int64_t)cdiv(M, BM) * cdiv(N, BN) * batch;
  double best = INFINITY;
  int bv = 1, bs = 1;
  for (int v = 1; v <= 4; ++v)
    for (int sp : {1, 2, 4, 8, 16, 32}) {
      if (sp > 1 && (K < 2 * BK * sp || tiles * sp > 4096 || batch * sp >= 65536)) continue;
      const double t = time(v, sp);
      if (t < best) { best = t; bv = v; bs = sp; }
    }
  DC2_HIP(hipEventDestroy(e0));
  DC2_HIP(hipEventDestroy(e1));
  DC2_HIP(hipStreamSynchronize(s));   // the scratch operands are freed next
  if (best < t_builtin) {
    gemm_set_choice(opA, opB, M, N, K, batch, bv, bs);
  } else {   // the built-in choice stays (it may split K by a factor outside the candidates)
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tune->erase(TuneKey{opA, opB, M, N, K, batch});
    g_tune_any.store(!g_tune->empty());
    bv = bs = 0;
    best = t_builtin;
  }
  out[0] = bv;
  out[1] = bs;
  out[2] = best;
  out[3] = t_builtin;
}

}  // namespace dc2
