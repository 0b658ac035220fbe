// Split-precision complex GEMM (Ozaki scheme) on CDNA4 INT8 matrix cores.
//
// C = A * B (complex FP64, alpha = 1, beta = 0) for the integral-transform passes,
// A[m][k] = in[k + lda*m] ('T') and B[k][j] = U[k + ldb*j] or its conjugate ('N'/'R').
//
// Every row m of A (and column j of B) is scaled by a power of two 2^-e so that its
// largest real or imaginary part lies in [1/2, 1), then cut into S signed 7-bit digits:
//   a = sum_{p<S} d_p 2^{-7(p+1)} + r 2^{-7S},  d_p in [-127, 127],  |r| < 1,
// each extraction exact in FP64.  A complex product is two real GEMMs over K' = 2K:
//   Re C = [Ar | Ai] [Br ; -Bi],   Im C = [Ar | Ai] [Bi ; Br]
// (conjugated B flips the sign of Bi; digit negation is exact).  The digit products
// of level L = p + q < S are accumulated exactly in INT32 with
// v_mfma_i32_16x16x64_i8 (|sum| <= S * K' * 127^2 < 2^31 for K' <= 8192) and
// recombined in FP64 as sum_L 2^{-7(L+2)} acc_L, smallest level first, times
// 2^{e_m + f_j}.  Dropped levels (L >= S) and the residuals bound the error by
// about (S + 1) 2^{-7S} K' max|row of A| max|column of B|.
#include "dc2_kernels.hpp"

namespace dc2 {
namespace {

typedef int v4i __attribute__((ext_vector_type(4)));

constexpr int OZ_KC = 64;        // K per INT8 MFMA
constexpr int OZ_BM = 64, OZ_BN = 32;

// exponent e with max(|re|, |im|) * 2^-e in [1/2, 1) (0 for a zero vector)
__device__ inline int scale_exponent(double amax) {
  if (!(amax > 0.0)) return 0;
  return ilogb(amax) + 1;
}

// One wave per row (A) or column (B) of K complex values x[k * inc] (inc = 1: contiguous).
// Writes S slices of 2*Kp digits: out[p][row][0..Kp) = digits of s1*Re, [Kp..2Kp) = digits
// of s2*Im (zero padded), where (s1, s2) = (+1, +1) or the signs that build [Br;-Bi] etc.
template <int S>
__global__ __launch_bounds__(256) void k_oz_split(int64_t nrows, int64_t K, int64_t Kp, const cplx* __restrict__ X,
                                                  int64_t ld, int first_im, int sign_first, int sign_second,
                                                  int8_t* __restrict__ out, int* __restrict__ expo) {
  const int lane = threadIdx.x & 63;
  const int64_t row = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (row >= nrows) return;
  const cplx* x = X + ld * row;
  double amax = 0.0;
  for (int64_t k = lane; k < K; k += 64) amax = fmax(amax, fmax(fabs(x[k].x), fabs(x[k].y)));
  for (int o = 32; o > 0; o >>= 1) amax = fmax(amax, __shfl_xor(amax, o, 64));
  const int e = scale_exponent(amax);
  if (lane == 0) expo[row] = e;
  const int64_t rowlen = 2 * Kp, slice = nrows * rowlen;
  int8_t* o = out + rowlen * row;
  for (int64_t k = lane; k < Kp; k += 64) {
    double v1 = 0.0, v2 = 0.0;
    if (k < K) {
      const cplx z = x[k];
      v1 = ldexp(first_im ? z.y : z.x, -e) * sign_first;
      v2 = ldexp(first_im ? z.x : z.y, -e) * sign_second;
    }
#pragma unroll
    for (int p = 0; p < S; ++p) {
      v1 *= 128.0;
      v2 *= 128.0;
      const double d1 = trunc(v1), d2 = trunc(v2);
      v1 -= d1;
      v2 -= d2;
      o[p * slice + k] = (int8_t)d1;
      o[p * slice + Kp + k] = (int8_t)d2;
    }
  }
}

// C[m + ldc*j] = sum_L 2^{-7(L+2)} (acc_L) * 2^{ea[m] + eb[j]} for the re and im GEMMs.
// Workgroup: 4 waves, tile 64 x 32; wave (wm, wn) owns rows wm*32..+32 and columns wn*16..+16
// (two 16x16 MFMA tiles along m).
template <int S>
__global__ __launch_bounds__(256) void k_oz_gemm(int64_t M, int64_t N, int64_t K2, const int8_t* __restrict__ A8,
                                                 const int* __restrict__ ea, const int8_t* __restrict__ Bre,
                                                 const int8_t* __restrict__ Bim, const int* __restrict__ eb,
                                                 cplx* __restrict__ C, int64_t ldc) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  int64_t mt, nt;
  xcd_tile(gridDim.x, gridDim.y, blockIdx.x, blockIdx.y, mt, nt);
  const int64_t m0 = mt * OZ_BM + (wave & 1) * 32;
  const int64_t n0 = nt * OZ_BN + (wave >> 1) * 16;
  const int r16 = lane & 15, g = lane >> 4;
  const int64_t sliceA = M * K2, sliceB = N * K2;
  v4i accr[S][2], acci[S][2];
#pragma unroll
  for (int L = 0; L < S; ++L)
#pragma unroll
    for (int t = 0; t < 2; ++t) {
      accr[L][t] = v4i{0, 0, 0, 0};
      acci[L][t] = v4i{0, 0, 0, 0};
    }
  const int64_t ma0 = m0 + r16, ma1 = m0 + 16 + r16, jb = n0 + r16;
  const bool va0 = ma0 < M, va1 = ma1 < M, vb = jb < N;
  for (int64_t kc = 0; kc < K2; kc += OZ_KC) {
    const int64_t ko = kc + 16 * g;   // the 16 consecutive k of this lane
    v4i a[S][2], br[S], bi[S];
#pragma unroll
    for (int p = 0; p < S; ++p) {
      a[p][0] = va0 ? *reinterpret_cast<const v4i*>(A8 + p * sliceA + ma0 * K2 + ko) : v4i{0, 0, 0, 0};
      a[p][1] = va1 ? *reinterpret_cast<const v4i*>(A8 + p * sliceA + ma1 * K2 + ko) : v4i{0, 0, 0, 0};
      br[p] = vb ? *reinterpret_cast<const v4i*>(Bre + p * sliceB + jb * K2 + ko) : v4i{0, 0, 0, 0};
      bi[p] = vb ? *reinterpret_cast<const v4i*>(Bim + p * sliceB + jb * K2 + ko) : v4i{0, 0, 0, 0};
    }
#pragma unroll
    for (int p = 0; p < S; ++p)
#pragma unroll
      for (int q = 0; q < S - p; ++q)
#pragma unroll
        for (int t = 0; t < 2; ++t) {
          accr[p + q][t] = __builtin_amdgcn_mfma_i32_16x16x64_i8(a[p][t], br[q], accr[p + q][t], 0, 0, 0);
          acci[p + q][t] = __builtin_amdgcn_mfma_i32_16x16x64_i8(a[p][t], bi[q], acci[p + q][t], 0, 0, 0);
        }
  }
  // D[i][j] of a 16x16 tile: lane j + 16*(i/4), element i%4 (measured on gfx950)
  const int64_t j = n0 + r16;
  if (j >= N) return;
  const int fj = eb[j];
#pragma unroll
  for (int t = 0; t < 2; ++t)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int64_t m = m0 + 16 * t + 4 * g + r;
      if (m >= M) continue;
      double re = 0.0, im = 0.0;
#pragma unroll
      for (int L = S - 1; L >= 0; --L) {
        re += ldexp((double)accr[L][t][r], -7 * (L + 2));
        im += ldexp((double)acci[L][t][r], -7 * (L + 2));
      }
      const int sc = ea[m] + fj;
      C[m + ldc * j] = mk(ldexp(re, sc), ldexp(im, sc));
    }
}

template <int S>
void oz_launch(hipStream_t s, char opB, int64_t M, int64_t N, int64_t K, const cplx* A, int64_t lda, const cplx* B,
               int64_t ldb, cplx* C, int64_t ldc, int8_t* work, int* expo) {
  const int64_t Kh = ((K + 31) / 32) * 32;   // half of K' = 2*Kh is a multiple of 64
  const int64_t K2 = 2 * Kh;
  int8_t* A8 = work;
  int8_t* Bre = A8 + S * M * K2;
  int8_t* Bim = Bre + S * N * K2;
  int* ea = expo;
  int* eb = expo + M;
  const int conj = opB == 'R' ? -1 : 1;
  // A rows: [Ar | Ai]
  hipLaunchKernelGGL((k_oz_split<S>), dim3(cdiv(M, 4)), dim3(256), 0, s, M, K, Kh, A, lda, 0, 1, 1, A8, ea);
  // B columns: Re part [Br ; -Bi] (conj: [Br ; Bi]), Im part [Bi ; Br] (conj: [-Bi ; Br])
  hipLaunchKernelGGL((k_oz_split<S>), dim3(cdiv(N, 4)), dim3(256), 0, s, N, K, Kh, B, ldb, 0, 1, -conj, Bre, eb);
  hipLaunchKernelGGL((k_oz_split<S>), dim3(cdiv(N, 4)), dim3(256), 0, s, N, K, Kh, B, ldb, 1, conj, 1, Bim, eb);
  hipLaunchKernelGGL((k_oz_gemm<S>), dim3(cdiv(M, OZ_BM), cdiv(N, OZ_BN)), dim3(256), 0, s, M, N, K2, A8, ea, Bre,
                     Bim, eb, C, ldc);
  DC2_HIP(hipGetLastError());
}

}  // namespace

int64_t zgemm_ozaki_work_bytes(int slices, int64_t M, int64_t N, int64_t K) {
  const int64_t K2 = 2 * (((K + 31) / 32) * 32);
  return (int64_t)slices * (M + 2 * N) * K2;
}

void zgemm_ozaki(hipStream_t s, int slices, char opB, int64_t M, int64_t N, int64_t K, const cplx* A, int64_t lda,
                 const cplx* B, int64_t ldb, cplx* C, int64_t ldc, int8_t* work, int* expo) {
  if (M <= 0 || N <= 0) return;
  DC2_CHECK(K >= 1 && K <= 4096, 1, "zgemm_ozaki: K must be in 1..4096");
  DC2_CHECK(opB == 'N' || opB == 'R', 1, "zgemm_ozaki: opB must be N or R");
  DC2_CHECK(cdiv(M, OZ_BM) < (1u << 31) && cdiv(N, OZ_BN) < 65536 && cdiv(M, 4) < (1u << 31), 1,
            "zgemm_ozaki: matrix too large for the grid");
  switch (slices) {
    case 3: oz_launch<3>(s, opB, M, N, K, A, lda, B, ldb, C, ldc, work, expo); break;
    case 4: oz_launch<4>(s, opB, M, N, K, A, lda, B, ldb, C, ldc, work, expo); break;
    case 5: oz_launch<5>(s, opB, M, N, K, A, lda, B, ldb, C, ldc, work, expo); break;
    case 6: oz_launch<6>(s, opB, M, N, K, A, lda, B, ldb, C, ldc, work, expo); break;
    case 7: oz_launch<7>(s, opB, M, N, K, A, lda, B, ldb, C, ldc, work, expo); break;
    case 8: oz_launch<8>(s, opB, M, N, K, A, lda, B, ldb, C, ldc, work, expo); break;
    default: throw Error(1, "zgemm_ozaki: slices must be in 3..8");
  }
}

}  // namespace dc2
