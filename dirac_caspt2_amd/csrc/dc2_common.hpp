// Shared definitions for the CDNA4 (gfx950) CASPT2 kernels.
//
// Complex numbers are double2 {x = re, y = im}, the memory layout of Fortran
// complex*16 / C double _Complex.  Matrices are column-major (Fortran order)
// with explicit leading dimensions, as the reference's arrays are.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

namespace dc2 {

typedef double2 cplx;

__host__ __device__ inline cplx mk(double r, double i) { return make_double2(r, i); }
__host__ __device__ inline cplx cadd(cplx a, cplx b) { return mk(a.x + b.x, a.y + b.y); }
__host__ __device__ inline cplx csub(cplx a, cplx b) { return mk(a.x - b.x, a.y - b.y); }
__host__ __device__ inline cplx cmul(cplx a, cplx b) { return mk(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
__host__ __device__ inline cplx cconj(cplx a) { return mk(a.x, -a.y); }
__host__ __device__ inline cplx cscale(cplx a, double s) { return mk(a.x * s, a.y * s); }
__host__ __device__ inline double cabs2(cplx a) { return a.x * a.x + a.y * a.y; }
// a += b*c
__host__ __device__ inline void cfma(cplx& a, cplx b, cplx c) {
  a.x = fma(b.x, c.x, fma(-b.y, c.y, a.x));
  a.y = fma(b.x, c.y, fma(b.y, c.x, a.y));
}

// XCD-aware tile order of a 2D grid of Mt x Nt output tiles (blockIdx.x over Mt, .y over Nt):
// workgroups are dealt to the 8 XCDs in turn by their linear id, so the i-th workgroup an XCD
// receives takes tile i of that XCD's contiguous share of the tiles in n-fastest order.  The
// tiles of one row block (one stretch of A) then run on one XCD one after another and reuse
// that XCD's L2 copy of it, instead of being spread over all eight.  A bijection on the tiles.
__device__ inline void xcd_tile(int64_t Mt, int64_t Nt, int64_t bx, int64_t by, int64_t& mt, int64_t& nt) {
  const int64_t T = Mt * Nt, lin = by * Mt + bx;
  const int64_t q = T / 8, r = T % 8, x = lin % 8, i = lin / 8;
  const int64_t g = x * q + (x < r ? x : r) + i;
  mt = g / Nt;
  nt = g - mt * Nt;
}

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define DC2_HIP(expr)                                                                  \
  do {                                                                                 \
    hipError_t e_ = (expr);                                                            \
    if (e_ != hipSuccess)                                                              \
      throw ::dc2::Error(2, std::string("HIP error: ") + hipGetErrorString(e_) + " at " + \
                                    __FILE__ + ":" + std::to_string(__LINE__));        \
  } while (0)

#define DC2_CHECK(cond, code, msg)                       \
  do {                                                   \
    if (!(cond)) throw ::dc2::Error((code), (msg));      \
  } while (0)

// Device allocation owned by the library context.
template <class T>
struct DevBuf {
  T* p = nullptr;
  size_t n = 0;
  DevBuf() = default;
  explicit DevBuf(size_t count) { alloc(count); }
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
  DevBuf& operator=(DevBuf&& o) noexcept {
    if (this != &o) { release(); p = o.p; n = o.n; o.p = nullptr; o.n = 0; }
    return *this;
  }
  ~DevBuf() { release(); }
  void alloc(size_t count) {
    release();
    n = count;
    if (count) DC2_HIP(hipMalloc(&p, count * sizeof(T)));
  }
  // keep the allocation if it is large enough
  void ensure(size_t count) {
    if (count > n || p == nullptr) alloc(count);
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  T* get() const { return p; }
};

inline unsigned cdiv(size_t a, size_t b) { return (unsigned)((a + b - 1) / b); }

// --- GEMM (dc2_gemm.hip) -----------------------------------------------------
// Optional accounting of GEMM launches: HIP events around every launch on the
// launch stream, algorithmic flops 8*M*N*K*batch per launch, grouped by a tag.
struct GemmRecord {
  hipEvent_t a, b;
  double flops;
  int tag;
  char opA, opB;   // the kernel instantiation k_zgemm<opA, opB>
  bool mfma;        // false: one of the kernels for skinny shapes
  int64_t M, N, K, batch;
};
struct GemmStats {
  bool on = false;
  std::vector<GemmRecord> rec;
};
void set_gemm_stats(GemmStats* st);
// tag of the GEMM launches issued by the calling host thread
void set_gemm_tag(int tag);
// while on, the GEMM launches of the calling host thread use no split-K workspace (no
// grow-only buffer that a later reallocation could move): set around stream capture
void set_gemm_no_workspace(bool on);
// C = A A^H (m x m) from the tiles on and above the diagonal, the lower triangle mirrored;
// split: k chunks summed by a reduction (0: chosen by size, 1: unsplit)
void zherk_upper(hipStream_t s, int64_t m, int64_t k, const cplx* A, int64_t lda, cplx* C, int64_t ldc,
                 int split = 0);
// C = A B^H (m x m) for a product known to be Hermitian (B = D A, D real diagonal over k):
// the same upper tiles, reduction and mirror
void zgemm_herm(hipStream_t s, int64_t m, int64_t k, const cplx* A, int64_t lda, const cplx* B, int64_t ldb, cplx* C,
                int64_t ldc, int split = 0);
// GEMM tuning table (exact shapes): the MFMA kernel variant (1..4: K tile 16/32 x LDS padding
// 1/16) and split-K factor (1: none) chosen for (opA, opB, M, N, K, batch) by measurement.
// gemm_tune times every candidate on scratch operands on stream s and installs the fastest;
// out: {best variant, best split, best seconds, seconds of the built-in choice}.
void gemm_tune(hipStream_t s, char opA, char opB, int64_t M, int64_t N, int64_t K, int64_t batch, int reps,
               double* out);
void gemm_set_choice(char opA, char opB, int64_t M, int64_t N, int64_t K, int64_t batch, int variant, int split);
void gemm_clear_choices();
int64_t gemm_choice_count();
// free the split-K workspace of stream s (call before destroying the stream)
void release_stream_workspace(hipStream_t s);

// C[m,n] = alpha * sum_k op(A)[m,k] op(B)[k,n] + beta * C[m,n], batched by strides.
// opA: 'N' A[m + lda*k], 'T' A[k + lda*m], 'C' conj(A[k + lda*m])
// opB: 'N' B[k + ldb*n], 'R' conj(B[k + ldb*n]), 'T' B[n + ldb*k], 'C' conj(B[n + ldb*k])
void zgemm(hipStream_t s, char opA, char opB, int64_t M, int64_t N, int64_t K, cplx alpha,
           const cplx* A, int64_t lda, int64_t sA, const cplx* B, int64_t ldb, int64_t sB, cplx beta,
           cplx* C, int64_t ldc, int64_t sC, int64_t batch);

}  // namespace dc2
