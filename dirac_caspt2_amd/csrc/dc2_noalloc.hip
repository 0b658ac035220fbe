// Split-precision complex GEMM that allocates nothing (no counterpart in the reference):
// the INT8 slices and per-row exponents of the Ozaki scheme live in one fixed device array
// of this library instead of the context's grow-only scratch, for callers that cannot
// allocate device memory (code under HIP graph capture, tight serving loops).
//
// The array `scratch` has a fixed size of 64 MiB on every device the entry point runs on,
// held until the process ends.  Its device address is looked up once per device
// (hipGetSymbolAddress with that device current, on the entry point's first call there);
// host code never passes the array's name to a kernel, only that address.  All calls on
// one device share the array: calls on one context are ordered by its stream, calls that
// overlap on one device (several contexts or host threads) race.
#include "dc2_kernels.hpp"
#include "dcaspt2_amd.h"

#include <functional>
#include <mutex>

namespace dc2 {
int run_guarded(const std::function<void()>& f);   // dc2_api.hip: status and message
hipStream_t context_stream(dc2_context* c);
int context_device(dc2_context* c);
}  // namespace dc2

using namespace dc2;

namespace {

constexpr int64_t kScratchBytes = 64ll << 20;
constexpr int kMaxDevices = 64;

__device__ __align__(256) int8_t scratch[kScratchBytes];

std::mutex g_scratch_mu;
int8_t* g_scratch_ptr[kMaxDevices] = {};   // address of `scratch` on each device

// the device address of `scratch` on `device`, which is the current device
int8_t* scratch_on(int device) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  DC2_CHECK(device >= 0 && device < kMaxDevices, 1, "device ordinal out of range");
  if (g_scratch_ptr[device] == nullptr) {
    void* p = nullptr;
    DC2_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    DC2_CHECK(p != nullptr, 2, "hipGetSymbolAddress returned a null address for the fixed scratch array");
    g_scratch_ptr[device] = static_cast<int8_t*>(p);
  }
  return g_scratch_ptr[device];
}

}  // namespace

extern "C" int dc2_zgemm_split_noalloc(dc2_context* c, int32_t slices, char opB, int64_t M, int64_t N, int64_t K,
                                       const double* A, int64_t lda, const double* B, int64_t ldb, double* C,
                                       int64_t ldc) {
  return run_guarded([&] {
    DC2_CHECK(c && A && B && C, 1, "null argument");
    DC2_CHECK(M >= 0 && N >= 0 && K >= 1 && K <= 4096, 1, "dc2_zgemm_split_noalloc: need M, N >= 0 and 1 <= K <= 4096");
    DC2_CHECK(slices >= 3 && slices <= 8, 1, "dc2_zgemm_split_noalloc: slices must be in 3..8");
    DC2_CHECK(lda >= K && ldb >= K && ldc >= std::max<int64_t>(M, 1), 1, "leading dimension too small");
    const int64_t wb = (zgemm_ozaki_work_bytes(slices, M, N, K) + 255) / 256 * 256;
    const int64_t need = wb + (M + N) * (int64_t)sizeof(int);
    DC2_CHECK(need <= kScratchBytes, 1,
              "dc2_zgemm_split_noalloc: the product needs " + std::to_string(need) +
                  " bytes of scratch, more than the fixed " + std::to_string(kScratchBytes) +
                  " (slices * (M + 2N) * 2 * ceil32(K) + 4 (M + N)); use dc2_zgemm_split");
    const int dev = context_device(c);
    DC2_HIP(hipSetDevice(dev));
    int8_t* base = scratch_on(dev);
    zgemm_ozaki(context_stream(c), slices, opB, M, N, K, reinterpret_cast<const cplx*>(A), lda,
                reinterpret_cast<const cplx*>(B), ldb, reinterpret_cast<cplx*>(C), ldc, base,
                reinterpret_cast<int*>(base + wb));
  });
}
