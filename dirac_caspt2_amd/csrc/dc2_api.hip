// C ABI and host orchestration of the CASCI / CASPT2 path on one GPU.
//
// All heavy arrays live in HBM for the lifetime of the context: integrals,
// CI vectors, determinant tables, densities, transformed blocks.  The host only
// moves scalars and a few eigenvalue vectors (to size the S^-1/2 projection).
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <array>
#include <cstdio>
#include <functional>
#include <atomic>
#include <exception>
#include <future>
#include <thread>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <string>

#include "../../include/dcaspt2_amd.h"
#include "dc2_batch.hpp"
#include "dc2_hostla.hpp"
#include "dc2_kernels.hpp"

#include <functional>

using namespace dc2;

namespace {

thread_local std::string g_last_error;

#define DC2_ROCSOLVER(expr)                                                                \
  do {                                                                                     \
    rocblas_status st_ = (expr);                                                           \
    if (st_ != rocblas_status_success)                                                     \
      throw Error(3, std::string("rocSOLVER/rocBLAS status ") + std::to_string((int)st_) + \
                         " at " + __FILE__ + ":" + std::to_string(__LINE__));              \
  } while (0)

struct Block {
  DevBuf<cplx> own;
  const cplx* ptr = nullptr;
  int64_t count = 0;
  bool derived = false;   // copied out of OOOO by the library
};

const char* kSub = "ABCDEFGH";

// An execution lane: a stream, its rocBLAS handle and grow-only scratch buffers.
// Subspace solves run on several lanes at once, one host thread per lane; the
// scratch buffers persist across calls, so repeated energies allocate nothing
// once the sizes have been seen.
struct Lane {
  hipStream_t stream = nullptr;
  rocblas_handle blas = nullptr;
  bool owned = false;
  std::map<std::string, DevBuf<cplx>> zb;
  std::map<std::string, DevBuf<double>> db;
  std::map<std::string, DevBuf<int64_t>> lb;
  std::map<std::string, DevBuf<int>> ib;
  DevBuf<int> info;
  cplx* z(const char* name, int64_t n) {
    DevBuf<cplx>& b = zb[name];
    b.ensure((size_t)std::max<int64_t>(n, 1));
    return b.get();
  }
  double* d(const char* name, int64_t n) {
    DevBuf<double>& b = db[name];
    b.ensure((size_t)std::max<int64_t>(n, 1));
    return b.get();
  }
  int64_t* l(const char* name, int64_t n) {
    DevBuf<int64_t>& b = lb[name];
    b.ensure((size_t)std::max<int64_t>(n, 1));
    return b.get();
  }
  int* i(const char* name, int64_t n) {
    DevBuf<int>& b = ib[name];
    b.ensure((size_t)std::max<int64_t>(n, 1));
    return b.get();
  }
  // grow-only page-locked host memory for the lane's transfers (pageable copies go through
  // the runtime's staging buffer, which concurrent threads share): slot 0 the Davidson's
  // per-iteration arrays, slot 1 the downloads of solver results (download())
  void* pinned[2] = {nullptr, nullptr};
  size_t pinned_bytes[2] = {0, 0};
  void* host(size_t bytes, int slot = 0) {
    if (bytes > pinned_bytes[slot]) {
      if (pinned[slot]) {
        DC2_HIP(hipStreamSynchronize(stream));   // no copy still reads or writes it
        DC2_HIP(hipHostFree(pinned[slot]));
        pinned[slot] = nullptr;
        pinned_bytes[slot] = 0;
      }
      DC2_HIP(hipHostMalloc(&pinned[slot], bytes, hipHostMallocDefault));
      pinned_bytes[slot] = bytes;
    }
    return pinned[slot];
  }
  // bytes from device memory to dst on the host through slot 1; synchronises the stream
  void download(void* dst, const void* src, size_t bytes) {
    if (bytes == 0) return;
    void* h = host(bytes, 1);
    DC2_HIP(hipMemcpyAsync(h, src, bytes, hipMemcpyDeviceToHost, stream));
    DC2_HIP(hipStreamSynchronize(stream));
    std::memcpy(dst, h, bytes);
  }
  Lane() = default;
  Lane(const Lane&) = delete;
  Lane& operator=(const Lane&) = delete;
  ~Lane() {
    if (stream) (void)hipStreamSynchronize(stream);
    zb.clear();
    db.clear();
    lb.clear();
    ib.clear();
    info.release();
    for (void* h : pinned)
      if (h) (void)hipHostFree(h);
    if (owned) {
      if (blas) rocblas_destroy_handle(blas);
      if (stream) {
        release_stream_workspace(stream);
        (void)hipStreamDestroy(stream);
      }
    }
  }
};

}  // namespace

struct dc2_context {
  int device = 0;
  hipStream_t stream = nullptr;
  rocblas_handle blas = nullptr;
  int ni = -1, na = 0, ns = 0, ne = 0;
  int nsym = 1, totsym = 1;
  std::vector<int> irpamo;                 // per spinor, 1-based
  std::vector<int> multb_d, multb_ds;      // nsym x nsym
  int ras[7] = {0, 0, 0, 0, 0, 0, 0};
  double ecore = 0.0;
  DevBuf<cplx> h;
  std::map<std::string, Block> blocks;
  // CASCI
  int64_t nfull = 0, ndet = 0;
  DevBuf<int64_t> full_dets, cas_dets;
  DevBuf<int> rank_full, rank_cas, cas2full, exc;
  DevBuf<cplx> civec, cimat;               // eigenvectors (ndet x nroots_avail), the CI matrix
  std::vector<double> ecas;
  int64_t nroots_avail = 0;                 // eigenpairs held in civec / ecas
  int casci_solver = 0;                     // 0 automatic, 1 dense eigensolver, 2 block Davidson
  int davidson_iters = 0;
  bool have_casci = false;
  // CASPT2 state
  DevBuf<cplx> fock, U, hp, cnew, c_full, cF, Y1, Y1o, Y1F, G1, G1F, G2, G2F, G3, G3F;
  DevBuf<double> eps, e0, out2, partial;
  std::map<std::string, DevBuf<cplx>> tblk;
  DevBuf<int> tpairs, ipairs, spairs;
  int64_t sub_info[4 * DC2_NSUBSPACE] = {0};
  double sub_margin[DC2_NSUBSPACE] = {0};
  double timings[12] = {0};
  bool have_pt2 = false;
  // scratch
  DevBuf<cplx> w1, w2, w3, w4;
  DevBuf<double> dw1, dw2;
  DevBuf<int> info;
  bool debug = false;
  // eigensolver: -1 automatic (divide and conquer; D&C with Jacobi leaves for the
  // rank-deficient nact^3 S matrices), else 0 QR (zheev), 1 D&C, 2 D&C + Jacobi, 3 Jacobi
  int eig_method = -1;
  int lane_plan = -1;        // 4 lanes: 0 D alone on lane 3, 1 D with B and F (their S/B solves
                             // run while the transforms do), -1 1 from 256 secondary spinors
  bool no_thin = false;      // option "c_thin" 0: C's metric always as the m x m Gram matrix
  bool early_v = true;       // option "early_v": V of A and C on lane 3 once their blocks exist
  int ac_route = 0;          // A/C metric: 0 automatic, 1 RDM formulas (nact^3), 2 determinant space
  int nlanes = 4;            // concurrent subspace lanes (1 = sequential on the library stream)
  bool schur = true;         // E2 of the determinant-space route through the compressed resolvent
  int split_slices = 0;      // > 0: transform passes with K >= split_min_k as Ozaki INT8 GEMMs with this many slices
  int split_min_k = 32;
  double split_flops = 0.0, fp64_flops = 0.0;   // transform flops of the last dc2_caspt2 by precision
  std::atomic<int64_t> eig_fallbacks{0};
  dc2::EigBatcher* batcher = nullptr;          // set by a request server: eigensolves batched across requests
  std::unique_ptr<Lane> lane0;                  // the library stream as a lane
  std::vector<std::unique_ptr<Lane>> lanes;     // worker lanes (created on first use)
  GemmStats gstats;        // GEMM launch accounting (dc2_set_profiling)
  // block Davidson: the launches between two host round trips of an iteration replayed from
  // captured graphs, keyed by the sizes and every buffer address they use
  bool graphs = true;
  // eigenproblems with n >= native_min (and n within this library's eigensolver) take it even
  // when the eigensolves are batched across requests (0: never)
  int native_min = 0;
  int dav_maxsub = 0;   // block-Davidson subspace before a thick restart (0: max(6 b, 24))
  int dav_block = 0;    // block-Davidson block size (0: max(nroot + 3, 4))
  int native_max = 2048;   // largest eigenproblem this library's eigensolver takes
  // the tridiagonal stage of this library's eigensolver: its own divide and conquer (true) or
  // rocSOLVER dstedc
  bool tri_own = true;
  std::map<std::vector<int64_t>, hipGraphExec_t> dav_graphs;
  int64_t graph_launches = 0, graph_captures = 0;

  // called from the lane that finished subspace k (0 = A .. 7 = H) with its e2 and sumc2
  void (*on_subspace)(void* user, int32_t k, double e2, double sumc2) = nullptr;
  void* on_subspace_user = nullptr;

  int nmo() const { return ni + na + ns; }
  int nmoc() const { return ni + na; }
};

namespace {

template <class F>
int guarded(F&& f) {
  try {
    f();
    return 0;
  } catch (const Error& e) {
    g_last_error = e.what();
    return e.code;
  } catch (const std::exception& e) {
    g_last_error = e.what();
    return 1;
  }
}

struct BlockShape {
  const char* name;
  const char* sp;
};
const BlockShape kShapes[] = {{"OOOO", "oooo"}, {"F1", "ssoo"}, {"F2", "soos"}, {"A1", "aiaa"}, {"A2", "aiii"},
                              {"B", "aiai"},    {"C1", "saaa"}, {"C2", "saii"}, {"C3", "siia"}, {"D1", "siaa"},
                              {"D2", "saai"},   {"D3", "siii"}, {"E", "siai"},  {"F", "sasa"},  {"G", "sisa"},
                              {"H", "sisi"}};

const char* shape_of(const std::string& name) {
  for (const auto& s : kShapes)
    if (name == s.name) return s.sp;
  throw Error(1, "unknown integral block name '" + name + "'");
}

int space_dim(const dc2_context* c, char s) {
  switch (s) {
    case 'i': return c->ni;
    case 'a': return c->na;
    case 's': return c->ns;
    default: return c->ni + c->na;
  }
}
int space_off(const dc2_context* c, char s) {
  switch (s) {
    case 'i': return 0;
    case 'a': return c->ni;
    case 's': return c->ni + c->na;
    default: return 0;
  }
}
int64_t block_count(const dc2_context* c, const char* sp) {
  int64_t n = 1;
  for (int k = 0; k < 4; ++k) n *= space_dim(c, sp[k]);
  return n;
}

void need_spaces(const dc2_context* c) { DC2_CHECK(c->ni >= 0, 1, "dc2_set_spaces has not been called"); }

// pointer to the original-basis block; A1/A2/B fall back to a copy out of OOOO
const cplx* get_block(dc2_context* c, const std::string& name) {
  auto it = c->blocks.find(name);
  if (it != c->blocks.end()) return it->second.ptr;
  if (name == "A1" || name == "A2" || name == "B") {
    auto o = c->blocks.find("OOOO");
    DC2_CHECK(o != c->blocks.end(), 1, "integral block OOOO is required");
    const char* sp = shape_of(name);
    Block& b = c->blocks[name];
    b.derived = true;
    b.count = block_count(c, sp);
    b.own.alloc(b.count);
    const int64_t m = c->nmoc();
    int64_t ld[3] = {m, m, m};
    int off[4], n[4];
    for (int k = 0; k < 4; ++k) { off[k] = space_off(c, sp[k]); n[k] = space_dim(c, sp[k]); }
    extract4(c->stream, o->second.ptr, ld, off, n, b.own.get());
    b.ptr = b.own.get();
    return b.ptr;
  }
  throw Error(1, "integral block " + name + " has not been set");
}

// copy ordered on the context's stream, then waited for.  Never the legacy stream: another
// context's thread may be capturing a graph on a blocking stream, and a legacy-stream copy
// would have to wait for it, which the runtime refuses.
void copy_sync(dc2_context* c, void* dst, const void* src, size_t bytes, hipMemcpyKind kind) {
  if (bytes == 0) return;
  DC2_HIP(hipMemcpyAsync(dst, src, bytes, kind, c->stream));
  DC2_HIP(hipStreamSynchronize(c->stream));
}

// diagnostics of a matrix handed to an eigensolver that failed
std::string matrix_report(dc2_context* c, int n, const cplx* A) {
  std::vector<cplx> h((size_t)n * n);
  copy_sync(c, h.data(), A, h.size() * sizeof(cplx), hipMemcpyDeviceToHost);
  int64_t nonfinite = 0;
  double herm = 0.0, amax = 0.0, dimag = 0.0;
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i) {
      const cplx a = h[i + (size_t)n * j], b = h[j + (size_t)n * i];
      if (!std::isfinite(a.x) || !std::isfinite(a.y)) ++nonfinite;
      amax = std::max(amax, std::hypot(a.x, a.y));
      herm = std::max(herm, std::hypot(a.x - b.x, a.y + b.y));
      if (i == j) dimag = std::max(dimag, std::fabs(a.y));
    }
  return " (n=" + std::to_string(n) + ", non-finite=" + std::to_string(nonfinite) + ", max|a|=" + std::to_string(amax) +
         ", max|a-a^H|=" + std::to_string(herm) + ", max|Im diag|=" + std::to_string(dimag) + ")";
}

// debug mode (DC2_DEBUG=1): fail at the first device array holding a non-finite value
template <class T>
void check_finite(dc2_context* c, hipStream_t s, const char* what, const T* p, int64_t count) {
  if (!c->debug || count <= 0 || p == nullptr) return;
  std::vector<T> h(count);
  DC2_HIP(hipMemcpyAsync(h.data(), p, count * sizeof(T), hipMemcpyDeviceToHost, s));
  DC2_HIP(hipStreamSynchronize(s));
  const double* d = reinterpret_cast<const double*>(h.data());
  const int64_t nd = count * (int64_t)(sizeof(T) / sizeof(double));
  for (int64_t k = 0; k < nd; ++k)
    if (!std::isfinite(d[k]))
      throw Error(4, std::string("non-finite value in ") + what + " at double index " + std::to_string(k) + " of " +
                         std::to_string(nd));
}
template <class T>
void check_finite(dc2_context* c, const char* what, const T* p, int64_t count) {
  check_finite(c, c->stream, what, p, count);
}

// every element of the eigensolver output finite: checked on the device (a 792 x 792
// result copied to pageable host memory for the scan took ~3 ms)
bool all_finite(Lane& L, const cplx* A, int64_t count) {
  return all_finite_dev(L.stream, reinterpret_cast<const double*>(A), 2 * count, L.i("finite_flag", 1),
                        reinterpret_cast<int*>(L.host(sizeof(int), 1)));
}

int run_eig(Lane& L, int method, int n, cplx* A, double* w) {
  auto* a = reinterpret_cast<rocblas_double_complex*>(A);
  double* e = L.d("eig_e", n);
  L.info.ensure(2);
  switch (method) {
    case 1:
      DC2_ROCSOLVER(rocsolver_zheevd(L.blas, rocblas_evect_original, rocblas_fill_upper, n, a, n, w, e,
                                     L.info.get()));
      break;
    case 2:
      DC2_ROCSOLVER(rocsolver_zheevdj(L.blas, rocblas_evect_original, rocblas_fill_upper, n, a, n, w, L.info.get()));
      break;
    case 3:
      DC2_ROCSOLVER(rocsolver_zheevj(L.blas, rocblas_esort_ascending, rocblas_evect_original, rocblas_fill_upper, n, a,
                                     n, 0.0, e, 100, L.info.get() + 1, w, L.info.get()));
      break;
    default:
      // Tridiagonal QR, the ZHEEV algorithm of diag.F90:109
      DC2_ROCSOLVER(rocsolver_zheev(L.blas, rocblas_evect_original, rocblas_fill_upper, n, a, n, w, e,
                                    L.info.get()));
  }
  int info = 0;
  L.download(&info, L.info.get(), sizeof(int));
  return info;
}

enum class EigKind { general, rank_deficient };

// Hermitian eigendecomposition (upper triangle), ascending eigenvalues, on lane L.
// The configured rocSOLVER variant runs first; a failure or a non-finite result (the
// divide-and-conquer variants returned NaN eigenvectors with info = 0 for
// rank-deficient S matrices with hundreds of zero eigenvalues) falls back to the
// tridiagonal QR algorithm on a saved copy.
// Eigenpairs of the Hermitian matrix whose upper triangle A (n x n, device) holds, as the
// device solvers read it, on the host (herm_eig): eigenvectors overwrite A, ascending
// eigenvalues go to w.  Synchronises the stream.
void host_heev_upper(Lane& L, int n, cplx* A, double* w) {
  const hipStream_t s = L.stream;
  const size_t nn = (size_t)n * n;
  std::vector<zc> hA(nn);
  L.download(hA.data(), A, nn * sizeof(cplx));
  for (int j = 0; j < n; ++j) {
    hA[j + (size_t)n * j] = hA[j + (size_t)n * j].real();
    for (int i = 0; i < j; ++i) hA[j + (size_t)n * i] = std::conj(hA[i + (size_t)n * j]);
  }
  // eigenvectors and eigenvalues written straight into page-locked memory, uploaded from it
  zc* hV = reinterpret_cast<zc*>(L.host(nn * sizeof(zc) + n * sizeof(double), 1));
  double* hw = reinterpret_cast<double*>(hV + nn);
  herm_eig(n, hA.data(), hw, hV);
  DC2_HIP(hipMemcpyAsync(A, hV, nn * sizeof(cplx), hipMemcpyHostToDevice, s));
  DC2_HIP(hipMemcpyAsync(w, hw, n * sizeof(double), hipMemcpyHostToDevice, s));
  DC2_HIP(hipStreamSynchronize(s));   // the host arrays are the sources of the copies
}

// matrices up to this size are diagonalised on the host by the automatic eigensolver choice:
// a device solve is a chain of dependent launches (about 0.3-0.5 ms at any size), the host
// solve of a 32 x 32 matrix takes about 0.1 ms
constexpr int kHostEigMax = 32;

void zheevd(dc2_context* c, Lane& L, int n, cplx* A, double* w, const char* what = "",
            EigKind kind = EigKind::general) {
  if (n <= 0) return;
  if (c->eig_method < 0 && n <= kHostEigMax) {
    host_heev_upper(L, n, A, w);
    check_finite(c, L.stream, what, A, (int64_t)n * n);
    return;
  }
  const int64_t nn = (int64_t)n * n;
  const bool native_ok = n >= 3 && n <= std::min(2048, c->native_max);
  const bool batched = c->batcher && !(c->native_min > 0 && n >= c->native_min);
  int method = c->eig_method >= 0 ? c->eig_method
                                  : (kind == EigKind::rank_deficient ? 2 : (native_ok && !batched ? 4 : 1));
  if (method == 4 && !native_ok) method = 1;
  cplx* keep = L.z("eig_keep", nn);
  DC2_HIP(hipMemcpyAsync(keep, A, nn * sizeof(cplx), hipMemcpyDeviceToDevice, L.stream));
  int info = 0;
  if (method == 4) {
    // tridiagonalisation and tridiagonal divide and conquer of this library (or rocSOLVER
    // dstedc); rocSOLVER's D&C if it cannot run or its result is not usable
    void* work = L.d("eig_native", heev_native_work_bytes(n) / 8 + 1);
    L.info.ensure(2);
    // on a worker lane other lanes' kernels run beside the reduction: it takes fewer units
    const bool narrow = &L != c->lane0.get();
    bool done = heev_native(L.stream, L.blas, n, A, n, w, work, L.info.get(),
                            reinterpret_cast<int*>(L.host(2 * sizeof(int), 1)), c->tri_own, c->debug, narrow);
    bool ran = done, finite = false;
    if (done) {
      L.download(&info, L.info.get(), sizeof(int));
      finite = all_finite(L, A, nn);
      done = info == 0 && finite;
    }
    if (!done && c->debug)
      std::fprintf(stderr, "dc2: native eigensolver rejected for %s (n=%d): %s, info %d, finite %d\n", what, n,
                   ran ? "ran" : "did not run", info, (int)finite);
    if (!done) {
      ++c->eig_fallbacks;
      DC2_HIP(hipMemcpyAsync(A, keep, nn * sizeof(cplx), hipMemcpyDeviceToDevice, L.stream));
      method = 1;
    }
  }
  if (method == 4) {
  } else if (c->batcher && (method == 1 || method == 2)) {
    DC2_HIP(hipStreamSynchronize(L.stream));   // the matrix is complete before another stream reads it
    info = c->batcher->solve(n, method, A, w);
  } else {
    info = run_eig(L, method, n, A, w);
  }
  // fallbacks: divide and conquer -> D&C with Jacobi leaves -> tridiagonal QR
  for (int next : {2, 0}) {
    if (method == 0 || method == 4 || (info == 0 && all_finite(L, A, nn))) break;
    if (next == 2 && method != 1) continue;   // only plain D&C falls back to D&C with Jacobi leaves
    ++c->eig_fallbacks;
    DC2_HIP(hipMemcpyAsync(A, keep, nn * sizeof(cplx), hipMemcpyDeviceToDevice, L.stream));
    info = run_eig(L, next, n, A, w);
    method = next;
  }
  check_finite(c, L.stream, what, A, nn);
  if (info != 0)
    throw Error(100 + info, std::string("Hermitian eigensolver failed for ") + what + ", info = " + std::to_string(info) +
                                (c->debug ? matrix_report(c, n, keep) : std::string(" (n=") + std::to_string(n) + ")"));
}

void zheevd(dc2_context* c, int n, cplx* A, double* w, const char* what = "") { zheevd(c, *c->lane0, n, A, w, what); }

struct Timer {
  hipEvent_t a, b;
  hipStream_t s;
  explicit Timer(hipStream_t st) : s(st) {
    DC2_HIP(hipEventCreate(&a));
    DC2_HIP(hipEventCreate(&b));
    DC2_HIP(hipEventRecord(a, s));
  }
  double stop() {
    DC2_HIP(hipEventRecord(b, s));
    DC2_HIP(hipEventSynchronize(b));
    float ms = 0;
    DC2_HIP(hipEventElapsedTime(&ms, a, b));
    return ms * 1e-3;
  }
  ~Timer() {
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
  }
};

std::vector<int> pair_list(int n) {
  std::vector<int> v;
  for (int t = 0; t < n; ++t)
    for (int u = 0; u < t; ++u) { v.push_back(t); v.push_back(u); }
  return v;
}

void upload_ints(hipStream_t s, DevBuf<int>& dst, const std::vector<int>& v) {
  dst.ensure(std::max<size_t>(v.size(), 1));
  if (!v.empty()) DC2_HIP(hipMemcpyAsync(dst.get(), v.data(), v.size() * sizeof(int), hipMemcpyHostToDevice, s));
}

// ---------------------------------------------------------------------------
// Lowest nroot eigenpairs of the Hermitian CI matrix H (n x n, device) by block Davidson with
// the diagonal preconditioner: block of b >= nroot + 3 vectors (Kramers-degenerate pairs stay
// inside the block), subspace of at most maxsub vectors (thick restart on the Ritz vectors),
// converged when every requested residual |H x - theta x| is below 1e-11.  The Rayleigh
// matrix V^H H V is kept on the host and extended by the new block only (after a restart it
// is diag(theta)); its eigenproblem runs on the host (herm_eig), the orthonormalisation of a
// block of four on the device (a second round trip to the host for larger blocks), so that
// an iteration has one host round trip.  Writes c->civec (n x nroot) and
// c->ecas; returns false (no result) when it does not converge, and the caller falls back to
// the dense eigensolver.
bool davidson(dc2_context* c, int64_t n, const cplx* H, int nroot) {
  Lane& L = *c->lane0;
  hipStream_t s = L.stream;
  const cplx one = mk(1.0, 0.0), zero = mk(0.0, 0.0), mone = mk(-1.0, 0.0);
  // b = 4, maxsub = 24 for one root: measured fastest on the 924-determinant CASCI (20
  // iterations, against 19 with 6 / 48, whose larger host eigenproblems cost more); from 1500
  // determinants 12 b: on 3003 determinants 48 iterations and 13.8 ms against 107 and 22.6 ms
  // with 24 (96: 38 iterations, 22.6 ms)
  const int b = (int)std::min<int64_t>(n, std::max(nroot + 3, c->dav_block > 0 ? c->dav_block : 4));
  const int auto_sub = n >= 1500 ? 12 * b : std::max(6 * b, 24);
  const int maxsub = (int)std::min<int64_t>(n, c->dav_maxsub > 0 ? std::max(c->dav_maxsub, 2 * b) : auto_sub);
  if (maxsub < 2 * b) return false;
  cplx* V = L.z("dv_V", n * maxsub);
  cplx* W = L.z("dv_W", n * maxsub);
  cplx* X = L.z("dv_X", n * b);
  cplx* AX = L.z("dv_AX", n * b);
  cplx* R = L.z("dv_R", n * b);
  cplx* T = L.z("dv_T", n * b);
  cplx* Sm = L.z("dv_small", (int64_t)maxsub * b);       // Gram-Schmidt coefficients
  cplx* Sy = L.z("dv_y", (int64_t)maxsub * b + b);       // Ritz coefficients, then theta
  cplx* Sc = L.z("dv_sc", (int64_t)b * b);
  double* d = L.d("dv_d", n);
  int* idx = L.i("dv_idx", b);
  // what an iteration downloads, contiguous so that it is one copy: R^H R and T^H T (RG), the
  // kept count of the next block (one element's room), V^H W_new (Cn, m + b rows)
  const size_t nCh = (size_t)maxsub * b, nrg = 2 * (size_t)b * b, ny = (size_t)maxsub * b + b, nsc = (size_t)b * b;
  cplx* const DL = L.z("dv_dl", (int64_t)(nrg + 1 + nCh));
  cplx* const RG = DL;
  int* const nk_dev = reinterpret_cast<int*>(DL + nrg);
  cplx* const Cn = DL + nrg + 1;
  // page-locked host side: the download's mirror (rg, nk_host, Ch); uploads up_y (Ritz
  // coefficients, then theta at maxsub * b: one copy), up_sc, up_idx, whose contents stay
  // untouched until the next synchronising download
  zc* const rg = reinterpret_cast<zc*>(L.host((nrg + 1 + nCh + ny + nsc) * sizeof(zc) + b * sizeof(int)));
  int* const nk_host = reinterpret_cast<int*>(rg + nrg);
  zc* const Ch = rg + nrg + 1;
  zc* const up_y = Ch + nCh;
  zc* const up_sc = up_y + ny;
  int* const up_idx = reinterpret_cast<int*>(up_sc + nsc);
  auto to_host = [&](void* dst, const void* src, size_t bytes) {
    DC2_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s));
    DC2_HIP(hipStreamSynchronize(s));
  };
  auto to_dev = [&](void* dst, const void* src, size_t bytes) {
    DC2_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s));
  };
  // start: unit vectors at the b smallest diagonal elements
  davidson_diag(s, n, H, d);
  std::vector<double> dh(n);
  to_host(dh.data(), d, n * sizeof(double));
  std::vector<int> order(n);
  std::iota(order.begin(), order.end(), 0);
  std::partial_sort(order.begin(), order.begin() + b, order.end(), [&](int x, int y) { return dh[x] < dh[y]; });
  std::copy(order.begin(), order.begin() + b, up_idx);
  to_dev(idx, up_idx, b * sizeof(int));
  davidson_unit(s, n, b, idx, V);
  zgemm(s, 'N', 'N', n, b, n, one, H, n, 0, V, n, 0, zero, W, n, 0, 1);
  std::vector<int> seq(b);
  std::iota(seq.begin(), seq.end(), 0);
  int* sel = L.i("dv_sel", b);
  to_dev(sel, seq.data(), b * sizeof(int));
  int m = b, nk = b;   // subspace size; the last nk columns are new
  int ldc = b;         // rows of the downloaded V^H W_new block (Ch)
  const double tol = 1e-11;
  std::vector<zc> Th((size_t)maxsub * maxsub), Tm, Yh, G, Q;
  std::vector<double> theta, gw;
  double* th = reinterpret_cast<double*>(Sy + (size_t)maxsub * b);
  // the block orthonormalisation on the device for blocks of four (one root), so that an
  // iteration needs one host round trip; the host's, in a round trip of its own, otherwise
  const bool dev_orth = b == 4;
  // the first block's Rayleigh columns
  zgemm(s, 'C', 'N', m, nk, n, one, V, n, 0, W, n, 0, zero, Cn, ldc, 0, 1);
  to_host(Ch, Cn, (size_t)ldc * nk * sizeof(cplx));
  for (int it = 1; it <= 200; ++it) {
    c->davidson_iters = it;
    // Rayleigh matrix: the new columns V^H W[:, m-nk:m], downloaded with ldc rows
    for (int j = 0; j < nk; ++j)
      for (int i = 0; i < m; ++i) {
        const zc x = Ch[i + (size_t)ldc * j];
        const int jj = m - nk + j;
        if (i < m - nk) {
          Th[i + (size_t)maxsub * jj] = x;
          Th[jj + (size_t)maxsub * i] = std::conj(x);
        } else if (i <= jj) {   // the new diagonal block: its upper triangle, mirrored
          const zc y = (i == jj) ? zc(x.real(), 0.0) : 0.5 * (x + std::conj(Ch[jj + (size_t)ldc * (i - (m - nk))]));
          Th[i + (size_t)maxsub * jj] = y;
          Th[jj + (size_t)maxsub * i] = std::conj(y);
        }
      }
    Tm.resize((size_t)m * m);
    for (int j = 0; j < m; ++j)
      for (int i = 0; i < m; ++i) Tm[i + (size_t)m * j] = Th[i + (size_t)maxsub * j];
    theta.resize(m);
    Yh.resize((size_t)m * b);
    herm_eig_low(m, Tm.data(), theta.data(), Yh.data(), b);   // every theta, the lowest b vectors
    std::copy(Yh.begin(), Yh.begin() + (size_t)m * b, up_y);     // first b Ritz vectors
    std::memcpy(up_y + (size_t)maxsub * b, theta.data(), b * sizeof(double));
    to_dev(Sy, up_y, (size_t)maxsub * b * sizeof(cplx) + b * sizeof(double));   // Sy, then theta
    const bool restart = m + b > maxsub;
    const int m_in = m;
    // Ritz vectors and residuals, preconditioned corrections (one launch up to 64 subspace
    // vectors), (restart,) two classical Gram-Schmidt passes against V, the Gram matrix of the
    // block and (device_orth) its orthonormal combinations: launches only
    auto segment = [&](int mm, bool rs) {
      if (mm <= 64) {   // sel is the identity
        davidson_ritz(s, n, b, mm, V, W, Sy, th, d, X, AX, R, T);
      } else {
        zgemm(s, 'N', 'N', n, b, mm, one, V, n, 0, Sy, mm, 0, zero, X, n, 0, 1);
        zgemm(s, 'N', 'N', n, b, mm, one, W, n, 0, Sy, mm, 0, zero, AX, n, 0, 1);
        davidson_residual(s, n, b, AX, X, th, R);
        davidson_precond(s, n, b, sel, R, th, d, T);
      }
      zgemm(s, 'C', 'N', b, b, n, one, R, n, 0, R, n, 0, zero, RG, b, 0, 1);   // residual norms
      davidson_normalize(s, n, b, T);
      if (rs) {   // thick restart on the b Ritz vectors: V^H H V = diag(theta)
        DC2_HIP(hipMemcpyAsync(V, X, n * b * sizeof(cplx), hipMemcpyDeviceToDevice, s));
        DC2_HIP(hipMemcpyAsync(W, AX, n * b * sizeof(cplx), hipMemcpyDeviceToDevice, s));
        mm = b;
      }
      for (int pass = 0; pass < 2; ++pass) {
        zgemm(s, 'C', 'N', mm, b, n, one, V, n, 0, T, n, 0, zero, Sm, mm, 0, 1);
        zgemm(s, 'N', 'N', n, b, mm, mone, V, n, 0, Sm, mm, 0, one, T, n, 0, 1);
      }
      zgemm(s, 'C', 'N', b, b, n, one, T, n, 0, T, n, 0, zero, RG + (size_t)b * b, b, 0, 1);
      if (dev_orth) davidson_block_orth(s, RG + (size_t)b * b, Sc, nk_dev);
    };
    // Replayed from a captured graph when no restart copy is involved and the launches are
    // not being timed; the captured products run without split-K (their workspace can be
    // reallocated later), and graphs are cached by sizes and buffer addresses
    const bool graph_ok = c->graphs && !c->gstats.on && !restart && n * b <= 65536 && (int64_t)m_in * b <= 2048;
    if (graph_ok) {
      const std::vector<int64_t> key = {n, b, m_in, (int64_t)V, (int64_t)W, (int64_t)X, (int64_t)AX, (int64_t)R,
                                        (int64_t)T, (int64_t)Sy, (int64_t)th, (int64_t)RG, (int64_t)Sm, (int64_t)d,
                                        (int64_t)sel, (int64_t)s, (int64_t)Sc, (int64_t)nk_dev, (int64_t)dev_orth};
      auto it = c->dav_graphs.find(key);
      if (it == c->dav_graphs.end()) {   // capture; on any failure the context stops using graphs
        hipGraph_t g = nullptr;
        hipGraphExec_t ge = nullptr;
        bool ok = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal) == hipSuccess;
        if (ok) {
          set_gemm_no_workspace(true);   // the graph must not hold the grow-only split-K buffer
          try {
            segment(m_in, false);
          } catch (...) {
            ok = false;
          }
          set_gemm_no_workspace(false);
          ok = hipStreamEndCapture(s, &g) == hipSuccess && ok && g != nullptr;
          ok = ok && hipGraphInstantiateWithFlags(&ge, g, 0) == hipSuccess;
          if (g) (void)hipGraphDestroy(g);
        }
        if (ok) {
          it = c->dav_graphs.emplace(key, ge).first;
          ++c->graph_captures;
        } else {
          (void)hipGetLastError();
          c->graphs = false;
        }
      }
      if (it != c->dav_graphs.end()) {
        DC2_HIP(hipGraphLaunch(it->second, s));
        ++c->graph_launches;
      } else {
        segment(m_in, false);
      }
    } else {
      segment(m_in, restart);
    }
    if (restart) m = b;
    int kept_host = 0;
    if (!dev_orth) {   // the host orthonormalises the block: a round trip of its own
      to_host(rg, RG, nrg * sizeof(cplx));
      G.assign(rg + (size_t)b * b, rg + nrg);
      Q.resize(G.size());
      gw.resize(b);
      herm_eig(b, G.data(), gw.data(), Q.data());
      int kept = 0;
      for (int k = b - 1; k >= 0; --k) {
        if (!(gw[k] > 1e-12)) continue;   // |projected unit vector|^2
        for (int i = 0; i < b; ++i) up_sc[i + (size_t)b * kept] = Q[i + (size_t)b * k] / std::sqrt(gw[k]);
        ++kept;
      }
      for (int k = kept; k < b; ++k)
        for (int i = 0; i < b; ++i) up_sc[i + (size_t)b * k] = zc(0.0, 0.0);
      kept_host = kept;
      to_dev(Sc, up_sc, (size_t)b * b * sizeof(cplx));
    }
    // the next block (b columns, the dropped ones zero) and its Rayleigh columns, then one
    // download of the residual norms, the kept count and V^H W_new; m + b <= maxsub here
    zgemm(s, 'N', 'N', n, b, b, one, T, n, 0, Sc, b, 0, zero, V + n * m, n, 0, 1);
    zgemm(s, 'N', 'N', n, b, n, one, H, n, 0, V + n * m, n, 0, zero, W + n * m, n, 0, 1);
    zgemm(s, 'C', 'N', m + b, b, n, one, V, n, 0, W + n * m, n, 0, zero, Cn, m + b, 0, 1);
    to_host(rg, DL, (nrg + 1 + (size_t)(m + b) * b) * sizeof(cplx));
    double worst = 0.0;
    for (int j = 0; j < nroot; ++j) worst = std::max(worst, std::sqrt(std::fabs(rg[j + (size_t)b * j].real())));
    if (worst < tol) {
      c->civec.ensure(n * nroot);
      DC2_HIP(hipMemcpyAsync(c->civec.get(), X, n * nroot * sizeof(cplx), hipMemcpyDeviceToDevice, s));
      DC2_HIP(hipStreamSynchronize(s));
      c->ecas.assign(theta.begin(), theta.begin() + nroot);
      c->nroots_avail = nroot;
      return true;
    }
    if (restart) {
      for (int j = 0; j < b; ++j)
        for (int i = 0; i < b; ++i) Th[i + (size_t)maxsub * j] = (i == j) ? zc(theta[j], 0.0) : zc(0.0, 0.0);
    }
    nk = dev_orth ? *nk_host : kept_host;
    if (nk <= 0) return false;
    ldc = m + b;
    m += nk;
  }
  return false;
}

// the CASCI determinants (search_cas_configuration.F90) and the full string space of the
// densities: c->ndet, cas_dets, full_dets, rank tables, cas2full
void enumerate_space(dc2_context* c) {
  need_spaces(c);
  hipStream_t s = c->stream;
  const int na = c->na;
  const int64_t N = 1ll << na;
  // scratch from the library lane's grow-only buffers: a hipFree would synchronise every
  // stream of the device, including other contexts' work
  Lane& L0 = *c->lane0;
  int* ff = L0.i("ci_ff", N);
  int* fc = L0.i("ci_fc", N);
  int* counts = L0.i("ci_counts", 2);
  int* irp = L0.i("ci_irp", std::max(na, 1));
  int* md = L0.i("ci_md", std::max<int64_t>((int64_t)c->multb_d.size(), 1));
  int* mds = L0.i("ci_mds", std::max<int64_t>((int64_t)c->multb_ds.size(), 1));
  std::vector<int> irp_act(std::max(na, 1), 1);
  for (int w = 0; w < na; ++w) irp_act[w] = c->irpamo.empty() ? 1 : c->irpamo[c->ni + w];
  DC2_HIP(hipMemcpyAsync(irp, irp_act.data(), irp_act.size() * sizeof(int), hipMemcpyHostToDevice, s));
  if (!c->multb_d.empty()) {
    DC2_HIP(hipMemcpyAsync(md, c->multb_d.data(), c->multb_d.size() * sizeof(int), hipMemcpyHostToDevice, s));
    DC2_HIP(hipMemcpyAsync(mds, c->multb_ds.data(), c->multb_ds.size() * sizeof(int), hipMemcpyHostToDevice, s));
  }
  DetParams p{na, c->ne, c->ni, c->ras[0], c->ras[1], c->ras[2], c->ras[3], c->ras[4], c->ras[5], c->ras[6],
              c->nsym, c->totsym, irp, md, mds};
  c->rank_full.ensure(N);
  c->rank_cas.ensure(N);
  int64_t nf = 0, nd = 0;
  enumerate_dets(s, p, ff, fc, c->rank_full.get(), c->rank_cas.get(), nullptr, nullptr, nullptr, &nf, &nd, counts);
  DC2_CHECK(nd > 0, 1, "the number of CASCI configurations is 0 (check totsym)");
  c->nfull = nf;
  c->ndet = nd;
  c->full_dets.ensure(nf);
  c->cas_dets.ensure(nd);
  c->cas2full.ensure(nd);
  enumerate_dets(s, p, ff, fc, c->rank_full.get(), c->rank_cas.get(), c->full_dets.get(), c->cas_dets.get(),
                 c->cas2full.get(), &nf, &nd, counts);
}

void run_casci(dc2_context* c, int nroot) {
  need_spaces(c);
  DC2_CHECK(c->h.get() != nullptr, 1, "dc2_set_one_electron has not been called");
  const cplx* O = get_block(c, "OOOO");
  hipStream_t s = c->stream;
  Timer tm(s);
  const int na = c->na;
  Lane& L0 = *c->lane0;
  enumerate_space(c);
  const int64_t nd = c->ndet;
  c->cimat.ensure(nd * nd);
  casmat(s, c->ni, na, c->ne, nd, c->cas_dets.get(), c->rank_cas.get(), O, c->nmoc(), c->h.get(), c->nmo(),
         c->cimat.get());
  const bool dav = c->casci_solver == 2 || (c->casci_solver == 0 && nd >= 300 && nroot <= 16);
  c->davidson_iters = 0;
  // odd electron counts: one more root, so that the Kramers partner of the last requested
  // root is held (root_vector needs the whole degenerate group)
  const int nwant = (int)std::min<int64_t>(std::max(nroot, 1) + (c->ne % 2), nd);
  if (!dav || !davidson(c, nd, c->cimat.get(), nwant)) {
    c->civec.ensure(nd * nd);
    DC2_HIP(hipMemcpyAsync(c->civec.get(), c->cimat.get(), nd * nd * sizeof(cplx), hipMemcpyDeviceToDevice, s));
    double* w = L0.d("ci_w", nd);
    zheevd(c, (int)nd, c->civec.get(), w, "the CASCI matrix");
    c->ecas.resize(nd);
    DC2_HIP(hipMemcpyAsync(c->ecas.data(), w, nd * sizeof(double), hipMemcpyDeviceToHost, s));
    DC2_HIP(hipStreamSynchronize(s));
    c->nroots_avail = nd;
  }
  c->have_casci = true;
  c->have_pt2 = false;
  c->timings[0] = tm.stop();
}

// four one-index passes, each contracting the leading index and rotating it to the end
void transform(dc2_context* c, const std::string& name) {
  const char* sp = shape_of(name);
  const cplx* src = get_block(c, name);
  const int64_t cnt = block_count(c, sp);
  DevBuf<cplx>& out = c->tblk[name];
  out.ensure(cnt);
  c->w1.ensure(cnt);
  int64_t dims[4];
  for (int k = 0; k < 4; ++k) dims[k] = space_dim(c, sp[k]);
  const cplx one = mk(1.0, 0.0), zero = mk(0.0, 0.0);
  const int nmo = c->nmo();
  const cplx* in = src;
  for (int pass = 0; pass < 4; ++pass) {
    const int64_t n = dims[0], rest = dims[1] * dims[2] * dims[3];
    cplx* dst = (pass % 2 == 0) ? c->w1.get() : out.get();
    const int off = space_off(c, sp[pass]);
    const cplx* Ub = c->U.get() + off + (int64_t)nmo * off;
    // (i'j'|k'l') = sum conj(U_ii') U_jj' conj(U_kk') U_ll' (ij|kl): conj on the 1st and 3rd index
    const char opB = (pass % 2 == 0) ? 'R' : 'N';
    if (c->split_slices > 0 && n >= c->split_min_k) {
      Lane& L0 = *c->lane0;
      const int64_t wb = zgemm_ozaki_work_bytes(c->split_slices, rest, n, n);
      int8_t* work = reinterpret_cast<int8_t*>(L0.l("oz_work", (wb + 7) / 8));
      int* expo = L0.i("oz_expo", rest + n);
      zgemm_ozaki(c->stream, c->split_slices, opB, rest, n, n, in, n, Ub, nmo, dst, rest, work, expo);
      c->split_flops += 8.0 * rest * n * n;
    } else {
      zgemm(c->stream, 'T', opB, rest, n, n, one, in, n, 0, Ub, nmo, 0, zero, dst, rest, 0, 1);
      c->fp64_flops += 8.0 * rest * n * n;
    }
    in = dst;
    const int64_t d0 = dims[0];
    dims[0] = dims[1]; dims[1] = dims[2]; dims[2] = dims[3]; dims[3] = d0;
  }
}

struct SolveOut {
  double e2 = 0, sumc2 = 0, margin = INFINITY;
  int64_t dimm = 0;
};

// e2 and sumc2 from the two-double result of a reduction on lane L
void read_out2(Lane& L, const double* out2, SolveOut& r) {
  double o[2];
  L.download(o, out2, sizeof(o));
  r.e2 = o[0];
  r.sumc2 = o[1];
}

// cdiag(B~) -> X = W^H Y -> E2 / sumc2 (solve_A_subspace.F90:197-257); bc1 is overwritten
void finish_block(dc2_context* c, Lane& L, int64_t dimm, int64_t npert, cplx* bc1, const cplx* Y, const double* alpha,
                  SolveOut& r) {
  hipStream_t s = L.stream;
  const cplx one = mk(1.0, 0.0), zero = mk(0.0, 0.0);
  cplx* X = L.z("fin_X", dimm * npert);
  double* wb = L.d("fin_wb", dimm);
  check_finite(c, s, "B~", bc1, dimm * dimm);
  zheevd(c, L, (int)dimm, bc1, wb, "a B~ matrix");
  set_gemm_tag(4);
  zgemm(s, 'C', 'N', dimm, npert, dimm, one, bc1, dimm, 0, Y, dimm, 0, zero, X, dimm, 0, 1);
  set_gemm_tag(0);
  const int64_t np = e2_partials_needed(dimm * npert);
  double* partial = L.d("partial", 2 * np);
  double* out2 = L.d("out2", 2);
  e2_reduce(s, dimm, npert, X, wb, alpha, partial, np, out2);
  read_out2(L, out2, r);
}

// count of retained eigenvalues (w >= 1e-8, cutoff.F90:64-69) and the distance of the
// closest eigenvalue to the threshold
// closest eigenvalue to the threshold; scale: the factor between these eigenvalues and those
// of the reference's S matrix (2 for the antisymmetric reduction of A)
int64_t count_retained(Lane& L, int64_t n, const double* w_dev, double& margin, double scale = 1.0) {
  std::vector<double> w(n);
  L.download(w.data(), w_dev, n * sizeof(double));
  const double thr = 1.0e-8;
  int64_t dimm = 0;
  for (int64_t k = 0; k < n; ++k) {
    if (scale * w[k] >= thr) ++dimm;
    margin = std::min(margin, std::fabs(scale * w[k] - thr));
  }
  return dimm;
}

// The V-independent half of a subspace solve (solve_A_subspace.F90:150-200): S -> cdiag(1e-8)
// -> ccutoff -> ulambda_s_half -> B~ -> cdiag.  Leaves U L^-1/2 (X_uc), the eigenvectors of
// B~ (X_bc1) and its eigenvalues (X_wb) in lane buffers named for subspace X; S and B are
// overwritten.  scale: the factor between these S eigenvalues and those of the reference's S
// matrix (2 for the antisymmetric reduction of A).
struct BlockPrep {
  int64_t n = 0, dimm = 0;
  double margin = INFINITY;
};
BlockPrep solve_prep(dc2_context* c, Lane& L, char X, int64_t n, cplx* S, cplx* B, EigKind kind, double scale = 1.0) {
  BlockPrep p;
  p.n = n;
  if (n == 0) return p;
  hipStream_t s = L.stream;
  auto nm = [X](const char* b) { return std::string(1, X) + b; };
  double* ws = L.d(nm("_ws").c_str(), n);
  zheevd(c, L, (int)n, S, ws, "an S matrix", kind);
  const int64_t dimm = count_retained(L, n, ws, p.margin, scale);
  p.dimm = dimm;
  if (dimm == 0) return p;
  const cplx one = mk(1.0, 0.0), zero = mk(0.0, 0.0);
  cplx* uc = L.z(nm("_uc").c_str(), n * dimm);
  cplx* T1 = L.z(nm("_T1").c_str(), dimm * n);
  cplx* bc1 = L.z(nm("_bc1").c_str(), dimm * dimm);
  double* wb = L.d(nm("_wb").c_str(), dimm);
  check_finite(c, s, "S eigenvectors", S, n * n);
  scale_cols_tail(s, n, dimm, S, ws, uc);
  set_gemm_tag(4);
  zgemm(s, 'C', 'N', dimm, n, n, one, uc, n, 0, B, n, 0, zero, T1, dimm, 0, 1);
  zgemm(s, 'N', 'N', dimm, dimm, n, one, T1, dimm, 0, uc, n, 0, zero, bc1, dimm, 0, 1);
  set_gemm_tag(0);
  check_finite(c, s, "B~", bc1, dimm * dimm);
  zheevd(c, L, (int)dimm, bc1, wb, "a B~ matrix");
  return p;
}

// The V half (solve_A_subspace.F90:197-257): Y = (U L^-1/2)^H V, X = W^H Y, E2 and sumc2
SolveOut solve_finish(dc2_context* c, Lane& L, char X, const BlockPrep& p, int64_t npert, const cplx* V,
                      const double* alpha) {
  SolveOut r;
  r.margin = p.margin;
  r.dimm = p.dimm;
  if (p.n == 0 || p.dimm == 0 || npert == 0) return r;
  hipStream_t s = L.stream;
  auto nm = [X](const char* b) { return std::string(1, X) + b; };
  const int64_t n = p.n, dimm = p.dimm;
  const cplx one = mk(1.0, 0.0), zero = mk(0.0, 0.0);
  const cplx* uc = L.z(nm("_uc").c_str(), n * dimm);
  const cplx* bc1 = L.z(nm("_bc1").c_str(), dimm * dimm);
  const double* wb = L.d(nm("_wb").c_str(), dimm);
  cplx* Y = L.z(nm("_Y").c_str(), dimm * npert);
  cplx* Xm = L.z(nm("_X").c_str(), dimm * npert);
  set_gemm_tag(4);
  zgemm(s, 'C', 'N', dimm, npert, n, one, uc, n, 0, V, n, 0, zero, Y, dimm, 0, 1);
  zgemm(s, 'C', 'N', dimm, npert, dimm, one, bc1, dimm, 0, Y, dimm, 0, zero, Xm, dimm, 0, 1);
  set_gemm_tag(0);
  const int64_t np = e2_partials_needed(dimm * npert);
  double* partial = L.d(nm("_partial").c_str(), 2 * np);
  double* out2 = L.d(nm("_out2").c_str(), 2);
  e2_reduce(s, dimm, npert, Xm, wb, alpha, partial, np, out2);
  read_out2(L, out2, r);
  return r;
}

struct GramMetric {
  bool valid = false;
  bool btilde = false;    // B~ and its eigendecomposition already in X_bc1 / X_wbt
  bool thin = false;      // C through its thin factor: T's eigenvectors in X_Tthin, t L^-1/2 in X_TSthin
  int64_t m = 0, dimm = 0;
  double margin = INFINITY;
  double seconds = 0.0;   // time spent on the lane before the subspace itself ran
};

// C's metric through its thin factor (build_thin_c) when that is smaller than the
// (nelec-1)-electron space and the compressed resolvent cannot be used anyway
bool gram_thin(const dc2_context* c, char X, int64_t m) {
  const int na = c->na, ne = c->ne;
  if (X != 'C' || ne < 2 || na < 2 || c->no_thin) return false;
  const int64_t nred = (int64_t)na * na * (na - 1) / 2;
  // the blocks of M M^T have the eigenvalue 2 + kappa (na - 1) along sigma: invertible only
  const double kappa = (double)(na - 2 * (ne - 1)) / ((double)(ne - 1) * (ne - 1));
  return nred < m && m - nred > e2_schur_max_dropped() && 2.0 + kappa * (na - 1) > 1e-6;
}

// Phi from the E_uv|0> vectors Y of the CASCI state in one orbital basis, G = Phi Phi^H,
// its eigendecomposition (lane buffers X_G, X_lam) and the retained count (w >= 1e-8,
// cutoff.F90:64-69).  eps_act: the orbital energies of f (zeros when f is not used).
GramMetric gram_eig(dc2_context* c, Lane& L, char X, int64_t n, const cplx* Y, const double* eps_act) {
  GramMetric g;
  hipStream_t s = L.stream;
  const bool create = X == 'A';
  auto nm = [X](const char* b) { return std::string(1, X) + b; };
  const int na = c->na;
  const int64_t m = binom_host(na, create ? c->ne + 1 : c->ne - 1);
  const double thr = 1.0e-8;
  g.valid = true;
  g.m = m;
  if (n > m) g.margin = thr;   // the n - m eigenvalues of S that vanish identically
  if (m == 0) return g;
  const cplx one = mk(1.0, 0.0), zero = mk(0.0, 0.0);
  int64_t* strings = L.l(nm("_strings").c_str(), m);
  cplx* Phi = L.z(nm("_Phi").c_str(), m * n);
  cplx* G = L.z(nm("_G").c_str(), m * m);
  double* f = L.d(nm("_f").c_str(), m);
  double* lam = L.d(nm("_lam").c_str(), m);
  build_phi(s, na, c->ne, create, c->nfull, c->rank_full.get(), Y, eps_act, m, strings, Phi, f);
  // a RAS-restricted CI vector leaves directions of the (nelec+-1) space unreached: G is then
  // rank-deficient, where plain divide and conquer can return NaN eigenvectors
  const bool ras = c->ras[0] + c->ras[1] + c->ras[2] != 0;
  const int64_t nred = (int64_t)na * na * (na - 1) / 2;
  if (gram_thin(c, X, m)) {
    g.thin = true;
    // G = Psi' Psi'^H with the thin factor Psi' (m x nred): the nonzero eigenvalues of G are
    // those of T = Psi'^H Psi' (nred x nred), its eigenvectors Psi' t / sqrt(l); only the
    // retained ones are formed (the compressed resolvent, which needs all of G, is not used
    // with m - nred > e2_schur_max_dropped() directions cut)
    cplx* Av = L.z(nm("_Av").c_str(), m * na);
    cplx* PsiH = L.z(nm("_PsiH").c_str(), nred * m);
    cplx* T = L.z(nm("_Tthin").c_str(), nred * nred);
    double* lt = L.d(nm("_lthin").c_str(), nred);
    build_thin_c(s, na, c->ne, m, Phi, Av, PsiH);
    zherk_upper(s, nred, m, PsiH, nred, T, nred);
    check_finite(c, s, "Psi'^H Psi'", T, nred * nred);
    zheevd(c, L, (int)nred, T, lt, "a determinant-space metric", ras ? EigKind::rank_deficient : EigKind::general);
    g.margin = std::min(g.margin, thr);   // the m - nred eigenvalues of G that vanish identically
    g.dimm = count_retained(L, nred, lt, g.margin);
    const int64_t nd = m - g.dimm;
    fill_zero(s, G, m * nd);
    if (nd > 0) DC2_HIP(hipMemsetAsync(lam, 0, nd * sizeof(double), s));
    if (g.dimm > 0) {
      cplx* TS = L.z(nm("_TSthin").c_str(), nred * g.dimm);
      scale_cols_tail(s, nred, g.dimm, T, lt, TS);
      zgemm(s, 'C', 'N', m, g.dimm, nred, one, PsiH, nred, 0, TS, nred, 0, zero, G + m * nd, m, 0, 1);
      DC2_HIP(hipMemcpyAsync(lam + nd, lt + (nred - g.dimm), g.dimm * sizeof(double), hipMemcpyDeviceToDevice, s));
    }
    return g;
  }
  zherk_upper(s, m, n, Phi, m, G, m);   // Hermitian: the upper tiles only
  check_finite(c, s, "Phi Phi^H", G, m * m);
  zheevd(c, L, (int)m, G, lam, "a determinant-space metric", ras ? EigKind::rank_deficient : EigKind::general);
  g.dimm = count_retained(L, m, lam, g.margin);
  return g;
}

// Eigenvectors of G computed for the CASCI vector in the original orbitals, carried to the
// canonical (rotated) orbitals.  G is the same operator in both bases (the sums over t, u, v
// of a+_t E_uv|0><0|E_vu a_t do not depend on the orbital basis of the active space), so its
// eigenvalues, hence the retained set, are those of the rotated G, and an eigenvector g in
// the original string basis is ds^H g in the rotated one, with ds(k, l) = det(U_act[occ_k,
// occ_l]) over the (nelec+-1)-strings: tracic's ds (trac.F90:105-141) on that space, whose
// coefficients transform as c = ds c~.
void gram_rotate(dc2_context* c, Lane& L, char X, const GramMetric& g) {
  if (g.m == 0) return;
  hipStream_t s = L.stream;
  auto nm = [X](const char* b) { return std::string(1, X) + b; };
  const int na = c->na, nmo = c->nmo();
  const int64_t m = g.m;
  const cplx one = mk(1.0, 0.0), zero = mk(0.0, 0.0);
  const int64_t* strings = L.l(nm("_strings").c_str(), m);
  cplx* G = L.z(nm("_G").c_str(), m * m);
  cplx* Ua = L.z(nm("_Ua").c_str(), (int64_t)na * na);
  cplx* M = L.z(nm("_M").c_str(), m * m);
  cplx* T = L.z(nm("_T").c_str(), m * m);
  int64_t ld[3] = {nmo, 1, 1};
  int off[4] = {c->ni, c->ni, 0, 0}, nn[4] = {na, na, 1, 1};
  extract4(s, c->U.get(), ld, off, nn, Ua);
  tracic_ds(s, na, X == 'A' ? c->ne + 1 : c->ne - 1, m, strings, Ua, M);
  // the cut directions are needed only by the compressed resolvent
  const int64_t nd = m - g.dimm, c0 = (c->schur && nd <= e2_schur_max_dropped()) ? 0 : nd;
  set_gemm_tag(4);
  zgemm(s, 'C', 'N', m, m - c0, m, one, M, m, 0, G + m * c0, m, 0, zero, T, m, 0, 1);
  set_gemm_tag(0);
  DC2_HIP(hipMemcpyAsync(G + m * c0, T, m * (m - c0) * sizeof(cplx), hipMemcpyDeviceToDevice, s));
}

// Phi and f of the rotated state (unless the metric was computed from it) and GL = G_k L^-1
void gram_rotated_parts(dc2_context* c, Lane& L, char X, int64_t n, const GramMetric& g, bool phi_current) {
  if (g.m == 0) return;
  hipStream_t s = L.stream;
  auto nm = [X](const char* b) { return std::string(1, X) + b; };
  const int64_t m = g.m;
  if (!phi_current)
    build_phi(s, c->na, c->ne, X == 'A', c->nfull, c->rank_full.get(), c->Y1.get(), c->eps.get() + c->ni, m,
              L.l(nm("_strings").c_str(), m), L.z(nm("_Phi").c_str(), m * n), L.d(nm("_f").c_str(), m));
  if (g.dimm > 0)
    scale_cols_tail(s, m, g.dimm, L.z(nm("_G").c_str(), m * m), L.d(nm("_lam").c_str(), m),
                    L.z(nm("_GL").c_str(), m * g.dimm), true);
}

// the whole metric from the rotated state (when it was not started early)
GramMetric gram_metric(dc2_context* c, Lane& L, char X, int64_t n) {
  GramMetric g = gram_eig(c, L, X, n, c->Y1.get(), c->eps.get() + c->ni);
  gram_rotated_parts(c, L, X, n, g, true);
  return g;
}

// B~ = G_k^H F G_k and its eigendecomposition (lane buffers X_bc1, X_wbt), which need the
// rotated metric and F only, not V: done as soon as those exist, so that the eigensolve
// overlaps the integral transforms.  Not needed when E2 will come from the compressed
// resolvent (at most e2_schur_max_dropped() directions dropped).
void gram_btilde(dc2_context* c, Lane& L, char X, GramMetric& g) {
  const int64_t m = g.m, dimm = g.dimm, nd = m - dimm;
  if (m == 0 || dimm == 0 || (c->schur && nd <= e2_schur_max_dropped())) return;
  hipStream_t s = L.stream;
  auto nm = [X](const char* b) { return std::string(1, X) + b; };
  const cplx one = mk(1.0, 0.0), zero = mk(0.0, 0.0);
  const cplx* Gk = L.z(nm("_G").c_str(), m * m) + m * nd;
  const double* f = L.d(nm("_f").c_str(), m);
  cplx* FG = L.z(nm("_FG").c_str(), m * dimm);
  cplx* bc1 = L.z(nm("_bc1").c_str(), dimm * dimm);
  double* wb = L.d(nm("_wbt").c_str(), dimm);
  scale_rows(s, m, dimm, f, Gk, FG);
  set_gemm_tag(4);
  zgemm(s, 'C', 'N', dimm, dimm, m, one, Gk, m, 0, FG, m, 0, zero, bc1, dimm, 0, 1);
  set_gemm_tag(0);
  check_finite(c, s, "B~", bc1, dimm * dimm);
  zheevd(c, L, (int)dimm, bc1, wb, "a B~ matrix");
  g.btilde = true;
}

// A and C through their determinant-space factor.  With Phi the (m x n) matrix of the
// vectors a+_t E_uv|0> (A) or a_t E_uv|0> (C) over the full (nelec+-1)-electron string
// space, the formula matrices of solve_A/C_subspace.F90 are S = Phi^H Phi and
// B = Phi^H F Phi, F = diag(sum_w eps_w n_w) (F commutes past the ladder operators).
// The nonzero spectrum of S is that of G = Phi Phi^H (m x m): for G g = l g, the S
// eigenvector is u = Phi^H g / sqrt(l), so the retained set (l >= 1e-8) and dimm are those
// of the reference, U L^-1/2 = Phi^H G_k L^-1, B~ = G_k^H F G_k and
// (U L^-1/2)^H V = L^-1 G_k^H (Phi V).  Used when m < n = nact^3.  When at most
// e2_schur_max_dropped() directions are dropped, E2 comes from the compressed resolvent
// of diag(F + alpha) (k_e2_schur) instead of an eigendecomposition of B~.  g: the metric
// of gram_metric on the same lane; lane buffers are named per subspace X (A or C), so both
// metrics can be held on one lane.
SolveOut solve_block_gram(dc2_context* c, Lane& L, char X, const GramMetric& g, int64_t n, int64_t npert,
                          const cplx* V, const double* alpha) {
  SolveOut r;
  hipStream_t s = L.stream;
  auto nm = [X](const char* b) { return std::string(1, X) + b; };
  const int64_t m = g.m, dimm = g.dimm;
  r.margin = g.margin;
  r.dimm = dimm;
  if (m == 0 || dimm == 0 || npert == 0) return r;
  const cplx one = mk(1.0, 0.0), zero = mk(0.0, 0.0);
  const cplx* Phi = L.z(nm("_Phi").c_str(), m * n);
  cplx* G = L.z(nm("_G").c_str(), m * m);
  const double* f = L.d(nm("_f").c_str(), m);
  const cplx* GL = L.z(nm("_GL").c_str(), m * dimm);
  const int64_t nd = m - dimm;
  const cplx* Gk = G + m * nd;
  cplx* Y = L.z(nm("_Y").c_str(), dimm * npert);
  if (g.thin) {
    // (U L^-1/2)^H V = L^-1/2 t^H (M M^T)^(-1/2) M V in the thin factor's space (no division
    // by the small eigenvalues beyond the reference's own L^-1/2)
    const int na = c->na;
    const int64_t nred = (int64_t)na * na * (na - 1) / 2;
    cplx* MV = L.z(nm("_MVthin").c_str(), nred * npert);
    cplx* W = L.z(nm("_Wthin").c_str(), nred * npert);
    const cplx* TS = L.z(nm("_TSthin").c_str(), nred * dimm);
    thin_c_w(s, na, c->ne, npert, V, MV, W);
    set_gemm_tag(4);
    zgemm(s, 'C', 'N', dimm, npert, nred, one, TS, nred, 0, W, nred, 0, zero, Y, dimm, 0, 1);
    set_gemm_tag(0);
  } else {
    cplx* P = L.z(nm("_P").c_str(), m * npert);
    set_gemm_tag(4);
    zgemm(s, 'N', 'N', m, npert, n, one, Phi, m, 0, V, n, 0, zero, P, m, 0, 1);
    zgemm(s, 'C', 'N', dimm, npert, m, one, GL, m, 0, P, m, 0, zero, Y, dimm, 0, 1);
    set_gemm_tag(0);
  }
  if (c->schur && nd <= e2_schur_max_dropped()) {
    cplx* Z = L.z(nm("_Z").c_str(), m * npert);
    set_gemm_tag(4);
    zgemm(s, 'N', 'N', m, npert, dimm, one, Gk, m, 0, Y, dimm, 0, zero, Z, m, 0, 1);
    set_gemm_tag(0);
    double* partial = L.d("partial", 2 * npert);
    double* out2 = L.d("out2", 2);
    e2_schur(s, m, (int)nd, npert, G, Z, f, alpha, partial, out2);
    read_out2(L, out2, r);
    if (std::isfinite(r.e2) && std::isfinite(r.sumc2)) return r;
    // a singular compressed system: fall through to the eigendecomposition of B~
  }
  if (g.btilde) {   // B~ diagonalised before V was known
    const cplx* bc1 = L.z(nm("_bc1").c_str(), dimm * dimm);
    const double* wb = L.d(nm("_wbt").c_str(), dimm);
    cplx* Xm = L.z(nm("_Xg").c_str(), dimm * npert);
    set_gemm_tag(4);
    zgemm(s, 'C', 'N', dimm, npert, dimm, one, bc1, dimm, 0, Y, dimm, 0, zero, Xm, dimm, 0, 1);
    set_gemm_tag(0);
    const int64_t np = e2_partials_needed(dimm * npert);
    double* partial = L.d("partial", 2 * np);
    double* out2 = L.d("out2", 2);
    e2_reduce(s, dimm, npert, Xm, wb, alpha, partial, np, out2);
    read_out2(L, out2, r);
    return r;
  }
  cplx* FG = L.z(nm("_FG").c_str(), m * dimm);
  cplx* bc1 = L.z(nm("_bc1").c_str(), dimm * dimm);
  scale_rows(s, m, dimm, f, Gk, FG);
  set_gemm_tag(4);
  zgemm(s, 'C', 'N', dimm, dimm, m, one, Gk, m, 0, FG, m, 0, zero, bc1, dimm, 0, 1);
  set_gemm_tag(0);
  finish_block(c, L, dimm, npert, bc1, Y, alpha, r);
  return r;
}

struct AcRoute {
  bool gram = false, reduced = false;
  bool phi = false;   // reduced: S and B as products of the determinant-space factor
};

// A and C: determinant-space metric when the (nelec+-1)-electron space is the smallest
// basis; A: the antisymmetric reduction (tuples t < u) when that is smaller still, its S
// and B formed from the determinant-space factor (no F-weighted 3-RDM, and no wait for the
// densities)
AcRoute ac_route_of(const dc2_context* c, char X) {
  AcRoute r;
  if (X != 'A' && X != 'C') return r;
  const int na = c->na;
  const int64_t n = (int64_t)na * na * na;
  const int64_t mdet = binom_host(na, X == 'A' ? c->ne + 1 : c->ne - 1);
  const int64_t nred = (int64_t)na * na * (na - 1) / 2;
  if (c->ac_route == 2) r.gram = true;
  else if (c->ac_route == 3) r.reduced = X == 'A';
  else if (c->ac_route == 4) {
    if (X == 'A') r.reduced = r.phi = true;
    else r.gram = true;
  } else if (c->ac_route == 0) {
    if (X == 'A' && nred < mdet && nred < n) r.reduced = r.phi = true;
    else r.gram = mdet < n;
  }
  return r;
}

// sizes of subspace X: the S/B dimension n (the formula matrices) and the perturbers
void subspace_dims(const dc2_context* c, char X, int64_t& n, int64_t& npert) {
  const int64_t ni = c->ni, na = c->na, ns = c->ns;
  const int64_t ntp = na * (na - 1) / 2, nip = ni * (ni - 1) / 2, nsp = ns * (ns - 1) / 2;
  n = 0;
  npert = 0;
  switch (X) {
    case 'A': n = na * na * na; npert = ni; break;
    case 'B': n = ntp; npert = nip; break;
    case 'C': n = na * na * na; npert = ns; break;
    case 'D': n = na * na; npert = ns * ni; break;
    case 'E': n = na; npert = nip * ns; break;
    case 'F': n = ntp; npert = nsp; break;
    case 'G': n = na; npert = nsp * ni; break;
    default: npert = nsp * nip; break;   // H
  }
}

// the V-independent half of subspace k: the S/B matrices from the densities (or the
// determinant-space metric) and their eigendecompositions
struct SubPrep {
  bool valid = false;
  BlockPrep bp;     // formula and antisymmetric routes
  GramMetric g;     // determinant route
  double seconds = 0.0;
};

SubPrep subspace_prep(dc2_context* c, Lane& L, int k) {
  SubPrep p;
  p.valid = true;
  const char X = kSub[k];
  if (X == 'H') return p;
  hipStream_t s = L.stream;
  auto nm = [X](const char* b) { return std::string(1, X) + b; };
  Timer tx(s);
  const int na = c->na;
  int64_t n, npert;
  subspace_dims(c, X, n, npert);
  const AcRoute route = ac_route_of(c, X);
  if (route.gram) {
    p.g = gram_metric(c, L, X, n);
  } else if (n > 0) {
    Dens d{na, c->G1.get(), c->G1F.get(), c->G2.get(), c->G2F.get(), c->G3.get(), c->G3F.get()};
    const int64_t mdet = (X == 'A' || X == 'C') ? binom_host(na, X == 'A' ? c->ne + 1 : c->ne - 1) : 0;
    const int64_t nred = (int64_t)na * na * (na - 1) / 2;
    const int64_t nb = route.reduced ? nred : n;   // dimension of the S/B matrices built
    cplx* S = L.z(nm("_S").c_str(), nb * nb);
    cplx* B = L.z(nm("_B").c_str(), nb * nb);
    const int* pairs = (X == 'B' || X == 'F') ? c->tpairs.get() : nullptr;
    int* amap = nullptr;
    if (route.reduced) {   // row k of the reduced matrices is tuple (t, u, v), t < u, of the formulas
      std::vector<int> hm;
      hm.reserve(nred);
      for (int t = 0; t < na; ++t)
        for (int u = t + 1; u < na; ++u)
          for (int v = 0; v < na; ++v) hm.push_back((t * na + u) * na + v);
      amap = L.i(nm("_map").c_str(), nred);
      DC2_HIP(hipMemcpyAsync(amap, hm.data(), nred * sizeof(int), hipMemcpyHostToDevice, s));
      DC2_HIP(hipStreamSynchronize(s));
    }
    if (route.phi) {
      // the reduced rows and columns of S = Phi^H Phi and B = Phi^H F Phi (solve_block_gram),
      // from the rows t < u of Phi^H over the (nelec+1)-electron strings
      int64_t* strings = L.l(nm("_strings").c_str(), mdet);
      cplx* Psi = L.z(nm("_Psi").c_str(), nred * mdet);
      cplx* Psif = L.z(nm("_Psif").c_str(), nred * mdet);
      build_phi_adj(s, na, c->ne, c->nfull, c->rank_full.get(), c->Y1.get(), c->eps.get() + c->ni, mdet, strings,
                    nred, amap, Psi, Psif);
      set_gemm_tag(4);
      zherk_upper(s, nred, mdet, Psi, nred, S, nred);
      zgemm_herm(s, nred, mdet, Psif, nred, Psi, nred, B, nred);
      set_gemm_tag(0);
    } else {
      build_SB(s, X, d, c->eps.get() + c->ni, c->e0.get(), nb, pairs, S, B, amap);
    }
    check_finite(c, s, (std::string("S") + X).c_str(), S, nb * nb);
    check_finite(c, s, (std::string("B") + X).c_str(), B, nb * nb);
    if (route.reduced) {
      // S_full = K^H S_red K with K K^H = 2 I (a+_u a+_t = -a+_t a+_u, a+_t a+_t = 0), so the
      // nonzero eigenvalues of the reference's S are twice those of S_red: cutoff at 0.5e-8
      const bool ras = c->ras[0] + c->ras[1] + c->ras[2] != 0;
      p.bp = solve_prep(c, L, X, nred, S, B, nred > mdet || (route.phi && ras) ? EigKind::rank_deficient
                                                                                 : EigKind::general, 2.0);
    } else {
      const EigKind kind = (X == 'A' || X == 'C') ? EigKind::rank_deficient : EigKind::general;
      p.bp = solve_prep(c, L, X, n, S, B, kind);
    }
  }
  p.seconds = tx.stop();
  return p;
}

// The 3-RDM part of V_A and V_C (solve_A_subspace.F90:571-593, solve_C_subspace.F90:547-564)
// from the E_qr|0> vectors Y1 of the rotated state, without the 3-RDM: for A,
// V(tuv, i) = -sum_pqr <E_vu E_pt E_qr> (p i|q r) with the block A1[p + na (i + ni (q + na r))];
// for C, V(tuv, a) = sum_jkl <E_vu E_tj E_kl> (a j|k l) with C1[a + ns (j + na (k + na l))].
// Zt = M Y1^T sums over the last index pair (M: the block as an (nx na) x na^2 matrix),
// v3_esum applies the excitation over the summed index through the E_pq table, and
// Y1^H Wt^T contracts with <0|E_vu: 2 * 8 nf na^3 nx flops against 8 nf na^6 for the 3-RDM.
void v3_part(dc2_context* c, Lane& L, char X, const cplx* blk, int nx, cplx* V) {
  hipStream_t s = L.stream;
  const int na = c->na;
  const int64_t nf = c->nfull, na2 = (int64_t)na * na, nr = (int64_t)nx * na;
  if (nx <= 0 || nf <= 0) return;
  const cplx one = mk(1.0, 0.0), zero = mk(0.0, 0.0);
  cplx* Zt = L.z("v3_Zt", nr * nf);
  cplx* Wt = L.z("v3_Wt", nr * nf);
  cplx* Out = L.z("v3_Out", na2 * nr);
  set_gemm_tag(3);
  zgemm(s, 'N', 'T', nr, nf, na2, one, blk, nr, 0, c->Y1.get(), nf, 0, zero, Zt, nr, 0, 1);
  if (X == 'C') v3_esum(s, true, na, nf, nx, 1, nx, nr, c->exc.get(), Zt, Wt);   // Zt[(a + ns j) + nr J]
  else v3_esum(s, false, na, nf, nx, na, 1, nr, c->exc.get(), Zt, Wt);          // Zt[(p + na i) + nr J]
  zgemm(s, 'C', 'T', na2, nr, nf, one, c->Y1.get(), nf, 0, Wt, nr, 0, zero, Out, na2, 0, 1);
  set_gemm_tag(0);
  v3_permute(s, na, nx, X == 'A' ? -1.0 : 1.0, Out, V);
}

// one subspace k (0..7 = A..H) on lane L: V and alpha, then the solve with the prepared
// S/B half (prepared here if prep is null or not valid)
// V = <X|H|0> of subspace k (not H) and its energy shifts alpha, in buffers of lane L (named
// after the subspace, so that a lane can hold those of several subspaces); k's transformed
// blocks, the densities and h' must be ready on L's stream
void subspace_v(dc2_context* c, Lane& L, int k, double eshift, cplx*& V_out, double*& alpha_out) {
  hipStream_t s = L.stream;
  const int ni = c->ni, na = c->na, ns = c->ns, nmo = c->nmo();
  Dens d{na, c->G1.get(), c->G1F.get(), c->G2.get(), c->G2F.get(), c->G3.get(), c->G3F.get()};
  auto blk = [&](const char* name) { return c->tblk.at(name).get(); };
  const char X = kSub[k];
  auto nm = [X](const char* b) { return std::string(1, X) + b; };
  int64_t n = 0, npert = 0;
  subspace_dims(c, X, n, npert);
  cplx* V = L.z(nm("_V").c_str(), n * npert);
  double* alpha = L.d(nm("_alpha").c_str(), npert);
  if (n > 0 && npert > 0) {
    switch (X) {
      case 'A': {
        cplx* eff = L.z("A_eff", na * ni);
        const cplx* A1 = blk("A1");
        effh_A(s, ni, na, c->hp.get(), nmo, blk("A2"), eff);
        v3_part(c, L, 'A', A1, ni, V);
        V_A_rest(s, na, ni, d, A1, eff, V);
      } break;
      case 'B': V_B(s, na, ni, d, blk("B"), c->tpairs.get(), c->ipairs.get(), V); break;
      case 'C': {
        cplx* eff = L.z("C_eff", (int64_t)ns * na);
        const cplx* C1 = blk("C1");
        if (ni) effh_C(s, ni, na, ns, c->hp.get(), nmo, C1, blk("C2"), blk("C3"), eff);
        else effh_C(s, 0, na, ns, c->hp.get(), nmo, C1, C1, C1, eff);
        v3_part(c, L, 'C', C1, ns, V);
        V_C_rest(s, na, ns, d, eff, V);
      } break;
      case 'D': {
        cplx* eff = L.z("D_eff", (int64_t)ns * ni);
        effh_D(s, ni, na, ns, c->hp.get(), nmo, blk("D3"), eff);
        V_D(s, na, ni, ns, d, blk("D1"), blk("D2"), eff, V);
      } break;
      case 'E': V_E(s, na, ni, ns, d, blk("E"), c->ipairs.get(), V); break;
      case 'F':
        V_F(s, na, ns, d, blk("F"), c->tpairs.get(), c->spairs.get(), V, L.z("F_vscr", V_F_scratch(na, ns)));
        break;
      case 'G': V_G(s, na, ni, ns, d, blk("G"), c->spairs.get(), V); break;
    }
    alphas(s, X, ni, na, ns, c->eps.get(), c->e0.get(), eshift, c->ipairs.get(), c->spairs.get(), npert, alpha);
  }
  check_finite(c, s, nm("V").c_str(), V, n * npert);
  check_finite(c, s, nm("alpha").c_str(), alpha, npert);
  V_out = V;
  alpha_out = alpha;
}

void run_subspace(dc2_context* c, Lane& L, int k, double eshift, double* e2, double* sumc2,
                  const SubPrep* prep = nullptr, const cplx* vpre = nullptr, const double* apre = nullptr,
                  hipEvent_t vpre_ready = nullptr) {
  hipStream_t s = L.stream;
  const int ni = c->ni, na = c->na, ns = c->ns;
  const int64_t nip = (int64_t)ni * (ni - 1) / 2, nsp = (int64_t)ns * (ns - 1) / 2;
  auto blk = [&](const char* name) { return c->tblk.at(name).get(); };
  const char X = kSub[k];
  auto nm = [X](const char* b) { return std::string(1, X) + b; };
  SubPrep local;
  if (!prep || !prep->valid) {
    local = subspace_prep(c, L, k);
    prep = &local;
  }
  Timer tx(s);
  SolveOut r;
  int64_t n = 0, npert = 0;
  subspace_dims(c, X, n, npert);
  if (X == 'H') {
    // eps laid out [secondary..., inactive...] for the kernel
    double* ep = L.d("H_eps", ns + ni);
    DC2_HIP(hipMemcpyAsync(ep, c->eps.get() + ni + na, ns * sizeof(double), hipMemcpyDeviceToDevice, s));
    DC2_HIP(hipMemcpyAsync(ep + ns, c->eps.get(), ni * sizeof(double), hipMemcpyDeviceToDevice, s));
    const int64_t np = e2_partials_needed(nsp * nip);
    double* partial = L.d("H_partial", 2 * np);
    double* out2 = L.d("H_out2", 2);
    e2_H(s, ni, ns, blk("H"), ep, eshift, c->ipairs.get(), c->spairs.get(), partial, np, out2);
    read_out2(L, out2, r);
    n = 0;
  } else {
    const AcRoute route = ac_route_of(c, X);
    cplx* V = nullptr;
    double* alpha = nullptr;
    if (vpre) {   // computed on another lane
      DC2_HIP(hipStreamWaitEvent(s, vpre_ready, 0));
      V = const_cast<cplx*>(vpre);
      alpha = const_cast<double*>(apre);
    } else {
      subspace_v(c, L, k, eshift, V, alpha);
    }
    if (route.gram) {
      r = solve_block_gram(c, L, X, prep->g, n, npert, V, alpha);
    } else if (route.reduced) {
      const int64_t nred = (int64_t)na * na * (na - 1) / 2;
      cplx* Vr = L.z("A_Vred", nred * npert);
      gather_rows(s, nred, npert, L.i(nm("_map").c_str(), nred), V, n, Vr);
      r = solve_finish(c, L, X, prep->bp, npert, Vr, alpha);
    } else {
      r = solve_finish(c, L, X, prep->bp, npert, V, alpha);
    }
  }
  e2[k] = r.e2;
  sumc2[k] = r.sumc2;
  if (c->on_subspace) c->on_subspace(c->on_subspace_user, k, r.e2, r.sumc2);
  c->sub_info[4 * k + 0] = n;
  c->sub_info[4 * k + 1] = r.dimm;
  c->sub_info[4 * k + 2] = npert;
  c->sub_info[4 * k + 3] = 1;
  c->sub_margin[k] = r.margin;
  c->timings[4 + k] = tx.stop() + prep->seconds;
}

// worker lane i (1-based), created on first use
Lane& worker_lane(dc2_context* c, int i) {
  while ((int)c->lanes.size() < i) {
    auto L = std::make_unique<Lane>();
    L->owned = true;
    DC2_HIP(hipStreamCreateWithFlags(&L->stream, hipStreamNonBlocking));
    c->lanes.push_back(std::move(L));
  }
  Lane& L = *c->lanes[i - 1];
  if (!L.blas) {   // streams made at context creation get their handle here
    DC2_ROCSOLVER(rocblas_create_handle(&L.blas));
    DC2_ROCSOLVER(rocblas_set_stream(L.blas, L.stream));
  }
  return L;
}

// first: subspaces before it (0 = A .. 8 = none computed) are skipped, as with a caspt2_restart
// file whose next subspace is `first` (module_restart_file.F90:15-24); e2_file/sumc2_file: the
// values read from that file (zeros where it lists nothing, may be null).  As in
// r4dcaspt2_tra.F90:168 and solve_X_subspace, the totals start from the sums of the file's
// values and every subspace computed adds its own.
// CI vector of root selectroot (1-based).  A root whose energy lies within 1e-8 Eh of other
// held roots (odd electron counts: the Kramers pair) is represented by the normalised
// projection P e_m of one determinant onto the group's eigenspace: m has the largest weight
// P_mm = sum_k |c_mk|^2, the lowest index among weights within a relative 1e-6 of it.  The
// eigensolver's basis of a degenerate eigenspace is arbitrary (ZHEEV's choice in the
// reference, diag.F90:109) and E2 depends on the vector; this one does not depend on that
// basis.  A group that may continue beyond the held roots is an error.
const cplx* root_vector(dc2_context* c, int selectroot) {
  const int64_t nd = c->ndet;
  const int k = selectroot - 1;
  std::vector<int> grp;
  for (int j = 0; j < c->nroots_avail; ++j)
    if (std::fabs(c->ecas[j] - c->ecas[k]) < 1.0e-8) grp.push_back(j);
  if (grp.size() == 1) return c->civec.get() + (int64_t)k * nd;
  DC2_CHECK(grp.back() < c->nroots_avail - 1 || c->nroots_avail == nd, 1,
            "the selected root is degenerate with the last root dc2_casci computed: request more roots");
  const int g = (int)grp.size();
  std::vector<zc> V((size_t)nd * g), out((size_t)nd, zc(0.0, 0.0));
  for (int q = 0; q < g; ++q)
    copy_sync(c, V.data() + (size_t)nd * q, c->civec.get() + (int64_t)grp[q] * nd, nd * sizeof(cplx),
              hipMemcpyDeviceToHost);
  std::vector<double> wt((size_t)nd, 0.0);
  double mx = 0.0;
  for (int64_t i = 0; i < nd; ++i) {
    for (int q = 0; q < g; ++q) wt[i] += std::norm(V[i + (size_t)nd * q]);
    mx = std::max(mx, wt[i]);
  }
  int64_t m = 0;
  while (wt[m] < mx * (1.0 - 1.0e-6)) ++m;
  const double inv = 1.0 / std::sqrt(wt[m]);
  for (int q = 0; q < g; ++q) {
    const zc a = std::conj(V[m + (size_t)nd * q]) * inv;
    for (int64_t i = 0; i < nd; ++i) out[i] += V[i + (size_t)nd * q] * a;
  }
  cplx* dst = c->lane0->z("root_vector", nd);
  copy_sync(c, dst, out.data(), nd * sizeof(cplx), hipMemcpyHostToDevice);
  return dst;
}

void run_caspt2(dc2_context* c, int selectroot, double eshift, double* e2, double* sumc2, double* scal,
                int first = 0, const double* e2_file = nullptr, const double* sumc2_file = nullptr) {
  DC2_CHECK(c->have_casci, 1, "dc2_casci must run before dc2_caspt2");
  DC2_CHECK(selectroot >= 1 && selectroot <= c->nroots_avail, 1,
            "selectroot out of range (dc2_casci must compute at least selectroot roots)");
  DC2_CHECK(!(c->ni == 0 && c->ns == 0), 1, "The CASPT2 energy cannot be defined when ninact = 0 and nsec = 0.");
  hipStream_t s = c->stream;
  const int ni = c->ni, na = c->na, ns = c->ns, nmo = c->nmo();
  const int64_t nf = c->nfull, nd = c->ndet, na2 = (int64_t)na * na;
  const cplx one = mk(1.0, 0.0), zero = mk(0.0, 0.0);
  for (double& t : c->timings) t = (&t == &c->timings[0]) ? t : 0.0;
  std::memset(c->sub_info, 0, sizeof(c->sub_info));
  for (int k = 0; k < DC2_NSUBSPACE; ++k) c->sub_margin[k] = INFINITY;

  Timer t1(s);
  // ---- E_pq tables and the CASCI 1-RDM in the original basis
  Lane& L0 = *c->lane0;
  c->exc.ensure(nf * na2);
  build_exc_table(s, na, nf, c->full_dets.get(), c->rank_full.get(), c->exc.get());
  c->eps.ensure(nmo);
  c->e0.ensure(1);
  c->c_full.ensure(nf);
  c->cF.ensure(nf);
  c->Y1.ensure(nf * na2);
  c->Y1o.ensure(nf * na2);
  c->G1.ensure(na2);
  double* zeros_act = L0.d("zeros_act", std::max(na, 1));
  DC2_HIP(hipMemsetAsync(zeros_act, 0, std::max(na, 1) * sizeof(double), s));
  const cplx* c0 = root_vector(c, selectroot);
  embed_ci(s, na, nf, nd, c->cas2full.get(), c0, c->full_dets.get(), zeros_act, c->c_full.get(), c->cF.get());
  apply_E(s, na, nf, 1, c->exc.get(), c->c_full.get(), c->Y1o.get());
  zgemm(s, 'T', 'R', na2, 1, nf, one, c->Y1o.get(), nf, 0, c->c_full.get(), nf, 0, zero, c->G1.get(), na2, 0, 1);
  check_finite(c, "the CASCI 1-RDM", c->G1.get(), na2);

  // Subspaces: A and C (the large metrics) on their own lanes, D on a third, the rest on the
  // library stream.  Without RAS limits the A/C lanes start now with the
  // eigendecompositions of their determinant-space metrics, which need only E_uv|0> of the
  // CASCI state in the original orbitals (the metric is the same operator in every orbital
  // basis of the active space), and carry the eigenvectors to the canonical orbitals once
  // the active rotation is known.  Every lane prepares the S/B halves of its subspaces once
  // the densities exist and takes the V halves once the transforms are done.
  const bool doA = ni > 0 && first <= 0, doB = ni > 0 && first <= 1, doC = ns > 0 && first <= 2;
  const bool doD = ni > 0 && ns > 0 && first <= 3, doE = ni > 0 && ns > 0 && first <= 4;
  const bool doF = ns > 0 && first <= 5, doG = ni > 0 && ns > 0 && first <= 6, doH = ni > 0 && ns > 0 && first <= 7;
  const bool do_sub[8] = {doA, doB, doC, doD, doE, doF, doG, doH};
  // lanes: 1 all on the library stream; 2: A and C on lane 1; 3: A on 1, C on 2; 4: D (the
  // next largest) on lane 3 as well, with B and F when the secondary space is large: their
  // S/B eigensolves then run beside the integral transforms instead of after them on the
  // library stream (324 spinors, 280 secondary: 90.1 against 92.7 ms per energy; 192
  // spinors, 164 secondary: 16.2 against 16.1 ms, where F on lane 3 made that lane the longest)
  const bool bf_on_d = c->lane_plan == 1 || (c->lane_plan < 0 && ns >= 256);
  std::vector<std::vector<int>> plan(std::max(1, std::min(c->nlanes, 4)));
  for (int k = 0; k < 8; ++k) {
    if (!do_sub[k]) continue;
    int lane = 0;
    if (plan.size() >= 2 && (k == 0 || k == 2)) lane = k == 0 ? 1 : std::min<int>(2, (int)plan.size() - 1);
    if (plan.size() >= 4 && (k == 3 || (bf_on_d && (k == 1 || k == 5)))) lane = 3;
    plan[lane].push_back(k);
  }
  const bool ras = c->ras[0] + c->ras[1] + c->ras[2] != 0;
  hipEvent_t y1o_ready = nullptr, rot_ready = nullptr, dens_ready = nullptr, ready = nullptr;
  DC2_HIP(hipEventCreateWithFlags(&y1o_ready, hipEventDisableTiming));
  DC2_HIP(hipEventCreateWithFlags(&rot_ready, hipEventDisableTiming));
  DC2_HIP(hipEventCreateWithFlags(&dens_ready, hipEventDisableTiming));
  DC2_HIP(hipEventCreateWithFlags(&ready, hipEventDisableTiming));
  DC2_HIP(hipEventRecord(y1o_ready, s));
  std::vector<std::exception_ptr> errs(plan.size());
  std::vector<Lane*> lane_of(plan.size(), c->lane0.get());
  for (size_t w = 1; w < plan.size(); ++w)   // all lanes exist before any thread starts
    if (!plan[w].empty()) lane_of[w] = &worker_lane(c, (int)w);
  SubPrep prep[8];
  // gates, each true when the stages before it succeeded and its event is recorded:
  // rotation known (rot_ready), densities (dens_ready), transforms (ready)
  std::promise<bool> go_rot, go_dens, go;
  std::shared_future<bool> go_rot_f = go_rot.get_future().share(), go_dens_f = go_dens.get_future().share(),
                           go_f = go.get_future().share();
  // V of A and C on lane 3 (D's lane, idle while the transforms run) as soon as their blocks
  // are transformed (ac_ready, go_ac), so that the A and C lanes, the longest, find it ready
  // after their two eigensolves (vpre_ready, go_vpre: set once by lane 3 on every path)
  const bool early_v = c->early_v && plan.size() >= 4 && !plan[3].empty() && !ras &&
                       (plan.size() > 1 && (doA || doC));
  const cplx* vpre[8] = {};
  const double* apre[8] = {};
  hipEvent_t ac_ready = nullptr, vpre_ready = nullptr;
  DC2_HIP(hipEventCreateWithFlags(&ac_ready, hipEventDisableTiming));
  DC2_HIP(hipEventCreateWithFlags(&vpre_ready, hipEventDisableTiming));
  std::promise<bool> go_ac, go_vpre;
  std::shared_future<bool> go_ac_f = go_ac.get_future().share(), go_vpre_f = go_vpre.get_future().share();
  std::atomic<bool> ac_set{false}, vpre_set{false};
  auto set_ac = [&](bool v) {
    if (!ac_set.exchange(true)) go_ac.set_value(v);
  };
  auto set_vpre = [&](bool v) {
    if (!vpre_set.exchange(true)) go_vpre.set_value(v);
  };
  if (!early_v) {
    set_ac(false);
    set_vpre(false);
  }
  std::vector<std::thread> threads;
  for (size_t w = 1; w < plan.size(); ++w) {
    if (plan[w].empty()) continue;
    threads.emplace_back([&, w]() {
      try {
        DC2_HIP(hipSetDevice(c->device));
        Lane& Lw = *lane_of[w];
        DC2_HIP(hipStreamWaitEvent(Lw.stream, y1o_ready, 0));
        const int64_t n3 = (int64_t)na * na2;
        auto gram_k = [&](int k) { return (k == 0 || k == 2) && ac_route_of(c, kSub[k]).gram; };
        // the original-basis metric describes the rotated state only when the CASCI space is
        // closed under the active rotation (no RAS limits); with RAS limits tracic projects
        // the rotated state back onto the restricted space, and the metric waits for it.  C's
        // thin factor is built from the rotated state (its eigenvectors t live in the tuple
        // basis, which the rotation mixes)
        if (!ras)
          for (int k : plan[w]) {
            if (!gram_k(k) || gram_thin(c, kSub[k], binom_host(na, c->ne - 1))) continue;
            Timer tm(Lw.stream);
            prep[k].g = gram_eig(c, Lw, kSub[k], n3, c->Y1o.get(), zeros_act);
            prep[k].seconds = tm.stop();
          }
        if (!go_rot_f.get()) return;
        DC2_HIP(hipStreamWaitEvent(Lw.stream, rot_ready, 0));
        for (int k : plan[w]) {
          if (!gram_k(k)) continue;
          Timer tm(Lw.stream);
          if (prep[k].g.valid) {
            gram_rotate(c, Lw, kSub[k], prep[k].g);
            gram_rotated_parts(c, Lw, kSub[k], n3, prep[k].g, false);
          } else {
            prep[k].g = gram_metric(c, Lw, kSub[k], n3);
          }
          gram_btilde(c, Lw, kSub[k], prep[k].g);
          prep[k].seconds += tm.stop();
          prep[k].valid = true;
        }
        // A's reduced metric from the determinant-space factor needs the rotated state only
        for (int k : plan[w])
          if (k == 0 && ac_route_of(c, 'A').phi) prep[k] = subspace_prep(c, Lw, k);
        if (!go_dens_f.get()) {
          if (w == 3) set_vpre(false);
          return;
        }
        DC2_HIP(hipStreamWaitEvent(Lw.stream, dens_ready, 0));
        for (int k : plan[w])
          if (!prep[k].valid) prep[k] = subspace_prep(c, Lw, k);
        if (w == 3 && early_v) {
          if (!go_ac_f.get()) {
            set_vpre(false);
            return;
          }
          DC2_HIP(hipStreamWaitEvent(Lw.stream, ac_ready, 0));
          cplx* V = nullptr;
          double* alpha = nullptr;
          for (int k : {0, 2})
            if (do_sub[k]) {
              subspace_v(c, Lw, k, eshift, V, alpha);
              vpre[k] = V;
              apre[k] = alpha;
            }
          DC2_HIP(hipEventRecord(vpre_ready, Lw.stream));
          set_vpre(true);
        }
        if (!go_f.get()) return;
        DC2_HIP(hipStreamWaitEvent(Lw.stream, ready, 0));
        for (int k : plan[w]) {
          if ((k == 0 || k == 2) && early_v && go_vpre_f.get())
            run_subspace(c, Lw, k, eshift, e2, sumc2, &prep[k], vpre[k], apre[k], vpre_ready);
          else
            run_subspace(c, Lw, k, eshift, e2, sumc2, &prep[k]);
        }
        DC2_HIP(hipStreamSynchronize(Lw.stream));
      } catch (...) {
        errs[w] = std::current_exception();
        if (w == 3) set_vpre(false);
      }
    });
  }
  auto finish_lanes = [&]() {
    set_ac(false);   // no effect once set; releases lane 3 when the transforms never ran
    for (auto& t : threads) t.join();
    (void)hipEventDestroy(ac_ready);
    (void)hipEventDestroy(vpre_ready);
    (void)hipEventDestroy(y1o_ready);
    (void)hipEventDestroy(rot_ready);
    (void)hipEventDestroy(dens_ready);
    (void)hipEventDestroy(ready);
  };
  // with the automatic solver choice, blocks of at most kHostFockEig spinors are diagonalised
  // on the host (a device solve is a chain of dependent launches); the canonical gauge makes
  // the eigenvectors independent of the solver
  constexpr int kHostFockEig = 64;
  auto diag_space = [&](int sp) {
    const int lo = sp == 0 ? 0 : (sp == 1 ? ni : ni + na), hi = sp == 0 ? ni : (sp == 1 ? ni + na : nmo);
    for (int isym = 1; isym <= c->nsym; ++isym) {
      std::vector<int> ind;
      for (int i = lo; i < hi; ++i)
        if ((c->irpamo.empty() ? 1 : c->irpamo[i]) == isym) ind.push_back(i);
      if (ind.empty()) continue;
      const int n = (int)ind.size();
      int* dind = L0.i("fd_ind", n);
      DC2_HIP(hipMemcpyAsync(dind, ind.data(), n * sizeof(int), hipMemcpyHostToDevice, s));
      cplx* blk = L0.z("fd_blk", (int64_t)n * n);
      cplx* scr = L0.z("fd_scr", (int64_t)n * n);
      double* w = L0.d("fd_w", n);
      gather_block(s, n, dind, c->fock.get(), nmo, blk);
      if (n <= kHostFockEig && c->eig_method < 0) host_heev_upper(L0, n, blk, w);   // not when a solver is chosen
      else zheevd(c, n, blk, w, "a Fock block");
      canonical_gauge(s, n, w, blk, scr, L0.i("fd_starts", n + 2));
      scatter_block(s, n, dind, blk, w, c->U.get(), nmo, c->eps.get());
      DC2_HIP(hipStreamSynchronize(s));
    }
  };
  const double* eps_act = c->eps.get() + ni;
  try {
    // ---- Fock matrix (fockcasci.F90) and its diagonalisation per space and irrep (fockdiag.F90)
    c->fock.ensure((int64_t)nmo * nmo);
    const cplx* F1 = ns ? get_block(c, "F1") : nullptr;
    const cplx* F2 = ns ? get_block(c, "F2") : nullptr;
    fock_build(s, ni, na, ns, c->h.get(), nmo, get_block(c, "OOOO"), F1, F2, c->G1.get(), c->fock.get());
    check_finite(c, "the Fock matrix", c->fock.get(), (int64_t)nmo * nmo);
    c->U.ensure((int64_t)nmo * nmo);
    fill_zero(s, c->U.get(), (int64_t)nmo * nmo);
    diag_space(0);
    diag_space(1);
    // ---- tracic: ds c~ = c (trac.F90:76-156)
    {
      cplx* Ua = L0.z("tr_Ua", na2);
      cplx* ds = L0.z("tr_ds", nd * nd);
      int64_t ld[3] = {nmo, 1, 1};
      int off[4] = {ni, ni, 0, 0}, nn[4] = {na, na, 1, 1};
      extract4(s, c->U.get(), ld, off, nn, Ua);
      tracic_ds(s, na, c->ne, nd, c->cas_dets.get(), Ua, ds);
      c->cnew.ensure(nd);
      if (!ras) {
        // without RAS limits the CASCI space of one total symmetry is closed under the active
        // rotation (U_act mixes spinors of one irrep only), so ds is the unitary image of U_act
        // on that space and ds^-1 = ds^H
        zgemm(s, 'C', 'N', nd, 1, nd, one, ds, nd, 0, c0, nd, 0, zero, c->cnew.get(), nd, 0, 1);
      } else {
        DC2_HIP(hipMemcpyAsync(c->cnew.get(), c0, nd * sizeof(cplx), hipMemcpyDeviceToDevice, s));
        int* ipiv = L0.i("tr_ipiv", nd);
        c->info.ensure(1);
        DC2_ROCSOLVER(rocsolver_zgetrf(c->blas, (int)nd, (int)nd, reinterpret_cast<rocblas_double_complex*>(ds),
                                       (int)nd, ipiv, c->info.get()));
        int info = 0;
        L0.download(&info, c->info.get(), sizeof(int));
        if (info != 0) throw Error(100 + info, "zgetrf (tracic) failed, info = " + std::to_string(info));
        DC2_ROCSOLVER(rocsolver_zgetrs(c->blas, rocblas_operation_none, (int)nd, 1,
                                       reinterpret_cast<rocblas_double_complex*>(ds), (int)nd, ipiv,
                                       reinterpret_cast<rocblas_double_complex*>(c->cnew.get()), (int)nd));
      }
    }
    check_finite(c, "the rotated CI vector", c->cnew.get(), nd);
    // ---- E_uv|0> of the rotated state
    embed_ci(s, na, nf, nd, c->cas2full.get(), c->cnew.get(), c->full_dets.get(), eps_act, c->c_full.get(),
             c->cF.get());
    apply_E(s, na, nf, 1, c->exc.get(), c->c_full.get(), c->Y1.get());

    DC2_HIP(hipEventRecord(rot_ready, s));
  } catch (...) {
    go_rot.set_value(false);
    go_dens.set_value(false);
    go.set_value(false);
    finish_lanes();
    throw;
  }
  go_rot.set_value(true);
  try {
    diag_space(2);
    c->timings[1] = t1.stop();
    check_finite(c, "U", c->U.get(), (int64_t)nmo * nmo);
    check_finite(c, "eps", c->eps.get(), nmo);

    // ---- densities of the rotated state (density_matrix.F90)
    Timer t2(s);
    set_gemm_tag(2);
    c->Y1F.ensure(nf * na2);
    apply_E(s, na, nf, 1, c->exc.get(), c->cF.get(), c->Y1F.get());
    c->G1F.ensure(na2);
    c->G2.ensure(na2 * na2);
    c->G2F.ensure(na2 * na2);
    zgemm(s, 'T', 'R', na2, 1, nf, one, c->Y1.get(), nf, 0, c->c_full.get(), nf, 0, zero, c->G1.get(), na2, 0, 1);
    zgemm(s, 'T', 'R', na2, 1, nf, one, c->Y1F.get(), nf, 0, c->c_full.get(), nf, 0, zero, c->G1F.get(), na2, 0, 1);
    calce0(s, na, eps_act, c->G1.get(), c->e0.get());
    zgemm(s, 'C', 'N', na2, na2, nf, one, c->Y1.get(), nf, 0, c->Y1.get(), nf, 0, zero, c->G2.get(), na2, 0, 1);
    zgemm(s, 'C', 'N', na2, na2, nf, one, c->Y1.get(), nf, 0, c->Y1F.get(), nf, 0, zero, c->G2F.get(), na2, 0, 1);
    // the F-weighted 3-RDM enters only the formula S/B matrices of A and C (not their
    // determinant-space route, not any V)
    // the 3-RDM and its F-weighted form enter only the formula S/B matrices of A and C (V
    // comes from the E_qr|0> vectors, the determinant-space routes from Phi)
    const bool need_g3f = (doA && !ac_route_of(c, 'A').gram && !ac_route_of(c, 'A').phi) ||
                          (doC && !ac_route_of(c, 'C').gram);
    const bool need_g3 = need_g3f;
    if (need_g3) {
      const int64_t na4 = na2 * na2;
      c->G3.ensure(na2 * na4);
      cplx* Y2 = L0.z("Y2", nf * na4);
      apply_E(s, na, nf, na2, c->exc.get(), c->Y1.get(), Y2);
      zgemm(s, 'C', 'N', na2, na4, nf, one, c->Y1.get(), nf, 0, Y2, nf, 0, zero, c->G3.get(), na2, 0, 1);
      if (need_g3f) {
        c->G3F.ensure(na2 * na4);
        apply_E(s, na, nf, na2, c->exc.get(), c->Y1F.get(), Y2);
        zgemm(s, 'C', 'N', na2, na4, nf, one, c->Y1.get(), nf, 0, Y2, nf, 0, zero, c->G3F.get(), na2, 0, 1);
      }
    }
    // h' = U^H h U (tramo1)
    c->hp.ensure((int64_t)nmo * nmo);
    {
      cplx* T = L0.z("hT", (int64_t)nmo * nmo);
      zgemm(s, 'C', 'N', nmo, nmo, nmo, one, c->U.get(), nmo, 0, c->h.get(), nmo, 0, zero, T, nmo, 0, 1);
      zgemm(s, 'N', 'N', nmo, nmo, nmo, one, T, nmo, 0, c->U.get(), nmo, 0, zero, c->hp.get(), nmo, 0, 1);
    }
    set_gemm_tag(0);
    c->timings[2] = t2.stop();
    check_finite(c, "e0", c->e0.get(), 1);
    check_finite(c, "G1", c->G1.get(), na2);
    check_finite(c, "G1F", c->G1F.get(), na2);
    check_finite(c, "G2", c->G2.get(), na2 * na2);
    check_finite(c, "G2F", c->G2F.get(), na2 * na2);
    if (need_g3) check_finite(c, "G3", c->G3.get(), na2 * na2 * na2);
    if (need_g3f) check_finite(c, "G3F", c->G3F.get(), na2 * na2 * na2);
    check_finite(c, "h'", c->hp.get(), (int64_t)nmo * nmo);
    upload_ints(s, c->tpairs, pair_list(na));
    upload_ints(s, c->ipairs, pair_list(ni));
    upload_ints(s, c->spairs, pair_list(ns));
    DC2_HIP(hipEventRecord(dens_ready, s));
  } catch (...) {
    go_dens.set_value(false);
    go.set_value(false);
    finish_lanes();
    throw;
  }
  go_dens.set_value(true);
  try {

    // ---- transforms of every needed subspace block (module_intra.F90)
    Timer t3(s);
    std::vector<std::string> need;
    if (doA) { need.push_back("A1"); need.push_back("A2"); }
    if (doC) { need.push_back("C1"); if (ni) { need.push_back("C2"); need.push_back("C3"); } }
    const size_t n_ac = need.size();   // A's and C's blocks first: lane 3 forms their V early
    if (doB) need.push_back("B");
    if (doD) { need.push_back("D1"); need.push_back("D2"); need.push_back("D3"); }
    if (doE) need.push_back("E");
    if (doF) need.push_back("F");
    if (doG) need.push_back("G");
    if (doH) need.push_back("H");
    set_gemm_tag(1);
    c->split_flops = c->fp64_flops = 0.0;
    // every block's map entry exists before lane 3 may look up A's and C's (transform then
    // only finds entries; inserting one while another thread reads the map would race)
    for (const auto& nm : need) (void)c->tblk[nm];
    for (size_t i = 0; i < need.size(); ++i) {
      transform(c, need[i]);
      if (i + 1 == n_ac && early_v) {
        DC2_HIP(hipEventRecord(ac_ready, s));
        set_ac(true);
      }
    }
    set_gemm_tag(0);
    c->timings[3] = t3.stop();
    for (const auto& nm : need) check_finite(c, ("transformed " + nm).c_str(), c->tblk[nm].get(), block_count(c, shape_of(nm)));

    for (int k = 0; k < 8; ++k) { e2[k] = 0.0; sumc2[k] = 0.0; }

    DC2_HIP(hipEventRecord(ready, s));
  } catch (...) {
    go.set_value(false);
    finish_lanes();
    throw;
  }
  go.set_value(true);
  try {
    for (int k : plan[0]) run_subspace(c, *c->lane0, k, eshift, e2, sumc2, &prep[k]);
  } catch (...) {
    errs[0] = std::current_exception();
  }
  finish_lanes();
  for (auto& e : errs)
    if (e) std::rethrow_exception(e);
  double e0h = 0.0;
  L0.download(&e0h, c->e0.get(), sizeof(double));
  double e2all = 0.0, s2 = 0.0;
  for (int k = 0; k < 8; ++k) {
    const double fe = e2_file ? e2_file[k] : 0.0, fs = sumc2_file ? sumc2_file[k] : 0.0;
    e2all += fe;
    s2 += fs;
    if (do_sub[k]) {
      e2all += e2[k];
      s2 += sumc2[k];
    } else {
      e2[k] = fe;
      sumc2[k] = fs;
    }
  }
  scal[0] = e0h;
  scal[1] = e2all;
  scal[2] = s2;
  scal[3] = 1.0 / (1.0 + s2);
  scal[4] = e2all - eshift * s2;
  scal[5] = e2all + (c->ecas[selectroot - 1] + c->ecore) - eshift * s2;
  c->have_pt2 = true;
}

}  // namespace

namespace dc2 {
// A context whose worker-lane streams (lanes - 1 of them) are made right after its own, before
// anything uses the null stream (creating a rocBLAS handle does): HIP spreads streams over its
// hardware queues in creation order, so with GPU_MAX_HW_QUEUES = 4 (HIP's default) the library
// stream and three lanes get a queue each, and the null stream shares one.  Two lanes on one
// queue wait for each other's stream waits and reductions (the C chain waited behind D's).
int context_create_lanes(int32_t device, int lanes, dc2_context** out) {
  return guarded([&] {
    DC2_CHECK(out != nullptr, 1, "null output pointer");
    DC2_CHECK(lanes >= 1 && lanes <= 4, 1, "lanes must be 1..4");
    auto c = std::make_unique<dc2_context>();
    c->device = device;
    c->nlanes = lanes;
    const char* dbg = std::getenv("DC2_DEBUG");
    c->debug = dbg && dbg[0] && dbg[0] != '0';
    if (const char* em = std::getenv("DC2_EIGEN")) {
      const std::string m(em);
      c->eig_method = m == "qr" ? 0 : m == "dc" ? 1 : m == "dcj" ? 2 : m == "jacobi" ? 3 : m == "native" ? 4 : -1;
    }
    if (const char* r = std::getenv("DC2_AC_METRIC")) {
      const std::string m(r);
      c->ac_route = m == "formula" ? 1 : m == "determinant" ? 2 : 0;
    }
    DC2_HIP(hipSetDevice(device));
    DC2_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    for (int i = 1; i < lanes; ++i) {
      auto L = std::make_unique<Lane>();
      L->owned = true;
      DC2_HIP(hipStreamCreateWithFlags(&L->stream, hipStreamNonBlocking));
      c->lanes.push_back(std::move(L));
    }
    DC2_ROCSOLVER(rocblas_create_handle(&c->blas));
    DC2_ROCSOLVER(rocblas_set_stream(c->blas, c->stream));
    c->lane0 = std::make_unique<Lane>();
    c->lane0->stream = c->stream;
    c->lane0->blas = c->blas;
    *out = c.release();
  });
}

void set_last_error(const std::string& msg) { g_last_error = msg; }
void set_batcher(dc2_context* c, EigBatcher* b) { c->batcher = b; }
int run_guarded(const std::function<void()>& f) { return guarded(f); }
hipStream_t context_stream(dc2_context* c) { return c->stream; }
int context_device(dc2_context* c) { return c->device; }
}  // namespace dc2

// ===================================================================== C ABI
extern "C" {

const char* dc2_last_error(void) { return g_last_error.c_str(); }
const char* dc2_version(void) { return "dcaspt2_amd 0.1 (gfx950)"; }

int dc2_context_create(int32_t device, dc2_context** out) {
  int lanes = 4;
  if (const char* nl = std::getenv("DC2_LANES")) lanes = std::max(1, std::min(4, std::atoi(nl)));
  return dc2::context_create_lanes(device, lanes, out);
}

int dc2_context_destroy(dc2_context* ctx) {
  return guarded([&] {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    if (ctx->gstats.on) set_gemm_stats(nullptr);
    for (auto& r : ctx->gstats.rec) { (void)hipEventDestroy(r.a); (void)hipEventDestroy(r.b); }
    for (auto& kv : ctx->dav_graphs) (void)hipGraphExecDestroy(kv.second);
    if (ctx->blas) rocblas_destroy_handle(ctx->blas);
    hipStream_t s = ctx->stream;
    delete ctx;
    if (s) {
      release_stream_workspace(s);
      (void)hipStreamDestroy(s);
    }
  });
}

int dc2_set_spaces(dc2_context* c, int32_t ninact, int32_t nact, int32_t nsec, int32_t nelec) {
  return guarded([&] {
    DC2_CHECK(c, 1, "null context");
    DC2_CHECK(ninact >= 0 && nact >= 1 && nsec >= 0 && nelec >= 1 && nelec <= nact, 1, "invalid space sizes");
    DC2_CHECK(nact <= 24, 1, "nact > 24 is not supported");
    c->ni = ninact; c->na = nact; c->ns = nsec; c->ne = nelec;
    c->blocks.clear();   // transformed-block buffers are kept and regrown on demand
    c->have_casci = c->have_pt2 = false;
    c->nsym = 1; c->totsym = 1; c->irpamo.clear(); c->multb_d.clear(); c->multb_ds.clear();
    for (int& r : c->ras) r = 0;
  });
}

int dc2_set_symmetry(dc2_context* c, int32_t nsymrpa, const int64_t* multb, const int64_t* irpamo, int32_t totsym) {
  return guarded([&] {
    DC2_CHECK(c, 1, "null context");
    need_spaces(c);
    const int n = nsymrpa;
    DC2_CHECK(n >= 1 && n <= 64, 1, "nsymrpa out of range");
    c->nsym = n; c->totsym = totsym;
    c->irpamo.assign(irpamo, irpamo + c->nmo());
    // MULTB2(i0, MULTB(i0,j0)) = j0; MULTB_D = MULTB2(1:n,1:n) - n; MULTB_DS(i,j) = MULTB(n+j, i)
    std::vector<int> m2(4 * n * n, 0);
    for (int i0 = 0; i0 < 2 * n; ++i0)
      for (int j0 = 0; j0 < 2 * n; ++j0) {
        const int k0 = (int)multb[i0 + 2 * n * j0];
        DC2_CHECK(k0 >= 1 && k0 <= 2 * n, 1, "MULTB entry out of range");
        m2[i0 + 2 * n * (k0 - 1)] = j0 + 1;
      }
    c->multb_d.assign(n * n, 0);
    c->multb_ds.assign(n * n, 0);
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j) {
        c->multb_d[i + n * j] = m2[i + 2 * n * j] - n;
        c->multb_ds[i + n * j] = (int)multb[(n + j) + 2 * n * i];
      }
    c->have_casci = c->have_pt2 = false;
  });
}

int dc2_set_ras(dc2_context* c, int32_t r1, int32_t r2, int32_t r3, int32_t h1, int32_t e3, int32_t mh1,
                int32_t me3) {
  return guarded([&] {
    DC2_CHECK(c, 1, "null context");
    need_spaces(c);
    DC2_CHECK(r1 + r2 + r3 == 0 || r1 + r2 + r3 == c->na, 1, "RAS spaces must add up to nact");
    int v[7] = {r1, r2, r3, h1, e3, mh1, me3};
    std::memcpy(c->ras, v, sizeof(v));
    c->have_casci = c->have_pt2 = false;
  });
}

int dc2_set_one_electron(dc2_context* c, const double* h, double ecore) {
  return guarded([&] {
    DC2_CHECK(c && h, 1, "null argument");
    need_spaces(c);
    const int64_t n = (int64_t)c->nmo() * c->nmo();
    c->h.ensure(n);
    // through the lane's page-locked memory: a pageable upload waits on the runtime's staging
    // buffer, which the slots of a request server share
    void* hp = c->lane0->host(n * sizeof(cplx), 1);
    std::memcpy(hp, h, n * sizeof(cplx));
    DC2_HIP(hipMemcpyAsync(c->h.get(), hp, n * sizeof(cplx), hipMemcpyHostToDevice, c->stream));
    DC2_HIP(hipStreamSynchronize(c->stream));
    c->ecore = ecore;
  });
}

int dc2_set_integral_block(dc2_context* c, const char* name, const double* data, int64_t count, int32_t where) {
  return guarded([&] {
    DC2_CHECK(c && name && data, 1, "null argument");
    need_spaces(c);
    const std::string nm(name);
    const int64_t need = block_count(c, shape_of(nm));
    DC2_CHECK(count == need, 1, "integral block " + nm + ": expected " + std::to_string(need) + " elements");
    Block& b = c->blocks[nm];
    b.derived = false;
    b.count = count;
    if (where == 2) {
      b.own.release();
      b.ptr = reinterpret_cast<const cplx*>(data);
    } else {
      b.own.alloc(count);
      DC2_HIP(hipMemcpyAsync(b.own.get(), data, count * sizeof(cplx),
                             where == 0 ? hipMemcpyHostToDevice : hipMemcpyDeviceToDevice, c->stream));
      DC2_HIP(hipStreamSynchronize(c->stream));
      b.ptr = b.own.get();
    }
    c->have_pt2 = false;
    if (nm == "OOOO") {
      // copies previously derived from OOOO are stale now
      for (const char* k : {"A1", "A2", "B"}) {
        auto it = c->blocks.find(k);
        if (it != c->blocks.end() && it->second.derived) c->blocks.erase(it);
      }
      c->have_casci = false;
    }
  });
}

int dc2_casci(dc2_context* c, int32_t nroot, double* eigen, int64_t* ndet_out) {
  return guarded([&] {
    DC2_CHECK(c, 1, "null context");
    DC2_HIP(hipSetDevice(c->device));
    run_casci(c, nroot);
    if (ndet_out) *ndet_out = c->ndet;
    for (int k = 0; k < nroot && k < c->nroots_avail; ++k) eigen[k] = c->ecas[k] + c->ecore;
  });
}

int dc2_get_dets(dc2_context* c, int64_t* dets) {
  return guarded([&] {
    DC2_CHECK(c && c->have_casci, 1, "no CASCI result");
    copy_sync(c, dets, c->cas_dets.get(), c->ndet * sizeof(int64_t), hipMemcpyDeviceToHost);
  });
}

int dc2_get_ci_vector(dc2_context* c, int32_t root, double* v) {
  return guarded([&] {
    DC2_CHECK(c && c->have_casci, 1, "no CASCI result");
    DC2_CHECK(root >= 1 && root <= c->nroots_avail, 1, "root out of range (not computed by dc2_casci)");
    copy_sync(c, v, c->civec.get() + (int64_t)(root - 1) * c->ndet, c->ndet * sizeof(cplx),
              hipMemcpyDeviceToHost);
  });
}

int dc2_caspt2(dc2_context* c, int32_t selectroot, double eshift, double* e2, double* sumc2, double* scalars) {
  return guarded([&] {
    DC2_CHECK(c && e2 && sumc2 && scalars, 1, "null argument");
    DC2_HIP(hipSetDevice(c->device));
    run_caspt2(c, selectroot, eshift, e2, sumc2, scalars);
  });
}

int dc2_set_casci_result(dc2_context* c, int64_t ndet, int32_t nroot, const int64_t* dets, const double* coeffs,
                         const double* ecas) {
  return guarded([&] {
    DC2_CHECK(c && dets && coeffs && ecas && ndet > 0 && nroot >= 1, 1, "invalid argument");
    DC2_HIP(hipSetDevice(c->device));
    enumerate_space(c);
    DC2_CHECK(c->ndet == ndet, 1,
              "ndet of the CASCI result (" + std::to_string(ndet) + ") differs from this problem's (" +
                  std::to_string(c->ndet) + ")");
    std::vector<int64_t> mine(ndet);
    DC2_HIP(hipMemcpyAsync(mine.data(), c->cas_dets.get(), ndet * sizeof(int64_t), hipMemcpyDeviceToHost, c->stream));
    DC2_HIP(hipStreamSynchronize(c->stream));
    for (int64_t k = 0; k < ndet; ++k)
      DC2_CHECK(mine[k] == dets[k], 1, "determinant " + std::to_string(k + 1) + " of the CASCI result differs");
    c->civec.ensure(ndet * nroot);
    DC2_HIP(hipMemcpyAsync(c->civec.get(), coeffs, ndet * nroot * sizeof(cplx), hipMemcpyHostToDevice, c->stream));
    DC2_HIP(hipStreamSynchronize(c->stream));
    c->ecas.assign(ecas, ecas + nroot);
    c->nroots_avail = nroot;
    c->davidson_iters = 0;
    c->have_casci = true;
    c->have_pt2 = false;
  });
}

int dc2_caspt2_restart(dc2_context* c, int32_t selectroot, double eshift, int32_t first, const double* e2_file,
                       const double* sumc2_file, double* e2, double* sumc2, double* scalars) {
  return guarded([&] {
    DC2_CHECK(c && e2 && sumc2 && scalars && e2_file && sumc2_file, 1, "null argument");
    DC2_CHECK(first >= 0 && first <= 8, 1, "first must be 0 (A) .. 8 (none computed)");
    DC2_HIP(hipSetDevice(c->device));
    run_caspt2(c, selectroot, eshift, e2, sumc2, scalars, first, e2_file, sumc2_file);
  });
}

int dc2_get_subspace_info(dc2_context* c, int64_t* info, double* margin) {
  return guarded([&] {
    DC2_CHECK(c && c->have_pt2, 1, "no CASPT2 result");
    std::memcpy(info, c->sub_info, sizeof(c->sub_info));
    std::memcpy(margin, c->sub_margin, sizeof(c->sub_margin));
  });
}

int dc2_get_fock_rotation(dc2_context* c, double* eps, double* U) {
  return guarded([&] {
    DC2_CHECK(c && c->have_pt2, 1, "no CASPT2 result");
    const int64_t n = c->nmo();
    if (eps) copy_sync(c, eps, c->eps.get(), n * sizeof(double), hipMemcpyDeviceToHost);
    if (U) copy_sync(c, U, c->U.get(), n * n * sizeof(cplx), hipMemcpyDeviceToHost);
  });
}

int dc2_get_rotated_ci(dc2_context* c, double* v) {
  return guarded([&] {
    DC2_CHECK(c && c->have_pt2, 1, "no CASPT2 result");
    copy_sync(c, v, c->cnew.get(), c->ndet * sizeof(cplx), hipMemcpyDeviceToHost);
  });
}

int dc2_get_density(dc2_context* c, int32_t which, double* out, int64_t count) {
  return guarded([&] {
    DC2_CHECK(c && c->have_pt2, 1, "no CASPT2 result");
    const int64_t na2 = (int64_t)c->na * c->na;
    const cplx* src = nullptr;
    int64_t n = 0;
    switch (which) {
      case 0: src = c->G1.get(); n = na2; break;
      case 1: src = c->G1F.get(); n = na2; break;
      case 2: src = c->G2.get(); n = na2 * na2; break;
      case 3: src = c->G2F.get(); n = na2 * na2; break;
      default: throw Error(1, "density index out of range");
    }
    DC2_CHECK(count == n, 1, "density: wrong element count");
    copy_sync(c, out, src, n * sizeof(cplx), hipMemcpyDeviceToHost);
  });
}

int dc2_get_transformed_block(dc2_context* c, const char* name, double* out, int64_t count) {
  return guarded([&] {
    DC2_CHECK(c && c->have_pt2 && name, 1, "no CASPT2 result");
    auto it = c->tblk.find(name);
    DC2_CHECK(it != c->tblk.end(), 1, std::string("block ") + name + " was not transformed");
    const int64_t n = block_count(c, shape_of(name));
    DC2_CHECK(count == n, 1, "transformed block: wrong element count");
    copy_sync(c, out, it->second.get(), n * sizeof(cplx), hipMemcpyDeviceToHost);
  });
}

int dc2_get_timings(dc2_context* c, double* t, int32_t n) {
  return guarded([&] {
    DC2_CHECK(c && t, 1, "null argument");
    for (int k = 0; k < n && k < 12; ++k) t[k] = c->timings[k];
  });
}

int dc2_casmat(dc2_context* c, double* mat) {
  return guarded([&] {
    DC2_CHECK(c && c->have_casci && mat, 1, "no CASCI result");
    DC2_HIP(hipMemcpyAsync(mat, c->cimat.get(), c->ndet * c->ndet * sizeof(cplx), hipMemcpyDeviceToHost, c->stream));
    DC2_HIP(hipStreamSynchronize(c->stream));
  });
}

int dc2_cdiag(dc2_context* c, double* a, int32_t n, int32_t* dimm, double* w, double thr) {
  return guarded([&] {
    DC2_CHECK(c && a && w && dimm && n > 0, 1, "invalid argument");
    DevBuf<cplx> m((int64_t)n * n);
    DevBuf<double> wd(n);
    DC2_HIP(hipMemcpyAsync(m.get(), a, (int64_t)n * n * sizeof(cplx), hipMemcpyHostToDevice, c->stream));
    zheevd(c, n, m.get(), wd.get());
    DC2_HIP(hipMemcpyAsync(a, m.get(), (int64_t)n * n * sizeof(cplx), hipMemcpyDeviceToHost, c->stream));
    DC2_HIP(hipMemcpyAsync(w, wd.get(), n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    DC2_HIP(hipStreamSynchronize(c->stream));
    if (thr == 0.0) *dimm = n;
    else { int k = 0; for (int i = 0; i < n; ++i) k += (w[i] >= thr); *dimm = k; }
  });
}

int dc2_zgemm(dc2_context* c, char opA, char opB, int64_t M, int64_t N, int64_t K, const double* alpha,
              const double* A, int64_t lda, int64_t sA, const double* B, int64_t ldb, int64_t sB, const double* beta,
              double* C, int64_t ldc, int64_t sC, int64_t batch) {
  return guarded([&] {
    DC2_CHECK(c && alpha && beta, 1, "null argument");
    zgemm(c->stream, opA, opB, M, N, K, mk(alpha[0], alpha[1]), reinterpret_cast<const cplx*>(A), lda, sA,
          reinterpret_cast<const cplx*>(B), ldb, sB, mk(beta[0], beta[1]), reinterpret_cast<cplx*>(C), ldc, sC,
          batch);
  });
}

int dc2_zherk(dc2_context* c, int64_t m, int64_t k, const double* A, int64_t lda, double* C, int64_t ldc,
              int32_t split) {
  return guarded([&] {
    DC2_CHECK(c && A && C, 1, "null argument");
    DC2_CHECK(m >= 0 && k >= 0 && lda >= std::max<int64_t>(m, 1) && ldc >= std::max<int64_t>(m, 1), 1,
              "dc2_zherk: bad size or leading dimension");
    zherk_upper(c->stream, m, k, reinterpret_cast<const cplx*>(A), lda, reinterpret_cast<cplx*>(C), ldc, split);
  });
}

int dc2_malloc(dc2_context* c, int64_t bytes, void** ptr) {
  return guarded([&] {
    DC2_CHECK(c && ptr && bytes >= 0, 1, "invalid argument");
    DC2_HIP(hipSetDevice(c->device));
    *ptr = nullptr;
    if (bytes) DC2_HIP(hipMalloc(ptr, bytes));
  });
}

int dc2_free(dc2_context* c, void* ptr) {
  return guarded([&] {
    DC2_CHECK(c, 1, "null context");
    if (ptr) DC2_HIP(hipFree(ptr));
  });
}

int dc2_memcpy(dc2_context* c, void* dst, const void* src, int64_t bytes, int32_t kind) {
  return guarded([&] {
    DC2_CHECK(c && (bytes == 0 || (dst && src)), 1, "invalid argument");
    const hipMemcpyKind k = kind == 1 ? hipMemcpyHostToDevice : kind == 2 ? hipMemcpyDeviceToHost
                                                                          : hipMemcpyDeviceToDevice;
    DC2_CHECK(kind >= 1 && kind <= 3, 1, "memcpy kind must be 1, 2 or 3");
    if (bytes) DC2_HIP(hipMemcpyAsync(dst, src, bytes, k, c->stream));
    DC2_HIP(hipStreamSynchronize(c->stream));
  });
}

int dc2_set_integral_factors(dc2_context* c, const double* B, int64_t naux, int32_t where) {
  return guarded([&] {
    DC2_CHECK(c && B && naux > 0, 1, "invalid argument");
    need_spaces(c);
    hipStream_t s = c->stream;
    const int64_t nmo = c->nmo();
    DevBuf<cplx> Bd;
    const cplx* Bp = reinterpret_cast<const cplx*>(B);
    if (where == 0) {
      Bd.alloc(naux * nmo * nmo);
      DC2_HIP(hipMemcpyAsync(Bd.get(), B, naux * nmo * nmo * sizeof(cplx), hipMemcpyHostToDevice, s));
      Bp = Bd.get();
    }
    const bool has_s = c->ns > 0, has_i = c->ni > 0;
    std::vector<std::string> names = {"OOOO"};
    if (has_s) {
      for (const char* n : {"F1", "F2", "C1", "F"}) names.push_back(n);
      if (has_i) for (const char* n : {"C2", "C3", "D1", "D2", "D3", "E", "G", "H"}) names.push_back(n);
    }
    const cplx one = mk(1.0, 0.0), zero = mk(0.0, 0.0);
    for (const auto& nm : names) {
      const char* sp = shape_of(nm);
      int64_t n[4];
      for (int k = 0; k < 4; ++k) n[k] = space_dim(c, sp[k]);
      const int64_t m1 = n[0] * n[1], m2 = n[2] * n[3];
      DevBuf<cplx> L(naux * m1), R(naux * m2);
      int64_t ld[3] = {naux, nmo, nmo};
      int offL[4] = {0, space_off(c, sp[0]), space_off(c, sp[1]), 0}, nL[4] = {(int)naux, (int)n[0], (int)n[1], 1};
      int offR[4] = {0, space_off(c, sp[2]), space_off(c, sp[3]), 0}, nR[4] = {(int)naux, (int)n[2], (int)n[3], 1};
      extract4(s, Bp, ld, offL, nL, L.get());
      extract4(s, Bp, ld, offR, nR, R.get());
      Block& b = c->blocks[nm];
      b.derived = false;
      b.count = m1 * m2;
      b.own.alloc(b.count);
      b.ptr = b.own.get();
      set_gemm_tag(5);
      zgemm(s, 'T', 'N', m1, m2, naux, one, L.get(), naux, 0, R.get(), naux, 0, zero, b.own.get(), m1, 0, 1);
      set_gemm_tag(0);
      DC2_HIP(hipStreamSynchronize(s));
    }
    for (const char* k : {"A1", "A2", "B"}) {
      auto it = c->blocks.find(k);
      if (it != c->blocks.end() && it->second.derived) c->blocks.erase(it);
    }
    c->have_casci = c->have_pt2 = false;
  });
}

int dc2_set_profiling(dc2_context* c, int32_t on) {
  return guarded([&] {
    DC2_CHECK(c, 1, "null context");
    DC2_HIP(hipStreamSynchronize(c->stream));
    for (auto& r : c->gstats.rec) { (void)hipEventDestroy(r.a); (void)hipEventDestroy(r.b); }
    c->gstats.rec.clear();
    c->gstats.on = on != 0;
    set_gemm_tag(0);
    set_gemm_stats(on ? &c->gstats : nullptr);
  });
}

int dc2_get_gemm_stats(dc2_context* c, int32_t tag, double* out) {
  return guarded([&] {
    DC2_CHECK(c && out, 1, "null argument");
    DC2_HIP(hipStreamSynchronize(c->stream));
    double n = 0, sec = 0, fl = 0, mx = 0;
    for (auto& r : c->gstats.rec) {
      if (tag >= 0 && r.tag != tag) continue;
      float ms = 0;
      DC2_HIP(hipEventElapsedTime(&ms, r.a, r.b));
      n += 1;
      sec += ms * 1e-3;
      fl += r.flops;
      mx = std::max(mx, r.flops);
    }
    out[0] = n; out[1] = sec; out[2] = fl; out[3] = mx;
  });
}

namespace dc2 {
// every integral block from n records already on the device (the 16 images of each)
void ingest_device(dc2_context* c, int64_t n, const int* dijkl, const double* dval) {
  hipStream_t s = c->stream;
  std::vector<std::string> names;
  std::vector<std::array<int, 4>> spaces, dims;
  std::vector<int64_t> counts;
  for (const auto& sh : kShapes) {
    const int64_t cnt = block_count(c, sh.sp);
    if (cnt <= 0) continue;
    std::array<int, 4> sp, dm;
    for (int k = 0; k < 4; ++k) {
      const char ch = sh.sp[k];
      sp[k] = ch == 'i' ? 0 : ch == 'a' ? 1 : ch == 's' ? 2 : 3;
      dm[k] = space_dim(c, ch);
    }
    names.push_back(sh.name);
    spaces.push_back(sp);
    dims.push_back(dm);
    counts.push_back(cnt);
  }
  std::vector<DevBuf<unsigned long long>> owner(names.size());
  std::vector<cplx*> data(names.size());
  std::vector<unsigned long long*> own(names.size());
  std::vector<const int*> psp(names.size()), pdm(names.size());
  for (size_t b = 0; b < names.size(); ++b) {
    Block& blk = c->blocks[names[b]];
    blk.derived = false;
    blk.count = counts[b];
    blk.own.alloc(counts[b]);
    blk.ptr = blk.own.get();
    owner[b].alloc(counts[b]);
    data[b] = blk.own.get();
    own[b] = owner[b].get();
    psp[b] = spaces[b].data();
    pdm[b] = dims[b].data();
  }
  ingest_records(s, c->ni, c->na, c->ns, (int)names.size(), psp.data(), pdm.data(), data.data(), own.data(),
                 counts.data(), n, dijkl, dval);
  DC2_HIP(hipStreamSynchronize(s));
  c->have_casci = c->have_pt2 = false;
}
}  // namespace dc2

int dc2_set_integral_records(dc2_context* c, int64_t n, const int32_t* ijkl, const double* values, int32_t where) {
  return guarded([&] {
    DC2_CHECK(c && (n == 0 || (ijkl && values)), 1, "null argument");
    DC2_CHECK(where == 0 || where == 1, 1, "where must be 0 (host) or 1 (device)");
    need_spaces(c);
    DevBuf<int> dijkl(std::max<int64_t>(4 * n, 1));
    DevBuf<double> dval(std::max<int64_t>(2 * n, 1));
    const hipMemcpyKind kind = where == 0 ? hipMemcpyHostToDevice : hipMemcpyDeviceToDevice;
    if (n > 0) {
      DC2_HIP(hipMemcpyAsync(dijkl.get(), ijkl, 4 * n * sizeof(int), kind, c->stream));
      DC2_HIP(hipMemcpyAsync(dval.get(), values, 2 * n * sizeof(double), kind, c->stream));
    }
    ingest_device(c, n, dijkl.get(), dval.get());
  });
}

int dc2_mdcint_records(dc2_context* c, int64_t nrec, const int64_t* ikr, const int64_t* jkr, const int64_t* nz,
                       const int64_t* indk, const int64_t* indl, const double* rklr, const double* rkli, int64_t nkr,
                       const int64_t* kr, int64_t nd2c, const int64_t* d2c, int32_t scheme, int64_t cap,
                       int64_t* nout, int32_t* ijkl, double* values, int32_t ingest) {
  return guarded([&] {
    DC2_CHECK(c && nout && nrec >= 0 && nkr >= 0 && nd2c >= 1 && kr && d2c, 1, "invalid argument");
    DC2_CHECK(nrec == 0 || (ikr && jkr && nz), 1, "null record arrays");
    DC2_CHECK((ijkl == nullptr) == (values == nullptr), 1, "ijkl and values are given together or not at all");
    if (ingest) need_spaces(c);
    // the labels are checked here, so that every device access stays inside kr and d2c
    for (int64_t x = 0; x < 2 * nkr + 1; ++x)
      DC2_CHECK(kr[x] >= 0 && kr[x] < nd2c, 1, "kr holds a spinor outside d2c");
    for (int64_t k = 0; k < nd2c; ++k)
      DC2_CHECK(d2c[k] >= 0 && d2c[k] < (1 << 30), 1, "d2c holds an index out of range");
    std::vector<int64_t> eoff(nrec + 1), coff(nrec + 1);
    int64_t nent = 0, ncand = 0;
    for (int64_t r = 0; r < nrec; ++r) {
      DC2_CHECK(nz[r] >= 0 && std::llabs(ikr[r]) <= nkr && std::llabs(jkr[r]) <= nkr, 1,
                "record " + std::to_string(r) + ": label or count out of range");
      eoff[r] = nent;
      coff[r] = ncand;
      nent += nz[r];
      ncand += nz[r] * ((scheme != 6 && std::llabs(ikr[r]) != std::llabs(jkr[r])) ? 4 : 2);
    }
    eoff[nrec] = nent;
    coff[nrec] = ncand;
    DC2_CHECK(nent == 0 || (indk && indl && rklr), 1, "null entry arrays");
    std::vector<int> kl((size_t)std::max<int64_t>(2 * nent, 1));
    for (int64_t e = 0; e < nent; ++e) {
      DC2_CHECK(std::llabs(indk[e]) <= nkr && std::llabs(indl[e]) <= nkr, 1,
                "entry " + std::to_string(e) + ": label out of range");
      kl[e] = (int)indk[e];
      kl[nent + e] = (int)indl[e];
    }
    hipStream_t s = c->stream;
    auto up = [&](auto* dst, const auto* src, int64_t count) {
      if (count > 0) DC2_HIP(hipMemcpyAsync(dst, src, count * sizeof(*src), hipMemcpyHostToDevice, s));
    };
    std::vector<int> cas(2 * nkr + 1);   // Kramers label -> CAS index, kr and d2c composed
    for (int64_t x = 0; x < 2 * nkr + 1; ++x) cas[x] = (int)d2c[kr[x]];
    DevBuf<int64_t> d_ikr(std::max<int64_t>(nrec, 1)), d_jkr(std::max<int64_t>(nrec, 1)),
        d_nz(std::max<int64_t>(nrec, 1)), d_eoff(nrec + 1), d_coff(nrec + 1);
    DevBuf<int> d_cas(2 * nkr + 1);
    DevBuf<int> d_kl((size_t)std::max<int64_t>(2 * nent, 1));
    DevBuf<double> d_rr(std::max<int64_t>(nent, 1)), d_ri(rkli ? std::max<int64_t>(nent, 1) : 1);
    up(d_ikr.get(), ikr, nrec);
    up(d_jkr.get(), jkr, nrec);
    up(d_nz.get(), nz, nrec);
    up(d_eoff.get(), eoff.data(), nrec + 1);
    up(d_coff.get(), coff.data(), nrec + 1);
    up(d_cas.get(), cas.data(), 2 * nkr + 1);
    up(d_kl.get(), kl.data(), 2 * nent);
    up(d_rr.get(), rklr, nent);
    if (rkli) up(d_ri.get(), rkli, nent);
    DevBuf<int> d_ijkl(std::max<int64_t>(4 * ncand, 1));
    DevBuf<double> d_val(std::max<int64_t>(2 * ncand, 1));
    const int64_t nsc = mdcint_convert_scratch(ncand) / 8;
    DevBuf<int64_t> d_sc(nsc);
    mdcint_convert(s, nrec, d_ikr.get(), d_jkr.get(), d_nz.get(), d_eoff.get(), d_coff.get(), d_kl.get(),
                   d_kl.get() + nent, d_rr.get(), rkli ? d_ri.get() : nullptr, nkr, d_cas.get(), scheme, ncand,
                   d_sc.get(), d_ijkl.get(), d_val.get());
    int64_t count = 0;
    copy_sync(c, &count, d_sc.get() + nsc - 1, sizeof(int64_t), hipMemcpyDeviceToHost);
    *nout = count;
    if (ijkl) {
      DC2_CHECK(cap >= count, 1, "dc2_mdcint_records: room for " + std::to_string(count) + " records is needed");
      copy_sync(c, ijkl, d_ijkl.get(), 4 * count * sizeof(int), hipMemcpyDeviceToHost);
      copy_sync(c, values, d_val.get(), 2 * count * sizeof(double), hipMemcpyDeviceToHost);
    }
    if (ingest) ingest_device(c, count, d_ijkl.get(), d_val.get());
  });
}

int dc2_get_integral_block_device(dc2_context* c, const char* name, const double** ptr, int64_t* count) {
  return guarded([&] {
    DC2_CHECK(c && name && ptr && count, 1, "null argument");
    auto it = c->blocks.find(name);
    DC2_CHECK(it != c->blocks.end() && it->second.ptr != nullptr, 1, std::string("integral block ") + name +
                                                                         " is not resident");
    *ptr = reinterpret_cast<const double*>(it->second.ptr);
    *count = it->second.count;
  });
}

int dc2_set_option(dc2_context* c, const char* key, int64_t value) {
  return guarded([&] {
    DC2_CHECK(c && key, 1, "null argument");
    const std::string k(key);
    if (k == "eigensolver") {
      DC2_CHECK(value >= -1 && value <= 4, 1, "eigensolver must be -1..4");
      c->eig_method = (int)value;
    } else if (k == "ac_metric") {
      DC2_CHECK(value >= 0 && value <= 4, 1, "ac_metric must be 0..4");
      c->ac_route = (int)value;
    } else if (k == "c_thin") {
      DC2_CHECK(value == 0 || value == 1, 1, "c_thin must be 0 or 1");
      c->no_thin = value == 0;
    } else if (k == "early_v") {
      DC2_CHECK(value == 0 || value == 1, 1, "early_v must be 0 or 1");
      c->early_v = value != 0;
    } else if (k == "lane_plan") {
      DC2_CHECK(value >= -1 && value <= 1, 1, "lane_plan must be -1, 0 or 1");
      c->lane_plan = (int)value;
    } else if (k == "lanes") {
      DC2_CHECK(value >= 1 && value <= 4, 1, "lanes must be 1..4");
      c->nlanes = (int)value;
    } else if (k == "schur") {
      DC2_CHECK(value == 0 || value == 1, 1, "schur must be 0 or 1");
      c->schur = value != 0;
    } else if (k == "graphs") {
      DC2_CHECK(value == 0 || value == 1, 1, "graphs must be 0 or 1");
      c->graphs = value != 0;
    } else if (k == "casci_solver") {
      DC2_CHECK(value >= 0 && value <= 2, 1, "casci_solver must be 0..2");
      c->casci_solver = (int)value;
    } else if (k == "split_slices") {
      DC2_CHECK(value == 0 || (value >= 3 && value <= 8), 1, "split_slices must be 0 or 3..8");
      c->split_slices = (int)value;
    } else if (k == "split_min_k") {
      DC2_CHECK(value >= 1 && value <= 4096, 1, "split_min_k must be in 1..4096");
      c->split_min_k = (int)value;
    } else if (k == "davidson_block") {
      DC2_CHECK(value >= 0 && value <= 64, 1, "davidson_block must be in 0..64");
      c->dav_block = (int)value;
    } else if (k == "davidson_maxsub") {
      DC2_CHECK(value >= 0 && value <= 512, 1, "davidson_maxsub must be in 0..512");
      c->dav_maxsub = (int)value;
    } else if (k == "native_max") {
      DC2_CHECK(value >= 0 && value <= 2048, 1, "native_max must be in 0..2048");
      c->native_max = (int)value;
    } else if (k == "native_min") {
      DC2_CHECK(value >= 0, 1, "native_min must be >= 0");
      c->native_min = (int)value;
    } else if (k == "tridiag_solver") {
      DC2_CHECK(value == 0 || value == 1, 1, "tridiag_solver must be 0 or 1");
      c->tri_own = value == 0;
    } else {
      throw Error(1, "unknown option '" + k + "'");
    }
  });
}

int dc2_get_gemm_op_stats(dc2_context* c, char opA, char opB, double* out) {
  return guarded([&] {
    DC2_CHECK(c && out, 1, "null argument");
    DC2_HIP(hipStreamSynchronize(c->stream));
    double n = 0, sec = 0, fl = 0, mx = 0;
    for (auto& r : c->gstats.rec) {
      if (r.opA != opA || r.opB != opB || !r.mfma) continue;
      float ms = 0;
      DC2_HIP(hipEventElapsedTime(&ms, r.a, r.b));
      n += 1;
      sec += ms * 1e-3;
      fl += r.flops;
      mx = std::max(mx, r.flops);
    }
    out[0] = n; out[1] = sec; out[2] = fl; out[3] = mx;
  });
}

int dc2_get_graph_stats(dc2_context* c, int64_t* out) {
  return guarded([&] {
    DC2_CHECK(c && out, 1, "null argument");
    out[0] = c->graph_captures;
    out[1] = c->graph_launches;
    out[2] = (int64_t)c->dav_graphs.size();
    out[3] = c->graphs ? 1 : 0;
  });
}

int dc2_get_gemm_shapes(dc2_context* c, int64_t* out, int64_t cap, int64_t* count) {
  return guarded([&] {
    DC2_CHECK(c && count && (cap == 0 || out), 1, "null argument");
    std::vector<std::array<int64_t, 6>> shapes;
    for (auto& r : c->gstats.rec) {
      if (!r.mfma) continue;
      std::array<int64_t, 6> k{(int64_t)r.opA, (int64_t)r.opB, r.M, r.N, r.K, r.batch};
      if (std::find(shapes.begin(), shapes.end(), k) == shapes.end()) shapes.push_back(k);
    }
    *count = (int64_t)shapes.size();
    for (int64_t i = 0; i < std::min<int64_t>(cap, *count); ++i)
      for (int j = 0; j < 6; ++j) out[6 * i + j] = shapes[i][j];
  });
}

int dc2_gemm_tune(dc2_context* c, char opA, char opB, int64_t M, int64_t N, int64_t K, int64_t batch, int32_t reps,
                  double* out) {
  return guarded([&] {
    DC2_CHECK(c && out, 1, "null argument");
    DC2_HIP(hipSetDevice(c->device));
    gemm_tune(c->stream, opA, opB, M, N, K, batch, reps, out);
  });
}

int dc2_gemm_set_choice(char opA, char opB, int64_t M, int64_t N, int64_t K, int64_t batch, int32_t variant,
                        int32_t split) {
  return guarded([&] { gemm_set_choice(opA, opB, M, N, K, batch, variant, split); });
}

int dc2_gemm_clear_choices(int64_t* count_before) {
  return guarded([&] {
    if (count_before) *count_before = gemm_choice_count();
    gemm_clear_choices();
  });
}

int dc2_zgemm_split(dc2_context* c, int32_t slices, char opB, int64_t M, int64_t N, int64_t K, const double* A,
                    int64_t lda, const double* B, int64_t ldb, double* C, int64_t ldc) {
  return guarded([&] {
    DC2_CHECK(c && A && B && C, 1, "null argument");
    DC2_CHECK(lda >= K && ldb >= K && ldc >= M, 1, "leading dimension too small");
    Lane& L0 = *c->lane0;
    const int64_t wb = zgemm_ozaki_work_bytes(slices, M, N, K);
    int8_t* work = reinterpret_cast<int8_t*>(L0.l("oz_work", (wb + 7) / 8));
    int* expo = L0.i("oz_expo", M + N);
    zgemm_ozaki(c->stream, slices, opB, M, N, K, reinterpret_cast<const cplx*>(A), lda,
                reinterpret_cast<const cplx*>(B), ldb, reinterpret_cast<cplx*>(C), ldc, work, expo);
    DC2_HIP(hipStreamSynchronize(c->stream));
  });
}

int dc2_get_transform_flops(dc2_context* c, double* out) {
  return guarded([&] {
    DC2_CHECK(c && out, 1, "null argument");
    out[0] = c->split_flops;
    out[1] = c->fp64_flops;
  });
}

int dc2_host_heev(int32_t n, double* a, double* w, int32_t method) {
  return guarded([&] {
    DC2_CHECK(a && w && n >= 1 && n <= 4096 && method >= 0 && method <= 4, 1, "invalid argument");
    std::vector<zc> A(reinterpret_cast<zc*>(a), reinterpret_cast<zc*>(a) + (size_t)n * n), V((size_t)n * n);
    if (method == 4) herm_eig_low(n, A.data(), w, V.data(), std::min(n, 4));   // the other columns stay zero
    else if (method == 3) DC2_CHECK(herm_eig_zheev(n, A.data(), w, V.data()), 3, "QL/QR iteration did not converge");
    else if (method == 0) herm_eig(n, A.data(), w, V.data());
    else if (method == 1) herm_jacobi(n, A.data(), w, V.data());
    else DC2_CHECK(herm_tql(n, A.data(), w, V.data()), 3, "QL iteration did not converge");
    std::copy(V.begin(), V.end(), reinterpret_cast<zc*>(a));
  });
}

int dc2_set_subspace_callback(dc2_context* c, void (*cb)(void* user, int32_t subspace, double e2, double sumc2),
                              void* user) {
  return guarded([&] {
    DC2_CHECK(c, 1, "null context");
    c->on_subspace = cb;
    c->on_subspace_user = user;
  });
}

int dc2_write_restart(const char* path, int32_t totsym, int32_t selectroot, int32_t next, const double* e2,
                      const double* sumc2) {
  return guarded([&] {
    DC2_CHECK(path && e2 && sumc2 && next >= 0 && next <= 8, 1, "invalid argument");
    const std::string tmp = std::string(path) + ".tmp";
    FILE* f = std::fopen(tmp.c_str(), "w");
    DC2_CHECK(f != nullptr, 1, std::string("cannot write ") + tmp);
    std::fprintf(f, "totsym: %d\nselectroot: %d\nNext subspace: %c\n", totsym, selectroot, (char)('A' + next));
    for (int k = 0; k < next; ++k)
      std::fprintf(f, "Subspace: %c Sumc2: %.17g Energy: %.17g\n", (char)('A' + k), sumc2[k], e2[k]);
    const bool ok = std::fclose(f) == 0;
    DC2_CHECK(ok && std::rename(tmp.c_str(), path) == 0, 1, std::string("cannot write ") + path);
  });
}

int dc2_fockivo(dc2_context* c, int32_t numh, const double* mo_energy, int32_t nirr, const int32_t* irreps,
                double* fock, double* w, double* U) {
  return guarded([&] {
    DC2_CHECK(c && mo_energy && w && U && (nirr == 0 || irreps), 1, "null argument");
    need_spaces(c);
    DC2_HIP(hipSetDevice(c->device));
    const int ni = c->ni, na = c->na, ns = c->ns, nmoc = ni + na;
    DC2_CHECK(ns > 0 && nmoc > 0, 1, "dc2_fockivo needs secondary and occupied spinors");
    if (numh <= 0) {   // fockivo.F90:40-42: the spinors within 0.1 Eh of the HOMO energy
      DC2_CHECK(c->ne >= 1 && ni + c->ne <= nmoc, 1, "dc2_fockivo: no HOMO (nelec)");
      const double eh = mo_energy[ni + c->ne - 1];
      numh = 0;
      for (int i = 0; i < nmoc; ++i) numh += std::fabs(mo_energy[i] - eh) < 1.0e-1;
    }
    DC2_CHECK(numh >= 1 && numh <= nmoc, 1, "dc2_fockivo: numh out of range");
    const cplx* F1 = get_block(c, "F1");
    const cplx* F2 = get_block(c, "F2");
    Lane& L0 = *c->lane0;
    hipStream_t s = c->stream;
    double* eps_d = L0.d("ivo_eps", ns);
    cplx* fock_d = L0.z("ivo_fock", (int64_t)ns * ns);
    DC2_HIP(hipMemcpyAsync(eps_d, mo_energy + nmoc, ns * sizeof(double), hipMemcpyHostToDevice, s));
    fockivo_build(s, nmoc, ns, numh, c->ne % 2 == 1, eps_d, F1, F2, fock_d);
    std::vector<zc> F((size_t)ns * ns);
    DC2_HIP(hipMemcpyAsync(F.data(), fock_d, F.size() * sizeof(cplx), hipMemcpyDeviceToHost, s));
    DC2_HIP(hipStreamSynchronize(s));
    if (fock) std::copy(F.begin(), F.end(), reinterpret_cast<zc*>(fock));
    zc* Uo = reinterpret_cast<zc*>(U);
    std::fill(Uo, Uo + (size_t)ns * ns, zc(0.0, 0.0));
    std::fill(w, w + ns, 0.0);
    // fockivo.F90:181-230: one block per irreducible representation, ZHEEV's eigenvectors
    for (int q = 0; q < nirr; ++q) {
      std::vector<int> mos;
      for (int a = 0; a < ns; ++a)
        if ((c->irpamo.empty() ? 1 : c->irpamo[nmoc + a]) == irreps[q]) mos.push_back(a);
      const int nv = (int)mos.size();
      if (nv == 0) continue;
      std::vector<zc> B((size_t)nv * nv), V((size_t)nv * nv);
      std::vector<double> wv(nv);
      for (int j = 0; j < nv; ++j)
        for (int i = 0; i < nv; ++i) B[i + (size_t)nv * j] = F[mos[i] + (size_t)ns * mos[j]];
      DC2_CHECK(herm_eig_zheev(nv, B.data(), wv.data(), V.data()), 3, "dc2_fockivo: eigensolver did not converge");
      for (int j = 0; j < nv; ++j) {
        w[mos[j]] = wv[j];
        for (int i = 0; i < nv; ++i) Uo[mos[i] + (size_t)ns * mos[j]] = V[i + (size_t)nv * j];
      }
    }
  });
}

int dc2_get_hetrd_stats(int32_t n, double* out) {
  return guarded([&] {
    DC2_CHECK(out && n >= 0, 1, "invalid argument");
    if (n == 0) hetrd_stats_reset();
    else hetrd_stats(n, out);
  });
}

int dc2_get_casci_info(dc2_context* c, int64_t* out) {
  return guarded([&] {
    DC2_CHECK(c && out, 1, "null argument");
    out[0] = c->ndet;
    out[1] = c->nroots_avail;
    out[2] = c->davidson_iters;
  });
}

int dc2_get_eig_fallbacks(dc2_context* c, int64_t* n) {
  return guarded([&] {
    DC2_CHECK(c && n, 1, "null argument");
    *n = c->eig_fallbacks;
  });
}

int dc2_synchronize(dc2_context* c) {
  return guarded([&] {
    DC2_CHECK(c, 1, "null context");
    DC2_HIP(hipStreamSynchronize(c->stream));
  });
}

}  // extern "C"
