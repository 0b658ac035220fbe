// Request server: continuous batching of CASPT2 energies on one GPU.
//
// Molecules (spinor spaces + two-electron integrals) are registered once; their
// integral blocks are formed on the device and stay resident in HBM, shared
// read-only by every request on that molecule (points of a potential-energy scan
// differ in their one-electron matrix).  Each of nslots slots owns a dc2_context
// (its streams, rocBLAS handles and grow-only scratch) and a host thread.  A slot's
// context keeps its scratch after a request, so admission counts the memory every slot
// holds (the largest footprint it has run): a free slot takes the oldest pending request as
// soon as the held memory plus the growth that request needs in this slot fits the budget.
// When nothing runs and the oldest request still does not fit, idle slots release their
// memory (each recreates its context).  Requests that arrive while others run are admitted
// as slots and memory free up, and their kernels overlap
// on the device.  No counterpart in the reference (its driver runs one problem per
// process, dcaspt2_run_subprograms.F90:57-62).
#include <hip/hip_runtime.h>

#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include <rocsolver/rocsolver.h>

#include "../../include/dcaspt2_amd.h"
#include "dc2_batch.hpp"

#include <sys/stat.h>

#include <cerrno>

namespace dc2 {
void set_last_error(const std::string& msg);   // dc2_api.hip: the message dc2_last_error() returns
void set_batcher(dc2_context* c, EigBatcher* b);   // dc2_api.hip: route the context's eigensolves
int context_create_lanes(int32_t device, int lanes, dc2_context** out);   // dc2_api.hip: with lane streams

// ---------------------------------------------------------------- EigBatcher
EigBatcher::EigBatcher(int device, int max_batch, double window_s, int executors)
    : device_(device), max_batch_(max_batch), window_(window_s) {
  DC2_HIP(hipSetDevice(device));
  for (int k = 0; k < executors; ++k) {
    auto x = std::make_unique<Exec>();
    DC2_HIP(hipStreamCreateWithFlags(&x->stream, hipStreamNonBlocking));
    if (rocblas_create_handle(&x->blas) != rocblas_status_success) throw Error(3, "rocblas_create_handle failed");
    rocblas_set_stream(x->blas, x->stream);
    execs_.push_back(std::move(x));
  }
  for (auto& x : execs_) x->th = std::thread(&EigBatcher::run, this, x.get());
}

EigBatcher::~EigBatcher() {
  {
    std::lock_guard<std::mutex> lk(mu_);
    stop_ = true;
  }
  cv_jobs_.notify_all();
  for (auto& x : execs_) x->th.join();
  for (auto& x : execs_) {
    (void)hipStreamSynchronize(x->stream);
    x->ptrs.release();
    x->D.release();
    x->E.release();
    x->info.release();
    rocblas_destroy_handle(x->blas);
    release_stream_workspace(x->stream);
    (void)hipStreamDestroy(x->stream);
  }
}

void EigBatcher::enter() {
  std::lock_guard<std::mutex> lk(mu_);
  ++participants_;
}

void EigBatcher::leave() {
  {
    std::lock_guard<std::mutex> lk(mu_);
    --participants_;
  }
  cv_jobs_.notify_all();   // the waiting jobs may now be everyone's
}

void EigBatcher::stats(int64_t* batches, int64_t* jobs) {
  std::lock_guard<std::mutex> lk(mu_);
  *batches = batches_;
  *jobs = solved_;
}

int EigBatcher::solve(int n, int method, cplx* A, double* w) {
  Job j;
  j.n = n;
  j.method = method;
  j.A = A;
  j.w = w;
  j.t = std::chrono::steady_clock::now();
  std::unique_lock<std::mutex> lk(mu_);
  jobs_.push_back(&j);
  cv_jobs_.notify_all();
  cv_done_.wait(lk, [&] { return j.done; });
  return j.info;
}

void EigBatcher::run(Exec* ex) {
  (void)hipSetDevice(device_);
  std::unique_lock<std::mutex> lk(mu_);
  for (;;) {
    cv_jobs_.wait(lk, [&] { return stop_ || !jobs_.empty(); });
    if (jobs_.empty()) return;   // stopping and drained
    // the group of the oldest job, by value: another executor may take it meanwhile.
    // Flush when full, when every participant waits, or late.
    const int key_n = jobs_.front()->n, key_m = jobs_.front()->method;
    const auto deadline = jobs_.front()->t + std::chrono::duration_cast<std::chrono::steady_clock::duration>(
                                                 std::chrono::duration<double>(window_));
    auto group_size = [&] {
      int g = 0;
      for (const Job* x : jobs_) g += (x->n == key_n && x->method == key_m);
      return g;
    };
    cv_jobs_.wait_until(lk, deadline, [&] {
      return stop_ || group_size() >= max_batch_ || (int)jobs_.size() >= participants_;
    });
    std::vector<Job*> group;
    for (auto it = jobs_.begin(); it != jobs_.end() && (int)group.size() < max_batch_;) {
      if ((*it)->n == key_n && (*it)->method == key_m) {
        group.push_back(*it);
        it = jobs_.erase(it);
      } else {
        ++it;
      }
    }
    if (group.empty()) continue;   // taken by another executor
    lk.unlock();
    try {
      execute(ex, group);
    } catch (const std::exception&) {
      for (Job* x : group) x->info = -1;   // reported as a failed solve; the caller falls back
    }
    lk.lock();
    ++batches_;
    solved_ += (int64_t)group.size();
    for (Job* x : group) x->done = true;
    cv_done_.notify_all();
    cv_jobs_.notify_all();
  }
}

void EigBatcher::execute(Exec* x, std::vector<Job*>& group) {
  const int b = (int)group.size(), n = group[0]->n;
  std::vector<cplx*> hp(b);
  for (int k = 0; k < b; ++k) hp[k] = group[k]->A;
  x->ptrs.ensure(max_batch_);
  x->D.ensure((size_t)max_batch_ * n);
  x->E.ensure((size_t)max_batch_ * n);
  x->info.ensure(max_batch_);
  DC2_HIP(hipMemcpyAsync(x->ptrs.get(), hp.data(), b * sizeof(cplx*), hipMemcpyHostToDevice, x->stream));
  rocblas_double_complex* const* A = reinterpret_cast<rocblas_double_complex* const*>(x->ptrs.get());
  rocblas_status st;
  if (group[0]->method == 2)
    st = rocsolver_zheevdj_batched(x->blas, rocblas_evect_original, rocblas_fill_upper, n, A, n, x->D.get(), n,
                                   x->info.get(), b);
  else
    st = rocsolver_zheevd_batched(x->blas, rocblas_evect_original, rocblas_fill_upper, n, A, n, x->D.get(), n,
                                  x->E.get(), n, x->info.get(), b);
  if (st != rocblas_status_success) throw Error(3, "batched eigensolver failed to launch");
  for (int k = 0; k < b; ++k)
    DC2_HIP(hipMemcpyAsync(group[k]->w, x->D.get() + (size_t)k * n, n * sizeof(double), hipMemcpyDeviceToDevice,
                           x->stream));
  std::vector<int> info(b);
  DC2_HIP(hipMemcpyAsync(info.data(), x->info.get(), b * sizeof(int), hipMemcpyDeviceToHost, x->stream));
  DC2_HIP(hipStreamSynchronize(x->stream));
  for (int k = 0; k < b; ++k) group[k]->info = info[k];
}
}  // namespace dc2

namespace {

struct Molecule {
  dc2_context* ctx = nullptr;
  int32_t ni = 0, na = 0, ns = 0, ne = 0;
  int64_t bytes = 0;
};

struct Request {
  int64_t id = 0;
  Molecule* mol = nullptr;
  int32_t mol_id = 0;
  std::vector<double> h;
  double ecore = 0.0, eshift = 0.0;
  int32_t root = 1, flags = 0, slot = -1;
  int32_t ras[7] = {0, 0, 0, 0, 0, 0, 0};
  int64_t footprint = 0;
  double t_submit = 0, t_start = 0, t_end = 0;
  int status = 0;
  std::string err;
  double e2[8] = {0}, sumc2[8] = {0}, scal[6] = {0};
  // resumed from a restart: subspaces before `first` taken from e2_file / sumc2_file
  int32_t first = 0;
  double e2_file[8] = {0}, sumc2_file[8] = {0};
  // checkpoint state (filled by the subspace callback of the slot's context)
  std::string ckpt_path;
  std::mutex ckpt_mu;
  bool sub_done[8] = {false};
  double ck_e2[8] = {0}, ck_sumc2[8] = {0};
  int ck_next = 0;
};

const char* kBlockNames[] = {"OOOO", "F1", "F2", "A1", "A2", "B", "C1", "C2", "C3",
                             "D1",   "D2", "D3", "E", "F",  "G", "H"};

int64_t binom(int n, int k) {
  if (k < 0 || k > n) return 0;
  int64_t r = 1;
  for (int i = 0; i < k; ++i) r = r * (n - i) / (i + 1);
  return r;
}

// device bytes one request holds while it runs (besides the molecule's resident blocks):
// transformed blocks + transform scratch, CI matrix and LU copy, E_pq tables and images,
// density tensors, the A/C metric work and the V matrices, doubled to cover what the sum
// leaves out (lane scratch, split-precision slices, solver workspaces): measured, one
// context with four lanes held 2.2 GB at 192 spinors (sum 1.3 GB) and 8.8 GB at 324
// spinors (sum 7.7 GB)
int64_t footprint_bytes(int64_t ni, int64_t na, int64_t ns, int64_t ne) {
  const int64_t z = 16;
  const int64_t blocks[] = {na * ni * na * na, na * ni * ni * ni, na * ni * na * ni, ns * na * na * na,
                            ns * na * ni * ni, ns * ni * ni * na, ns * ni * na * na, ns * na * na * ni,
                            ns * ni * ni * ni, ns * ni * na * ni, ns * na * ns * na, ns * ni * ns * na,
                            ns * ni * ns * ni};
  int64_t tb = 0, mx = 0;
  for (int64_t b : blocks) { tb += b; mx = b > mx ? b : mx; }
  const int64_t nf = binom((int)na, (int)ne), na2 = na * na, na3 = na2 * na;
  const int64_t mA = binom((int)na, (int)ne + 1), mC = binom((int)na, (int)ne - 1);
  const int64_t ci = 3 * nf * nf;
  const int64_t dens = 4 * nf * na2 + nf * na2 * na2 + 6 * na3 * na3;
  const int64_t ac = 2 * (std::min(mA, na3) * na3 + 2 * std::min(mA, na3) * std::min(mA, na3)) +
                     na3 * (ni + ns) * 3;
  const int64_t other = 8 * (ns + ni) * (ns + ni) + 4 * ns * ns * ni + 4 * ns * ni * ni * na;
  return 2 * (z * (tb + mx + ci + dens + ac + other) + (64ll << 20));
}

}  // namespace

struct dc2_server {
  int32_t device = 0, nslots = 1, lanes = 1;
  std::unique_ptr<dc2::EigBatcher> batcher;   // cross-request eigensolve batching (null: off)
  int64_t budget = 0, in_use = 0, peak = 0, completed = 0;   // in_use: memory the slots hold
  int running = 0;
  std::vector<int64_t> held;   // per slot: the largest footprint it has run (its retained scratch)
  bool stopping = false;
  int64_t next_id = 1;
  std::vector<std::unique_ptr<Molecule>> mols;
  std::vector<dc2_context*> slot_ctx;
  std::vector<Molecule*> slot_mol;
  std::deque<Request*> pending, done;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::vector<std::thread> threads;
  std::string ckpt_dir;   // per-request checkpoint files (empty: off)
  std::vector<std::pair<std::string, int64_t>> options;   // applied to every slot context
  std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
  double now() const { return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(); }
};

namespace {

int fail(int code, const std::string& msg) {
  dc2::set_last_error(msg);
  return code;
}

// point the slot's context at the molecule's resident blocks (borrowed, no copy)
int bind_molecule(dc2_context* c, const Molecule& m) {
  int st = dc2_set_spaces(c, m.ni, m.na, m.ns, m.ne);
  for (const char* name : kBlockNames) {
    if (st) break;
    const double* p = nullptr;
    int64_t n = 0;
    if (dc2_get_integral_block_device(m.ctx, name, &p, &n) != 0) continue;   // not needed for this problem
    st = dc2_set_integral_block(c, name, p, n, 2);
  }
  return st;
}

// subspace k of request r finished: extend the completed prefix and rewrite the checkpoint
void on_subspace_done(void* user, int32_t k, double e2, double sumc2) {
  Request* r = static_cast<Request*>(user);
  std::lock_guard<std::mutex> lk(r->ckpt_mu);
  r->sub_done[k] = true;
  r->ck_e2[k] = e2;
  r->ck_sumc2[k] = sumc2;
  const int before = r->ck_next;
  while (r->ck_next < 8 && r->sub_done[r->ck_next]) ++r->ck_next;
  if (r->ck_next != before) (void)dc2_write_restart(r->ckpt_path.c_str(), 1, r->root, r->ck_next, r->ck_e2, r->ck_sumc2);
}

void run_request(dc2_server* S, int k, Request* r) {
  dc2_context* c = S->slot_ctx[k];
  int st = 0;
  if (!r->ckpt_path.empty()) {
    for (int q = 0; q < 8; ++q)
      if (q < r->first) {   // the resumed prefix is done already
        r->sub_done[q] = true;
        r->ck_e2[q] = r->e2_file[q];
        r->ck_sumc2[q] = r->sumc2_file[q];
      }
    // subspaces the request skips (no inactive or secondary spinors) count as done
    const Molecule& m = *r->mol;
    const bool skip[8] = {m.ni == 0, m.ni == 0, m.ns == 0, m.ni == 0 || m.ns == 0, m.ni == 0 || m.ns == 0,
                          m.ns == 0, m.ni == 0 || m.ns == 0, m.ni == 0 || m.ns == 0};
    for (int q = 0; q < 8; ++q) r->sub_done[q] = r->sub_done[q] || skip[q];
    r->ck_next = 0;
    while (r->ck_next < 8 && r->sub_done[r->ck_next]) ++r->ck_next;
    st = dc2_write_restart(r->ckpt_path.c_str(), 1, r->root, r->ck_next, r->ck_e2, r->ck_sumc2);
    if (!st) st = dc2_set_subspace_callback(c, on_subspace_done, r);
  }
  if (S->slot_mol[k] != r->mol) {
    S->slot_mol[k] = nullptr;
    st = bind_molecule(c, *r->mol);
    if (!st) S->slot_mol[k] = r->mol;
  }
  if (!st) st = dc2_set_ras(c, r->ras[0], r->ras[1], r->ras[2], r->ras[3], r->ras[4], r->ras[5], r->ras[6]);
  if (!st) st = dc2_set_one_electron(c, r->h.data(), r->ecore);
  if (!st) st = dc2_set_option(c, "split_slices", r->flags);
  std::vector<double> eig(r->root);
  int64_t ndet = 0;
  if (!st) st = dc2_casci(c, r->root, eig.data(), &ndet);
  if (!st) {
    st = r->first > 0 ? dc2_caspt2_restart(c, r->root, r->eshift, r->first, r->e2_file, r->sumc2_file, r->e2,
                                           r->sumc2, r->scal)
                      : dc2_caspt2(c, r->root, r->eshift, r->e2, r->sumc2, r->scal);
  }
  r->status = st;
  if (st) r->err = dc2_last_error();
  if (!r->ckpt_path.empty()) (void)dc2_set_subspace_callback(c, nullptr, nullptr);
}

// a fresh context for slot k (its scratch released); the slot's thread only
int renew_context(dc2_server* S, int k) {
  dc2_context_destroy(S->slot_ctx[k]);
  S->slot_ctx[k] = nullptr;
  S->slot_mol[k] = nullptr;
  dc2_context* c = nullptr;
  int st = dc2::context_create_lanes(S->device, S->lanes, &c);
  if (st) {
    if (c) dc2_context_destroy(c);
    return st;
  }
  if (S->batcher) dc2::set_batcher(c, S->batcher.get());
  std::vector<std::pair<std::string, int64_t>> opts;
  {
    std::lock_guard<std::mutex> lk(S->mu);
    opts = S->options;
  }
  for (const auto& o : opts)
    if (!st) st = dc2_set_option(c, o.first.c_str(), o.second);
  if (st) {
    dc2_context_destroy(c);
    return st;
  }
  S->slot_ctx[k] = c;
  return 0;
}

void slot_main(dc2_server* S, int k) {
  (void)hipSetDevice(S->device);
  for (;;) {
    Request* r = nullptr;
    bool trim = false;
    {
      std::unique_lock<std::mutex> lk(S->mu);
      auto growth = [&] { return std::max<int64_t>(0, S->pending.front()->footprint - S->held[k]); };
      S->cv_work.wait(lk, [&] {
        if (!S->pending.empty()) {
          if (S->in_use + growth() <= S->budget) return true;
          return S->running == 0 && S->held[k] > 0;   // stuck: release this slot's memory
        }
        return S->stopping;
      });
      if (S->pending.empty()) return;   // stopping and drained
      const int64_t g = growth();
      if (S->in_use + g <= S->budget) {
        r = S->pending.front();
        S->pending.pop_front();
        S->in_use += g;
        S->held[k] = std::max(S->held[k], r->footprint);
        S->peak = std::max(S->peak, S->in_use);
        ++S->running;
        r->t_start = S->now();
        r->slot = k;
      } else {
        trim = true;
        S->in_use -= S->held[k];
        S->held[k] = 0;
      }
    }
    if (trim) {
      const int st = S->slot_ctx[k] ? renew_context(S, k) : 0;
      if (st || !S->slot_ctx[k]) {   // no context: this slot stops taking requests
        S->cv_work.notify_all();
        return;
      }
      S->cv_work.notify_all();
      continue;
    }
    if (S->batcher) S->batcher->enter();
    run_request(S, k, r);
    if (S->batcher) S->batcher->leave();
    {
      std::lock_guard<std::mutex> lk(S->mu);
      --S->running;
      ++S->completed;
      r->t_end = S->now();
      S->done.push_back(r);
    }
    S->cv_done.notify_all();
    S->cv_work.notify_all();
  }
}

}  // namespace

extern "C" {

int dc2_request_footprint(int32_t ninact, int32_t nact, int32_t nsec, int32_t nelec, int64_t* bytes) {
  if (!bytes || ninact < 0 || nact < 1 || nsec < 0 || nelec < 1 || nelec > nact) return fail(1, "invalid sizes");
  *bytes = footprint_bytes(ninact, nact, nsec, nelec);
  return 0;
}

int dc2_server_create(int32_t device, int32_t nslots, int64_t mem_budget, int32_t lanes_per_slot, dc2_server** out) {
  return dc2_server_create_batched(device, nslots, mem_budget, lanes_per_slot, 0, 0.0, out);
}

int dc2_server_create_batched(int32_t device, int32_t nslots, int64_t mem_budget, int32_t lanes_per_slot,
                              int32_t eig_batch, double eig_window_s, dc2_server** out) {
  if (!out || nslots < 1 || nslots > 256 || mem_budget <= 0 || lanes_per_slot < 1 || lanes_per_slot > 4 ||
      eig_batch < 0 || eig_batch > 256 || eig_window_s < 0.0 || eig_window_s > 1.0)
    return fail(1, "dc2_server_create: invalid argument");
  auto S = std::make_unique<dc2_server>();
  S->device = device;
  S->nslots = nslots;
  S->lanes = lanes_per_slot;
  S->budget = mem_budget;
  if (eig_batch > 1) {
    try {
      S->batcher = std::make_unique<dc2::EigBatcher>(device, eig_batch, eig_window_s);
    } catch (const std::exception& e) {
      return fail(2, std::string("dc2_server_create: ") + e.what());
    }
  }
  for (int k = 0; k < nslots; ++k) {
    dc2_context* c = nullptr;
    int st = dc2::context_create_lanes(device, lanes_per_slot, &c);
    if (st) {
      const std::string msg = dc2_last_error();
      if (c) dc2_context_destroy(c);
      for (auto* x : S->slot_ctx) dc2_context_destroy(x);
      return fail(st, "dc2_server_create: " + msg);
    }
    if (S->batcher) dc2::set_batcher(c, S->batcher.get());
    S->slot_ctx.push_back(c);
    S->slot_mol.push_back(nullptr);
    S->held.push_back(0);
  }
  dc2_server* raw = S.get();
  for (int k = 0; k < nslots; ++k) S->threads.emplace_back(slot_main, raw, k);
  *out = S.release();
  return 0;
}

int dc2_server_destroy(dc2_server* S) {
  if (!S) return 0;
  {
    std::lock_guard<std::mutex> lk(S->mu);
    S->stopping = true;
  }
  S->cv_work.notify_all();
  for (auto& t : S->threads) t.join();
  for (auto* c : S->slot_ctx) dc2_context_destroy(c);
  for (auto& m : S->mols) dc2_context_destroy(m->ctx);
  S->batcher.reset();
  for (auto* r : S->done) delete r;
  delete S;
  return 0;
}

int dc2_server_add_molecule(dc2_server* S, int32_t ninact, int32_t nact, int32_t nsec, int32_t nelec, const double* B,
                            int64_t naux, int32_t* mol_id) {
  if (!S || !B || !mol_id || naux < 1) return fail(1, "dc2_server_add_molecule: invalid argument");
  auto m = std::make_unique<Molecule>();
  m->ni = ninact; m->na = nact; m->ns = nsec; m->ne = nelec;
  int st = dc2::context_create_lanes(S->device, 1, &m->ctx);   // holds the blocks only
  if (!st) st = dc2_set_spaces(m->ctx, ninact, nact, nsec, nelec);
  if (!st) st = dc2_set_integral_factors(m->ctx, B, naux, 0);
  if (st) {
    const std::string msg = dc2_last_error();
    if (m->ctx) dc2_context_destroy(m->ctx);
    return fail(st, "dc2_server_add_molecule: " + msg);
  }
  for (const char* name : kBlockNames) {
    const double* p = nullptr;
    int64_t n = 0;
    if (dc2_get_integral_block_device(m->ctx, name, &p, &n) == 0) m->bytes += 16 * n;
  }
  std::lock_guard<std::mutex> lk(S->mu);
  *mol_id = (int32_t)S->mols.size();
  S->mols.push_back(std::move(m));
  return 0;
}

int dc2_server_set_option(dc2_server* S, const char* key, int64_t value) {
  if (!S || !key) return fail(1, "dc2_server_set_option: invalid argument");
  std::lock_guard<std::mutex> lk(S->mu);
  for (dc2_context* c : S->slot_ctx) {
    if (!c) continue;
    const int st = dc2_set_option(c, key, value);
    if (st) return st;
  }
  S->options.emplace_back(key, value);
  return 0;
}

int dc2_server_set_checkpoint_dir(dc2_server* S, const char* dir) {
  if (!S) return fail(1, "dc2_server_set_checkpoint_dir: null server");
  std::lock_guard<std::mutex> lk(S->mu);
  S->ckpt_dir = dir ? dir : "";
  return 0;
}

int dc2_server_submit(dc2_server* S, int32_t mol_id, const double* h, double ecore, int32_t selectroot, double eshift,
                      const int32_t* ras, int32_t flags, int64_t* req_id) {
  return dc2_server_submit_restart(S, mol_id, h, ecore, selectroot, eshift, ras, flags, 0, nullptr, nullptr, req_id);
}

int dc2_server_submit_restart(dc2_server* S, int32_t mol_id, const double* h, double ecore, int32_t selectroot,
                              double eshift, const int32_t* ras, int32_t flags, int32_t first, const double* e2_file,
                              const double* sumc2_file, int64_t* req_id) {
  if (!S || !h || !req_id || selectroot < 1 || !(flags == 0 || (flags >= 3 && flags <= 8)) || first < 0 ||
      first > 8 || (first > 0 && (!e2_file || !sumc2_file)))
    return fail(1, "dc2_server_submit: invalid argument");
  auto r = std::make_unique<Request>();
  {
    std::lock_guard<std::mutex> lk(S->mu);
    if (mol_id < 0 || mol_id >= (int32_t)S->mols.size()) return fail(1, "dc2_server_submit: unknown molecule");
    r->mol = S->mols[mol_id].get();
  }
  const Molecule& m = *r->mol;
  const int64_t nmo = (int64_t)m.ni + m.na + m.ns;
  r->mol_id = mol_id;
  r->h.assign(h, h + 2 * nmo * nmo);
  r->ecore = ecore;
  r->root = selectroot;
  r->eshift = eshift;
  r->flags = flags;
  if (ras) std::memcpy(r->ras, ras, sizeof(r->ras));
  r->first = first;
  if (first > 0) {
    std::memcpy(r->e2_file, e2_file, sizeof(r->e2_file));
    std::memcpy(r->sumc2_file, sumc2_file, sizeof(r->sumc2_file));
  }
  r->footprint = footprint_bytes(m.ni, m.na, m.ns, m.ne);
  if (r->footprint > S->budget) return fail(1, "dc2_server_submit: request exceeds the memory budget");
  {
    std::lock_guard<std::mutex> lk(S->mu);
    r->id = S->next_id++;
    if (!S->ckpt_dir.empty()) {
      const std::string d = S->ckpt_dir + "/" + std::to_string(r->id);
      if (::mkdir(d.c_str(), 0755) != 0 && errno != EEXIST) return fail(1, "dc2_server_submit: cannot create " + d);
      r->ckpt_path = d + "/caspt2_restart_1_" + std::to_string(selectroot);
    }
    r->t_submit = S->now();
    *req_id = r->id;
    S->pending.push_back(r.release());
  }
  S->cv_work.notify_all();   // every free slot re-checks (the one that fits may not be the first)
  return 0;
}

int dc2_server_wait(dc2_server* S, double timeout_s, int64_t* req_id, int32_t* req_status, double* out) {
  if (!S || !req_id || !req_status || !out) return fail(1, "dc2_server_wait: invalid argument");
  Request* r = nullptr;
  {
    std::unique_lock<std::mutex> lk(S->mu);
    const auto until = std::chrono::steady_clock::now() + std::chrono::duration<double>(std::max(0.0, timeout_s));
    S->cv_done.wait_until(lk, until, [&] { return !S->done.empty(); });
    if (S->done.empty()) {
      *req_id = 0;
      return 0;
    }
    r = S->done.front();
    S->done.pop_front();
  }
  *req_id = r->id;
  *req_status = r->status;
  std::memcpy(out, r->e2, 8 * sizeof(double));
  std::memcpy(out + 8, r->sumc2, 8 * sizeof(double));
  std::memcpy(out + 16, r->scal, 6 * sizeof(double));
  out[22] = r->t_submit;
  out[23] = r->t_start;
  out[24] = r->t_end;
  out[25] = r->slot;
  if (r->status) dc2::set_last_error(r->err);
  delete r;
  return 0;
}

int dc2_server_stats(dc2_server* S, int64_t* out) {
  if (!S || !out) return fail(1, "dc2_server_stats: invalid argument");
  std::lock_guard<std::mutex> lk(S->mu);
  out[0] = (int64_t)S->pending.size();
  out[1] = S->running;
  out[2] = S->completed;
  out[3] = S->in_use;
  out[4] = S->peak;
  out[5] = out[6] = 0;
  if (S->batcher) S->batcher->stats(&out[5], &out[6]);
  return 0;
}

}  // extern "C"
