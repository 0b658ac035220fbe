// Cross-request batching of Hermitian eigensolves (used by the request server).
//
// Threads that run requests hand their eigenproblems to one batcher thread, which
// groups the jobs of equal size and algorithm and solves each group with one
// pointer-array batched rocSOLVER call; several executors (each with its own stream)
// run different groups at the same time.  rocSOLVER's eigensolvers are bound by
// kernel launches at these sizes (n <= ~1000), so a batch costs little more than a
// single solve.  A group is flushed when every participating thread is waiting in
// the batcher, when it reaches max_batch, or window seconds after its oldest job.
#pragma once

#include <rocblas/rocblas.h>

#include <chrono>
#include <condition_variable>
#include <deque>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "dc2_common.hpp"

namespace dc2 {

class EigBatcher {
 public:
  EigBatcher(int device, int max_batch, double window_s, int executors = 3);
  ~EigBatcher();
  EigBatcher(const EigBatcher&) = delete;
  EigBatcher& operator=(const EigBatcher&) = delete;
  // Eigendecomposition of the Hermitian matrix A (n x n, upper triangle, on the device,
  // complete: the caller has synchronised the stream that produced it): eigenvectors
  // overwrite A, ascending eigenvalues go to w.  method 1: divide and conquer, 2: divide
  // and conquer with Jacobi leaves.  Blocks until solved; returns the solver's info.
  int solve(int n, int method, cplx* A, double* w);
  // a thread that may call solve() starts / stops running a request
  void enter();
  void leave();
  void stats(int64_t* batches, int64_t* jobs);

 private:
  struct Job {
    int n, method;
    cplx* A;
    double* w;
    int info = 0;
    bool done = false;
    std::chrono::steady_clock::time_point t;
  };
  // one executor: its own stream, rocBLAS handle and staging buffers
  struct Exec {
    hipStream_t stream = nullptr;
    rocblas_handle blas = nullptr;
    DevBuf<cplx*> ptrs;
    DevBuf<double> D, E;
    DevBuf<int> info;
    std::thread th;
  };
  void run(Exec* x);
  void execute(Exec* x, std::vector<Job*>& group);
  int device_, max_batch_;
  double window_;
  std::mutex mu_;
  std::condition_variable cv_jobs_, cv_done_;
  std::deque<Job*> jobs_;
  int participants_ = 0;
  bool stop_ = false;
  int64_t batches_ = 0, solved_ = 0;
  std::vector<std::unique_ptr<Exec>> execs_;
};

}  // namespace dc2
