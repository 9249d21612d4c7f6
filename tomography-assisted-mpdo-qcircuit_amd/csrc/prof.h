// Launch timing by kernel class with HIP events recorded on each launch's own stream. Off by
// default; mpdo_debug_prof_enable turns it on. Each record carries the launch's algorithmic work
// (flops of the arithmetic it is asked to do, bytes of its operands) for roofline accounting.
#pragma once
#include <hip/hip_runtime.h>

namespace mpdo {
namespace prof {

// EIG_SOLVE: bj_solve; EIG_ROTATE: bj_rotate_v_direct; EIG_UPDATE: bj_update
enum Cls { GEMM = 0, EIG_SOLVE = 1, EIG_ROTATE = 2, EIG_SMALL = 3, PERMUTE = 4, EIG_UPDATE = 5, NCLS = 6 };

bool enabled();
// one record per eigensolver call while timing is on: order, batch, sweeps
void eig_log(int n, int batch, int sweeps);
// work a class's launches actually executed (counted on the host from device flags read back
// while timing is on), next to the work each launch was asked to do (the Scope's figures)
void add_executed(int cls, double flops, double bytes);

// RAII: records a start event on construction and a stop event on destruction; `shape` (five
// ints, e.g. M, N, K, batch, precision of a GEMM) keys a per-shape breakdown of the class
class Scope {
 public:
  Scope(int cls, hipStream_t s, double flops, double bytes, const int* shape = nullptr);
  ~Scope();
  Scope(const Scope&) = delete;
  Scope& operator=(const Scope&) = delete;

 private:
  int cls_;
  hipStream_t s_;
  double flops_, bytes_;
  int shape_[5] = {0, 0, 0, 0, 0};
  bool has_shape_ = false;
  hipEvent_t e0_ = nullptr, e1_ = nullptr;
};

}  // namespace prof
}  // namespace mpdo
