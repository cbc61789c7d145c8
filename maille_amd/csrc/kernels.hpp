// Device kernels of the solve path (gfx950). Every kernel is launched by the host functions in
// solver.hip; this header only shares the small types both sides need.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mlk {

// Values never produced by arithmetic (signalling-NaN payloads): a slot holding one has not been
// written yet. Dataflow kernels initialise their outputs with it and wait for it to change.
constexpr uint64_t SENTINEL64 = 0xFFF4DEADBEEF0001ull;
constexpr uint32_t SENTINEL32 = 0xFFBEEF01u;

// PCG scalars kept on the device so that an iteration never waits for the host.
struct PcgScalars {
  double bnorm;      // |b|
  double rz;         // r.z of the current iteration
  double alpha, beta;
  double rnorm;      // |r| of the last update
  double rscale;     // convergence threshold tol * |b|
  int64_t it;        // completed iterations
  int64_t maxit;
  int32_t done;      // 1 once converged or failed: every later kernel returns at once
  int32_t status;    // 0 ok, 1 singular, 2 not converged, 3 stalled (stall_stop)
  // reliable updates (FP32 operator): the recursive residual is replaced by the FP64 true
  // residual b - A x whenever it has dropped by `delta` since the last replacement
  int32_t update;    // 1 while a replacement is due in the current iteration
  int32_t allow;     // replacements still allowed (cleared once they stop reducing |r|)
  double delta;
  double rmax;       // |r| at the last replacement
  double prev_true;  // true |r| at the last replacement
  int64_t updates;   // replacements done
  int32_t stall_stop;  // 1: stop with status 3 when a replacement fails to halve |b - A x|
  int32_t pad;
};

}  // namespace mlk
