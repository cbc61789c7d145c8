// Device kernels of the solve path (gfx950). Every kernel is launched by the host functions in
// solver.hip; this header only shares the small types both sides need.
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

namespace mlk {

// num / d from rd = RN(1 / d), as the FP64 triangular solves divide each row by its diagonal:
// q = num * rd, then two exact-remainder corrections q = fma(fma(-q, d, num), rd, q). After
// the first, q is within one ulp of num / d; with a faithful q and the correctly rounded
// reciprocal, the second gives the correctly rounded quotient (Markstein's theorem). The
// remainders num - q d are exact while they stay in the normal range, so the result is the
// division's bit for bit for d > 0 with a normal rd, |num| >= 2^-969 and |num / d| < 2^1020,
// and (through the copysign, which keeps the numerator's sign, the quotient's for d > 0) for
// num = +-0. Smaller nonzero numerators and quotients near overflow may differ from the
// division by one ulp: the solves of a positive-definite factor do not produce them, and a
// range test on the row chain, whose latency bounds these solves, costs more than it guards.
__host__ __device__ inline double quotient64(double num, double d, double rd) {
  double q = num * rd;
  double r = fma(-q, d, num);
  q = fma(r, rd, q);
  r = fma(-q, d, num);
  return copysign(fma(r, rd, q), num);
}

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
