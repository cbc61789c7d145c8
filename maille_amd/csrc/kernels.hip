// Device kernels of the 3D linear-elastic solve path, written for gfx950 (CDNA4, 64-wide
// wavefronts). Included once, by solver.hip.
//
// Floating-point contraction is switched off in the kernels that restate the reference's
// arithmetic (element stiffness, assembly, SpMV, IC(0), triangular solves, PCG updates): the
// reference is Rust, which never fuses a*b+c, and these kernels reproduce its rounding
// operation for operation. Only the reductions (dot products) use a different summation order.
#pragma once

#include "kernels.hpp"

namespace mlk {

// ------------------------------------------------------------------ small device helpers
__device__ inline uint64_t bits_of(double v) { return __double_as_longlong(v); }
__device__ inline uint32_t bits_of(float v) { return __float_as_uint(v); }
__device__ inline bool is_sentinel(double v) { return bits_of(v) == SENTINEL64; }
__device__ inline bool is_sentinel(float v) { return bits_of(v) == SENTINEL32; }

template <typename T> struct Bits;
template <> struct Bits<double> { typedef unsigned long long U; static constexpr U S = SENTINEL64; };
template <> struct Bits<float> { typedef unsigned int U; static constexpr U S = SENTINEL32; };

template <typename T>
__device__ inline T from_bits(typename Bits<T>::U u);
template <>
__device__ inline double from_bits<double>(unsigned long long u) { return __longlong_as_double(u); }
template <>
__device__ inline float from_bits<float>(unsigned int u) { return __uint_as_float(u); }

// A store other wavefronts of this kernel may wait for: device-scope, so it is written through
// to the point of coherence shared by every XCD.
template <typename T>
__device__ inline void publish(T* p, T v) {
  typedef typename Bits<T>::U U;
  __hip_atomic_store(reinterpret_cast<U*>(p), (U)bits_of(v), __ATOMIC_RELAXED,
                     __HIP_MEMORY_SCOPE_AGENT);
}

// Value of a slot written once per launch by publish(), polled with device-scope loads (a
// plain load may hit a stale line of this XCD's L2 that still holds the sentinel). The wait
// is bounded: past `limit` ticks of the device's steady wall clock (or once another wave has
// given up) it raises *abort and returns 0, so every wavefront reaches the end of the kernel.
// The clock and the abort flag are read every 16th poll only: each is another round trip, and
// polls are on the critical path of the dataflow kernels.
template <typename T>
__device__ inline T await_value(T* p, int* abort_flag, long long limit) {
  typedef typename Bits<T>::U U;
  long long t0 = -1;
  for (int it = 0;; ++it) {
    U u = __hip_atomic_load(reinterpret_cast<U*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (u != Bits<T>::S) return from_bits<T>(u);
    if ((it & 15) == 0) {
      if (__hip_atomic_load(abort_flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return T(0);
      const long long now = wall_clock64();
      if (t0 < 0) t0 = now;
      if (now - t0 > limit) {
        __hip_atomic_store(abort_flag, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        return T(0);
      }
    }
    __builtin_amdgcn_s_sleep(2);
  }
}

// Tagged FP64 levels: a value published for other workgroups carries an 8-bit launch tag in
// the low bits of its significand, so that a slot needs no reset between launches (a reader
// waits for its own launch's tag instead of for the end of a sentinel). The value itself is
// level_value(v), v with those bits cleared: v rounded toward zero to 44 fraction bits, a
// deterministic rounding every reader and writer applies alike.
constexpr unsigned long long LEVEL_TAG_MASK = 0xffull;
__device__ inline double level_value(double v) {
  return __longlong_as_double((long long)(bits_of(v) & ~LEVEL_TAG_MASK));
}

// Value of a tagged level slot once it holds `tag`, polled like await_value (bounded; 0 on
// giving up), with the tag cleared.
__device__ inline double await_tagged(double* p, unsigned tag, int* abort_flag, long long limit) {
  typedef unsigned long long U;
  long long t0 = -1;
  for (int it = 0;; ++it) {
    U u = __hip_atomic_load(reinterpret_cast<U*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if ((u & LEVEL_TAG_MASK) == tag) return __longlong_as_double((long long)(u & ~LEVEL_TAG_MASK));
    if ((it & 15) == 0) {
      if (__hip_atomic_load(abort_flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return 0.0;
      const long long now = wall_clock64();
      if (t0 < 0) t0 = now;
      if (now - t0 > limit) {
        __hip_atomic_store(abort_flag, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        return 0.0;
      }
    }
    __builtin_amdgcn_s_sleep(2);
  }
}

// Co-residency probe of a persistent kernel's grid (a launch that does nothing else): thread
// 0 of every workgroup counts it in probe[0], then waits, at most `limit` ticks of the steady
// wall clock, until all gridDim.x workgroups have arrived; one that gives up sets probe[1]. The
// other waves wait at the barrier meanwhile, so each workgroup holds the resources of the
// kernel that calls this (its registers and LDS) while it is counted. Every wait ends, so the
// launch always drains: workgroups that were not resident run once others have exited.
__device__ inline void residency_probe(int* probe, long long limit) {
  if (threadIdx.x == 0) {
    __hip_atomic_fetch_add(probe, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const long long t0 = wall_clock64();
    while (__hip_atomic_load(probe, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) <
           (int)gridDim.x) {
      if (wall_clock64() - t0 > limit) {
        __hip_atomic_store(probe + 1, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        break;
      }
      __builtin_amdgcn_s_sleep(8);
    }
  }
  __syncthreads();
}

__device__ inline double lane_broadcast(double v, int lane) {
  int lo = __builtin_amdgcn_readlane((int)(__double_as_longlong(v) & 0xffffffffll), lane);
  int hi = __builtin_amdgcn_readlane((int)(__double_as_longlong(v) >> 32), lane);
  return __longlong_as_double(((long long)(unsigned)hi << 32) | (unsigned)lo);
}
__device__ inline float lane_broadcast(float v, int lane) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}

// Orders the calling wave's LDS accesses before and after this point with respect to the other
// lanes of the same wave (the wave-level counterpart of __syncthreads).
__device__ inline void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Deterministic block sum (fixed tree), result valid in thread 0.
template <int BS>
__device__ inline double block_sum(double v) {
  __shared__ double red[BS];
  red[threadIdx.x] = v;
  __syncthreads();
#pragma unroll
  for (int s = BS / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  double r = red[0];
  __syncthreads();
  return r;
}

// ------------------------------------------------------------------ fills and copies
template <typename T>
__global__ void k_fill_bits(T* p, int64_t n, typename Bits<T>::U value) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) reinterpret_cast<typename Bits<T>::U*>(p)[i] = value;
}

__global__ void k_fill_f64(double* p, int64_t n, double v) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

__global__ void k_fill_i32(int* p, int64_t n, int v) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// p[i] = i (the identity connectivity of a batch of separate elements)
__global__ void k_iota_i32(int* p, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = (int)i;
}

// Element matrices from the entry-major layout of k_stiffness (slot j * ne + e) to one
// row-major m x m block per element (out[e * mm + j]), mm = m * m
__global__ void k_entry_to_element_major(int64_t ne, int mm, const double* __restrict__ in,
                                         double* __restrict__ out) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < ne * mm) {
    const int64_t e = t / mm, j = t - e * mm;
    out[t] = in[j * ne + e];
  }
}

// Values uniform in [-1, 1) from SplitMix64 of seed + index: a fixed pseudo-random vector.
__global__ void k_fill_hash(double* p, int64_t n, unsigned long long seed) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    unsigned long long z = seed + 0x9E3779B97F4A7C15ull * (unsigned long long)(i + 1);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    p[i] = (double)(z >> 11) * (2.0 / 9007199254740992.0) - 1.0;
  }
}

template <typename S, typename D>
__global__ void k_gather(int64_t n, const int* __restrict__ src, const S* __restrict__ from,
                         D* __restrict__ to) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    int s = src[i];
    to[i] = s >= 0 ? (D)from[s] : D(0);
  }
}

// 1 / (double)d, correctly rounded, of each diagonal entry of a triangular-solve layout
template <typename T>
__global__ void k_recip64(int64_t n, const T* __restrict__ d, double* __restrict__ r) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) r[i] = 1.0 / (double)d[i];
}

template <typename S, typename D>
__global__ void k_convert(int64_t n, const S* __restrict__ from, D* __restrict__ to) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) to[i] = (D)from[i];
}

// ------------------------------------------------------------------ element stiffness
// Tet shape functions (shape.rs:95-153) and Mapping3d (shape.rs:157-190) at one Gauss point;
// g receives dN/dx. Returns det J.
template <int NEN>
__device__ inline double tet_mapping(const double (&x)[NEN][3], double xi, double eta,
                                     double zeta, double (&g)[NEN][3]) {
#pragma clang fp contract(off)
  double dn[NEN][3];
  const double dl[4][3] = {{-1.0, -1.0, -1.0}, {1.0, 0.0, 0.0}, {0.0, 1.0, 0.0}, {0.0, 0.0, 1.0}};
  if constexpr (NEN == 4) {
    for (int a = 0; a < 4; ++a)
      for (int c = 0; c < 3; ++c) dn[a][c] = dl[a][c];
  } else {
    const double l[4] = {1.0 - xi - eta - zeta, xi, eta, zeta};
    const int ed[6][2] = {{0, 1}, {1, 2}, {2, 0}, {0, 3}, {1, 3}, {2, 3}};
    for (int a = 0; a < 4; ++a)
      for (int c = 0; c < 3; ++c) dn[a][c] = (4.0 * l[a] - 1.0) * dl[a][c];
    for (int k = 0; k < 6; ++k)
      for (int c = 0; c < 3; ++c)
        dn[4 + k][c] = 4.0 * (l[ed[k][0]] * dl[ed[k][1]][c] + l[ed[k][1]] * dl[ed[k][0]][c]);
  }
  double j[3][3] = {{0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}};
  for (int k = 0; k < NEN; ++k)
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) j[r][c] += dn[k][r] * x[k][c];
  double cof[3][3];
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) {
      int r1 = (r + 1) % 3, r2 = (r + 2) % 3, c1 = (c + 1) % 3, c2 = (c + 2) % 3;
      cof[r][c] = j[r1][c1] * j[r2][c2] - j[r1][c2] * j[r2][c1];
    }
  double det = 0.0;
  for (int c = 0; c < 3; ++c) det += j[0][c] * cof[0][c];
  double inv[3][3];
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) inv[r][c] = cof[c][r] / det;
  for (int k = 0; k < NEN; ++k)
    for (int r = 0; r < 3; ++r) {
      double s = 0.0;
      for (int c = 0; c < 3; ++c) s += inv[r][c] * dn[k][c];
      g[k][r] = s;
    }
  return det;
}

// Column i of B (element.rs:209-220), Voigt (xx, yy, zz, yz, xz, xy).
template <int NEN>
__device__ inline void b_column(const double (&g)[NEN][3], int i, double (&col)[6]) {
  // the node's gradient picked by selection: a runtime index into g would move the whole
  // array to scratch memory
  const int node = i / 3;
  double gx = 0.0, gy = 0.0, gz = 0.0;
#pragma unroll
  for (int a = 0; a < NEN; ++a)
    if (a == node) {
      gx = g[a][0];
      gy = g[a][1];
      gz = g[a][2];
    }
  const int c = i % 3;
  col[0] = c == 0 ? gx : 0.0;
  col[1] = c == 1 ? gy : 0.0;
  col[2] = c == 2 ? gz : 0.0;
  col[3] = c == 0 ? 0.0 : (c == 1 ? gz : gy);
  col[4] = c == 0 ? gz : (c == 1 ? 0.0 : gx);
  col[5] = c == 0 ? gy : (c == 1 ? gx : 0.0);
}

// Row `a` of a per-node array picked by selection (no runtime index into register arrays).
template <int NEN>
__device__ inline void pick_row(const double (&g)[NEN][3], int a, double (&out)[3]) {
  out[0] = out[1] = out[2] = 0.0;
#pragma unroll
  for (int k = 0; k < NEN; ++k)
    if (k == a) {
      out[0] = g[k][0];
      out[1] = g[k][1];
      out[2] = g[k][2];
    }
}
template <int NEN>
__device__ inline double pick(const double (&v)[NEN], int a) {
  double out = 0.0;
#pragma unroll
  for (int k = 0; k < NEN; ++k)
    if (k == a) out = v[k];
  return out;
}

// solid_d (element.rs:193-206)
__device__ inline void solid_d(double young, double nu, double (&d)[6][6]) {
#pragma clang fp contract(off)
  double c = young / ((1.0 + nu) * (1.0 - 2.0 * nu));
  double gm = young / (2.0 * (1.0 + nu));
  for (int r = 0; r < 6; ++r)
    for (int k = 0; k < 6; ++k) d[r][k] = 0.0;
  for (int r = 0; r < 3; ++r)
    for (int k = 0; k < 3; ++k) d[r][k] = (r == k) ? c * (1.0 - nu) : c * nu;
  for (int r = 3; r < 6; ++r) d[r][r] = gm;
}

template <int NEN> struct Rule;
template <> struct Rule<4> { static constexpr int N = 1; };
template <> struct Rule<10> { static constexpr int N = 4; };

// Gauss points: TETRA_1 (quadrature.rs:100-105) for Tet4, TETRA_4 (:108-138) for Tet10.
__device__ inline void gauss_point(int nen, int q, double& xi, double& eta, double& zeta,
                                   double& w) {
  const double A = 0.585410196624968500, B = 0.138196601125010500;
  if (nen == 4) { xi = eta = zeta = 0.25; w = 1.0 / 6.0; return; }
  xi = q == 1 ? A : B;
  eta = q == 2 ? A : B;
  zeta = q == 3 ? A : B;
  w = 1.0 / 24.0;
}

// solid_stiffness (element.rs:223-243): one thread computes column i of Ke for element e,
// i.e. ke[j][i] for every j, accumulated over the Gauss points in the reference's order.
// Output slot of ke[j][i] is (j*M + i) * ne + e (entry-major, elements contiguous).
template <int NEN>
__global__ __launch_bounds__(256) void k_stiffness(int64_t ne, const double* __restrict__ xyz,
                                                   const int* __restrict__ conn,
                                                   const double* __restrict__ young,
                                                   const double* __restrict__ nu,
                                                   double* __restrict__ ke) {
#pragma clang fp contract(off)
  constexpr int M = 3 * NEN;
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)M * ne) return;
  const int i = (int)(t / ne);
  const int64_t e = t - (int64_t)i * ne;
  double x[NEN][3];
  for (int a = 0; a < NEN; ++a) {
    int node = conn[e * NEN + a];
    for (int c = 0; c < 3; ++c) x[a][c] = xyz[3 * (int64_t)node + c];
  }
  double d[6][6];
  solid_d(young[e], nu[e], d);
  double acc[M];
  for (int j = 0; j < M; ++j) acc[j] = 0.0;
  for (int q = 0; q < Rule<NEN>::N; ++q) {
    double xi, eta, zeta, w;
    gauss_point(NEN, q, xi, eta, zeta, w);
    double g[NEN][3];
    double det = tet_mapping<NEN>(x, xi, eta, zeta, g);
    double factor = w * fabs(det);
    double bi[6], db[6];
    b_column<NEN>(g, i, bi);
    for (int r = 0; r < 6; ++r) {
      double s = 0.0;
      for (int c = 0; c < 6; ++c) s += d[r][c] * bi[c];
      db[r] = s;
    }
    for (int j = 0; j < M; ++j) {
      double bj[6];
      b_column<NEN>(g, j, bj);
      double s = 0.0;
      for (int r = 0; r < 6; ++r) s += bj[r] * db[r];
      acc[j] += factor * s;
    }
  }
  for (int j = 0; j < M; ++j) ke[((int64_t)j * M + i) * ne + e] = acc[j];
}

// tet_mapping for element e at one Gauss point, with the same operations in the same order,
// but coordinates read as they are used and the gradients stored to g_out as they are formed
// (fewer live registers); returns det J.
template <int NEN, typename Out>
__device__ inline double tet_mapping_to(const double* __restrict__ xyz, const int* __restrict__ en,
                                        double xi, double eta, double zeta, Out&& g_out) {
#pragma clang fp contract(off)
  double dn[NEN][3];
  const double dl[4][3] = {{-1.0, -1.0, -1.0}, {1.0, 0.0, 0.0}, {0.0, 1.0, 0.0}, {0.0, 0.0, 1.0}};
  if constexpr (NEN == 4) {
    for (int a = 0; a < 4; ++a)
      for (int c = 0; c < 3; ++c) dn[a][c] = dl[a][c];
  } else {
    const double l[4] = {1.0 - xi - eta - zeta, xi, eta, zeta};
    const int ed[6][2] = {{0, 1}, {1, 2}, {2, 0}, {0, 3}, {1, 3}, {2, 3}};
    for (int a = 0; a < 4; ++a)
      for (int c = 0; c < 3; ++c) dn[a][c] = (4.0 * l[a] - 1.0) * dl[a][c];
    for (int k = 0; k < 6; ++k)
      for (int c = 0; c < 3; ++c)
        dn[4 + k][c] = 4.0 * (l[ed[k][0]] * dl[ed[k][1]][c] + l[ed[k][1]] * dl[ed[k][0]][c]);
  }
  double j[3][3] = {{0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}};
  for (int k = 0; k < NEN; ++k) {
    const int64_t node = en[k];
    const double xk[3] = {xyz[3 * node], xyz[3 * node + 1], xyz[3 * node + 2]};
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) j[r][c] += dn[k][r] * xk[c];
  }
  double cof[3][3];
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) {
      int r1 = (r + 1) % 3, r2 = (r + 2) % 3, c1 = (c + 1) % 3, c2 = (c + 2) % 3;
      cof[r][c] = j[r1][c1] * j[r2][c2] - j[r1][c2] * j[r2][c1];
    }
  double det = 0.0;
  for (int c = 0; c < 3; ++c) det += j[0][c] * cof[0][c];
  double inv[3][3];
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) inv[r][c] = cof[c][r] / det;
  for (int k = 0; k < NEN; ++k)
    for (int r = 0; r < 3; ++r) {
      double s = 0.0;
      for (int c = 0; c < 3; ++c) s += inv[r][c] * dn[k][c];
      g_out(k, r, s);
    }
  return det;
}

// solid_stiffness (element.rs:223-243) with each Gauss point's mapping computed once per
// element: a workgroup takes EB consecutive elements. First one thread per (element, Gauss
// point) computes the shape gradients dN/dx and w |det J| (tet_mapping) and, for the first
// point, D (solid_d), into LDS; then one thread per (column i, half of the rows j, element)
// accumulates ke[j][i] += w |det J| (b_j . D b_i) over the Gauss points from LDS, with the same
// functions, the same zero entries of B and the same order as k_stiffness, so the results are
// identical bit for bit (each half recomputes D b_i, which keeps 15 sums live instead of 30).
// Consecutive lanes take consecutive elements, so that every store of slot
// (j*M + i) * ne + e covers EB contiguous words.
template <int NEN, int EB = 16>
__global__ __launch_bounds__(256) void k_stiffness_lds(int64_t ne, const double* __restrict__ xyz,
                                                       const int* __restrict__ conn,
                                                       const double* __restrict__ young,
                                                       const double* __restrict__ nu,
                                                       double* __restrict__ ke) {
#pragma clang fp contract(off)
  constexpr int M = 3 * NEN, NG = Rule<NEN>::N;
  static_assert(EB * NG <= 256, "one thread per (element, Gauss point)");
  __shared__ double sg[EB][NG][NEN][3];
  __shared__ double sf[EB][NG];
  __shared__ double sd[EB][6][6];
  const int64_t e0 = (int64_t)blockIdx.x * EB;
  const int t = threadIdx.x;
  if (t < EB * NG) {
    const int el = t / NG, q = t - el * NG;
    const int64_t e = e0 + el;
    if (e < ne) {
      double xi, eta, zeta, w;
      gauss_point(NEN, q, xi, eta, zeta, w);
      const double det = tet_mapping_to<NEN>(xyz, conn + e * NEN, xi, eta, zeta,
                                             [&](int a, int c, double v) { sg[el][q][a][c] = v; });
      sf[el][q] = w * fabs(det);
      if (q == 0) {
        double d[6][6];
        solid_d(young[e], nu[e], d);
        for (int r = 0; r < 6; ++r)
          for (int c = 0; c < 6; ++c) sd[el][r][c] = d[r][c];
      }
    }
  }
  __syncthreads();
  // column i of B from the staged gradients, as b_column builds it
  auto column = [&](int el, int q, int i, double (&col)[6]) {
    const int node = i / 3, c = i % 3;
    const double gx = sg[el][q][node][0], gy = sg[el][q][node][1], gz = sg[el][q][node][2];
    col[0] = c == 0 ? gx : 0.0;
    col[1] = c == 1 ? gy : 0.0;
    col[2] = c == 2 ? gz : 0.0;
    col[3] = c == 0 ? 0.0 : (c == 1 ? gz : gy);
    col[4] = c == 0 ? gz : (c == 1 ? 0.0 : gx);
    col[5] = c == 0 ? gy : (c == 1 ? gx : 0.0);
  };
  constexpr int H = M / 2;  // rows j per task
  static_assert(M % 2 == 0, "two halves of the rows");
  for (int task = t; task < EB * M * 2; task += blockDim.x) {
    const int el = task % EB, rest = task / EB;
    const int i = rest >> 1, j0 = (rest & 1) * H;
    const int64_t e = e0 + el;
    if (e >= ne) continue;
    double acc[H];
    for (int jj = 0; jj < H; ++jj) acc[jj] = 0.0;
    for (int q = 0; q < NG; ++q) {
      const double factor = sf[el][q];
      double bi[6], db[6];
      column(el, q, i, bi);
      for (int r = 0; r < 6; ++r) {
        double s = 0.0;
        for (int c = 0; c < 6; ++c) s += sd[el][r][c] * bi[c];
        db[r] = s;
      }
      for (int jj = 0; jj < H; ++jj) {
        double bj[6];
        column(el, q, j0 + jj, bj);
        double s = 0.0;
        for (int r = 0; r < 6; ++r) s += bj[r] * db[r];
        acc[jj] += factor * s;
      }
    }
    for (int jj = 0; jj < H; ++jj) ke[((int64_t)(j0 + jj) * M + i) * ne + e] = acc[jj];
  }
}

// solid_stresses (element.rs:270-302) at Gauss point q of element e, thread per (e, q):
// x, y, z, sxx, syy, szz, tyz, txz, txy.
template <int NEN>
__global__ __launch_bounds__(256) void k_stresses(int64_t ne, const double* __restrict__ xyz,
                                                  const int* __restrict__ conn,
                                                  const double* __restrict__ young,
                                                  const double* __restrict__ nu,
                                                  const double* __restrict__ u,
                                                  double* __restrict__ out) {
#pragma clang fp contract(off)
  constexpr int M = 3 * NEN, NG = Rule<NEN>::N;
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ne * NG) return;
  const int64_t e = t / NG;
  const int q = (int)(t - e * NG);
  double x[NEN][3], ue[M];
  for (int a = 0; a < NEN; ++a) {
    int64_t node = conn[e * NEN + a];
    for (int c = 0; c < 3; ++c) {
      x[a][c] = xyz[3 * node + c];
      ue[3 * a + c] = u[3 * node + c];
    }
  }
  double d[6][6];
  solid_d(young[e], nu[e], d);
  double xi, eta, zeta, w;
  gauss_point(NEN, q, xi, eta, zeta, w);
  double g[NEN][3];
  tet_mapping<NEN>(x, xi, eta, zeta, g);
  double eps[6];
  for (int r = 0; r < 6; ++r) {
    double s = 0.0;
    for (int k = 0; k < M; ++k) {
      double col[6];
      b_column<NEN>(g, k, col);
      s += col[r] * ue[k];
    }
    eps[r] = s;
  }
  // shape values for the Gauss-point position
  double n[NEN];
  {
    const double l[4] = {1.0 - xi - eta - zeta, xi, eta, zeta};
    if constexpr (NEN == 4) {
      for (int a = 0; a < 4; ++a) n[a] = l[a];
    } else {
      const int ed[6][2] = {{0, 1}, {1, 2}, {2, 0}, {0, 3}, {1, 3}, {2, 3}};
      for (int a = 0; a < 4; ++a) n[a] = l[a] * (2.0 * l[a] - 1.0);
      for (int k = 0; k < 6; ++k) n[4 + k] = 4.0 * l[ed[k][0]] * l[ed[k][1]];
    }
  }
  double* o = out + t * 9;
  for (int k = 0; k < 3; ++k) {
    double s = 0.0;
    for (int a = 0; a < NEN; ++a) s += n[a] * x[a][k];
    o[k] = s;
  }
  for (int r = 0; r < 6; ++r) {
    double s = 0.0;
    for (int c = 0; c < 6; ++c) s += d[r][c] * eps[c];
    o[3 + r] = s;
  }
}

// ------------------------------------------------------------------ elasto-plastic response
// VonMises::update (plasticity.rs:45-112): closed-form radial return from the state (ep, p) at
// the start of the step, for the total strain `strain`. Writes the stress, the end state and
// the consistent tangent dsigma/deps; returns whether the point yielded. Infinite yield stress:
// elastic. Shear strains are engineering shears, shear stresses tensor components.
__device__ inline bool von_mises_update(double young, double poisson, double yield0,
                                        double hardening, const double (&strain)[6],
                                        const double (&ep0)[6], double p0, double (&stress)[6],
                                        double (&ep)[6], double& p, double (&c)[6][6]) {
#pragma clang fp contract(off)
  const double g = young / (2.0 * (1.0 + poisson));
  const double bulk = young / (3.0 * (1.0 - 2.0 * poisson));
  double el[6];
  for (int k = 0; k < 6; ++k) el[k] = strain[k] - ep0[k];
  const double volume = el[0] + el[1] + el[2];
  double trial[6];
  for (int k = 0; k < 6; ++k) trial[k] = k < 3 ? 2.0 * g * (el[k] - volume / 3.0) : g * el[k];
  const double root_3_2 = sqrt(1.5);
  double contract = 0.0;  // s : s, shears counted twice (plasticity.rs:115-126)
  for (int k = 0; k < 6; ++k) contract += k < 3 ? trial[k] * trial[k] : 2.0 * trial[k] * trial[k];
  const double norm = sqrt(contract);
  const double q = root_3_2 * norm;
  const double ys = yield0 + hardening * p0;
  const double excess = q - ys;
  double theta = 1.0, theta_bar = 0.0, dp = 0.0;
  if (excess > 1e-10 * ys) {  // YIELD_TOLERANCE (plasticity.rs:11)
    dp = excess / (3.0 * g + hardening);
    theta = 1.0 - 3.0 * g * dp / q;
    theta_bar = 3.0 * g / (3.0 * g + hardening) - (1.0 - theta);
  }
  double n[6];
  for (int k = 0; k < 6; ++k) n[k] = dp > 0.0 ? trial[k] / norm : 0.0;
  for (int k = 0; k < 6; ++k) {
    const double tensor = root_3_2 * dp * n[k];
    ep[k] = ep0[k] + (k < 3 ? tensor : 2.0 * tensor);
  }
  p = p0 + dp;
  for (int k = 0; k < 6; ++k) stress[k] = theta * trial[k] + (k < 3 ? bulk * volume : 0.0);
  for (int i = 0; i < 6; ++i)
    for (int j = 0; j < 6; ++j) {
      const double delta = i == j ? 1.0 : 0.0;
      const double dev = (i < 3 && j < 3) ? delta - 1.0 / 3.0 : ((i >= 3 && i == j) ? 0.5 : 0.0);
      const double vol = (i < 3 && j < 3) ? bulk : 0.0;
      c[i][j] = vol + 2.0 * g * theta * dev - 2.0 * g * theta_bar * n[i] * n[j];
    }
  return dp > 0.0;
}

// solid_response (element.rs:395-452) on the global displacements u, from the committed states,
// one thread per (element e, column i) like k_stiffness: each thread repeats the Gauss-point
// strain and return mapping (cheap beside the column) and accumulates, Gauss point after Gauss
// point in the reference's order, column i of the tangent (tangent[j][i] = sum w b_j.(C b_i),
// written to slot (j*M + i)*ne + e, the stiffness layout the gather assembly reads) and the
// internal force fint[i] = sum w b_i.sigma (slot i*ne + e). The thread of column 0 writes the
// end states (component k of point q at (k*NG + q)*ne + e; 6 plastic strains, then p), the
// Gauss-point records (position, then the Voigt stress: 9 doubles at (e*NG + q)*9) and, if a
// point yielded, raises *yielding.
template <int NEN>
__global__ __launch_bounds__(256) void k_response(
    int64_t ne, const double* __restrict__ xyz, const int* __restrict__ conn,
    const double* __restrict__ young, const double* __restrict__ nu,
    const double* __restrict__ yield0, const double* __restrict__ hardening,
    const double* __restrict__ u, const double* __restrict__ committed,
    double* __restrict__ ke, double* __restrict__ fint, double* __restrict__ trial,
    double* __restrict__ records, int* __restrict__ yielding) {
#pragma clang fp contract(off)
  constexpr int M = 3 * NEN, NG = Rule<NEN>::N;
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)M * ne) return;
  const int i = (int)(t / ne);
  const int64_t e = t - (int64_t)i * ne;
  double x[NEN][3], ue[M];
  for (int a = 0; a < NEN; ++a) {
    const int64_t node = conn[e * NEN + a];
    for (int c = 0; c < 3; ++c) {
      x[a][c] = xyz[3 * node + c];
      ue[3 * a + c] = u[3 * node + c];
    }
  }
  const double E = young[e], v = nu[e], sy = yield0[e], h = hardening[e];
  double acc[M];
  for (int j = 0; j < M; ++j) acc[j] = 0.0;
  double fi = 0.0;
  bool yielded = false;
  for (int q = 0; q < NG; ++q) {
    double xi, eta, zeta, w;
    gauss_point(NEN, q, xi, eta, zeta, w);
    double g[NEN][3];
    const double det = tet_mapping<NEN>(x, xi, eta, zeta, g);
    const double weight = w * fabs(det);
    double strain[6];
    for (int r = 0; r < 6; ++r) {
      double s = 0.0;
      for (int k = 0; k < M; ++k) {
        double col[6];
        b_column<NEN>(g, k, col);
        s += col[r] * ue[k];
      }
      strain[r] = s;
    }
    double ep0[6], ep[6], p, stress[6], c[6][6];
    for (int k = 0; k < 6; ++k) ep0[k] = committed[((int64_t)k * NG + q) * ne + e];
    const double p0 = committed[((int64_t)6 * NG + q) * ne + e];
    yielded |= von_mises_update(E, v, sy, h, strain, ep0, p0, stress, ep, p, c);
    double bi[6], cb[6];
    b_column<NEN>(g, i, bi);
    double f = 0.0;
    for (int r = 0; r < 6; ++r) f += bi[r] * stress[r];
    fi += weight * f;
    for (int r = 0; r < 6; ++r) {
      double s = 0.0;
      for (int k = 0; k < 6; ++k) s += c[r][k] * bi[k];
      cb[r] = s;
    }
    for (int j = 0; j < M; ++j) {
      double bj[6];
      b_column<NEN>(g, j, bj);
      double s = 0.0;
      for (int r = 0; r < 6; ++r) s += bj[r] * cb[r];
      acc[j] += weight * s;
    }
    if (i == 0) {
      for (int k = 0; k < 6; ++k) trial[((int64_t)k * NG + q) * ne + e] = ep[k];
      trial[((int64_t)6 * NG + q) * ne + e] = p;
      double n[NEN];
      const double l[4] = {1.0 - xi - eta - zeta, xi, eta, zeta};
      if constexpr (NEN == 4) {
        for (int a = 0; a < 4; ++a) n[a] = l[a];
      } else {
        const int ed[6][2] = {{0, 1}, {1, 2}, {2, 0}, {0, 3}, {1, 3}, {2, 3}};
        for (int a = 0; a < 4; ++a) n[a] = l[a] * (2.0 * l[a] - 1.0);
        for (int k = 0; k < 6; ++k) n[4 + k] = 4.0 * l[ed[k][0]] * l[ed[k][1]];
      }
      double* o = records + (e * NG + q) * 9;
      for (int k = 0; k < 3; ++k) {
        double s = 0.0;
        for (int a = 0; a < NEN; ++a) s += n[a] * x[a][k];
        o[k] = s;
      }
      for (int k = 0; k < 6; ++k) o[3 + k] = stress[k];
    }
  }
  for (int j = 0; j < M; ++j) ke[((int64_t)j * M + i) * ne + e] = acc[j];
  fint[(int64_t)i * ne + e] = fi;
  if (i == 0 && yielded) *yielding = 1;
}

// large_strain_response (finite_strain.rs:119-214) with LargeStrainLaw::evaluate (:39-86), one
// thread per (element e, column c = 3 nb + k) like k_response: per Gauss point, F = I + sum u_a
// (x) grad N_a, the second Piola-Kirchhoff stress S and modulus of the law (Saint-Venant-Kirchhoff,
// or Neo-Hookean when neo[e]), P = F S and A_ijkl = delta_ik S_jl + sum F_i,ii C_ii,j,kk,l F_k,kk
// for the thread's k; column c of the tangent (rows (na, i): sum_j sum_l g_na,j A_ijkl g_nb,l) to
// slot (row*M + c)*ne + e, and fint[c] = sum w P_k,j g_nb,j. The thread of column 0 writes zero
// states and the Cauchy stress P F^T / J at the current Gauss-point position.
template <int NEN>
__global__ __launch_bounds__(256) void k_large_response(
    int64_t ne, const double* __restrict__ xyz, const int* __restrict__ conn,
    const double* __restrict__ young, const double* __restrict__ nu,
    const unsigned char* __restrict__ neo, const double* __restrict__ u,
    double* __restrict__ ke, double* __restrict__ fint, double* __restrict__ trial,
    double* __restrict__ records) {
#pragma clang fp contract(off)
  constexpr int M = 3 * NEN, NG = Rule<NEN>::N;
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)M * ne) return;
  const int col = (int)(t / ne);
  const int64_t e = t - (int64_t)col * ne;
  const int nb = col / 3, kc = col % 3;
  double x[NEN][3], ue[M];
  for (int a = 0; a < NEN; ++a) {
    const int64_t node = conn[e * NEN + a];
    for (int c = 0; c < 3; ++c) {
      x[a][c] = xyz[3 * node + c];
      ue[3 * a + c] = u[3 * node + c];
    }
  }
  const double E = young[e], v = nu[e];
  const double lambda = E * v / ((1.0 + v) * (1.0 - 2.0 * v)), mu = E / (2.0 * (1.0 + v));
  const bool nh = neo && neo[e];
  double acc[M];
  for (int r = 0; r < M; ++r) acc[r] = 0.0;
  double fi = 0.0;
  auto kd = [](int a, int b) { return a == b ? 1.0 : 0.0; };
  for (int q = 0; q < NG; ++q) {
    double xi, eta, zeta, w;
    gauss_point(NEN, q, xi, eta, zeta, w);
    double g[NEN][3];
    const double weight = w * fabs(tet_mapping<NEN>(x, xi, eta, zeta, g));
    double f[3][3];
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) f[i][j] = i == j ? 1.0 : 0.0;
    for (int a = 0; a < NEN; ++a)
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) f[i][j] += ue[3 * a + i] * g[a][j];
    double c[3][3];
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        double s = 0.0;
        for (int k = 0; k < 3; ++k) s += f[k][i] * f[k][j];
        c[i][j] = s;
      }
    const double jdet = f[0][0] * (f[1][1] * f[2][2] - f[1][2] * f[2][1]) -
                        f[0][1] * (f[1][0] * f[2][2] - f[1][2] * f[2][0]) +
                        f[0][2] * (f[1][0] * f[2][1] - f[1][1] * f[2][0]);
    double s2[3][3], inv[3][3], log_j = 0.0;
    if (nh) {
      log_j = log(jdet);
      const double dc = c[0][0] * (c[1][1] * c[2][2] - c[1][2] * c[2][1]) -
                        c[0][1] * (c[1][0] * c[2][2] - c[1][2] * c[2][0]) +
                        c[0][2] * (c[1][0] * c[2][1] - c[1][1] * c[2][0]);
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
          const int r0 = (j + 1) % 3, r1 = (j + 2) % 3, c0 = (i + 1) % 3, c1 = (i + 2) % 3;
          inv[i][j] = (c[r0][c0] * c[r1][c1] - c[r0][c1] * c[r1][c0]) / dc;
        }
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) s2[i][j] = mu * (kd(i, j) - inv[i][j]) + lambda * log_j * inv[i][j];
    } else {
      double ee[3][3];
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) ee[i][j] = 0.5 * (c[i][j] - kd(i, j));
      const double trace = ee[0][0] + ee[1][1] + ee[2][2];
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) s2[i][j] = 2.0 * mu * ee[i][j] + lambda * trace * kd(i, j);
    }
    auto modulus = [&](int i, int j, int k, int l) {
      if (nh)
        return lambda * inv[i][j] * inv[k][l] +
               (mu - lambda * log_j) * (inv[i][k] * inv[j][l] + inv[i][l] * inv[j][k]);
      return lambda * kd(i, j) * kd(k, l) + mu * (kd(i, k) * kd(j, l) + kd(i, l) * kd(j, k));
    };
    double pk1[3][3];
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        double s = 0.0;
        for (int k = 0; k < 3; ++k) s += f[i][k] * s2[k][j];
        pk1[i][j] = s;
      }
    // A_{i j kc l} for this thread's kc
    double a4[3][3][3];
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j)
        for (int l = 0; l < 3; ++l) {
          double material = 0.0;
          for (int ii = 0; ii < 3; ++ii)
            for (int kk = 0; kk < 3; ++kk) material += f[i][ii] * modulus(ii, j, kk, l) * f[kc][kk];
          const double geometric = i == kc ? s2[j][l] : 0.0;
          a4[i][j][l] = geometric + material;
        }
    double gb[3];
    pick_row<NEN>(g, nb, gb);
    {
      double s = 0.0;
      for (int j = 0; j < 3; ++j) s += pk1[kc][j] * gb[j];
      fi += weight * s;
    }
#pragma unroll
    for (int na = 0; na < NEN; ++na)
      for (int i = 0; i < 3; ++i) {
        double vv = 0.0;
        for (int j = 0; j < 3; ++j)
          for (int l = 0; l < 3; ++l) vv += g[na][j] * a4[i][j][l] * gb[l];
        acc[3 * na + i] += weight * vv;
      }
    if (col == 0) {
      for (int k = 0; k < 7; ++k) trial[((int64_t)k * NG + q) * ne + e] = 0.0;
      double n[NEN];
      const double l4[4] = {1.0 - xi - eta - zeta, xi, eta, zeta};
      if constexpr (NEN == 4) {
        for (int a = 0; a < 4; ++a) n[a] = l4[a];
      } else {
        const int ed[6][2] = {{0, 1}, {1, 2}, {2, 0}, {0, 3}, {1, 3}, {2, 3}};
        for (int a = 0; a < 4; ++a) n[a] = l4[a] * (2.0 * l4[a] - 1.0);
        for (int k = 0; k < 6; ++k) n[4 + k] = 4.0 * l4[ed[k][0]] * l4[ed[k][1]];
      }
      double sigma[3][3];
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
          double s = 0.0;
          for (int k = 0; k < 3; ++k) s += pk1[i][k] * f[j][k];
          sigma[i][j] = s;
        }
      double* o = records + (e * NG + q) * 9;
      for (int k = 0; k < 3; ++k) {
        double p0 = 0.0;
        for (int a = 0; a < NEN; ++a) p0 += n[a] * x[a][k];
        double s = 0.0;
        for (int a = 0; a < NEN; ++a) s += n[a] * ue[3 * a + k];
        o[k] = p0 + s;
      }
      o[3] = sigma[0][0] / jdet;
      o[4] = sigma[1][1] / jdet;
      o[5] = sigma[2][2] / jdet;
      o[6] = sigma[1][2] / jdet;
      o[7] = sigma[0][2] / jdet;
      o[8] = sigma[0][1] / jdet;
    }
  }
  for (int r = 0; r < M; ++r) ke[((int64_t)r * M + col) * ne + e] = acc[r];
  fint[(int64_t)col * ne + e] = fi;
}

// Element tangents corrected in place, group after group (a group is one element's list of
// (slot, value) additions in the caller's order): ke[slot] += value.
__global__ __launch_bounds__(256) void k_add_groups(int64_t ngroups, const int64_t* __restrict__ off,
                                                    const int64_t* __restrict__ slot,
                                                    const double* __restrict__ val,
                                                    double* __restrict__ ke) {
#pragma clang fp contract(off)
  const int64_t gidx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (gidx >= ngroups) return;
  for (int64_t c = off[gidx]; c < off[gidx + 1]; ++c) ke[slot[c]] += val[c];
}

// internal_force (model.rs:1241-1250), thread per node: F_int[3n + c] sums the node's element
// forces over its incidences in element order (f[g] += v, starting from 0.0).
template <int NEN>
__global__ __launch_bounds__(256) void k_internal_force(int64_t n_nodes, int64_t ne,
                                                        const int64_t* __restrict__ inc_off,
                                                        const int* __restrict__ inc_item,
                                                        const double* __restrict__ fint,
                                                        double* __restrict__ out) {
#pragma clang fp contract(off)
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= n_nodes) return;
  double s[3] = {0.0, 0.0, 0.0};
  for (int64_t k = inc_off[v]; k < inc_off[v + 1]; ++k) {
    const int64_t t = inc_item[k];
    const int64_t e = t / NEN;
    const int a = (int)(t - e * NEN);
    for (int c = 0; c < 3; ++c) s[c] += fint[(int64_t)(3 * a + c) * ne + e];
  }
  for (int c = 0; c < 3; ++c) out[3 * v + c] = s[c];
}

// The Newton residual of equilibrium (model.rs:950-967), thread per DOF: external = lambda (F +
// follower) + 0 (follower forces per unit lambda under large strains, else 0; no contact force on
// this path), r = F_int - external, -r for the solve, and
// per-block sums of F_int^2, external^2 and r^2 over the free DOFs (fixed-tree, deterministic)
// for the three norms; the reactions are r on the fixed DOFs (model.rs:981-986).
__global__ __launch_bounds__(256) void k_newton_residual(int64_t ndof, double lambda,
                                                         const double* __restrict__ f,
                                                         const double* __restrict__ follower,
                                                         const double* __restrict__ fint,
                                                         const int* __restrict__ rank_of_dof,
                                                         double* __restrict__ neg_r,
                                                         double* __restrict__ reactions,
                                                         double* __restrict__ partial) {
#pragma clang fp contract(off)
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double a = 0.0, b = 0.0, c = 0.0;
  if (g < ndof) {
    const double ext = lambda * (f[g] + (follower ? follower[g] : 0.0)) + 0.0;
    const double fi = fint[g];
    const double r = fi - ext;
    neg_r[g] = -r;
    const bool free_dof = rank_of_dof[g] >= 0;
    reactions[g] = free_dof ? 0.0 : r;
    a = fi * fi;
    b = ext * ext;
    c = free_dof ? r * r : 0.0;
  }
  a = block_sum<256>(a);
  b = block_sum<256>(b);
  c = block_sum<256>(c);
  if (threadIdx.x == 0) {
    partial[3 * blockIdx.x] = a;
    partial[3 * blockIdx.x + 1] = b;
    partial[3 * blockIdx.x + 2] = c;
  }
}

// One block: the three sums of k_newton_residual's partials, in a fixed order.
__global__ __launch_bounds__(256) void k_sum3(const double* __restrict__ partial, int count,
                                              double* __restrict__ out) {
  for (int w = 0; w < 3; ++w) {
    double v = 0.0;
    for (int k = threadIdx.x; k < count; k += 256) v += partial[3 * k + w];
    v = block_sum<256>(v);
    if (threadIdx.x == 0) out[w] = v;
    __syncthreads();
  }
}

// du at the start of a Newton iteration (model.rs:998-1004): lambda * BC value - u on the fixed
// DOFs, zero elsewhere.
__global__ __launch_bounds__(256) void k_newton_du(int64_t ndof, double lambda,
                                                   const int* __restrict__ rank_of_dof,
                                                   const double* __restrict__ imposed,
                                                   const double* __restrict__ u,
                                                   double* __restrict__ du) {
#pragma clang fp contract(off)
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= ndof) return;
  du[g] = rank_of_dof[g] < 0 ? lambda * imposed[g] - u[g] : 0.0;
}

// FreeSystem::solve's scatter into du (model.rs:1509-1514) and u += du (model.rs:1017-1019).
__global__ __launch_bounds__(256) void k_newton_update(int64_t ndof,
                                                       const int* __restrict__ rank_of_dof,
                                                       const double* __restrict__ x,
                                                       const double* __restrict__ du,
                                                       double* __restrict__ u) {
#pragma clang fp contract(off)
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= ndof) return;
  const int r = rank_of_dof[g];
  u[g] += r >= 0 ? x[r] : du[g];
}

// ------------------------------------------------------------------ transient heat conduction
// TETRA_14 (quadrature.rs:139-177), the rule of the Tet10 capacity
__device__ inline void tetra14_point(int q, double& xi, double& eta, double& zeta, double& w) {
  const double A1 = 0.09273525031089142, W1 = 0.01224884051939372;
  const double A2 = 0.3108859192633008, W2 = 0.01878132095300283;
  const double B = 0.04550370412564851, W3 = 0.00709100346284674;
  const double C1 = 1.0 - 3.0 * A1, C2 = 1.0 - 3.0 * A2, D = 0.5 - B;
  if (q < 8) {
    const double a = q < 4 ? A1 : A2, c = q < 4 ? C1 : C2;
    const int r = q & 3;
    xi = r == 1 ? c : a;
    eta = r == 2 ? c : a;
    zeta = r == 3 ? c : a;
    w = q < 4 ? W1 : W2;
    return;
  }
  const double pts[6][3] = {{B, B, D}, {B, D, B}, {D, B, B}, {D, D, B}, {D, B, D}, {B, D, D}};
  xi = pts[q - 8][0];
  eta = pts[q - 8][1];
  zeta = pts[q - 8][2];
  w = W3;
}

// TETRA_4 at point q (Tet4's capacity rule)
__device__ inline void tetra4_point(int q, double& xi, double& eta, double& zeta, double& w) {
  gauss_point(10, q, xi, eta, zeta, w);
}

template <int NEN>
__device__ inline void tet_values(double xi, double eta, double zeta, double (&n)[NEN]) {
#pragma clang fp contract(off)
  const double l[4] = {1.0 - xi - eta - zeta, xi, eta, zeta};
  if constexpr (NEN == 4) {
    for (int a = 0; a < 4; ++a) n[a] = l[a];
  } else {
    const int ed[6][2] = {{0, 1}, {1, 2}, {2, 0}, {0, 3}, {1, 3}, {2, 3}};
    for (int a = 0; a < 4; ++a) n[a] = l[a] * (2.0 * l[a] - 1.0);
    for (int k = 0; k < 6; ++k) n[4 + k] = 4.0 * l[ed[k][0]] * l[ed[k][1]];
  }
}

// conduction (element.rs:450-462, the element's own rule) and consistent capacity
// (element.rs:465-477 with TETRA_4 for Tet4, TETRA_14 for Tet10; assembly.rs:392-431), one
// thread per (element e, row i): K[i][j] += w|det| k (g_i . g_j), C[i][j] += w|det| rho_c N_i N_j,
// Gauss point after Gauss point. Entry (i, j) at slot (i*NEN + j)*ne + e.
template <int NEN>
__global__ __launch_bounds__(256) void k_thermal_matrices(int64_t ne, const double* __restrict__ xyz,
                                                          const int* __restrict__ conn,
                                                          const double* __restrict__ k,
                                                          const double* __restrict__ rho_c,
                                                          double* __restrict__ kmat,
                                                          double* __restrict__ cmat) {
#pragma clang fp contract(off)
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)NEN * ne) return;
  const int i = (int)(t / ne);
  const int64_t e = t - (int64_t)i * ne;
  double x[NEN][3];
  for (int a = 0; a < NEN; ++a) {
    const int64_t node = conn[e * NEN + a];
    for (int c = 0; c < 3; ++c) x[a][c] = xyz[3 * node + c];
  }
  const double ke = k[e], rc = rho_c[e];
  double acc[NEN];
  for (int j = 0; j < NEN; ++j) acc[j] = 0.0;
  for (int q = 0; q < Rule<NEN>::N; ++q) {
    double xi, eta, zeta, w;
    gauss_point(NEN, q, xi, eta, zeta, w);
    double g[NEN][3];
    const double weight = w * fabs(tet_mapping<NEN>(x, xi, eta, zeta, g));
    double gi[3];
    pick_row<NEN>(g, i, gi);
#pragma unroll
    for (int j = 0; j < NEN; ++j) {
      double s = 0.0;
      for (int c = 0; c < 3; ++c) s += gi[c] * g[j][c];
      acc[j] += weight * ke * s;
    }
  }
  for (int j = 0; j < NEN; ++j) kmat[((int64_t)i * NEN + j) * ne + e] = acc[j];
  for (int j = 0; j < NEN; ++j) acc[j] = 0.0;
  constexpr int NC = NEN == 4 ? 4 : 14;
  for (int q = 0; q < NC; ++q) {
    double xi, eta, zeta, w;
    if constexpr (NEN == 4) tetra4_point(q, xi, eta, zeta, w);
    else tetra14_point(q, xi, eta, zeta, w);
    double g[NEN][3], n[NEN];
    const double weight = w * fabs(tet_mapping<NEN>(x, xi, eta, zeta, g));
    tet_values<NEN>(xi, eta, zeta, n);
    const double ni = pick<NEN>(n, i);
#pragma unroll
    for (int j = 0; j < NEN; ++j) acc[j] += weight * rc * ni * n[j];
  }
  for (int j = 0; j < NEN; ++j) cmat[((int64_t)i * NEN + j) * ne + e] = acc[j];
}

// Convection face matrices added into K (thermal_matrix, model.rs:1320-1326): element e adds its
// list (slot, value) in order, the faces in the order the model lists them.
__global__ __launch_bounds__(256) void k_add_faces(int64_t ne, const int64_t* __restrict__ off,
                                                   const int64_t* __restrict__ slot,
                                                   const double* __restrict__ val,
                                                   double* __restrict__ kmat) {
#pragma clang fp contract(off)
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= ne) return;
  for (int64_t c = off[e]; c < off[e + 1]; ++c) kmat[slot[c]] += val[c];
}

// model.rs:652-663 (combine): left = C/dt + theta K, right = C/dt + (theta - 1) K
__global__ __launch_bounds__(256) void k_theta_combine(int64_t n, double dt, double s_left,
                                                       double s_right,
                                                       const double* __restrict__ kmat,
                                                       const double* __restrict__ cmat,
                                                       double* __restrict__ left,
                                                       double* __restrict__ right) {
#pragma clang fp contract(off)
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  const double c = cmat[q], k = kmat[q];
  left[q] = c / dt + s_left * k;
  right[q] = c / dt + s_right * k;
}

// product (model.rs:1353-1369) over every DOF, then + F (model.rs:681-682), thread per node: the
// node's incidences (element e, local node a) in element order each add sum_b M[a][b] T[node b].
template <int NEN>
__global__ __launch_bounds__(256) void k_thermal_carry(int64_t n_nodes, int64_t ne,
                                                       const int64_t* __restrict__ inc_off,
                                                       const int* __restrict__ inc_item,
                                                       const int* __restrict__ conn,
                                                       const double* __restrict__ mat,
                                                       const double* __restrict__ temp,
                                                       const double* __restrict__ f,
                                                       double* __restrict__ out) {
#pragma clang fp contract(off)
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= n_nodes) return;
  double y = 0.0;
  for (int64_t k = inc_off[v]; k < inc_off[v + 1]; ++k) {
    const int64_t t = inc_item[k];
    const int64_t e = t / NEN;
    const int a = (int)(t - e * NEN);
    double s = 0.0;
    for (int b = 0; b < NEN; ++b) s += mat[((int64_t)a * NEN + b) * ne + e] * temp[conn[e * NEN + b]];
    y += s;
  }
  out[v] = y + f[v];
}

// Reactions of a step (model.rs:684-692): product of the left matrix over the fixed DOFs (the
// rows (e, a) of fixed DOF k in element order), minus the step's right-hand side.
template <int NEN>
__global__ __launch_bounds__(256) void k_thermal_reactions(int64_t nfixed, const int* __restrict__ rdof,
                                                           const int64_t* __restrict__ roff,
                                                           const int64_t* __restrict__ item,
                                                           const int* __restrict__ conn, int64_t ne,
                                                           const double* __restrict__ left,
                                                           const double* __restrict__ temp,
                                                           const double* __restrict__ rhs,
                                                           double* __restrict__ reactions) {
#pragma clang fp contract(off)
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nfixed) return;
  double y = 0.0;
  for (int64_t c = roff[k]; c < roff[k + 1]; ++c) {
    const int64_t e = item[c] / NEN;
    const int a = (int)(item[c] - e * NEN);
    double s = 0.0;
    for (int b = 0; b < NEN; ++b) s += left[((int64_t)a * NEN + b) * ne + e] * temp[conn[e * NEN + b]];
    y += s;
  }
  const int g = rdof[k];
  reactions[g] = y - rhs[g];
}

// heat_fluxes (element.rs:478-499), thread per (element, Gauss point): x, y, z, then
// q = -k grad T.
template <int NEN>
__global__ __launch_bounds__(256) void k_heat_flux(int64_t ne, const double* __restrict__ xyz,
                                                   const int* __restrict__ conn,
                                                   const double* __restrict__ k,
                                                   const double* __restrict__ temp,
                                                   double* __restrict__ out) {
#pragma clang fp contract(off)
  constexpr int NG = Rule<NEN>::N;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ne * NG) return;
  const int64_t e = t / NG;
  const int q = (int)(t - e * NG);
  double x[NEN][3], te[NEN];
  for (int a = 0; a < NEN; ++a) {
    const int64_t node = conn[e * NEN + a];
    for (int c = 0; c < 3; ++c) x[a][c] = xyz[3 * node + c];
    te[a] = temp[node];
  }
  double xi, eta, zeta, w;
  gauss_point(NEN, q, xi, eta, zeta, w);
  double g[NEN][3], n[NEN];
  tet_mapping<NEN>(x, xi, eta, zeta, g);
  tet_values<NEN>(xi, eta, zeta, n);
  double* o = out + t * 6;
  for (int c = 0; c < 3; ++c) {
    double s = 0.0;
    for (int a = 0; a < NEN; ++a) s += n[a] * x[a][c];
    o[c] = s;
  }
  for (int c = 0; c < 3; ++c) {
    double s = 0.0;
    for (int a = 0; a < NEN; ++a) s += g[a][c] * te[a];
    o[3 + c] = -k[e] * s;
  }
}

// ------------------------------------------------------------------ assembly
// Csr::add in the reference's scatter order (model.rs:1480-1492): entry q sums its element
// contributions in element order, starting from 0.0.
__global__ __launch_bounds__(256) void k_assemble(int64_t nnz, const int64_t* __restrict__ off,
                                                  const uint32_t* __restrict__ slot,
                                                  const double* __restrict__ ke,
                                                  double* __restrict__ vals) {
#pragma clang fp contract(off)
  int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nnz) return;
  double v = 0.0;
  for (int64_t c = off[q]; c < off[q + 1]; ++c) v += ke[slot[c]];
  vals[q] = v;
}

// coupling[a] = sum of ke * u_c over fixed columns, element order (model.rs:1489)
__global__ __launch_bounds__(256) void k_coupling(int64_t nfree, const int64_t* __restrict__ off,
                                                  const uint32_t* __restrict__ slot,
                                                  const int* __restrict__ dof,
                                                  const double* __restrict__ ke,
                                                  const double* __restrict__ imposed,
                                                  double* __restrict__ coupling) {
#pragma clang fp contract(off)
  int64_t a = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= nfree) return;
  double c = 0.0;
  for (int64_t k = off[a]; k < off[a + 1]; ++k) c += ke[slot[k]] * imposed[dof[k]];
  coupling[a] = c;
}

// FreeSystem::solve's right-hand side (model.rs:1503-1508): rhs[rank] = -coupling + f
__global__ __launch_bounds__(256) void k_rhs(int64_t ndof, const int* __restrict__ rank_of_dof,
                                             const double* __restrict__ f,
                                             const double* __restrict__ coupling,
                                             double* __restrict__ rhs) {
#pragma clang fp contract(off)
  int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= ndof) return;
  int r = rank_of_dof[g];
  if (r < 0) return;
  double c = coupling ? coupling[r] : 0.0;
  rhs[r] = -c + f[g];
}

// model.rs:1509-1514: u[g] = x[rank] on free DOFs, the imposed value on fixed ones
__global__ __launch_bounds__(256) void k_scatter_u(int64_t ndof, const int* __restrict__ rank_of_dof,
                                                   const double* __restrict__ x,
                                                   const double* __restrict__ imposed,
                                                   double* __restrict__ u) {
  int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= ndof) return;
  int r = rank_of_dof[g];
  u[g] = r >= 0 ? x[r] : imposed[g];
}

// product (model.rs:1353-1369) on the fixed DOFs, then R -= F (model.rs:586-593)
template <int NEN>
__global__ __launch_bounds__(256) void k_reactions(int64_t nfixed, const int* __restrict__ rdof,
                                                   const int64_t* __restrict__ roff,
                                                   const int64_t* __restrict__ item,
                                                   const double* __restrict__ ke,
                                                   const int* __restrict__ conn, int64_t ne,
                                                   const double* __restrict__ u,
                                                   const double* __restrict__ f,
                                                   double* __restrict__ reactions) {
#pragma clang fp contract(off)
  constexpr int M = 3 * NEN;
  int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nfixed) return;
  double y = 0.0;
  for (int64_t c = roff[k]; c < roff[k + 1]; ++c) {
    int64_t e = item[c] / M;
    int li = (int)(item[c] - e * M);
    double s = 0.0;
    for (int lj = 0; lj < M; ++lj) {
      int64_t h = 3 * (int64_t)conn[e * NEN + lj / 3] + lj % 3;
      s += ke[((int64_t)li * M + lj) * ne + e] * u[h];
    }
    y += s;
  }
  int g = rdof[k];
  reactions[g] = y - f[g];
}

// Compliance sensitivity, thread per element: C = f.u with K(E) u = f and homogeneous essential
// conditions is self-adjoint (the adjoint state is u), and K_e is linear in E_e, so
// dC/dE_e = -u_e^T (K_e / E_e) u_e.
template <int NEN>
__global__ __launch_bounds__(256) void k_compliance_gradient(int64_t ne,
                                                             const int* __restrict__ conn,
                                                             const double* __restrict__ ke,
                                                             const double* __restrict__ young,
                                                             const double* __restrict__ u,
                                                             double* __restrict__ out) {
#pragma clang fp contract(off)
  constexpr int M = 3 * NEN;
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= ne) return;
  double ue[M];
#pragma unroll
  for (int a = 0; a < NEN; ++a) {
    const int64_t node = conn[e * NEN + a];
#pragma unroll
    for (int c = 0; c < 3; ++c) ue[3 * a + c] = u[3 * node + c];
  }
  double s = 0.0;
#pragma unroll
  for (int i = 0; i < M; ++i) {
    double t = 0.0;
#pragma unroll
    for (int j = 0; j < M; ++j) t += ke[((int64_t)i * M + j) * ne + e] * ue[j];
    s += ue[i] * t;
  }
  out[e] = -s / young[e];
}

// Node-averaged stresses (nodal_stresses + smooth, results.rs:138-166,184-221), thread per node:
// its incidences (element e, local node k), in element order, each add the element's Gauss-point
// values extrapolated to node k, Gauss point after Gauss point (s += w * v); the sum is divided
// by the number of incidences (an isolated node keeps zeros). Tet4: one Gauss point, weight 1.
// Tet10: TETRA_4 point g lies near vertex g, and a vertex takes (3 sqrt5 + 1)/4 of its own point
// and -(sqrt5 - 1)/4 of the others, an edge mid-node (sqrt5 + 1)/4 of its two end points' and
// -(sqrt5 - 1)/4 of the others (results.rs:113-133). out per node: x, y, z, then the six
// components in the Gauss-point layout's order.
template <int NEN>
__device__ inline double extrapolation_weight(int k, int g) {
  constexpr double VERTEX = 1.9270509831248424, MIDDLE = 0.8090169943749475,
                   OTHER = -0.30901699437494745;
  if (NEN == 4) return 1.0;
  if (k < 4) return k == g ? VERTEX : OTHER;
  // mid-node k: edge (0-1), (1-2), (2-0), (0-3), (1-3), (2-3) for k = 4 .. 9
  const int a = (k == 4 || k == 6 || k == 7) ? 0 : (k == 5 || k == 8) ? 1 : 2;
  const int b = k == 4 ? 1 : k == 5 ? 2 : k == 6 ? 2 : 3;
  return (g == a || g == b) ? MIDDLE : OTHER;
}

template <int NEN>
__global__ __launch_bounds__(256) void k_nodal_stresses(int64_t n_nodes,
                                                        const double* __restrict__ xyz,
                                                        const int64_t* __restrict__ inc_off,
                                                        const int* __restrict__ inc_item,
                                                        const double* __restrict__ gp,
                                                        double* __restrict__ out) {
#pragma clang fp contract(off)
  constexpr int NGP = NEN == 4 ? 1 : 4;
  const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= n_nodes) return;
  double s[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  const int64_t t0 = inc_off[n], t1 = inc_off[n + 1];
  // four incidences at a time: their items, then all their Gauss-point values, are loaded
  // before the sums take them in order (Tet4; one at a time for the four-point Tet10)
  constexpr int U = NEN == 4 ? 4 : 1;
  int64_t t = t0;
  for (; t + U <= t1; t += U) {
    int it[U];
    double v[U][NGP][6];
#pragma unroll
    for (int u = 0; u < U; ++u) it[u] = inc_item[t + u];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const double* p = gp + (int64_t)(it[u] / NEN) * NGP * 9;
#pragma unroll
      for (int g = 0; g < NGP; ++g)
#pragma unroll
        for (int c = 0; c < 6; ++c) v[u][g][c] = p[g * 9 + 3 + c];
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int k = it[u] % NEN;
#pragma unroll
      for (int g = 0; g < NGP; ++g) {
        const double w = extrapolation_weight<NEN>(k, g);
#pragma unroll
        for (int c = 0; c < 6; ++c) s[c] += w * v[u][g][c];
      }
    }
  }
  for (; t < t1; ++t) {
    const int it = inc_item[t];
    const int64_t e = it / NEN;
    const int k = it - (int)e * NEN;
    const double* p = gp + e * NGP * 9;
#pragma unroll
    for (int g = 0; g < NGP; ++g) {
      const double w = extrapolation_weight<NEN>(k, g);
#pragma unroll
      for (int c = 0; c < 6; ++c) s[c] += w * p[g * 9 + 3 + c];
    }
  }
  const double cnt = (double)(t1 > t0 ? t1 - t0 : 1);
#pragma unroll
  for (int d = 0; d < 3; ++d) out[9 * n + d] = xyz[3 * n + d];
#pragma unroll
  for (int c = 0; c < 6; ++c) out[9 * n + 3 + c] = s[c] / cnt;
}

// A load of data read once per launch (matrix entries, factor coefficients): non-temporal, so
// that it does not evict from the L2 the vector entries that neighbouring rows gather again.
template <typename T>
__device__ inline T stream_load(const T* p) {
  return __builtin_nontemporal_load(p);
}
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));

// ------------------------------------------------------------------ SpMV (sliced ELL)
// y = A x (Csr::multiply, iterative.rs:131-142). Lane j of a slice sums row j in the order
// diagonal, lower entries, upper entries: the reference's accumulation order, so the FP64
// result is bit-identical. Optionally accumulates x.y per block (the PCG curvature p.Ap).
// Entries are taken BATCH at a time, so that their loads are in flight together. D16: the
// columns are read as 16-bit offsets from the row (col = j + sdel), else as int32 indices.
template <typename T, int BATCH = 8, bool D16 = false>
__global__ __launch_bounds__(256) void k_spmv(int n, const int* __restrict__ soff,
                                              const int* __restrict__ scol,
                                              const int16_t* __restrict__ sdel,
                                              const int* __restrict__ slen,
                                              const T* __restrict__ sval,
                                              const T* __restrict__ x, T* __restrict__ y,
                                              double* __restrict__ partial,
                                              const PcgScalars* __restrict__ S,
                                              int only_on_update = 0) {
#pragma clang fp contract(off)
  if (S && (S->done || (only_on_update && !S->update))) return;
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  double dotv = 0.0;
  if (j < n) {
    const int base = soff[j >> 6] + (j & 63);
    const int len = slen[j];
    auto column = [&](int at) {
      return D16 ? j + (int)stream_load(sdel + at) : stream_load(scol + at);
    };
    T s = stream_load(sval + base) * x[column(base)];
    int k = 1;
    for (; k + BATCH <= len; k += BATCH) {
      int c[BATCH];
      T v[BATCH], xv[BATCH];
#pragma unroll
      for (int u = 0; u < BATCH; ++u) {
        c[u] = column(base + ((k + u) << 6));
        v[u] = stream_load(sval + base + ((k + u) << 6));
      }
#pragma unroll
      for (int u = 0; u < BATCH; ++u) xv[u] = x[c[u]];
#pragma unroll
      for (int u = 0; u < BATCH; ++u) s += v[u] * xv[u];
    }
    for (; k < len; ++k) {
      const int at = base + (k << 6);
      s += stream_load(sval + at) * x[column(at)];
    }
    y[j] = s;
    dotv = (double)x[j] * (double)s;
  }
  if (partial) {
    double b = block_sum<256>(dotv);
    if (threadIdx.x == 0) partial[blockIdx.x] = b;
  }
}

// k_spmv with the grouped column streams: the rows of a group (consecutive rows of the slice
// with equal column sets, the DOFs of one node) read one shared list of columns, ascending, as
// 16-bit offsets from the group's first row, so a wave loads each group's index once instead of
// once per row. Row j walks the list and skips its own column, which it took first as the
// diagonal: list entry q < pos is its sliced-ELL entry q + 1, and q > pos is entry q, so the
// values, the products and the order of the sum are those of k_spmv (bit-identical).
template <typename T, int BATCH = 8>
__global__ __launch_bounds__(256) void k_spmv_grouped(int n, const int* __restrict__ soff,
                                                      const int* __restrict__ goff,
                                                      const int* __restrict__ gnum,
                                                      const int16_t* __restrict__ gdel,
                                                      const uint32_t* __restrict__ ginfo,
                                                      const T* __restrict__ sval,
                                                      const T* __restrict__ x, T* __restrict__ y,
                                                      double* __restrict__ partial,
                                                      const PcgScalars* __restrict__ S,
                                                      int only_on_update = 0) {
#pragma clang fp contract(off)
  if (S && (S->done || (only_on_update && !S->update))) return;
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  double dotv = 0.0;
  if (j < n) {
    const int sl = j >> 6;
    const int base = soff[sl] + (j & 63);
    const uint32_t gi = ginfo[j];
    const int len = gi & 255, pos = (gi >> 8) & 255, g = (gi >> 16) & 63;
    const int first = j - (int)((gi >> 24) & 63);
    const int G = gnum[sl];
    const int16_t* list = gdel + goff[sl] + g;
    T s = stream_load(sval + base) * x[j];
    int q = 0;
    for (; q + BATCH <= len; q += BATCH) {
      int c[BATCH];
      T v[BATCH], xv[BATCH];
#pragma unroll
      for (int u = 0; u < BATCH; ++u) {
        const int qq = q + u;
        c[u] = first + (int)stream_load(list + qq * G);
        // the own position loads an entry that is not summed
        v[u] = stream_load(sval + base + ((qq < pos ? qq + 1 : qq) << 6));
      }
#pragma unroll
      for (int u = 0; u < BATCH; ++u) xv[u] = x[c[u]];
#pragma unroll
      for (int u = 0; u < BATCH; ++u) s = q + u == pos ? s : s + v[u] * xv[u];
    }
    for (; q < len; ++q) {
      if (q == pos) continue;
      const int c = first + (int)stream_load(list + q * G);
      s += stream_load(sval + base + ((q < pos ? q + 1 : q) << 6)) * x[c];
    }
    y[j] = s;
    dotv = (double)x[j] * (double)s;
  }
  if (partial) {
    double b = block_sum<256>(dotv);
    if (threadIdx.x == 0) partial[blockIdx.x] = b;
  }
}

// ------------------------------------------------------------------ Jacobi sweeps
// One Jacobi sweep of a triangular system of the IC(0) factor (no counterpart in the
// reference, which solves it exactly): out_i = (rhs_i - sum_j m_ij in_j) * rd_i over the
// strict triangle m of row i (l_ij, j < i, for L y = r; l_ji, j > i, for L^T z = y), with
// rd = RN64(1 / l_ii). The sum runs sequentially in ascending j from 0, with the factor's FP32
// coefficients widened exactly and FP64 vectors and arithmetic (contraction off, so a float64
// restatement reproduces it bit for bit). FIRST: the previous iterate is in = rhs * rd, formed
// as it is read. Lane i of a 64-row slice owns row i; entries are column-major in the slice.
// D16: columns are read as 16-bit distances from the row, j = i + sign * d (sign -1 for the
// lower triangle, +1 for the upper one), else as int32 indices. With out_rd, out * rd is also
// written (the next system's starting iterate, the same product FIRST would form). With
// `partial`, the block's share of dot.out (the PCG's r.z) is written, in a fixed tree. With
// `level`, out is an intermediate iterate and is written as level_value(out), the rounding the
// fused sweeps give the iterates they exchange. With a weight w != 1 the sweep is damped:
// out_i = c1 * in_i + w * t_i, with t_i the undamped result above and c1 = 1 - w (each product
// and the sum rounded, in that order), the same in every form of the sweeps.
template <bool FIRST, bool D16, int BATCH = 8>
__global__ __launch_bounds__(256) void k_sweep(int n, const int* __restrict__ soff,
                                               const int* __restrict__ scol,
                                               const uint16_t* __restrict__ sdel, int sign,
                                               const int* __restrict__ slen,
                                               const float* __restrict__ sval,
                                               const double* __restrict__ rd,
                                               const double* __restrict__ rhs,
                                               const double* __restrict__ in,
                                               double* __restrict__ out, int level,
                                               double* __restrict__ out_rd,
                                               const double* __restrict__ dot,
                                               double* __restrict__ partial,
                                               const PcgScalars* __restrict__ S, double w,
                                               double c1) {
#pragma clang fp contract(off)
  if (S && S->done) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double dv = 0.0;
  if (i < n) {
    const int base = soff[i >> 6] + (i & 63);
    const int len = slen[i];
    auto column = [&](int at) { return D16 ? i + sign * (int)sdel[at] : scol[at]; };
    double s = 0.0;
    int k = 0;
    // BATCH entries at a time: columns and coefficients, then the gathers, so that a wave
    // keeps many loads in flight; the sum stays in entry order
    for (; k + BATCH <= len; k += BATCH) {
      int c[BATCH];
      float v[BATCH];
      double x[BATCH];
#pragma unroll
      for (int u = 0; u < BATCH; ++u) {
        c[u] = column(base + ((k + u) << 6));
        v[u] = sval[base + ((k + u) << 6)];
      }
#pragma unroll
      for (int u = 0; u < BATCH; ++u) x[u] = FIRST ? rhs[c[u]] * rd[c[u]] : in[c[u]];
#pragma unroll
      for (int u = 0; u < BATCH; ++u) s = s + (double)v[u] * x[u];
    }
    for (; k < len; ++k) {
      const int at = base + (k << 6);
      const int c = column(at);
      s = s + (double)sval[at] * (FIRST ? rhs[c] * rd[c] : in[c]);
    }
    const double tv = (rhs[i] - s) * rd[i];
    const double y = w != 1.0 ? c1 * (FIRST ? rhs[i] * rd[i] : in[i]) + w * tv : tv;
    out[i] = level ? level_value(y) : y;
    if (out_rd) out_rd[i] = y * rd[i];
    if (dot) dv = dot[i] * y;
  }
  if (partial) {
    const double b = block_sum<256>(dv);
    if (threadIdx.x == 0) partial[blockIdx.x] = b;
  }
}

// ------------------------------------------------------------------ interleaved vectors
// nrep vectors of n entries are stacked (entry i of vector b at b n + i) in the batched PCG
// and interleaved (at i nrep + b) in the batched SpMV and Jacobi sweeps below, where the W
// lanes of a group take W vectors of one row: the row's entries are one load for the group
// and each gathered column is W consecutive values.

// out = in moved between the two layouts (TO_IL: stacked -> interleaved), in tiles of 32
// vectors x 32 entries through LDS so that both sides are read and written in runs. With
// out_rd (TO_IL only), out_rd[i nrep + b] = in * rd[i] is written as well.
template <bool TO_IL>
__global__ __launch_bounds__(256) void k_interleave(int n, int nrep, const double* __restrict__ in,
                                                    double* __restrict__ out,
                                                    const double* __restrict__ rd = nullptr,
                                                    double* __restrict__ out_rd = nullptr) {
  __shared__ double tile[32][33];
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  const int i0 = blockIdx.x * 32, b0 = blockIdx.y * 32;
#pragma unroll
  for (int r = 0; r < 32; r += 8) {
    if (TO_IL) {  // rows of the tile: vectors, read along the entries
      const int b = b0 + ty + r, i = i0 + tx;
      if (b < nrep && i < n) tile[ty + r][tx] = in[(int64_t)b * n + i];
    } else {      // rows of the tile: entries, read along the vectors
      const int i = i0 + ty + r, b = b0 + tx;
      if (i < n && b < nrep) tile[ty + r][tx] = in[(int64_t)i * nrep + b];
    }
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < 32; r += 8) {
    if (TO_IL) {
      const int i = i0 + ty + r, b = b0 + tx;
      if (i < n && b < nrep) {
        const double v = tile[tx][ty + r];
        out[(int64_t)i * nrep + b] = v;
        if (out_rd) out_rd[(int64_t)i * nrep + b] = v * rd[i];
      }
    } else {
      const int b = b0 + ty + r, i = i0 + tx;
      if (b < nrep && i < n) out[(int64_t)b * n + i] = tile[tx][ty + r];
    }
  }
}

// y = A x for nrep interleaved vectors: group of W lanes per (row j, W consecutive vectors),
// lane b of the group sums row j in k_spmv's order (sliced-ELL entry order: the diagonal, then
// the lower and the upper entries; contraction off), so each vector's result is bit-identical
// to its own SpMV.
template <typename T, int W>
__global__ __launch_bounds__(256) void k_spmm_il(int n, int nrep, const int* __restrict__ soff,
                                                 const int* __restrict__ scol,
                                                 const int* __restrict__ slen,
                                                 const T* __restrict__ sval,
                                                 const T* __restrict__ x, T* __restrict__ y) {
#pragma clang fp contract(off)
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int chunks = (nrep + W - 1) / W;
  const int64_t item = g / W;
  const int j = (int)(item / chunks);
  const int b = (int)(item % chunks) * W + (int)(g % W);
  if (j >= n || b >= nrep) return;
  const int base = soff[j >> 6] + (j & 63);
  const int len = slen[j];
  T s = sval[base] * x[(int64_t)scol[base] * nrep + b];
  int k = 1;
  for (; k + 8 <= len; k += 8) {
    int c[8];
    T v[8], xv[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      c[u] = scol[base + ((k + u) << 6)];
      v[u] = sval[base + ((k + u) << 6)];
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) xv[u] = x[(int64_t)c[u] * nrep + b];
#pragma unroll
    for (int u = 0; u < 8; ++u) s += v[u] * xv[u];
  }
  for (; k < len; ++k) {
    const int at = base + (k << 6);
    s += sval[at] * x[(int64_t)scol[at] * nrep + b];
  }
  y[(int64_t)j * nrep + b] = s;
}

// One Jacobi sweep of nrep interleaved vectors (rhs, in and out alike): group of W lanes per
// (row i, W consecutive vectors). For each vector the arithmetic is k_sweep's (the row summed
// in ascending column from 0, FP32 coefficients widened exactly, FP64 vectors, contraction
// off; `level`: the result rounded to level_value), so each is bit-identical to its own
// k_sweep; a first sweep reads in = rhs * rd written beforehand (the product k_sweep forms),
// and a weight w != 1 damps the sweep as k_sweep does. With out_rd, out * rd is written as well
// (the next system's first iterate).
template <bool D16, int W>
__global__ __launch_bounds__(256) void k_sweep_il(int n, int nrep, const int* __restrict__ soff,
                                                  const int* __restrict__ scol,
                                                  const uint16_t* __restrict__ sdel, int sign,
                                                  const int* __restrict__ slen,
                                                  const float* __restrict__ sval,
                                                  const double* __restrict__ rd,
                                                  const double* __restrict__ rhs,
                                                  const double* __restrict__ in,
                                                  double* __restrict__ out, int level,
                                                  double* __restrict__ out_rd, double w,
                                                  double c1) {
#pragma clang fp contract(off)
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int chunks = (nrep + W - 1) / W;
  const int64_t item = g / W;
  const int i = (int)(item / chunks);
  const int b = (int)(item % chunks) * W + (int)(g % W);
  if (i >= n || b >= nrep) return;
  const int base = soff[i >> 6] + (i & 63);
  const int len = slen[i];
  auto column = [&](int at) { return D16 ? i + sign * (int)sdel[at] : scol[at]; };
  auto iterate = [&](int c) { return in[(int64_t)c * nrep + b]; };
  double s = 0.0;
  int k = 0;
  for (; k + 8 <= len; k += 8) {
    int c[8];
    float v[8];
    double x[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      c[u] = column(base + ((k + u) << 6));
      v[u] = sval[base + ((k + u) << 6)];
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) x[u] = iterate(c[u]);
#pragma unroll
    for (int u = 0; u < 8; ++u) s = s + (double)v[u] * x[u];
  }
  for (; k < len; ++k) {
    const int at = base + (k << 6);
    s = s + (double)sval[at] * iterate(column(at));
  }
  const int64_t o = (int64_t)i * nrep + b;
  const double tv = (rhs[o] - s) * rd[i];
  const double y = w != 1.0 ? c1 * in[o] + w * tv : tv;
  out[o] = level ? level_value(y) : y;
  if (out_rd) out_rd[o] = y * rd[i];
}

// ------------------------------------------------------------------ fused Jacobi sweeps
constexpr int SWEEP_LANES = 256;  // threads of a workgroup; rows of an entry sub-block
constexpr int SWEEP_EXT = 1792;   // LDS slots for a block's external columns
constexpr int SWEEP_STAGE = 4;    // external values a lane has in flight while staging them

// All k Jacobi sweeps of one triangular system in one launch, with the arithmetic of k
// launches of k_sweep (each row summed in ascending column from 0, FP32 coefficients widened
// exactly, FP64 vectors, contraction off), so the result is bit-identical to theirs, and the
// factor's coefficients are read from memory once instead of k times.
//
// A persistent grid (every workgroup co-resident, shown by a probe of this kernel) takes the
// blocks of R = 256 RPL rows at positions w, w + G, ...: the lower system (L y = r) in
// ascending block order and the upper one (REV, L^T z = y) in descending order, so that every
// column a block reads from outside lies in a block at a lower position. Lane t owns rows
// b R + 256 h + t (h < RPL) and keeps the first MAXLEN coefficients and LDS slots of each in
// registers (a longer row re-reads the rest each sweep). The sweeps run over an LDS block
// holding the previous iterate at the block's own rows (slots 0..R-1) and at its distinct
// external columns (slots R..; the columns themselves are kept in LDS too): sweep 1 starts
// from rhs * rd; the result of sweep j < k is
// rounded to level_value and published (device-scope) into level j of lev with this launch's
// `tag` (1..255) in its low bits, and the external columns of that level are then awaited
// (bounded waits for the tag, as the triangular solves wait on the sentinel) and staged for
// sweep j + 1. A slot holds another launch's tag until this launch writes it, so the levels
// need no reset between launches as long as the host gives every launch since the levels
// were last cleared a tag of its own. The last sweep writes out (not rounded) and, with
// partial, each 256-row sub-block's share of dot.out (the partials of k_sweep's grid). A weight
// w != 1 damps every sweep as k_sweep does, from the row's previous iterate held in LDS.
// MINW: waves per SIMD the register allocation must allow (1: no bound).
// nblk < 0: co-residency probe of the grid.
template <int MAXLEN, bool REV, int MINW = 1, int RPL = 1>
__global__ __launch_bounds__(SWEEP_LANES) __attribute__((amdgpu_waves_per_eu(MINW))) void
k_sweeps_fused(int n, int nblk, int k, const int* __restrict__ eoff,
               const uint8_t* __restrict__ elen, const float* __restrict__ eval,
               const uint16_t* __restrict__ eidx, const int* __restrict__ xoff,
               const int* __restrict__ xcol, const double* __restrict__ rd,
               const double* __restrict__ rhs, double* __restrict__ lev, unsigned tag,
               double* __restrict__ out, const double* __restrict__ dot,
               double* __restrict__ partial, const PcgScalars* __restrict__ S,
               int* abort_flag, long long limit, double w, double c1) {
#pragma clang fp contract(off)
  static_assert(MAXLEN % 4 == 0, "entries are held in quads");
  constexpr int R = SWEEP_LANES * RPL;
  if (nblk < 0) {
    residency_probe(abort_flag, limit);
    return;
  }
  if (S && S->done) return;
  __shared__ double buf[R + SWEEP_EXT];
  __shared__ int cols[SWEEP_EXT];
  const int t = threadIdx.x;
  const int nsub = (n + SWEEP_LANES - 1) / SWEEP_LANES;
  for (int p = blockIdx.x; p < nblk; p += gridDim.x) {
    const int b = REV ? nblk - 1 - p : p;
    float v[RPL][MAXLEN];
    uint32_t ix[RPL][MAXLEN / 2];  // two 16-bit slots each
    int len[RPL], base[RPL];
    double ri[RPL], rdi[RPL], y[RPL];
#pragma unroll
    for (int h = 0; h < RPL; ++h) {
      const int i = b * R + h * SWEEP_LANES + t;
      len[h] = i < n ? (int)elen[i] : 0;
      // entry q of row i: base + (q >> 2) * 4 * SWEEP_LANES + (q & 3); a sub-block past the
      // last one has no live rows and is never read
      base[h] = eoff[min(b * RPL + h, nsub)] + 4 * t;
#pragma unroll
      for (int g = 0; g < MAXLEN / 4; ++g) {
        float4 f = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        uint2 d = make_uint2(0u, 0u);
        if (4 * g < len[h]) {
          const f32x4 fv =
              stream_load(reinterpret_cast<const f32x4*>(eval + base[h] + g * 4 * SWEEP_LANES));
          const u32x2 dv =
              stream_load(reinterpret_cast<const u32x2*>(eidx + base[h] + g * 4 * SWEEP_LANES));
          f = make_float4(fv.x, fv.y, fv.z, fv.w);
          d = make_uint2(dv.x, dv.y);
        }
        v[h][4 * g] = f.x;
        v[h][4 * g + 1] = f.y;
        v[h][4 * g + 2] = f.z;
        v[h][4 * g + 3] = f.w;
        ix[h][2 * g] = d.x;
        ix[h][2 * g + 1] = d.y;
      }
      rdi[h] = i < n ? rd[i] : 0.0;
      ri[h] = i < n ? rhs[i] : 0.0;
      buf[h * SWEEP_LANES + t] = ri[h] * rdi[h];
    }
    const int xb = xoff[b], nx = xoff[b + 1] - xb;
    // the block's external columns, kept in LDS for every sweep's staging
    for (int e = t; e < nx; e += SWEEP_LANES) {
      const int c = xcol[xb + e];
      cols[e] = c;
      buf[R + e] = rhs[c] * rd[c];
    }
    __syncthreads();
    for (int j = 1;; ++j) {
      // an empty statement that "changes" the rows' registers each sweep: otherwise the
      // widened coefficients and the LDS addresses, the same in every sweep, are computed
      // once before the loop and held, which triples the registers per entry
#pragma unroll
      for (int h = 0; h < RPL; ++h) {
#pragma unroll
        for (int q = 0; q < MAXLEN; ++q) asm volatile("" : "+v"(v[h][q]));
#pragma unroll
        for (int q = 0; q < MAXLEN / 2; ++q) asm volatile("" : "+v"(ix[h][q]));
      }
#pragma unroll
      for (int h = 0; h < RPL; ++h) {
        double s = 0.0;
#pragma unroll
        for (int q = 0; q < MAXLEN; ++q)
          if (q < len[h])
            s = s + (double)v[h][q] * buf[(ix[h][q >> 1] >> (16 * (q & 1))) & 0xffffu];
        for (int q = MAXLEN; q < len[h]; ++q) {
          const int at = base[h] + (q >> 2) * 4 * SWEEP_LANES + (q & 3);
          s = s + (double)eval[at] * buf[eidx[at]];
        }
        const double tv = (ri[h] - s) * rdi[h];
        y[h] = w != 1.0 ? c1 * buf[h * SWEEP_LANES + t] + w * tv : tv;
      }
      if (j >= k) break;
      double* level = lev + (int64_t)(j - 1) * n;
      double yl[RPL];
#pragma unroll
      for (int h = 0; h < RPL; ++h) {
        const int i = b * R + h * SWEEP_LANES + t;
        yl[h] = level_value(y[h]);
        if (i < n)
          __hip_atomic_store(reinterpret_cast<unsigned long long*>(level + i),
                             (unsigned long long)bits_of(yl[h]) | tag, __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_AGENT);
      }
      __syncthreads();  // every lane has read the previous iterate from LDS
#pragma unroll
      for (int h = 0; h < RPL; ++h) buf[h * SWEEP_LANES + t] = yl[h];
      for (int e0 = 0; e0 < nx; e0 += SWEEP_STAGE * SWEEP_LANES) {
        int c[SWEEP_STAGE];
        unsigned long long got[SWEEP_STAGE];
#pragma unroll
        for (int u = 0; u < SWEEP_STAGE; ++u) {
          const int e = e0 + u * SWEEP_LANES + t;
          c[u] = e < nx ? cols[e] : -1;
        }
#pragma unroll
        for (int u = 0; u < SWEEP_STAGE; ++u)
          got[u] = c[u] >= 0 ? __hip_atomic_load(reinterpret_cast<unsigned long long*>(level + c[u]),
                                                 __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                             : 0ull;
#pragma unroll
        for (int u = 0; u < SWEEP_STAGE; ++u)
          if (c[u] >= 0)
            buf[R + e0 + u * SWEEP_LANES + t] =
                (got[u] & LEVEL_TAG_MASK) == tag
                    ? __longlong_as_double((long long)(got[u] & ~LEVEL_TAG_MASK))
                    : await_tagged(level + c[u], tag, abort_flag, limit);
      }
      __syncthreads();
    }
#pragma unroll
    for (int h = 0; h < RPL; ++h) {
      const int i = b * R + h * SWEEP_LANES + t;
      if (i < n) out[i] = y[h];
      if (partial) {
        const double bs = block_sum<SWEEP_LANES>(i < n && dot ? dot[i] * y[h] : 0.0);
        if (t == 0 && b * RPL + h < nsub) partial[b * RPL + h] = bs;
      }
    }
    __syncthreads();  // the next block overwrites buf
  }
}

// ------------------------------------------------------------------ the reference's dot
constexpr int DOT8_THREADS = 256;
constexpr int DOT8_PER_THREAD = 16;                        // elements a thread stages per tile
constexpr int DOT8_TILE = DOT8_THREADS * DOT8_PER_THREAD;  // a multiple of 8

// dot(a, b) in the reference's order (solve.rs:100-110): eight lane sums over the chunks of
// eight, s_k += a[8c + k] * b[8c + k] for c in order, combined as ((s0 + s1) + (s2 + s3)) +
// ((s4 + s5) + (s6 + s7)), plus the products of the tail (n mod 8 elements) summed in order,
// added last. The eight sums are sequential chains, so one workgroup computes them: lanes 0-7
// of the first wave run the chains over the products of a tile held in LDS, while every
// thread loads the next tile and, once the chains are through, stores its products in the
// other LDS buffer. Each product is rounded once and each sum in the reference's order, so
// the result equals the reference's bit for bit. Written to out[0]; gated off by S->done.
__global__ __launch_bounds__(DOT8_THREADS) void k_dot8(int64_t n, const double* __restrict__ a,
                                                      const double* __restrict__ b,
                                                      double* __restrict__ out,
                                                      const PcgScalars* __restrict__ S) {
#pragma clang fp contract(off)
  if (S && S->done) return;
  __shared__ double prod[2][DOT8_TILE];
  const int t = threadIdx.x;
  const int64_t body = n / 8 * 8;  // elements in full chunks
  const int64_t tiles = (body + DOT8_TILE - 1) / DOT8_TILE;
  double s = 0.0;  // lane t < 8: the sum of lane t
  double va[DOT8_PER_THREAD], vb[DOT8_PER_THREAD];
  auto load = [&](int64_t tile) {
#pragma unroll
    for (int u = 0; u < DOT8_PER_THREAD; ++u) {
      const int64_t i = tile * DOT8_TILE + (int64_t)u * DOT8_THREADS + t;
      va[u] = i < body ? a[i] : 0.0;
      vb[u] = i < body ? b[i] : 0.0;
    }
  };
  auto store = [&](int buf) {
#pragma unroll
    for (int u = 0; u < DOT8_PER_THREAD; ++u) prod[buf][u * DOT8_THREADS + t] = va[u] * vb[u];
  };
  if (tiles > 0) {
    load(0);
    store(0);
  }
  __syncthreads();
  for (int64_t tile = 0; tile < tiles; ++tile) {
    const int buf = (int)(tile & 1);
    if (tile + 1 < tiles) load(tile + 1);
    if (t < 8) {
      const int64_t len = min<int64_t>(DOT8_TILE, body - tile * DOT8_TILE);
      const double* pr = prod[buf];
      for (int64_t c = 0; c < len / 8; ++c) s = s + pr[8 * c + t];
    }
    if (tile + 1 < tiles) store(buf ^ 1);  // the other buffer, read by no lane this tile
    __syncthreads();
  }
  __shared__ double lane[8];
  if (t < 8) lane[t] = s;
  __syncthreads();
  if (t == 0) {
    double tail = 0.0;
    for (int64_t i = body; i < n; ++i) tail = tail + a[i] * b[i];
    out[0] = ((lane[0] + lane[1]) + (lane[2] + lane[3])) + ((lane[4] + lane[5]) + (lane[6] + lane[7])) +
             tail;
  }
}

// ------------------------------------------------------------------ reductions
template <typename A, typename B>
__global__ __launch_bounds__(256) void k_dot(int64_t n, const A* __restrict__ a,
                                             const B* __restrict__ b, double* __restrict__ partial) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double v = i < n ? (double)a[i] * (double)b[i] : 0.0;
  double s = block_sum<256>(v);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// Sum of `count` partials in a fixed order, by one block of SUM_THREADS threads: each thread
// adds its strided share four loads at a time (four running sums, then combined), then a fixed
// tree over the block.
constexpr int SUM_THREADS = 1024;
__device__ inline double sum_partials(const double* __restrict__ partial, int count) {
  double v0 = 0.0, v1 = 0.0, v2 = 0.0, v3 = 0.0;
  int k = threadIdx.x;
  for (; k + 3 * SUM_THREADS < count; k += 4 * SUM_THREADS) {
    v0 += partial[k];
    v1 += partial[k + SUM_THREADS];
    v2 += partial[k + 2 * SUM_THREADS];
    v3 += partial[k + 3 * SUM_THREADS];
  }
  for (; k < count; k += SUM_THREADS) v0 += partial[k];
  return block_sum<SUM_THREADS>((v0 + v1) + (v2 + v3));
}

// r.r of b for the initial norm; b == 0 => done with x = 0 (iterative.rs:211-216)
__global__ __launch_bounds__(SUM_THREADS) void k_init_norm(const double* __restrict__ partial, int count,
                                                   PcgScalars* S, double tol) {
  double s = sum_partials(partial, count);
  if (threadIdx.x == 0) {
    S->bnorm = sqrt(s);
    S->rscale = tol * S->bnorm;
    S->it = 0;
    S->status = 0;
    S->done = S->bnorm == 0.0 ? 1 : 0;
    S->rmax = S->bnorm;
    S->prev_true = S->bnorm;
    S->update = 0;
    S->updates = 0;
  }
}

__global__ __launch_bounds__(SUM_THREADS) void k_init_rz(const double* __restrict__ partial, int count,
                                                 PcgScalars* S) {
  double s = sum_partials(partial, count);
  if (threadIdx.x == 0 && !S->done) S->rz = s;
}

// curvature p.Ap -> alpha, or Singular (iterative.rs:224-228)
__global__ __launch_bounds__(SUM_THREADS) void k_alpha(const double* __restrict__ partial, int count,
                                               PcgScalars* S) {
  if (S->done) return;
  double curvature = sum_partials(partial, count);
  if (threadIdx.x == 0) {
    if (isnan(curvature) || curvature <= 0.0) {
      S->status = 1;
      S->done = 1;
      S->it += 1;
    } else {
      S->alpha = S->rz / curvature;
    }
  }
}

// |r| after the update; converged (iterative.rs:233-235) or out of iterations (:244-247)
__global__ __launch_bounds__(SUM_THREADS) void k_converged(const double* __restrict__ partial, int count,
                                                   PcgScalars* S) {
  if (S->done) return;
  double rr = sum_partials(partial, count);
  if (threadIdx.x == 0) {
    S->rnorm = sqrt(rr);
    S->it += 1;
    if (S->allow && S->rnorm < S->delta * S->rmax) {
      S->update = 1;  // decide after the true residual (k_replaced)
    } else if (S->rnorm <= S->rscale) {
      S->done = 1;
      S->status = 0;
    } else if (S->it >= S->maxit) {
      S->done = 1;
      S->status = 2;
    }
  }
}

// After a residual replacement: the true |r| becomes the recursive one; replacements stop
// once one fails to halve the previous true residual (the FP64 floor is reached).
__global__ __launch_bounds__(SUM_THREADS) void k_replaced(const double* __restrict__ partial, int count,
                                                  PcgScalars* S) {
  if (S->done || !S->update) return;
  double rr = sum_partials(partial, count);
  if (threadIdx.x == 0) {
    const double t = sqrt(rr);
    if (t > 0.5 * S->prev_true) {
      S->allow = 0;
      if (S->stall_stop && t > S->rscale) {
        // the low-precision operator no longer reduces the true residual: hand the iterate
        // back so that the solve continues with the FP64 operator
        S->rnorm = t;
        S->updates += 1;
        S->update = 0;
        S->done = 1;
        S->status = 3;
        return;
      }
    }
    S->prev_true = t;
    S->rmax = t;
    S->rnorm = t;
    S->updates += 1;
    S->update = 0;
    if (t <= S->rscale) {
      S->done = 1;
      S->status = 0;
    } else if (S->it >= S->maxit) {
      S->done = 1;
      S->status = 2;
    }
  }
}

// r = b - A x (the FP64 true residual), its low-precision copy and partial r.r, only when a
// replacement is due
template <typename TM>
__global__ __launch_bounds__(256) void k_replace_residual(int n, const PcgScalars* __restrict__ S,
                                                          const double* __restrict__ b,
                                                          const double* __restrict__ ax,
                                                          double* __restrict__ r,
                                                          TM* __restrict__ r_low,
                                                          double* __restrict__ partial) {
  if (S->done || !S->update) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double v = 0.0;
  if (i < n) {
    const double ri = b[i] - ax[i];
    r[i] = ri;
    if (r_low) r_low[i] = (TM)ri;
    v = ri * ri;
  }
  double s = block_sum<256>(v);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// beta = rz' / rz (iterative.rs:237-239)
__global__ __launch_bounds__(SUM_THREADS) void k_beta(const double* __restrict__ partial, int count,
                                              PcgScalars* S) {
  if (S->done) return;
  double rz_next = sum_partials(partial, count);
  if (threadIdx.x == 0) {
    S->beta = rz_next / S->rz;
    S->rz = rz_next;
  }
}

// x += alpha p; r -= alpha Ap (iterative.rs:229-232); partial r.r; fills the two triangular
// solve outputs (when given) with the sentinel for the coming preconditioner application, or
// writes r * rd, the starting iterate of the coming Jacobi sweeps.
template <typename TV, typename TM>
__global__ __launch_bounds__(256) void k_update_xr(int n, const PcgScalars* __restrict__ S,
                                                   double* __restrict__ x, double* __restrict__ r,
                                                   const TV* __restrict__ p,
                                                   const TV* __restrict__ ap,
                                                   TM* __restrict__ r_low, TM* __restrict__ fill_a,
                                                   TM* __restrict__ fill_b,
                                                   double* __restrict__ partial,
                                                   const double* __restrict__ rd = nullptr,
                                                   double* __restrict__ r_rd = nullptr) {
#pragma clang fp contract(off)
  if (S->done) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double v = 0.0;
  if (i < n) {
    const double alpha = S->alpha;
    x[i] += alpha * (double)p[i];
    double ri = r[i] - alpha * (double)ap[i];
    r[i] = ri;
    if (r_low) r_low[i] = (TM)ri;
    if (r_rd) r_rd[i] = ri * rd[i];  // the first Jacobi sweep's starting iterate
    v = ri * ri;
    typedef typename Bits<TM>::U U;
    if (fill_a) reinterpret_cast<U*>(fill_a)[i] = Bits<TM>::S;
    if (fill_b) reinterpret_cast<U*>(fill_b)[i] = Bits<TM>::S;
  }
  double s = block_sum<256>(v);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// p = z + beta p (iterative.rs:240-242)
template <typename TV, typename TM>
__global__ __launch_bounds__(256) void k_update_p(int n, const PcgScalars* __restrict__ S,
                                                  TV* __restrict__ p, const TM* __restrict__ z) {
#pragma clang fp contract(off)
  if (S->done) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = (TV)((double)z[i] + S->beta * (double)p[i]);
}

// ------------------------------------------------------------------ batched PCG
// Several independent systems solved together: system s owns the vector entries
// [seg_off[s], seg_off[s+1]) and its own scalars S[s]; seg_of[i] is the system of entry i.
// Every system follows iterative.rs:210-248 on its own (its own alpha, beta and stop), so its
// iterates are those of a solve of that system alone.

// a.b over each system's entries: one workgroup per system, fixed order. Each thread keeps
// four sums (entries i, i + 256, i + 512, i + 768 of every 1,024), so that eight loads are in
// flight, and adds them as (s0 + s1) + (s2 + s3).
template <typename A, typename B>
__global__ __launch_bounds__(256) void k_seg_dot(const int64_t* __restrict__ seg_off,
                                                 const A* __restrict__ a,
                                                 const B* __restrict__ b,
                                                 double* __restrict__ out) {
  const int s = blockIdx.x;
  const int64_t hi = seg_off[s + 1];
  double v0 = 0.0, v1 = 0.0, v2 = 0.0, v3 = 0.0;
  int64_t i = seg_off[s] + threadIdx.x;
  for (; i + 768 < hi; i += 1024) {
    v0 += (double)a[i] * (double)b[i];
    v1 += (double)a[i + 256] * (double)b[i + 256];
    v2 += (double)a[i + 512] * (double)b[i + 512];
    v3 += (double)a[i + 768] * (double)b[i + 768];
  }
  for (; i < hi; i += 256) v0 += (double)a[i] * (double)b[i];
  const double t = block_sum<256>((v0 + v1) + (v2 + v3));
  if (threadIdx.x == 0) out[s] = t;
}

__global__ void k_seg_init(int nseg, const double* __restrict__ bb, PcgScalars* S, double tol,
                           int64_t maxit) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nseg) return;
  PcgScalars& q = S[s];
  q.bnorm = sqrt(bb[s]);
  q.rscale = tol * q.bnorm;
  q.it = 0;
  q.maxit = maxit;
  q.status = 0;
  q.done = q.bnorm == 0.0 ? 1 : 0;  // b == 0: x = 0 exactly
  q.rnorm = q.bnorm;
}

__global__ void k_seg_rz(int nseg, const double* __restrict__ rz, PcgScalars* S) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s < nseg && !S[s].done) S[s].rz = rz[s];
}

__global__ void k_seg_alpha(int nseg, const double* __restrict__ pap, PcgScalars* S) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nseg || S[s].done) return;
  const double curvature = pap[s];
  if (isnan(curvature) || curvature <= 0.0) {
    S[s].status = 1;
    S[s].done = 1;
    S[s].it += 1;
  } else {
    S[s].alpha = S[s].rz / curvature;
  }
}

__global__ void k_seg_converged(int nseg, const double* __restrict__ rr, PcgScalars* S) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nseg || S[s].done) return;
  S[s].rnorm = sqrt(rr[s]);
  S[s].it += 1;
  if (S[s].rnorm <= S[s].rscale) {
    S[s].done = 1;
  } else if (S[s].it >= S[s].maxit) {
    S[s].done = 1;
    S[s].status = 2;
  }
}

__global__ void k_seg_beta(int nseg, const double* __restrict__ rz, PcgScalars* S) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nseg || S[s].done) return;
  S[s].beta = rz[s] / S[s].rz;
  S[s].rz = rz[s];
}

// x += alpha p; r -= alpha Ap on the systems still running; refills the triangular-solve
// outputs with the sentinel everywhere.
template <typename TM>
__global__ __launch_bounds__(256) void k_seg_update_xr(int64_t n, const int* __restrict__ seg_of,
                                                       const PcgScalars* __restrict__ S,
                                                       double* __restrict__ x,
                                                       double* __restrict__ r,
                                                       const double* __restrict__ p,
                                                       const double* __restrict__ ap,
                                                       TM* __restrict__ r_low,
                                                       TM* __restrict__ fill_a,
                                                       TM* __restrict__ fill_b) {
#pragma clang fp contract(off)
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  typedef typename Bits<TM>::U U;
  reinterpret_cast<U*>(fill_a)[i] = Bits<TM>::S;
  reinterpret_cast<U*>(fill_b)[i] = Bits<TM>::S;
  const PcgScalars& q = S[seg_of[i]];
  if (q.done) return;
  const double alpha = q.alpha;
  x[i] += alpha * p[i];
  const double ri = r[i] - alpha * ap[i];
  r[i] = ri;
  if (r_low) r_low[i] = (TM)ri;
}

// p = z + beta p on the systems still running
template <typename TM>
__global__ __launch_bounds__(256) void k_seg_update_p(int64_t n, const int* __restrict__ seg_of,
                                                      const PcgScalars* __restrict__ S,
                                                      double* __restrict__ p,
                                                      const TM* __restrict__ z) {
#pragma clang fp contract(off)
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const PcgScalars& q = S[seg_of[i]];
  if (q.done) return;
  p[i] = (double)z[i] + q.beta * p[i];
}

// r = b - A x in FP64 (the refinement's true residual), partial r.r
__global__ __launch_bounds__(256) void k_residual(int n, const double* __restrict__ b,
                                                  const double* __restrict__ ax,
                                                  double* __restrict__ r,
                                                  double* __restrict__ partial) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double v = 0.0;
  if (i < n) {
    double ri = b[i] - ax[i];
    r[i] = ri;
    v = ri * ri;
  }
  double s = block_sum<256>(v);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

__global__ __launch_bounds__(256) void k_axpy_add(int n, double* __restrict__ x,
                                                  const double* __restrict__ d) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] += d[i];
}

// ------------------------------------------------------------------ triangular solves
// Dataflow solve of a lower-triangular system in "logical" row order. Vectors are indexed
// physically, phys(l) = REV ? n-1-l : l, so the backward solve L^T z = y runs through the same
// kernel on the transposed layout.
//
//   SUBSEQ = false: out_l = (rhs_l - sum_p val_p out_col) / diag_l  (forward loop,
//                   iterative.rs:175-179: sum in ascending column order, then subtract)
//   SUBSEQ = true:  out_l = ((rhs_l - val_1 out_1) - val_2 out_2 ...) / diag_l (backward loop,
//                   iterative.rs:180-186: subtractions in decreasing row order)
//
// Rows are grouped on the host into blocks: chains of at most 64 rows along each row's critical
// dependency, stored as consecutive slots (bstart[]; brow[] is the row of a slot). Each slot's
// entries are split into those whose column lies in another block ("external": a coefficient
// and the index xu of its column in the block's list ucol of distinct external columns) and
// those inside it (ik = position of the row in the block, ival); the internal columns are the
// row's last, so externals then internals is the row's own order. Both parts are stored slot
// after slot.
//
// Workgroup w of the G-workgroup grid takes the blocks at positions w, w+G, ... of the host's
// processing order border[], which is topological (a block's external columns all lie in blocks
// earlier in border[]): with every workgroup resident the lowest unfinished block always has
// its inputs and a running workgroup, so the solve progresses. A workgroup is three waves:
//  - two pollers, half a round apart, load at device scope the block's distinct external
//    columns that are still missing, round after round, and drop each value into LDS as it
//    arrives;
//  - the solver never waits on global memory: every lane consumes, in order, its external
//    values as they appear in LDS, and the rows are resolved one after the other. Lane k
//    finishes its row, publishes and broadcasts it, and each lane whose next internal entry is
//    that row consumes it at once. Rows whose external values are all in run back to back in a
//    branch-free loop; a lane that completes late catches up on the rows solved meanwhile.
// Splitting the roles keeps the publishing stores of the solver out of the pollers' waits (one
// counter covers both loads and stores of a wave) and keeps the polls going while rows run.
// Every row is summed in the reference's order.
template <typename T, bool SUBSEQ>
__device__ inline T accumulate(T acc, T a, T v) {
#pragma clang fp contract(off)
  return SUBSEQ ? acc - a * v : acc + a * v;
}

constexpr int TRI_XCAP = 512;  // external entries a block may hold (host-enforced)
constexpr int TRI_ICAP = 512;  // internal entries a block may hold (host-enforced)
constexpr int TRI_UCAP = 256;  // distinct external columns a block may hold (host-enforced)
constexpr int TRI_THREADS = 192;  // three waves: the solver (lanes 0-63) and two pollers

// LDS words shared by the waves of a workgroup, read and written whole.
template <typename U>
__device__ inline U lds_load(const U* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
template <typename U>
__device__ inline void lds_store(U* p, U v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// TRACE: for diagnostics, each block records its timeline in trace[10 b .. 10 b + 9] and the
// publish time of each slot's row in trace[10 nblocks + slot] (first right-hand side only).
// T is the arithmetic of the vectors and sums, TL that of the stored coefficients (FP32 for the
// mixed modes: each coefficient is widened exactly to T as it is used, so the solve is the
// reference's FP64 solve applied to the FP32-rounded factor).
template <typename T, bool REV, bool SUBSEQ, bool TRACE = false, typename TL = T>
__global__ __launch_bounds__(TRI_THREADS) void k_trisolve(
    int n, int nblocks, const int* __restrict__ border, const int* __restrict__ bstart,
    const int* __restrict__ brow, const int* __restrict__ xoff, const int* __restrict__ uoff,
    const int* __restrict__ ucol, const unsigned char* __restrict__ xu,
    const TL* __restrict__ xval, const int* __restrict__ ioff, const unsigned char* __restrict__ ik,
    const TL* __restrict__ ival, const TL* __restrict__ diag, const double* __restrict__ rdiag,
    const T* __restrict__ rhs, T* out, const PcgScalars* __restrict__ S, int* abort_flag,
    long long limit, long long* __restrict__ trace, int nrep) {
#pragma clang fp contract(off)
  static_assert(std::is_same<T, double>::value, "the solves run on FP64 vectors and sums");
  if (nrep < 0) {  // a co-residency probe of this kernel (abort_flag: the probe's two words)
    residency_probe(abort_flag, limit);
    return;
  }
  if (S && S->done) return;
  typedef typename Bits<T>::U U;
  __shared__ U suv[TRI_UCAP];    // values of the distinct external columns (sentinel: not yet)
  __shared__ int suc[TRI_UCAP];  // their physical indices
  __shared__ TL sxa[TRI_XCAP];   // external coefficients
  __shared__ unsigned char sxu[TRI_XCAP];  // their columns, as indices into suv
  // one more internal entry than the layout can hold: position 64 (no row of a block), read
  // ahead past the end of a lane's list
  __shared__ TL sia[TRI_ICAP + 1];
  __shared__ unsigned char sik[TRI_ICAP + 1];
  __shared__ T own[64];     // the rows of the block solved so far, copied when the solver waits
  __shared__ int s_state;   // 0: rows running; 1: rows done; 2: a wait gave up
  __shared__ long long s_poll[3];  // TRACE: the poller's rounds, load ticks and loads
  const int lane = threadIdx.x & 63;
  const int wv = threadIdx.x >> 6;  // 0: the solver; 1, 2: the pollers
  const bool poller = wv > 0;
  T* const out0 = out;
  const T* const rhs0 = rhs;
  for (int pos = blockIdx.x; pos < nblocks * nrep; pos += gridDim.x) {
    // positions run block-major over the nrep right-hand sides (stored one after the other),
    // which keeps the processing order topological for each of them
    const int kk = pos / nrep;
    const int rep = pos - kk * nrep;
    out = out0 + (int64_t)rep * n;
    rhs = rhs0 + (int64_t)rep * n;
    const int b = border[kk];
    const int s = bstart[b];
    const int e = bstart[b + 1];
    const int rows = e - s;
    const int xb = xoff[s], nx = xoff[e] - xb;
    const int ib = ioff[s], ni = ioff[e] - ib;
    const int ub = uoff[b], nu = uoff[b + 1] - ub;
    const long long tb = TRACE ? wall_clock64() : 0;
    // Waits of the block are bounded from its start.
    const long long t0 = wall_clock64();
    // per-lane state of the solver
    int p = 0, x1 = 0, ip = 0, i1 = 0, pl = 0, l = 0;
    T acc = T(0), rl = T(0), dl = T(1);
    double rd = 1.0;  // RN64(1 / diag), see the row loop
    // stage the block: the poller the distinct external columns with the values published so
    // far, read at device scope (a plain load could hit a stale line of this XCD's L2 that
    // still holds the sentinel); the solver the external entries (coefficient, column index)
    // and the internal ones
    if (wv == 1) {
#pragma unroll 4
      for (int q = lane; q < nu; q += 64) {
        const int c = ucol[ub + q];
        const int pc = REV ? n - 1 - c : c;
        suc[q] = pc;
        suv[q] = __hip_atomic_load(reinterpret_cast<U*>(out + pc), __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
      }
    } else if (wv == 0) {
#pragma unroll 4
      for (int q = lane; q < nx; q += 64) {
        sxu[q] = xu[xb + q];
        sxa[q] = xval[xb + q];
      }
#pragma unroll 4
      for (int q = lane; q < ni; q += 64) {
        sik[q] = ik[ib + q];
        sia[q] = ival[ib + q];
      }
      if (lane == 0) {
        sik[TRI_ICAP] = 64;
        sia[TRI_ICAP] = TL(0);
        s_state = 0;
      }
      if (lane < rows) {
        l = s + lane;  // slot
        const int row = brow[l];
        pl = REV ? n - 1 - row : row;
        p = xoff[l] - xb;
        x1 = xoff[l + 1] - xb;
        ip = ioff[l] - ib;
        i1 = ioff[l + 1] - ib;
        rl = rhs[pl];
        rd = rdiag[l];
        dl = (T)diag[l];
        acc = SUBSEQ ? rl : T(0);
      }
    }
    __syncthreads();
    if (poller) {
      // One round trip per round: every distinct column still missing is loaded again, until
      // all are in or the rows are done. The clock and the abort flag are read on the 1st,
      // 17th, 33rd ... round only (each is another round trip).
      long long t_load = 0;
      int rounds = 0, n_polled = 0;
      // the second poller starts half a round later, so that between them a missing column is
      // looked at twice per round trip
      if (wv == 2) __builtin_amdgcn_s_sleep(16);
      for (int it = 0;; ++it) {
        const long long ta = TRACE ? wall_clock64() : 0;
        bool missing = false;
#pragma unroll 2
        for (int q = lane; q < nu; q += 64) {
          if (lds_load(&suv[q]) == Bits<T>::S) {
            if (TRACE) ++n_polled;
            const U u = __hip_atomic_load(reinterpret_cast<U*>(out + suc[q]), __ATOMIC_RELAXED,
                                          __HIP_MEMORY_SCOPE_AGENT);
            if (u != Bits<T>::S)
              lds_store(&suv[q], u);
            else
              missing = true;
          }
        }
        if (TRACE) {
          t_load += wall_clock64() - ta;
          ++rounds;
        }
        if (!__ballot(missing) || lds_load(&s_state) != 0) break;
        if ((it & 15) == 0) {
          const bool give_up =
              __hip_atomic_load(abort_flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) ||
              wall_clock64() - t0 > limit;
          if (__builtin_amdgcn_readfirstlane((int)give_up)) {
            // give up: the host reports ML_DEVICE_TIMEOUT; the solver stops waiting
            __hip_atomic_store(abort_flag, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (lane == 0) lds_store(&s_state, 2);
            break;
          }
        }
        __builtin_amdgcn_s_sleep(1);
      }
      if (TRACE) {
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) n_polled += __shfl_xor(n_polled, o);
        if (wv == 1 && lane == 0) {
          s_poll[0] = rounds;
          s_poll[1] = t_load;
          s_poll[2] = n_polled;
        }
      }
    } else {
      // Each lane consumes its external entries, in order, as far as their values are in; the
      // entries are looked up eight at a time so that the LDS reads overlap. The index of the
      // column a lane waits on is kept, so that a lane still waiting costs one LDS read.
      int wu = p < x1 ? (int)sxu[p] : 0;
      auto advance = [&]() {
        for (bool go = p < x1 && lds_load(&suv[wu]) != Bits<T>::S; go;) {
          int ju[8];
          T aa[8];
          U vv[8];
#pragma unroll
          for (int u = 0; u < 8; ++u) ju[u] = p + u < x1 ? (int)sxu[p + u] : -1;
#pragma unroll
          for (int u = 0; u < 8; ++u) {
            vv[u] = ju[u] >= 0 ? lds_load(&suv[ju[u]]) : Bits<T>::S;
            aa[u] = ju[u] >= 0 ? (T)sxa[p + u] : T(0);
          }
#pragma unroll
          for (int u = 0; u < 8; ++u) {
            if (go && ju[u] >= 0 && vv[u] != Bits<T>::S) {
              acc = accumulate<T, SUBSEQ>(acc, aa[u], from_bits<T>(vv[u]));
              ++p;
            } else {
              go = false;
            }
          }
          go = go && p < x1;
        }
        wu = p < x1 ? (int)sxu[p] : 0;
      };
      advance();
      long long tp = 0;
      int late = 0;
      if (TRACE) {
        tp = wall_clock64();
        late = __popcll(__ballot(p < x1));
      }
      // The lane's next internal entry (nk: the block position of the row it reads, nv: its
      // coefficient) and the one after, read ahead so that consuming a row never waits on LDS.
      // A lane whose external part is still pending holds nk = 64, which matches no row, until
      // it catches up.
      int nk = 64, nk2 = 64;
      // kept in the stored precision and widened where used, so that the read-ahead never
      // waits on its LDS load inside the row loop
      TL nv = TL(0), nv2 = TL(0);
      T yown = T(0);  // the lane's solved value, once its row is done
      auto refill = [&]() {
        nk = ip < i1 ? (int)sik[ip] : 64;
        nv = ip < i1 ? sia[ip] : TL(0);
        nk2 = ip + 1 < i1 ? (int)sik[ip + 1] : 64;
        nv2 = ip + 1 < i1 ? sia[ip + 1] : TL(0);
      };
      if (p == x1) refill();
      long long waited = 0;
      int spins = 0;
      // lanes whose external part is still pending: none below the current row, whose
      // predecessors in the block have all been solved
      unsigned long long pend = __ballot(p < x1);
      int k = 0;
      while (k < rows) {
        if ((pend >> k) & 1ull) {
          // Row k waits for external values: the lanes take in what the poller has dropped
          // into LDS, and a lane whose external part completes catches up, in order, on its
          // internal entries of the rows solved so far (four at a time, from own[]).
          const long long tw = TRACE ? wall_clock64() : 0;
          own[lane] = yown;
          wave_sync();
          do {
            const unsigned long long before = pend;
            advance();
            if (((before >> lane) & 1ull) && p == x1) {
              for (bool go = ip < i1; go;) {
                int kq[4];
                T aa[4], vv[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) kq[u] = ip + u < i1 ? (int)sik[ip + u] : 64;
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                  vv[u] = kq[u] < k ? own[kq[u]] : T(0);
                  aa[u] = kq[u] < k ? (T)sia[ip + u] : T(0);
                }
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                  if (go && kq[u] < k) {
                    acc = accumulate<T, SUBSEQ>(acc, aa[u], vv[u]);
                    ++ip;
                  } else {
                    go = false;
                  }
                }
                go = go && ip < i1;
              }
              refill();
            }
            pend = __ballot(p < x1);
            if ((pend >> k) & 1ull) {
              // the poller gave up, or (every 16th spin) this wait has outlived the bound
              bool give_up = lds_load(&s_state) == 2;
              if (!give_up && (spins & 15) == 0) give_up = wall_clock64() - t0 > limit;
              ++spins;
              if (__builtin_amdgcn_readfirstlane((int)give_up)) {
                // the host reports ML_DEVICE_TIMEOUT; every lane stops waiting
                __hip_atomic_store(abort_flag, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                if (p < x1) {
                  p = x1;
                  while (ip < i1 && (int)sik[ip] < k) ++ip;
                  refill();
                }
                pend = 0;
              } else {
                __builtin_amdgcn_s_sleep(1);
              }
            }
          } while ((pend >> k) & 1ull);
          if (TRACE) waited += wall_clock64() - tw;
        }
        // Rows k .. kend-1 have all their external values: solve them back to back. Loads left
        // in flight are drained first, so that the row loop only ever waits on its own
        // read-ahead, issued a row earlier. The loop has no branch but the one around the
        // division: every lane takes the same steps, selecting on whether it reads row k.
        const int kend = pend ? min(rows, (int)__builtin_ctzll(pend)) : rows;
        __builtin_amdgcn_s_waitcnt(0x0070);  // vmcnt(0) lgkmcnt(0)
        for (; k < kend; ++k) {
          T y = acc;
          if (lane == k) {
            // the row's quotient, the division's bit for bit, from the reciprocal of the
            // diagonal and two exact-remainder corrections (quotient64): five dependent
            // operations on the row chain instead of the division sequence
            y = quotient64(SUBSEQ ? acc : rl - acc, dl, rd);
            yown = y;
            publish(out + pl, y);
            if (TRACE && rep == 0) trace[10 * (long long)nblocks + l] = wall_clock64();
          }
          y = lane_broadcast(y, k);
          const bool hit = nk == k;
          const T upd = accumulate<T, SUBSEQ>(acc, (T)nv, y);
          acc = hit ? upd : acc;
          ip += hit ? 1 : 0;
          nk = hit ? nk2 : nk;
          nv = hit ? nv2 : nv;
          // the entry after next (the same one again when the lane did not advance), or the
          // terminal position 64 past the end of the list
          const int q = ip + 1 < i1 ? ip + 1 : TRI_ICAP;
          nk2 = sik[q];
          nv2 = sia[q];
        }
      }
      if (lane == 0) lds_store(&s_state, 1);  // the poller stops
      if (TRACE && lane == 0 && rep == 0) {
        trace[10 * (long long)b + 0] = tb;
        trace[10 * (long long)b + 1] = tp;
        trace[10 * (long long)b + 3] = ((long long)blockIdx.x << 8) | late;
        trace[10 * (long long)b + 4] = waited;
        trace[10 * (long long)b + 7] = spins;
        trace[10 * (long long)b + 9] = nu;
      }
    }
    __syncthreads();
    if (TRACE && threadIdx.x == 0 && rep == 0) {
      trace[10 * (long long)b + 2] = wall_clock64();
      trace[10 * (long long)b + 5] = s_poll[0];
      trace[10 * (long long)b + 6] = s_poll[1];
      trace[10 * (long long)b + 8] = s_poll[2];
    }
  }
}

// ------------------------------------------------------------------ IC(0)
// Csr::incomplete_cholesky (iterative.rs:146-170), one wavefront per row. Wave w of the G-wave
// grid takes the rows at positions w, w+G, ... of order[], the rows sorted by their level in the
// dependency DAG: a row only reads rows at earlier positions, so with every wave resident the
// launch progresses, and the rows a wave takes one after the other are a few levels apart, so
// it seldom holds a row long before that row's inputs exist.
// Row i: for each entry p (column j, ascending) the sum over row j's entries whose column k is
// also a computed column of row i is taken in ascending k order, exactly the reference's sum
// over its dense work vector; l_ij = (a_ij - s) / l_jj; then the pivot a_ii(1+shift) - sum of
// squares must exceed 1e-12 a_ii. Rows of L are published entry by entry; a failed pivot
// raises *fail and every later row only writes zeros so the launch drains quickly.
// Loads run ahead of the sums, so that their round trips overlap instead of adding up along
// the dependency chain: at the row's start every entry's row bounds and diagonal l_jj are
// loaded at once, and while entry p is summed the columns and values of the next entry's row
// are loaded. A slot still holding the sentinel when it was loaded is awaited as before; slots
// are written once, so the values, and the sums, are those of the loads done in order.
template <typename T, int MAXROW>
__global__ __launch_bounds__(64) void k_ic0(int n, const int* __restrict__ order,
                                            const int* __restrict__ off,
                                            const int* __restrict__ col,
                                            const T* __restrict__ a, T* l, T shift,
                                            int* fail, int* abort_flag, long long limit,
                                            int* where, const int* __restrict__ seg,
                                            const T* __restrict__ seg_shift) {
#pragma clang fp contract(off)
  if (n < 0) {  // a co-residency probe of this kernel (fail: the probe's two words)
    residency_probe(fail, limit);
    return;
  }
  typedef typename Bits<T>::U U;
  constexpr int PRE = MAXROW < 64 ? MAXROW : 64;  // entries whose rows are looked up at the start
  __shared__ int scol[MAXROW];
  __shared__ T sw[MAXROW];
  __shared__ T sa[MAXROW];
  __shared__ int sjs[PRE];  // first strict entry of each column's row
  __shared__ int sjd[PRE];  // diagonal entry of each column's row
  __shared__ U sdg[PRE];    // that diagonal as loaded at the row's start (maybe the sentinel)
  const int lane = threadIdx.x;
  auto row_js = [&](int q) { return q < PRE ? sjs[q] : off[scol[q]]; };
  auto row_jd = [&](int q) { return q < PRE ? sjd[q] : off[scol[q] + 1] - 1; };
  // the first 64 strict entries of row scol[q]: lane's column (-1 past the row) and value bits
  auto fetch = [&](int q, int& k, U& v) {
    const int idx = row_js(q) + lane;
    k = -1;
    v = Bits<T>::S;
    if (idx < row_jd(q)) {
      k = col[idx];
      v = __hip_atomic_load(reinterpret_cast<U*>(l + idx), __ATOMIC_RELAXED,
                            __HIP_MEMORY_SCOPE_AGENT);
    }
  };
  for (int pos = blockIdx.x; pos < n; pos += gridDim.x) {
    const int i = order[pos];
    const int start = off[i], dg = off[i + 1] - 1, len = dg - start;
    // batched systems: each has its own shift and failure flag
    int* const my_fail = seg ? fail + seg[i] : fail;
    const T my_shift = seg ? seg_shift[seg[i]] : shift;
    const bool failed = __builtin_amdgcn_readfirstlane(
                            __hip_atomic_load(my_fail, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) != 0;
    if (failed || len > MAXROW) {
      if (len > MAXROW && lane == 0)
        __hip_atomic_store(my_fail, 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      for (int t = lane; t <= len; t += 64) publish(l + start + t, T(0));
    } else {
      for (int t = lane; t < len; t += 64) {
        const int j = col[start + t];
        scol[t] = j;
        sa[t] = a[start + t];
        if (t < PRE) {
          const int jd = off[j + 1] - 1;
          sjs[t] = off[j];
          sjd[t] = jd;
          sdg[t] = __hip_atomic_load(reinterpret_cast<U*>(l + jd), __ATOMIC_RELAXED,
                                     __HIP_MEMORY_SCOPE_AGENT);
        }
      }
      wave_sync();
      int ck = -1, nk = -1;
      U cv = Bits<T>::S, nv = Bits<T>::S;
      if (len > 0) fetch(0, ck, cv);
      for (int q = 0; q < len; ++q) {
        if (q + 1 < len) fetch(q + 1, nk, nv);  // used by the next entry
        const int j = scol[q];
        const int js = row_js(q), jd = row_jd(q);
        T sacc = T(0);
        for (int base = js; base < jd; base += 64) {
          const int idx = base + lane;
          const bool first = base == js;  // the prefetched chunk
          T term = T(0);
          bool hit = false;
          if (idx < jd) {
            const int k = first ? ck : col[idx];
            int lo = 0, hi = q;  // search k among the computed columns scol[0..q)
            while (lo < hi) {
              int mid = (lo + hi) >> 1;
              if (scol[mid] < k) lo = mid + 1;
              else hi = mid;
            }
            if (lo < q && scol[lo] == k) {
              const T ljk = first && cv != Bits<T>::S ? from_bits<T>(cv)
                                                      : await_value(l + idx, abort_flag, limit);
              term = ljk * sw[lo];
              hit = true;
            }
          }
          unsigned long long mask = __ballot(hit);
          while (mask) {
            const int bl = __builtin_ctzll(mask);
            sacc = sacc + lane_broadcast(term, bl);
            mask &= mask - 1;
          }
        }
        const U dg = q < PRE ? sdg[q] : Bits<T>::S;
        const T ljj = dg != Bits<T>::S ? from_bits<T>(dg) : await_value(l + jd, abort_flag, limit);
        if (lane == 0 && __hip_atomic_load(abort_flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) &&
            __hip_atomic_load(where, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < 0) {
          __hip_atomic_store(where, i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          __hip_atomic_store(where + 1, j, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          __hip_atomic_store(where + 2, (int)blockIdx.x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          __hip_atomic_store(where + 3, (int)gridDim.x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        const T lij = (sa[q] - sacc) / ljj;
        if (lane == 0) sw[q] = lij;
        wave_sync();
        ck = nk;
        cv = nv;
      }
      T squares = T(0);
      for (int q = 0; q < len; ++q) squares = squares + sw[q] * sw[q];
      const T pivot = a[dg] * (T(1) + my_shift) - squares;
      const bool bad = isnan(pivot) || pivot <= T(1e-12) * a[dg];
      if (bad && lane == 0) __hip_atomic_store(my_fail, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      for (int t = lane; t < len; t += 64) publish(l + start + t, sw[t]);
      if (lane == 0) publish(l + dg, bad ? T(0) : (T)sqrt(pivot));
    }
    wave_sync();
  }
}

}  // namespace mlk
