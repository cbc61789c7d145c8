// Host orchestration of the device solve path and the C ABI declared in include/maille_hip.h.
//
//   Operator : K_ff on the device (reference lower CSR, sliced-ELL copy for the SpMV, forward
//              and transposed layouts of the IC(0) factor) + the PCG driver (FP64 drop-in, or
//              FP32 operator and/or FP32 preconditioner with FP64 residual replacement).
//   Model    : mesh -> host plan (pattern.cpp, plan.cpp) -> element kernels and gather
//              assembly on the device -> Operator -> reactions and stresses.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstring>
#include <future>
#include <limits>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "../../include/maille_hip.h"
#include "kernels.hip"
#include "msh.hpp"
#include "pattern.hpp"
#include "plan.hpp"

namespace ml {

namespace {

thread_local std::string g_last_error;

struct HipError : std::runtime_error {
  using std::runtime_error::runtime_error;
};
struct ArgError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

#define HIP_CHECK(expr)                                                                 \
  do {                                                                                  \
    hipError_t e_ = (expr);                                                             \
    if (e_ != hipSuccess)                                                               \
      throw HipError(std::string(#expr) + ": " + hipGetErrorString(e_));                \
  } while (0)

double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// Device allocation owned by a handle.
template <typename T>
struct DBuf {
  T* p = nullptr;
  size_t n = 0;
  DBuf() = default;
  DBuf(const DBuf&) = delete;
  DBuf& operator=(const DBuf&) = delete;
  ~DBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  // n is the requested count, also when it is zero (one element is still allocated so that
  // p is a valid pointer, but nothing may be read from it)
  void alloc(size_t count) {
    if (count == n && p) return;
    release();
    HIP_CHECK(hipMalloc((void**)&p, std::max<size_t>(count, 1) * sizeof(T)));
    n = count;
  }
  void upload(const T* host, size_t count, hipStream_t s) {
    alloc(count);
    if (count) HIP_CHECK(hipMemcpyAsync(p, host, count * sizeof(T), hipMemcpyHostToDevice, s));
  }
  template <class V>
  void upload_vec(const V& v, hipStream_t s) {
    upload(v.data(), v.size(), s);
  }
};

inline unsigned grid_for(int64_t n, int bs = 256) { return (unsigned)((n + bs - 1) / bs); }

// per-block capacities of the triangular-solve layouts, as the kernel's LDS holds them
constexpr TriCaps kTriCaps = {mlk::TRI_XCAP, mlk::TRI_ICAP, mlk::TRI_UCAP};
// the fused sweep layouts are built for the kernel's block and LDS capacities
static_assert(FUSED_LANES == mlk::SWEEP_LANES && FUSED_EXT == mlk::SWEEP_EXT,
              "fused sweep layout and kernel disagree");

// Kernel-duration accounting with HIP events on the handle's stream (enabled on demand).
enum ProfKind {
  P_SPMV64, P_SPMV32, P_TRI64, P_TRI32, P_IC0, P_STIFF, P_ASM, P_VEC, P_STRESS, P_NODAL, P_COUNT
};

struct Profiler {
  bool on = false;
  struct Rec { int kind; hipEvent_t a, b; };
  std::vector<Rec> pending;
  double ms[P_COUNT] = {0};
  int64_t count[P_COUNT] = {0};
  hipEvent_t begin(hipStream_t s) {
    if (!on) return nullptr;
    hipEvent_t e;
    HIP_CHECK(hipEventCreate(&e));
    HIP_CHECK(hipEventRecord(e, s));
    return e;
  }
  void end(int kind, hipEvent_t a, hipStream_t s) {
    if (!on || !a) return;
    hipEvent_t b;
    HIP_CHECK(hipEventCreate(&b));
    HIP_CHECK(hipEventRecord(b, s));
    pending.push_back({kind, a, b});
  }
  void resolve() {
    for (auto& r : pending) {
      HIP_CHECK(hipEventSynchronize(r.b));
      float t = 0;
      HIP_CHECK(hipEventElapsedTime(&t, r.a, r.b));
      ms[r.kind] += t;
      count[r.kind] += 1;
      (void)hipEventDestroy(r.a);
      (void)hipEventDestroy(r.b);
    }
    pending.clear();
  }
  void reset() {
    resolve();
    for (int k = 0; k < P_COUNT; ++k) { ms[k] = 0; count[k] = 0; }
  }
};

int device_cus() {
  int dev = 0, cus = 0;
  HIP_CHECK(hipGetDevice(&dev));
  HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  return cus;
}

int occupancy_per_cu(const void* kernel, int threads) {
  int per_cu = 0;
  HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, threads, 0));
  return std::max(1, per_cu);
}

// The bound of every dataflow wait: 2 s of the device's steady wall clock (wall_clock64 ticks
// at hipDeviceAttributeWallClockRate kHz).
long long wait_limit() {
  int dev = 0, khz = 0;
  HIP_CHECK(hipGetDevice(&dev));
  HIP_CHECK(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev));
  if (khz <= 0) khz = 100000;
  return 2000ll * khz;
}

// Workgroups of a persistent dataflow kernel to launch per CU. The dataflow kernels need every
// workgroup of their grid resident (the lowest unfinished block must always have a running
// workgroup), and the runtime's occupancy figure over-reports what gfx950 holds: the probes
// below found 28 of a reported 32 one-wave IC(0) workgroups per CU, 25 of 26 at 6 KB of LDS,
// 7 of 8 three-wave triangular-solve workgroups, and one count that fitted in one run and not
// in another. A cooperative launch admits the reported grid, which explains a dataflow launch
// that stalled until its waits expired. So the figure is only where the search starts: the
// kernel itself is launched as a co-residency probe at that many workgroups per CU, then one
// fewer, until every workgroup of three probes in a row was resident at once (the probe's
// waits are bounded, so a failed probe drains too); the grid then keeps one workgroup per CU
// below that count. The result is cached per device and kernel. It assumes no other kernel
// occupies the device during a dataflow launch; a launch that stalls anyway ends through its
// bounded waits and is repeated on a smaller grid (Operator::shrink), counted in
// ml_solve_info.device_retries.
struct ProbeKey {
  int device;
  const void* kernel;
  bool operator<(const ProbeKey& o) const {
    return device != o.device ? device < o.device : kernel < o.kernel;
  }
};
std::mutex g_probe_mu;
std::map<ProbeKey, int> g_resident;

template <class Launch>
int resident_per_cu(const void* kernel, int threads, hipStream_t stream, Launch launch_probe) {
  int dev = 0;
  HIP_CHECK(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lock(g_probe_mu);
  auto it = g_resident.find({dev, kernel});
  if (it != g_resident.end()) return it->second;
  const int cus = device_cus();
  const long long limit = wait_limit() / 100;  // 20 ms: a resident grid gathers in microseconds
  DBuf<int> probe;
  probe.alloc(2);
  int per_cu = occupancy_per_cu(kernel, threads);
  auto resident = [&](int count) {
    mlk::k_fill_i32<<<1, 64, 0, stream>>>(probe.p, 2, 0);
    launch_probe(count * cus, probe.p, limit);
    HIP_CHECK(hipGetLastError());
    int h[2] = {0, 0};
    HIP_CHECK(hipMemcpyAsync(h, probe.p, sizeof(h), hipMemcpyDeviceToHost, stream));
    HIP_CHECK(hipStreamSynchronize(stream));
    return h[1] == 0 && h[0] == count * cus;
  };
  // at the very limit, whether every workgroup finds a place varies from one launch to the
  // next: a count must pass three probes in a row, and the grid keeps one workgroup per CU
  // below it
  for (; per_cu > 1; --per_cu)
    if (resident(per_cu) && resident(per_cu) && resident(per_cu)) break;
  per_cu = std::max(1, per_cu - 1);
  g_resident[{dev, kernel}] = per_cu;
  return per_cu;
}

// Probes of the dataflow kernels (normal launches: the probe measures what the hardware holds,
// not what the runtime admits).
template <typename... TArgs>
int trisolve_resident(void (*kernel)(TArgs...), hipStream_t stream) {
  return resident_per_cu((const void*)kernel, mlk::TRI_THREADS, stream,
                         [&](int grid, int* probe, long long limit) {
    kernel<<<grid, mlk::TRI_THREADS, 0, stream>>>(0, 0, nullptr, nullptr, nullptr, nullptr,
                                                  nullptr, nullptr, nullptr, nullptr, nullptr,
                                                  nullptr, nullptr, nullptr, nullptr, nullptr,
                                                  nullptr, nullptr, probe, limit, nullptr, -1);
  });
}
template <typename... TArgs>
int ic0_resident(void (*kernel)(TArgs...), hipStream_t stream) {
  return resident_per_cu((const void*)kernel, 64, stream,
                         [&](int grid, int* probe, long long limit) {
    kernel<<<grid, 64, 0, stream>>>(-1, nullptr, nullptr, nullptr, nullptr, nullptr, 0, probe,
                                    nullptr, limit, nullptr, nullptr, nullptr);
  });
}

template <typename... TArgs>
int fused_resident(void (*kernel)(TArgs...), hipStream_t stream) {
  return resident_per_cu((const void*)kernel, mlk::SWEEP_LANES, stream,
                         [&](int grid, int* probe, long long limit) {
    kernel<<<grid, mlk::SWEEP_LANES, 0, stream>>>(0, -1, 0, nullptr, nullptr, nullptr, nullptr,
                                                 nullptr, nullptr, nullptr, nullptr, nullptr, 0u,
                                                 nullptr, nullptr, nullptr, nullptr, probe, limit,
                                                 1.0, 0.0);
  });
}

// Launches a persistent dataflow kernel of `threads`-thread workgroups on a grid whose
// co-residency a probe of the same kernel has shown (resident_per_cu). An ordinary launch: the
// runtime's admission check of a cooperative one rests on the occupancy figure the probe
// replaces.
template <typename... Args>
void launch_dataflow(void (*kernel)(Args...), int& grid, int threads, hipStream_t stream,
                     Args... args) {
  kernel<<<grid, threads, 0, stream>>>(args...);
  HIP_CHECK(hipGetLastError());
}

}  // namespace

// Variants of k_sweeps_fused: coefficients per row kept in registers (a longer row re-reads
// the rest each sweep), the waves per SIMD its registers must allow, and the rows per lane
// (blocks of 256 rows, or 512 with two rows per lane, for rows of at most 32 strict entries).
struct FusedVariant {
  int len, waves, rpl;
};
constexpr int FUSED_VARIANTS = 8;
constexpr FusedVariant kFused[FUSED_VARIANTS] = {
    {16, 8, 1}, {24, 1, 1}, {32, 1, 1}, {48, 1, 1}, {80, 1, 1}, {16, 5, 2}, {24, 1, 2}, {32, 1, 2}};
// The code a caller names a variant by: 100 * rows per lane + coefficients held per row.
int fused_code(int v) { return 100 * kFused[v].rpl + kFused[v].len; }
// The variants worth timing for a pattern whose fused layouts exist as have[rows per lane - 1]
// and whose longest row holds `width` strict entries: their layout is built, and they hold no
// more coefficients in registers than the longest row needs (the narrowest variant of their
// shape holding every row; all of that shape when none does).
bool fused_variant_candidate(int v, const bool have[2], int width) {
  if (!have[kFused[v].rpl - 1]) return false;
  int fit = 1000;
  for (int u = 0; u < FUSED_VARIANTS; ++u)
    if (kFused[u].rpl == kFused[v].rpl && kFused[u].len >= width) fit = std::min(fit, kFused[u].len);
  return kFused[v].len <= fit;
}
// The variant a handle uses before the autotuner has timed them: one row per lane before two,
// then one compiled without a register bound (waves 1) before a bounded one, then the widest;
// -1 when no variant is a candidate (the sweeps then run one launch each).
int default_fused_variant(const bool have[2], int width) {
  int best = -1;
  auto better = [&](int v, int b) {
    if (kFused[v].rpl != kFused[b].rpl) return kFused[v].rpl < kFused[b].rpl;
    if ((kFused[v].waves == 1) != (kFused[b].waves == 1)) return kFused[v].waves == 1;
    return kFused[v].len > kFused[b].len;
  };
  for (int v = 0; v < FUSED_VARIANTS; ++v)
    if (fused_variant_candidate(v, have, width) && (best < 0 || better(v, best))) best = v;
  return best;
}

// The host layouts of the Jacobi sweeps of one lower pattern, built and checked together: the
// per-sweep layouts and the fused ones in blocks of 256 and of 512 rows (two rows per lane,
// kept only for rows of at most 32 strict entries), the three on threads of their own.
struct SweepHostLayouts {
  std::string err;  // the first failed check, or empty
  SweepLayout lo, up;
  FusedSweepLayout fused[2][2];  // [rows per lane - 1][0 lower, 1 upper]
  bool have[2] = {false, false};
};

static SweepHostLayouts build_sweep_layouts(const std::vector<int64_t>& offsets,
                                            const std::vector<int32_t>& cols) {
  LowerPattern lp;
  lp.offsets = offsets;
  lp.cols = cols;
  SweepHostLayouts s;
  std::string ferr[2];
  auto fused = [&](int h) {
    FusedSweepLayout& flo = s.fused[h][0];
    FusedSweepLayout& fup = s.fused[h][1];
    // a pattern that does not fit the layout keeps the other one, or one launch per sweep
    if (!fused_sweep_layouts(lp, FUSED_LANES * (h + 1), &flo, &fup).empty()) return;
    if (h == 1 && std::max(flo.width, fup.width) > 32) return;
    ferr[h] = check_fused_sweep_layouts(lp, flo, fup);
    s.have[h] = ferr[h].empty();
  };
  ThreadGroup fused_builds;
  fused_builds.run([&] { fused(0); });
  fused_builds.run([&] { fused(1); });
  s.err = sweep_layouts(lp, &s.lo, &s.up);
  if (s.err.empty()) s.err = check_sweep_layouts(lp, s.lo, s.up);
  fused_builds.join();
  for (int h = 0; h < 2 && s.err.empty(); ++h) s.err = ferr[h];
  return s;
}

// ====================================================================== Operator
struct Operator {
  int n = 0;
  int64_t nnz = 0;
  int precision = ML_FP64;
  hipStream_t stream = nullptr;
  Profiler prof;
  int64_t depth = 0;
  int max_row = 0;
  int shift_index = -1;
  bool factored = false;

  // reference lower CSR
  std::vector<int64_t> h_off;
  std::vector<int32_t> h_col;
  // the sweep layouts under construction from h_off and h_col (a handle whose precision runs
  // sweeps from the start begins them while its other layouts are built); declared after
  // them, so that a handle destroyed early waits for the build before they go
  std::future<SweepHostLayouts> sweep_build;
  DBuf<int> d_off, d_col;
  DBuf<int> d_icorder;   // IC(0) processing order (rows by dependency level)
  DBuf<double> d_a;      // values
  DBuf<float> d_a32;
  // sliced ELL of the full symmetric matrix
  int slices = 0;
  DBuf<int> d_soff, d_scol, d_slen, d_ssrc;
  DBuf<int16_t> d_sdel;  // 16-bit column offsets of the sliced ELL, when the band allows
  bool spmv_d16 = false;
  // the grouped column streams (one per group of rows with equal column sets), when they fit
  DBuf<int> d_goff, d_gnum;
  DBuf<int16_t> d_gdel;
  DBuf<uint32_t> d_ginfo;
  bool have_grouped = false, spmv_grouped = false;  // built; in use (autotuned)
  int64_t grouped_bytes = 0;  // the grouped streams' bytes (columns + per-row fields)
  DBuf<double> d_sval;
  DBuf<float> d_sval32;
  // the IC(0) factor and the layouts of its two triangular solves
  DBuf<double> d_l;
  DBuf<float> d_l32;
  struct TriDev {
    int nblk = 0;
    DBuf<int> border, bstart, brow, xoff, xsrc, ioff, isrc, dsrc, uoff, ucol;
    DBuf<unsigned char> ik, xu;
    DBuf<double> xval, ival, diag;
    DBuf<float> xval32, ival32, diag32;
    DBuf<double> rdiag32, rdiag64;  // 1 / (double)diag, what the solves multiply by
    void set(const TriLayout& t, hipStream_t st) {
      nblk = (int)t.bstart.size() - 1;
      border.upload_vec(t.border, st);
      bstart.upload_vec(t.bstart, st);
      brow.upload_vec(t.brow, st);
      xoff.upload_vec(t.xoff, st);
      xsrc.upload_vec(t.xsrc, st);
      uoff.upload_vec(t.uoff, st);
      ucol.upload_vec(t.ucol, st);
      xu.upload_vec(t.xu, st);
      ioff.upload_vec(t.ioff, st);
      isrc.upload_vec(t.isrc, st);
      dsrc.upload_vec(t.dsrc, st);
      std::vector<unsigned char> k8(t.ik.begin(), t.ik.end());
      ik.upload_vec(k8, st);
      HIP_CHECK(hipStreamSynchronize(st));
      n_ext = t.xcol.size();
      n_int = t.ik.size();
      n_slots = t.dsrc.size();
    }
    // the factor's values in the layout, in the precision the handle applies it
    size_t n_ext = 0, n_int = 0, n_slots = 0;
    void alloc64() {
      xval.alloc(n_ext);
      ival.alloc(n_int);
      diag.alloc(n_slots);
      rdiag64.alloc(n_slots);
    }
    void alloc32() {
      xval32.alloc(n_ext);
      ival32.alloc(n_int);
      diag32.alloc(n_slots);
      rdiag32.alloc(n_slots);
    }
  } fw, bw;
  int64_t nf = 0;  // strictly lower entries
  // workspace
  DBuf<double> w_x, w_r, w_p, w_ap, w_y, w_z, w_b, w_d, w_ax, w_partial;
  DBuf<float> w_p32, w_ap32;  // p and Ap of the FP32 operator (ML_MIXED)
  DBuf<mlk::PcgScalars> d_S, d_Sin;
  DBuf<int> d_flags;  // [0] abort, [1] IC(0) failure, [2..5] where an IC(0) wait gave up
  // Jacobi sweeps of the FP32 factor (ML_MIXED_JACOBI, or a mixed handle given sweeps > 0):
  // sliced ELL of the strict lower triangle and of the strict upper one (rows of L^T), their
  // FP32 coefficients gathered after each factorisation, and RN64(1 / l_ii)
  struct SweepDev {
    DBuf<int> soff, scol, slen, ssrc;
    DBuf<uint16_t> sdel;  // 16-bit column distances, when they fit (then scol is not kept)
    DBuf<float> sval;
    bool d16 = false;
    int64_t entries = 0;
    void set(const SweepLayout& l, hipStream_t st) {
      entries = (int64_t)l.scol.size();
      d16 = !l.sdelta.empty();
      soff.upload_vec(l.soff, st);
      if (d16) sdel.upload_vec(l.sdelta, st);
      else scol.upload_vec(l.scol, st);
      slen.upload_vec(l.slen, st);
      ssrc.upload_vec(l.ssrc, st);
      sval.alloc(l.scol.size());
    }
  } sw_lo, sw_up;
  DBuf<int> d_sw_dsrc;
  DBuf<float> d_sw_diag;
  DBuf<double> d_sw_rd, w_s1, w_s2, w_s3, w_y0;  // w_s3: z0 = y rd, w_y0: r rd
  // the same sweeps laid out for k_sweeps_fused (all sweeps of a system in one launch), when
  // the pattern fits its capacities, and the levels each system publishes between sweeps
  struct FusedDev {
    int nblk = 0;
    int64_t entries = 0;
    DBuf<int> eoff, esrc, xoff, xcol;
    DBuf<uint8_t> elen;
    DBuf<uint16_t> eidx;
    DBuf<float> eval;
    void set(const FusedSweepLayout& f, hipStream_t st) {
      nblk = (int)f.blocks;
      entries = (int64_t)f.esrc.size();
      eoff.upload_vec(f.eoff, st);
      esrc.upload_vec(f.esrc, st);
      xoff.upload_vec(f.xoff, st);
      xcol.upload_vec(f.xcol, st);
      elen.upload_vec(f.elen, st);
      eidx.upload_vec(f.eidx, st);
      eval.alloc(f.esrc.size());
    }
  } fz[2][2];  // [rows per lane - 1][0 lower, 1 upper]: blocks of 256 and of 512 rows
  bool have_fz[2] = {false, false};  // the pattern fits blocks of 256 / 512 rows
  bool have_fused = false;           // some fused layout is built
  bool fused_on = true;              // the caller's choice (ml_pcg_set_sweeps_fused)
  int fused_var = 1;        // the variant in use
  int fused_width = 0;      // the longest row of the fused layouts
  int fused_per_cu = 0, full_fused = 0, fused_grid = 0;
  DBuf<double> w_lev_lo, w_lev_up;  // sweeps - 1 levels of n values per system
  int lev_sweeps = 0;               // the sweep count the level buffers are sized for
  int lev_tag = 1;                  // the next fused launch's level tag
  bool have_sweep = false;
  int sweeps = 0;  // Jacobi sweeps per triangular system; 0: the exact dataflow solves
  double sweep_w = 1.0;  // the sweeps' weight (1: undamped)
  int sweep_batch = 8, spmv_batch = 8;  // entries per batch of loads (autotuned: 8 or 16)
  bool sweeps_pinned = false;  // set by the caller: the autotuner keeps the count
  bool ref_dots = false;       // PCG dot products in the reference's order (k_dot8)
  // grids of the dataflow kernels: the full ones, and the ones in use (full >> shrinks)
  int full_tri64 = 0, full_tri32 = 0, full_ic = 0, full_ic64 = 0;
  int cap_tri64 = 0, cap_tri32 = 0;  // dataflow grids before the cap by block count
  int tri_per_cu = 1, ic_per_cu = 1, ic64_per_cu = 1;  // probed co-resident workgroups per CU
  // reported and probed workgroups per CU: forward solve, backward solve, IC(0) (512-entry
  // rows), fused sweeps (the tighter of the two systems; zeros without fused layouts)
  int residency[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  int rhs_per_pass = 256;            // right-hand sides per batched pass (autotuned)
  int tri_grid64 = 0, tri_grid32 = 0, ic_grid = 0, ic_grid64 = 0;
  int shrinks = 0;   // grid halvings after a stalled dataflow launch
  int retries = 0;   // work repeated on a smaller grid (reported in ml_solve_info)
  long long wait_ticks = 0;

  ~Operator() {
    if (stream) (void)hipStreamDestroy(stream);
  }

  void init_stream() {
    if (!stream) HIP_CHECK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    if (!wait_ticks) wait_ticks = wait_limit();
  }

  // Upload the pattern and every layout derived from it. The layouts are checked on the host
  // against every bound the kernels rely on before any of them reaches the device.
  void set_pattern(const LowerPattern& lp) {
    init_stream();
    h_off = lp.offsets;
    h_col = lp.cols;
    if (precision == ML_MIXED_JACOBI)
      sweep_build = std::async(std::launch::async, build_sweep_layouts, std::cref(h_off),
                               std::cref(h_col));
    DeviceLayouts d;
    std::string err = device_layouts(lp, kTriCaps, &d);
    if (err.empty()) err = check_layouts(lp, d, kTriCaps);
    if (!err.empty()) throw ArgError(err);
    n = (int)lp.n();
    nnz = lp.nnz();
    depth = d.depth;
    max_row = d.max_row;
    slices = (int)d.slices;
    d_off.upload_vec(d.off32, stream);
    d_col.upload_vec(d.col32, stream);
    d_icorder.upload_vec(d.ic_order, stream);
    d_soff.upload_vec(d.soff, stream);
    d_scol.upload_vec(d.scol, stream);
    spmv_d16 = !d.sdel.empty();
    if (spmv_d16) d_sdel.upload_vec(d.sdel, stream);
    else d_sdel.release();
    have_grouped = !d.ginfo.empty();
    spmv_grouped = have_grouped;
    if (have_grouped) {
      d_goff.upload_vec(d.goff, stream);
      d_gnum.upload_vec(d.gnum, stream);
      d_gdel.upload_vec(d.gdel, stream);
      d_ginfo.upload_vec(d.ginfo, stream);
      grouped_bytes = (int64_t)d.gdel.size() * 2 + (int64_t)d.ginfo.size() * 4 +
                      (int64_t)(d.goff.size() + d.gnum.size()) * 4;
    } else {
      d_goff.release();
      d_gnum.release();
      d_gdel.release();
      d_ginfo.release();
      grouped_bytes = 0;
    }
    d_slen.upload_vec(d.slen, stream);
    d_ssrc.upload_vec(d.ssrc, stream);
    nf = nnz - n;
    fw.set(d.fw, stream);
    bw.set(d.bw, stream);
    if (precision == ML_FP64) {
      fw.alloc64();
      bw.alloc64();
    }
    // values and workspace
    d_a.alloc(nnz);
    d_l.alloc(nnz);
    d_sval.alloc(d_scol.n);
    for (auto* b : {&w_x, &w_r, &w_p, &w_ap, &w_y, &w_z, &w_b, &w_d, &w_ax}) b->alloc(n);
    const int nblk = std::max(fw.nblk, bw.nblk);
    // every dataflow grid is one whose co-residency was shown by a probe of the same kernel;
    // the forward and backward solves share one grid, so it fits the tighter of the two
    const int cus = device_cus();
    auto note = [&](int slot, const void* kernel, int threads, int probed) {
      residency[2 * slot] = occupancy_per_cu(kernel, threads);
      residency[2 * slot + 1] = probed;
      return probed;
    };
    if (precision == ML_FP64) {
      auto fw_k = mlk::k_trisolve<double, false, false>;
      auto bw_k = mlk::k_trisolve<double, true, true>;
      tri_per_cu = std::min(note(0, (const void*)fw_k, mlk::TRI_THREADS, trisolve_resident(fw_k, stream)),
                            note(1, (const void*)bw_k, mlk::TRI_THREADS, trisolve_resident(bw_k, stream)));
      ic_per_cu = note(2, (const void*)mlk::k_ic0<double, 512>, 64,
                       ic0_resident(mlk::k_ic0<double, 512>, stream));
      ic64_per_cu = ic0_resident(mlk::k_ic0<double, 64>, stream);
    } else {
      auto fw_k = mlk::k_trisolve<double, false, false, false, float>;
      auto bw_k = mlk::k_trisolve<double, true, true, false, float>;
      tri_per_cu = std::min(note(0, (const void*)fw_k, mlk::TRI_THREADS, trisolve_resident(fw_k, stream)),
                            note(1, (const void*)bw_k, mlk::TRI_THREADS, trisolve_resident(bw_k, stream)));
      ic_per_cu = note(2, (const void*)mlk::k_ic0<float, 512>, 64,
                       ic0_resident(mlk::k_ic0<float, 512>, stream));
      ic64_per_cu = ic0_resident(mlk::k_ic0<float, 64>, stream);
    }
    cap_tri64 = cap_tri32 = tri_per_cu * cus;
    full_tri64 = std::min<int>(nblk, cap_tri64);
    full_tri32 = std::min<int>(nblk, cap_tri32);
    full_ic = std::min<int>(n, ic_per_cu * cus);
    full_ic64 = std::min<int>(n, ic64_per_cu * cus);
    shrinks = 0;
    set_grids();
    w_partial.alloc(std::max<int64_t>(grid_for(n), 1));
    d_S.alloc(1);
    d_Sin.alloc(1);
    d_flags.alloc(6);
    mlk::k_fill_i32<<<1, 64, 0, stream>>>(d_flags.p, 2, 0);
    mlk::k_fill_i32<<<1, 64, 0, stream>>>(d_flags.p + 2, 4, -1);
    HIP_CHECK(hipGetLastError());
    if (precision != ML_FP64) alloc_mixed();
    have_sweep = false;
    sweeps = 0;
    if (precision == ML_MIXED_JACOBI) set_sweeps(4);
    factored = false;
    // the host layouts above are temporaries: wait for their uploads before they go away
    HIP_CHECK(hipStreamSynchronize(stream));
  }

  // Jacobi sweeps per triangular system for the FP32 factor (0: exact dataflow solves); the
  // sweep layouts are built and checked on the host the first time sweeps are asked for.
  void set_sweeps(int k) {
    if (k < 0 || k > 64) throw ArgError("sweeps must be in 0..64");
    if (k > 0 && precision == ML_FP64)
      throw ArgError("the FP64 drop-in applies its factor exactly (sweeps need a mixed handle)");
    if (k > 0 && !have_sweep) {
      SweepHostLayouts L = sweep_build.valid() ? sweep_build.get()
                                               : build_sweep_layouts(h_off, h_col);
      if (!L.err.empty()) throw ArgError(L.err);
      sw_lo.set(L.lo, stream);
      sw_up.set(L.up, stream);
      d_sw_dsrc.upload_vec(L.lo.dsrc, stream);
      d_sw_diag.alloc(n);
      d_sw_rd.alloc(n);
      for (auto* b : {&w_s1, &w_s2, &w_s3, &w_y0}) b->alloc(n);
      // the fused layouts, where the pattern fits them (else every sweep is its own launch):
      // blocks of 256 rows, and of 512 (two rows per lane) for rows short enough to hold two
      for (int h = 0; h < 2; ++h) {
        if (!L.have[h]) continue;
        fz[h][0].set(L.fused[h][0], stream);
        fz[h][1].set(L.fused[h][1], stream);
        fused_width = std::max(L.fused[h][0].width, L.fused[h][1].width);
        have_fz[h] = true;
      }
      // a pattern without a candidate variant keeps one launch per sweep
      have_fused = default_fused() >= 0;
      if (have_fused) select_fused(default_fused());
      HIP_CHECK(hipStreamSynchronize(stream));  // the host layouts go out of scope
      have_sweep = true;
      if (factored) gather_sweeps();
    }
    sweeps = k;
    if (have_fused && k > 1 && lev_sweeps != k) {
      for (auto* b : {&w_lev_lo, &w_lev_up}) b->alloc((size_t)(k - 1) * n);
      lev_sweeps = k;
      fill_levels();
    }
  }

  typedef void (*FusedKernel)(int, int, int, const int*, const uint8_t*, const float*,
                              const uint16_t*, const int*, const int*, const double*,
                              const double*, double*, unsigned, double*, const double*, double*,
                              const mlk::PcgScalars*, int*, long long, double, double);
  // k_sweeps_fused in one of its variants (default: the one in use)
  FusedKernel fused_kernel(bool upper, int var = -1) const {
    switch (var < 0 ? fused_var : var) {
      case 0: return upper ? mlk::k_sweeps_fused<16, true, 8> : mlk::k_sweeps_fused<16, false, 8>;
      case 1: return upper ? mlk::k_sweeps_fused<24, true> : mlk::k_sweeps_fused<24, false>;
      case 2: return upper ? mlk::k_sweeps_fused<32, true> : mlk::k_sweeps_fused<32, false>;
      case 3: return upper ? mlk::k_sweeps_fused<48, true> : mlk::k_sweeps_fused<48, false>;
      case 4: return upper ? mlk::k_sweeps_fused<80, true> : mlk::k_sweeps_fused<80, false>;
      case 5: return upper ? mlk::k_sweeps_fused<16, true, 5, 2> : mlk::k_sweeps_fused<16, false, 5, 2>;
      case 6: return upper ? mlk::k_sweeps_fused<24, true, 1, 2> : mlk::k_sweeps_fused<24, false, 1, 2>;
      default: return upper ? mlk::k_sweeps_fused<32, true, 1, 2> : mlk::k_sweeps_fused<32, false, 1, 2>;
    }
  }
  bool fused_candidate(int v) const { return fused_variant_candidate(v, have_fz, fused_width); }
  int default_fused() const { return default_fused_variant(have_fz, fused_width); }
  const FusedDev& fused_layout(bool upper) const { return fz[kFused[fused_var].rpl - 1][upper]; }
  // puts a variant in use: its grid is bounded by probes of both of its kernels
  void select_fused(int var) {
    fused_var = var;
    fused_per_cu = std::min(fused_resident(fused_kernel(false), stream),
                            fused_resident(fused_kernel(true), stream));
    residency[6] = occupancy_per_cu((const void*)fused_kernel(false), mlk::SWEEP_LANES);
    residency[7] = fused_per_cu;
    full_fused = std::min<int>(fused_layout(false).nblk, fused_per_cu * device_cus());
    set_grids();
  }
  bool fused_active() const { return have_fused && fused_on && sweeps > 0; }
  // the fused kernel's variant in use as 100 * rows per lane + coefficients per row held in
  // registers; 0 when the sweeps are launched one by one
  int fused_capacity() const { return fused_active() ? fused_code(fused_var) : 0; }
  // 0: one launch per sweep; 1: the fused kernel in its chosen variant; a variant code
  void set_fused(int fused) {
    if (fused == 0 || fused == 1) {
      fused_on = fused == 1;
      return;
    }
    for (int v = 0; v < FUSED_VARIANTS; ++v)
      if (fused_code(v) == fused) {
        if (!have_fz[kFused[v].rpl - 1])
          throw ArgError("the pattern does not fit that fused sweep variant");
        fused_on = true;
        select_fused(v);
        fill_levels();
        return;
      }
    throw ArgError("fused must be 0, 1 or a variant code 116, 124, 132, 148, 180, 216, 224, 232");
  }

  // every level slot of both systems cleared to tag 0, which no launch uses; the tags start
  // again from 1
  void fill_levels() {
    lev_tag = 1;
    if (lev_sweeps < 2) return;
    const int64_t m = (int64_t)(lev_sweeps - 1) * n;
    for (auto* b : {&w_lev_lo, &w_lev_up})
      mlk::k_fill_bits<double><<<grid_for(m), 256, 0, stream>>>(b->p, m, 0ull);
    HIP_CHECK(hipGetLastError());
  }
  // the tag of the next fused launch: each launch since the levels were last cleared has a tag
  // of its own (1..255), so a slot holds the current launch's tag only once that launch has
  // written it; after 255 launches the levels are cleared (stream-ordered, a fill kernel)
  unsigned next_level_tag() {
    if (lev_tag > (int)mlk::LEVEL_TAG_MASK) fill_levels();
    return (unsigned)lev_tag++;
  }

  // the FP32 factor's values into the sweep layouts, and the reciprocals of its diagonal
  void gather_sweeps() {
    for (SweepDev* t : {&sw_lo, &sw_up})
      mlk::k_gather<float, float><<<grid_for(t->entries), 256, 0, stream>>>(
          t->entries, t->ssrc.p, d_l32.p, t->sval.p);
    for (int h = 0; h < 2; ++h)
      for (FusedDev* t : {&fz[h][0], &fz[h][1]})
        if (have_fz[h] && t->entries)
          mlk::k_gather<float, float><<<grid_for(t->entries), 256, 0, stream>>>(
              t->entries, t->esrc.p, d_l32.p, t->eval.p);
    mlk::k_gather<float, float><<<grid_for(n), 256, 0, stream>>>((int64_t)n, d_sw_dsrc.p, d_l32.p,
                                                                 d_sw_diag.p);
    mlk::k_recip64<float><<<grid_for(n), 256, 0, stream>>>((int64_t)n, d_sw_diag.p, d_sw_rd.p);
    HIP_CHECK(hipGetLastError());
  }

  void alloc_mixed() {
    d_a32.alloc(nnz);
    d_l32.alloc(nnz);
    d_sval32.alloc(d_scol.n);
    fw.alloc32();
    bw.alloc32();
    if (precision == ML_MIXED)
      for (auto* b : {&w_p32, &w_ap32}) b->alloc(n);
  }

  // After d_a holds the lower values: refresh the SpMV copies.
  void values_changed() {
    mlk::k_gather<double, double><<<grid_for(d_scol.n), 256, 0, stream>>>(
        (int64_t)d_scol.n, d_ssrc.p, d_a.p, d_sval.p);
    HIP_CHECK(hipGetLastError());
    if (precision != ML_FP64) {
      mlk::k_convert<double, float><<<grid_for(nnz), 256, 0, stream>>>(nnz, d_a.p, d_a32.p);
      mlk::k_gather<double, float><<<grid_for(d_scol.n), 256, 0, stream>>>(
          (int64_t)d_scol.n, d_ssrc.p, d_a.p, d_sval32.p);
      HIP_CHECK(hipGetLastError());
    }
    factored = false;
  }

  int64_t read_flag(int k) {
    int v = 0;
    HIP_CHECK(hipMemcpyAsync(&v, d_flags.p + k, sizeof(int), hipMemcpyDeviceToHost, stream));
    HIP_CHECK(hipStreamSynchronize(stream));
    return v;
  }

  // One IC(0) factorisation with the given shift; with `seg` (batched systems: the system of
  // each row), each system uses seg_shift[system] and reports its failure in seg_fail[system].
  template <typename T>
  int run_ic0(const T* a, T* l, T shift, const int* seg = nullptr, const T* seg_shift = nullptr,
              int* seg_fail = nullptr, int nseg = 0) {
    mlk::k_fill_i32<<<1, 64, 0, stream>>>(d_flags.p, 2, 0);
    mlk::k_fill_i32<<<1, 64, 0, stream>>>(d_flags.p + 2, 4, -1);
    if (seg) mlk::k_fill_i32<<<grid_for(nseg), 256, 0, stream>>>(seg_fail, nseg, 0);
    mlk::k_fill_bits<T><<<grid_for(nnz), 256, 0, stream>>>(l, nnz, mlk::Bits<T>::S);
    hipEvent_t ev = prof.begin(stream);
    const int* off = d_off.p;
    const int* col = d_col.p;
    int* fail = seg ? seg_fail : d_flags.p + 1;
    int* abort_flag = d_flags.p;
    int* where = d_flags.p + 2;
    long long limit = wait_ticks;
    int nn = n;
    const int* order = d_icorder.p;
    if (max_row <= 64)
      launch_dataflow(mlk::k_ic0<T, 64>, ic_grid64, 64, stream, nn, order, off, col, a, l, shift,
                      fail, abort_flag, limit, where, seg, seg_shift);
    else
      launch_dataflow(mlk::k_ic0<T, 512>, ic_grid, 64, stream, nn, order, off, col, a, l, shift,
                      fail, abort_flag, limit, where, seg, seg_shift);
    prof.end(P_IC0, ev, stream);
    int flags[6];
    HIP_CHECK(hipMemcpyAsync(flags, d_flags.p, sizeof(flags), hipMemcpyDeviceToHost, stream));
    HIP_CHECK(hipStreamSynchronize(stream));
    if (flags[0]) {
      g_last_error = "IC(0) wait expired: row " + std::to_string(flags[2]) + " waiting on row " +
                     std::to_string(flags[3]) + ", wave " + std::to_string(flags[4]) + " of " +
                     std::to_string(flags[5]) + ", limit " + std::to_string(wait_ticks) + " ticks";
      return ML_DEVICE_TIMEOUT;
    }
    if (flags[1] == 2) throw ArgError("matrix row longer than 512 entries");
    return flags[1] ? ML_SINGULAR : ML_OK;
  }

  // Pcg::new (iterative.rs:201-207): the first shift of the ladder whose IC(0) succeeds.
  void set_grids() {
    tri_grid64 = std::max(1, full_tri64 >> shrinks);
    tri_grid32 = std::max(1, full_tri32 >> shrinks);
    ic_grid = std::max(1, full_ic >> shrinks);
    ic_grid64 = std::max(1, full_ic64 >> shrinks);
    fused_grid = std::max(1, full_fused >> shrinks);
  }

  // After a dataflow launch stalled (its bounded waits expired): halve the dataflow grids so
  // that the work can be repeated; false once they have been halved three times.
  bool shrink() {
    if (shrinks >= 3) return false;
    ++shrinks;
    ++retries;
    set_grids();
    return true;
  }

  int factor() {
    static const double SHIFTS[6] = {0.0, 1e-3, 1e-2, 1e-1, 1.0, 10.0};
    factored = false;
    shift_index = -1;
    int st = ML_SINGULAR;
    if (precision == ML_FP64) {
      for (int s = 0; s < 6; ++s) {
        st = run_ic0<double>(d_a.p, d_l.p, SHIFTS[s]);
        if (st == ML_DEVICE_TIMEOUT && shrink()) { --s; continue; }
        if (st == ML_DEVICE_TIMEOUT) return st;
        if (st == ML_OK) { shift_index = s; break; }
      }
      if (st != ML_OK) return ML_SINGULAR;
      gather_factor<double>(d_l.p);
    } else {
      // the mixed modes only apply the FP32 factor: the ladder runs in FP32 arithmetic
      for (int s = 0; s < 6; ++s) {
        st = run_ic0<float>(d_a32.p, d_l32.p, (float)SHIFTS[s]);
        if (st == ML_DEVICE_TIMEOUT && shrink()) { --s; continue; }
        if (st == ML_DEVICE_TIMEOUT) return st;
        if (st == ML_OK) { shift_index = s; break; }
      }
      if (st != ML_OK) return ML_SINGULAR;
      gather_factor<float>(d_l32.p);
      if (have_sweep) gather_sweeps();
    }
    factored = true;
    return ML_OK;
  }

  template <typename T>
  void gather_factor(const T* l) {
    const bool f32 = std::is_same<T, float>::value;
    for (TriDev* t : {&fw, &bw}) {
      T* xv = f32 ? (T*)t->xval32.p : (T*)t->xval.p;
      T* iv = f32 ? (T*)t->ival32.p : (T*)t->ival.p;
      T* dv = f32 ? (T*)t->diag32.p : (T*)t->diag.p;
      // a layout may have no external or no internal entries at all (a single block)
      if (t->n_ext)
        mlk::k_gather<T, T><<<grid_for(t->n_ext), 256, 0, stream>>>((int64_t)t->n_ext, t->xsrc.p, l, xv);
      if (t->n_int)
        mlk::k_gather<T, T><<<grid_for(t->n_int), 256, 0, stream>>>((int64_t)t->n_int, t->isrc.p, l, iv);
      mlk::k_gather<T, T><<<grid_for(n), 256, 0, stream>>>((int64_t)n, t->dsrc.p, l, dv);
      if (f32)
        mlk::k_recip64<float><<<grid_for(n), 256, 0, stream>>>((int64_t)n, t->diag32.p,
                                                               t->rdiag32.p);
      else
        mlk::k_recip64<double><<<grid_for(n), 256, 0, stream>>>((int64_t)n, t->diag.p,
                                                                t->rdiag64.p);
    }
    HIP_CHECK(hipGetLastError());
  }

  // ---------------------------------------------------------------- kernels
  // z = (L L^T)^-1 r through y by the two dataflow triangular solves, with FP64 vectors and
  // sums: the FP64 factor (drop-in), or the mixed handles' FP32 factor, each coefficient
  // widened as it is used. With nrep > 1, r, y and z hold nrep vectors one after the other,
  // solved in one launch.
  void precondition(const double* r, double* y, double* z, const mlk::PcgScalars* S,
                    long long* trace_fw = nullptr, long long* trace_bw = nullptr, int nrep = 1) {
    const bool c32 = precision != ML_FP64;
    int& grid1 = c32 ? tri_grid32 : tri_grid64;
    int grid_rep = std::min<int64_t>((int64_t)std::max(fw.nblk, bw.nblk) * nrep,
                                     std::max(1, (c32 ? cap_tri32 : cap_tri64) >> shrinks));
    // the traced kernels have their own footprint, so their grid is bounded by their own probe
    int grid_trace = grid1;
    if (trace_fw || trace_bw) {
      const int per_cu =
          c32 ? std::min(trisolve_resident(mlk::k_trisolve<double, false, false, true, float>, stream),
                         trisolve_resident(mlk::k_trisolve<double, true, true, true, float>, stream))
              : std::min(trisolve_resident(mlk::k_trisolve<double, false, false, true>, stream),
                         trisolve_resident(mlk::k_trisolve<double, true, true, true>, stream));
      grid_trace = std::min(grid1, per_cu * device_cus());
    }
    int& grid = nrep > 1 ? grid_rep : ((trace_fw || trace_bw) ? grid_trace : grid1);
    hipEvent_t ev = prof.begin(stream);
    int nn = n;
    long long limit = wait_ticks;
    auto run = [&](TriDev& t, const double* rhs, double* out, bool fwd, long long* trace) {
      int nb = t.nblk;
      auto go = [&](auto kernel, auto* xv, auto* iv, auto* dg, const double* rd) {
        launch_dataflow(kernel, grid, mlk::TRI_THREADS, stream, nn, nb, (const int*)t.border.p,
                        (const int*)t.bstart.p, (const int*)t.brow.p, (const int*)t.xoff.p,
                        (const int*)t.uoff.p, (const int*)t.ucol.p, (const unsigned char*)t.xu.p,
                        xv, (const int*)t.ioff.p, (const unsigned char*)t.ik.p, iv, dg, rd, rhs,
                        out, S, d_flags.p, limit, trace, nrep);
      };
      if (c32) {
        const float* xv = t.xval32.p;
        const float* iv = t.ival32.p;
        const float* dg = t.diag32.p;
        const double* rd = t.rdiag32.p;
        if (fwd)
          trace ? go(mlk::k_trisolve<double, false, false, true, float>, xv, iv, dg, rd)
                : go(mlk::k_trisolve<double, false, false, false, float>, xv, iv, dg, rd);
        else
          trace ? go(mlk::k_trisolve<double, true, true, true, float>, xv, iv, dg, rd)
                : go(mlk::k_trisolve<double, true, true, false, float>, xv, iv, dg, rd);
      } else {
        const double* xv = t.xval.p;
        const double* iv = t.ival.p;
        const double* dg = t.diag.p;
        const double* rd = t.rdiag64.p;
        if (fwd)
          trace ? go(mlk::k_trisolve<double, false, false, true>, xv, iv, dg, rd)
                : go(mlk::k_trisolve<double, false, false>, xv, iv, dg, rd);
        else
          trace ? go(mlk::k_trisolve<double, true, true, true>, xv, iv, dg, rd)
                : go(mlk::k_trisolve<double, true, true>, xv, iv, dg, rd);
      }
    };
    run(fw, r, y, true, trace_fw);
    run(bw, (const double*)y, z, false, trace_bw);
    prof.end(c32 ? P_TRI32 : P_TRI64, ev, stream);
  }

  // z ~ (L L^T)^-1 r by `sweeps` Jacobi sweeps of L y = r from y = D^-1 r, then as many of
  // L^T z = y from z = D^-1 y (the transposed iteration, so the preconditioner stays
  // symmetric), on the FP32 factor with FP64 vectors. With `dot`, the last sweep also leaves
  // per-block partial sums of dot.z in `partial` (the PCG's r.z). With y0 (= r * rd, written
  // by the PCG's residual update) the first forward sweep reads it instead of forming it; the
  // last forward sweep writes z0 = y * rd for the first backward one likewise.
  void apply_sweeps(const double* r, double* z, const mlk::PcgScalars* S,
                    const double* dot = nullptr, double* partial = nullptr,
                    const double* y0 = nullptr) {
    hipEvent_t ev = prof.begin(stream);
    const unsigned g = grid_for(n);
    const int k = sweeps;
    double* a = w_s1.p;
    double* b = w_s2.p;
    double* z0 = w_s3.p;
    if (fused_active()) {
      // one launch per system, each with a level tag of its own
      double* lo = k > 1 ? w_lev_lo.p : nullptr;
      double* up = k > 1 ? w_lev_up.p : nullptr;
      const FusedDev& fl = fused_layout(false);
      const FusedDev& fu = fused_layout(true);
      const unsigned tag_lo = k > 1 ? next_level_tag() : 0u;
      fused_kernel(false)<<<fused_grid, mlk::SWEEP_LANES, 0, stream>>>(
          n, fl.nblk, k, fl.eoff.p, fl.elen.p, fl.eval.p, fl.eidx.p, fl.xoff.p, fl.xcol.p,
          d_sw_rd.p, r, lo, tag_lo, a, nullptr, nullptr, S, d_flags.p, wait_ticks, sweep_w,
          1.0 - sweep_w);
      const unsigned tag_up = k > 1 ? next_level_tag() : 0u;
      fused_kernel(true)<<<fused_grid, mlk::SWEEP_LANES, 0, stream>>>(
          n, fu.nblk, k, fu.eoff.p, fu.elen.p, fu.eval.p, fu.eidx.p, fu.xoff.p, fu.xcol.p,
          d_sw_rd.p, a, up, tag_up, z, dot, partial, S, d_flags.p, wait_ticks, sweep_w,
          1.0 - sweep_w);
      HIP_CHECK(hipGetLastError());
      prof.end(P_TRI32, ev, stream);
      return;
    }
    // `level`: an intermediate iterate, rounded as the fused sweeps round the ones they exchange
    auto sweep = [&](const SweepDev& t, const double* rhs, const double* in, double* out,
                     double* out_rd, bool last, bool level) {
      const double* dv = last ? dot : nullptr;
      double* pv = last ? partial : nullptr;
      const int sign = &t == &sw_lo ? -1 : 1;
      const bool first = in == nullptr;
      auto go = [&](auto kernel) {
        kernel<<<g, 256, 0, stream>>>(n, t.soff.p, t.scol.p, t.sdel.p, sign, t.slen.p, t.sval.p,
                                      d_sw_rd.p, rhs, in, out, level ? 1 : 0, out_rd, dv, pv,
                                      S, sweep_w, 1.0 - sweep_w);
      };
      if (sweep_batch == 16) {
        if (t.d16) first ? go(mlk::k_sweep<true, true, 16>) : go(mlk::k_sweep<false, true, 16>);
        else first ? go(mlk::k_sweep<true, false, 16>) : go(mlk::k_sweep<false, false, 16>);
      } else {
        if (t.d16) first ? go(mlk::k_sweep<true, true>) : go(mlk::k_sweep<false, true>);
        else first ? go(mlk::k_sweep<true, false>) : go(mlk::k_sweep<false, false>);
      }
    };
    // forward, ping-pong between a and b; y ends in a, and z0 = y rd in z0
    sweep(sw_lo, r, y0, a, k == 1 ? z0 : nullptr, false, k > 1);
    for (int m = 2; m <= k; ++m) {
      sweep(sw_lo, r, a, b, m == k ? z0 : nullptr, false, m < k);
      std::swap(a, b);
    }
    // backward from y = a and z0; the sweeps alternate between b and z so that the last lands
    // in z
    const double* y = a;
    const double* prev = z0;
    for (int m = 1; m <= k; ++m) {
      double* out = ((k - m) & 1) ? b : z;
      sweep(sw_up, y, prev, out, nullptr, m == k, m < k);
      prev = out;
    }
    HIP_CHECK(hipGetLastError());
    prof.end(P_TRI32, ev, stream);
  }

  // ---------------------------------------------------------------- interleaved vectors
  // Scratch for nrep vectors interleaved (entry i of vector b at i nrep + b): the batched
  // SpMV and Jacobi sweeps run in that layout, between transposes of the stacked vectors.
  DBuf<double> il[5];
  void ensure_interleaved(int64_t N) {
    for (auto& b : il)
      if ((int64_t)b.n < N) b.alloc(N);
  }
  // with inter_rd, inter * rd (the reciprocal diagonal of the sweeps) is written as well
  void interleave(const double* stacked, double* inter, int nrep, double* inter_rd = nullptr) {
    const dim3 grid((unsigned)((n + 31) / 32), (unsigned)((nrep + 31) / 32));
    mlk::k_interleave<true><<<grid, 256, 0, stream>>>(n, nrep, stacked, inter,
                                                      inter_rd ? d_sw_rd.p : nullptr, inter_rd);
  }
  void deinterleave(const double* inter, double* stacked, int nrep) {
    const dim3 grid((unsigned)((n + 31) / 32), (unsigned)((nrep + 31) / 32));
    mlk::k_interleave<false><<<grid, 256, 0, stream>>>(n, nrep, inter, stacked);
  }
  // lanes per row of the interleaved kernels (a wave takes 64 / lanes rows): the autotuned
  // width when set, else the vectors of a pass up to a wave; never wider than the pass needs
  int il_lanes = 0;
  int lanes_for(int nrep) const {
    const int fit = nrep > 32 ? 64 : nrep > 16 ? 32 : nrep > 8 ? 16 : 8;
    return il_lanes > 0 ? std::min(il_lanes, fit) : fit;
  }
  unsigned groups_grid(int nrep, int w) const {
    const int64_t threads = (int64_t)n * ((nrep + w - 1) / w) * w;
    return (unsigned)((threads + 255) / 256);
  }

  // y = A x (FP64) for nrep stacked vectors: the SpMV of every vector, bit-identical to its own
  // (k_spmm_il on the interleaved copies).
  void spmv_many(const double* x, double* y, int nrep) {
    ensure_interleaved((int64_t)n * nrep);
    interleave(x, il[0].p, nrep);
    const int w = lanes_for(nrep);
    const unsigned g = groups_grid(nrep, w);
    auto go = [&](auto kernel) {
      kernel<<<g, 256, 0, stream>>>(n, nrep, d_soff.p, d_scol.p, d_slen.p, d_sval.p, il[0].p,
                                    il[1].p);
    };
    switch (w) {
      case 8: go(mlk::k_spmm_il<double, 8>); break;
      case 16: go(mlk::k_spmm_il<double, 16>); break;
      case 32: go(mlk::k_spmm_il<double, 32>); break;
      default: go(mlk::k_spmm_il<double, 64>); break;
    }
    deinterleave(il[1].p, y, nrep);
    HIP_CHECK(hipGetLastError());
  }

  // z ~ (L L^T)^-1 r for nrep stacked right-hand sides: the sweeps of apply_sweeps, one
  // launch per sweep over every right-hand side (k_sweep_il on interleaved copies), each
  // system's result bit-identical to apply_sweeps on it alone. The first iterate of each
  // system (rhs * rd, the product the single system's forms take) is written beforehand: the
  // forward one by the transpose of r, the backward one by the last forward sweep.
  void apply_sweeps_rep(const double* r, double* z, int nrep) {
    hipEvent_t ev = prof.begin(stream);
    ensure_interleaved((int64_t)n * nrep);
    double* ri = il[0].p;   // r
    double* a = il[1].p;    // the forward first iterate, then backward scratch
    double* b = il[2].p;    // forward scratch, then z
    double* yi = il[3].p;   // y
    double* z0 = il[4].p;   // the backward first iterate
    interleave(r, ri, nrep, a);
    const int w = lanes_for(nrep);
    const unsigned g = groups_grid(nrep, w);
    const int k = sweeps;
    auto sweep = [&](const SweepDev& t, const double* rhs, const double* in, double* out,
                     bool level, double* out_rd) {
      const int sign = &t == &sw_lo ? -1 : 1;
      auto go = [&](auto kernel) {
        kernel<<<g, 256, 0, stream>>>(n, nrep, t.soff.p, t.scol.p, t.sdel.p, sign, t.slen.p,
                                      t.sval.p, d_sw_rd.p, rhs, in, out, level ? 1 : 0, out_rd,
                                      sweep_w, 1.0 - sweep_w);
      };
      auto width = [&](auto d16) {
        constexpr bool D = decltype(d16)::value;
        switch (w) {
          case 8: go(mlk::k_sweep_il<D, 8>); break;
          case 16: go(mlk::k_sweep_il<D, 16>); break;
          case 32: go(mlk::k_sweep_il<D, 32>); break;
          default: go(mlk::k_sweep_il<D, 64>); break;
        }
      };
      if (t.d16) width(std::true_type());
      else width(std::false_type());
    };
    // forward from a into yi (alternating with b), z0 = y rd; backward from z0 into b (with a)
    const double* prev = a;
    for (int m = 1; m <= k; ++m) {
      double* out = ((k - m) & 1) ? b : yi;
      sweep(sw_lo, ri, prev, out, m < k, m == k ? z0 : nullptr);
      prev = out;
    }
    prev = z0;
    for (int m = 1; m <= k; ++m) {
      double* out = ((k - m) & 1) ? a : b;
      sweep(sw_up, yi, prev, out, m < k, nullptr);
      prev = out;
    }
    deinterleave(b, z, nrep);
    HIP_CHECK(hipGetLastError());
    prof.end(P_TRI32, ev, stream);
  }

  template <typename T>
  void fill_sentinel(T* p) {
    mlk::k_fill_bits<T><<<grid_for(n), 256, 0, stream>>>(p, n, mlk::Bits<T>::S);
    HIP_CHECK(hipGetLastError());
  }

  // The handle's preconditioner, w_z = M^-1 w_r, as its PCG applies it: the exact solves with
  // the FP64 factor or the mixed handles' FP32 one, or the handle's Jacobi sweeps. A trace is
  // taken of the exact solves.
  void apply_precondition(long long* trace_fw = nullptr, long long* trace_bw = nullptr) {
    mlk::k_fill_i32<<<1, 64, 0, stream>>>(d_flags.p, 2, 0);
    if (sweeps > 0 && !trace_fw && !trace_bw) {
      if (fused_active()) fill_levels();
      apply_sweeps(w_r.p, w_z.p, nullptr);
    } else {
      fill_sentinel<double>(w_y.p);
      fill_sentinel<double>(w_z.p);
      precondition(w_r.p, w_y.p, w_z.p, nullptr, trace_fw, trace_bw);
    }
    HIP_CHECK(hipGetLastError());
  }

  void init_scalars(mlk::PcgScalars* S, int64_t maxit, bool reliable, bool stall_stop) {
    mlk::PcgScalars h;
    std::memset(&h, 0, sizeof(h));
    h.maxit = maxit;
    h.allow = reliable ? 1 : 0;
    h.stall_stop = stall_stop ? 1 : 0;
    h.delta = 0.1;
    // blocking copy from the stack: the stream must not hold kernels still reading S
    HIP_CHECK(hipStreamSynchronize(stream));
    HIP_CHECK(hipMemcpy(S, &h, sizeof(h), hipMemcpyHostToDevice));
  }

  mlk::PcgScalars read_scalars(const mlk::PcgScalars* S) {
    mlk::PcgScalars h;
    HIP_CHECK(hipMemcpyAsync(&h, S, sizeof(h), hipMemcpyDeviceToHost, stream));
    HIP_CHECK(hipStreamSynchronize(stream));
    return h;
  }

  template <typename TA>
  void spmv(const TA* x, TA* y, double* partial, const mlk::PcgScalars* S, int gated = 0) {
    const bool f32 = std::is_same<TA, float>::value;
    hipEvent_t ev = prof.begin(stream);
    const TA* val = f32 ? (const TA*)d_sval32.p : (const TA*)d_sval.p;
    if (spmv_grouped) {
      auto kernel = spmv_batch == 16 ? mlk::k_spmv_grouped<TA, 16> : mlk::k_spmv_grouped<TA, 8>;
      kernel<<<grid_for(n), 256, 0, stream>>>(n, d_soff.p, d_goff.p, d_gnum.p, d_gdel.p,
                                              d_ginfo.p, val, x, y, partial, S, gated);
    } else {
      auto kernel = spmv_d16 ? (spmv_batch == 16 ? mlk::k_spmv<TA, 16, true> : mlk::k_spmv<TA, 8, true>)
                             : (spmv_batch == 16 ? mlk::k_spmv<TA, 16> : mlk::k_spmv<TA, 8>);
      kernel<<<grid_for(n), 256, 0, stream>>>(n, d_soff.p, d_scol.p, d_sdel.p, d_slen.p, val, x,
                                              y, partial, S, gated);
    }
    HIP_CHECK(hipGetLastError());
    prof.end(f32 ? P_SPMV32 : P_SPMV64, ev, stream);
  }

  static constexpr int PCG_STALLED = 100;  // internal: the FP32 operator stopped helping

  // Preconditioned CG (iterative.rs:210-248) on the device. TA: arithmetic of the operator and
  // of p and Ap; TM: arithmetic of z (FP64: the mixed handles store their factor in FP32 but
  // apply it with FP64 vectors and sums, since FP32 intermediate vectors of the triangular
  // solves cost PCG iterations). x, r and every dot product are FP64. With an FP32 operator the recursive residual is replaced by the FP64 true residual
  // b - A x each time it has dropped tenfold (reliable updates); when a replacement fails to
  // halve the true residual the solve returns PCG_STALLED with its iterate, for the caller to
  // continue with the FP64 operator. Convergence is the reference's |r| <= tol |b| on the
  // residual in hand. x0 = 0 (the reference), or the x given when `warm` (r = b - A x in FP64).
  // b, x, r are device vectors in solver numbering; returns a status and fills *hS.
  template <typename TA, typename TM>
  int pcg(const double* b, double* x, double* r, double tol, int64_t maxit, bool warm,
          mlk::PcgScalars* S, mlk::PcgScalars* hS) {
    const bool a32 = std::is_same<TA, float>::value;
    static_assert(std::is_same<TM, double>::value, "the preconditioner runs on FP64 vectors");
    TA* p = a32 ? (TA*)w_p32.p : (TA*)w_p.p;
    TA* ap = a32 ? (TA*)w_ap32.p : (TA*)w_ap.p;
    TM* y = w_y.p;
    TM* z = w_z.p;
    TM* rlow = nullptr;
    const TM* rin = r;
    const int vb = grid_for(n);
    double* part = w_partial.p;
    const bool jac = sweeps > 0;
    // the fused sweeps form r * rd themselves (reading it from a vector written by the residual
    // update measured slower: the extra writes cost more than the gathers they save, which
    // mostly hit the L2); the per-launch ones read it from the update
    const bool y0 = jac && !fused_active();
    // the reference's dot products (eight sequential lane sums, k_dot8) instead of the tree
    // sums of block partials: each scalar is then the reference's, bit for bit
    const bool d8 = ref_dots && !a32;
    const int parts = d8 ? 1 : vb;  // the partials the scalar kernels reduce
    auto dot8 = [&](const double* u, const double* v, const mlk::PcgScalars* gate) {
      mlk::k_dot8<<<1, mlk::DOT8_THREADS, 0, stream>>>(n, u, v, part, gate);
    };
    mlk::k_fill_i32<<<1, 64, 0, stream>>>(d_flags.p, 1, 0);  // abort flag
    if (fused_active()) fill_levels();
    init_scalars(S, maxit, a32, a32);
    if (warm) {
      spmv<double>(x, w_ax.p, nullptr, nullptr);
      mlk::k_residual<<<vb, 256, 0, stream>>>(n, b, w_ax.p, r, part);
    } else {
      mlk::k_fill_f64<<<vb, 256, 0, stream>>>(x, n, 0.0);
      HIP_CHECK(hipMemcpyAsync(r, b, sizeof(double) * n, hipMemcpyDeviceToDevice, stream));
    }
    if (d8) dot8(b, b, nullptr);
    else mlk::k_dot<double, double><<<vb, 256, 0, stream>>>(n, b, b, part);
    mlk::k_init_norm<<<1, mlk::SUM_THREADS, 0, stream>>>(part, parts, S, tol);
    HIP_CHECK(hipGetLastError());
    if (jac) {
      apply_sweeps(rin, z, S, d8 ? nullptr : r, d8 ? nullptr : part);
    } else {
      fill_sentinel<TM>(y);
      fill_sentinel<TM>(z);
      precondition(rin, y, z, S);
      if (!d8) mlk::k_dot<double, TM><<<vb, 256, 0, stream>>>(n, r, z, part);
    }
    if (d8) dot8(r, z, S);
    mlk::k_init_rz<<<1, mlk::SUM_THREADS, 0, stream>>>(part, parts, S);
    mlk::k_convert<TM, TA><<<vb, 256, 0, stream>>>(n, z, p);
    HIP_CHECK(hipGetLastError());
    *hS = read_scalars(S);
    if (read_flag(0)) return ML_DEVICE_TIMEOUT;
    if (hS->done) return ML_OK;  // b == 0: x = 0 exactly
    int64_t launched = 0, batch = 4;
    for (;;) {
      for (int64_t k = 0; k < batch; ++k) {
        spmv<TA>(p, ap, d8 ? nullptr : part, S);
        if (d8) dot8((const double*)p, (const double*)ap, S);
        mlk::k_alpha<<<1, mlk::SUM_THREADS, 0, stream>>>(part, parts, S);
        hipEvent_t ev = prof.begin(stream);
        mlk::k_update_xr<TA, TM><<<vb, 256, 0, stream>>>(n, S, x, r, p, ap, rlow,
                                                         jac ? nullptr : y, jac ? nullptr : z,
                                                         part, y0 ? d_sw_rd.p : nullptr,
                                                         y0 ? w_y0.p : nullptr);
        prof.end(P_VEC, ev, stream);
        if (d8) dot8(r, r, S);
        mlk::k_converged<<<1, mlk::SUM_THREADS, 0, stream>>>(part, parts, S);
        if (a32) {
          spmv<double>(x, w_ax.p, nullptr, S, 1);
          mlk::k_replace_residual<TM><<<vb, 256, 0, stream>>>(n, S, b, w_ax.p, r, rlow, part);
          mlk::k_replaced<<<1, mlk::SUM_THREADS, 0, stream>>>(part, vb, S);
        }
        if (jac) {
          // the residual replacement (ML_MIXED) rewrites r after k_update_xr: then y0 is
          // formed by the first sweep itself
          apply_sweeps(rin, z, S, d8 ? nullptr : r, d8 ? nullptr : part,
                       a32 || !y0 ? nullptr : w_y0.p);
        } else {
          precondition(rin, y, z, S);
          if (!d8) {
            ev = prof.begin(stream);
            mlk::k_dot<double, TM><<<vb, 256, 0, stream>>>(n, r, z, part);
            prof.end(P_VEC, ev, stream);
          }
        }
        if (d8) dot8(r, z, S);
        mlk::k_beta<<<1, mlk::SUM_THREADS, 0, stream>>>(part, parts, S);
        ev = prof.begin(stream);
        mlk::k_update_p<TA, TM><<<vb, 256, 0, stream>>>(n, S, p, z);
        prof.end(P_VEC, ev, stream);
        HIP_CHECK(hipGetLastError());
      }
      launched += batch;
      *hS = read_scalars(S);
      if (read_flag(0)) return ML_DEVICE_TIMEOUT;
      if (hS->done) break;
      if (launched > maxit + 8) break;  // cannot happen: k_converged stops at maxit
      batch = std::min<int64_t>(32, batch * 2);
    }
    if (hS->status == 3) return PCG_STALLED;
    if (hS->status == 1) return ML_SINGULAR;
    if (hS->status == 2) return ML_NOT_CONVERGED;
    return ML_OK;
  }

  // ---------------------------------------------------------------- autotuner
  // Average duration (ms) of `reps` preconditioner applications to w_r (or nrep copies of it)
  // on the current grids; negative if a launch stalled.
  double time_precondition(int reps, int nrep, double* r, double* y, double* z) {
    typedef double T;
    const int64_t N = (int64_t)n * nrep;
    hipEvent_t a, b;
    HIP_CHECK(hipEventCreate(&a));
    HIP_CHECK(hipEventCreate(&b));
    mlk::k_fill_i32<<<1, 64, 0, stream>>>(d_flags.p, 1, 0);
    float total = 0.0f;
    for (int k = 0; k <= reps; ++k) {  // the first application warms up
      mlk::k_fill_bits<T><<<grid_for(N), 256, 0, stream>>>(y, N, mlk::Bits<T>::S);
      mlk::k_fill_bits<T><<<grid_for(N), 256, 0, stream>>>(z, N, mlk::Bits<T>::S);
      HIP_CHECK(hipEventRecord(a, stream));
      precondition(r, y, z, nullptr, nullptr, nullptr, nrep);
      HIP_CHECK(hipEventRecord(b, stream));
      HIP_CHECK(hipEventSynchronize(b));
      float t = 0.0f;
      HIP_CHECK(hipEventElapsedTime(&t, a, b));
      if (k > 0) total += t;
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    if (read_flag(0)) return -1.0;
    return total / reps;
  }

  // Seconds of one run of `work` (stream-ordered device work), after one that warms up.
  template <class F>
  double seconds_of(F work) {
    hipEvent_t a, b;
    HIP_CHECK(hipEventCreate(&a));
    HIP_CHECK(hipEventCreate(&b));
    work();
    HIP_CHECK(hipEventRecord(a, stream));
    work();
    HIP_CHECK(hipEventRecord(b, stream));
    HIP_CHECK(hipEventSynchronize(b));
    float ms = 0.0f;
    HIP_CHECK(hipEventElapsedTime(&ms, a, b));
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    return 1e-3 * ms;
  }

  // Seconds of one application of the handle's preconditioner as its PCG applies it (the
  // exact solves of the factor, or the handle's Jacobi sweeps), after one that warms up;
  // -1 when a wait gave up.
  double application_seconds() {
    if (sweeps == 0) {
      const double ms = time_precondition(1, 1, w_r.p, w_y.p, w_z.p);
      return ms < 0.0 ? -1.0 : 1e-3 * ms;
    }
    mlk::k_fill_i32<<<1, 64, 0, stream>>>(d_flags.p, 1, 0);
    const double t = seconds_of([&] { apply_sweeps(w_r.p, w_z.p, nullptr); });
    if (read_flag(0)) {
      mlk::k_fill_i32<<<1, 64, 0, stream>>>(d_flags.p, 1, 0);
      return -1.0;
    }
    return t;
  }

  // Picks, by timing, the waves per CU of the triangular solves (single right-hand side) and,
  // for nrhs > 1, the right-hand sides per pass. Results do not change: the solves are
  // bit-exact under any schedule. chosen[3]: waves per CU, right-hand sides per pass, and the
  // per-CU occupancy the runtime reports. The entries per batch of loads of the SpMV and the
  // sweeps are timed too (bit-exact either way). An ML_MIXED_JACOBI handle whose sweep count
  // the caller has not set then times whole solves of a fixed pseudo-random right-hand side
  // with 0 (the exact solves), 2, 3, 4, 6 and 8 sweeps and keeps the fastest that converges
  // (which changes its iterates, within the solve's tolerance).
  int autotune(int64_t nrhs, int32_t* chosen) {
    if (!factored) {
      int st = factor();
      if (st != ML_OK) return st;
    }
    const bool c32 = precision != ML_FP64;
    const int occ =
        c32 ? occupancy_per_cu((const void*)mlk::k_trisolve<double, false, false, false, float>,
                               mlk::TRI_THREADS)
            : occupancy_per_cu((const void*)mlk::k_trisolve<double, false, false>, mlk::TRI_THREADS);
    const int cus = device_cus();
    const int nblk = std::max(fw.nblk, bw.nblk);
    const int top = tri_per_cu;  // the probed co-resident bound
    mlk::k_fill_f64<<<grid_for(n), 256, 0, stream>>>(w_r.p, n, 1.0);
    // entries per batch of loads of the SpMV and of the sweeps: the faster of 8 and 16
    {
      hipEvent_t a, b;
      HIP_CHECK(hipEventCreate(&a));
      HIP_CHECK(hipEventCreate(&b));
      auto time_of = [&](auto&& work) {
        float total = 0.0f;
        for (int k = 0; k <= 3; ++k) {  // the first run warms up
          HIP_CHECK(hipEventRecord(a, stream));
          work();
          HIP_CHECK(hipEventRecord(b, stream));
          HIP_CHECK(hipEventSynchronize(b));
          float t = 0.0f;
          HIP_CHECK(hipEventElapsedTime(&t, a, b));
          if (k > 0) total += t;
        }
        return total;
      };
      // the SpMV's column streams (grouped, when built, or per row) and loads in flight
      float best_t = 1e30f;
      int best_b = 8;
      bool best_g = false;
      for (int grouped = have_grouped ? 1 : 0; grouped >= 0; --grouped)
        for (int bsz : {8, 16}) {
          spmv_grouped = grouped == 1;
          spmv_batch = bsz;
          const float t = time_of([&] { spmv<double>(w_r.p, w_ap.p, nullptr, nullptr); });
          if (t < best_t) { best_t = t; best_b = bsz; best_g = spmv_grouped; }
        }
      spmv_batch = best_b;
      spmv_grouped = best_g;
      if (have_sweep && !(have_fused && fused_on)) {  // the fused kernel has no batches
        const int k0 = sweeps;
        sweeps = std::max(1, sweeps);
        best_t = 1e30f;
        for (int bsz : {8, 16}) {
          sweep_batch = bsz;
          const float t = time_of([&] { apply_sweeps(w_r.p, w_z.p, nullptr); });
          if (t < best_t) { best_t = t; best_b = bsz; }
        }
        sweep_batch = best_b;
        sweeps = k0;
      }
      (void)hipEventDestroy(a);
      (void)hipEventDestroy(b);
    }
    double best = 1e300;
    int best_pc = tri_per_cu;
    std::vector<int> cands;  // even counts below the bound, and the bound itself
    for (int pc = 2; pc < top; pc += 2) cands.push_back(pc);
    cands.push_back(top);
    for (int pc : cands) {
      (c32 ? full_tri32 : full_tri64) = std::min(nblk, pc * cus);
      shrinks = 0;
      set_grids();
      const double t = time_precondition(3, 1, w_r.p, w_y.p, w_z.p);
      if (t > 0 && t < best) {
        best = t;
        best_pc = pc;
      }
    }
    // the replicated solves of batched passes keep their own grid (cap_tri): with many copies
    // of every block they are bound by throughput, not by the dependency chain
    (c32 ? full_tri32 : full_tri64) = std::min(nblk, best_pc * cus);
    shrinks = 0;
    set_grids();
    if (nrhs > 1) {
      // right-hand sides per pass: the cheapest application per right-hand side
      double best_per = 1e300;
      int best_k = rhs_per_pass;
      for (int k : {8, 16, 32, 64, 128, 256}) {
        if (k > nrhs && k != 8) break;
        const int64_t N = (int64_t)n * k;
        DBuf<double> r64, y64, z64;
        r64.alloc(N); y64.alloc(N); z64.alloc(N);
        mlk::k_fill_f64<<<grid_for(N), 256, 0, stream>>>(r64.p, N, 1.0);
        double t;
        if (sweeps > 0) {  // the handle's Jacobi sweeps over every right-hand side
          t = 1e3 * seconds_of([&] { apply_sweeps_rep(r64.p, z64.p, k); });
        } else {
          t = time_precondition(2, k, r64.p, y64.p, z64.p);
        }
        if (t > 0 && t / k < best_per) {
          best_per = t / k;
          best_k = k;
        }
      }
      rhs_per_pass = best_k;
      // rows per wave of the interleaved SpMV and sweeps (64 / lanes per row), at that pass
      const int k = (int)std::min<int64_t>(rhs_per_pass, std::max<int64_t>(nrhs, 8));
      const int64_t N = (int64_t)n * k;
      DBuf<double> r64, z64;
      r64.alloc(N);
      z64.alloc(N);
      mlk::k_fill_f64<<<grid_for(N), 256, 0, stream>>>(r64.p, N, 1.0);
      double best_w = 1e300;
      int lanes = 0;
      for (int w : {64, 32, 16, 8}) {
        il_lanes = w;
        if (lanes_for(k) != w) continue;
        const double t = seconds_of([&] {
          spmv_many(r64.p, z64.p, k);
          if (sweeps > 0) apply_sweeps_rep(r64.p, z64.p, k);
        });
        if (t < best_w) {
          best_w = t;
          lanes = w;
        }
      }
      il_lanes = lanes;
    }
    if (have_fused && fused_on) {
      // the fused kernel's variant: the fastest application, among the variants no wider
      // than the one that holds every row in registers
      const int k0 = sweeps;
      if (sweeps < 1) set_sweeps(4);
      hipEvent_t a, b;
      HIP_CHECK(hipEventCreate(&a));
      HIP_CHECK(hipEventCreate(&b));
      float best_t = 1e30f;
      int best_v = fused_var;
      for (int v = 0; v < FUSED_VARIANTS; ++v) {
        if (!fused_candidate(v)) continue;
        select_fused(v);
        fill_levels();
        float total = 0.0f;
        for (int r = 0; r <= 3; ++r) {  // the first run warms up
          HIP_CHECK(hipEventRecord(a, stream));
          apply_sweeps(w_r.p, w_z.p, nullptr);
          HIP_CHECK(hipEventRecord(b, stream));
          HIP_CHECK(hipEventSynchronize(b));
          float t = 0.0f;
          HIP_CHECK(hipEventElapsedTime(&t, a, b));
          if (r > 0) total += t;
        }
        if (read_flag(0)) total = 1e30f;  // a wait gave up: not a candidate
        mlk::k_fill_i32<<<1, 64, 0, stream>>>(d_flags.p, 1, 0);
        if (total < best_t) {
          best_t = total;
          best_v = v;
        }
      }
      (void)hipEventDestroy(a);
      (void)hipEventDestroy(b);
      select_fused(best_v);
      fill_levels();
      if (k0 != sweeps) set_sweeps(k0);
    }
    if (precision == ML_MIXED_JACOBI && !sweeps_pinned) {
      DBuf<double> tb, tx;
      tb.alloc(n);
      tx.alloc(n);
      mlk::k_fill_hash<<<grid_for(n), 256, 0, stream>>>(tb.p, n, 0x6d61696c6c65ull);
      HIP_CHECK(hipGetLastError());
      // whole solves at each sweep count, a likely winner first: a later count runs only as
      // many iterations as could still beat the fastest solve so far, estimated from one timed
      // application of its preconditioner and the rest of an iteration of that fastest solve
      // (with a 10% margin); a count that reaches that cap is slower and is dropped
      double best_t = 1e300, other = -1.0;
      int best_k = sweeps;
      double best_w = sweep_w;
      // (sweeps, weight): undamped first, then damped by 0.8, which shortens the solves whose
      // factor's off-diagonal part is large against its diagonal (Tet10); when the first damped
      // count does not beat the fastest undamped one, the other damped counts are skipped
      static const std::pair<int, double> kCandidates[] = {
          {6, 1.0}, {4, 1.0}, {8, 1.0}, {5, 1.0}, {3, 1.0}, {2, 1.0},
          {4, 0.8}, {6, 0.8}, {3, 0.8}, {8, 0.8}, {2, 0.8}, {0, 1.0}};
      bool damped_seen = false, skip_damped = false;
      for (const auto& cand : kCandidates) {
        const int k = cand.first;
        const bool damped = cand.second != 1.0;
        if (damped && skip_damped) continue;
        const bool first_damped = damped && !damped_seen;
        damped_seen = damped_seen || damped;
        set_sweeps(k);
        sweep_w = cand.second;
        const double app = application_seconds();
        if (app < 0.0) continue;  // a wait gave up: not a candidate
        int64_t cap = ML_MAX_ITERATIONS;
        if (other >= 0.0)
          cap = std::min<int64_t>(cap, (int64_t)(1.1 * best_t / (other + app)) + 10);
        mlk::PcgScalars hS;
        HIP_CHECK(hipStreamSynchronize(stream));
        const double t0 = now_s();
        const int st = pcg<double, double>(tb.p, tx.p, w_r.p, ML_TOLERANCE, cap, false, d_S.p,
                                           &hS);
        const double t = now_s() - t0;
        if (st == ML_DEVICE_TIMEOUT) return st;
        if (st == ML_OK && t < best_t) {
          best_t = t;
          best_k = k;
          best_w = cand.second;
          other = std::max(0.0, t / (double)std::max<int64_t>(hS.it, 1) - app);
        } else if (first_damped) {
          skip_damped = true;
        }
      }
      set_sweeps(best_k);
      sweep_w = best_w;
    }
    if (chosen) {
      chosen[0] = best_pc;
      chosen[1] = rhs_per_pass;
      chosen[2] = occ;
    }
    return ML_OK;
  }

  // ---------------------------------------------------------------- batched PCG
  // Workspace of the batched solves (allocated on first use, grown as needed).
  DBuf<double> wb_r, wb_p, wb_ap, wb_dot;
  DBuf<double> wb_y64, wb_z64;
  DBuf<mlk::PcgScalars> wb_S;

  // Per-system shift ladder of a block-diagonal operator (Pcg::new, iterative.rs:201-207, for
  // each system on its own): one IC(0) launch per round over every system, each at its own
  // shift; the systems that failed climb one step and the round repeats. shift_idx receives
  // each system's shift index.
  template <typename T>
  int factor_segments(int nseg, const int* seg_of, std::vector<int>* shift_idx) {
    static const double SHIFTS[6] = {0.0, 1e-3, 1e-2, 1e-1, 1.0, 10.0};
    const bool f32 = std::is_same<T, float>::value;
    factored = false;
    shift_idx->assign(nseg, 0);
    DBuf<T> d_shift;
    DBuf<int> d_fail;
    d_shift.alloc(nseg);
    d_fail.alloc(nseg);
    std::vector<T> h_shift(nseg);
    std::vector<int> h_fail(nseg);
    T* l = f32 ? (T*)d_l32.p : (T*)d_l.p;
    const T* a = f32 ? (const T*)d_a32.p : (const T*)d_a.p;
    for (int round = 0;; ++round) {
      for (int s = 0; s < nseg; ++s) h_shift[s] = (T)SHIFTS[(*shift_idx)[s]];
      HIP_CHECK(hipMemcpyAsync(d_shift.p, h_shift.data(), sizeof(T) * nseg, hipMemcpyHostToDevice, stream));
      int st = run_ic0<T>(a, l, T(0), seg_of, d_shift.p, d_fail.p, nseg);
      if (st == ML_DEVICE_TIMEOUT && shrink()) continue;
      if (st != ML_OK) return st;
      HIP_CHECK(hipMemcpyAsync(h_fail.data(), d_fail.p, sizeof(int) * nseg, hipMemcpyDeviceToHost, stream));
      HIP_CHECK(hipStreamSynchronize(stream));
      bool any = false;
      for (int s = 0; s < nseg; ++s) {
        if (h_fail[s] == 2) throw ArgError("matrix row longer than 512 entries");
        if (h_fail[s]) {
          any = true;
          if (++(*shift_idx)[s] >= 6) return ML_SINGULAR;
        }
      }
      if (!any) break;
    }
    gather_factor<T>(l);
    if (f32 && have_sweep) gather_sweeps();
    factored = true;
    return ML_OK;
  }

  // nseg independent PCG solves (iterative.rs:210-248 each) in one set of launches: either
  // nrep right-hand sides of this operator stored one after the other (nseg = nrep), or the
  // systems of this block-diagonal operator (nrep = 1). System s owns the vector entries
  // [seg_off[s], seg_off[s+1]) and seg_of[i] is the system of entry i. FP64 operator,
  // iterate and preconditioner vectors (the factor's coefficients in the handle's precision).
  // A handle with Jacobi sweeps applies its factor by the sweeps: on a block-diagonal operator
  // (nrep = 1) they act on each model's rows alone, since no entry couples two models, and
  // replicated right-hand sides take them one launch per sweep over all of them
  // (apply_sweeps_rep). The scalars of every system end in *hS.
  template <typename TM>
  int pcg_batch(int nrep, int nseg, const int64_t* seg_off, const int* seg_of, const double* b,
                double* x, double tol, int64_t maxit, std::vector<mlk::PcgScalars>* hS) {
    static_assert(std::is_same<TM, double>::value, "the preconditioner runs on FP64 vectors");
    const int64_t N = (int64_t)n * nrep;
    for (auto* v : {&wb_r, &wb_p, &wb_ap}) if ((int64_t)v->n < N) v->alloc(N);
    if ((int64_t)wb_dot.n < 4 * nseg) wb_dot.alloc(4 * nseg);
    if ((int64_t)wb_S.n < nseg) wb_S.alloc(nseg);
    for (auto* v : {&wb_y64, &wb_z64}) if ((int64_t)v->n < N) v->alloc(N);
    const bool jac = sweeps > 0;
    TM* y = wb_y64.p;
    TM* z = wb_z64.p;
    TM* rlow = nullptr;
    double* r = wb_r.p;
    double* p = wb_p.p;
    double* ap = wb_ap.p;
    double* bb = wb_dot.p;
    double* pap = wb_dot.p + nseg;
    double* rr = wb_dot.p + 2 * nseg;
    double* rz = wb_dot.p + 3 * nseg;
    mlk::PcgScalars* S = wb_S.p;
    const TM* rin = r;
    const unsigned vb = grid_for(N), sb = grid_for(nseg);
    auto apply_m = [&]() {
      if (jac && nrep == 1) apply_sweeps(rin, z, nullptr);
      else if (jac) apply_sweeps_rep(rin, z, nrep);
      else precondition(rin, y, z, nullptr, nullptr, nullptr, nrep);
    };
    mlk::k_fill_i32<<<1, 64, 0, stream>>>(d_flags.p, 1, 0);  // abort flag
    if (jac && nrep == 1 && fused_active()) fill_levels();
    mlk::k_fill_f64<<<vb, 256, 0, stream>>>(x, N, 0.0);
    HIP_CHECK(hipMemcpyAsync(r, b, sizeof(double) * N, hipMemcpyDeviceToDevice, stream));
    mlk::k_seg_dot<double, double><<<nseg, 256, 0, stream>>>(seg_off, b, b, bb);
    mlk::k_seg_init<<<sb, 256, 0, stream>>>(nseg, bb, S, tol, maxit);
    mlk::k_fill_bits<TM><<<vb, 256, 0, stream>>>(y, N, mlk::Bits<TM>::S);
    mlk::k_fill_bits<TM><<<vb, 256, 0, stream>>>(z, N, mlk::Bits<TM>::S);
    apply_m();
    mlk::k_seg_dot<double, TM><<<nseg, 256, 0, stream>>>(seg_off, r, z, rz);
    mlk::k_seg_rz<<<sb, 256, 0, stream>>>(nseg, rz, S);
    mlk::k_convert<TM, double><<<vb, 256, 0, stream>>>(N, z, p);
    HIP_CHECK(hipGetLastError());
    auto spmv_all = [&](const double* in, double* out) {
      hipEvent_t ev = prof.begin(stream);
      if (nrep > 1)
        spmv_many(in, out, nrep);
      else
        mlk::k_spmv<double><<<vb, 256, 0, stream>>>(n, d_soff.p, d_scol.p, nullptr, d_slen.p,
                                                     d_sval.p,
                                                     in, out, nullptr, nullptr, 0);
      prof.end(P_SPMV64, ev, stream);
    };
    hS->resize(nseg);
    auto all_done = [&]() {
      HIP_CHECK(hipMemcpyAsync(hS->data(), S, sizeof(mlk::PcgScalars) * nseg, hipMemcpyDeviceToHost, stream));
      HIP_CHECK(hipStreamSynchronize(stream));
      for (const auto& q : *hS) if (!q.done) return false;
      return true;
    };
    if (read_flag(0)) return ML_DEVICE_TIMEOUT;
    if (all_done()) return ML_OK;
    int64_t launched = 0, batch = 4;
    for (;;) {
      for (int64_t k = 0; k < batch; ++k) {
        spmv_all(p, ap);
        mlk::k_seg_dot<double, double><<<nseg, 256, 0, stream>>>(seg_off, p, ap, pap);
        mlk::k_seg_alpha<<<sb, 256, 0, stream>>>(nseg, pap, S);
        hipEvent_t ev = prof.begin(stream);
        mlk::k_seg_update_xr<TM><<<vb, 256, 0, stream>>>(N, seg_of, S, x, r, p, ap, rlow, y, z);
        prof.end(P_VEC, ev, stream);
        mlk::k_seg_dot<double, double><<<nseg, 256, 0, stream>>>(seg_off, r, r, rr);
        mlk::k_seg_converged<<<sb, 256, 0, stream>>>(nseg, rr, S);
        apply_m();
        mlk::k_seg_dot<double, TM><<<nseg, 256, 0, stream>>>(seg_off, r, z, rz);
        mlk::k_seg_beta<<<sb, 256, 0, stream>>>(nseg, rz, S);
        ev = prof.begin(stream);
        mlk::k_seg_update_p<TM><<<vb, 256, 0, stream>>>(N, seg_of, S, p, z);
        prof.end(P_VEC, ev, stream);
        HIP_CHECK(hipGetLastError());
      }
      launched += batch;
      if (read_flag(0)) return ML_DEVICE_TIMEOUT;
      if (all_done()) break;
      if (launched > maxit + 8) break;  // cannot happen: every system stops at maxit
      batch = std::min<int64_t>(32, batch * 2);
    }
    return ML_OK;
  }

  // The batched solve of nrep right-hand sides b (device, one after the other) of this
  // operator; per-system status and scalars in *hS. A stalled dataflow launch repeats the
  // solve on smaller grids.
  int solve_many(int nrep, const double* b, double* x, double tol, int64_t maxit,
                 std::vector<mlk::PcgScalars>* hS) {
    if (!factored) {
      int st = factor();
      if (st != ML_OK) return st;
    }
    if (tol <= 0) tol = ML_TOLERANCE;
    if (maxit <= 0) maxit = ML_MAX_ITERATIONS;
    std::vector<int64_t> off(nrep + 1);
    std::vector<int> seg((int64_t)n * nrep);
    for (int k = 0; k <= nrep; ++k) off[k] = (int64_t)k * n;
    for (int k = 0; k < nrep; ++k) std::fill(seg.begin() + (int64_t)k * n, seg.begin() + (int64_t)(k + 1) * n, k);
    DBuf<int64_t> d_off_seg;
    DBuf<int> d_seg;
    d_off_seg.upload_vec(off, stream);
    d_seg.upload_vec(seg, stream);
    int st;
    for (;;) {
      // batched passes apply the factor by the exact solves (FP32 coefficients for the mixed
      // handles), whose replicated blocks fill the device, or by the handle's Jacobi sweeps
      st = pcg_batch<double>(nrep, nrep, d_off_seg.p, d_seg.p, b, x, tol, maxit, hS);
      if (st == ML_DEVICE_TIMEOUT && shrink()) continue;
      break;
    }
    HIP_CHECK(hipStreamSynchronize(stream));
    return st;
  }

  // |b - A x| / |b| in FP64 (uses w_ax, w_d as scratch)
  double true_residual(const double* b, const double* x, double bnorm) {
    const int vb = grid_for(n);
    spmv<double>(x, w_ax.p, nullptr, nullptr);
    mlk::k_residual<<<vb, 256, 0, stream>>>(n, b, w_ax.p, w_d.p, w_partial.p);
    mlk::k_init_norm<<<1, mlk::SUM_THREADS, 0, stream>>>(w_partial.p, vb, d_Sin.p, 1.0);
    HIP_CHECK(hipGetLastError());
    mlk::PcgScalars h = read_scalars(d_Sin.p);
    return bnorm > 0 ? h.bnorm / bnorm : 0.0;
  }

  // Solve A x = b (device vectors, solver numbering).
  int solve(const double* b, double* x, double tol, int64_t maxit, ml_solve_info* info) {
    if (!factored) {
      int st = factor();
      if (st != ML_OK) return st;
    }
    if (tol <= 0) tol = ML_TOLERANCE;
    if (maxit <= 0) maxit = ML_MAX_ITERATIONS;
    mlk::PcgScalars hS;
    int st;
    int64_t it0 = 0, updates0 = 0;
    int escalated = 0;
    const int retries0 = retries;
    for (;;) {
      it0 = updates0 = 0;
      escalated = 0;
      if (precision == ML_FP64) {
        st = pcg<double, double>(b, x, w_r.p, tol, maxit, false, d_S.p, &hS);
      } else if (precision == ML_MIXED) {
        st = pcg<float, double>(b, x, w_r.p, tol, maxit, false, d_S.p, &hS);
        if (st == PCG_STALLED) {
          // continue from the iterate with the FP64 operator (the preconditioner stays FP32)
          it0 = hS.it;
          updates0 = hS.updates;
          escalated = 1;
          st = pcg<double, double>(b, x, w_r.p, tol, std::max<int64_t>(1, maxit - it0), true,
                                   d_S.p, &hS);
        }
      } else {
        // ML_MIXED_IC0 and ML_MIXED_JACOBI: FP64 operator, the FP32 factor
        st = pcg<double, double>(b, x, w_r.p, tol, maxit, false, d_S.p, &hS);
      }
      // a stalled triangular solve: solve again from x0 = 0 on smaller grids
      if (st == ML_DEVICE_TIMEOUT && shrink()) continue;
      break;
    }
    if (info) {
      info->device_retries = retries - retries0;
      info->iterations = it0 + hS.it;
      info->residual = hS.bnorm > 0 ? hS.rnorm / hS.bnorm : 0.0;
      info->shift_index = shift_index;
      info->refinements = (int32_t)(updates0 + hS.updates);
      info->escalated = escalated;
      info->true_residual = (st == ML_OK && hS.bnorm > 0) ? true_residual(b, x, hS.bnorm)
                                                          : info->residual;
    }
    return st;
  }
};

// ====================================================================== Model
struct Model {
  MeshView mesh;
  std::vector<double> nodes, young, nu, imposed;
  std::vector<int64_t> conn;
  std::vector<uint8_t> fixed;
  Plan plan;
  Operator op;
  int ngp = 1;
  DBuf<double> d_xyz, d_young, d_nu, d_ke, d_imposed, d_coupling, d_f, d_u, d_R, d_x, d_rhs;
  DBuf<double> d_ulast;  // displacements of the last solve (adjoint state of the compliance)
  // node -> (element, local node) incidences in element order, built on first use
  DBuf<int64_t> d_inc_off;
  DBuf<int> d_inc_item;
  bool have_inc = false;
  DBuf<int> d_conn, d_rank, d_cpl_dof, d_rea_dof;
  DBuf<int64_t> d_asm_off, d_cpl_off, d_rea_off, d_rea_item;
  DBuf<uint32_t> d_asm_slot, d_cpl_slot;
  bool assembled = false;
  bool have_u = false;

  hipStream_t s() { return op.stream; }

  int create(int64_t n_nodes, const double* xyz, int64_t ne, int nen_, const int64_t* cn,
             const double* E, const double* v, const uint8_t* fx, const double* imp,
             int precision, int solver, ml_solve_info* info) {
    if (nen_ != 4 && nen_ != 10) throw ArgError("nen must be 4 (Tet4) or 10 (Tet10)");
    nodes.assign(xyz, xyz + 3 * n_nodes);
    conn.assign(cn, cn + ne * nen_);
    young.assign(E, E + ne);
    nu.assign(v, v + ne);
    fixed.assign(fx, fx + 3 * n_nodes);
    imposed.assign(3 * n_nodes, 0.0);
    if (imp) imposed.assign(imp, imp + 3 * n_nodes);
    mesh = MeshView{n_nodes, ne, nen_, nodes.data(), conn.data(), young.data(), nu.data()};
    ngp = nen_ == 4 ? 1 : 4;
    i64 where = -1;
    int merr = check_mesh(mesh, &where);
    if (info) { info->mesh_error = merr; info->bad_element = where; }
    if (merr) return ML_INVALID_MESH;
    std::string err;
    int st = build_plan(mesh, fixed.data(), imposed.data(), plan, &err);
    if (st) throw ArgError(err);
    if (info) info->seconds[5] = plan.seconds;
    // SolverKind::Auto (solve.rs:185-194) is Direct below ITERATIVE_FROM free DOFs: the
    // caller's skyline arm, not this library's
    if (solver == ML_AUTO && plan.nfree < ML_ITERATIVE_FROM) {
      g_last_error = "ML_AUTO with " + std::to_string(plan.nfree) + " free DOFs (< " +
                     std::to_string(ML_ITERATIVE_FROM) + "): SolverKind::Auto is Direct here";
      return ML_DIRECT_REQUIRED;
    }
    op.precision = precision;
    // with every DOF imposed there is no free system: u is the imposed field and only the
    // reactions are computed (FreeSystem::new and ::solve with zero unknowns)
    if (plan.nfree > 0) op.set_pattern(plan.lower);
    else op.init_stream();
    upload();
    assemble();
    return ML_OK;
  }
  bool all_fixed() const { return plan.nfree == 0; }

  void upload() {
    hipStream_t st = s();
    d_xyz.upload_vec(nodes, st);
    std::vector<int> c32(conn.begin(), conn.end());
    d_conn.upload_vec(c32, st);
    d_young.upload_vec(young, st);
    d_nu.upload_vec(nu, st);
    d_imposed.upload_vec(imposed, st);
    d_rank.upload_vec(plan.rank_of_dof, st);
    d_asm_off.upload_vec(plan.asm_off, st);
    d_asm_slot.upload_vec(plan.asm_slot, st);
    d_cpl_off.upload_vec(plan.cpl_off, st);
    d_cpl_slot.upload_vec(plan.cpl_slot, st);
    d_cpl_dof.upload_vec(plan.cpl_dof, st);
    d_rea_dof.upload_vec(plan.rea_dof, st);
    d_rea_off.upload_vec(plan.rea_off, st);
    d_rea_item.upload_vec(plan.rea_item, st);
    d_ke.alloc((size_t)plan.m * plan.m * plan.n_elements);
    d_coupling.alloc(plan.nfree);
    d_x.alloc(plan.nfree);
    d_rhs.alloc(plan.nfree);
    d_f.alloc(plan.ndof);
    d_u.alloc(plan.ndof);
    d_ulast.alloc(plan.ndof);
    d_R.alloc(plan.ndof);
    HIP_CHECK(hipStreamSynchronize(st));
    // the host copies of the big lists are no longer needed
    decltype(plan.asm_slot)().swap(plan.asm_slot);
  }

  // Tet10: the kernel that stages each Gauss point's mapping in LDS (k_stiffness_lds), or the
  // one that recomputes it per column (k_stiffness); bit-identical, the autotuner keeps the
  // faster
  bool stiffness_lds = true;

  void stiffness() {
    const int64_t ne = plan.n_elements;
    hipEvent_t ev = op.prof.begin(s());
    if (plan.nen == 4)
      mlk::k_stiffness<4><<<grid_for(12 * ne), 256, 0, s()>>>(ne, d_xyz.p, d_conn.p, d_young.p,
                                                             d_nu.p, d_ke.p);
    else if (stiffness_lds)
      mlk::k_stiffness_lds<10><<<grid_for(ne, 16), 256, 0, s()>>>(ne, d_xyz.p, d_conn.p, d_young.p,
                                                                 d_nu.p, d_ke.p);
    else
      mlk::k_stiffness<10><<<grid_for(30 * ne), 256, 0, s()>>>(ne, d_xyz.p, d_conn.p, d_young.p,
                                                              d_nu.p, d_ke.p);
    HIP_CHECK(hipGetLastError());
    op.prof.end(P_STIFF, ev, s());
  }

  // Tet10: time both stiffness kernels (three runs each after a warm-up), keep the faster
  void tune_stiffness() {
    if (plan.nen != 10) return;
    hipEvent_t a, b;
    HIP_CHECK(hipEventCreate(&a));
    HIP_CHECK(hipEventCreate(&b));
    float best = 1e30f;
    bool pick = true;
    for (bool lds : {true, false}) {
      stiffness_lds = lds;
      stiffness();
      HIP_CHECK(hipEventRecord(a, s()));
      for (int k = 0; k < 3; ++k) stiffness();
      HIP_CHECK(hipEventRecord(b, s()));
      HIP_CHECK(hipEventSynchronize(b));
      float t = 0.0f;
      HIP_CHECK(hipEventElapsedTime(&t, a, b));
      if (t < best) { best = t; pick = lds; }
    }
    stiffness_lds = pick;
    stiffness();  // d_ke holds the chosen kernel's (identical) element matrices
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
  }

  // model.rs:1480-1492 on the device: Ke of every element, gathered into K_ff and K_fc u_c.
  void assemble() {
    stiffness();
    if (all_fixed()) {
      assembled = true;
      return;
    }
    hipEvent_t ev = op.prof.begin(s());
    mlk::k_assemble<<<grid_for(op.nnz), 256, 0, s()>>>(op.nnz, d_asm_off.p, d_asm_slot.p, d_ke.p,
                                                       op.d_a.p);
    HIP_CHECK(hipGetLastError());
    op.prof.end(P_ASM, ev, s());
    if (!plan.cpl_slot.empty()) {
      mlk::k_coupling<<<grid_for(plan.nfree), 256, 0, s()>>>(plan.nfree, d_cpl_off.p,
                                                             d_cpl_slot.p, d_cpl_dof.p, d_ke.p,
                                                             d_imposed.p, d_coupling.p);
      HIP_CHECK(hipGetLastError());
    }
    op.values_changed();
    assembled = true;
  }

  int solve_device(const double* f, double* u, double* R, ml_solve_info* info) {
    if (info) {
      std::memset(info, 0, sizeof(*info));
      info->shift_index = -1;  // no factorisation (every DOF imposed) until one runs
      info->bad_element = -1;
    }
    double t0 = now_s();
    const int retries0 = op.retries;
    if (all_fixed()) {
      mlk::k_scatter_u<<<grid_for(plan.ndof), 256, 0, s()>>>(plan.ndof, d_rank.p, d_x.p,
                                                             d_imposed.p, u);
      HIP_CHECK(hipGetLastError());
      if (R) reactions(u, f, R);
      if (u != d_ulast.p)
        HIP_CHECK(hipMemcpyAsync(d_ulast.p, u, sizeof(double) * plan.ndof, hipMemcpyDeviceToDevice, s()));
      HIP_CHECK(hipStreamSynchronize(s()));
      have_u = true;
      return ML_OK;
    }
    if (!op.factored) {
      int st = op.factor();
      if (info) info->shift_index = op.shift_index;
      if (st != ML_OK) return st;
    }
    const int factor_retries = op.retries - retries0;
    double t1 = now_s();
    mlk::k_rhs<<<grid_for(plan.ndof), 256, 0, s()>>>(
        plan.ndof, d_rank.p, f, plan.cpl_slot.empty() ? nullptr : d_coupling.p, d_rhs.p);
    HIP_CHECK(hipGetLastError());
    int st = op.solve(d_rhs.p, d_x.p, 0, 0, info);
    double t2 = now_s();
    if (st != ML_OK && st != ML_NOT_CONVERGED) return st;
    mlk::k_scatter_u<<<grid_for(plan.ndof), 256, 0, s()>>>(plan.ndof, d_rank.p, d_x.p,
                                                           d_imposed.p, u);
    HIP_CHECK(hipGetLastError());
    if (R) reactions(u, f, R);
    if (u != d_ulast.p)
      HIP_CHECK(hipMemcpyAsync(d_ulast.p, u, sizeof(double) * plan.ndof, hipMemcpyDeviceToDevice, s()));
    HIP_CHECK(hipStreamSynchronize(s()));
    have_u = true;
    if (info) {
      info->device_retries += factor_retries;
      info->seconds[1] = t1 - t0;
      info->seconds[2] = t2 - t1;
      info->seconds[3] = now_s() - t2;
    }
    return st;
  }

  void reactions(const double* u, const double* f, double* R) {
    mlk::k_fill_f64<<<grid_for(plan.ndof), 256, 0, s()>>>(R, plan.ndof, 0.0);
    int64_t nfixed = (int64_t)plan.rea_dof.size();
    if (nfixed) {
      if (plan.nen == 4)
        mlk::k_reactions<4><<<grid_for(nfixed), 256, 0, s()>>>(
            nfixed, d_rea_dof.p, d_rea_off.p, d_rea_item.p, d_ke.p, d_conn.p, plan.n_elements, u,
            f, R);
      else
        mlk::k_reactions<10><<<grid_for(nfixed), 256, 0, s()>>>(
            nfixed, d_rea_dof.p, d_rea_off.p, d_rea_item.p, d_ke.p, d_conn.p, plan.n_elements, u,
            f, R);
    }
    HIP_CHECK(hipGetLastError());
  }

  // dC/dE_e of C = f.u at the last solution (homogeneous essential conditions only)
  void compliance_gradient(double* out_dev) {
    const int64_t ne = plan.n_elements;
    if (plan.nen == 4)
      mlk::k_compliance_gradient<4><<<grid_for(ne), 256, 0, s()>>>(ne, d_conn.p, d_ke.p, d_young.p,
                                                                  d_ulast.p, out_dev);
    else
      mlk::k_compliance_gradient<10><<<grid_for(ne), 256, 0, s()>>>(ne, d_conn.p, d_ke.p, d_young.p,
                                                                   d_ulast.p, out_dev);
    HIP_CHECK(hipGetLastError());
  }

  // Gauss-point stresses (gp_dev: n_elements x ngp x 9) averaged at the nodes (n_nodes x 9).
  // node -> (element, local node) incidences in element order (t = e*nen + a)
  void ensure_incidences() {
    const int64_t nn = mesh.n_nodes, ne = plan.n_elements;
    const int nen = plan.nen;
    if (!have_inc) {
      if (ne * nen > INT32_MAX) throw ArgError("too many element nodes for the incidence list");
      std::vector<int64_t> off(nn + 1, 0);
      for (int64_t t = 0; t < ne * nen; ++t) off[conn[t] + 1] += 1;
      for (int64_t v = 0; v < nn; ++v) off[v + 1] += off[v];
      std::vector<int> item(ne * nen);
      std::vector<int64_t> at(off.begin(), off.end() - 1);
      for (int64_t t = 0; t < ne * nen; ++t) item[at[conn[t]]++] = (int)t;  // element order
      d_inc_off.upload_vec(off, s());
      d_inc_item.upload_vec(item, s());
      HIP_CHECK(hipStreamSynchronize(s()));
      have_inc = true;
    }
  }

  void nodal_stresses(const double* gp_dev, double* out_dev) {
    const int64_t nn = mesh.n_nodes;
    const int nen = plan.nen;
    ensure_incidences();
    hipEvent_t ev = op.prof.begin(s());
    if (nen == 4)
      mlk::k_nodal_stresses<4><<<grid_for(nn), 256, 0, s()>>>(nn, d_xyz.p, d_inc_off.p,
                                                                d_inc_item.p, gp_dev, out_dev);
    else
      mlk::k_nodal_stresses<10><<<grid_for(nn), 256, 0, s()>>>(nn, d_xyz.p, d_inc_off.p,
                                                                 d_inc_item.p, gp_dev, out_dev);
    op.prof.end(P_NODAL, ev, s());
    HIP_CHECK(hipGetLastError());
  }

  void stresses(const double* u_dev, double* out_dev) {
    const int64_t ne = plan.n_elements;
    hipEvent_t ev = op.prof.begin(s());
    if (plan.nen == 4)
      mlk::k_stresses<4><<<grid_for(ne), 256, 0, s()>>>(ne, d_xyz.p, d_conn.p, d_young.p, d_nu.p,
                                                       u_dev, out_dev);
    else
      mlk::k_stresses<10><<<grid_for(4 * ne), 256, 0, s()>>>(ne, d_xyz.p, d_conn.p, d_young.p,
                                                            d_nu.p, u_dev, out_dev);
    op.prof.end(P_STRESS, ev, s());
    HIP_CHECK(hipGetLastError());
  }
};

// follower_pressure (assembly.rs:732-862) for a tetrahedron face, on the host: positions of the
// face's nodes in the element, forces[3m] per unit pressure factor and stiffness[3m x 3m]
// (row-major) of the pressure `value` at the displacements u. The face's outward normal comes
// from face_rule on the reference coordinates (assembly.rs:555-584); n dS = x,xi x x,eta on the
// current ones, with TRIANGLE_3 / TRIANGLE_6 weights and tri3 / tri6 derivatives (shape.rs:1-47).
int follower_terms(int nen, const int64_t* en, const double* xyz, int face, double value,
                   const double* u, int* pos, double* forces, double* stiffness) {
  static const int FACES[4][3] = {{0, 1, 2}, {0, 3, 1}, {1, 3, 2}, {2, 3, 0}};
  static const int EDGES[6][2] = {{0, 1}, {1, 2}, {2, 0}, {0, 3}, {1, 3}, {2, 3}};
  auto mid = [](int a, int b) {
    for (int k = 0; k < 6; ++k)
      if ((EDGES[k][0] == a && EDGES[k][1] == b) || (EDGES[k][0] == b && EDGES[k][1] == a))
        return 4 + k;
    return 4;
  };
  const int* fc = FACES[face];
  pos[0] = fc[0]; pos[1] = fc[1]; pos[2] = fc[2];
  int m = 3;
  if (nen == 10) {
    pos[3] = mid(fc[0], fc[1]); pos[4] = mid(fc[1], fc[2]); pos[5] = mid(fc[2], fc[0]);
    m = 6;
  }
  double ref[6][3], cur[6][3];
  for (int a = 0; a < m; ++a)
    for (int c = 0; c < 3; ++c) {
      ref[a][c] = xyz[3 * en[pos[a]] + c];
      cur[a][c] = ref[a][c] + u[3 * en[pos[a]] + c];
    }
  double e1[3], e2[3];
  for (int k = 0; k < 3; ++k) { e1[k] = ref[1][k] - ref[0][k]; e2[k] = ref[2][k] - ref[0][k]; }
  double cr[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2],
                  e1[0] * e2[1] - e1[1] * e2[0]};
  double sq = 0.0;
  for (int k = 0; k < 3; ++k) sq += cr[k] * cr[k];
  const double twice_area = std::sqrt(sq);
  double outward[3];
  for (int k = 0; k < 3; ++k)
    outward[k] = (ref[0][k] + ref[1][k] + ref[2][k]) / 3.0 -
                 (xyz[3 * en[0] + k] + xyz[3 * en[1] + k] + xyz[3 * en[2] + k] + xyz[3 * en[3] + k]) / 4.0;
  double dn0 = 0.0;
  for (int k = 0; k < 3; ++k) dn0 += cr[k] * outward[k];
  const double nsign = dn0 < 0.0 ? -1.0 : 1.0;
  double normal[3];
  for (int k = 0; k < 3; ++k) normal[k] = nsign * cr[k] / twice_area;
  const double A6 = 0.445948490915965, WA = 0.223381589678011 / 2.0;
  const double B6 = 0.091576213509771, WB = 0.109951743655322 / 2.0;
  const double tri3[3][3] = {{1.0 / 6.0, 1.0 / 6.0, 1.0 / 6.0},
                             {2.0 / 3.0, 1.0 / 6.0, 1.0 / 6.0},
                             {1.0 / 6.0, 2.0 / 3.0, 1.0 / 6.0}};
  const double tri6[6][3] = {{A6, A6, WA}, {1.0 - 2.0 * A6, A6, WA}, {A6, 1.0 - 2.0 * A6, WA},
                             {B6, B6, WB}, {1.0 - 2.0 * B6, B6, WB}, {B6, 1.0 - 2.0 * B6, WB}};
  const int ng = m == 3 ? 3 : 6;
  double nv[6][6], dn[6][6][2], wq[6];
  for (int q = 0; q < ng; ++q) {
    const double* gp = m == 3 ? tri3[q] : tri6[q];
    const double xi = gp[0], eta = gp[1];
    wq[q] = gp[2];
    if (m == 3) {
      nv[q][0] = 1.0 - xi - eta; nv[q][1] = xi; nv[q][2] = eta;
      dn[q][0][0] = -1.0; dn[q][0][1] = -1.0;
      dn[q][1][0] = 1.0; dn[q][1][1] = 0.0;
      dn[q][2][0] = 0.0; dn[q][2][1] = 1.0;
    } else {
      const double l1 = 1.0 - xi - eta, l2 = xi, l3 = eta;
      nv[q][0] = l1 * (2.0 * l1 - 1.0);
      nv[q][1] = l2 * (2.0 * l2 - 1.0);
      nv[q][2] = l3 * (2.0 * l3 - 1.0);
      nv[q][3] = 4.0 * l1 * l2;
      nv[q][4] = 4.0 * l2 * l3;
      nv[q][5] = 4.0 * l3 * l1;
      dn[q][0][0] = 1.0 - 4.0 * l1; dn[q][0][1] = 1.0 - 4.0 * l1;
      dn[q][1][0] = 4.0 * l2 - 1.0; dn[q][1][1] = 0.0;
      dn[q][2][0] = 0.0; dn[q][2][1] = 4.0 * l3 - 1.0;
      dn[q][3][0] = 4.0 * (l1 - l2); dn[q][3][1] = -4.0 * l2;
      dn[q][4][0] = 4.0 * l3; dn[q][4][1] = 4.0 * l2;
      dn[q][5][0] = -4.0 * l3; dn[q][5][1] = 4.0 * (l1 - l3);
    }
  }
  auto tangents = [&](const double (*xx)[3], int q, double* a, double* b) {
    for (int c = 0; c < 3; ++c) {
      double s0 = 0.0, s1 = 0.0;
      for (int k = 0; k < m; ++k) s0 += dn[q][k][0] * xx[k][c];
      for (int k = 0; k < m; ++k) s1 += dn[q][k][1] * xx[k][c];
      a[c] = s0;
      b[c] = s1;
    }
  };
  auto area_vector = [](const double* a, const double* b, double* o) {
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
  };
  auto levi = [](int i, int j, int k) {
    const double a = i, b = j, c = k;
    return (a - b) * (b - c) * (c - a) / 2.0;
  };
  double ta[3], tb[3], av[3];
  tangents(ref, 0, ta, tb);
  area_vector(ta, tb, av);
  double along = 0.0;
  for (int c = 0; c < 3; ++c) along += av[c] * normal[c];
  const double sign = along < 0.0 ? -1.0 : 1.0;
  std::fill(forces, forces + 3 * m, 0.0);
  std::fill(stiffness, stiffness + 9 * m * m, 0.0);
  for (int q = 0; q < ng; ++q) {
    double xi[3], eta[3], area[3];
    tangents(cur, q, xi, eta);
    area_vector(xi, eta, area);
    const double w = -value * sign * wq[q];
    for (int a = 0; a < m; ++a)
      for (int c = 0; c < 3; ++c) forces[3 * a + c] += w * nv[q][a] * area[c];
    for (int b = 0; b < m; ++b)
      for (int c = 0; c < 3; ++c)
        for (int k = 0; k < 3; ++k) {
          double d = 0.0;
          for (int l = 0; l < 3; ++l)
            d += levi(c, k, l) * dn[q][b][0] * eta[l] + levi(c, l, k) * xi[l] * dn[q][b][1];
          for (int a = 0; a < m; ++a) stiffness[(3 * a + c) * 3 * m + 3 * b + k] += w * nv[q][a] * d;
        }
  }
  return m;
}

template <class T>
void swap_bufs(DBuf<T>& a, DBuf<T>& b) {
  std::swap(a.p, b.p);
  std::swap(a.n, b.n);
}

// Model::solve_quasi_static (model.rs:715-797) for small strains on tetrahedra without contact,
// follower pressures or temperatures: increments of a Step, each brought to equilibrium by
// Newton-Raphson (equilibrium, model.rs:923-1022) and cut in two halves, at most MAX_CUTS deep,
// when Newton fails (advance, model.rs:800-857). Each Newton iteration runs on the device: the
// element responses (k_response: return mapping, consistent tangent, internal forces), the
// residual and its norms, the gather assembly of the tangents into the model's pattern, IC(0)
// with the shift ladder and the PCG solve. The pattern, RCM numbering and assembly plan are the
// model's: FreeSystem::new rebuilds the same ones every iteration (same element DOFs, same fixed
// set), so only the values change.
struct QuasiStatic {
  static constexpr double NEWTON_TOLERANCE = 1e-8;  // model.rs:1220
  static constexpr int NEWTON_ITERATIONS = 25;       // model.rs:1222
  static constexpr int MAX_CUTS = 5;                 // model.rs:1224

  Model& m;
  std::vector<double> amp_t, amp_v;
  int ngp;
  int64_t ne, ndof, nn;
  unsigned nblk;
  DBuf<double> d_yield, d_hard, d_committed, d_trial, d_fint_e, d_fint, d_negr, d_react,
      d_records, d_du, d_u, d_ut, d_f, d_partial, d_sums;
  DBuf<int> d_flag;
  // Kinematics::Large: per-element law (1 Neo-Hookean, 0 Saint-Venant-Kirchhoff) and follower
  // pressures (element, face, value), whose forces and stiffness follow u (follower_force,
  // model.rs:868-903)
  bool large = false;
  DBuf<unsigned char> d_neo;
  std::vector<int64_t> pres_el, pres_face;
  std::vector<double> pres_val;
  DBuf<double> d_follow, d_gval;
  DBuf<int64_t> d_goff, d_gslot;
  std::vector<double> h_u;
  double scale = 0.0;
  std::vector<double> residuals;
  int64_t newton = 0, cuts = 0, pcg_iterations = 0;
  int device_retries = 0;
  double t_resp = 0, t_factor = 0, t_solve = 0;

  QuasiStatic(Model& model, const double* yield0, const double* hardening, const double* f)
      : m(model) {
    ne = m.plan.n_elements;
    ndof = m.plan.ndof;
    nn = m.mesh.n_nodes;
    ngp = m.plan.nen == 4 ? 1 : 4;
    nblk = grid_for(ndof);
    hipStream_t st = m.s();
    d_yield.upload(yield0, ne, st);
    d_hard.upload(hardening, ne, st);
    d_f.upload(f, ndof, st);
    d_committed.alloc((size_t)7 * ngp * ne);
    d_trial.alloc((size_t)7 * ngp * ne);
    d_fint_e.alloc((size_t)3 * m.plan.nen * ne);
    d_records.alloc((size_t)9 * ngp * ne);
    for (auto* b : {&d_fint, &d_negr, &d_react, &d_du, &d_u, &d_ut}) b->alloc(ndof);
    d_partial.alloc((size_t)3 * nblk);
    d_sums.alloc(3);
    d_flag.alloc(1);
    mlk::k_fill_f64<<<grid_for(d_committed.n), 256, 0, st>>>(d_committed.p, d_committed.n, 0.0);
    mlk::k_fill_f64<<<grid_for(ndof), 256, 0, st>>>(d_u.p, ndof, 0.0);
    HIP_CHECK(hipGetLastError());
    m.ensure_incidences();
    // the model's K values are replaced by tangents from here on
    m.assembled = false;
    m.op.factored = false;
    m.have_u = false;
  }

  // Amplitude::at (model.rs:236-246)
  double amplitude(double t) const {
    size_t after = 0;
    while (after < amp_t.size() && amp_t[after] <= t) ++after;
    if (after >= 1 && after < amp_t.size())
      return amp_v[after - 1] +
             (amp_v[after] - amp_v[after - 1]) * (t - amp_t[after - 1]) / (amp_t[after] - amp_t[after - 1]);
    if (after >= 1) return amp_v[after - 1];
    return amp_v.empty() ? 0.0 : amp_v[0];
  }

  // responses (model.rs:1058-1067) at u from the committed states, then internal_force
  // (model.rs:1241-1250): tangents into the model's Ke array, trial states, Gauss-point records.
  void responses(const double* u) {
    hipStream_t st = m.s();
    mlk::k_fill_i32<<<1, 256, 0, st>>>(d_flag.p, 1, 0);
    hipEvent_t ev = m.op.prof.begin(st);
    if (large) {
      if (m.plan.nen == 4)
        mlk::k_large_response<4><<<grid_for(12 * ne), 256, 0, st>>>(
            ne, m.d_xyz.p, m.d_conn.p, m.d_young.p, m.d_nu.p, d_neo.p, u, m.d_ke.p, d_fint_e.p,
            d_trial.p, d_records.p);
      else
        mlk::k_large_response<10><<<grid_for(30 * ne), 256, 0, st>>>(
            ne, m.d_xyz.p, m.d_conn.p, m.d_young.p, m.d_nu.p, d_neo.p, u, m.d_ke.p, d_fint_e.p,
            d_trial.p, d_records.p);
    } else if (m.plan.nen == 4)
      mlk::k_response<4><<<grid_for(12 * ne), 256, 0, st>>>(
          ne, m.d_xyz.p, m.d_conn.p, m.d_young.p, m.d_nu.p, d_yield.p, d_hard.p, u,
          d_committed.p, m.d_ke.p, d_fint_e.p, d_trial.p, d_records.p, d_flag.p);
    else
      mlk::k_response<10><<<grid_for(30 * ne), 256, 0, st>>>(
          ne, m.d_xyz.p, m.d_conn.p, m.d_young.p, m.d_nu.p, d_yield.p, d_hard.p, u,
          d_committed.p, m.d_ke.p, d_fint_e.p, d_trial.p, d_records.p, d_flag.p);
    HIP_CHECK(hipGetLastError());
    m.op.prof.end(P_STIFF, ev, st);
    if (m.plan.nen == 4)
      mlk::k_internal_force<4><<<grid_for(nn), 256, 0, st>>>(nn, ne, m.d_inc_off.p,
                                                            m.d_inc_item.p, d_fint_e.p, d_fint.p);
    else
      mlk::k_internal_force<10><<<grid_for(nn), 256, 0, st>>>(nn, ne, m.d_inc_off.p,
                                                             m.d_inc_item.p, d_fint_e.p, d_fint.p);
    HIP_CHECK(hipGetLastError());
  }

  // follower_force (model.rs:868-903) at u: the follower forces per unit lambda into d_follow
  // (pressure after pressure, as the reference adds them) and -lambda sym(K_p) added into each
  // element's tangent, its pressures in order. The terms are computed on the host from u.
  void follower(const double* u, double lambda) {
    hipStream_t st = m.s();
    const int nen = m.plan.nen, mm = 3 * nen;
    const int64_t ne = m.plan.n_elements;
    h_u.resize(ndof);
    HIP_CHECK(hipMemcpyAsync(h_u.data(), u, sizeof(double) * ndof, hipMemcpyDeviceToHost, st));
    HIP_CHECK(hipStreamSynchronize(st));
    std::vector<double> fol(ndof, 0.0);
    // per element (in element order), its pressures' tangent corrections in pressure order
    std::vector<std::vector<int64_t>> by_el_idx;
    std::vector<int64_t> els;
    {
      std::vector<int64_t> order(pres_el.size());
      for (size_t k = 0; k < order.size(); ++k) order[k] = (int64_t)k;
      std::stable_sort(order.begin(), order.end(),
                       [&](int64_t a, int64_t b) { return pres_el[a] < pres_el[b]; });
      for (int64_t k : order) {
        if (els.empty() || els.back() != pres_el[k]) {
          els.push_back(pres_el[k]);
          by_el_idx.emplace_back();
        }
        by_el_idx.back().push_back(k);
      }
    }
    std::vector<int> npos(pres_el.size());
    std::vector<std::array<int, 6>> pos(pres_el.size());
    std::vector<std::array<double, 18>> forces(pres_el.size());
    std::vector<std::array<double, 324>> stiff(pres_el.size());
    for (size_t k = 0; k < pres_el.size(); ++k)
      npos[k] = follower_terms(nen, m.conn.data() + pres_el[k] * nen, m.nodes.data(),
                               (int)pres_face[k], pres_val[k], h_u.data(), pos[k].data(),
                               forces[k].data(), stiff[k].data());
    for (size_t k = 0; k < pres_el.size(); ++k) {
      const int64_t* en = m.conn.data() + pres_el[k] * nen;
      for (int a = 0; a < npos[k]; ++a)
        for (int c = 0; c < 3; ++c) fol[3 * en[pos[k][a]] + c] += forces[k][3 * a + c];
    }
    std::vector<int64_t> goff(1, 0), gslot;
    std::vector<double> gval;
    for (size_t gi = 0; gi < els.size(); ++gi) {
      const int64_t e = els[gi];
      for (int64_t k : by_el_idx[gi]) {
        const int n3 = 3 * npos[k];
        for (int r = 0; r < n3; ++r)
          for (int c = 0; c < n3; ++c) {
            const double sym = 0.5 * (stiff[k][r * n3 + c] + stiff[k][c * n3 + r]);
            const int row = 3 * pos[k][r / 3] + r % 3, col = 3 * pos[k][c / 3] + c % 3;
            gslot.push_back(((int64_t)row * mm + col) * ne + e);
            gval.push_back(-(lambda * sym));
          }
      }
      goff.push_back((int64_t)gslot.size());
    }
    d_follow.upload_vec(fol, st);
    d_goff.upload_vec(goff, st);
    d_gslot.upload_vec(gslot, st);
    d_gval.upload_vec(gval, st);
    mlk::k_add_groups<<<grid_for((int64_t)els.size()), 256, 0, st>>>((int64_t)els.size(), d_goff.p,
                                                                    d_gslot.p, d_gval.p, m.d_ke.p);
    HIP_CHECK(hipGetLastError());
  }

  // equilibrium (model.rs:923-1022) from u (device, updated in place)
  int equilibrium(double* u, double lambda, double& sc) {
    residuals.clear();
    hipStream_t st = m.s();
    for (;;) {
      double t0 = now_s();
      responses(u);
      const bool follows = large && !pres_el.empty();
      if (follows) follower(u, lambda);
      mlk::k_newton_residual<<<nblk, 256, 0, st>>>(ndof, lambda, d_f.p,
                                                  follows ? d_follow.p : nullptr, d_fint.p,
                                                  m.d_rank.p, d_negr.p, d_react.p, d_partial.p);
      mlk::k_sum3<<<1, 256, 0, st>>>(d_partial.p, (int)nblk, d_sums.p);
      HIP_CHECK(hipGetLastError());
      double sums[3];
      int yielding = 0;
      HIP_CHECK(hipMemcpyAsync(sums, d_sums.p, sizeof(sums), hipMemcpyDeviceToHost, st));
      HIP_CHECK(hipMemcpyAsync(&yielding, d_flag.p, sizeof(int), hipMemcpyDeviceToHost, st));
      HIP_CHECK(hipStreamSynchronize(st));
      t_resp += now_s() - t0;
      sc = std::fmax(std::fmax(sc, std::sqrt(sums[0])), std::sqrt(sums[1]));
      const double free_norm = std::sqrt(sums[2]);
      const double relative = sc > 0.0 ? free_norm / sc : free_norm;
      if (!std::isfinite(relative)) return ML_NEWTON_NOT_CONVERGED;
      residuals.push_back(relative);
      if (residuals.size() > 1 && relative <= NEWTON_TOLERANCE) return ML_OK;
      if ((int)residuals.size() > NEWTON_ITERATIONS) return ML_NEWTON_NOT_CONVERGED;
      // K_ff du_f = -r_f - K_fc du_c with du_c = lambda * BC value - u_c
      t0 = now_s();
      mlk::k_newton_du<<<grid_for(ndof), 256, 0, st>>>(ndof, lambda, m.d_rank.p, m.d_imposed.p,
                                                       u, d_du.p);
      if (m.all_fixed()) {  // no free DOF: the step is the imposed increment
        ++newton;
        mlk::k_newton_update<<<grid_for(ndof), 256, 0, st>>>(ndof, m.d_rank.p, m.d_x.p, d_du.p, u);
        HIP_CHECK(hipGetLastError());
        continue;
      }
      hipEvent_t ev = m.op.prof.begin(st);
      mlk::k_assemble<<<grid_for(m.op.nnz), 256, 0, st>>>(m.op.nnz, m.d_asm_off.p,
                                                          m.d_asm_slot.p, m.d_ke.p, m.op.d_a.p);
      HIP_CHECK(hipGetLastError());
      m.op.prof.end(P_ASM, ev, st);
      const bool coupled = !m.plan.cpl_slot.empty();
      if (coupled) {
        mlk::k_coupling<<<grid_for(m.plan.nfree), 256, 0, st>>>(
            m.plan.nfree, m.d_cpl_off.p, m.d_cpl_slot.p, m.d_cpl_dof.p, m.d_ke.p, d_du.p,
            m.d_coupling.p);
        HIP_CHECK(hipGetLastError());
      }
      m.op.values_changed();
      const int retries0 = m.op.retries;
      int s = m.op.factor();
      device_retries += m.op.retries - retries0;
      t_factor += now_s() - t0;
      if (s == ML_SINGULAR && yielding) return ML_PLASTIC_COLLAPSE;
      if (s != ML_OK) return s;
      t0 = now_s();
      mlk::k_rhs<<<grid_for(ndof), 256, 0, st>>>(ndof, m.d_rank.p, d_negr.p,
                                                 coupled ? m.d_coupling.p : nullptr, m.d_rhs.p);
      HIP_CHECK(hipGetLastError());
      ml_solve_info info;
      std::memset(&info, 0, sizeof(info));
      s = m.op.solve(m.d_rhs.p, m.d_x.p, 0, 0, &info);
      pcg_iterations += info.iterations;
      device_retries += info.device_retries;
      ++newton;
      t_solve += now_s() - t0;
      if (s != ML_OK) return s;
      mlk::k_newton_update<<<grid_for(ndof), 256, 0, st>>>(ndof, m.d_rank.p, m.d_x.p, d_du.p, u);
      HIP_CHECK(hipGetLastError());
    }
  }

  // advance (model.rs:800-857): the state changes only when an equilibrium converges
  int advance(double t0, double t1, int depth) {
    HIP_CHECK(hipMemcpyAsync(d_ut.p, d_u.p, sizeof(double) * ndof, hipMemcpyDeviceToDevice, m.s()));
    double sc = scale;
    int st = equilibrium(d_ut.p, amplitude(t1), sc);
    if (st == ML_OK) {
      swap_bufs(d_u, d_ut);
      swap_bufs(d_committed, d_trial);
      scale = sc;
      return ML_OK;
    }
    if ((st == ML_NEWTON_NOT_CONVERGED || st == ML_SINGULAR || st == ML_PLASTIC_COLLAPSE) &&
        depth < MAX_CUTS) {
      const double middle = (t0 + t1) / 2.0;
      ++cuts;
      st = advance(t0, middle, depth + 1);
      if (st != ML_OK) return st;
      return advance(middle, t1, depth + 1);
    }
    return st;
  }

  // increment k's field into the caller's host arrays (any may be null): u, reactions, the
  // Gauss-point records and p per Gauss point (layout (element, Gauss point))
  void record(int64_t k, double* u, double* R, double* records, double* p) {
    hipStream_t st = m.s();
    if (u) HIP_CHECK(hipMemcpyAsync(u + k * ndof, d_u.p, sizeof(double) * ndof, hipMemcpyDeviceToHost, st));
    if (R) {
      if (k == 0) std::fill(R, R + ndof, 0.0);
      else HIP_CHECK(hipMemcpyAsync(R + k * ndof, d_react.p, sizeof(double) * ndof, hipMemcpyDeviceToHost, st));
    }
    const size_t ns = (size_t)ngp * ne;
    if (records)
      HIP_CHECK(hipMemcpyAsync(records + k * ns * 9, d_records.p, sizeof(double) * ns * 9,
                               hipMemcpyDeviceToHost, st));
    std::vector<double> pq;
    if (p) {
      pq.resize(ns);
      HIP_CHECK(hipMemcpyAsync(pq.data(), d_committed.p + 6 * ns, sizeof(double) * ns,
                               hipMemcpyDeviceToHost, st));
    }
    HIP_CHECK(hipStreamSynchronize(st));
    if (p)
      for (int64_t e = 0; e < ne; ++e)
        for (int q = 0; q < ngp; ++q) p[k * ns + e * ngp + q] = pq[(size_t)q * ne + e];
  }
};

}  // namespace ml

// edge_load + FacePressure::add_load (assembly.rs:680-710, 638-644) for one tetrahedron edge on
// the host: f_a = q int N_a ds over the straight edge between its two vertices (TET10_EDGES,
// mid-node 4 + edge on a Tet10; mesh.rs:204-211), LINE_2 with line2 or LINE_3 with line3
// (quadrature.rs:62-90, shape.rs:115-125), terms q_c * w * length * N_a added point after point.
void edge_load_add(int nen, const int64_t* enodes, const double* xyz, int edge, const double* q,
                   double* f) {
  static const int EDGES[6][2] = {{0, 1}, {1, 2}, {2, 0}, {0, 3}, {1, 3}, {2, 3}};
  static const double S2[2] = {0.2113248654051871, 0.7886751345948129}, W2[2] = {0.5, 0.5};
  static const double S3[3] = {0.1127016653792583, 0.5, 0.8872983346207417};
  static const double W3[3] = {5.0 / 18.0, 4.0 / 9.0, 5.0 / 18.0};
  const int pos[3] = {EDGES[edge][0], EDGES[edge][1], 4 + edge};
  const int m = nen == 4 ? 2 : 3;
  const double* c0 = xyz + 3 * enodes[pos[0]];
  const double* c1 = xyz + 3 * enodes[pos[1]];
  double len2 = 0.0;
  for (int c = 0; c < 3; ++c) len2 += (c1[c] - c0[c]) * (c1[c] - c0[c]);
  const double length = std::sqrt(len2);
  double forces[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  for (int g = 0; g < m; ++g) {
    const double sv = m == 2 ? S2[g] : S3[g], w = m == 2 ? W2[g] : W3[g];
    double na[3];
    if (m == 2) {
      na[0] = 1.0 - sv;
      na[1] = sv;
    } else {
      na[0] = (1.0 - sv) * (1.0 - 2.0 * sv);
      na[1] = sv * (2.0 * sv - 1.0);
      na[2] = 4.0 * sv * (1.0 - sv);
    }
    for (int a = 0; a < m; ++a)
      for (int c = 0; c < 3; ++c) forces[3 * a + c] += q[c] * w * length * na[a];
  }
  for (int a = 0; a < m; ++a)
    for (int c = 0; c < 3; ++c) f[3 * enodes[pos[a]] + c] += forces[3 * a + c];
}

// face_rule + face_pressure + FacePressure::add_load (assembly.rs:527-593, 638-674) for one
// tetrahedron face on the host: f_a = -p sum_g w_g 2A N_a n, n outward (centroid test),
// TRIANGLE_3 on a Tet4 face, TRIANGLE_6 on a Tet10 face, added to f in face-node order.
void face_pressure_add(int nen, const int64_t* enodes, const double* xyz, int face, double value,
                       double* f) {
  static const int FACES[4][3] = {{0, 1, 2}, {0, 3, 1}, {1, 3, 2}, {2, 3, 0}};
  static const int EDGES[6][2] = {{0, 1}, {1, 2}, {2, 0}, {0, 3}, {1, 3}, {2, 3}};
  auto mid = [](int a, int b) {
    for (int k = 0; k < 6; ++k)
      if ((EDGES[k][0] == a && EDGES[k][1] == b) || (EDGES[k][0] == b && EDGES[k][1] == a))
        return 4 + k;
    return 4;
  };
  int pos[6];
  const int* fc = FACES[face];
  pos[0] = fc[0]; pos[1] = fc[1]; pos[2] = fc[2];
  int nn = 3;
  if (nen == 10) {
    pos[3] = mid(fc[0], fc[1]); pos[4] = mid(fc[1], fc[2]); pos[5] = mid(fc[2], fc[0]);
    nn = 6;
  }
  auto X = [&](int local, int k) { return xyz[3 * enodes[local] + k]; };
  double u[3], v[3];
  for (int k = 0; k < 3; ++k) {
    u[k] = X(pos[1], k) - X(pos[0], k);
    v[k] = X(pos[2], k) - X(pos[0], k);
  }
  double cross[3] = {u[1] * v[2] - u[2] * v[1], u[2] * v[0] - u[0] * v[2], u[0] * v[1] - u[1] * v[0]};
  double sq = 0.0;
  for (int k = 0; k < 3; ++k) sq += cross[k] * cross[k];
  const double twice_area = std::sqrt(sq);
  double outward[3];
  for (int k = 0; k < 3; ++k)
    outward[k] = (X(pos[0], k) + X(pos[1], k) + X(pos[2], k)) / 3.0 -
                 (X(0, k) + X(1, k) + X(2, k) + X(3, k)) / 4.0;
  double dot = 0.0;
  for (int k = 0; k < 3; ++k) dot += cross[k] * outward[k];
  const double sign = dot < 0.0 ? -1.0 : 1.0;
  double normal[3];
  for (int k = 0; k < 3; ++k) normal[k] = sign * cross[k] / twice_area;
  // quadrature.rs:19-54
  const double A6 = 0.445948490915965, WA = 0.223381589678011 / 2.0;
  const double B6 = 0.091576213509771, WB = 0.109951743655322 / 2.0;
  const double tri3[3][3] = {{1.0 / 6.0, 1.0 / 6.0, 1.0 / 6.0},
                             {2.0 / 3.0, 1.0 / 6.0, 1.0 / 6.0},
                             {1.0 / 6.0, 2.0 / 3.0, 1.0 / 6.0}};
  const double tri6[6][3] = {{A6, A6, WA}, {1.0 - 2.0 * A6, A6, WA}, {A6, 1.0 - 2.0 * A6, WA},
                             {B6, B6, WB}, {1.0 - 2.0 * B6, B6, WB}, {B6, 1.0 - 2.0 * B6, WB}};
  double forces[18] = {0};
  const int ng = nn == 3 ? 3 : 6;
  for (int g = 0; g < ng; ++g) {
    const double* gp = nn == 3 ? tri3[g] : tri6[g];
    const double w = gp[2] * twice_area;
    const double xi = gp[0], eta = gp[1];
    double nv[6];
    if (nn == 3) {
      nv[0] = 1.0 - xi - eta; nv[1] = xi; nv[2] = eta;
    } else {
      const double l1 = 1.0 - xi - eta, l2 = xi, l3 = eta;
      nv[0] = l1 * (2.0 * l1 - 1.0);
      nv[1] = l2 * (2.0 * l2 - 1.0);
      nv[2] = l3 * (2.0 * l3 - 1.0);
      nv[3] = 4.0 * l1 * l2;
      nv[4] = 4.0 * l2 * l3;
      nv[5] = 4.0 * l3 * l1;
    }
    for (int a = 0; a < nn; ++a)
      for (int c = 0; c < 3; ++c) forces[3 * a + c] -= value * w * nv[a] * normal[c];
  }
  for (int a = 0; a < nn; ++a)
    for (int c = 0; c < 3; ++c) f[3 * enodes[pos[a]] + c] += forces[3 * a + c];
}

// tet4 / tet10 values and Mapping3d's determinant at (xi, eta, zeta) on the host (shape.rs:95-190)
void tet_point_host(int nen, const double (*x)[3], double xi, double eta, double zeta, double* n,
                    double* det) {
  const double dl[4][3] = {{-1.0, -1.0, -1.0}, {1.0, 0.0, 0.0}, {0.0, 1.0, 0.0}, {0.0, 0.0, 1.0}};
  const double l[4] = {1.0 - xi - eta - zeta, xi, eta, zeta};
  const int ed[6][2] = {{0, 1}, {1, 2}, {2, 0}, {0, 3}, {1, 3}, {2, 3}};
  double dn[10][3];
  if (nen == 4) {
    for (int a = 0; a < 4; ++a) {
      n[a] = l[a];
      for (int c = 0; c < 3; ++c) dn[a][c] = dl[a][c];
    }
  } else {
    for (int a = 0; a < 4; ++a) {
      n[a] = l[a] * (2.0 * l[a] - 1.0);
      for (int c = 0; c < 3; ++c) dn[a][c] = (4.0 * l[a] - 1.0) * dl[a][c];
    }
    for (int k2 = 0; k2 < 6; ++k2) {
      n[4 + k2] = 4.0 * l[ed[k2][0]] * l[ed[k2][1]];
      for (int c = 0; c < 3; ++c)
        dn[4 + k2][c] = 4.0 * (l[ed[k2][0]] * dl[ed[k2][1]][c] + l[ed[k2][1]] * dl[ed[k2][0]][c]);
    }
  }
  double j[3][3] = {{0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}};
  for (int a = 0; a < nen; ++a)
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) j[r][c] += dn[a][r] * x[a][c];
  double d = 0.0;
  for (int c = 0; c < 3; ++c) {
    const int c1 = (c + 1) % 3, c2 = (c + 2) % 3;
    d += j[0][c] * (j[1][c1] * j[2][c2] - j[1][c2] * j[2][c1]);
  }
  *det = d;
}

// face_convection (assembly.rs:598-624) on the face rule of assembly.rs:527-593 (TRIANGLE_3 on a
// Tet4 face, TRIANGLE_6 on a Tet10 face, weights w 2A), on the host: the face's node positions
// in the element, matrix[a][b] = sum film w N_a N_b and load[a] = sum film sink w N_a.
int face_convection_terms(int nen, const int64_t* enodes, const double* xyz, int face,
                          double film, double sink, int* pos, double* matrix, double* load) {
  static const int FACES[4][3] = {{0, 1, 2}, {0, 3, 1}, {1, 3, 2}, {2, 3, 0}};
  static const int EDGES[6][2] = {{0, 1}, {1, 2}, {2, 0}, {0, 3}, {1, 3}, {2, 3}};
  auto mid = [](int a, int b) {
    for (int k = 0; k < 6; ++k)
      if ((EDGES[k][0] == a && EDGES[k][1] == b) || (EDGES[k][0] == b && EDGES[k][1] == a))
        return 4 + k;
    return 4;
  };
  const int* fc = FACES[face];
  pos[0] = fc[0]; pos[1] = fc[1]; pos[2] = fc[2];
  int nn = 3;
  if (nen == 10) {
    pos[3] = mid(fc[0], fc[1]); pos[4] = mid(fc[1], fc[2]); pos[5] = mid(fc[2], fc[0]);
    nn = 6;
  }
  auto X = [&](int local, int k) { return xyz[3 * enodes[local] + k]; };
  double u[3], v[3];
  for (int k = 0; k < 3; ++k) {
    u[k] = X(pos[1], k) - X(pos[0], k);
    v[k] = X(pos[2], k) - X(pos[0], k);
  }
  double cross[3] = {u[1] * v[2] - u[2] * v[1], u[2] * v[0] - u[0] * v[2], u[0] * v[1] - u[1] * v[0]};
  double sq = 0.0;
  for (int k = 0; k < 3; ++k) sq += cross[k] * cross[k];
  const double twice_area = std::sqrt(sq);
  const double A6 = 0.445948490915965, WA = 0.223381589678011 / 2.0;
  const double B6 = 0.091576213509771, WB = 0.109951743655322 / 2.0;
  const double tri3[3][3] = {{1.0 / 6.0, 1.0 / 6.0, 1.0 / 6.0},
                             {2.0 / 3.0, 1.0 / 6.0, 1.0 / 6.0},
                             {1.0 / 6.0, 2.0 / 3.0, 1.0 / 6.0}};
  const double tri6[6][3] = {{A6, A6, WA}, {1.0 - 2.0 * A6, A6, WA}, {A6, 1.0 - 2.0 * A6, WA},
                             {B6, B6, WB}, {1.0 - 2.0 * B6, B6, WB}, {B6, 1.0 - 2.0 * B6, WB}};
  std::fill(matrix, matrix + nn * nn, 0.0);
  std::fill(load, load + nn, 0.0);
  const int ng = nn == 3 ? 3 : 6;
  for (int g = 0; g < ng; ++g) {
    const double* gp = nn == 3 ? tri3[g] : tri6[g];
    const double w = gp[2] * twice_area;
    const double xi = gp[0], eta = gp[1];
    double nv[6];
    if (nn == 3) {
      nv[0] = 1.0 - xi - eta; nv[1] = xi; nv[2] = eta;
    } else {
      const double l1 = 1.0 - xi - eta, l2 = xi, l3 = eta;
      nv[0] = l1 * (2.0 * l1 - 1.0);
      nv[1] = l2 * (2.0 * l2 - 1.0);
      nv[2] = l3 * (2.0 * l3 - 1.0);
      nv[3] = 4.0 * l1 * l2;
      nv[4] = 4.0 * l2 * l3;
      nv[5] = 4.0 * l3 * l1;
    }
    for (int a = 0; a < nn; ++a) {
      load[a] += film * sink * w * nv[a];
      for (int b = 0; b < nn; ++b) matrix[a * nn + b] += film * w * nv[a] * nv[b];
    }
  }
  return nn;
}

namespace ml {
// Model::solve_step_with (model.rs:628-706) for a thermal model on tetrahedra: one DOF
// (temperature) per node, the theta method (C/dt + theta K) T_{n+1} = (C/dt - (1-theta) K) T_n + F
// with K = conduction + convection faces (thermal_matrix, model.rs:1320-1326) and the consistent
// capacity C. The element matrices, the left matrix's gather assembly and its IC(0) are built on
// the device once per (dt, theta); every step is a carried-term product, a PCG solve on the
// device and the step's reactions and heat fluxes.
struct Thermal {
  MeshView mesh;
  std::vector<double> nodes, k, rho_c, zeros, imposed;
  std::vector<int64_t> conn;
  std::vector<uint8_t> fixed;
  Plan plan;
  Operator op;
  DBuf<double> d_xyz, d_k, d_rhoc, d_kmat, d_cmat, d_left, d_right, d_imposed, d_coupling, d_f,
      d_t, d_carry, d_x, d_rhs, d_R, d_q, d_fval;
  DBuf<int> d_conn, d_rank, d_rea_dof, d_inc_item;
  DBuf<int64_t> d_asm_off, d_rea_off, d_rea_item, d_cpl_off, d_inc_off, d_foff, d_fslot;
  DBuf<uint32_t> d_asm_slot, d_cpl_slot;
  DBuf<int> d_cpl_dof;
  double dt = 0.0, theta = 0.0;
  bool factored = false;

  hipStream_t s() { return op.stream; }

  int create(int64_t n_nodes, const double* xyz, int64_t ne, int nen, const int64_t* cn,
             const double* kk, const double* rc, int64_t nconv, const int64_t* conv_el,
             const int64_t* conv_face, const double* film, const double* sink,
             const uint8_t* fx, const double* imp, int precision, ml_solve_info* info) {
    if (nen != 4 && nen != 10) throw ArgError("nen must be 4 (Tet4) or 10 (Tet10)");
    nodes.assign(xyz, xyz + 3 * n_nodes);
    conn.assign(cn, cn + ne * nen);
    k.assign(kk, kk + ne);
    rho_c.assign(rc, rc + ne);
    zeros.assign(ne, 0.0);
    fixed.assign(fx, fx + n_nodes);
    imposed.assign(n_nodes, 0.0);
    if (imp) imposed.assign(imp, imp + n_nodes);
    for (int64_t e = 0; e < ne; ++e)
      if (!(std::isfinite(rc[e]) && rc[e] > 0.0))
        throw ArgError("density and specific heat must be given (rho c > 0) for every element");
    // conductivity plays the part of E in the mesh checks (finite, > 0), with nu = 0
    mesh = MeshView{n_nodes, ne, nen, nodes.data(), conn.data(), k.data(), zeros.data(), 1};
    i64 where = -1;
    int merr = check_mesh(mesh, &where);
    if (info) { info->mesh_error = merr; info->bad_element = where; }
    if (merr) return ML_INVALID_MESH;
    std::string err;
    if (build_plan(mesh, fixed.data(), imposed.data(), plan, &err)) throw ArgError(err);
    if (plan.nfree == 0) throw ArgError("every temperature is fixed: nothing to solve for");
    if (info) info->seconds[5] = plan.seconds;
    op.precision = precision;
    op.set_pattern(plan.lower);
    hipStream_t st = s();
    d_xyz.upload_vec(nodes, st);
    std::vector<int> c32(conn.begin(), conn.end());
    d_conn.upload_vec(c32, st);
    d_k.upload_vec(k, st);
    d_rhoc.upload_vec(rho_c, st);
    d_imposed.upload_vec(imposed, st);
    d_rank.upload_vec(plan.rank_of_dof, st);
    d_asm_off.upload_vec(plan.asm_off, st);
    d_asm_slot.upload_vec(plan.asm_slot, st);
    d_cpl_off.upload_vec(plan.cpl_off, st);
    d_cpl_slot.upload_vec(plan.cpl_slot, st);
    d_cpl_dof.upload_vec(plan.cpl_dof, st);
    d_rea_dof.upload_vec(plan.rea_dof, st);
    d_rea_off.upload_vec(plan.rea_off, st);
    d_rea_item.upload_vec(plan.rea_item, st);
    // node -> (element, local node) incidences in element order
    std::vector<int64_t> off(n_nodes + 1, 0);
    for (int64_t t = 0; t < ne * nen; ++t) off[conn[t] + 1] += 1;
    for (int64_t v = 0; v < n_nodes; ++v) off[v + 1] += off[v];
    std::vector<int> item(ne * nen);
    std::vector<int64_t> at(off.begin(), off.end() - 1);
    for (int64_t t = 0; t < ne * nen; ++t) item[at[conn[t]]++] = (int)t;
    d_inc_off.upload_vec(off, st);
    d_inc_item.upload_vec(item, st);
    // convection faces grouped per element, in the given order: (K slot, value) additions
    std::vector<int64_t> foff(ne + 1, 0), fslot;
    std::vector<double> fval;
    {
      std::vector<std::vector<int64_t>> per(ne);
      for (int64_t c = 0; c < nconv; ++c) {
        if (conv_el[c] < 0 || conv_el[c] >= ne || conv_face[c] < 0 || conv_face[c] > 3)
          throw ArgError("convection face out of range");
        if (!(std::isfinite(film[c]) && std::isfinite(sink[c])))
          throw ArgError("convection film and sink must be finite");
        per[conv_el[c]].push_back(c);
      }
      int pos[6];
      double mat[36], load[6];
      for (int64_t e = 0; e < ne; ++e) {
        for (int64_t c : per[e]) {
          const int nn = face_convection_terms(nen, conn.data() + e * nen, nodes.data(),
                                               (int)conv_face[c], film[c], sink[c], pos, mat, load);
          for (int a = 0; a < nn; ++a)
            for (int b = 0; b < nn; ++b) {
              fslot.push_back(((int64_t)pos[a] * nen + pos[b]) * ne + e);
              fval.push_back(mat[a * nn + b]);
            }
        }
        foff[e + 1] = (int64_t)fslot.size();
      }
    }
    d_foff.upload_vec(foff, st);
    d_fslot.upload_vec(fslot, st);
    d_fval.upload_vec(fval, st);
    const size_t mm = (size_t)nen * nen * ne;
    for (auto* b : {&d_kmat, &d_cmat, &d_left, &d_right}) b->alloc(mm);
    d_coupling.alloc(plan.nfree);
    d_x.alloc(plan.nfree);
    d_rhs.alloc(plan.nfree);
    for (auto* b : {&d_f, &d_t, &d_carry, &d_R}) b->alloc(n_nodes);
    d_q.alloc((size_t)6 * (nen == 4 ? 1 : 4) * ne);
    if (nen == 4)
      mlk::k_thermal_matrices<4><<<grid_for(4 * ne), 256, 0, st>>>(ne, d_xyz.p, d_conn.p, d_k.p,
                                                                  d_rhoc.p, d_kmat.p, d_cmat.p);
    else
      mlk::k_thermal_matrices<10><<<grid_for(10 * ne), 256, 0, st>>>(ne, d_xyz.p, d_conn.p, d_k.p,
                                                                    d_rhoc.p, d_kmat.p, d_cmat.p);
    HIP_CHECK(hipGetLastError());
    if (!fslot.empty()) {
      mlk::k_add_faces<<<grid_for(ne), 256, 0, st>>>(ne, d_foff.p, d_fslot.p, d_fval.p, d_kmat.p);
      HIP_CHECK(hipGetLastError());
    }
    HIP_CHECK(hipStreamSynchronize(st));
    decltype(plan.asm_slot)().swap(plan.asm_slot);
    return ML_OK;
  }

  // FreeSystem::new on the left matrix (model.rs:664-671), once per (dt, theta)
  int prepare(double step_dt, double step_theta) {
    if (factored && step_dt == dt && step_theta == theta) return ML_OK;
    factored = false;
    hipStream_t st = s();
    const int64_t ne = plan.n_elements, mm = (int64_t)plan.nen * plan.nen * ne;
    mlk::k_theta_combine<<<grid_for(mm), 256, 0, st>>>(mm, step_dt, step_theta, step_theta - 1.0,
                                                       d_kmat.p, d_cmat.p, d_left.p, d_right.p);
    mlk::k_assemble<<<grid_for(op.nnz), 256, 0, st>>>(op.nnz, d_asm_off.p, d_asm_slot.p, d_left.p,
                                                      op.d_a.p);
    HIP_CHECK(hipGetLastError());
    if (!plan.cpl_slot.empty()) {
      mlk::k_coupling<<<grid_for(plan.nfree), 256, 0, st>>>(plan.nfree, d_cpl_off.p, d_cpl_slot.p,
                                                            d_cpl_dof.p, d_left.p, d_imposed.p,
                                                            d_coupling.p);
      HIP_CHECK(hipGetLastError());
    }
    op.values_changed();
    const int r = op.factor();
    if (r != ML_OK) return r;
    dt = step_dt;
    theta = step_theta;
    factored = true;
    return ML_OK;
  }

  void fluxes(double* out_dev) {
    const int64_t ne = plan.n_elements;
    if (plan.nen == 4)
      mlk::k_heat_flux<4><<<grid_for(ne), 256, 0, s()>>>(ne, d_xyz.p, d_conn.p, d_k.p, d_t.p, out_dev);
    else
      mlk::k_heat_flux<10><<<grid_for(4 * ne), 256, 0, s()>>>(ne, d_xyz.p, d_conn.p, d_k.p, d_t.p,
                                                              out_dev);
    HIP_CHECK(hipGetLastError());
  }

  // one step from d_t: carried + F, the solve, the new temperatures, reactions into d_R
  int step(ml_solve_info* info) {
    hipStream_t st = s();
    const int64_t nn = plan.n_nodes, ne = plan.n_elements;
    if (plan.nen == 4)
      mlk::k_thermal_carry<4><<<grid_for(nn), 256, 0, st>>>(nn, ne, d_inc_off.p, d_inc_item.p,
                                                           d_conn.p, d_right.p, d_t.p, d_f.p,
                                                           d_carry.p);
    else
      mlk::k_thermal_carry<10><<<grid_for(nn), 256, 0, st>>>(nn, ne, d_inc_off.p, d_inc_item.p,
                                                            d_conn.p, d_right.p, d_t.p, d_f.p,
                                                            d_carry.p);
    mlk::k_rhs<<<grid_for(nn), 256, 0, st>>>(nn, d_rank.p, d_carry.p,
                                             plan.cpl_slot.empty() ? nullptr : d_coupling.p,
                                             d_rhs.p);
    HIP_CHECK(hipGetLastError());
    const int r = op.solve(d_rhs.p, d_x.p, 0, 0, info);
    if (r != ML_OK) return r;
    mlk::k_scatter_u<<<grid_for(nn), 256, 0, st>>>(nn, d_rank.p, d_x.p, d_t.p, d_t.p);
    mlk::k_fill_f64<<<grid_for(nn), 256, 0, st>>>(d_R.p, nn, 0.0);
    const int64_t nfixed = (int64_t)plan.rea_dof.size();
    if (nfixed) {
      if (plan.nen == 4)
        mlk::k_thermal_reactions<4><<<grid_for(nfixed), 256, 0, st>>>(
            nfixed, d_rea_dof.p, d_rea_off.p, d_rea_item.p, d_conn.p, ne, d_left.p, d_t.p,
            d_carry.p, d_R.p);
      else
        mlk::k_thermal_reactions<10><<<grid_for(nfixed), 256, 0, st>>>(
            nfixed, d_rea_dof.p, d_rea_off.p, d_rea_item.p, d_conn.p, ne, d_left.p, d_t.p,
            d_carry.p, d_R.p);
    }
    HIP_CHECK(hipGetLastError());
    return ML_OK;
  }
};
}  // namespace ml

// ====================================================================== C ABI
using namespace ml;

struct ml_pcg {
  Operator op;
};
struct ml_model {
  Model m;
};
struct ml_msh {
  msh::File f;
};
// Several models solved together: their K_ff concatenated block-diagonally into one operator
// (one launch per kernel for all of them), each model a system of the batched PCG.
struct ml_batch {
  std::vector<ml_model*> models;
  Operator op;
  int nseg = 0;
  std::vector<int64_t> row_off;  // model k owns rows [row_off[k], row_off[k+1])
  std::vector<int> shift_idx;
  DBuf<int64_t> d_seg_off;
  DBuf<int> d_seg;
  DBuf<double> d_rhs, d_x;
};
struct ml_pattern {
  Graph graph;
  std::vector<int64_t> order;
  LowerPattern lower;
};

namespace {

template <class F>
ml_status guarded(F body) {
  try {
    return (ml_status)body();
  } catch (const ArgError& e) {
    g_last_error = e.what();
    return ML_INVALID_ARGUMENT;
  } catch (const HipError& e) {
    g_last_error = e.what();
    return ML_HIP_ERROR;
  } catch (const std::bad_alloc&) {
    g_last_error = "host allocation failed";
    return ML_HIP_ERROR;
  } catch (const std::exception& e) {
    g_last_error = e.what();
    return ML_HIP_ERROR;
  }
}

// Validates a lower-triangle CSR in the reference layout (ascending columns, diagonal last).
LowerPattern pattern_from_csr(int64_t n, const int64_t* off, const int64_t* cols) {
  if (n <= 0 || n >= INT32_MAX || !off || !cols) throw ArgError("bad CSR arguments");
  LowerPattern lp;
  lp.offsets.assign(off, off + n + 1);
  if (lp.offsets[0] != 0 || lp.offsets[n] >= INT32_MAX) throw ArgError("bad CSR offsets");
  lp.cols.resize(lp.offsets[n]);
  for (int64_t i = 0; i < n; ++i) {
    if (lp.offsets[i + 1] <= lp.offsets[i]) throw ArgError("empty CSR row");
    for (int64_t p = lp.offsets[i]; p < lp.offsets[i + 1]; ++p) {
      int64_t c = cols[p];
      bool last = p == lp.offsets[i + 1] - 1;
      if (last ? c != i : (c < 0 || c >= i || (p > lp.offsets[i] && c <= cols[p - 1])))
        throw ArgError("CSR is not lower-triangular with ascending columns and diagonal last");
      lp.cols[p] = (int32_t)c;
    }
  }
  return lp;
}

}  // namespace

extern "C" {

const char* ml_last_error(void) { return g_last_error.c_str(); }

int ml_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

ml_status ml_solid_d(double young, double nu, double* d) {
  if (!d) return ML_INVALID_ARGUMENT;
  // element.rs:193-206, host
  double c = young / ((1.0 + nu) * (1.0 - 2.0 * nu));
  double g = young / (2.0 * (1.0 + nu));
  for (int i = 0; i < 36; ++i) d[i] = 0.0;
  for (int i = 0; i < 3; ++i)
    for (int k = 0; k < 3; ++k) d[6 * i + k] = (i == k) ? c * (1.0 - nu) : c * nu;
  for (int i = 3; i < 6; ++i) d[6 * i + i] = g;
  return ML_OK;
}

ml_status ml_stiffness_batch(int nen, int64_t ne, const double* coords, const double* young,
                             const double* nu, double* ke) {
  return guarded([&]() -> int {
    if ((nen != 4 && nen != 10) || ne <= 0 || !coords || !young || !nu || !ke)
      throw ArgError("bad stiffness batch arguments");
    const int m = 3 * nen;
    if ((int64_t)m * m * ne >= (int64_t)INT32_MAX * 4) throw ArgError("batch too large");
    hipStream_t st;
    HIP_CHECK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    std::unique_ptr<void, void (*)(void*)> guard(st, [](void* p) { (void)hipStreamDestroy((hipStream_t)p); });
    DBuf<double> dx, dE, dv, dk;
    DBuf<int> dc;
    dx.upload(coords, (size_t)ne * nen * 3, st);
    std::vector<int> c(ne * nen);
    for (int64_t k = 0; k < ne * nen; ++k) c[k] = (int)k;
    dc.upload_vec(c, st);
    dE.upload(young, ne, st);
    dv.upload(nu, ne, st);
    dk.alloc((size_t)m * m * ne);
    if (nen == 4)
      mlk::k_stiffness<4><<<grid_for(12 * ne), 256, 0, st>>>(ne, dx.p, dc.p, dE.p, dv.p, dk.p);
    else
      mlk::k_stiffness_lds<10><<<grid_for(ne, 16), 256, 0, st>>>(ne, dx.p, dc.p, dE.p, dv.p, dk.p);
    HIP_CHECK(hipGetLastError());
    std::vector<double> h((size_t)m * m * ne);
    HIP_CHECK(hipMemcpyAsync(h.data(), dk.p, h.size() * sizeof(double), hipMemcpyDeviceToHost, st));
    HIP_CHECK(hipStreamSynchronize(st));
    for (int64_t e = 0; e < ne; ++e)
      for (int j = 0; j < m * m; ++j) ke[e * m * m + j] = h[(size_t)j * ne + e];
    return ML_OK;
  });
}

}  // extern "C"

namespace {

// Staging of ml_stiffness_batch_noalloc: a file-scope device array, one copy per GPU, held
// until the process ends. Host code never uses its name as an address: each device's address
// comes from hipGetSymbolAddress, taken once per device on its first call with that device
// current, and kept in g_scratch.
static __device__ __align__(256) unsigned char scratch[ML_STIFFNESS_SCRATCH_BYTES];

std::mutex g_scratch_mu;
void* g_scratch[64] = {};

void* scratch_address() {
  int dev = 0;
  HIP_CHECK(hipGetDevice(&dev));
  if (dev < 0 || dev >= 64) throw ArgError("device index out of range for the scratch table");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch[dev]) {
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      throw HipError(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch[dev] = p;
  }
  return g_scratch[dev];
}

// Bytes of scratch a batch of ne elements of nen nodes needs: coordinates, E, nu, the identity
// connectivity, the entry-major stiffnesses and their element-major copy, each region 256-byte
// aligned.
size_t noalloc_bytes(int nen, int64_t ne, size_t* off) {
  const size_t m = 3 * (size_t)nen;
  const size_t sizes[6] = {sizeof(double) * 3 * nen * (size_t)ne, sizeof(double) * (size_t)ne,
                           sizeof(double) * (size_t)ne, sizeof(int) * nen * (size_t)ne,
                           sizeof(double) * m * m * (size_t)ne, sizeof(double) * m * m * (size_t)ne};
  size_t at = 0;
  for (int k = 0; k < 6; ++k) {
    off[k] = at;
    at += (sizes[k] + 255) / 256 * 256;
  }
  return at;
}

}  // namespace

extern "C" {

ml_status ml_stiffness_batch_noalloc(int nen, int64_t ne, const double* coords,
                                     const double* young, const double* nu, double* ke) {
  return guarded([&]() -> int {
    if ((nen != 4 && nen != 10) || ne <= 0 || !coords || !young || !nu || !ke)
      throw ArgError("bad stiffness batch arguments");
    size_t off[6];
    if (ne > (int64_t)ML_STIFFNESS_SCRATCH_BYTES || noalloc_bytes(nen, ne, off) > ML_STIFFNESS_SCRATCH_BYTES) {
      size_t probe[6];
      int64_t most = 0;
      for (int64_t lo = 1, hi = ML_STIFFNESS_SCRATCH_BYTES; lo <= hi;) {
        const int64_t mid = lo + (hi - lo) / 2;
        if (noalloc_bytes(nen, mid, probe) <= ML_STIFFNESS_SCRATCH_BYTES) { most = mid; lo = mid + 1; }
        else hi = mid - 1;
      }
      throw ArgError("batch of " + std::to_string(ne) + " elements exceeds the " +
                     std::to_string(ML_STIFFNESS_SCRATCH_BYTES >> 20) + " MiB scratch: at most " +
                     std::to_string(most) + " " + (nen == 4 ? "Tet4" : "Tet10") +
                     " elements per call");
    }
    unsigned char* base = (unsigned char*)scratch_address();
    double* dx = (double*)(base + off[0]);
    double* dE = (double*)(base + off[1]);
    double* dv = (double*)(base + off[2]);
    int* dc = (int*)(base + off[3]);
    double* dk = (double*)(base + off[4]);
    double* dke = (double*)(base + off[5]);
    const int m = 3 * nen;
    hipStream_t st = nullptr;  // the null stream: nothing is created
    HIP_CHECK(hipMemcpyAsync(dx, coords, sizeof(double) * 3 * nen * ne, hipMemcpyHostToDevice, st));
    HIP_CHECK(hipMemcpyAsync(dE, young, sizeof(double) * ne, hipMemcpyHostToDevice, st));
    HIP_CHECK(hipMemcpyAsync(dv, nu, sizeof(double) * ne, hipMemcpyHostToDevice, st));
    mlk::k_iota_i32<<<grid_for(ne * nen), 256, 0, st>>>(dc, ne * nen);
    if (nen == 4)
      mlk::k_stiffness<4><<<grid_for(12 * ne), 256, 0, st>>>(ne, dx, dc, dE, dv, dk);
    else
      mlk::k_stiffness_lds<10><<<grid_for(ne, 16), 256, 0, st>>>(ne, dx, dc, dE, dv, dk);
    mlk::k_entry_to_element_major<<<grid_for(ne * m * m), 256, 0, st>>>(ne, m * m, dk, dke);
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipMemcpyAsync(ke, dke, sizeof(double) * m * m * ne, hipMemcpyDeviceToHost, st));
    HIP_CHECK(hipStreamSynchronize(st));
    return ML_OK;
  });
}

ml_status ml_quotient64(int64_t n, const double* num, const double* den, double* out) {
  return guarded([&]() -> int {
    if (n < 0 || (n > 0 && (!num || !den || !out))) throw ArgError("bad quotient arguments");
    for (int64_t i = 0; i < n; ++i) out[i] = mlk::quotient64(num[i], den[i], 1.0 / den[i]);
    return ML_OK;
  });
}

ml_status ml_pattern_build(int64_t n, int64_t ne, const int64_t* eoff, const int64_t* eit,
                           ml_pattern** out) {
  return guarded([&]() -> int {
    if (!out || n <= 0 || n >= INT32_MAX || ne < 0 || !eoff || (ne > 0 && !eit))
      throw ArgError("bad pattern arguments");
    for (int64_t e = 0; e < ne; ++e)
      for (int64_t p = eoff[e]; p < eoff[e + 1]; ++p)
        if (eit[p] < 0 || eit[p] >= n) throw ArgError("element vertex out of range");
    auto* h = new ml_pattern;
    h->graph = graph_from_elements(n, ne, eoff, eit);
    h->order = reverse_cuthill_mckee(h->graph);
    std::vector<int64_t> rank(n);
    for (int64_t r = 0; r < n; ++r) rank[h->order[r]] = r;
    h->lower = lower_pattern(h->graph, h->order, rank);
    *out = h;
    return ML_OK;
  });
}

int64_t ml_pattern_n(const ml_pattern* p) { return p ? p->lower.n() : 0; }
int64_t ml_pattern_nnz(const ml_pattern* p) { return p ? p->lower.nnz() : 0; }
int64_t ml_pattern_ntargets(const ml_pattern* p) { return p ? (int64_t)p->graph.targets.size() : 0; }

ml_status ml_pattern_get(const ml_pattern* p, int64_t* order, int64_t* offsets, int64_t* cols) {
  if (!p) return ML_INVALID_ARGUMENT;
  if (order) std::copy(p->order.begin(), p->order.end(), order);
  if (offsets) std::copy(p->lower.offsets.begin(), p->lower.offsets.end(), offsets);
  if (cols) std::copy(p->lower.cols.begin(), p->lower.cols.end(), cols);
  return ML_OK;
}

ml_status ml_pattern_graph(const ml_pattern* p, int64_t* offsets, int64_t* targets) {
  if (!p) return ML_INVALID_ARGUMENT;
  if (offsets) std::copy(p->graph.offsets.begin(), p->graph.offsets.end(), offsets);
  if (targets) std::copy(p->graph.targets.begin(), p->graph.targets.end(), targets);
  return ML_OK;
}

void ml_pattern_free(ml_pattern* p) { delete p; }

ml_status ml_layout_verify(int64_t n, const int64_t* offsets, const int64_t* cols,
                           const double* lvalues, const double* r, double* z, int64_t* stats) {
  return guarded([&]() -> int {
    LowerPattern lp = pattern_from_csr(n, offsets, cols);
    DeviceLayouts d;
    std::string err = device_layouts(lp, kTriCaps, &d);
    if (err.empty()) err = check_layouts(lp, d, kTriCaps);
    SweepLayout lo, up;
    if (err.empty()) err = sweep_layouts(lp, &lo, &up);
    if (err.empty()) err = check_sweep_layouts(lp, lo, up);
    for (int rows : {FUSED_LANES, 2 * FUSED_LANES}) {
      FusedSweepLayout flo, fup;
      if (err.empty() && fused_sweep_layouts(lp, rows, &flo, &fup).empty())
        err = check_fused_sweep_layouts(lp, flo, fup);
    }
    if (!err.empty()) throw ArgError(err);
    if (stats) {
      stats[0] = (int64_t)d.fw.bstart.size() - 1;
      stats[1] = (int64_t)d.bw.bstart.size() - 1;
      stats[2] = d.fw.max_block_ext;
      stats[3] = d.fw.max_block_int;
      stats[4] = d.bw.max_block_ext;
      stats[5] = d.bw.max_block_int;
      stats[6] = d.depth;
      stats[7] = d.max_row;
    }
    if (lvalues && r && z) layout_precondition(d, lvalues, r, z);
    return ML_OK;
  });
}

ml_status ml_layout_sweeps(int64_t n, const int64_t* offsets, const int64_t* cols, int rows,
                           const double* lvalues, const double* r, int sweeps, double* z,
                           int64_t* stats) {
  return guarded([&]() -> int {
    if (sweeps < 1 || sweeps > 64) throw ArgError("sweeps must be in 1..64");
    LowerPattern lp = pattern_from_csr(n, offsets, cols);
    FusedSweepLayout flo, fup;
    std::string err = fused_sweep_layouts(lp, rows, &flo, &fup);
    if (err.empty()) err = check_fused_sweep_layouts(lp, flo, fup);
    if (!err.empty()) throw ArgError(err);
    if (stats) {
      stats[0] = flo.blocks;
      stats[1] = std::max(flo.width, fup.width);
      stats[2] = flo.max_ext;
      stats[3] = fup.max_ext;
      // the fused variant a handle on this pattern starts with (0: one launch per sweep)
      bool have[2] = {false, false};
      for (int h = 0; h < 2; ++h) {
        FusedSweepLayout a, b;
        have[h] = fused_sweep_layouts(lp, FUSED_LANES * (h + 1), &a, &b).empty() &&
                  (h == 0 || std::max(a.width, b.width) <= 32);
      }
      const int v = default_fused_variant(have, (int)stats[1]);
      stats[4] = v < 0 ? 0 : fused_code(v);
    }
    if (lvalues && r && z) fused_sweeps_host(lp, flo, fup, lvalues, r, sweeps, z);
    return ML_OK;
  });
}

ml_status ml_pcg_create(int64_t n, const int64_t* offsets, const int64_t* cols,
                        const double* values, ml_precision precision, ml_pcg** out) {
  return guarded([&]() -> int {
    if (!out || !values) throw ArgError("null argument");
    if (precision < ML_FP64 || precision > ML_MIXED_JACOBI) throw ArgError("bad precision");
    LowerPattern lp = pattern_from_csr(n, offsets, cols);
    std::unique_ptr<ml_pcg> h(new ml_pcg);
    h->op.precision = precision;
    h->op.set_pattern(lp);
    h->op.d_a.upload(values, lp.nnz(), h->op.stream);
    h->op.values_changed();
    int st = h->op.factor();
    if (st != ML_OK) return st;
    *out = h.release();
    return ML_OK;
  });
}

ml_status ml_pcg_solve_device(ml_pcg* h, const double* b, double* x, double tol, int64_t maxit,
                              ml_solve_info* info) {
  return guarded([&]() -> int {
    if (!h || !b || !x) throw ArgError("null argument");
    double t0 = now_s();
    int st = h->op.solve(b, x, tol, maxit, info);
    if (info) info->seconds[2] = now_s() - t0;
    return st;
  });
}

namespace {

int status_of(const mlk::PcgScalars& q) {
  return q.status == 1 ? ML_SINGULAR : q.status == 2 ? ML_NOT_CONVERGED : ML_OK;
}

void info_of(const mlk::PcgScalars& q, int shift_index, ml_solve_info* info) {
  std::memset(info, 0, sizeof(*info));
  info->iterations = q.it;
  info->residual = q.bnorm > 0 ? q.rnorm / q.bnorm : 0.0;
  info->true_residual = info->residual;
  info->shift_index = shift_index;
  info->bad_element = -1;
}

}  // namespace

ml_status ml_pcg_solve(ml_pcg* h, const double* b, double* x, int64_t nrhs, double tol,
                       int64_t maxit, ml_solve_info* info) {
  return guarded([&]() -> int {
    if (!h || !b || !x || nrhs <= 0) throw ArgError("bad arguments");
    Operator& op = h->op;
    const int n = op.n;
    if (nrhs > 1) {
      // batched: up to MAX_BATCH_RHS right-hand sides per set of launches
      int result = ML_OK;
      const int retries0 = op.retries;
      mlk::PcgScalars worst;
      std::memset(&worst, 0, sizeof(worst));
      for (int64_t k0 = 0; k0 < nrhs; k0 += op.rhs_per_pass) {
        const int k = (int)std::min<int64_t>(op.rhs_per_pass, nrhs - k0);
        DBuf<double> bd, xd;
        bd.upload(b + k0 * n, (size_t)k * n, op.stream);
        xd.alloc((size_t)k * n);
        std::vector<mlk::PcgScalars> hS;
        int st = op.solve_many(k, bd.p, xd.p, tol, maxit, &hS);
        if (st != ML_OK) return st;
        HIP_CHECK(hipMemcpyAsync(x + k0 * n, xd.p, sizeof(double) * k * n, hipMemcpyDeviceToHost,
                                 op.stream));
        HIP_CHECK(hipStreamSynchronize(op.stream));
        for (const auto& q : hS) {
          if (status_of(q) != ML_OK && result == ML_OK) result = status_of(q);
          if (q.it >= worst.it) worst = q;
        }
      }
      if (info) {
        info_of(worst, op.shift_index, info);
        info->device_retries = op.retries - retries0;
      }
      return result;
    }
    ml_solve_info worst;
    std::memset(&worst, 0, sizeof(worst));
    int result = ML_OK;
    for (int64_t k = 0; k < nrhs; ++k) {
      HIP_CHECK(hipMemcpyAsync(op.w_b.p, b + k * n, sizeof(double) * n, hipMemcpyHostToDevice,
                               op.stream));
      ml_solve_info one;
      std::memset(&one, 0, sizeof(one));
      DBuf<double> bcopy;
      bcopy.alloc(n);
      HIP_CHECK(hipMemcpyAsync(bcopy.p, op.w_b.p, sizeof(double) * n, hipMemcpyDeviceToDevice,
                               op.stream));
      DBuf<double> xd;
      xd.alloc(n);
      int st = op.solve(bcopy.p, xd.p, tol, maxit, &one);
      HIP_CHECK(hipMemcpyAsync(x + k * n, xd.p, sizeof(double) * n, hipMemcpyDeviceToHost,
                               op.stream));
      HIP_CHECK(hipStreamSynchronize(op.stream));
      if (st != ML_OK && result == ML_OK) result = st;
      if (one.iterations >= worst.iterations) worst = one;
    }
    if (info) *info = worst;
    return result;
  });
}

ml_status ml_pcg_multiply(ml_pcg* h, const double* x, double* y) {
  return guarded([&]() -> int {
    if (!h || !x || !y) throw ArgError("null argument");
    Operator& op = h->op;
    HIP_CHECK(hipMemcpyAsync(op.w_p.p, x, sizeof(double) * op.n, hipMemcpyHostToDevice, op.stream));
    op.spmv<double>(op.w_p.p, op.w_ap.p, nullptr, nullptr);
    HIP_CHECK(hipMemcpyAsync(y, op.w_ap.p, sizeof(double) * op.n, hipMemcpyDeviceToHost, op.stream));
    HIP_CHECK(hipStreamSynchronize(op.stream));
    return ML_OK;
  });
}

ml_status ml_pcg_multiply_many(ml_pcg* h, int64_t nrhs, const double* x, double* y) {
  return guarded([&]() -> int {
    if (!h || !x || !y || nrhs < 1) throw ArgError("bad arguments");
    Operator& op = h->op;
    if ((int64_t)op.n * nrhs > INT32_MAX) throw ArgError("too many vectors at once");
    const int64_t N = (int64_t)op.n * nrhs;
    DBuf<double> xd, yd;
    xd.alloc(N);
    yd.alloc(N);
    HIP_CHECK(hipMemcpyAsync(xd.p, x, sizeof(double) * N, hipMemcpyHostToDevice, op.stream));
    op.spmv_many(xd.p, yd.p, (int)nrhs);
    HIP_CHECK(hipMemcpyAsync(y, yd.p, sizeof(double) * N, hipMemcpyDeviceToHost, op.stream));
    HIP_CHECK(hipStreamSynchronize(op.stream));
    return ML_OK;
  });
}

ml_status ml_pcg_precondition(ml_pcg* h, const double* r, double* z) {
  return guarded([&]() -> int {
    if (!h || !r || !z) throw ArgError("null argument");
    Operator& op = h->op;
    if (!op.factored) {
      int st = op.factor();
      if (st != ML_OK) return st;
    }
    HIP_CHECK(hipMemcpyAsync(op.w_r.p, r, sizeof(double) * op.n, hipMemcpyHostToDevice, op.stream));
    op.apply_precondition();
    HIP_CHECK(hipMemcpyAsync(z, op.w_z.p, sizeof(double) * op.n, hipMemcpyDeviceToHost, op.stream));
    HIP_CHECK(hipStreamSynchronize(op.stream));
    if (op.read_flag(0)) return ML_DEVICE_TIMEOUT;
    return ML_OK;
  });
}

ml_status ml_pcg_precondition_many(ml_pcg* h, int64_t nrhs, const double* r, double* z) {
  return guarded([&]() -> int {
    if (!h || !r || !z || nrhs < 1) throw ArgError("bad arguments");
    Operator& op = h->op;
    if ((int64_t)op.n * nrhs > INT32_MAX) throw ArgError("too many right-hand sides at once");
    if (!op.factored) {
      int st = op.factor();
      if (st != ML_OK) return st;
    }
    const int k = (int)nrhs;
    const int64_t N = (int64_t)op.n * k;
    DBuf<double> rd, yd, zd;
    rd.alloc(N);
    yd.alloc(N);
    zd.alloc(N);
    HIP_CHECK(hipMemcpyAsync(rd.p, r, sizeof(double) * N, hipMemcpyHostToDevice, op.stream));
    mlk::k_fill_i32<<<1, 64, 0, op.stream>>>(op.d_flags.p, 2, 0);
    if (op.sweeps > 0) {
      op.apply_sweeps_rep(rd.p, zd.p, k);
    } else {
      mlk::k_fill_bits<double><<<grid_for(N), 256, 0, op.stream>>>(yd.p, N, mlk::Bits<double>::S);
      mlk::k_fill_bits<double><<<grid_for(N), 256, 0, op.stream>>>(zd.p, N, mlk::Bits<double>::S);
      op.precondition(rd.p, yd.p, zd.p, nullptr, nullptr, nullptr, k);
    }
    HIP_CHECK(hipMemcpyAsync(z, zd.p, sizeof(double) * N, hipMemcpyDeviceToHost, op.stream));
    HIP_CHECK(hipStreamSynchronize(op.stream));
    if (op.read_flag(0)) return ML_DEVICE_TIMEOUT;
    return ML_OK;
  });
}

ml_status ml_pcg_factor(ml_pcg* h, double* l, int32_t* shift_index) {
  return guarded([&]() -> int {
    if (!h) throw ArgError("null handle");
    Operator& op = h->op;
    if (!op.factored) {
      int st = op.factor();
      if (st != ML_OK) return st;
    }
    if (l) {
      if (op.precision != ML_FP64)  // the FP32 factor, widened
        mlk::k_convert<float, double><<<grid_for(op.nnz), 256, 0, op.stream>>>(op.nnz, op.d_l32.p, op.d_l.p);
      HIP_CHECK(hipMemcpyAsync(l, op.d_l.p, sizeof(double) * op.nnz, hipMemcpyDeviceToHost, op.stream));
      HIP_CHECK(hipStreamSynchronize(op.stream));
    }
    if (shift_index) *shift_index = op.shift_index;
    return ML_OK;
  });
}

void ml_pcg_destroy(ml_pcg* h) { delete h; }

ml_status ml_pcg_set_wait_limit(ml_pcg* h, int64_t ticks) {
  return guarded([&]() -> int {
    if (!h || ticks < 0) throw ArgError("bad arguments");
    h->op.wait_ticks = ticks > 0 ? ticks : wait_limit();
    return ML_OK;
  });
}

ml_status ml_pcg_set_sweeps(ml_pcg* h, int sweeps) {
  return guarded([&]() -> int {
    if (!h) throw ArgError("null handle");
    h->op.set_sweeps(sweeps);
    h->op.sweeps_pinned = true;
    return ML_OK;
  });
}

int ml_pcg_sweeps(const ml_pcg* h) { return h ? h->op.sweeps : 0; }

ml_status ml_model_set_sweeps(ml_model* h, int sweeps) {
  return guarded([&]() -> int {
    if (!h) throw ArgError("null handle");
    if (h->m.all_fixed()) {
      if (sweeps < 0 || sweeps > 64) throw ArgError("sweeps must be in 0..64");
      return ML_OK;  // no free system
    }
    h->m.op.set_sweeps(sweeps);
    h->m.op.sweeps_pinned = true;
    return ML_OK;
  });
}

int ml_model_sweeps(const ml_model* h) { return h ? h->m.op.sweeps : 0; }

static ml_status set_sweep_weight(Operator* op, double w) {
  return guarded([&]() -> int {
    if (!op) throw ArgError("null handle");
    if (!(w > 0.0 && w <= 1.0)) throw ArgError("the sweep weight must be in (0, 1]");
    if (op->precision == ML_FP64) throw ArgError("the FP64 drop-in has no sweeps");
    op->sweep_w = w;
    op->sweeps_pinned = true;
    return ML_OK;
  });
}

ml_status ml_pcg_set_sweep_weight(ml_pcg* h, double w) {
  return set_sweep_weight(h ? &h->op : nullptr, w);
}

double ml_pcg_sweep_weight(const ml_pcg* h) { return h ? h->op.sweep_w : 0.0; }

ml_status ml_model_set_sweep_weight(ml_model* h, double w) {
  return set_sweep_weight(h ? &h->m.op : nullptr, w);
}

double ml_model_sweep_weight(const ml_model* h) { return h ? h->m.op.sweep_w : 0.0; }

static ml_status set_batch_lanes(Operator* op, int lanes) {
  return guarded([&]() -> int {
    if (!op) throw ArgError("null handle");
    if (lanes != 0 && lanes != 8 && lanes != 16 && lanes != 32 && lanes != 64)
      throw ArgError("lanes per row must be 0, 8, 16, 32 or 64");
    op->il_lanes = lanes;
    return ML_OK;
  });
}

ml_status ml_pcg_set_batch_lanes(ml_pcg* h, int lanes) {
  return set_batch_lanes(h ? &h->op : nullptr, lanes);
}

ml_status ml_model_set_batch_lanes(ml_model* h, int lanes) {
  return set_batch_lanes(h ? &h->m.op : nullptr, lanes);
}

int ml_pcg_batch_lanes(const ml_pcg* h) { return h ? h->op.il_lanes : 0; }

int ml_model_batch_lanes(const ml_model* h) { return h ? h->m.op.il_lanes : 0; }

ml_status ml_pcg_set_sweeps_fused(ml_pcg* h, int fused) {
  return guarded([&]() -> int {
    if (!h) throw ArgError("null handle");
    h->op.set_fused(fused);
    return ML_OK;
  });
}

int ml_pcg_sweeps_fused(const ml_pcg* h) { return h ? h->op.fused_capacity() : 0; }

ml_status ml_model_set_sweeps_fused(ml_model* h, int fused) {
  return guarded([&]() -> int {
    if (!h) throw ArgError("null handle");
    if (h->m.all_fixed()) return ML_OK;  // no free system
    h->m.op.set_fused(fused);
    return ML_OK;
  });
}

int ml_model_sweeps_fused(const ml_model* h) { return h ? h->m.op.fused_capacity() : 0; }

ml_status ml_pcg_set_reference_dots(ml_pcg* h, int on) {
  return guarded([&]() -> int {
    if (!h) throw ArgError("null handle");
    if (on && h->op.precision == ML_MIXED) throw ArgError("ML_MIXED keeps its tree sums");
    h->op.ref_dots = on != 0;
    return ML_OK;
  });
}

ml_status ml_model_set_reference_dots(ml_model* h, int on) {
  return guarded([&]() -> int {
    if (!h) throw ArgError("null handle");
    if (on && h->m.op.precision == ML_MIXED) throw ArgError("ML_MIXED keeps its tree sums");
    h->m.op.ref_dots = on != 0;
    return ML_OK;
  });
}

ml_status ml_model_residency(const ml_model* h, int32_t* out) {
  return guarded([&]() -> int {
    if (!h || !out) throw ArgError("null argument");
    for (int k = 0; k < 8; ++k) out[k] = h->m.op.residency[k];
    return ML_OK;
  });
}

ml_status ml_model_spmv_index(const ml_model* h, int64_t* out) {
  return guarded([&]() -> int {
    if (!h || !out) throw ArgError("null argument");
    const Operator& op = h->m.op;
    const int64_t full = 2 * op.nnz - op.n;
    out[0] = op.spmv_grouped ? 1 : op.spmv_d16 ? 2 : 4;
    out[1] = op.spmv_grouped ? op.grouped_bytes
                             : full * (op.spmv_d16 ? 2 : 4) + (int64_t)op.n * 4 + (op.slices + 1) * 4;
    return ML_OK;
  });
}

ml_status ml_model_index_bytes(const ml_model* h, int32_t* out) {
  return guarded([&]() -> int {
    if (!h || !out) throw ArgError("null argument");
    const Operator& op = h->m.op;
    out[0] = op.spmv_grouped ? 0 : op.spmv_d16 ? 2 : 4;
    out[1] = !op.have_sweep ? 0 : (op.fused_active() || (op.sw_lo.d16 && op.sw_up.d16)) ? 2 : 4;
    return ML_OK;
  });
}

ml_status ml_pcg_trace(ml_pcg* h, const double* r, int64_t* nblocks, int64_t* trace_fw,
                       int64_t* trace_bw, int64_t capacity) {
  return guarded([&]() -> int {
    if (!h || !r || !nblocks) throw ArgError("null argument");
    Operator& op = h->op;
    nblocks[0] = op.fw.nblk;
    nblocks[1] = op.bw.nblk;
    if (!trace_fw || !trace_bw) return ML_OK;
    if (capacity < 10 * (int64_t)std::max(op.fw.nblk, op.bw.nblk) + op.n)
      throw ArgError("trace too small");
    if (!op.factored) {
      int st = op.factor();
      if (st != ML_OK) return st;
    }
    DBuf<long long> tf, tb;
    tf.alloc(10 * op.fw.nblk + op.n);
    tb.alloc(10 * op.bw.nblk + op.n);
    HIP_CHECK(hipMemcpyAsync(op.w_r.p, r, sizeof(double) * op.n, hipMemcpyHostToDevice, op.stream));
    op.apply_precondition(tf.p, tb.p);
    HIP_CHECK(hipMemcpyAsync(trace_fw, tf.p, sizeof(long long) * (10 * op.fw.nblk + op.n),
                             hipMemcpyDeviceToHost, op.stream));
    HIP_CHECK(hipMemcpyAsync(trace_bw, tb.p, sizeof(long long) * (10 * op.bw.nblk + op.n),
                             hipMemcpyDeviceToHost, op.stream));
    HIP_CHECK(hipStreamSynchronize(op.stream));
    return ML_OK;
  });
}

ml_status ml_model_create(int64_t n_nodes, const double* nodes, int64_t n_elements, int nen,
                          const int64_t* conn, const double* young, const double* nu,
                          const uint8_t* fixed, const double* imposed, ml_solver_kind solver,
                          ml_precision precision, ml_model** out, ml_solve_info* info) {
  return guarded([&]() -> int {
    if (!out || !nodes || !conn || !young || !nu || !fixed) throw ArgError("null argument");
    if (solver == ML_DIRECT)
      throw ArgError("the direct (skyline) solver is not part of the device library");
    if (precision < ML_FP64 || precision > ML_MIXED_JACOBI) throw ArgError("bad precision");
    if (info) {
      std::memset(info, 0, sizeof(*info));
      info->bad_element = -1;
    }
    std::unique_ptr<ml_model> h(new ml_model);
    int st = h->m.create(n_nodes, nodes, n_elements, nen, conn, young, nu, fixed, imposed,
                         precision, solver, info);
    if (st != ML_OK) return st;
    *out = h.release();
    return ML_OK;
  });
}

ml_status ml_model_assemble(ml_model* h) {
  return guarded([&]() -> int {
    if (!h) throw ArgError("null handle");
    h->m.assemble();
    HIP_CHECK(hipStreamSynchronize(h->m.s()));
    return ML_OK;
  });
}

ml_status ml_model_set_young(ml_model* h, const double* young) {
  return guarded([&]() -> int {
    if (!h || !young) throw ArgError("null argument");
    Model& m = h->m;
    for (int64_t e = 0; e < m.plan.n_elements; ++e)
      if (!(std::isfinite(young[e]) && young[e] > 0.0)) throw ArgError("Young's modulus must be > 0");
    m.young.assign(young, young + m.plan.n_elements);
    m.d_young.upload_vec(m.young, m.s());
    m.op.factored = false;
    m.assembled = false;
    return ML_OK;
  });
}

ml_status ml_model_factor(ml_model* h, ml_solve_info* info) {
  return guarded([&]() -> int {
    if (!h) throw ArgError("null handle");
    if (!h->m.assembled) h->m.assemble();
    if (h->m.all_fixed()) return ML_OK;  // no free system to factor
    double t0 = now_s();
    const int retries0 = h->m.op.retries;
    int st = h->m.op.factor();
    if (info) {
      info->shift_index = h->m.op.shift_index;
      info->device_retries = h->m.op.retries - retries0;
      info->seconds[1] = now_s() - t0;
    }
    return st;
  });
}

ml_status ml_model_solve_device(ml_model* h, const double* f, double* u, double* R,
                                ml_solve_info* info) {
  return guarded([&]() -> int {
    if (!h || !f || !u) throw ArgError("null argument");
    if (!h->m.assembled) h->m.assemble();
    return h->m.solve_device(f, u, R, info);
  });
}

ml_status ml_model_solve(ml_model* h, const double* f, double* u, double* R, ml_solve_info* info) {
  return guarded([&]() -> int {
    if (!h || !f || !u) throw ArgError("null argument");
    Model& m = h->m;
    if (!m.assembled) m.assemble();
    HIP_CHECK(hipMemcpyAsync(m.d_f.p, f, sizeof(double) * m.plan.ndof, hipMemcpyHostToDevice, m.s()));
    int st = m.solve_device(m.d_f.p, m.d_u.p, R ? m.d_R.p : nullptr, info);
    if (st != ML_OK && st != ML_NOT_CONVERGED) return st;
    HIP_CHECK(hipMemcpyAsync(u, m.d_u.p, sizeof(double) * m.plan.ndof, hipMemcpyDeviceToHost, m.s()));
    if (R)
      HIP_CHECK(hipMemcpyAsync(R, m.d_R.p, sizeof(double) * m.plan.ndof, hipMemcpyDeviceToHost, m.s()));
    HIP_CHECK(hipStreamSynchronize(m.s()));
    return st;
  });
}

ml_status ml_model_stresses(ml_model* h, const double* u, double* out) {
  return guarded([&]() -> int {
    if (!h || !u || !out) throw ArgError("null argument");
    Model& m = h->m;
    HIP_CHECK(hipMemcpyAsync(m.d_u.p, u, sizeof(double) * m.plan.ndof, hipMemcpyHostToDevice, m.s()));
    DBuf<double> d_out;
    size_t count = (size_t)m.plan.n_elements * m.ngp * 9;
    d_out.alloc(count);
    m.stresses(m.d_u.p, d_out.p);
    HIP_CHECK(hipMemcpyAsync(out, d_out.p, count * sizeof(double), hipMemcpyDeviceToHost, m.s()));
    HIP_CHECK(hipStreamSynchronize(m.s()));
    return ML_OK;
  });
}

ml_status ml_model_stresses_device(ml_model* h, const double* u, double* out) {
  return guarded([&]() -> int {
    if (!h || !u || !out) throw ArgError("null argument");
    Model& m = h->m;
    m.stresses(u, out);
    HIP_CHECK(hipStreamSynchronize(m.s()));
    return ML_OK;
  });
}

ml_status ml_model_nodal_stresses(ml_model* h, const double* u, double* out) {
  return guarded([&]() -> int {
    if (!h || !u || !out) throw ArgError("null argument");
    Model& m = h->m;
    HIP_CHECK(hipMemcpyAsync(m.d_u.p, u, sizeof(double) * m.plan.ndof, hipMemcpyHostToDevice, m.s()));
    DBuf<double> d_gp, d_out;
    d_gp.alloc((size_t)m.plan.n_elements * m.ngp * 9);
    d_out.alloc((size_t)m.mesh.n_nodes * 9);
    m.stresses(m.d_u.p, d_gp.p);
    m.nodal_stresses(d_gp.p, d_out.p);
    HIP_CHECK(hipMemcpyAsync(out, d_out.p, sizeof(double) * 9 * m.mesh.n_nodes,
                             hipMemcpyDeviceToHost, m.s()));
    HIP_CHECK(hipStreamSynchronize(m.s()));
    return ML_OK;
  });
}

int64_t ml_model_nfree(const ml_model* h) { return h ? h->m.plan.nfree : 0; }
int64_t ml_model_nnz(const ml_model* h) { return h ? h->m.plan.lower.nnz() : 0; }
int64_t ml_model_levels(const ml_model* h) { return h ? h->m.plan.depth : 0; }

ml_status ml_model_csr(ml_model* h, int64_t* order, int64_t* offsets, int64_t* cols,
                       double* values) {
  return guarded([&]() -> int {
    if (!h) throw ArgError("null handle");
    Model& m = h->m;
    if (order) std::copy(m.plan.order.begin(), m.plan.order.end(), order);
    if (offsets) std::copy(m.plan.lower.offsets.begin(), m.plan.lower.offsets.end(), offsets);
    if (cols) std::copy(m.plan.lower.cols.begin(), m.plan.lower.cols.end(), cols);
    if (values) {
      HIP_CHECK(hipMemcpyAsync(values, m.op.d_a.p, sizeof(double) * m.op.nnz, hipMemcpyDeviceToHost, m.s()));
      HIP_CHECK(hipStreamSynchronize(m.s()));
    }
    return ML_OK;
  });
}

ml_status ml_model_compliance_gradient(ml_model* h, double* out) {
  return guarded([&]() -> int {
    if (!h || !out) throw ArgError("null argument");
    Model& m = h->m;
    if (!m.have_u) throw ArgError("no solution yet: solve the model first");
    for (double v : m.imposed)
      if (v != 0.0)
        throw ArgError("the compliance gradient needs homogeneous essential conditions");
    DBuf<double> d_out;
    d_out.alloc(m.plan.n_elements);
    m.compliance_gradient(d_out.p);
    HIP_CHECK(hipMemcpyAsync(out, d_out.p, sizeof(double) * m.plan.n_elements,
                             hipMemcpyDeviceToHost, m.s()));
    HIP_CHECK(hipStreamSynchronize(m.s()));
    return ML_OK;
  });
}

ml_status ml_model_solve_quasi_static(ml_model* h, const double* yield_stress,
                                      const double* hardening, const double* f,
                                      double time_period, int64_t increments, int64_t npoints,
                                      const double* amp_t, const double* amp_v, double* u,
                                      double* reactions, double* stresses, double* p,
                                      double* residuals, int64_t* residual_counts,
                                      ml_quasi_static_info* info) {
  return guarded([&]() -> int {
    if (!h || !yield_stress || !hardening || !f || !amp_t || !amp_v) throw ArgError("null argument");
    Model& m = h->m;
    const int64_t ne = m.plan.n_elements, ndof = m.plan.ndof;
    if (!(std::isfinite(time_period) && time_period > 0.0))
      throw ArgError("the Step's time period must be finite and > 0");
    if (increments <= 0) throw ArgError("a Step needs at least one increment");
    if (npoints <= 0) throw ArgError("amplitude point 0 is invalid");
    for (int64_t k = 0; k < npoints; ++k) {
      bool ok = std::isfinite(amp_t[k]) && std::isfinite(amp_v[k]);
      const double eps = std::numeric_limits<double>::epsilon();
      ok = ok && (k == 0 ? (std::fabs(amp_t[k]) <= eps && std::fabs(amp_v[k]) <= eps)
                         : amp_t[k] > amp_t[k - 1]);
      if (!ok) throw ArgError("amplitude point " + std::to_string(k) + " is invalid");
    }
    for (int64_t e = 0; e < ne; ++e) {
      if (!(yield_stress[e] > 0.0) || std::isnan(yield_stress[e]))
        throw ArgError("yield stress must be > 0 (infinite: elastic)");
      if (!(std::isfinite(hardening[e]) && hardening[e] >= 0.0))
        throw ArgError("hardening must be finite and >= 0");
    }
    for (int64_t g = 0; g < ndof; ++g)
      if (!std::isfinite(f[g])) throw ArgError("loads must be finite");
    if (info) {
      std::memset(info, 0, sizeof(*info));
      info->failed_increment = -1;
    }
    const double t_start = now_s();
    QuasiStatic qs(m, yield_stress, hardening, f);
    qs.amp_t.assign(amp_t, amp_t + npoints);
    qs.amp_v.assign(amp_v, amp_v + npoints);
    qs.responses(qs.d_u.p);  // model.rs:767 (the mesh at rest)
    qs.record(0, u, reactions, stresses, p);
    const double dt = time_period / (double)increments;  // Step::time_step (model.rs:171-173)
    int st = ML_OK;
    for (int64_t inc = 1; inc <= increments; ++inc) {
      const double time = (double)inc * dt;
      const double start = time - dt;
      st = qs.advance(start, time, 0);
      if (st != ML_OK) {
        if (info) {
          info->failed_increment = inc;
          info->failed_time = time;
        }
        break;
      }
      qs.record(inc, u, reactions, stresses, p);
      if (residual_counts) residual_counts[inc - 1] = (int64_t)qs.residuals.size();
      if (residuals)
        std::copy(qs.residuals.begin(), qs.residuals.end(),
                  residuals + (inc - 1) * (QuasiStatic::NEWTON_ITERATIONS + 1));
    }
    HIP_CHECK(hipStreamSynchronize(m.s()));
    if (info) {
      info->newton_iterations = qs.newton;
      info->cuts = qs.cuts;
      info->pcg_iterations = qs.pcg_iterations;
      info->device_retries = qs.device_retries;
      info->seconds[0] = qs.t_resp;
      info->seconds[1] = qs.t_factor;
      info->seconds[2] = qs.t_solve;
      info->seconds[3] = now_s() - t_start;
    }
    return st;
  });
}

// ------------------------------------------------------------------ transient heat conduction
struct ml_thermal {
  Thermal t;
};

ml_status ml_thermal_create(int64_t n_nodes, const double* nodes, int64_t n_elements, int nen,
                            const int64_t* conn, const double* conductivity, const double* rho_c,
                            int64_t nconv, const int64_t* conv_elements, const int64_t* conv_faces,
                            const double* film, const double* sink, const uint8_t* fixed,
                            const double* imposed, ml_precision precision, ml_thermal** out,
                            ml_solve_info* info) {
  return guarded([&]() -> int {
    if (!out || !nodes || !conn || !conductivity || !rho_c || !fixed) throw ArgError("null argument");
    if (nconv > 0 && (!conv_elements || !conv_faces || !film || !sink)) throw ArgError("null argument");
    if (precision < ML_FP64 || precision > ML_MIXED_JACOBI) throw ArgError("bad precision");
    if (info) {
      std::memset(info, 0, sizeof(*info));
      info->bad_element = -1;
    }
    std::unique_ptr<ml_thermal> h(new ml_thermal);
    int st = h->t.create(n_nodes, nodes, n_elements, nen, conn, conductivity, rho_c, nconv,
                         conv_elements, conv_faces, film, sink, fixed, imposed, precision, info);
    if (st != ML_OK) return st;
    *out = h.release();
    return ML_OK;
  });
}

ml_status ml_thermal_add_loads(ml_thermal* h, int64_t nsources, const int64_t* elements,
                               const double* values, int64_t nconv, const int64_t* conv_elements,
                               const int64_t* conv_faces, const double* film, const double* sink,
                               double* f) {
  return guarded([&]() -> int {
    if (!h || !f) throw ArgError("null argument");
    Thermal& t = h->t;
    const int nen = t.plan.nen;
    const int64_t ne = t.plan.n_elements;
    // assembly.rs:465-484 (element_heat_source): f_i += w |det J| N_i q on the element's rule
    for (int64_t s2 = 0; s2 < nsources; ++s2) {
      const int64_t e = elements[s2];
      if (e < 0 || e >= ne) throw ArgError("heat source element out of range");
      if (!std::isfinite(values[s2])) throw ArgError("heat source must be finite");
      const int64_t* en = t.conn.data() + e * nen;
      double x[10][3];
      for (int a = 0; a < nen; ++a)
        for (int c = 0; c < 3; ++c) x[a][c] = t.nodes[3 * en[a] + c];
      double fe[10] = {0};
      const int ng = nen == 4 ? 1 : 4;
      const double A = 0.585410196624968500, B = 0.138196601125010500;
      for (int q = 0; q < ng; ++q) {
        double xi = 0.25, eta = 0.25, zeta = 0.25, w = 1.0 / 6.0;
        if (nen == 10) {
          xi = q == 1 ? A : B; eta = q == 2 ? A : B; zeta = q == 3 ? A : B; w = 1.0 / 24.0;
        }
        double n[10], det;
        tet_point_host(nen, x, xi, eta, zeta, n, &det);
        const double weight = w * std::fabs(det);
        for (int a = 0; a < nen; ++a) fe[a] += weight * n[a] * values[s2];
      }
      for (int a = 0; a < nen; ++a) f[en[a]] += fe[a];
    }
    // model.rs:1122-1126: convection loads, faces grouped per element in element order
    std::vector<int64_t> order(nconv);
    for (int64_t c = 0; c < nconv; ++c) {
      order[c] = c;
      if (conv_elements[c] < 0 || conv_elements[c] >= ne || conv_faces[c] < 0 || conv_faces[c] > 3)
        throw ArgError("convection face out of range");
    }
    std::stable_sort(order.begin(), order.end(),
                     [&](int64_t a, int64_t b) { return conv_elements[a] < conv_elements[b]; });
    int pos[6];
    double mat[36], load[6];
    for (int64_t c : order) {
      const int64_t* en = t.conn.data() + conv_elements[c] * nen;
      const int nn = face_convection_terms(nen, en, t.nodes.data(), (int)conv_faces[c], film[c],
                                           sink[c], pos, mat, load);
      for (int a = 0; a < nn; ++a) f[en[pos[a]]] += load[a];
    }
    return ML_OK;
  });
}

ml_status ml_thermal_solve_step(ml_thermal* h, double time_period, int64_t increments,
                                double theta, const double* f, const double* initial,
                                double* temperatures, double* reactions, double* fluxes,
                                ml_solve_info* info) {
  return guarded([&]() -> int {
    if (!h || !f || !initial || !temperatures) throw ArgError("null argument");
    if (!(std::isfinite(time_period) && time_period > 0.0))
      throw ArgError("the Step's time period must be finite and > 0");
    if (increments <= 0) throw ArgError("a Step needs at least one increment");
    if (!(theta >= 0.5 && theta <= 1.0)) throw ArgError("theta must lie in [0.5, 1]");
    Thermal& t = h->t;
    const int64_t nn = t.plan.n_nodes, ne = t.plan.n_elements;
    for (int64_t v = 0; v < nn; ++v)
      if (!std::isfinite(initial[v])) throw ArgError("initial temperatures must be finite");
    if (info) {
      std::memset(info, 0, sizeof(*info));
      info->bad_element = -1;
    }
    const double t0 = now_s();
    const double dt = time_period / (double)increments;  // Step::time_step (model.rs:171-173)
    int st = t.prepare(dt, theta);
    if (info) info->shift_index = t.op.shift_index;
    if (st != ML_OK) return st;
    const double t1 = now_s();
    hipStream_t s = t.s();
    // model.rs:666-669: the initial field, the BC values at the fixed nodes
    std::vector<double> t_init(initial, initial + nn);
    for (int64_t v = 0; v < nn; ++v)
      if (t.fixed[v]) t_init[v] = t.imposed[v];
    t.d_t.upload_vec(t_init, s);
    t.d_f.upload(f, nn, s);
    const size_t nq = (size_t)6 * (t.plan.nen == 4 ? 1 : 4) * ne;
    auto record = [&](int64_t k) {
      HIP_CHECK(hipMemcpyAsync(temperatures + k * nn, t.d_t.p, sizeof(double) * nn,
                               hipMemcpyDeviceToHost, s));
      if (reactions) {
        if (k == 0) std::fill(reactions, reactions + nn, 0.0);
        else HIP_CHECK(hipMemcpyAsync(reactions + k * nn, t.d_R.p, sizeof(double) * nn,
                                      hipMemcpyDeviceToHost, s));
      }
      if (fluxes) {
        t.fluxes(t.d_q.p);
        HIP_CHECK(hipMemcpyAsync(fluxes + k * nq, t.d_q.p, sizeof(double) * nq,
                                 hipMemcpyDeviceToHost, s));
      }
    };
    record(0);
    int64_t iters = 0;
    for (int64_t k = 1; k <= increments; ++k) {
      ml_solve_info si;
      std::memset(&si, 0, sizeof(si));
      st = t.step(&si);
      iters += si.iterations;
      if (info) {
        info->device_retries += si.device_retries;
        info->residual = si.residual;
      }
      if (st != ML_OK) {
        if (info) info->iterations = iters;
        return st;
      }
      record(k);
    }
    HIP_CHECK(hipStreamSynchronize(s));
    if (info) {
      info->iterations = iters;
      info->seconds[1] = t1 - t0;
      info->seconds[2] = now_s() - t1;
    }
    return ML_OK;
  });
}

ml_status ml_thermal_solve_steady(ml_thermal* h, const double* f, double* temperatures,
                                  double* reactions, double* fluxes, ml_solve_info* info) {
  return guarded([&]() -> int {
    if (!h || !f || !temperatures) throw ArgError("null argument");
    Thermal& t = h->t;
    const int64_t nn = t.plan.n_nodes, ne = t.plan.n_elements;
    if (info) {
      std::memset(info, 0, sizeof(*info));
      info->bad_element = -1;
    }
    // K T = F (Model::solve, Analysis::Thermal): the left matrix with C / dt = 0 and theta = 1
    // is K itself (0 + 1 K rounds to K)
    const double t0 = now_s();
    int st = t.prepare(std::numeric_limits<double>::infinity(), 1.0);
    if (info) info->shift_index = t.op.shift_index;
    if (st != ML_OK) return st;
    const double t1 = now_s();
    hipStream_t s = t.s();
    t.d_t.upload_vec(t.imposed, s);  // the BC temperatures at the fixed nodes
    t.d_f.upload(f, nn, s);
    // the step's carried term is (C/dt - 0 K) T = 0: F alone on the right
    mlk::k_fill_f64<<<grid_for(nn), 256, 0, s>>>(t.d_carry.p, nn, 0.0);
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipMemcpyAsync(t.d_carry.p, t.d_f.p, sizeof(double) * nn, hipMemcpyDeviceToDevice, s));
    ml_solve_info si;
    std::memset(&si, 0, sizeof(si));
    mlk::k_rhs<<<grid_for(nn), 256, 0, s>>>(nn, t.d_rank.p, t.d_carry.p,
                                            t.plan.cpl_slot.empty() ? nullptr : t.d_coupling.p,
                                            t.d_rhs.p);
    HIP_CHECK(hipGetLastError());
    st = t.op.solve(t.d_rhs.p, t.d_x.p, 0, 0, &si);
    if (info) {
      info->iterations = si.iterations;
      info->residual = si.residual;
      info->true_residual = si.true_residual;
      info->device_retries = si.device_retries;
    }
    if (st != ML_OK) return st;
    mlk::k_scatter_u<<<grid_for(nn), 256, 0, s>>>(nn, t.d_rank.p, t.d_x.p, t.d_t.p, t.d_t.p);
    mlk::k_fill_f64<<<grid_for(nn), 256, 0, s>>>(t.d_R.p, nn, 0.0);
    const int64_t nfixed = (int64_t)t.plan.rea_dof.size();
    if (nfixed) {
      if (t.plan.nen == 4)
        mlk::k_thermal_reactions<4><<<grid_for(nfixed), 256, 0, s>>>(
            nfixed, t.d_rea_dof.p, t.d_rea_off.p, t.d_rea_item.p, t.d_conn.p, ne, t.d_left.p,
            t.d_t.p, t.d_carry.p, t.d_R.p);
      else
        mlk::k_thermal_reactions<10><<<grid_for(nfixed), 256, 0, s>>>(
            nfixed, t.d_rea_dof.p, t.d_rea_off.p, t.d_rea_item.p, t.d_conn.p, ne, t.d_left.p,
            t.d_t.p, t.d_carry.p, t.d_R.p);
    }
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipMemcpyAsync(temperatures, t.d_t.p, sizeof(double) * nn, hipMemcpyDeviceToHost, s));
    if (reactions)
      HIP_CHECK(hipMemcpyAsync(reactions, t.d_R.p, sizeof(double) * nn, hipMemcpyDeviceToHost, s));
    if (fluxes) {
      t.fluxes(t.d_q.p);
      HIP_CHECK(hipMemcpyAsync(fluxes, t.d_q.p,
                               sizeof(double) * 6 * (t.plan.nen == 4 ? 1 : 4) * ne,
                               hipMemcpyDeviceToHost, s));
    }
    HIP_CHECK(hipStreamSynchronize(s));
    if (info) {
      info->seconds[1] = t1 - t0;
      info->seconds[2] = now_s() - t1;
    }
    return ML_OK;
  });
}

void ml_thermal_destroy(ml_thermal* h) { delete h; }

ml_status ml_model_solve_large_strain(ml_model* h, const uint8_t* neo_hookean, const double* f,
                                      int64_t npressures, const int64_t* pres_elements,
                                      const int64_t* pres_faces, const double* pres_values,
                                      double time_period, int64_t increments, int64_t npoints,
                                      const double* amp_t, const double* amp_v, double* u,
                                      double* reactions, double* stresses, double* residuals,
                                      int64_t* residual_counts, ml_quasi_static_info* info) {
  return guarded([&]() -> int {
    if (!h || !f || !amp_t || !amp_v) throw ArgError("null argument");
    if (npressures > 0 && (!pres_elements || !pres_faces || !pres_values))
      throw ArgError("null argument");
    Model& m = h->m;
    const int64_t ne = m.plan.n_elements, ndof = m.plan.ndof;
    if (!(std::isfinite(time_period) && time_period > 0.0))
      throw ArgError("the Step's time period must be finite and > 0");
    if (increments <= 0) throw ArgError("a Step needs at least one increment");
    if (npoints <= 0) throw ArgError("amplitude point 0 is invalid");
    for (int64_t k = 0; k < npoints; ++k) {
      bool ok = std::isfinite(amp_t[k]) && std::isfinite(amp_v[k]);
      const double eps = std::numeric_limits<double>::epsilon();
      ok = ok && (k == 0 ? (std::fabs(amp_t[k]) <= eps && std::fabs(amp_v[k]) <= eps)
                         : amp_t[k] > amp_t[k - 1]);
      if (!ok) throw ArgError("amplitude point " + std::to_string(k) + " is invalid");
    }
    for (int64_t g = 0; g < ndof; ++g)
      if (!std::isfinite(f[g])) throw ArgError("loads must be finite");
    for (int64_t k = 0; k < npressures; ++k) {
      if (pres_elements[k] < 0 || pres_elements[k] >= ne || pres_faces[k] < 0 || pres_faces[k] > 3)
        throw ArgError("pressure face out of range");
      if (!std::isfinite(pres_values[k])) throw ArgError("pressure must be finite");
    }
    if (info) {
      std::memset(info, 0, sizeof(*info));
      info->failed_increment = -1;
    }
    const double t_start = now_s();
    std::vector<double> inf(ne, std::numeric_limits<double>::infinity()), zero(ne, 0.0);
    QuasiStatic qs(m, inf.data(), zero.data(), f);
    qs.large = true;
    std::vector<unsigned char> neo(ne, 0);
    if (neo_hookean)
      for (int64_t e = 0; e < ne; ++e) neo[e] = neo_hookean[e] ? 1 : 0;
    qs.d_neo.upload_vec(neo, m.s());
    qs.pres_el.assign(pres_elements, pres_elements + npressures);
    qs.pres_face.assign(pres_faces, pres_faces + npressures);
    qs.pres_val.assign(pres_values, pres_values + npressures);
    qs.amp_t.assign(amp_t, amp_t + npoints);
    qs.amp_v.assign(amp_v, amp_v + npoints);
    qs.responses(qs.d_u.p);  // model.rs:767 (the mesh at rest)
    qs.record(0, u, reactions, stresses, nullptr);
    const double dt = time_period / (double)increments;
    int st = ML_OK;
    for (int64_t inc = 1; inc <= increments; ++inc) {
      const double time = (double)inc * dt;
      st = qs.advance(time - dt, time, 0);
      if (st != ML_OK) {
        if (info) {
          info->failed_increment = inc;
          info->failed_time = time;
        }
        break;
      }
      qs.record(inc, u, reactions, stresses, nullptr);
      if (residual_counts) residual_counts[inc - 1] = (int64_t)qs.residuals.size();
      if (residuals)
        std::copy(qs.residuals.begin(), qs.residuals.end(),
                  residuals + (inc - 1) * (QuasiStatic::NEWTON_ITERATIONS + 1));
    }
    HIP_CHECK(hipStreamSynchronize(m.s()));
    if (info) {
      info->newton_iterations = qs.newton;
      info->cuts = qs.cuts;
      info->pcg_iterations = qs.pcg_iterations;
      info->device_retries = qs.device_retries;
      info->seconds[0] = qs.t_resp;
      info->seconds[1] = qs.t_factor;
      info->seconds[2] = qs.t_solve;
      info->seconds[3] = now_s() - t_start;
    }
    return st;
  });
}

ml_status ml_model_add_pressures(ml_model* h, int64_t count, const int64_t* elements,
                                 const int64_t* faces, const double* values, double* f) {
  return guarded([&]() -> int {
    if (!h || (count > 0 && (!elements || !faces || !values)) || !f) throw ArgError("null argument");
    Model& m = h->m;
    for (int64_t k = 0; k < count; ++k) {
      if (elements[k] < 0 || elements[k] >= m.plan.n_elements || faces[k] < 0 || faces[k] > 3)
        throw ArgError("pressure on a face that does not exist");
      face_pressure_add(m.plan.nen, m.conn.data() + elements[k] * m.plan.nen, m.nodes.data(),
                        (int)faces[k], values[k], f);
    }
    return ML_OK;
  });
}

ml_status ml_model_add_edge_loads(ml_model* h, int64_t count, const int64_t* elements,
                                  const int64_t* edges, const double* q, double* f) {
  return guarded([&]() -> int {
    if (!h || (count > 0 && (!elements || !edges || !q)) || !f) throw ArgError("null argument");
    Model& m = h->m;
    for (int64_t k = 0; k < count; ++k) {
      if (elements[k] < 0 || elements[k] >= m.plan.n_elements || edges[k] < 0 || edges[k] > 5)
        throw ArgError("line load on an edge that does not exist");
      for (int c = 0; c < 3; ++c)
        if (!std::isfinite(q[3 * k + c])) throw ArgError("line load not finite");
    }
    for (int64_t k = 0; k < count; ++k)
      edge_load_add(m.plan.nen, m.conn.data() + elements[k] * m.plan.nen, m.nodes.data(),
                    (int)edges[k], q + 3 * k, f);
    return ML_OK;
  });
}

ml_status ml_model_solve_many(ml_model* h, int64_t nrhs, const double* f, double* u,
                              ml_solve_info* infos) {
  return guarded([&]() -> int {
    if (!h || !f || !u || nrhs <= 0) throw ArgError("bad arguments");
    Model& m = h->m;
    if (m.all_fixed()) throw ArgError("every DOF is fixed: nothing to solve for");
    if (!m.assembled) m.assemble();
    Operator& op = m.op;
    const int64_t ndof = m.plan.ndof, nf = m.plan.nfree;
    int result = ML_OK;
    for (int64_t k0 = 0; k0 < nrhs; k0 += op.rhs_per_pass) {
      const int k = (int)std::min<int64_t>(op.rhs_per_pass, nrhs - k0);
      DBuf<double> fd, rhs, xd, ud;
      fd.upload(f + k0 * ndof, (size_t)k * ndof, op.stream);
      rhs.alloc((size_t)k * nf);
      xd.alloc((size_t)k * nf);
      ud.alloc((size_t)k * ndof);
      for (int j = 0; j < k; ++j)
        mlk::k_rhs<<<grid_for(ndof), 256, 0, op.stream>>>(
            ndof, m.d_rank.p, fd.p + j * ndof, m.plan.cpl_slot.empty() ? nullptr : m.d_coupling.p,
            rhs.p + j * nf);
      HIP_CHECK(hipGetLastError());
      std::vector<mlk::PcgScalars> hS;
      int st = op.solve_many(k, rhs.p, xd.p, 0, 0, &hS);
      if (st != ML_OK) return st;
      for (int j = 0; j < k; ++j)
        mlk::k_scatter_u<<<grid_for(ndof), 256, 0, op.stream>>>(ndof, m.d_rank.p, xd.p + j * nf,
                                                                m.d_imposed.p, ud.p + j * ndof);
      HIP_CHECK(hipMemcpyAsync(u + k0 * ndof, ud.p, sizeof(double) * k * ndof,
                               hipMemcpyDeviceToHost, op.stream));
      HIP_CHECK(hipStreamSynchronize(op.stream));
      for (int j = 0; j < k; ++j) {
        if (status_of(hS[j]) != ML_OK && result == ML_OK) result = status_of(hS[j]);
        if (infos) info_of(hS[j], op.shift_index, infos + k0 + j);
      }
    }
    return result;
  });
}

ml_status ml_batch_create(int64_t count, ml_model* const* models, ml_precision precision,
                          ml_batch** out) {
  return guarded([&]() -> int {
    if (!out || count <= 0 || !models) throw ArgError("bad arguments");
    if (precision != ML_FP64 && precision != ML_MIXED_IC0 && precision != ML_MIXED_JACOBI)
      throw ArgError("batched solves take ML_FP64, ML_MIXED_IC0 or ML_MIXED_JACOBI");
    for (int64_t k = 0; k < count; ++k)
      if (!models[k] || models[k]->m.all_fixed())
        throw ArgError("every model of a batch needs a free DOF");
    std::unique_ptr<ml_batch> b(new ml_batch);
    b->models.assign(models, models + count);
    b->nseg = (int)count;
    // block-diagonal concatenation of the models' lower patterns
    LowerPattern cat;
    cat.offsets.push_back(0);
    b->row_off.push_back(0);
    for (ml_model* h : b->models) {
      if (!h) throw ArgError("null model");
      Model& m = h->m;
      if (!m.assembled) m.assemble();
      const LowerPattern& lp = m.plan.lower;
      const int64_t r0 = cat.n(), e0 = cat.nnz();
      for (int64_t i = 0; i < lp.n(); ++i) cat.offsets.push_back(e0 + lp.offsets[i + 1]);
      for (int32_t c : lp.cols) cat.cols.push_back((int32_t)(c + r0));
      b->row_off.push_back(cat.n());
      if (cat.n() >= INT32_MAX || cat.nnz() >= INT32_MAX) throw ArgError("batch too large");
    }
    Operator& op = b->op;
    op.precision = precision;
    op.set_pattern(cat);
    int64_t e0 = 0;
    for (ml_model* h : b->models) {
      Model& m = h->m;
      HIP_CHECK(hipStreamSynchronize(m.s()));
      HIP_CHECK(hipMemcpyAsync(op.d_a.p + e0, m.op.d_a.p, sizeof(double) * m.plan.lower.nnz(),
                               hipMemcpyDeviceToDevice, op.stream));
      e0 += m.plan.lower.nnz();
    }
    op.values_changed();
    std::vector<int> seg(cat.n());
    for (int k = 0; k < b->nseg; ++k)
      std::fill(seg.begin() + b->row_off[k], seg.begin() + b->row_off[k + 1], k);
    b->d_seg_off.upload_vec(b->row_off, op.stream);
    b->d_seg.upload_vec(seg, op.stream);
    b->d_rhs.alloc(cat.n());
    b->d_x.alloc(cat.n());
    int st = precision == ML_FP64 ? op.factor_segments<double>(b->nseg, b->d_seg.p, &b->shift_idx)
                                  : op.factor_segments<float>(b->nseg, b->d_seg.p, &b->shift_idx);
    HIP_CHECK(hipStreamSynchronize(op.stream));
    if (st != ML_OK) return st;
    *out = b.release();
    return ML_OK;
  });
}

ml_status ml_batch_solve(ml_batch* b, const double* const* f, double* const* u,
                         ml_solve_info* infos) {
  return guarded([&]() -> int {
    if (!b || !f || !u) throw ArgError("null argument");
    Operator& op = b->op;
    for (int k = 0; k < b->nseg; ++k) {
      Model& m = b->models[k]->m;
      if (!f[k] || !u[k]) throw ArgError("null load or displacement array");
      HIP_CHECK(hipMemcpyAsync(m.d_f.p, f[k], sizeof(double) * m.plan.ndof,
                               hipMemcpyHostToDevice, op.stream));
      mlk::k_rhs<<<grid_for(m.plan.ndof), 256, 0, op.stream>>>(
          m.plan.ndof, m.d_rank.p, m.d_f.p, m.plan.cpl_slot.empty() ? nullptr : m.d_coupling.p,
          b->d_rhs.p + b->row_off[k]);
    }
    HIP_CHECK(hipGetLastError());
    std::vector<mlk::PcgScalars> hS;
    int st;
    for (;;) {
      st = op.pcg_batch<double>(1, b->nseg, b->d_seg_off.p, b->d_seg.p, b->d_rhs.p, b->d_x.p,
                                ML_TOLERANCE, ML_MAX_ITERATIONS, &hS);
      if (st == ML_DEVICE_TIMEOUT && op.shrink()) continue;
      break;
    }
    if (st != ML_OK) return st;
    int result = ML_OK;
    for (int k = 0; k < b->nseg; ++k) {
      Model& m = b->models[k]->m;
      mlk::k_scatter_u<<<grid_for(m.plan.ndof), 256, 0, op.stream>>>(
          m.plan.ndof, m.d_rank.p, b->d_x.p + b->row_off[k], m.d_imposed.p, m.d_ulast.p);
      HIP_CHECK(hipMemcpyAsync(u[k], m.d_ulast.p, sizeof(double) * m.plan.ndof,
                               hipMemcpyDeviceToHost, op.stream));
      m.have_u = true;
      if (status_of(hS[k]) != ML_OK && result == ML_OK) result = status_of(hS[k]);
      if (infos) info_of(hS[k], b->shift_idx[k], infos + k);
    }
    HIP_CHECK(hipStreamSynchronize(op.stream));
    return result;
  });
}

void ml_batch_destroy(ml_batch* b) { delete b; }

ml_status ml_pcg_autotune(ml_pcg* h, int64_t nrhs, int32_t* chosen) {
  return guarded([&]() -> int {
    if (!h) throw ArgError("null handle");
    return h->op.autotune(nrhs, chosen);
  });
}

ml_status ml_model_autotune(ml_model* h, int64_t nrhs, int32_t* chosen) {
  return guarded([&]() -> int {
    if (!h) throw ArgError("null handle");
    if (!h->m.assembled) h->m.assemble();
    h->m.tune_stiffness();
    if (h->m.all_fixed()) return ML_OK;  // nothing to tune
    return h->m.op.autotune(nrhs, chosen);
  });
}

void ml_model_destroy(ml_model* h) { delete h; }

// Profiling hooks for the benchmark: accumulate kernel durations measured with HIP events on
// the handle's stream. kinds: 0 spmv64, 1 spmv32, 2 precondition64 (fwd+bwd), 3
// precondition32, 4 ic0, 5 stiffness or elasto-plastic response, 6 assembly, 7 vector updates,
// 8 Gauss-point stresses,
// 9 nodal averaging.
ml_status ml_model_profile(ml_model* h, int enable, double* ms, int64_t* counts) {
  return guarded([&]() -> int {
    if (!h) throw ArgError("null handle");
    Profiler& p = h->m.op.prof;
    p.resolve();
    if (ms)
      for (int k = 0; k < P_COUNT; ++k) ms[k] = p.ms[k];
    if (counts)
      for (int k = 0; k < P_COUNT; ++k) counts[k] = p.count[k];
    if (enable >= 0) {
      p.reset();
      p.on = enable != 0;
    }
    return ML_OK;
  });
}

// ------------------------------------------------------------------ Gmsh .msh input
ml_status ml_msh_read(const char* text, int64_t length, ml_msh** out, int32_t* error_kind,
                      int64_t* error_line) {
  if (error_kind) *error_kind = 0;
  if (error_line) *error_line = 0;
  return guarded([&]() -> int {
    if (!out || length < 0 || (length > 0 && !text)) throw ArgError("bad msh arguments");
    *out = nullptr;
    try {
      auto* h = new ml_msh;
      try {
        h->f = msh::read(text, (size_t)length);
      } catch (...) {
        delete h;
        throw;
      }
      *out = h;
    } catch (const msh::Error& e) {
      if (error_kind) *error_kind = e.kind;
      if (error_line) *error_line = e.line;
      g_last_error = e.line > 0 ? "line " + std::to_string(e.line) + ": " + e.what() : e.what();
      return ML_INVALID_FILE;
    }
    return ML_OK;
  });
}

ml_status ml_msh_sizes(const ml_msh* m, int64_t* sizes) {
  if (!m || !sizes) return ML_INVALID_ARGUMENT;
  sizes[0] = m->f.n_nodes();
  sizes[1] = m->f.nen;
  sizes[2] = m->f.n_cells();
  sizes[3] = (int64_t)m->f.sets.size();
  sizes[4] = m->f.dim;
  return ML_OK;
}

ml_status ml_msh_geometry(const ml_msh* m, double* xyz, int64_t* cells) {
  if (!m) return ML_INVALID_ARGUMENT;
  if (xyz) std::copy(m->f.xyz.begin(), m->f.xyz.end(), xyz);
  if (cells) std::copy(m->f.cells.begin(), m->f.cells.end(), cells);
  return ML_OK;
}

static const msh::Set* msh_set(const ml_msh* m, int64_t k, const std::string** name) {
  if (!m || k < 0 || k >= (int64_t)m->f.sets.size()) return nullptr;
  auto it = m->f.sets.begin();
  std::advance(it, k);
  if (name) *name = &it->first;
  return &it->second;
}

ml_status ml_msh_set_info(const ml_msh* m, int64_t k, char* name, int64_t capacity,
                          int64_t* counts) {
  const std::string* nm = nullptr;
  const msh::Set* s = msh_set(m, k, &nm);
  if (!s) {
    g_last_error = "no such set";
    return ML_INVALID_ARGUMENT;
  }
  if (counts) {
    counts[0] = (int64_t)s->nodes.size();
    counts[1] = (int64_t)s->faces.size();
    counts[2] = (int64_t)s->edges.size();
    counts[3] = (int64_t)s->elements.size();
    counts[4] = (int64_t)nm->size();
  }
  if (name) {
    if (capacity < (int64_t)nm->size() + 1) {
      g_last_error = "set name buffer too small";
      return ML_INVALID_ARGUMENT;
    }
    std::memcpy(name, nm->data(), nm->size());
    name[nm->size()] = 0;
  }
  return ML_OK;
}

ml_status ml_msh_set_items(const ml_msh* m, int64_t k, int64_t* nodes, int64_t* faces,
                           int64_t* edges, int64_t* elements) {
  const msh::Set* s = msh_set(m, k, nullptr);
  if (!s) {
    g_last_error = "no such set";
    return ML_INVALID_ARGUMENT;
  }
  if (nodes) std::copy(s->nodes.begin(), s->nodes.end(), nodes);
  if (elements) std::copy(s->elements.begin(), s->elements.end(), elements);
  if (faces)
    for (size_t i = 0; i < s->faces.size(); ++i) {
      faces[2 * i] = s->faces[i].first;
      faces[2 * i + 1] = s->faces[i].second;
    }
  if (edges)
    for (size_t i = 0; i < s->edges.size(); ++i) {
      edges[2 * i] = s->edges[i].first;
      edges[2 * i + 1] = s->edges[i].second;
    }
  return ML_OK;
}

void ml_msh_destroy(ml_msh* m) { delete m; }

}  // extern "C"
