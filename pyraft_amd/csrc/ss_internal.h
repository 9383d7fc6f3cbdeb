// ss_internal.h -- declarations shared by the host-side translation units of libraft
// (plans and operators in ss_api.cpp, batches and the autotuner in ss_batch.cpp).
#ifndef PYRAFT_AMD_SS_INTERNAL_H
#define PYRAFT_AMD_SS_INTERNAL_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <map>
#include <mutex>
#include <thread>
#include <utility>
#include <vector>

#include "../../include/raft_slantstack.h"
#include "ss_kernels.h"

// ------------------------------------------------------------------ errors

// records the message for raft_ss_last_error() on this thread and returns code
int ss_set_error(int code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));

#define HIP_TRY(expr)                                                                                     \
  do {                                                                                                    \
    hipError_t e_ = (expr);                                                                               \
    if (e_ != hipSuccess)                                                                                 \
      return ss_set_error(RAFT_SS_EHIP, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), __FILE__, \
                          __LINE__);                                                                      \
  } while (0)

// ------------------------------------------------------------------ plan

struct raft_ss_plan;
// Backprojection variant for launches over angles [j0, j1) (OSEM subsets): the fastest
// variant of a timed plain launch over that range, cached per process by device,
// geometry and range length.  The plan's own variant when it was set explicitly, when
// the range is empty, or while `stream` is being captured and no earlier call has
// timed this range (no timing under capture).
int ss_plan_range_variant(raft_ss_plan* p, int j0, int j1, hipStream_t stream, int* variant);
// An OSEM run split into its choices (ss_em_prepare: slab rows per subset, the subsets'
// backprojection variant, workspace bytes; may time launches) and its kernels
// (ss_em_launch: launches only, capturable; LL of iteration it to d_ll[it], or to the
// workspace when d_ll is null).
struct SSEmLaunch {
  int S = 1;
  SSDevPlan dp{}, dpb{};
  std::vector<SSDevPlan> dps;  // per subset: the Radon with its slab choice
  std::vector<int> poff;       // LL partial offsets of the subsets (S + 1)
  // byte offsets in the workspace (the gather layouts start it) and its size
  size_t meas_off = 0, rho_off = 0, weights_off = 0, partials_off = 0, ll_off = 0, slabs_off = 0, bp_off = 0;
  size_t bytes = 0;
};
int ss_em_prepare(const raft_ss_plan* p, int nsubsets, int nslices, int niter, hipStream_t stream, SSEmLaunch* L);
int ss_em_launch(const raft_ss_plan* p, const SSEmLaunch& L, const float* d_sino, float* d_x, int niter,
                 int nslices, void* ws, double* d_ll, hipStream_t stream);
// Launch choice of the plan's Radon over angle ranges of j1 - j0 angles on nslices slices:
// the block shape and the sample rows per slab (SSDevPlan::radon_shape, radon_slab_rows),
// set in *dp.  Candidates: both shapes without row slabs and, where the default rule
// divides the launch, both shapes at the rule's slab rows and shape 0 at twice and half of
// them.  The fastest is timed once per process (gather layouts built first, not timed)
// and cached with the backprojection choices (mesh plans: per plan only, since the cache
// key does not name their tables); no timing while `stream` is captured, and
// with `may_time` false only the plan's earlier choice or the rule with shape 0 (no host
// allocation).  Rows pinned by raft_ss_plan_set_radon_slab_rows apply to every launch,
// with shape 0.
int ss_plan_radon_rows(const raft_ss_plan* p, int j0, int j1, int nslices, hipStream_t stream, bool may_time,
                       SSDevPlan* dp);

struct raft_ss_plan {
  int device = 0;
  SSGeom g{};
  double img_corners[4] = {0, 0, 0, 0};   // tl_x, tl_y, br_x, br_y
  double sino_corners[4] = {0, 0, 0, 0};
  std::vector<SSAngle> h_angles;
  SSAngle* d_angles = nullptr;
  int2* d_bounds = nullptr;
  int4* d_wbounds = nullptr;
  int* d_irregular = nullptr;
  double* d_ray_t = nullptr;       // mesh plans: detector coordinates [R]
  int64_t* d_ray_alpha = nullptr;  // mesh plans: trajectory offsets [A][R]
  bool mesh = false;               // arbitrary (theta, t) tables: Radon only
  int nirregular = 0;
  int ndegenerate = 0;
  int nwaves = 0;
  int bp_variant = SS_BP_DEFAULT_VARIANT;
  bool bp_variant_fixed = false;  // set by raft_ss_plan_set_bp_variant: no per-range tuning
  int bp_window_v[SS_BP_NVARIANTS] = {0};  // per backprojection variant: staged window
  int bp_angles_v[SS_BP_NVARIANTS] = {0};  //   and angles per stage (0: not runnable)
  int rt_variant = -1;                     // pixel-driven transpose: tile variant (-1: line form)
  int rt_window_v[SS_BP_NVARIANTS] = {0};  //   its staged window and angles per stage per variant
  int rt_angles_v[SS_BP_NVARIANTS] = {0};
  float* d_ramp = nullptr;  // FBP convolution kernel, built on first use
  void* d_ramp_fft = nullptr;  // its FFT-form spectrum and twiddles (if R fits the FFT form)
  SSFftTaps ramp_fft{};
  std::mutex ramp_mutex;
  int radon_rows_fixed = 0;  // raft_ss_plan_set_radon_slab_rows (0: tuned / rule)
  mutable std::mutex radon_mutex;
  mutable std::map<long long, int> radon_rows;  // (range length, slices) -> tuned slab rows
  SSDevPlan dev() const { return dev(bp_variant); }
  // the plan as the kernels see it, with backprojection variant v
  SSDevPlan dev(int v) const {
    SSDevPlan p;
    p.g = g;
    p.angles = d_angles;
    p.bounds = d_bounds;
    p.wbounds = d_wbounds;
    p.irregular = d_irregular;
    p.ray_t = d_ray_t;
    p.ray_alpha = d_ray_alpha;
    p.nirregular = nirregular;
    p.nwaves = nwaves;
    p.bp_variant = v;
    p.bp_window = bp_window_v[v];
    p.bp_angles = bp_angles_v[v];
    p.radon_slab_rows = 0;
    p.radon_shape = 0;
    p.rt_variant = rt_variant >= 0 ? rt_variant : 0;
    p.rt_window = rt_variant >= 0 ? rt_window_v[rt_variant] : 0;
    p.rt_angles = rt_variant >= 0 ? rt_angles_v[rt_variant] : 0;
    return p;
  }
  ~raft_ss_plan() {
    if (d_angles) (void)hipFree(d_angles);
    if (d_bounds) (void)hipFree(d_bounds);
    if (d_wbounds) (void)hipFree(d_wbounds);
    if (d_irregular) (void)hipFree(d_irregular);
    if (d_ray_t) (void)hipFree(d_ray_t);
    if (d_ray_alpha) (void)hipFree(d_ray_alpha);
    if (d_ramp_fft) (void)hipFree(d_ramp_fft);
    if (d_ramp) (void)hipFree(d_ramp);
  }
};

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

// status of a drop-in (void) symbol: kept for raft_ss_legacy_status(), printed on failure
void ss_report_legacy(const char* fn, int rc);

// FP32 circular-convolution kernel of the numpy FBP filters (kind RAFT_SS_FILTER_*)
void ss_filter_kernel(int R, double tmin, double tmax, int kind, double param, float* g);

// ------------------------------------------------------------------ host <-> raft_image staging

// Pinned staging buffers reused across calls: pinning tens of MB costs milliseconds, more
// than the copies.  Released buffers go back to the pool, which keeps the 4 largest; the
// pool itself is never destroyed, so no HIP call runs at process exit.  A release first
// waits for the device, as hipHostFree does, so that no copy into or out of the buffer
// is in flight when another call takes it (the error paths return before their own wait).
struct SSPinnedPool {
  std::mutex mu;
  std::vector<std::pair<void*, size_t>> free;
};
inline SSPinnedPool& ss_pinned_pool() {
  static SSPinnedPool* pool = new SSPinnedPool();
  return *pool;
}

struct PinnedF32 {
  float* p = nullptr;
  size_t bytes = 0;
  PinnedF32() = default;
  PinnedF32(const PinnedF32&) = delete;
  PinnedF32& operator=(const PinnedF32&) = delete;
  ~PinnedF32() {
    if (!p) return;
    (void)hipDeviceSynchronize();
    SSPinnedPool& pool = ss_pinned_pool();
    std::lock_guard<std::mutex> lock(pool.mu);
    pool.free.emplace_back(p, bytes);
    if (pool.free.size() > 4) {
      auto it = std::min_element(pool.free.begin(), pool.free.end(),
                                 [](const std::pair<void*, size_t>& a, const std::pair<void*, size_t>& b) {
                                   return a.second < b.second;
                                 });
      (void)hipHostFree(it->first);
      pool.free.erase(it);
    }
  }
  // n floats: the smallest pooled buffer that holds them, else a new one
  hipError_t alloc(size_t n) {
    const size_t need = n * sizeof(float) + 4;
    SSPinnedPool& pool = ss_pinned_pool();
    {
      std::lock_guard<std::mutex> lock(pool.mu);
      auto best = pool.free.end();
      for (auto it = pool.free.begin(); it != pool.free.end(); ++it)
        if (it->second >= need && (best == pool.free.end() || it->second < best->second)) best = it;
      if (best != pool.free.end()) {
        p = static_cast<float*>(best->first);
        bytes = best->second;
        pool.free.erase(best);
        return hipSuccess;
      }
    }
    void* q = nullptr;
    const hipError_t e = hipHostMalloc(&q, need, 0);
    if (e != hipSuccess) return e;
    p = static_cast<float*>(q);
    bytes = need;
    return hipSuccess;
  }
};

// f(i0, i1) over rows [0, lines) in contiguous parts on up to 16 threads (the host's
// hardware threads); matrices under 2^20 elements stay on the calling thread, and parts
// whose thread cannot start run on it too
template <class F>
inline void ss_for_rows(int lines, size_t cols, F f) {
  const unsigned hw = std::thread::hardware_concurrency();
  int t = (int)std::min<unsigned>(hw ? hw : 1u, 16u);
  if ((size_t)lines * cols < ((size_t)1 << 20)) t = 1;
  t = std::max(1, std::min(t, lines));
  const int per = (lines + t - 1) / t;
  std::vector<std::thread> th;
  int a = per;
  for (; a < lines; a += per) {
    try {
      th.emplace_back(f, a, std::min(lines, a + per));
    } catch (...) {
      break;
    }
  }
  f(0, std::min(lines, per));
  if (a < lines) f(a, lines);
  for (auto& x : th) x.join();
}

// strided FP64 matrix -> dense row-major FP32
inline void gather_f64(const raft_matrix& m, float* dst) {
  ss_for_rows(m.lines, (size_t)m.columns, [&m, dst](int i0, int i1) {
    for (int i = i0; i < i1; ++i)
      for (int j = 0; j < m.columns; ++j) dst[(size_t)i * m.columns + j] = (float)RAFT_MATRIX_AT(m, i, j);
  });
}

inline void scatter_f64(const float* src, const raft_matrix& m) {
  ss_for_rows(m.lines, (size_t)m.columns, [src, &m](int i0, int i1) {
    for (int i = i0; i < i1; ++i)
      for (int j = 0; j < m.columns; ++j) RAFT_MATRIX_AT(m, i, j) = (double)src[(size_t)i * m.columns + j];
  });
}

inline void corners_of(const raft_image& im, double* c) {
  c[0] = im.tl_x;
  c[1] = im.tl_y;
  c[2] = im.br_x;
  c[3] = im.br_y;
}

// ------------------------------------------------------------------ workspace

// stream-ordered scratch allocation, freed on the same stream
struct DevBuf {
  void* p = nullptr;
  hipStream_t s = nullptr;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() {
    if (p) (void)hipFreeAsync(p, s);
  }
  hipError_t alloc(size_t bytes, hipStream_t stream) {
    s = stream;
    return hipMallocAsync(&p, bytes ? bytes : 4, stream);
  }
  template <class T>
  T* as() const {
    return reinterpret_cast<T*>(p);
  }
};

// Scratch of one call of a device-pointer operator.  An eager call takes the cached
// workspace of its device and stream: grow-only, reused by every later call on that
// stream, since calls on one stream run in stream order and never overlap; a larger
// request releases the old buffer on the stream (after the work that used it).  The call
// holds the workspace's lock while it enqueues, so host threads sharing a stream take
// turns.  A call made while its stream is being captured allocates stream-ordered memory
// of its own instead (released at the end of the call, the graph holding the allocation),
// so that no graph refers to a cached buffer that a later call may release.
hipError_t ss_stream_scratch(hipStream_t stream, size_t bytes, void** p, std::unique_lock<std::mutex>* lock);
struct Scratch {
  DevBuf own;
  void* p = nullptr;
  std::unique_lock<std::mutex> lock;
  hipError_t alloc(size_t bytes, hipStream_t stream) {
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    hipError_t e = hipStreamIsCapturing(stream, &cs);
    if (e != hipSuccess) return e;
    if (cs != hipStreamCaptureStatusNone) {
      e = own.alloc(bytes, stream);
      p = own.p;
      return e;
    }
    return ss_stream_scratch(stream, bytes, &p, &lock);
  }
  template <class T>
  T* as() const {
    return reinterpret_cast<T*>(p);
  }
};

inline int check_plan(const raft_ss_plan* p, int nslices) {
  if (!p) return ss_set_error(RAFT_SS_EINVAL, "null plan");
  if (nslices < 1) return ss_set_error(RAFT_SS_EINVAL, "nslices must be >= 1 (got %d)", nslices);
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev != p->device)
    return ss_set_error(RAFT_SS_EINVAL, "plan was built on device %d but device %d is current", p->device, dev);
  return 0;
}

// operators other than the Radon need the uniform detector grid
inline int check_uniform(const raft_ss_plan* p, const char* what) {
  if (p && p->mesh) return ss_set_error(RAFT_SS_EINVAL, "%s needs a uniform-grid plan (mesh plans do the Radon only)", what);
  return 0;
}

inline int check_range(const raft_ss_plan* p, int j0, int j1) {
  if (j0 < 0 || j1 > p->g.nangles || j0 > j1)
    return ss_set_error(RAFT_SS_EINVAL, "angle range [%d, %d) outside [0, %d)", j0, j1, p->g.nangles);
  return 0;
}

#endif
