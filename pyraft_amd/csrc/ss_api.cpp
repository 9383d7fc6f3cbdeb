// ss_api.cpp -- the libraft C ABI: plans, device-pointer entry points and the
// reference-compatible host-pointer entry points.
//
// Every function returns 0 on success or a negative RAFT_SS_E* code; the message
// of the last failure on the calling thread is available from raft_ss_last_error().
// The legacy `void` symbols print that message to stderr instead.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdarg.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <cstdlib>
#include <list>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "ss_internal.h"
#include "ss_plan_logic.h"

// ------------------------------------------------------------------ errors

static thread_local std::string g_last_error;

int ss_set_error(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_last_error = buf;
  return code;
}

#define set_error ss_set_error

extern "C" const char* raft_ss_last_error(void) { return g_last_error.c_str(); }

// ------------------------------------------------------------------ per-stream scratch

namespace {
struct StreamWorkspace {
  std::mutex mu;
  void* p = nullptr;
  size_t bytes = 0;
};
std::mutex g_ws_mu;
// (device, stream) -> workspace; never destroyed, so no HIP call runs at process exit
std::map<std::pair<int, hipStream_t>, StreamWorkspace*>* g_ws = new std::map<std::pair<int, hipStream_t>, StreamWorkspace*>();
}  // namespace

hipError_t ss_stream_scratch(hipStream_t stream, size_t bytes, void** p, std::unique_lock<std::mutex>* lock) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  StreamWorkspace* w;
  {
    std::lock_guard<std::mutex> g(g_ws_mu);
    StreamWorkspace*& slot = (*g_ws)[std::make_pair(dev, stream)];
    if (!slot) slot = new StreamWorkspace();
    w = slot;
  }
  *lock = std::unique_lock<std::mutex>(w->mu);
  if (bytes > w->bytes) {
    if (w->p) {
      e = hipFreeAsync(w->p, stream);
      w->p = nullptr;
      w->bytes = 0;
      if (e != hipSuccess) return e;
    }
    const size_t want = std::max(bytes, (size_t)1 << 20);
    e = hipMallocAsync(&w->p, want, stream);
    if (e != hipSuccess) return e;
    w->bytes = want;
  }
  *p = w->p;
  return hipSuccess;
}

// Geometry in FP64 exactly as the reference derives it from the corners.
static SSGeom make_geom(int n, int nrays, int nangles, const double* ic, const double* sc) {
  SSGeom g;
  g.n = n;
  g.nrays = nrays;
  g.nangles = nangles;
  g.pitch = (n + 2 + 15) & ~15;  // n + 2 rounded up to whole 128-byte lines of records
  g.xmin = ic[0];
  g.ymin = ic[3];
  g.dx = (ic[2] - ic[0]) / (n - 1);   // raft/raft_radon_slantstack.c:150
  g.dy = (ic[1] - ic[3]) / (n - 1);   // :151
  g.tmin = sc[3];
  g.dt = (sc[1] - sc[3]) / (nrays - 1);  // :170
  g.th0 = sc[0];
  g.dth = (sc[2] - sc[0]) / (nangles);   // :156
  return g;
}

static int64_t to_fix(double v) { return (int64_t)llrint(v * SS_FIX_ONE); }

// Host-side angle table (glibc cos/sin/tan, as the reference calls them).  `thetas`
// (optional) gives the angles of a mesh plan (raft/raft_radon_slantstack_withmesh.c:193).
static void build_angles(const SSGeom& g, std::vector<SSAngle>& out, int* ndeg, const double* thetas = nullptr) {
  const double tol = 1.0 / sqrt(2.0);  // raft/raft_radon_slantstack.c:139
  out.resize(g.nangles);
  *ndeg = 0;
  for (int j = 0; j < g.nangles; ++j) {
    SSAngle a;
    memset(&a, 0, sizeof(a));
    const double theta = thetas ? thetas[j] : g.th0 + j * g.dth;  // :158
    a.cost = cos(theta);
    a.sint = sin(theta);
    a.tant = tan(theta);
    const double fs = fabs(a.sint), fc = fabs(a.cost);
    a.branch_a = fc < tol ? 1 : 0;  // :174
    a.delta = (float)(a.branch_a ? (g.dy / fs) : (g.dx / fc));
    const SSTraj tr = ss_trajectory(g, a);
    a.alpha0 = to_fix(tr.alpha0);
    a.gamma = to_fix(tr.gamma);
    a.beta = to_fix(tr.beta);
    a.degenerate = fabs(tr.beta) < SS_DEGENERATE_SLOPE ? 1 : 0;
    a.regular = !a.degenerate;
    *ndeg += a.degenerate;
    a.bp_a = (g.xmin * a.cost + g.ymin * a.sint - g.tmin) / g.dt;
    a.bp_b = (g.dx * a.cost) / g.dt;
    a.bp_c = (g.dy * a.sint) / g.dt;
    a.rt_ig = a.gamma != 0 ? SS_FIX_ONE / (double)a.gamma : 0.0;
    a.rt_a0 = a.gamma != 0 ? (double)a.alpha0 / (double)a.gamma : 0.0;
    a.rt_b = a.gamma != 0 ? (double)a.beta / (double)a.gamma : 0.0;
    a.rt_w = fmin(-a.rt_ig, g.n * a.rt_ig);
    a.rt_wn = (int)ceil((g.n + 1) * fabs(a.rt_ig));
    a.rt_i0 = fmin(a.rt_ig, (g.n - 2) * a.rt_ig);
    a.rt_i1 = fmax(a.rt_ig, (g.n - 2) * a.rt_ig);
    out[j] = a;
  }
}

static int validate_geom(const SSGeom& g, const std::vector<SSAngle>& angles) {
  if (g.n < 2) return set_error(RAFT_SS_EINVAL, "image side must be >= 2 (got %d)", g.n);
  if (g.nrays < 2) return set_error(RAFT_SS_EINVAL, "nrays must be >= 2 (got %d)", g.nrays);
  if (g.nangles < 1) return set_error(RAFT_SS_EINVAL, "nangles must be >= 1 (got %d)", g.nangles);
  if (g.n > 16384 || g.nrays > (1 << 24) || g.nangles > (1 << 24))
    return set_error(RAFT_SS_EINVAL, "geometry too large (n=%d R=%d A=%d)", g.n, g.nrays, g.nangles);
  const double vals[] = {g.xmin, g.ymin, g.dx, g.dy, g.tmin, g.dt, g.th0, g.dth};
  for (double v : vals)
    if (!std::isfinite(v)) return set_error(RAFT_SS_EINVAL, "non-finite geometry (check the corners)");
  if (g.dx == 0.0 || g.dy == 0.0 || g.dt == 0.0)
    return set_error(RAFT_SS_EINVAL, "zero-extent field of view (dx=%g dy=%g dt=%g)", g.dx, g.dy, g.dt);
  // the Q32.32 trajectory must stay far inside int64
  for (const SSAngle& a : angles) {
    const double amax = fabs((double)a.alpha0) + fabs((double)a.gamma) * g.nrays + fabs((double)a.beta) * g.n;
    if (!(amax < 4.0e18))
      return set_error(RAFT_SS_EINVAL, "ray coordinates out of fixed-point range; unsupported field of view");
  }
  return 0;
}

// Backprojection staging for kernel variant v: rays a 32 x (8*rows) pixel tile can
// touch at one angle (+margins for the floor of the tile minimum and the second tap),
// and angles per LDS stage.  A variant whose window does not fit its LDS stage gets
// per_stage = 0 (not runnable for this geometry).
static void bp_staging(const std::vector<SSAngle>& angles, int v, int* window, int* per_stage) {
  const double th = 8.0 * kSSBpRows[v] - 1.0;
  double span = 0.0;
  for (const SSAngle& a : angles) span = std::max(span, 31.0 * fabs(a.bp_b) + th * fabs(a.bp_c));
  int W = (int)ceil(span) + 5;
  W = (W + 3) & ~3;
  *window = W;
  *per_stage = W <= kSSBpStage[v] ? std::min(64, kSSBpStage[v] / W) : 0;
}

// Pixel-driven transpose staging for tile variant v: rays a tile can reach at one angle
// (kappa per column and per row from the angle's transpose constants, + margins for the
// floor of the tile minimum, the second candidate and FP32 rounding) and angles per stage;
// 0 angles when some angle's rays are closer than one cell (the form needs >= 1).
static void rt_staging(const std::vector<SSAngle>& angles, int v, int* window, int* per_stage) {
  const double th = 8.0 * kSSBpRows[v] - 1.0;
  double span = 0.0;
  for (const SSAngle& a : angles) {
    if (!(fabs((double)a.gamma) >= SS_FIX_ONE)) {
      *window = 0;
      *per_stage = 0;
      return;
    }
    const double kb = a.branch_a ? a.rt_b : a.rt_ig, kc = a.branch_a ? a.rt_ig : a.rt_b;
    span = std::max(span, 31.0 * fabs(kb) + th * fabs(kc));
  }
  int W = (int)ceil(span) + 5;
  W = (W + 3) & ~3;
  *window = W;
  // the stage holds 2 STAGE floats: per angle a window of W + 1 rays, W + 4 floats apart
  *per_stage = W + 4 <= 2 * kSSBpStage[v] ? std::min(64, 2 * kSSBpStage[v] / (W + 4)) : 0;
}

// all variants' staging; the default variant, or the first runnable one
static int bp_variants(raft_ss_plan* p) {
  p->rt_variant = -1;
  for (int v = 0; v < SS_BP_NVARIANTS; ++v) rt_staging(p->h_angles, v, &p->rt_window_v[v], &p->rt_angles_v[v]);
  for (int v : {SS_RT_DEFAULT_VARIANT, 0, 1, 2, 4, 5})
    if (p->rt_variant < 0 && p->rt_angles_v[v] > 0) p->rt_variant = v;
  int first = -1;
  for (int v = 0; v < SS_BP_NVARIANTS; ++v) {
    bp_staging(p->h_angles, v, &p->bp_window_v[v], &p->bp_angles_v[v]);
    if (p->bp_angles_v[v] > 0 && first < 0) first = v;
  }
  if (first < 0)
    return set_error(RAFT_SS_EINVAL, "too many detector bins per pixel (window %d)",
                     p->bp_window_v[SS_BP_DEFAULT_VARIANT]);
  p->bp_variant = p->bp_angles_v[SS_BP_DEFAULT_VARIANT] > 0 ? SS_BP_DEFAULT_VARIANT : first;
  // an image with fewer than 4 default tiles per CU starts from the 2-rows-per-thread
  // tiles (twice the blocks); the autotuner may still pick another variant
  const int kSmallVariant = 1;  // 2 pixel rows per thread, 2048 staged pairs
  if (p->bp_variant == SS_BP_DEFAULT_VARIANT && kSSBpRows[kSmallVariant] == 2 && p->bp_angles_v[kSmallVariant] > 0 &&
      ss_bp_tiles(p->g.n, SS_BP_DEFAULT_VARIANT) < 1024)
    p->bp_variant = kSmallVariant;
  return 0;
}

// Mesh plans: every trajectory offset must stay far inside Q32.32's int64 range.
static int validate_mesh(const SSGeom& g, const std::vector<SSAngle>& angles, const double* t) {
  double tmax = 0.0;
  for (int k = 0; k < g.nrays; ++k) {
    if (!std::isfinite(t[k])) return set_error(RAFT_SS_EINVAL, "non-finite detector coordinate t[%d]", k);
    tmax = std::max(tmax, fabs(t[k]));
  }
  for (const SSAngle& a : angles) {
    const double u = std::max(fabs(ss_ray_alpha(g, a, tmax)), fabs(ss_ray_alpha(g, a, -tmax)));
    const double amax = u * SS_FIX_ONE + fabs((double)a.beta) * g.n;
    if (!(amax < 4.0e18))
      return set_error(RAFT_SS_EINVAL, "ray coordinates out of fixed-point range; unsupported mesh");
  }
  return 0;
}

static int plan_create(raft_ss_plan** out, int n, int nrays, int nangles, const double* img_corners,
                       const double* sino_corners, const double* thetas, const double* ts, void* stream_) {
  if (!out || !img_corners || !sino_corners) return set_error(RAFT_SS_EINVAL, "null argument");
  *out = nullptr;
  hipStream_t stream = as_stream(stream_);
  std::unique_ptr<raft_ss_plan> p(new raft_ss_plan());
  HIP_TRY(hipGetDevice(&p->device));
  memcpy(p->img_corners, img_corners, sizeof(p->img_corners));
  memcpy(p->sino_corners, sino_corners, sizeof(p->sino_corners));
  p->g = make_geom(n, nrays, nangles, img_corners, sino_corners);
  p->mesh = ts != nullptr;
  if (n < 2 || nrays < 2 || nangles < 1) return validate_geom(p->g, p->h_angles);
  build_angles(p->g, p->h_angles, &p->ndegenerate, thetas);
  int rc = validate_geom(p->g, p->h_angles);
  if (rc) return rc;
  if (p->mesh && (rc = validate_mesh(p->g, p->h_angles, ts))) return rc;

  const SSGeom& g = p->g;
  p->nwaves = (nrays + 63) / 64;
  if (!p->mesh && (rc = bp_variants(p.get()))) return rc;

  const int capacity = (p->ndegenerate + 2) * nrays + 4096;
  HIP_TRY(hipMalloc(&p->d_angles, sizeof(SSAngle) * nangles));
  HIP_TRY(hipMalloc(&p->d_bounds, sizeof(int2) * (size_t)nangles * nrays));
  HIP_TRY(hipMalloc(&p->d_wbounds, sizeof(int4) * (size_t)nangles * p->nwaves));
  HIP_TRY(hipMalloc(&p->d_irregular, sizeof(int) * (capacity + 1)));
  if (p->mesh) {
    HIP_TRY(hipMalloc(&p->d_ray_t, sizeof(double) * nrays));
    HIP_TRY(hipMalloc(&p->d_ray_alpha, sizeof(int64_t) * (size_t)nangles * nrays));
    HIP_TRY(hipMemcpyAsync(p->d_ray_t, ts, sizeof(double) * nrays, hipMemcpyHostToDevice, stream));
  }
  int* d_count = p->d_irregular + capacity;
  HIP_TRY(hipMemcpyAsync(p->d_angles, p->h_angles.data(), sizeof(SSAngle) * nangles, hipMemcpyHostToDevice, stream));
  HIP_TRY(hipMemsetAsync(d_count, 0, sizeof(int), stream));
  HIP_TRY(ss_launch_plan_bounds(g, p->d_angles, p->d_bounds, p->d_irregular, d_count, capacity, p->d_ray_t,
                                p->d_ray_alpha, stream));
  HIP_TRY(ss_launch_plan_wave_bounds(g, p->d_bounds, p->d_wbounds, stream));
  int count = 0;
  HIP_TRY(hipMemcpyAsync(&count, d_count, sizeof(int), hipMemcpyDeviceToHost, stream));
  HIP_TRY(hipStreamSynchronize(stream));
  if (count > capacity) return set_error(RAFT_SS_EINTERNAL, "irregular ray list overflow (%d > %d)", count, capacity);
  p->nirregular = count;
  if (count > 0) {
    // angles with irregular rays are not regular
    std::vector<int> irr(count);
    HIP_TRY(hipMemcpy(irr.data(), p->d_irregular, sizeof(int) * count, hipMemcpyDeviceToHost));
    for (int id : irr) p->h_angles[id / nrays].regular = 0;
    HIP_TRY(hipMemcpy(p->d_angles, p->h_angles.data(), sizeof(SSAngle) * nangles, hipMemcpyHostToDevice));
  }
  *out = p.release();
  return 0;
}

extern "C" int raft_ss_plan_create(raft_ss_plan** out, int n, int nrays, int nangles, const double* img_corners,
                                   const double* sino_corners, void* stream) {
  return plan_create(out, n, nrays, nangles, img_corners, sino_corners, nullptr, nullptr, stream);
}

extern "C" int raft_ss_plan_create_mesh(raft_ss_plan** out, int n, const double* img_corners, int nrays,
                                        const double* t, int nangles, const double* theta, void* stream) {
  if (!t || !theta) return set_error(RAFT_SS_EINVAL, "null argument");
  if (nrays < 2 || nangles < 1) return set_error(RAFT_SS_EINVAL, "a mesh needs nrays >= 2 and nangles >= 1");
  // the kernels take t and theta from the tables; the uniform-grid fields of the
  // geometry record are unused and get a valid placeholder extent
  const double sc[4] = {0.0, 1.0, 1.0, -1.0};
  return plan_create(out, n, nrays, nangles, img_corners, sc, theta, t, stream);
}

extern "C" int raft_ss_plan_destroy(raft_ss_plan* plan) {
  delete plan;
  return 0;
}

extern "C" int raft_ss_plan_info(const raft_ss_plan* p, int* info, int ninfo) {
  if (!p || !info) return set_error(RAFT_SS_EINVAL, "null argument");
  const int vals[] = {p->g.n,
                      p->g.nrays,
                      p->g.nangles,
                      p->nirregular,
                      p->ndegenerate,
                      p->bp_window_v[p->bp_variant],
                      p->bp_angles_v[p->bp_variant],
                      p->bp_variant,
                      p->mesh ? 1 : ss_bp_groups(p->dev(), 0, p->g.nangles, 1)};
  for (int i = 0; i < ninfo && i < (int)(sizeof(vals) / sizeof(vals[0])); ++i) info[i] = vals[i];
  return 0;
}

extern "C" int raft_ss_plan_bounds(const raft_ss_plan* p, int* host_bounds) {
  if (!p || !host_bounds) return set_error(RAFT_SS_EINVAL, "null argument");
  HIP_TRY(hipMemcpy(host_bounds, p->d_bounds, sizeof(int2) * (size_t)p->g.nangles * p->g.nrays,
                    hipMemcpyDeviceToHost));
  return 0;
}

extern "C" int raft_ss_plan_wave_bounds(const raft_ss_plan* p, int* host_wbounds) {
  if (!p || !host_wbounds) return set_error(RAFT_SS_EINVAL, "null argument");
  HIP_TRY(hipMemcpy(host_wbounds, p->d_wbounds, sizeof(int4) * (size_t)p->g.nangles * p->nwaves,
                    hipMemcpyDeviceToHost));
  return 0;
}

// Host evaluation of the same plan logic (no GPU needed): per-ray ranges in the
// layout of raft_ss_plan_bounds plus the per-angle Q32.32 constants, so tests can
// verify index ranges of every kernel before anything runs on a device.
extern "C" int raft_ss_plan_host_bounds(int n, int nrays, int nangles, const double* img_corners,
                                        const double* sino_corners, int* host_bounds, long long* host_fix,
                                        int* host_branch) {
  SSGeom g = make_geom(n, nrays, nangles, img_corners, sino_corners);
  std::vector<SSAngle> angles;
  int ndeg = 0;
  if (n < 2 || nrays < 2 || nangles < 1) return validate_geom(g, angles);
  build_angles(g, angles, &ndeg);
  int rc = validate_geom(g, angles);
  if (rc) return rc;
  const int nth = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  std::vector<std::thread> th;
  for (int t = 0; t < nth; ++t)
    th.emplace_back([&, t]() {
      for (int j = t; j < nangles; j += nth) {
        const SSAngle& a = angles[j];
        const SSTraj tr = ss_trajectory(g, a);
        for (int k = 0; k < nrays; ++k) {
          int lo, hi;
          if (ss_ray_range(g, a, tr, k, &lo, &hi)) {
            lo = 0;
            hi = -2;
          }
          host_bounds[2 * ((size_t)j * nrays + k)] = lo;
          host_bounds[2 * ((size_t)j * nrays + k) + 1] = hi;
        }
      }
    });
  for (auto& t : th) t.join();
  for (int j = 0; j < nangles; ++j) {
    if (host_fix) {
      host_fix[3 * j] = angles[j].alpha0;
      host_fix[3 * j + 1] = angles[j].gamma;
      host_fix[3 * j + 2] = angles[j].beta;
    }
    if (host_branch) host_branch[j] = angles[j].branch_a;
  }
  return 0;
}

// Host evaluation of a mesh plan's ranges and per-ray Q32.32 trajectory offsets, in the
// layouts of raft_ss_plan_bounds ([A][R] x (lo, hi)) and [A][R] int64.
extern "C" int raft_ss_plan_host_bounds_mesh(int n, const double* img_corners, int nrays, const double* t,
                                             int nangles, const double* theta, int* host_bounds,
                                             long long* host_alpha) {
  if (!img_corners || !t || !theta || !host_bounds) return set_error(RAFT_SS_EINVAL, "null argument");
  if (nrays < 2 || nangles < 1) return set_error(RAFT_SS_EINVAL, "a mesh needs nrays >= 2 and nangles >= 1");
  const double sc[4] = {0.0, 1.0, 1.0, -1.0};
  SSGeom g = make_geom(n, nrays, nangles, img_corners, sc);
  std::vector<SSAngle> angles;
  int ndeg = 0;
  if (n < 2) return validate_geom(g, angles);
  build_angles(g, angles, &ndeg, theta);
  int rc = validate_geom(g, angles);
  if (rc || (rc = validate_mesh(g, angles, t))) return rc;
  for (int j = 0; j < nangles; ++j) {
    const SSAngle& a = angles[j];
    const SSTraj tr = ss_trajectory(g, a);
    for (int k = 0; k < nrays; ++k) {
      const double u0 = ss_ray_alpha(g, a, t[k]);
      int lo, hi;
      if (ss_ray_range_t(g, a, tr.beta, t[k], u0, &lo, &hi)) {
        lo = 0;
        hi = -2;
      }
      host_bounds[2 * ((size_t)j * nrays + k)] = lo;
      host_bounds[2 * ((size_t)j * nrays + k) + 1] = hi;
      if (host_alpha) host_alpha[(size_t)j * nrays + k] = (long long)llrint(u0 * SS_FIX_ONE);
    }
  }
  return 0;
}

// Host-side plan summary without a device: info[0..4] = #degenerate angles,
// bp window, bp angles per stage, 64-ray groups, plan status.
extern "C" int raft_ss_plan_host_info(int n, int nrays, int nangles, const double* img_corners,
                                      const double* sino_corners, int* info) {
  if (!img_corners || !sino_corners || !info) return set_error(RAFT_SS_EINVAL, "null argument");
  SSGeom g = make_geom(n, nrays, nangles, img_corners, sino_corners);
  std::vector<SSAngle> angles;
  int ndeg = 0;
  if (n < 2 || nrays < 2 || nangles < 1) return validate_geom(g, angles);
  build_angles(g, angles, &ndeg);
  int rc = validate_geom(g, angles);
  if (rc) return rc;
  int W = 0, J = 0;
  bp_staging(angles, SS_BP_DEFAULT_VARIANT, &W, &J);
  if (J == 0) {
    int v = 0;
    for (; v < SS_BP_NVARIANTS; ++v) {
      bp_staging(angles, v, &W, &J);
      if (J > 0) break;
    }
    if (J == 0) return set_error(RAFT_SS_EINVAL, "too many detector bins per pixel (window %d)", W);
  }
  info[0] = ndeg;
  info[1] = W;
  info[2] = J;
  info[3] = (nrays + 63) / 64;
  info[4] = 0;
  return 0;
}

// Host-side backprojection staging of every kernel variant (no device needed):
// windows[v] = staged rays per angle, per_stage[v] = angles per stage (0: the window
// does not fit the variant's LDS stage).
extern "C" int raft_ss_plan_host_bp_staging(int n, int nrays, int nangles, const double* img_corners,
                                            const double* sino_corners, int* windows, int* per_stage) {
  if (!img_corners || !sino_corners || !windows || !per_stage) return set_error(RAFT_SS_EINVAL, "null argument");
  SSGeom g = make_geom(n, nrays, nangles, img_corners, sino_corners);
  std::vector<SSAngle> angles;
  int ndeg = 0;
  if (n < 2 || nrays < 2 || nangles < 1) return validate_geom(g, angles);
  build_angles(g, angles, &ndeg);
  int rc = validate_geom(g, angles);
  if (rc) return rc;
  for (int v = 0; v < SS_BP_NVARIANTS; ++v) bp_staging(angles, v, &windows[v], &per_stage[v]);
  return 0;
}

extern "C" int raft_ss_plan_host_rt_staging(int n, int nrays, int nangles, const double* img_corners,
                                            const double* sino_corners, int* windows, int* per_stage,
                                            double* table) {
  if (!img_corners || !sino_corners || !windows || !per_stage) return set_error(RAFT_SS_EINVAL, "null argument");
  SSGeom g = make_geom(n, nrays, nangles, img_corners, sino_corners);
  std::vector<SSAngle> angles;
  int ndeg = 0;
  if (n < 2 || nrays < 2 || nangles < 1) return validate_geom(g, angles);
  build_angles(g, angles, &ndeg);
  int rc = validate_geom(g, angles);
  if (rc) return rc;
  for (int v = 0; v < SS_BP_NVARIANTS; ++v) rt_staging(angles, v, &windows[v], &per_stage[v]);
  if (table)
    for (int j = 0; j < nangles; ++j) {
      const SSAngle& a = angles[j];
      double* t = table + 5 * (size_t)j;
      t[0] = a.rt_ig;
      t[1] = a.rt_a0;
      t[2] = a.rt_b;
      t[3] = a.branch_a;
      t[4] = fabs((double)a.gamma) / SS_FIX_ONE;
    }
  return 0;
}

namespace {
// g = Re(ifft(fftshift(h))) with h = |w| cos(pi w dt), w = linspace(-wc, wc, R), wc = 1/(2 dt)
// (pyraft/filters/filters.py:44-55), evaluated in FP64 by a direct real DFT.
void compute_ramp(int R, double tmin, double tmax, float* gk) { ss_filter_kernel(R, tmin, tmax, 0, 0.0, gk); }
}  // namespace

// The circular-convolution kernel g = Re(ifft(fftshift(h))) of the numpy FBP filters:
// kind 0, h = |w| cos(pi w dt) (lowpassino_fbp, pyraft/filters/filters.py:49-52); kind 1,
// h = |w| / (1 + L |w|^2) (lowpassino_reg_fbp, pyraft/filters/filters.py:93-96), with
// w = linspace(-wc, wc, R), wc = 1/(2 dt).  g is real because s is real: only the even
// part of the (non-Hermitian) fftshift(h) acts.
void ss_filter_kernel(int R, double tmin, double tmax, int kind, double param, float* gk) {
  // numpy.linspace(tmin, tmax, R) stores the endpoint exactly, so t[R-1] - t[0] = tmax - tmin
  const double dt = (tmax - tmin) / (R - 1);
  const double wc = 1 / (2 * dt);
  const double step = (wc - (-wc)) / (R - 1);
  std::vector<double> h(R), G(R);
  const double pi = 3.141592653589793;
  for (int i = 0; i < R; ++i) {
    const double w = (i == R - 1) ? wc : (i * step + (-wc));
    h[i] = kind == 1 ? fabs(w) / (1.0 + param * (fabs(w) * fabs(w))) : fabs(w) * cos(pi * w * dt);
  }
  const int shift = R / 2;  // numpy.fft.fftshift = roll by R//2
  for (int f = 0; f < R; ++f) G[f] = h[((f - shift) % R + R) % R];
  for (int m = 0; m < R; ++m) {
    double acc = 0.0;
    for (int f = 0; f < R; ++f) {
      const long long ph = ((long long)f * m) % R;
      acc += G[f] * cos(2.0 * pi * (double)ph / R);
    }
    gk[m] = (float)(acc / R);
  }
}

namespace {

int ensure_ramp(raft_ss_plan* p, hipStream_t stream) {
  std::lock_guard<std::mutex> lock(p->ramp_mutex);
  if (p->d_ramp) return 0;
  const int R = p->g.nrays;
  std::vector<float> gk(R);
  compute_ramp(R, p->sino_corners[3], p->sino_corners[1], gk.data());
  struct Owned {  // freed on every early return
    void* p = nullptr;
    ~Owned() {
      if (p) (void)hipFree(p);
    }
  } taps_mem, fft_mem;
  HIP_TRY(hipMalloc(&taps_mem.p, sizeof(float) * R));
  HIP_TRY(hipMemcpyAsync(taps_mem.p, gk.data(), sizeof(float) * R, hipMemcpyHostToDevice, stream));
  SSFftTaps taps{};
  if (const size_t bytes = ss_fft_taps_bytes(R)) {
    HIP_TRY(hipMalloc(&fft_mem.p, bytes));
    HIP_TRY(ss_fft_taps_build(static_cast<const float*>(taps_mem.p), R, false, fft_mem.p, &taps, stream));
  }
  HIP_TRY(hipStreamSynchronize(stream));
  p->d_ramp = static_cast<float*>(taps_mem.p);
  p->d_ramp_fft = fft_mem.p;
  p->ramp_fft = taps;
  taps_mem.p = fft_mem.p = nullptr;
  return 0;
}

// Sinograms of at most this many values are filtered straight from / to the (R, A)
// layout (strided rows; for small launches the layout copies cost more than the strides)
constexpr size_t kFilterInPlaceLayout = (size_t)1 << 20;

// the plan's FBP filter over `nrows` angle rows: the cached FFT form where it fits
hipError_t plan_filter(const raft_ss_plan* p, const float* in, float* out, int nrows, hipStream_t stream) {
  if (p->d_ramp_fft) return ss_launch_fft_filter(in, out, p->ramp_fft, nrows, stream);
  return ss_launch_ramp_filter(in, out, p->d_ramp, p->g.nrays, nrows, stream);
}
}  // namespace

// ------------------------------------------------------------------ device-pointer entry points

extern "C" int raft_ss_ramp_kernel(int R, double tmin, double tmax, float* host_g) {
  return raft_ss_filter_kernel(R, tmin, tmax, RAFT_SS_FILTER_COS_RAMP, 0.0, host_g);
}

extern "C" int raft_ss_filter_kernel(int R, double tmin, double tmax, int kind, double param, float* host_g) {
  if (R < 2 || !host_g) return set_error(RAFT_SS_EINVAL, "filter kernel needs R >= 2 and an output buffer");
  if (!std::isfinite(tmin) || !std::isfinite(tmax) || tmin == tmax)
    return set_error(RAFT_SS_EINVAL, "degenerate detector extent");
  if (kind != RAFT_SS_FILTER_COS_RAMP && kind != RAFT_SS_FILTER_TIKHONOV_RAMP)
    return set_error(RAFT_SS_EINVAL, "unknown filter kind %d", kind);
  ss_filter_kernel(R, tmin, tmax, kind, param, host_g);
  return 0;
}

extern "C" int raft_ss_filter_rows(const float* d_in, float* d_out, const float* d_g, int R, int nrows,
                                   void* stream) {
  if (!d_in || !d_out || !d_g || R < 2 || nrows < 0) return set_error(RAFT_SS_EINVAL, "invalid filter arguments");
  if (d_in == d_out) return set_error(RAFT_SS_EINVAL, "filter cannot run in place");
  if (nrows == 0) return 0;
  HIP_TRY(ss_launch_ramp_filter(d_in, d_out, d_g, R, nrows, as_stream(stream)));
  return 0;
}

extern "C" int raft_ss_radon_range(const raft_ss_plan* p, const float* d_image, float* d_out, int angle_major, int j0,
                                   int j1, int nslices, void* stream_) {
  int rc = check_plan(p, nslices);
  if (rc) return rc;
  if ((rc = check_range(p, j0, j1))) return rc;
  if (!d_image || !d_out) return set_error(RAFT_SS_EINVAL, "null device pointer");
  hipStream_t stream = as_stream(stream_);
  SSDevPlan dp = p->dev();
  if ((rc = ss_plan_radon_rows(p, j0, j1, nslices, stream, true, &dp))) return rc;
  const size_t layout = (size_t)nslices * p->g.n * p->g.pitch;
  Scratch ws;
  HIP_TRY(ws.alloc(2 * layout * sizeof(float2) + ss_radon_scratch(dp, j0, j1, nslices), stream));
  float2* gA = ws.as<float2>();
  float2* gB = gA + layout;
  float* slabs = reinterpret_cast<float*>(gB + layout);
  HIP_TRY(ss_launch_gather_layouts(dp, d_image, gA, gB, nslices, stream));
  SSEmRatioArgs none{nullptr, nullptr};
  HIP_TRY(ss_launch_radon(dp, gA, gB, j0, j1, nslices, angle_major ? SS_OUT_SINOT : SS_OUT_SINO, d_out, none, slabs,
                          stream));
  return 0;
}

extern "C" int raft_ss_radon(const raft_ss_plan* p, const float* d_image, float* d_sino, int nslices, void* stream) {
  if (!p) return set_error(RAFT_SS_EINVAL, "null plan");
  return raft_ss_radon_range(p, d_image, d_sino, 0, 0, p->g.nangles, nslices, stream);
}

extern "C" int raft_ss_backprojection_range(const raft_ss_plan* p, const float* d_sino, float* d_image,
                                            int angle_major, int j0, int j1, int nslices, void* stream_) {
  int rc = check_plan(p, nslices);
  if (rc) return rc;
  if ((rc = check_uniform(p, "backprojection"))) return rc;
  if ((rc = check_range(p, j0, j1))) return rc;
  if (!d_image || !d_sino) return set_error(RAFT_SS_EINVAL, "null device pointer");
  hipStream_t stream = as_stream(stream_);
  const SSDevPlan dp = p->dev();
  const float* sT = d_sino;
  const size_t sino_floats = angle_major ? 0 : (size_t)nslices * p->g.nrays * p->g.nangles;
  Scratch ws;
  HIP_TRY(ws.alloc(sino_floats * sizeof(float) + ss_bp_scratch(dp, j0, j1, nslices), stream));
  if (!angle_major) {
    HIP_TRY(ss_launch_sino_to_t(d_sino, ws.as<float>(), p->g.nrays, p->g.nangles, nslices, stream));
    sT = ws.as<float>();
  }
  SSEmUpdateArgs none{nullptr, nullptr};
  HIP_TRY(ss_launch_backprojection(dp, sT, j0, j1, nslices, SS_BP_PLAIN, d_image, none, ws.as<float>() + sino_floats,
                                   stream));
  return 0;
}

extern "C" int raft_ss_backprojection(const raft_ss_plan* p, const float* d_sino, float* d_image, int nslices,
                                      void* stream) {
  if (!p) return set_error(RAFT_SS_EINVAL, "null plan");
  return raft_ss_backprojection_range(p, d_sino, d_image, 0, 0, p->g.nangles, nslices, stream);
}

extern "C" int raft_ss_radon_transpose_range(const raft_ss_plan* p, const float* d_sino, float* d_image,
                                             int angle_major, int j0, int j1, int nslices, void* stream_) {
  int rc = check_plan(p, nslices);
  if (rc) return rc;
  if ((rc = check_uniform(p, "radon_transpose"))) return rc;
  if ((rc = check_range(p, j0, j1))) return rc;
  if (!d_image || !d_sino) return set_error(RAFT_SS_EINVAL, "null device pointer");
  hipStream_t stream = as_stream(stream_);
  const SSDevPlan dp = p->dev();
  const float* sT = d_sino;
  const size_t sino_floats = angle_major ? 0 : (size_t)nslices * p->g.nrays * p->g.nangles;
  const size_t scratch = ss_radon_transpose_scratch(dp, nslices);
  Scratch ws;
  HIP_TRY(ws.alloc(sino_floats * sizeof(float) + scratch, stream));
  if (!angle_major) {
    HIP_TRY(ss_launch_sino_to_t(d_sino, ws.as<float>(), p->g.nrays, p->g.nangles, nslices, stream));
    sT = ws.as<float>();
  }
  HIP_TRY(ss_launch_radon_transpose(dp, sT, j0, j1, nslices, d_image, scratch ? ws.as<float>() + sino_floats : nullptr,
                                    stream));
  return 0;
}

extern "C" int raft_ss_radon_transpose(const raft_ss_plan* p, const float* d_sino, float* d_image, int nslices,
                                       void* stream) {
  if (!p) return set_error(RAFT_SS_EINVAL, "null plan");
  return raft_ss_radon_transpose_range(p, d_sino, d_image, 0, 0, p->g.nangles, nslices, stream);
}

extern "C" int raft_ss_fbp_filter(raft_ss_plan* p, const float* d_sino, float* d_out, int angle_major, int nslices,
                                  void* stream_) {
  int rc = check_plan(p, nslices);
  if (rc) return rc;
  if ((rc = check_uniform(p, "fbp_filter"))) return rc;
  if (!d_sino || !d_out) return set_error(RAFT_SS_EINVAL, "null device pointer");
  hipStream_t stream = as_stream(stream_);
  if ((rc = ensure_ramp(p, stream))) return rc;
  const int R = p->g.nrays, A = p->g.nangles;
  const size_t total = (size_t)nslices * R * A;
  if (angle_major) {
    HIP_TRY(plan_filter(p, d_sino, d_out, nslices * A, stream));
    return 0;
  }
  if (p->d_ramp_fft && total <= kFilterInPlaceLayout) {
    HIP_TRY(ss_launch_fft_filter(d_sino, d_out, p->ramp_fft, nslices * A, stream, A, A));
    return 0;
  }
  Scratch ws;
  HIP_TRY(ws.alloc(2 * total * sizeof(float), stream));
  float* t0 = ws.as<float>();
  float* t1 = t0 + total;
  HIP_TRY(ss_launch_sino_to_t(d_sino, t0, R, A, nslices, stream));
  HIP_TRY(plan_filter(p, t0, t1, nslices * A, stream));
  HIP_TRY(ss_launch_t_to_sino(t1, d_out, R, A, nslices, stream));
  return 0;
}

extern "C" int raft_ss_fbp(raft_ss_plan* p, const float* d_sino, float* d_image, int nslices, void* stream_) {
  int rc = check_plan(p, nslices);
  if (rc) return rc;
  if ((rc = check_uniform(p, "fbp"))) return rc;
  if (!d_sino || !d_image) return set_error(RAFT_SS_EINVAL, "null device pointer");
  hipStream_t stream = as_stream(stream_);
  if ((rc = ensure_ramp(p, stream))) return rc;
  const int R = p->g.nrays, A = p->g.nangles;
  const size_t total = (size_t)nslices * R * A;
  const SSDevPlan dp = p->dev();
  Scratch ws;
  HIP_TRY(ws.alloc(2 * total * sizeof(float) + ss_bp_scratch(dp, 0, A, nslices), stream));
  float* t0 = ws.as<float>();
  float* t1 = t0 + total;
  if (p->d_ramp_fft && total <= kFilterInPlaceLayout) {
    HIP_TRY(ss_launch_fft_filter(d_sino, t1, p->ramp_fft, nslices * A, stream, A, 0));
  } else {
    HIP_TRY(ss_launch_sino_to_t(d_sino, t0, R, A, nslices, stream));
    HIP_TRY(plan_filter(p, t0, t1, nslices * A, stream));
  }
  SSEmUpdateArgs none{nullptr, nullptr};
  HIP_TRY(ss_launch_backprojection(dp, t1, 0, A, nslices, SS_BP_PLAIN, d_image, none, t1 + total, stream));
  return 0;
}

// Launch geometry of an OSEM run: each subset's Radon with the plan's slab choice for its
// angle count, the LL partial offsets of the subsets, the subsets' backprojection variant
// (tuned for their angle count: launches of few angles favour smaller tiles than a
// full-range backprojection) and the workspace it needs.  Timing choices are made here,
// so a later call of ss_em_launch (possibly under stream capture) only looks them up.
int ss_em_prepare(const raft_ss_plan* p, int nsubsets, int nslices, int niter, hipStream_t stream, SSEmLaunch* L) {
  const int R = p->g.nrays, A = p->g.nangles, n = p->g.n;
  int rc = 0;
  L->S = nsubsets;
  L->dp = p->dev();
  L->poff.assign(nsubsets + 1, 0);
  L->dps.assign(nsubsets, L->dp);
  size_t slab_bytes = 0;  // the largest subset's slab partial sums
  for (int b = 0; b < nsubsets; ++b) {
    const int j0 = (int)(((long long)b * A) / nsubsets), j1 = (int)(((long long)(b + 1) * A) / nsubsets);
    if ((rc = ss_plan_radon_rows(p, j0, j1, nslices, stream, true, &L->dps[b]))) return rc;
    L->poff[b + 1] = L->poff[b] + ss_radon_partials(L->dps[b], j0, j1, nslices);
    slab_bytes = std::max(slab_bytes, ss_radon_scratch(L->dps[b], j0, j1, nslices));
  }
  int bv = p->bp_variant;
  if ((rc = ss_plan_range_variant(const_cast<raft_ss_plan*>(p), 0, (int)(A / nsubsets), stream, &bv))) return rc;
  L->dpb = p->dev(bv);
  size_t bp_bytes = 0;  // the largest subset's backprojection partial images
  for (int b = 0; b < nsubsets; ++b) {
    const int j0 = (int)(((long long)b * A) / nsubsets), j1 = (int)(((long long)(b + 1) * A) / nsubsets);
    bp_bytes = std::max(bp_bytes, ss_bp_scratch(L->dpb, j0, j1, nslices));
  }
  // workspace: two gather layouts, the measured and ratio sinograms (detector-major),
  // the S weight images, the LL partials, the LL trace, slab sums, BP partial images
  const size_t sino_sz = (size_t)nslices * R * A, img_sz = (size_t)nslices * n * n;
  const size_t layout = (size_t)nslices * n * p->g.pitch;
  size_t off = sizeof(float2) * 2 * layout;
  L->meas_off = off;
  off += sizeof(float) * sino_sz;
  L->rho_off = off;
  off += sizeof(float) * sino_sz;
  L->weights_off = off;
  off += sizeof(float) * (size_t)nsubsets * img_sz;
  off = (off + 7) & ~(size_t)7;
  L->partials_off = off;
  off += sizeof(double) * (size_t)L->poff[nsubsets];
  L->ll_off = off;
  off += sizeof(double) * (size_t)std::max(niter, 1);
  L->slabs_off = off;
  off += slab_bytes;
  L->bp_off = off;
  L->bytes = off + bp_bytes;
  return 0;
}

// The OSEM kernels of a prepared run on workspace `ws` (L.bytes): W_b = BP_b(1), then
// per iteration and subset the gather, the Radon with the ratio and LL partials in its
// epilogue and the backprojection with the update in its epilogue; the LL of iteration
// it goes to d_ll[it] (device).  Kernel launches only: no allocation, copy or wait.
int ss_em_launch(const raft_ss_plan* p, const SSEmLaunch& L, const float* d_sino, float* d_x, int niter,
                 int nslices, void* ws, double* d_ll, hipStream_t stream) {
  const int R = p->g.nrays, A = p->g.nangles, n = p->g.n, S = L.S;
  const size_t sino_sz = (size_t)nslices * R * A, img_sz = (size_t)nslices * n * n;
  const size_t layout = (size_t)nslices * n * p->g.pitch;
  const int npartials = L.poff[S];
  char* base = static_cast<char*>(ws);
  float2* gA = reinterpret_cast<float2*>(base);
  float2* gB = gA + layout;
  float* measT = reinterpret_cast<float*>(base + L.meas_off);
  float* rhoT = reinterpret_cast<float*>(base + L.rho_off);
  float* weights = reinterpret_cast<float*>(base + L.weights_off);
  double* partials = reinterpret_cast<double*>(base + L.partials_off);
  float* slabs = reinterpret_cast<float*>(base + L.slabs_off);
  float* bp_part = reinterpret_cast<float*>(base + L.bp_off);
  if (!d_ll) d_ll = reinterpret_cast<double*>(base + L.ll_off);

  HIP_TRY(ss_launch_sino_to_t(d_sino, measT, R, A, nslices, stream));
  // W_b = BP_b(1) with the same backprojection kernel
  HIP_TRY(ss_launch_fill(rhoT, 1.0f, sino_sz, stream));
  SSEmUpdateArgs none{nullptr, nullptr};
  for (int b = 0; b < S; ++b) {
    const int j0 = (int)(((long long)b * A) / S), j1 = (int)(((long long)(b + 1) * A) / S);
    HIP_TRY(ss_launch_backprojection(L.dpb, rhoT, j0, j1, nslices, SS_BP_PLAIN, weights + (size_t)b * img_sz, none,
                                     bp_part, stream));
  }
  for (int it = 0; it < niter; ++it) {
    for (int b = 0; b < S; ++b) {
      const int j0 = (int)(((long long)b * A) / S), j1 = (int)(((long long)(b + 1) * A) / S);
      // only the layouts the subset's angles read (its Radon reads no other)
      int need = 0;
      for (int j = j0; j < j1; ++j) need |= p->h_angles[j].branch_a ? SS_GATHER_A : SS_GATHER_B;
      HIP_TRY(ss_launch_gather_layouts(L.dp, d_x, gA, gB, nslices, stream, need));
      SSEmRatioArgs ea{measT, partials + L.poff[b]};
      HIP_TRY(ss_launch_radon(L.dps[b], gA, gB, j0, j1, nslices, SS_OUT_EM_RATIO, rhoT, ea, slabs, stream));
      SSEmUpdateArgs ua{d_x, weights + (size_t)b * img_sz};
      HIP_TRY(ss_launch_backprojection(L.dpb, rhoT, j0, j1, nslices, SS_BP_EM_UPDATE, nullptr, ua, bp_part, stream));
    }
    HIP_TRY(ss_launch_sum_partials(partials, npartials, d_ll + it, stream));
  }
  return 0;
}

// Ordered-subsets EM.  d_x holds the starting image on entry (the reference starts
// from ones, raft/raft_em_gpu.cu:41-46) and the result on exit.  Subset b covers
// angle columns [b*A/S, (b+1)*A/S); S = 1 is the plain EM of raft_em_gpu.
// ll_host (optional, niter doubles) receives the per-iteration log-likelihood.
extern "C" int raft_ss_em(raft_ss_plan* p, const float* d_sino, float* d_x, int niter, int nsubsets, int nslices,
                          double* ll_host, void* stream_) {
  int rc = check_plan(p, nslices);
  if (rc) return rc;
  if ((rc = check_uniform(p, "em"))) return rc;
  if (!d_sino || !d_x) return set_error(RAFT_SS_EINVAL, "null device pointer");
  if (niter < 0) return set_error(RAFT_SS_EINVAL, "niter must be >= 0");
  const int A = p->g.nangles;
  if (nsubsets < 1 || nsubsets > A) return set_error(RAFT_SS_EINVAL, "nsubsets must be in [1, %d]", A);
  hipStream_t stream = as_stream(stream_);
  if (ll_host) {  // the host copy needs a wait, which a stream capture cannot hold
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    HIP_TRY(hipStreamIsCapturing(stream, &cs));
    if (cs != hipStreamCaptureStatusNone)
      return set_error(RAFT_SS_EINVAL, "em: ll_host must be NULL while the stream is being captured");
  }
  SSEmLaunch L;
  if ((rc = ss_em_prepare(p, nsubsets, nslices, niter, stream, &L))) return rc;
  Scratch ws;
  HIP_TRY(ws.alloc(L.bytes, stream));
  if ((rc = ss_em_launch(p, L, d_sino, d_x, niter, nslices, ws.p, nullptr, stream))) return rc;
  if (ll_host && niter > 0) {
    HIP_TRY(hipMemcpyAsync(ll_host, static_cast<char*>(ws.p) + L.ll_off, sizeof(double) * niter,
                           hipMemcpyDeviceToHost, stream));
    HIP_TRY(hipStreamSynchronize(stream));
  }
  return 0;
}

// Transmission EM (raft/raft_transmission_em.cu:16-66): W = BP(raw), then per iteration
// x <- x (BP(flat exp(-R x)) / W)^2.
extern "C" int raft_ss_tr_em(raft_ss_plan* p, const float* d_raw, const float* d_flat, float* d_x, int niter,
                             int nslices, void* stream_) {
  int rc = check_plan(p, nslices);
  if (rc) return rc;
  if ((rc = check_uniform(p, "tr_em"))) return rc;
  if (!d_raw || !d_flat || !d_x) return set_error(RAFT_SS_EINVAL, "null device pointer");
  if (niter < 0) return set_error(RAFT_SS_EINVAL, "niter must be >= 0");
  const int R = p->g.nrays, A = p->g.nangles, n = p->g.n;
  hipStream_t stream = as_stream(stream_);
  SSDevPlan dp = p->dev();
  if ((rc = ss_plan_radon_rows(p, 0, A, nslices, stream, true, &dp))) return rc;
  const size_t sino_sz = (size_t)nslices * R * A;
  const size_t img_sz = (size_t)nslices * n * n;
  const size_t layout = (size_t)nslices * n * p->g.pitch;
  Scratch ws;
  const size_t slab_bytes = ss_radon_scratch(dp, 0, A, nslices);
  HIP_TRY(ws.alloc(sizeof(float2) * 2 * layout + sizeof(float) * (3 * sino_sz + 2 * img_sz) + slab_bytes +
                       ss_bp_scratch(dp, 0, A, nslices),
                   stream));
  float2* gA = ws.as<float2>();
  float2* gB = gA + layout;
  float* rawT = reinterpret_cast<float*>(gB + layout);
  float* flatT = rawT + sino_sz;
  float* pT = flatT + sino_sz;
  float* w = pT + sino_sz;
  float* b = w + img_sz;
  float* slabs = b + img_sz;
  float* bp_part = slabs + slab_bytes / sizeof(float);
  SSEmUpdateArgs none{nullptr, nullptr};
  SSEmRatioArgs no_ratio{nullptr, nullptr};
  HIP_TRY(ss_launch_sino_to_t(d_raw, rawT, R, A, nslices, stream));
  HIP_TRY(ss_launch_sino_to_t(d_flat, flatT, R, A, nslices, stream));
  HIP_TRY(ss_launch_backprojection(dp, rawT, 0, A, nslices, SS_BP_PLAIN, w, none, bp_part, stream));
  for (int it = 0; it < niter; ++it) {
    HIP_TRY(ss_launch_gather_layouts(dp, d_x, gA, gB, nslices, stream));
    HIP_TRY(ss_launch_radon(dp, gA, gB, 0, A, nslices, SS_OUT_SINOT, pT, no_ratio, slabs, stream));
    HIP_TRY(ss_launch_tr_expected(pT, flatT, sino_sz, stream));
    HIP_TRY(ss_launch_backprojection(dp, pT, 0, A, nslices, SS_BP_PLAIN, b, none, bp_part, stream));
    HIP_TRY(ss_launch_tr_update(d_x, b, w, img_sz, stream));
  }
  return 0;
}

// ------------------------------------------------------------------ phantoms

static int upload_ellipses(const raft_matrix& desc, std::vector<SSEllipse>& ell) {
  if (desc.columns < 6) return set_error(RAFT_SS_EINVAL, "phantom description needs 6 columns (got %d)", desc.columns);
  ell.resize(desc.lines);
  for (int e = 0; e < desc.lines; ++e) {
    SSEllipse E;
    E.val = RAFT_MATRIX_AT(desc, e, 0);
    E.a = RAFT_MATRIX_AT(desc, e, 1);
    E.b = RAFT_MATRIX_AT(desc, e, 2);
    E.x0 = RAFT_MATRIX_AT(desc, e, 3);
    E.y0 = RAFT_MATRIX_AT(desc, e, 4);
    const double rot = RAFT_MATRIX_AT(desc, e, 5);
    E.c = cos(rot);
    E.s = sin(rot);
    ell[e] = E;
  }
  return 0;
}

// Rasterise an ellipse description into a device FP32 image (row 0 = tl_y, the
// rasteriser's own orientation, raft/raft_image_phantom_fromdesc.c:18-41).
extern "C" int raft_ss_phantom(const double* desc_rows, int nell, int lines, int cols, const double* corners,
                               float* d_out, void* stream_) {
  if (!desc_rows || !corners || !d_out || nell < 0 || lines < 2 || cols < 2)
    return set_error(RAFT_SS_EINVAL, "invalid phantom arguments");
  hipStream_t stream = as_stream(stream_);
  raft_matrix desc{const_cast<double*>(desc_rows), nell, 1, 6, 6};
  std::vector<SSEllipse> ell;
  int rc = upload_ellipses(desc, ell);
  if (rc) return rc;
  DevBuf d_ell;
  HIP_TRY(d_ell.alloc(sizeof(SSEllipse) * std::max(1, nell), stream));
  if (nell) HIP_TRY(hipMemcpyAsync(d_ell.p, ell.data(), sizeof(SSEllipse) * nell, hipMemcpyHostToDevice, stream));
  const double hs = (corners[2] - corners[0]) / (cols - 1), vs = (corners[3] - corners[1]) / (lines - 1);
  HIP_TRY(ss_launch_phantom(d_ell.as<SSEllipse>(), nell, lines, cols, corners[0], corners[1], hs, vs, nullptr, d_out,
                            stream));
  HIP_TRY(hipStreamSynchronize(stream));
  return 0;
}

// ------------------------------------------------------------------ host <-> raft_image staging

namespace {

// plan cache for the host-pointer entry points
struct PlanKey {
  int device, n, R, A;
  double c[8];
  bool operator==(const PlanKey& o) const {
    return device == o.device && n == o.n && R == o.R && A == o.A && memcmp(c, o.c, sizeof(c)) == 0;
  }
};
std::mutex g_cache_mutex;
std::list<std::pair<PlanKey, std::shared_ptr<raft_ss_plan>>> g_cache;

int cached_plan(int n, int R, int A, const double* ic, const double* sc, std::shared_ptr<raft_ss_plan>* out) {
  PlanKey key;
  memset(&key, 0, sizeof(key));
  HIP_TRY(hipGetDevice(&key.device));
  key.n = n;
  key.R = R;
  key.A = A;
  memcpy(key.c, ic, 4 * sizeof(double));
  memcpy(key.c + 4, sc, 4 * sizeof(double));
  {
    std::lock_guard<std::mutex> lock(g_cache_mutex);
    for (auto it = g_cache.begin(); it != g_cache.end(); ++it)
      if (it->first == key) {
        g_cache.splice(g_cache.begin(), g_cache, it);
        *out = g_cache.front().second;
        return 0;
      }
  }
  raft_ss_plan* raw = nullptr;
  int rc = raft_ss_plan_create(&raw, n, R, A, ic, sc, nullptr);
  if (rc) return rc;
  std::shared_ptr<raft_ss_plan> sp(raw);
  std::lock_guard<std::mutex> lock(g_cache_mutex);
  g_cache.emplace_front(key, sp);
  while (g_cache.size() > 8) g_cache.pop_back();
  *out = sp;
  return 0;
}


int host_radon(const raft_image& phantom, const raft_image& radon) {
  if (!phantom.data.p_data || !radon.data.p_data) return set_error(RAFT_SS_EINVAL, "empty image");
  const int n = phantom.data.lines;
  if (phantom.data.columns != n)
    return set_error(RAFT_SS_EINVAL, "slant-stack needs a square image (got %dx%d)", n, phantom.data.columns);
  const int R = radon.data.lines, A = radon.data.columns;
  double ic[4], sc[4];
  corners_of(phantom, ic);
  corners_of(radon, sc);
  std::shared_ptr<raft_ss_plan> plan;
  int rc = cached_plan(n, R, A, ic, sc, &plan);
  if (rc) return rc;
  PinnedF32 hi, ho;
  HIP_TRY(hi.alloc((size_t)n * n));
  HIP_TRY(ho.alloc((size_t)R * A));
  gather_f64(phantom.data, hi.p);
  DevBuf di, dso;
  HIP_TRY(di.alloc(sizeof(float) * n * n, nullptr));
  HIP_TRY(dso.alloc(sizeof(float) * (size_t)R * A, nullptr));
  HIP_TRY(hipMemcpyAsync(di.p, hi.p, sizeof(float) * n * n, hipMemcpyHostToDevice, nullptr));
  rc = raft_ss_radon(plan.get(), di.as<float>(), dso.as<float>(), 1, nullptr);
  if (rc) return rc;
  HIP_TRY(hipMemcpyAsync(ho.p, dso.p, sizeof(float) * (size_t)R * A, hipMemcpyDeviceToHost, nullptr));
  HIP_TRY(hipStreamSynchronize(nullptr));
  scatter_f64(ho.p, radon.data);
  return 0;
}

// Radon on arbitrary (t, theta) tables into an nrays x nangles strided matrix, element
// (k, j) = ray k, angle j (raft/raft_radon_slantstack_withmesh.c:118-121)
int host_radon_mesh(const raft_image& phantom, const raft_matrix& out, const raft_vector& t,
                    const raft_vector& theta) {
  if (!phantom.data.p_data || !out.p_data || !t.p_data || !theta.p_data)
    return set_error(RAFT_SS_EINVAL, "empty image or vector");
  const int n = phantom.data.lines;
  if (phantom.data.columns != n)
    return set_error(RAFT_SS_EINVAL, "slant-stack needs a square image (got %dx%d)", n, phantom.data.columns);
  const int R = out.lines, A = out.columns;
  if (t.size < R || theta.size < A)
    return set_error(RAFT_SS_EINVAL, "mesh tables too short (t: %d < %d rays or theta: %d < %d angles)", t.size, R,
                     theta.size, A);
  if (R < 1 || A < 1) return 0;
  // a one-ray table is run as two rays at the same t
  const int Rp = std::max(R, 2);
  std::vector<double> ht(Rp), hth(A);
  for (int k = 0; k < Rp; ++k) ht[k] = t.p_data[(long)std::min(k, R - 1) * t.stride];
  for (int j = 0; j < A; ++j) hth[j] = theta.p_data[(long)j * theta.stride];
  double ic[4];
  corners_of(phantom, ic);
  raft_ss_plan* raw = nullptr;
  int rc = raft_ss_plan_create_mesh(&raw, n, ic, Rp, ht.data(), A, hth.data(), nullptr);
  if (rc) return rc;
  std::unique_ptr<raft_ss_plan, int (*)(raft_ss_plan*)> plan(raw, raft_ss_plan_destroy);
  PinnedF32 hi, ho;
  HIP_TRY(hi.alloc((size_t)n * n));
  HIP_TRY(ho.alloc((size_t)Rp * A));
  gather_f64(phantom.data, hi.p);
  DevBuf di, dso;
  HIP_TRY(di.alloc(sizeof(float) * n * n, nullptr));
  HIP_TRY(dso.alloc(sizeof(float) * (size_t)Rp * A, nullptr));
  HIP_TRY(hipMemcpyAsync(di.p, hi.p, sizeof(float) * n * n, hipMemcpyHostToDevice, nullptr));
  rc = raft_ss_radon(plan.get(), di.as<float>(), dso.as<float>(), 1, nullptr);
  if (rc) return rc;
  HIP_TRY(hipMemcpyAsync(ho.p, dso.p, sizeof(float) * (size_t)Rp * A, hipMemcpyDeviceToHost, nullptr));
  HIP_TRY(hipStreamSynchronize(nullptr));
  for (int k = 0; k < R; ++k)
    for (int j = 0; j < A; ++j) RAFT_MATRIX_AT(out, k, j) = (double)ho.p[(size_t)k * A + j];
  return 0;
}

int host_backprojection(const raft_image& sino, const raft_image& img) {
  if (!sino.data.p_data || !img.data.p_data) return set_error(RAFT_SS_EINVAL, "empty image");
  const int n = img.data.lines;
  if (img.data.columns != n)
    return set_error(RAFT_SS_EINVAL, "slant-stack needs a square image (got %dx%d)", n, img.data.columns);
  const int R = sino.data.lines, A = sino.data.columns;
  double ic[4], sc[4];
  corners_of(img, ic);
  corners_of(sino, sc);
  std::shared_ptr<raft_ss_plan> plan;
  int rc = cached_plan(n, R, A, ic, sc, &plan);
  if (rc) return rc;
  PinnedF32 hs, ho;
  HIP_TRY(hs.alloc((size_t)R * A));
  HIP_TRY(ho.alloc((size_t)n * n));
  gather_f64(sino.data, hs.p);
  DevBuf ds, di;
  HIP_TRY(ds.alloc(sizeof(float) * (size_t)R * A, nullptr));
  HIP_TRY(di.alloc(sizeof(float) * n * n, nullptr));
  HIP_TRY(hipMemcpyAsync(ds.p, hs.p, sizeof(float) * (size_t)R * A, hipMemcpyHostToDevice, nullptr));
  rc = raft_ss_backprojection(plan.get(), ds.as<float>(), di.as<float>(), 1, nullptr);
  if (rc) return rc;
  HIP_TRY(hipMemcpyAsync(ho.p, di.p, sizeof(float) * n * n, hipMemcpyDeviceToHost, nullptr));
  HIP_TRY(hipStreamSynchronize(nullptr));
  scatter_f64(ho.p, img.data);
  return 0;
}

int host_radon_transpose(const raft_image& sino, const raft_image& img) {
  if (!sino.data.p_data || !img.data.p_data) return set_error(RAFT_SS_EINVAL, "empty image");
  const int n = img.data.lines;
  if (img.data.columns != n)
    return set_error(RAFT_SS_EINVAL, "slant-stack needs a square image (got %dx%d)", n, img.data.columns);
  const int R = sino.data.lines, A = sino.data.columns;
  double ic[4], sc[4];
  corners_of(img, ic);
  corners_of(sino, sc);
  std::shared_ptr<raft_ss_plan> plan;
  int rc = cached_plan(n, R, A, ic, sc, &plan);
  if (rc) return rc;
  PinnedF32 hs, ho;
  HIP_TRY(hs.alloc((size_t)R * A));
  HIP_TRY(ho.alloc((size_t)n * n));
  gather_f64(sino.data, hs.p);
  DevBuf ds, di;
  HIP_TRY(ds.alloc(sizeof(float) * (size_t)R * A, nullptr));
  HIP_TRY(di.alloc(sizeof(float) * n * n, nullptr));
  HIP_TRY(hipMemcpyAsync(ds.p, hs.p, sizeof(float) * (size_t)R * A, hipMemcpyHostToDevice, nullptr));
  rc = raft_ss_radon_transpose(plan.get(), ds.as<float>(), di.as<float>(), 1, nullptr);
  if (rc) return rc;
  HIP_TRY(hipMemcpyAsync(ho.p, di.p, sizeof(float) * n * n, hipMemcpyDeviceToHost, nullptr));
  HIP_TRY(hipStreamSynchronize(nullptr));
  scatter_f64(ho.p, img.data);
  return 0;
}

const double kDefaultImg[4] = {-1.0, 1.0, 1.0, -1.0};
const double kDefaultSino[4] = {0.0, 1.0, 3.14159265358979323846, -1.0};

int host_f32_op(int which, float* h_out, const float* h_in, int n, int R, int A, int niter,
                const float* h_in2 = nullptr) {
  if (!h_out || !h_in) return set_error(RAFT_SS_EINVAL, "null host pointer");
  std::shared_ptr<raft_ss_plan> plan;
  int rc = cached_plan(n, R, A, kDefaultImg, kDefaultSino, &plan);
  if (rc) return rc;
  const size_t nimg = (size_t)n * n, nsino = (size_t)R * A;
  DevBuf da, db;
  HIP_TRY(da.alloc(sizeof(float) * nimg, nullptr));
  HIP_TRY(db.alloc(sizeof(float) * nsino, nullptr));
  if (which == 0) {  // radon: in = image, out = sino
    HIP_TRY(hipMemcpyAsync(da.p, h_in, sizeof(float) * nimg, hipMemcpyHostToDevice, nullptr));
    if ((rc = raft_ss_radon(plan.get(), da.as<float>(), db.as<float>(), 1, nullptr))) return rc;
    HIP_TRY(hipMemcpyAsync(h_out, db.p, sizeof(float) * nsino, hipMemcpyDeviceToHost, nullptr));
  } else if (which == 1) {  // backprojection: in = sino, out = image
    HIP_TRY(hipMemcpyAsync(db.p, h_in, sizeof(float) * nsino, hipMemcpyHostToDevice, nullptr));
    if ((rc = raft_ss_backprojection(plan.get(), db.as<float>(), da.as<float>(), 1, nullptr))) return rc;
    HIP_TRY(hipMemcpyAsync(h_out, da.p, sizeof(float) * nimg, hipMemcpyDeviceToHost, nullptr));
  } else if (which == 2) {  // em from ones
    HIP_TRY(hipMemcpyAsync(db.p, h_in, sizeof(float) * nsino, hipMemcpyHostToDevice, nullptr));
    HIP_TRY(ss_launch_fill(da.as<float>(), 1.0f, nimg, nullptr));
    if ((rc = raft_ss_em(plan.get(), db.as<float>(), da.as<float>(), niter, 1, 1, nullptr, nullptr))) return rc;
    HIP_TRY(hipMemcpyAsync(h_out, da.p, sizeof(float) * nimg, hipMemcpyDeviceToHost, nullptr));
  } else {  // transmission em from ones: h_in = raw sinogram, h_in2 = flat field
    if (!h_in2) return set_error(RAFT_SS_EINVAL, "null host pointer");
    DevBuf dflat;
    HIP_TRY(dflat.alloc(sizeof(float) * nsino, nullptr));
    HIP_TRY(hipMemcpyAsync(db.p, h_in, sizeof(float) * nsino, hipMemcpyHostToDevice, nullptr));
    HIP_TRY(hipMemcpyAsync(dflat.p, h_in2, sizeof(float) * nsino, hipMemcpyHostToDevice, nullptr));
    HIP_TRY(ss_launch_fill(da.as<float>(), 1.0f, nimg, nullptr));
    if ((rc = raft_ss_tr_em(plan.get(), db.as<float>(), dflat.as<float>(), da.as<float>(), niter, 1, nullptr)))
      return rc;
    HIP_TRY(hipMemcpyAsync(h_out, da.p, sizeof(float) * nimg, hipMemcpyDeviceToHost, nullptr));
  }
  HIP_TRY(hipStreamSynchronize(nullptr));
  return 0;
}

thread_local int g_legacy_status = 0;

void report(const char* fn, int rc) { ss_report_legacy(fn, rc); }
}  // namespace

void ss_report_legacy(const char* fn, int rc) {
  g_legacy_status = rc;
  if (rc) fprintf(stderr, "libraft %s: %s\n", fn, g_last_error.c_str());
}

extern "C" int raft_ss_legacy_status(void) { return g_legacy_status; }

// ------------------------------------------------------------------ legacy symbols (reference ABI)

extern "C" void raft_radon_slantstack(raft_image phantom, raft_image radon, int nthreads) {
  (void)nthreads;
  report("raft_radon_slantstack", host_radon(phantom, radon));
}

extern "C" void raft_backprojection_slantstack(raft_image sinogram, raft_image backprojection, int nthreads) {
  (void)nthreads;
  report("raft_backprojection_slantstack", host_backprojection(sinogram, backprojection));
}

extern "C" void raft_radon_transpose_slantstack(raft_image sinogram, raft_image image, int nthreads) {
  (void)nthreads;
  report("raft_radon_transpose_slantstack", host_radon_transpose(sinogram, image));
}

extern "C" void raft_radon_slantstack_withmesh(raft_image phantom, raft_image radon, raft_vector t, raft_vector theta,
                                               int nthreads) {
  (void)nthreads;
  report("raft_radon_slantstack_withmesh", host_radon_mesh(phantom, radon.data, t, theta));
}

extern "C" void raft_radon_slantstack_view(raft_vector view, raft_image phantom, raft_vector t, double theta) {
  // one angle: the view is a t.size x 1 matrix over the view vector
  raft_matrix out;
  out.p_data = view.p_data;
  out.lines = t.size;
  out.columns = 1;
  out.column_stride = view.stride;
  out.line_stride = 1;
  raft_vector th;
  th.p_data = &theta;
  th.size = 1;
  th.stride = 1;
  int rc = 0;
  if (view.size < t.size) rc = set_error(RAFT_SS_EINVAL, "view vector shorter than t (%d < %d)", view.size, t.size);
  if (!rc) rc = host_radon_mesh(phantom, out, t, th);
  report("raft_radon_slantstack_view", rc);
}

extern "C" void raft_tr_em_gpu(float* output, float* rawsino, float* flat, int sizeImage, int nrays, int nangles,
                               int niter) {
  report("raft_tr_em_gpu", host_f32_op(3, output, rawsino, sizeImage, nrays, nangles, niter, flat));
}

extern "C" void raft_radon_slantstack_gpu(float* h_output, float* h_input, int sizeImage, int nrays, int nangles) {
  report("raft_radon_slantstack_gpu", host_f32_op(0, h_output, h_input, sizeImage, nrays, nangles, 0));
}

extern "C" void raft_backprojection_slantstack_gpu(float* image, float* sino, int sizeImage, int nrays,
                                                   int nangles) {
  report("raft_backprojection_slantstack_gpu", host_f32_op(1, image, sino, sizeImage, nrays, nangles, 0));
}

extern "C" void raft_em_gpu(float* output, float* sino, int sizeImage, int nrays, int nangles, int niter) {
  report("raft_em_gpu", host_f32_op(2, output, sino, sizeImage, nrays, nangles, niter));
}

static int host_phantom(const raft_image& image, const raft_matrix& desc) {
  if (!image.data.p_data) return 0;  // nothing to do (raft/raft_image_phantom_fromdesc.c:7-8)
  std::vector<SSEllipse> ell;
  int rc = upload_ellipses(desc, ell);
  if (rc) return rc;
  const int lines = image.data.lines, cols = image.data.columns;
  if (lines < 2 || cols < 2) return set_error(RAFT_SS_EINVAL, "phantom image must be at least 2x2");
  const size_t total = (size_t)lines * cols;
  std::vector<double> h(total);
  for (int i = 0; i < lines; ++i)
    for (int j = 0; j < cols; ++j) h[(size_t)i * cols + j] = RAFT_MATRIX_AT(image.data, i, j);
  DevBuf d_ell, d_img;
  HIP_TRY(d_ell.alloc(sizeof(SSEllipse) * std::max<size_t>(1, ell.size()), nullptr));
  HIP_TRY(d_img.alloc(sizeof(double) * total, nullptr));
  if (!ell.empty())
    HIP_TRY(hipMemcpyAsync(d_ell.p, ell.data(), sizeof(SSEllipse) * ell.size(), hipMemcpyHostToDevice, nullptr));
  HIP_TRY(hipMemcpyAsync(d_img.p, h.data(), sizeof(double) * total, hipMemcpyHostToDevice, nullptr));
  const double hs = (image.br_x - image.tl_x) / (cols - 1), vs = (image.br_y - image.tl_y) / (lines - 1);
  HIP_TRY(ss_launch_phantom(d_ell.as<SSEllipse>(), (int)ell.size(), lines, cols, image.tl_x, image.tl_y, hs, vs,
                            d_img.as<double>(), nullptr, nullptr));
  HIP_TRY(hipMemcpyAsync(h.data(), d_img.p, sizeof(double) * total, hipMemcpyDeviceToHost, nullptr));
  HIP_TRY(hipStreamSynchronize(nullptr));
  for (int i = 0; i < lines; ++i)
    for (int j = 0; j < cols; ++j) RAFT_MATRIX_AT(image.data, i, j) = h[(size_t)i * cols + j];
  return 0;
}

extern "C" void raft_image_phantom_fromdesc(raft_image image, raft_matrix desc) {
  report("raft_image_phantom_fromdesc", host_phantom(image, desc));
}

static int host_radon_fromdesc(const raft_image& sino, const raft_matrix& desc) {
  if (!sino.data.p_data) return 0;
  if (desc.columns < 6) return set_error(RAFT_SS_EINVAL, "phantom description needs 6 columns");
  const int lines = sino.data.lines, cols = sino.data.columns, nell = desc.lines;
  if (lines < 2 || cols < 2) return set_error(RAFT_SS_EINVAL, "sinogram must be at least 2x2");
  // per (ellipse, column) constants, FP64 with glibc trig (raft/raft_radon_fromdesc.c:9-50)
  const double ths = (sino.br_x - sino.tl_x) / (cols - 1);
  const double tss = (sino.br_y - sino.tl_y) / (lines - 1);
  std::vector<SSEllipseView> views((size_t)std::max(1, nell) * cols);
  for (int e = 0; e < nell; ++e) {
    const double val = RAFT_MATRIX_AT(desc, e, 0), a = RAFT_MATRIX_AT(desc, e, 1), b = RAFT_MATRIX_AT(desc, e, 2);
    const double rot = RAFT_MATRIX_AT(desc, e, 5), ex = RAFT_MATRIX_AT(desc, e, 3), ey = RAFT_MATRIX_AT(desc, e, 4);
    for (int j = 0; j < cols; ++j) {
      double th = sino.tl_x + j * ths;
      const double c = cos(th), s = sin(th);
      th -= rot;
      const double cp = cos(th), sp = sin(th);
      double m = 1.0 / sqrt(((cp * cp) / (b * b)) + ((sp * sp) / (a * a)));
      double mab = m / (a * b);
      m *= 2.0;
      m *= val;
      const double delt = mab * ((ex * c) + (ey * s) - sino.tl_y);
      mab *= tss;
      views[(size_t)e * cols + j] = SSEllipseView{m, mab, delt};
    }
  }
  const size_t total = (size_t)lines * cols;
  std::vector<double> h(total);
  for (int i = 0; i < lines; ++i)
    for (int j = 0; j < cols; ++j) h[(size_t)i * cols + j] = RAFT_MATRIX_AT(sino.data, i, j);
  DevBuf d_views, d_out;
  HIP_TRY(d_views.alloc(sizeof(SSEllipseView) * views.size(), nullptr));
  HIP_TRY(d_out.alloc(sizeof(double) * total, nullptr));
  HIP_TRY(hipMemcpyAsync(d_views.p, views.data(), sizeof(SSEllipseView) * views.size(), hipMemcpyHostToDevice,
                         nullptr));
  HIP_TRY(hipMemcpyAsync(d_out.p, h.data(), sizeof(double) * total, hipMemcpyHostToDevice, nullptr));
  HIP_TRY(ss_launch_radon_fromdesc(d_views.as<SSEllipseView>(), nell, lines, cols, d_out.as<double>(), nullptr));
  HIP_TRY(hipMemcpyAsync(h.data(), d_out.p, sizeof(double) * total, hipMemcpyDeviceToHost, nullptr));
  HIP_TRY(hipStreamSynchronize(nullptr));
  for (int i = 0; i < lines; ++i)
    for (int j = 0; j < cols; ++j) RAFT_MATRIX_AT(sino.data, i, j) = h[(size_t)i * cols + j];
  return 0;
}

extern "C" void raft_radon_fromdesc(raft_image sino, raft_matrix desc) {
  report("raft_radon_fromdesc", host_radon_fromdesc(sino, desc));
}

extern "C" int raft_ss_version(void) { return 1; }
