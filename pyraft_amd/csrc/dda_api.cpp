// dda_api.cpp -- plans, device-pointer operators and drop-in symbols of the
// line-traversal (Bresenham / DDA) projector pair.
//
// The per-angle table repeats raft_ddaiterator_set_image / set_theta
// (raft/raft_ddaiterator.c:16-106) in FP64 with glibc cos/sin and no fused
// multiply-adds (this file is compiled with -ffp-contract=off), so the device traces
// from the same numbers the reference computes.
#include <hip/hip_runtime.h>
#include <math.h>
#include <string.h>

#include <algorithm>
#include <list>
#include <memory>
#include <mutex>
#include <vector>

#include "../../include/raft_bresenham.h"
#include "dda_kernels.h"
#include "ss_internal.h"

struct raft_dda_plan {
  int device = 0;
  SSDdaGeom g{};
  double img_corners[4] = {0, 0, 0, 0};
  double sino_corners[4] = {0, 0, 0, 0};
  std::vector<SSDdaAngle> h_angles;
  SSDdaAngle* d_angles = nullptr;
  std::vector<SSDdaGatherAngle> h_gather;  // angles of the pixel-driven transpose
  std::vector<int> h_walk;                 // angles traced ray by ray in the transpose
  std::vector<SSDdaStraight> h_straight;   // straight angles of the transpose (then no traced ones)
  std::vector<int> h_ranges;               // their rays per slow index
  SSDdaStraight* d_straight = nullptr;
  int* d_ranges = nullptr;
  SSDdaGatherAngle* d_gather = nullptr;
  int* d_walk = nullptr;
  float* d_row_scale = nullptr;  // [A] lmax of the gather angles
  int kmax = 0;
  int gather_ws = 0;  // segment length of the two-candidate gather, 0 when it does not apply
  double* d_weights = nullptr;  // R^T 1 in FP64, built on first use
  std::mutex weights_mutex;
  SSDdaTransposePlan tplan() const {
    return SSDdaTransposePlan{d_angles, d_gather, (int)h_gather.size(), kmax, gather_ws, d_straight,
                              (int)h_straight.size(), d_ranges, d_row_scale, d_walk, (int)h_walk.size()};
  }
  ~raft_dda_plan() {
    if (d_angles) (void)hipFree(d_angles);
    if (d_row_scale) (void)hipFree(d_row_scale);
    if (d_gather) (void)hipFree(d_gather);
    if (d_walk) (void)hipFree(d_walk);
    if (d_straight) (void)hipFree(d_straight);
    if (d_ranges) (void)hipFree(d_ranges);
    if (d_weights) (void)hipFree(d_weights);
  }
};

namespace {

double sgn(double v) { return v > 0.0 ? 1.0 : (v < 0.0 ? -1.0 : 0.0); }

SSDdaGeom make_geom(int lines, int cols, int nrays, int nangles, const double* ic, const double* sc) {
  SSDdaGeom g;
  g.lines = lines;
  g.cols = cols;
  g.nrays = nrays;
  g.nangles = nangles;
  g.hx = (ic[2] - ic[0]) / (cols - 1);
  g.hy = (ic[3] - ic[1]) / (lines - 1);
  g.ox = ic[0];
  g.oy = ic[1];
  g.bx0 = ic[0] - (g.hx / 2.0);
  g.bx1 = ic[2] + (g.hx / 2.0);
  g.by0 = ic[1] - (g.hy / 2.0);
  g.by1 = ic[3] + (g.hy / 2.0);
  g.t0 = sc[1];
  g.dt = (sc[3] - sc[1]) / (nrays - 1);
  return g;
}

// raft/raft_ddaiterator.c:38-106 for theta
SSDdaAngle make_angle(const SSDdaGeom& g, double theta) {
  SSDdaAngle a;
  a.c = cos(theta);
  a.s = sin(theta);
  int sg = (int)sgn(a.s);
  double p = sg * g.bx1, q = sg * g.bx0;
  a.xe = sg * (p > q ? p : q);
  a.dj = (int)(-sgn(g.hx) * sg);
  sg = (int)sgn(a.c);
  p = sg * g.by1;
  q = sg * g.by0;
  a.ye = sg * (p < q ? p : q);
  a.di = (int)(sgn(g.hy) * sg);
  a.i0 = a.di < 0 ? g.lines - 1 : 0;
  a.j0 = a.dj < 0 ? g.cols - 1 : 0;
  // the reference's fast-axis test (fast = line when |hx sin| <= |hy cos|): for
  // square pixels the axis crossed more often; kept as is for other pixel shapes
  a.fast_i = fabs(g.hx * a.s) <= fabs(g.hy * a.c);
  if (a.fast_i) {
    a.e_step = fabs(g.hy * a.s / a.c);
    a.e_cap = fabs(g.hx);
    a.dslow = a.dj;
    a.dfast = a.di;
    a.len_full = fabs(g.hy / a.c);
  } else {
    a.e_step = fabs(g.hx * a.c / a.s);
    a.e_cap = fabs(g.hy);
    a.dslow = a.di;
    a.dfast = a.dj;
    a.len_full = fabs(g.hx / a.s);
  }
  a.axis = a.s == 0.0 ? 1 : (a.c == 0.0 ? 2 : 0);
  a.k_len = a.e_step > 0.0 ? a.len_full / a.e_step : 0.0;
  a.fixed = 0;
  a.tr_wstage = a.tr_lowlane = a.tr_wgap = 0;
  a.tr_qchunk = 0;
  a.tr_inv = a.tr_off = a.tr_d = a.tr_lmax = 0.0;
  a.tr_dhi = 0;
  a.tr_dlo = 0u;
  a.tr_c1 = a.tr_kz = a.tr_pz = a.tr_pz1 = 0.f;
  return a;
}

// Splits the transpose's angles: the traversal equals exact w x h pixel geometry when
// its slow-axis drift per move does not exceed the pixel width across (always for
// square pixels; for other shapes the reference's fast-axis test can pick the axis
// crossed less often) and the angle is not within 1e-7 of an axis, where the
// traversal's tie rule at pixel edges decides which pixel a ray on an edge belongs to.
bool pixel_driven(const SSDdaGeom& g, const SSDdaAngle& a, SSDdaGatherAngle* out, int index) {
  const double ac = fabs(a.c), as = fabs(a.s);
  if (ac < 1e-7 || as < 1e-7 || !(a.e_step <= a.e_cap)) return false;
  const double w = fabs(g.hx), h = fabs(g.hy);
  SSDdaGatherAngle ga;
  ga.c = a.c;
  ga.s = a.s;
  ga.hys = g.hy * a.s;
  ga.p = 0.5 * (w * ac + h * as);
  const double q = 0.5 * fabs(w * ac - h * as);
  ga.inv_pq = 1.0 / (ga.p - q);
  ga.lmax = fmin(w / as, h / ac);
  ga.half = ga.p / fabs(g.dt);
  ga.k0h = (g.ox * a.c + g.oy * a.s - g.t0) / g.dt - ga.half;
  ga.kx = g.hx * a.c / g.dt;
  ga.ky = g.hy * a.s / g.dt;
  const double kz = fabs(g.dt) * ga.inv_pq, pz = ga.p * ga.inv_pq;
  ga.c1 = (float)(1.0 - ga.half);
  ga.kz = (float)kz;
  ga.pz = (float)pz;
  ga.pz1 = (float)(pz - kz * (2.0 - ga.half));
  ga.kyf = (float)ga.ky;
  ga.a = index;
  // rays with a nonzero length have |t - t_c| < p: the integers of the open interval
  // (rc - half, rc + half), at most ceil(2 half) of them, from floor(rc - half) + 1
  ga.k = (int)fmax(1.0, ceil(2.0 * ga.half));
  if (!isfinite(ga.inv_pq) || !isfinite(ga.half) || ga.k > 1 << 20) return false;
  *out = ga;
  return true;
}

// Whether every ray of the (traced) angle runs straight: walked as the transpose's
// scatter walks it, each ray inside the image visits the nf fast indices from its entry
// edge in order at one slow index, each with length len_full, and the rays of each slow
// index are consecutive.  Then appends the angle and its ray range per slow index.
bool straight_angle(const SSDdaGeom& g, const SSDdaAngle& a, int index, std::vector<SSDdaStraight>& out,
                    std::vector<int>& ranges) {
  const int nf = a.fast_i ? g.lines : g.cols;
  const int ns = a.fast_i ? g.cols : g.lines;
  const int f_edge = a.dfast > 0 ? 0 : nf - 1;
  std::vector<int> lo(ns, 0), hi(ns, 0);
  int last_s = -1;
  for (int r = 0; r < g.nrays; ++r) {
    SSDdaRay ray = dda_ray_start(g, a, g.t0 + (r * g.dt));
    if (!((unsigned)ray.f < (unsigned)nf && (unsigned)ray.s < (unsigned)ns)) continue;
    const int s0 = ray.s;
    int count = 0;
    for (int it = 0; it <= nf; ++it) {
      if (!((unsigned)ray.f < (unsigned)nf && (unsigned)ray.s < (unsigned)ns)) break;
      if (ray.s != s0 || ray.len != a.len_full || ray.f != f_edge + count * a.dfast || ray.slow_next) return false;
      ++count;
      dda_fast_move(a, ray);
    }
    if (count != nf) return false;
    if (s0 != last_s) {
      if (hi[s0] != 0) return false;  // this slow index had rays before another one
      lo[s0] = r;
      last_s = s0;
    }
    hi[s0] = r + 1;
  }
  SSDdaStraight st;
  st.a = index;
  st.slow_col = a.fast_i;
  st.off = (int)ranges.size();
  st.len = a.len_full;
  for (int s = 0; s < ns; ++s) {
    ranges.push_back(lo[s]);
    ranges.push_back(hi[s]);
  }
  out.push_back(st);
  return true;
}

// Trapezoid terms of an exact-geometry angle (ga: its pixel-driven record): the slow
// pitch times the trig factor (sigma per unit t is 1 / step), sigma per fast index, the
// half-width of the trapezoid's support in slow pixels, and z = clamp(pz - kz |u|).
struct TrapTerms {
  double step, d, half_s, kz, pz;
};
TrapTerms trap_terms(const SSDdaGeom& g, const SSDdaGatherAngle& ga, const SSDdaAngle& a) {
  TrapTerms r;
  r.step = a.fast_i ? g.hx * a.c : g.hy * a.s;
  r.d = -(a.fast_i ? g.hy * a.s : g.hx * a.c) / r.step;
  r.half_s = ga.p / fabs(r.step);
  r.kz = fabs(r.step) * ga.inv_pq;
  r.pz = ga.p * ga.inv_pq;
  return r;
}

// Marks an exact-geometry angle for the Radon's trapezoid form and sets its terms
void set_trapezoid(const SSDdaGeom& g, const SSDdaGatherAngle& ga, SSDdaAngle& a) {
  a.fixed = 0;
  const TrapTerms tt = trap_terms(g, ga, a);
  // exact geometry gives |d| <= 1 and half_s <= 1 (two candidates, u1 >= 0) up to rounding
  if (!isfinite(tt.d) || !(fabs(tt.d) <= 2.0) || !(tt.half_s <= 1.0 + 1e-9) || !isfinite(tt.kz)) return;
  const long long dfix = llround(tt.d * 4294967296.0);
  a.tr_inv = 1.0 / tt.step;
  a.tr_off = -(g.ox * a.c + g.oy * a.s) / tt.step - tt.half_s + (DDA_PADW + 1.0);
  a.tr_d = tt.d;
  a.tr_lmax = ga.lmax;
  a.tr_dhi = (int)(dfix >> 32);
  a.tr_dlo = (unsigned)(dfix & 0xffffffffLL);
  a.tr_c1 = (float)((1.0 - tt.half_s) * 4294967296.0);
  a.tr_kz = (float)(tt.kz / 4294967296.0);
  a.tr_pz = (float)tt.pz;
  a.tr_pz1 = (float)(tt.pz - tt.kz * (2.0 - tt.half_s));
  const double spread = 63.0 * fabs(g.dt * a.tr_inv);  // tau spread over a wave's rays
  a.tr_wstage = spread + (DDA_WS_ROWS - 1) * fabs(a.tr_d) + 9.0 <= DDA_WS_SEG - 2.0 ? 1 : 0;
  a.tr_lowlane = g.dt * a.tr_inv >= 0.0 ? 0 : 63;
  a.tr_wgap = a.tr_d < 0.0 ? (int)floor((DDA_WS_ROWS - 1) * fabs(a.tr_d)) + 1 : 0;
  a.tr_qchunk = dfix * DDA_WS_ROWS;
  a.fixed = 1;
}

int check_dda(const raft_dda_plan* p, int nslices) {
  if (!p) return ss_set_error(RAFT_SS_EINVAL, "null plan");
  if (nslices < 1 || nslices > 65535)
    return ss_set_error(RAFT_SS_EINVAL, "nslices must be in [1, 65535] (got %d)", nslices);
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev != p->device)
    return ss_set_error(RAFT_SS_EINVAL, "plan was built on device %d but device %d is current", p->device, dev);
  return 0;
}

// R^T 1 for the plan in FP64, built once (atomics of the ray-traced angles: the first
// call fixes their rounding)
int ensure_weights(raft_dda_plan* p, hipStream_t stream) {
  std::lock_guard<std::mutex> lock(p->weights_mutex);
  if (p->d_weights) return 0;
  const size_t npix = (size_t)p->g.lines * p->g.cols;
  double* w = nullptr;
  HIP_TRY(hipMalloc(&w, sizeof(double) * npix));
  DevBuf sinoT, acc, accT;
  hipError_t e = sinoT.alloc(sizeof(float) * (size_t)(p->g.nrays + 4) * p->g.nangles, stream);
  if (e == hipSuccess) e = acc.alloc(sizeof(double) * npix, stream);
  if (e == hipSuccess) e = accT.alloc(sizeof(double) * npix, stream);
  if (e == hipSuccess)
    e = dda_launch_transpose<double, double>(p->g, p->tplan(), sinoT.as<float>(), true, sinoT.as<float>(),
                                             acc.as<double>(), accT.as<double>(), nullptr, w, 1, stream);
  if (e == hipSuccess) e = hipStreamSynchronize(stream);
  if (e != hipSuccess) {
    (void)hipFree(w);
    return ss_set_error(RAFT_SS_EHIP, "building the backprojection weights failed: %s", hipGetErrorString(e));
  }
  p->d_weights = w;
  return 0;
}

// out[s] = R^T sino[s] [/ w] with accumulator T
template <typename T, typename Out>
int transpose_into(const raft_dda_plan* p, const float* sino, const double* w, Out* out, int nslices,
                   hipStream_t stream) {
  const size_t n = (size_t)p->g.lines * p->g.cols * nslices;
  DevBuf sinoT, acc, accT;
  HIP_TRY(sinoT.alloc(sizeof(float) * (size_t)(p->g.nrays + 4) * p->g.nangles * nslices, stream));
  HIP_TRY(acc.alloc(sizeof(T) * n, stream));
  HIP_TRY(accT.alloc(sizeof(T) * n, stream));
  HIP_TRY((dda_launch_transpose<T, Out>(p->g, p->tplan(), sino, false, sinoT.as<float>(), acc.as<T>(), accT.as<T>(),
                                        w, out, nslices, stream)));
  return 0;
}

}  // namespace

extern "C" int raft_dda_plan_create(raft_dda_plan** out, int lines, int cols, int nrays, int nangles,
                                    const double* img_corners, const double* sino_corners, void* stream_) {
  if (!out || !img_corners || !sino_corners) return ss_set_error(RAFT_SS_EINVAL, "null argument");
  *out = nullptr;
  if (lines < 2 || cols < 2 || nrays < 2 || nangles < 2)
    return ss_set_error(RAFT_SS_EINVAL, "line traversal needs lines, cols, nrays, nangles >= 2 (got %d, %d, %d, %d)",
                        lines, cols, nrays, nangles);
  if (lines > 65536 || cols > 65536 || nrays > (1 << 24) || nangles > 65535)
    return ss_set_error(RAFT_SS_EINVAL, "geometry too large (%d x %d image, %d x %d sinogram)", lines, cols, nrays,
                        nangles);
  for (int i = 0; i < 4; ++i)
    if (!isfinite(img_corners[i]) || !isfinite(sino_corners[i]))
      return ss_set_error(RAFT_SS_EINVAL, "non-finite corner");
  hipStream_t stream = as_stream(stream_);
  std::unique_ptr<raft_dda_plan> p(new raft_dda_plan());
  HIP_TRY(hipGetDevice(&p->device));
  memcpy(p->img_corners, img_corners, sizeof(p->img_corners));
  memcpy(p->sino_corners, sino_corners, sizeof(p->sino_corners));
  p->g = make_geom(lines, cols, nrays, nangles, img_corners, sino_corners);
  const SSDdaGeom& g = p->g;
  if (!(g.hx != 0.0) || !(g.hy != 0.0) || !isfinite(g.hx) || !isfinite(g.hy) || !isfinite(g.dt))
    return ss_set_error(RAFT_SS_EINVAL, "degenerate image or detector extent");
  const double dth = (sino_corners[2] - sino_corners[0]) / (nangles - 1);
  p->h_angles.resize(nangles);
  for (int a = 0; a < nangles; ++a) {
    p->h_angles[a] = make_angle(g, sino_corners[0] + (a * dth));
    if (!isfinite(p->h_angles[a].c) || !isfinite(p->h_angles[a].s))
      return ss_set_error(RAFT_SS_EINVAL, "non-finite angle %d", a);
  }
  for (int a = 0; a < nangles; ++a) {
    SSDdaGatherAngle ga;
    SSDdaAngle& an = p->h_angles[a];
    if (pixel_driven(g, an, &ga, a)) {
      p->h_gather.push_back(ga);
      set_trapezoid(g, ga, an);
    } else {
      p->h_walk.push_back(a);
    }
  }
  // traced angles that all run straight (the axis angles of square pixels) become
  // vectors over the slow index; the most tilted are tried first, as they fail soonest
  if (!p->h_walk.empty()) {
    std::vector<int> order = p->h_walk;
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
      const SSDdaAngle &ax = p->h_angles[x], &ay = p->h_angles[y];
      return fmin(fabs(ax.c), fabs(ax.s)) > fmin(fabs(ay.c), fabs(ay.s));
    });
    bool all = true;
    for (int a : order)
      if (!straight_angle(g, p->h_angles[a], a, p->h_straight, p->h_ranges)) {
        all = false;
        break;
      }
    if (all) {
      p->h_walk.clear();
    } else {
      p->h_straight.clear();
      p->h_ranges.clear();
    }
  }
  std::vector<float> row_scale(nangles, 0.f);
  int ws = 8;
  for (const SSDdaGatherAngle& ga : p->h_gather) {
    row_scale[ga.a] = (float)ga.lmax;
    p->kmax = std::max(p->kmax, ga.k);
    ws = std::max(ws, dda_gather2_ws(ga));
  }
  p->gather_ws = p->kmax <= 2 && ws <= DDA_G2_WSMAX ? ws : 0;
  HIP_TRY(hipMalloc(&p->d_row_scale, sizeof(float) * nangles));
  HIP_TRY(hipMemcpyAsync(p->d_row_scale, row_scale.data(), sizeof(float) * nangles, hipMemcpyHostToDevice, stream));
  HIP_TRY(hipMalloc(&p->d_angles, sizeof(SSDdaAngle) * nangles));
  HIP_TRY(hipMalloc(&p->d_gather, sizeof(SSDdaGatherAngle) * (p->h_gather.size() + 1)));
  HIP_TRY(hipMalloc(&p->d_walk, sizeof(int) * (p->h_walk.size() + 1)));
  HIP_TRY(hipMemcpyAsync(p->d_angles, p->h_angles.data(), sizeof(SSDdaAngle) * nangles, hipMemcpyHostToDevice,
                         stream));
  if (!p->h_gather.empty())
    HIP_TRY(hipMemcpyAsync(p->d_gather, p->h_gather.data(), sizeof(SSDdaGatherAngle) * p->h_gather.size(),
                           hipMemcpyHostToDevice, stream));
  if (!p->h_walk.empty())
    HIP_TRY(hipMemcpyAsync(p->d_walk, p->h_walk.data(), sizeof(int) * p->h_walk.size(), hipMemcpyHostToDevice,
                           stream));
  if (!p->h_straight.empty()) {
    HIP_TRY(hipMalloc(&p->d_straight, sizeof(SSDdaStraight) * p->h_straight.size()));
    HIP_TRY(hipMalloc(&p->d_ranges, sizeof(int) * p->h_ranges.size()));
    HIP_TRY(hipMemcpyAsync(p->d_straight, p->h_straight.data(), sizeof(SSDdaStraight) * p->h_straight.size(),
                           hipMemcpyHostToDevice, stream));
    HIP_TRY(hipMemcpyAsync(p->d_ranges, p->h_ranges.data(), sizeof(int) * p->h_ranges.size(), hipMemcpyHostToDevice,
                           stream));
  }
  HIP_TRY(hipStreamSynchronize(stream));
  *out = p.release();
  return 0;
}

extern "C" int raft_dda_plan_destroy(raft_dda_plan* p) {
  delete p;
  return 0;
}

extern "C" int raft_dda_plan_info(const raft_dda_plan* p, int* info, int ninfo) {
  if (!p || !info) return ss_set_error(RAFT_SS_EINVAL, "null argument");
  int nfast = 0;
  for (const SSDdaAngle& a : p->h_angles) nfast += a.fast_i;
  const int vals[] = {p->g.lines, p->g.cols, p->g.nrays, p->g.nangles, nfast, (int)p->h_walk.size(),
                      (int)p->h_straight.size()};
  for (int i = 0; i < ninfo && i < (int)(sizeof(vals) / sizeof(vals[0])); ++i) info[i] = vals[i];
  return 0;
}

extern "C" int raft_dda_radon(const raft_dda_plan* p, const float* img, float* sino, int nslices, void* stream_) {
  int rc = check_dda(p, nslices);
  if (rc) return rc;
  if (!img || !sino) return ss_set_error(RAFT_SS_EINVAL, "null buffer");
  hipStream_t stream = as_stream(stream_);
  DevBuf pad;
  HIP_TRY(pad.alloc(sizeof(float) * dda_radon_scratch_floats(p->g, nslices), stream));
  HIP_TRY(dda_launch_radon(p->g, p->d_angles, img, pad.as<float>(), sino, nslices, stream));
  return 0;
}

extern "C" int raft_dda_radon_transpose(const raft_dda_plan* p, const float* sino, float* img, int nslices,
                                        void* stream_) {
  int rc = check_dda(p, nslices);
  if (rc) return rc;
  if (!img || !sino) return ss_set_error(RAFT_SS_EINVAL, "null buffer");
  return transpose_into<float, float>(p, sino, nullptr, img, nslices, as_stream(stream_));
}

extern "C" int raft_dda_backprojection(raft_dda_plan* p, const float* sino, float* img, int nslices, void* stream_) {
  int rc = check_dda(p, nslices);
  if (rc) return rc;
  if (!img || !sino) return ss_set_error(RAFT_SS_EINVAL, "null buffer");
  hipStream_t stream = as_stream(stream_);
  if ((rc = ensure_weights(p, stream))) return rc;
  return transpose_into<double, float>(p, sino, p->d_weights, img, nslices, stream);
}

extern "C" int raft_dda_weights(raft_dda_plan* p, float* out, void* stream_) {
  int rc = check_dda(p, 1);
  if (rc) return rc;
  if (!out) return ss_set_error(RAFT_SS_EINVAL, "null buffer");
  hipStream_t stream = as_stream(stream_);
  if ((rc = ensure_weights(p, stream))) return rc;
  HIP_TRY(dda_launch_to_f32(p->d_weights, out, (size_t)p->g.lines * p->g.cols, stream));
  return 0;
}

// One ray of the Radon exactly as k_dda_radon sums it, evaluated on the host (for
// checking index logic without a device): the (line, column) pairs and lengths of the
// pixels it adds, in order.  Angles traced move by move: the visited pixels with their
// lengths.  Exact-geometry angles (trapezoid form): per fast index of the ray's range,
// the candidates inside the image with a nonzero weight, the Q32 stepping as on the
// device, the weights evaluated in FP64 from the Q32 fraction (the device evaluates
// them in FP32).  Returns the pixel count (at most cap are written).
extern "C" int raft_dda_host_ray(int lines, int cols, int nrays, int nangles, const double* img_corners,
                                 const double* sino_corners, int angle, int ray, int* ij, double* len, int cap) {
  if (!img_corners || !sino_corners || (cap > 0 && (!ij || !len))) return ss_set_error(RAFT_SS_EINVAL, "null argument");
  if (lines < 2 || cols < 2 || nrays < 2 || nangles < 2 || angle < 0 || angle >= nangles || ray < 0 || ray >= nrays)
    return ss_set_error(RAFT_SS_EINVAL, "bad geometry or ray index");
  const SSDdaGeom g = make_geom(lines, cols, nrays, nangles, img_corners, sino_corners);
  const double dth = (sino_corners[2] - sino_corners[0]) / (nangles - 1);
  SSDdaAngle a = make_angle(g, sino_corners[0] + (angle * dth));
  SSDdaGatherAngle ga;
  if (pixel_driven(g, a, &ga, angle)) set_trapezoid(g, ga, a);
  const int nf = a.fast_i ? lines : cols, ns = a.fast_i ? cols : lines;
  const double t = g.t0 + (ray * g.dt);
  int count = 0;
  auto emit = [&](int f, int s, double l) {
    if (count < cap) {
      ij[2 * count] = a.fast_i ? f : s;
      ij[2 * count + 1] = a.fast_i ? s : f;
      len[count] = l;
    }
    ++count;
  };
  if (a.fixed) {
    const TrapTerms tt = trap_terms(g, ga, a);
    const int pitch = dda_pad_pitch(ns);
    auto tau_of = [&](int rr) { return fma(g.t0 + (rr * g.dt), a.tr_inv, a.tr_off); };
    // the wave's fast range: the union over its 64 rays (those below nrays)
    const int w0 = ray & ~63;
    int klo = nf, khi = -1;
    for (int l = 0; l < 64 && w0 + l < nrays; ++l) {
      const double tl = tau_of(w0 + l);
      const double ka = (DDA_PADW - 1.0 - tl) / a.tr_d, kb = (ns + DDA_PADW - tl) / a.tr_d;
      const double kl = fmax(fmin(ka, kb), -1.0), kh = fmin(fmax(ka, kb), (double)nf);
      if (!(kl <= kh)) continue;
      klo = std::min(klo, (int)fmax(0.0, floor(kl) - 1.0));
      khi = std::max(khi, (int)fmin(nf - 1.0, ceil(kh) + 1.0));
    }
    if (klo > khi) return 0;
    auto start = [&](double tau) {
      const double v = fmin(fmax(fma((double)klo, a.tr_d, tau), -1073741824.0), 1073741824.0);
      const double fv = floor(v);
      return std::make_pair((int)fv, (unsigned)((v - fv) * 4294967296.0));
    };
    auto step = [&](std::pair<int, unsigned>& st) {
      const unsigned nlo = st.second + a.tr_dlo;
      st.first += a.tr_dhi + (nlo < st.second ? 1 : 0);
      st.second = nlo;
    };
    std::pair<int, unsigned> st = start(tau_of(ray)), low = start(tau_of(w0 + a.tr_lowlane));
    int sb = 0;
    for (int k = klo; k <= khi; ++k) {
      int idx = st.first;  // padded slow index of the first candidate
      if (a.tr_wstage) {
        // the chunk's staged segment holds both candidates
        if ((k - klo) % DDA_WS_ROWS == 0) sb = (low.first - 1 - a.tr_wgap) & ~3;
        if (idx - sb < 0 || idx + 1 - sb > DDA_WS_SEG - 1 || sb < 0 || sb + DDA_WS_SEG > pitch)
          return ss_set_error(RAFT_SS_EINTERNAL, "ray %d angle %d: candidate %d outside the staged segment [%d, %d)",
                              ray, angle, idx, sb, sb + DDA_WS_SEG);
      } else {
        idx = (int)std::min((unsigned)idx, (unsigned)(pitch - 2));  // clamped into the padding
      }
      const double fr = st.second / 4294967296.0;
      const double u0 = fabs(1.0 - tt.half_s - fr), u1 = 2.0 - tt.half_s - fr;
      const double z0 = fmin(1.0, fmax(0.0, tt.pz - tt.kz * u0)), z1 = fmin(1.0, fmax(0.0, tt.pz - tt.kz * u1));
      if ((unsigned)(idx - DDA_PADW) < (unsigned)ns && z0 > 0.0) emit(k, idx - DDA_PADW, a.tr_lmax * z0);
      if ((unsigned)(idx + 1 - DDA_PADW) < (unsigned)ns && z1 > 0.0) emit(k, idx + 1 - DDA_PADW, a.tr_lmax * z1);
      step(st);
      step(low);
    }
    return count;
  }
  SSDdaRay r = dda_ray_start(g, a, t);
  auto inside = [&](void) { return (unsigned)r.f < (unsigned)nf && (unsigned)r.s < (unsigned)ns; };
  for (int it = 0; it <= nf; ++it) {
    if (!inside()) break;
    emit(r.f, r.s, r.len);
    if (r.slow_next) {
      r.s = dda_step(r.s, a.dslow);
      r.len = a.len_full - r.len;
      r.slow_next = 0;
      if (!inside()) break;
      emit(r.f, r.s, r.len);
    }
    dda_fast_move(a, r);
  }
  return count;
}

// ------------------------------------------------------------------ drop-in symbols

namespace {

struct DdaKey {
  int device, lines, cols, R, A;
  double c[8];
  bool operator==(const DdaKey& o) const {
    return device == o.device && lines == o.lines && cols == o.cols && R == o.R && A == o.A &&
           memcmp(c, o.c, sizeof(c)) == 0;
  }
};
std::mutex g_dda_cache_mutex;
std::list<std::pair<DdaKey, std::shared_ptr<raft_dda_plan>>> g_dda_cache;

// plans (with their weights) of recent geometries, shared by the drop-in calls
int cached_dda_plan(const raft_image& img, const raft_image& sino, std::shared_ptr<raft_dda_plan>* out) {
  DdaKey key;
  memset(&key, 0, sizeof(key));
  HIP_TRY(hipGetDevice(&key.device));
  key.lines = img.data.lines;
  key.cols = img.data.columns;
  key.R = sino.data.lines;
  key.A = sino.data.columns;
  corners_of(img, key.c);
  corners_of(sino, key.c + 4);
  {
    std::lock_guard<std::mutex> lock(g_dda_cache_mutex);
    for (auto it = g_dda_cache.begin(); it != g_dda_cache.end(); ++it)
      if (it->first == key) {
        g_dda_cache.splice(g_dda_cache.begin(), g_dda_cache, it);
        *out = g_dda_cache.front().second;
        return 0;
      }
  }
  raft_dda_plan* raw = nullptr;
  int rc = raft_dda_plan_create(&raw, key.lines, key.cols, key.R, key.A, key.c, key.c + 4, nullptr);
  if (rc) return rc;
  std::shared_ptr<raft_dda_plan> sp(raw);
  std::lock_guard<std::mutex> lock(g_dda_cache_mutex);
  g_dda_cache.emplace_front(key, sp);
  while (g_dda_cache.size() > 8) g_dda_cache.pop_back();
  *out = sp;
  return 0;
}

int check_images(const raft_image& img, const raft_image& sino) {
  if (!img.data.p_data || !sino.data.p_data) return ss_set_error(RAFT_SS_EINVAL, "empty image");
  return 0;
}

int host_dda_radon(const raft_image& img, const raft_image& sino) {
  int rc = check_images(img, sino);
  if (rc) return rc;
  std::shared_ptr<raft_dda_plan> plan;
  if ((rc = cached_dda_plan(img, sino, &plan))) return rc;
  const size_t npix = (size_t)img.data.lines * img.data.columns, nsino = (size_t)sino.data.lines * sino.data.columns;
  PinnedF32 hi, hs;
  HIP_TRY(hi.alloc(npix));
  HIP_TRY(hs.alloc(nsino));
  gather_f64(img.data, hi.p);
  DevBuf di, ds;
  HIP_TRY(di.alloc(sizeof(float) * npix, nullptr));
  HIP_TRY(ds.alloc(sizeof(float) * nsino, nullptr));
  HIP_TRY(hipMemcpyAsync(di.p, hi.p, sizeof(float) * npix, hipMemcpyHostToDevice, nullptr));
  if ((rc = raft_dda_radon(plan.get(), di.as<float>(), ds.as<float>(), 1, nullptr))) return rc;
  HIP_TRY(hipMemcpyAsync(hs.p, ds.p, sizeof(float) * nsino, hipMemcpyDeviceToHost, nullptr));
  HIP_TRY(hipStreamSynchronize(nullptr));
  scatter_f64(hs.p, sino.data);
  return 0;
}

// image += R^T sino
int host_dda_transpose(const raft_image& sino, const raft_image& img) {
  int rc = check_images(img, sino);
  if (rc) return rc;
  std::shared_ptr<raft_dda_plan> plan;
  if ((rc = cached_dda_plan(img, sino, &plan))) return rc;
  const size_t npix = (size_t)img.data.lines * img.data.columns, nsino = (size_t)sino.data.lines * sino.data.columns;
  PinnedF32 hs, ho;
  HIP_TRY(hs.alloc(nsino));
  HIP_TRY(ho.alloc(npix));
  gather_f64(sino.data, hs.p);
  DevBuf ds, dout;
  HIP_TRY(ds.alloc(sizeof(float) * nsino, nullptr));
  HIP_TRY(dout.alloc(sizeof(float) * npix, nullptr));
  HIP_TRY(hipMemcpyAsync(ds.p, hs.p, sizeof(float) * nsino, hipMemcpyHostToDevice, nullptr));
  if ((rc = raft_dda_radon_transpose(plan.get(), ds.as<float>(), dout.as<float>(), 1, nullptr))) return rc;
  HIP_TRY(hipMemcpyAsync(ho.p, dout.p, sizeof(float) * npix, hipMemcpyDeviceToHost, nullptr));
  HIP_TRY(hipStreamSynchronize(nullptr));
  const raft_matrix& m = img.data;
  for (int i = 0; i < m.lines; ++i)
    for (int j = 0; j < m.columns; ++j) RAFT_MATRIX_AT(m, i, j) += (double)ho.p[(size_t)i * m.columns + j];
  return 0;
}

// image = (image + R^T sino) / R^T 1, both sums in FP64
int host_dda_backprojection(const raft_image& sino, const raft_image& img) {
  int rc = check_images(img, sino);
  if (rc) return rc;
  std::shared_ptr<raft_dda_plan> plan;
  if ((rc = cached_dda_plan(img, sino, &plan))) return rc;
  if ((rc = check_dda(plan.get(), 1))) return rc;
  if ((rc = ensure_weights(plan.get(), nullptr))) return rc;
  const size_t npix = (size_t)img.data.lines * img.data.columns, nsino = (size_t)sino.data.lines * sino.data.columns;
  PinnedF32 hs;
  HIP_TRY(hs.alloc(nsino));
  gather_f64(sino.data, hs.p);
  std::vector<double> num(npix), w(npix);
  DevBuf ds, dnum;
  HIP_TRY(ds.alloc(sizeof(float) * nsino, nullptr));
  HIP_TRY(dnum.alloc(sizeof(double) * npix, nullptr));
  HIP_TRY(hipMemcpyAsync(ds.p, hs.p, sizeof(float) * nsino, hipMemcpyHostToDevice, nullptr));
  if ((rc = transpose_into<double, double>(plan.get(), ds.as<float>(), nullptr, dnum.as<double>(), 1, nullptr)))
    return rc;
  HIP_TRY(hipMemcpyAsync(num.data(), dnum.p, sizeof(double) * npix, hipMemcpyDeviceToHost, nullptr));
  HIP_TRY(hipMemcpyAsync(w.data(), plan->d_weights, sizeof(double) * npix, hipMemcpyDeviceToHost, nullptr));
  HIP_TRY(hipStreamSynchronize(nullptr));
  const raft_matrix& m = img.data;
  for (int i = 0; i < m.lines; ++i)
    for (int j = 0; j < m.columns; ++j) {
      const size_t k = (size_t)i * m.columns + j;
      RAFT_MATRIX_AT(m, i, j) = (RAFT_MATRIX_AT(m, i, j) + num[k]) / w[k];
    }
  return 0;
}

}  // namespace

extern "C" void raft_radon_bresenham(raft_image image, raft_image sino, int nthreads) {
  (void)nthreads;
  ss_report_legacy("raft_radon_bresenham", host_dda_radon(image, sino));
}

extern "C" void raft_radon_transpose_bresenham(raft_image sino, raft_image image, int nthreads) {
  (void)nthreads;
  ss_report_legacy("raft_radon_transpose_bresenham", host_dda_transpose(sino, image));
}

extern "C" void raft_backprojection_bresenham(raft_image sino, raft_image image, int nthreads) {
  (void)nthreads;
  ss_report_legacy("raft_backprojection_bresenham", host_dda_backprojection(sino, image));
}
