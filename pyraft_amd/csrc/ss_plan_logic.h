// ss_plan_logic.h -- FP64 geometry logic shared by the device plan kernel and the
// host-side self-check (both compile the same functions; no fused multiply-adds).
//
// The reference decides whether sample s of ray (j,k) contributes with
//   branch A: x = xmin + s*dx;  y = t/sin - x/tan;  m = floor((y - ymin)/dy)
//   branch B: y = xmin + s*dy;  x = t/cos - y*tan;  m = floor((x - xmin)/dx)
//   contributes iff 0 <= m <= N-2                 (raft/raft_radon_slantstack.c:179-209)
// and pixel (r,c) takes angle j in the backprojection iff
//   t = x*cos + y*sin;  m = floor((t - tmin)/dt);  0 <= m <= R-2
//                                                (raft/raft_backprojection_slantstack.c:155-177)
// These functions repeat those FP64 expressions operation for operation.
#ifndef PYRAFT_AMD_SS_PLAN_LOGIC_H
#define PYRAFT_AMD_SS_PLAN_LOGIC_H

#include <math.h>
#include <stdint.h>

#include "ss_geometry.h"

#if defined(__HIPCC__)
#define SS_HD __host__ __device__
#else
#define SS_HD
#endif

#if defined(__clang__)
#define SS_NOCONTRACT _Pragma("clang fp contract(off)")
#else
#define SS_NOCONTRACT
#endif

// the detector coordinate of ray k on the uniform grid (raft/raft_radon_slantstack.c:170-172)
static inline SS_HD double ss_ray_t(const SSGeom& g, int k) {
  SS_NOCONTRACT
  return g.tmin + k * g.dt;
}

// sample s of the ray at detector coordinate t; also the rule of the arbitrary-mesh
// forms, which take t and theta from tables (raft/raft_radon_slantstack_withmesh.c:176-224).
// floor() result compared in double: equivalent to the reference's (int) cast plus
// integer compare for every in-range value, and excludes NaN / out-of-range values
// (the reference's cvttsd2si yields INT_MIN for those, which also fails the test).
static inline SS_HD bool ss_radon_sample_valid_t(const SSGeom& g, const SSAngle& a, double t, int s) {
  SS_NOCONTRACT
  double v;
  if (a.branch_a) {
    const double x = g.xmin + s * g.dx;
    const double y = t / a.sint - x / a.tant;
    v = floor((y - g.ymin) / g.dy);
  } else {
    const double y = g.xmin + s * g.dy;
    const double x = t / a.cost - y * a.tant;
    v = floor((x - g.xmin) / g.dx);
  }
  return v >= 0.0 && v <= (double)(g.n - 2);
}

static inline SS_HD bool ss_radon_sample_valid(const SSGeom& g, const SSAngle& a, int k, int s) {
  return ss_radon_sample_valid_t(g, a, ss_ray_t(g, k), s);
}

static inline SS_HD bool ss_bp_sample_valid(const SSGeom& g, double cost, double sint, int row, int col) {
  SS_NOCONTRACT
  const double x = g.xmin + col * g.dx;
  const double y = g.ymin + row * g.dy;
  const double t = x * cost + y * sint;
  const double v = floor((t - g.tmin) / g.dt);
  return v >= 0.0 && v <= (double)(g.nrays - 2);
}

// Real-valued trajectory u(k, s) = alpha0 + k*gamma + s*beta of the interpolation
// coordinate (in grid units) along ray (j,k).
struct SSTraj {
  double alpha0, gamma, beta;
};

// u at s = 0 of the ray at detector coordinate t (any t: the mesh forms)
static inline SS_HD double ss_ray_alpha(const SSGeom& g, const SSAngle& a, double t) {
  SS_NOCONTRACT
  if (a.branch_a) return (t / a.sint - g.xmin / a.tant - g.ymin) / g.dy;
  return (t / a.cost - g.xmin * a.tant - g.xmin) / g.dx;
}

static inline SS_HD SSTraj ss_trajectory(const SSGeom& g, const SSAngle& a) {
  SS_NOCONTRACT
  SSTraj r;
  if (a.branch_a) {
    r.alpha0 = (g.tmin / a.sint - g.xmin / a.tant - g.ymin) / g.dy;
    r.gamma = g.dt / (a.sint * g.dy);
    r.beta = -g.dx / (a.tant * g.dy);
  } else {
    r.alpha0 = (g.tmin / a.cost - g.xmin * a.tant - g.xmin) / g.dx;
    r.gamma = g.dt / (a.cost * g.dx);
    r.beta = -(g.dy * a.tant) / g.dx;
  }
  return r;
}

// Slopes below this (grid units per sample) are "degenerate": the reference's FP64
// rounding noise (~1e-12 grid units) can then make the valid set non-contiguous.
#define SS_DEGENERATE_SLOPE 1e-7

// Exact contiguous valid range [lo, hi] of the ray at detector coordinate t whose
// real trajectory is u = u0 + s*beta.  Returns 0 on success with hi < lo for an
// empty ray; returns 1 when the valid set is not a single interval (the ray is then
// handled sample by sample).
static inline SS_HD int ss_ray_range_t(const SSGeom& g, const SSAngle& a, double beta, double t, double u0,
                                       int* lo_out, int* hi_out);

static inline SS_HD int ss_ray_range(const SSGeom& g, const SSAngle& a, const SSTraj& tr, int k, int* lo_out,
                                     int* hi_out) {
  return ss_ray_range_t(g, a, tr.beta, ss_ray_t(g, k), tr.alpha0 + k * tr.gamma, lo_out, hi_out);
}

static inline SS_HD int ss_ray_range_t(const SSGeom& g, const SSAngle& a, double beta, double t, double u0,
                                       int* lo_out, int* hi_out) {
  const int n = g.n;
  *lo_out = 0;
  *hi_out = -1;
  if (a.degenerate) {
    int first = -1, last = -1, count = 0;
    for (int s = 0; s < n; ++s) {
      if (ss_radon_sample_valid_t(g, a, t, s)) {
        if (first < 0) first = s;
        last = s;
        ++count;
      }
    }
    if (count == 0) return 0;
    if (count != last - first + 1) return 1;
    *lo_out = first;
    *hi_out = last;
    return 0;
  }
  // valid iff 0 <= u < n-1 with u = u0 + s*beta (beta != 0); real-valued interval first
  const double sa = (0.0 - u0) / beta;
  const double sb = ((double)(n - 1) - u0) / beta;
  double s_lo = sa < sb ? sa : sb;
  double s_hi = sa < sb ? sb : sa;
  if (!(s_hi >= -3.0) || !(s_lo <= (double)n + 2.0)) return 0;  // misses the image entirely
  if (s_lo < -3.0) s_lo = -3.0;
  if (s_hi > (double)n + 2.0) s_hi = (double)n + 2.0;
  int glo = (int)ceil(s_lo);
  int ghi = (int)floor(s_hi);
  if (glo < 0) glo = 0;
  if (ghi > n - 1) ghi = n - 1;
  // refine each end inside a +-3 window with the exact predicate
  int lo = -1, hi = -1;
  {
    int b = glo - 3 < 0 ? 0 : glo - 3;
    int e = glo + 3 > n - 1 ? n - 1 : glo + 3;
    for (int s = b; s <= e; ++s)
      if (ss_radon_sample_valid_t(g, a, t, s)) {
        lo = s;
        break;
      }
  }
  {
    int b = ghi + 3 > n - 1 ? n - 1 : ghi + 3;
    int e = ghi - 3 < 0 ? 0 : ghi - 3;
    for (int s = b; s >= e; --s)
      if (ss_radon_sample_valid_t(g, a, t, s)) {
        hi = s;
        break;
      }
  }
  if (lo < 0 && hi < 0) {
    // nothing found near either end: empty unless the estimate is badly off
    const int mid = (glo + ghi) / 2;
    if (glo <= ghi && mid >= 0 && mid < n && ss_radon_sample_valid_t(g, a, t, mid)) return 1;
    return 0;
  }
  if (lo < 0 || hi < 0 || hi < lo) return 1;
  // verify the interval is maximal and its interior holds
  if (lo > 0 && ss_radon_sample_valid_t(g, a, t, lo - 1)) return 1;
  if (hi < n - 1 && ss_radon_sample_valid_t(g, a, t, hi + 1)) return 1;
  if (!ss_radon_sample_valid_t(g, a, t, (lo + hi) / 2)) return 1;
  *lo_out = lo;
  *hi_out = hi;
  return 0;
}

#endif
