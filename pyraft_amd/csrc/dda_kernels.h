// dda_kernels.h -- line-traversal (Bresenham / DDA) projector pair: geometry records,
// the FP64 ray-entry logic shared by host and device, and the kernel launchers.
//
// Reference: raft/raft_ddaiterator.c:16-257 (traversal), raft/raft_radon_bresenham.cpp,
// raft/raft_radon_transpose_bresenham.cpp, raft/raft_backprojection_bresenham.cpp.
//
// Geometry (raft_image corners, signed pitches; raft/raft_image.h:56-61):
//   image line i <-> y = tl_y + i*hy, hy = (br_y - tl_y)/(lines - 1)
//   image column j <-> x = tl_x + j*hx, hx = (br_x - tl_x)/(cols - 1)
//   sinogram column a <-> theta = stl_x + a*(sbr_x - stl_x)/(A - 1)
//   sinogram row r    <-> t     = stl_y + r*(sbr_y - stl_y)/(R - 1)
// Ray {x cos(theta) + y sin(theta) = t} visits pixels along a "fast" axis (one index
// step per move) and a "slow" axis (a step whenever the accumulated drift e passes
// the pixel width across); each visited pixel receives its intersection length.
//
// Kernels work in (fast, slow) coordinates: a per-angle choice of the row-major image
// (fast = line) or its transpose (fast = column), so the 64 parallel rays of a wave
// read and write along a contiguous line of pixels.
#ifndef PYRAFT_AMD_DDA_KERNELS_H
#define PYRAFT_AMD_DDA_KERNELS_H

#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#if defined(__HIPCC__)
#define DDA_HD __host__ __device__
#else
#define DDA_HD
#endif

#if defined(__clang__)
#define DDA_NOCONTRACT _Pragma("clang fp contract(off)")
#else
#define DDA_NOCONTRACT
#endif

struct SSDdaGeom {
  int lines, cols, nrays, nangles;
  double hx, hy;            // signed pixel pitches
  double ox, oy;            // tl corner: centre of line 0 / column 0
  double bx0, bx1, by0, by1;  // pixel-edge box: tl_x - hx/2, br_x + hx/2, tl_y - hy/2, br_y + hy/2
  double t0, dt;            // ray r at t = t0 + r*dt
};

struct SSDdaAngle {
  double c, s;              // glibc cos / sin of theta (host)
  double xe, ye;            // edges a ray may enter through (x = xe, y = ye)
  double e_step;            // slow-axis drift per fast move
  double e_cap;             // pixel width across the slow axis
  double len_full;          // intersection length of a fast move without slow move
  double k_len;             // len_full / e_step (0 for axis-parallel angles)
  int di, dj;               // index steps along the ray
  int i0, j0;               // entry line / column on those edges
  int fast_i;               // 1: the line index is the fast axis
  int dfast, dslow;
  int axis;                 // 1: sin == 0, 2: cos == 0, 0: oblique
  int fixed;                // exact-geometry angle: the Radon takes the trapezoid form
  // Trapezoid form of the Radon on these angles: along fast index k the ray crosses
  // the slow axis at sigma = t tr_inv + (tr_off - DDA_PADW - 1) + half_s + k tr_d (slow-pixel units),
  // and its length in the pixel at slow index j is lmax z(|sigma - j|) with
  // z(u) = clamp((p - u |step|) / (p - q)) (p, q, lmax of SSDdaGatherAngle; |step| the
  // slow pitch times the trig factor).  With tau = sigma - half_s + DDA_PADW + 1 the
  // candidates are the padded slow indices floor(tau), floor(tau) + 1, at
  // u0 = 1 - half_s - fr and u1 = 2 - half_s - fr > 0 for fr the fraction of tau.
  double tr_inv, tr_off, tr_d, tr_lmax;
  int tr_dhi;               // tr_d in Q32: tr_dhi + tr_dlo 2^-32
  unsigned tr_dlo;
  float tr_c1, tr_kz, tr_pz, tr_pz1;  // (1 - half_s) 2^32, kz 2^-32, pz, pz - kz (2 - half_s)
  // Wave-staged form: the 64 rays of a wave span tau differences up to 63 |dt tr_inv|,
  // and over the DDA_WS_ROWS fast indices of a chunk tau moves by at most
  // (DDA_WS_ROWS - 1) |tr_d|, so one DDA_WS_SEG-pixel segment per fast index, starting at
  // the lowest ray's floor(tau) at the chunk's first index, less 1 and tr_wgap, aligned
  // down to 4, holds every lane's candidates on every row of the chunk when
  // spread + (DDA_WS_ROWS - 1) |tr_d| + 9 fits.
  int tr_wstage;            // 1: the wave-staged form applies
  int tr_lowlane;           // the wave's lane with the lowest tau (0, or 63 when tau falls with the ray)
  int tr_wgap;              // floor((DDA_WS_ROWS - 1) |tr_d|) + 1 when tr_d < 0, else 0
  long long tr_qchunk;      // DDA_WS_ROWS tr_d in Q32 (2^32 units): tau's step over a chunk
};

// Pixel-driven form of the transpose for one angle.  Where the traversal is exact
// geometry (its slow-axis drift per move e_step <= e_cap, and the angle is not within
// 1e-7 of an axis, where the reference's tie rule at pixel edges decides), the length
// of ray t in pixel (i, j) is the trapezoid L(|t - t_c|) of the line's distance from the
// pixel centre, t_c = x_j cos + y_i sin:
//   L = Lmax                          for |d| <= q
//   L = Lmax (p - |d|) / (p - q)      for q < |d| < p
//   L = 0                             for |d| >= p
// with p = (w|cos| + h|sin|)/2, q = |w|cos| - h|sin||/2, Lmax = min(w/|sin|, h/|cos|)
// for w x h pixels.  Other angles are traced ray by ray.
struct SSDdaGatherAngle {
  double c, s;
  double hys;               // hy * sin: t_c step per image line
  double p, inv_pq, lmax;
  double half;              // p / |dt|: candidate rays either side of t_c
  // Two-candidate form (k <= 2) in ray units: rc - half = k0h + j kx + i ky for pixel
  // (i, j), where rc = (t_c - t0) / dt; with fr the fraction of rc - half, the
  // candidates r0 = floor(rc - half) + 1 and r0 + 1 lie at u0 = c1 - fr and
  // u1 = 2 - half - fr > 0 from rc, so z0 = clamp(pz - kz |u0|), z1 = clamp(pz1 + kz fr).
  double k0h, kx, ky;
  float c1, kz, pz, pz1;    // 1 - half, |dt| / (p - q), p / (p - q), pz - kz (2 - half)
  float kyf;                // ky
  int a;                    // sinogram column
  int k;                    // candidate rays per pixel: ceil(2 half) + 1
};

// ray state in (fast, slow) coordinates
struct SSDdaRay {
  int f, s;                 // fast / slow index
  double len;               // intersection length of the current pixel
  double e;                 // drift since the last slow-axis boundary
  int slow_next;            // the next move is along the slow axis
};

// double -> int as the reference's assignments compile on x86-64: NaN and values
// outside the int range give INT_MIN, an index outside the image
static inline DDA_HD int dda_to_index(double v) {
  if (!(v >= -2147483648.0 && v < 2147483648.0)) return (int)0x80000000u;
  return (int)v;
}

// index step with two's-complement wrap-around: an entry index of INT_MIN (a ray that
// misses the image) stays far outside the image instead of overflowing
static inline DDA_HD int dda_step(int v, int d) { return (int)((unsigned)v + (unsigned)d); }

// pixel index of coordinate v on an axis from v0 with pitch h, moving in direction
// dir (raft/raft_ddaiterator.c:10-14)
static inline DDA_HD int dda_axis_index(double v, double v0, double h, int dir) {
  DDA_NOCONTRACT
  if (dir >= 0) return dda_to_index(floor(((v - v0) / h) - 0.5) + 1);
  return dda_to_index(ceil(((v - v0) / h) - 0.5));
}

// one fast move (raft/raft_ddaiterator.c:212-247).  The intersection length of a move
// that crosses a slow-axis boundary is len_full*(e_cap - e_prev)/e_step, the
// reference's alpha/(1+alpha) with alpha = (e_cap - e_prev)/e written without its
// two divisions (same value up to rounding of the length; the move sequence, which
// depends on e only, is identical).
static inline DDA_HD void dda_fast_move(const SSDdaAngle& a, SSDdaRay& r) {
  DDA_NOCONTRACT
  r.f = dda_step(r.f, a.dfast);
  const double e_prev = r.e;
  r.e += a.e_step;
  r.slow_next = r.e > a.e_cap;
  if (r.slow_next) {
    r.e -= a.e_cap;
    r.len = r.e != 0.0 ? (a.e_cap - e_prev) * a.k_len : a.len_full;
  } else {
    r.len = a.len_full;
  }
}

// entry pixel and first intersection length of the ray at t
// (raft/raft_ddaiterator.c:108-194).  Axis-parallel rays start with len_full: the
// reference leaves that first length as the previous ray left it (uninitialised for a
// thread's first ray) and carries len_full into every later ray of the angle.
static inline DDA_HD SSDdaRay dda_ray_start(const SSDdaGeom& g, const SSDdaAngle& a, double t) {
  DDA_NOCONTRACT
  SSDdaRay r;
  int i, j;
  r.slow_next = 0;
  r.len = a.len_full;
  r.e = 0.0;
  if (a.axis == 1) {
    j = dda_axis_index(t, g.ox, g.hx, a.dj);
    i = a.i0;
  } else if (a.axis == 2) {
    i = dda_axis_index(t, g.oy, g.hy, a.di);
    j = a.j0;
  } else {
    const double sx = (t * a.c - a.xe) / a.s;
    const double sy = (a.ye - t * a.s) / a.c;
    if (sx > sy) {
      // enters through the vertical edge x = xe
      const double xy = (sx * a.c) + (t * a.s);
      i = dda_axis_index(xy, g.oy, g.hy, a.di);
      j = a.j0;
      r.e = fabs(xy - (g.oy + (i - (0.5 * a.di)) * g.hy));
      if (!a.fast_i) {
        // fast = column: step back one column and take a fast move into the image
        r.f = dda_step(j, -a.dj);
        r.s = i;
        dda_fast_move(a, r);
        return r;
      }
      r.e -= fabs(g.hy);
      r.e /= a.c;
      r.len = fabs(r.e);
      r.e = fabs(r.len * a.s);
    } else {
      // enters through the horizontal edge y = ye
      const double xy = (-sy * a.s) + (t * a.c);
      j = dda_axis_index(xy, g.ox, g.hx, a.dj);
      i = a.i0;
      r.e = fabs(xy - (g.ox + (j - (0.5 * a.dj)) * g.hx));
      if (a.fast_i) {
        r.f = dda_step(i, -a.di);
        r.s = j;
        dda_fast_move(a, r);
        return r;
      }
      r.e -= fabs(g.hx);
      r.e /= a.s;
      r.len = fabs(r.e);
      r.e = fabs(r.len * a.c);
    }
  }
  r.f = a.fast_i ? i : j;
  r.s = a.fast_i ? j : i;
  return r;
}

// ---- launchers (dda_kernels.hip)

// Radon: sino[slice][R][A] (overwritten) of img[slice][lines][cols]; pad
// (dda_radon_scratch_floats) is scratch that the launcher fills with the image's padded
// (line, column) and (column, line) layouts: DDA_PADW zeros either side of every line,
// the line length rounded up to 4 floats (dda_pad_pitch).  The padding covers a whole
// staged segment (DDA_WS_SEG) of a wave whose lowest ray lies left of the image while
// another of its rays reaches in.
#define DDA_PADW 132
#define DDA_WS_ROWS 8   // fast indices per wave-staged chunk (even)
#define DDA_WS_SEG 128  // padded pixels staged per fast index and wave
static inline DDA_HD int dda_pad_pitch(int ns) { return (ns + 2 * DDA_PADW + 3) & ~3; }
static inline size_t dda_radon_scratch_floats(const SSDdaGeom& g, int nslices) {
  return (size_t)nslices * ((size_t)g.lines * dda_pad_pitch(g.cols) + (size_t)g.cols * dda_pad_pitch(g.lines));
}
hipError_t dda_launch_radon(const SSDdaGeom& g, const SSDdaAngle* angles, const float* img, float* pad, float* sino,
                            int nslices, hipStream_t stream);
// The transpose's angles split into the pixel-driven set (gather table, ngather) and the
// ray-traced set (walk list of sinogram columns, nwalk).
// Two-candidate gather (k_dda_gather2): block of DDA_G2_COLS columns x 4 G lines (G
// lines per thread: DDA_G2_LINES, or DDA_G2_LINES_SMALL where the larger tiles would
// leave the CUs short of blocks), angles staged DDA_G2_ANGLES at a time, each in a
// segment of at most DDA_G2_WSMAX rays
#define DDA_G2_COLS 64
#define DDA_G2_LINES 16
#define DDA_G2_LINES_SMALL 8
#define DDA_G2_ANGLES 16
#define DDA_G2_WSMAX 160

// segment length of the two-candidate gather for an angle (either tile height): the
// block's span of rc - half, (cols - 1)|kx| + (lines - 1)|ky|, plus 8
static inline int dda_gather2_ws(const SSDdaGatherAngle& ga) {
  const double span = (DDA_G2_COLS - 1) * fabs(ga.kx) + (4 * DDA_G2_LINES - 1) * fabs(ga.ky);
  return span < DDA_G2_WSMAX ? (int)ceil(span) + 8 : DDA_G2_WSMAX + 1;
}

// An angle of the transpose whose every ray runs straight along the fast axis: one slow
// index, all nf fast indices, each with length len (the axis angles).  Its R^T is a
// vector over the slow index, q summed over the rays of each slow index s (a contiguous
// range, ranges[off + 2 s] .. ranges[off + 2 s + 1] - 1) times len, added to every pixel
// of that column (slow_col: slow index = column) or line.
struct SSDdaStraight {
  int a;
  int slow_col;
  int off;
  double len;
};

struct SSDdaTransposePlan {
  const SSDdaAngle* angles;
  const SSDdaGatherAngle* gather;
  int ngather;
  int kmax;                 // largest candidate count of the gather table
  int gather_ws;            // segment length of k_dda_gather2 (0: not usable)
  const SSDdaStraight* straight;  // straight angles (only when no angle is traced)
  int nstraight;
  const int* ranges;
  const float* row_scale;   // [A]: lmax of the gather angles (0 for traced ones)
  const int* walk;
  int nwalk;
};

// out[slice] = R^T sino[slice] [/ w] (T: accumulator type of the ray-traced part, float or
// double).  sinoT ([nslices][A][R + 4] floats), acc, accT ([nslices][lines*cols] T) are
// scratch.  `ones`: R^T 1 for one slice (sino unused).
template <typename T, typename Out>
hipError_t dda_launch_transpose(const SSDdaGeom& g, const SSDdaTransposePlan& tp, const float* sino, bool ones,
                                float* sinoT, T* acc, T* accT, const double* w, Out* out, int nslices,
                                hipStream_t stream);
hipError_t dda_launch_to_f32(const double* src, float* dst, size_t n, hipStream_t stream);

#endif
