// ss_geometry.h -- slant-stack geometry shared by host code and HIP kernels.
//
// Grid conventions are the reference slant-stack's (raft/raft_radon_slantstack.c:143-172,
// raft/raft_backprojection_slantstack.c:138-157):
//   image  N x N, row r <-> y = ymin + r*dy (row 0 is the BOTTOM, ymin = br_y),
//          column c <-> x = xmin + c*dx, dx = (br_x - tl_x)/(N-1), dy = (tl_y - br_y)/(N-1)
//   sinogram R x A, row k <-> t = tmin + k*dt (tmin = br_y, dt = (tl_y - br_y)/(R-1)),
//          column j <-> theta = th0 + j*dth, dth = (br_x - tl_x)/A (endpoint excluded).
//
// Radon sample positions are kept in Q32.32 fixed point: for ray (j,k) and sample s
// the interpolation coordinate is u = alpha0_j + k*gamma_j + s*beta_j (exact integer
// arithmetic).  The forward kernel and the exact-transpose kernel evaluate the same
// integer expression, so the transpose sees bit-identical indices and weights.
#ifndef PYRAFT_AMD_SS_GEOMETRY_H
#define PYRAFT_AMD_SS_GEOMETRY_H

#include <stdint.h>

struct SSGeom {
  int n;        // image side
  int nrays;    // R
  int nangles;  // A
  int pitch;    // row length of the Radon gather layouts: n + 2 rounded up to a multiple of 16
  double xmin, ymin, dx, dy;
  double tmin, dt;
  double th0, dth;
};

// Per-angle constants.  cost/sint/tant are glibc's FP64 values for theta_j computed
// on the host (the same calls the reference makes), the rest are derived from them.
struct SSAngle {
  double cost, sint, tant;
  // Radon trajectory in Q32.32 (see header comment).  branch_a != 0 when
  // |cos| < 1/sqrt(2): samples walk image COLUMNS and interpolate between ROWS
  // (raft/raft_radon_slantstack.c:174-192); otherwise samples walk ROWS and
  // interpolate between COLUMNS (:193-211).
  int64_t alpha0, gamma, beta;
  float delta;   // dy/|sin| or dx/|cos|
  int branch_a;
  // Backprojection ray coordinate u = (x cos + y sin - tmin)/dt = bp_a + c*bp_b + r*bp_c.
  double bp_a, bp_b, bp_c;
  // Transpose: the ray index whose Q32.32 trajectory reaches cell position x (in cells)
  // at sample s is k = x*rt_ig - rt_a0 - s*rt_b  (rt_ig = 2^32/gamma, rt_a0 = alpha0/gamma,
  // rt_b = beta/gamma), used to bound the candidate rays of a pixel tile.
  double rt_ig, rt_a0, rt_b;
  // Candidate rays of image line s, with a0(s) = rt_a0 + s*rt_b: the rays reaching cells
  // [-1, n] lie in [rt_w - a0(s), rt_w - a0(s) + rt_wn], those reaching cells [1, n-2] in
  // [rt_i0 - a0(s), rt_i1 - a0(s)] (real-valued; zero when gamma = 0)
  double rt_w, rt_i0, rt_i1;
  int rt_wn;
  int rt_pad_;
  // |beta| below this threshold: sample validity is evaluated sample by sample
  // (knife-edge angles where the FP64 predicate need not be monotone).
  int degenerate;
  // 1: not degenerate and every ray has a contiguous exact range (no irregular rays);
  // set when a plan's ranges are built
  int regular;
};

// Q32.32 scale
#define SS_FIX_ONE 4294967296.0

#endif
