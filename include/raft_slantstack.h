/*
 * raft_slantstack.h -- exported C ABI of pyraft_amd/lib/libraft<EXT_SUFFIX>.
 *
 * Two groups of symbols:
 *
 *  1. Drop-in symbols with the reference's names and signatures (void, by-value
 *     structs / host float pointers).  They run on the current HIP device and
 *     print the failure reason to stderr (the reference is silent, see
 *     raft/raft_radon_gpu.cu:16-35).
 *
 *  2. New stream-ordered device-pointer entry points (prefix raft_ss_) that return
 *     0 or a negative RAFT_SS_E* code; raft_ss_last_error() gives the message.
 *     They have no reference counterpart; they are what the drop-in symbols and the
 *     Python package are built on.
 *
 * Layouts (row-major, FP32 on the device):
 *   image    [nslices][N][N], row r <-> y = ymin + r*dy (row 0 at the BOTTOM, as in
 *            raft/raft_radon_slantstack.c:146,184), column c <-> x = xmin + c*dx
 *   sinogram [nslices][R][A] (angle contiguous, as pyraft's (rays, angles) arrays);
 *            the *_range variants accept/produce the angle-major [nslices][A][R] layout
 *            when angle_major != 0.
 * Corners are {tl_x, tl_y, br_x, br_y} exactly as in raft_image (raft/raft_image.h:11-19);
 * the sinogram's x axis is theta (theta_j = tl_x + j*(br_x-tl_x)/A) and its y axis is t
 * (t_k = br_y + k*(tl_y-br_y)/(R-1)).
 *
 * Streams are hipStream_t passed as void* (NULL = the default stream).  A plan may be
 * shared by concurrent calls on one device; only raft_ss_plan_set_bp_variant and
 * raft_ss_plan_tune change it.
 */
#ifndef RAFT_SLANTSTACK_H
#define RAFT_SLANTSTACK_H

#include "raft_types.h"

#ifdef __cplusplus
extern "C" {
#endif

#if defined(__GNUC__)
#pragma GCC visibility push(default)
#endif

/* ---------------------------------------------------------------- status codes */
#define RAFT_SS_OK 0
#define RAFT_SS_EINVAL (-1)    /* bad argument / unsupported geometry                */
#define RAFT_SS_EHIP (-2)      /* a HIP runtime call failed                           */
#define RAFT_SS_EINTERNAL (-3) /* internal consistency check failed                   */

const char *raft_ss_last_error(void);
int raft_ss_version(void);

/* ================================================================ drop-in symbols */

/* Replaces raft_radon_slantstack (raft/raft_radon_slantstack.h:10-12), bound at
 * pyraft/raftypes.py:159-160 and called from pyraft.radon(..., method=SLANTSTACK)
 * (pyraft/sinogram/sino.py:201-202).  FP64 strided input/output, FP32 on the GPU;
 * nthreads is accepted and ignored. */
void raft_radon_slantstack(raft_image phantom, raft_image radon, int nthreads);

/* Replaces raft_backprojection_slantstack (raft/raft_backprojection_slantstack.h:10-13),
 * bound at pyraft/raftypes.py:161-162, called from pyraft.backprojection(sino, plan)
 * (pyraft/backprojection/backprojection.py:206-207). */
void raft_backprojection_slantstack(raft_image sinogram, raft_image backprojection, int nthreads);

/* Exact transpose of raft_radon_slantstack in the same by-value style.  No reference
 * counterpart: pyraft.radon_transpose supports only BRESENHAM
 * (pyraft/backprojection/backprojection.py:229-232); this is its SLANTSTACK case. */
void raft_radon_transpose_slantstack(raft_image sinogram, raft_image image, int nthreads);

/* Replaces raft_radon_slantstack_withmesh (raft/raft_radon_slantstack_withmesh.c:49-97;
 * the header raft/raft_radon_slantstack_withmesh.h:10-14 names the first two arguments
 * the other way round, the definition and pyraft/sinogram/sino.py:274 pass the phantom
 * first), called by pyraft.radon_withmesh and radon_3d_stack (pyraft/sinogram/sino.py:251-311):
 * slant-stack ray sums for angle theta[j] and detector coordinate t[k] read from the
 * vectors, element (k, j) of radon.  Every output is computed (the reference's thread
 * split leaves the last nviews*nrays mod nthreads outputs unwritten,
 * raft/raft_radon_slantstack_withmesh.c:65-77). */
void raft_radon_slantstack_withmesh(raft_image phantom, raft_image radon, raft_vector t, raft_vector theta,
                                    int nthreads);

/* Replaces raft_radon_slantstack_view (raft/raft_radon_slantstack_view.c:22-38), called
 * by pyraft.radon_view and snapshot (pyraft/sinogram/sino.py:220-343): one angle, view[k]
 * = ray sum at t[k]. */
void raft_radon_slantstack_view(raft_vector view, raft_image phantom, raft_vector t, double theta);

/* Status (0 or RAFT_SS_E*) of the last drop-in call above on this thread: the drop-in
 * symbols return void like the reference's, so callers that want to raise use this. */
int raft_ss_legacy_status(void);

/* Replaces raft_radon_slantstack_gpu (raft/raft_radon_gpu.h:8; pyraft/raftypes.py:208),
 * called by pyraft.radon_gpu (pyraft/sinogram/sino.py:377).  Host float buffers,
 * default corners (image [-1,1]^2, theta in [0,pi), t in [-1,1]).  Computes the
 * CPU slant-stack semantics, not the reference CUDA kernel's theta = pi - j*pi/A
 * texture variant. */
void raft_radon_slantstack_gpu(float *h_output, float *h_input, int sizeImage, int nrays, int nangles);

/* Replaces raft_backprojection_slantstack_gpu (raft/raft_backprojection_gpu.h:8;
 * pyraft/raftypes.py:206), called by backprojection_gpu and fbp_gpu
 * (pyraft/backprojection/backprojection.py:254, pyraft/inversion/inversion.py:56).
 * Linear interpolation, no column mirroring (see raft_radon_slantstack_gpu). */
void raft_backprojection_slantstack_gpu(float *image, float *sino, int sizeImage, int nrays, int nangles);

/* Replaces raft_em_gpu (raft/raft_em_gpu.cu:20; pyraft/raftypes.py:211), called by
 * pyraft.em_gpu (pyraft/iterative/em.py:202).  x0 = 1, niter plain EM iterations
 * with the slant-stack pair; zero projections give a zero ratio instead of Inf/NaN. */
void raft_em_gpu(float *output, float *sino, int sizeImage, int nrays, int nangles, int niter);

/* Replaces raft_tr_em_gpu (raft/raft_transmission_em.cu:16-66; pyraft/raftypes.py:212-213),
 * called by pyraft.em_transmission_gpu (pyraft/iterative/em.py:212-248).  Transmission EM
 * from x0 = 1 with the slant-stack pair and default corners: W = BP(rawsino), then per
 * iteration x <- x * (BP(flat * exp(-R x)) / W)^2 -- the reference applies its
 * multiplicative update twice (raft/raft_transmission_em.cu:57-59), kept here; pixels with
 * W = 0 keep their value instead of the reference's unguarded division. */
void raft_tr_em_gpu(float *output, float *rawsino, float *flat, int sizeImage, int nrays, int nangles, int niter);

/* Replaces raft_image_phantom_fromdesc (raft/raft_image_phantom_fromdesc.h:26;
 * pyraft/raftypes.py:138) used by pyraft.shepplogan (pyraft/phantom/phantom.py:49).
 * Adds into the image.  A description with < 6 columns is an error (the reference
 * frees the caller's buffer there, raft/raft_image_phantom_fromdesc.c:11-15). */
void raft_image_phantom_fromdesc(raft_image image, raft_matrix desc);

/* Replaces raft_radon_fromdesc (raft/raft_radon_fromdesc.h:23-24; pyraft/raftypes.py:140),
 * the exact Radon transform of an ellipse description used by pyraft.radon(ndarray)
 * (pyraft/sinogram/sino.py:209-212).  Adds into the sinogram. */
void raft_radon_fromdesc(raft_image sino, raft_matrix desc);

/* ================================================================ plans */

typedef struct raft_ss_plan raft_ss_plan;

/* Build the geometry plan for an N x N image and an R x A sinogram: the FP64 angle
 * table (glibc cos/sin/tan, as the reference evaluates them), the exact per-ray
 * sample ranges of the slant-stack predicate and the backprojection staging
 * parameters.  Synchronises `stream`. */
int raft_ss_plan_create(raft_ss_plan **plan, int n, int nrays, int nangles, const double *img_corners,
                        const double *sino_corners, void *stream);
int raft_ss_plan_destroy(raft_ss_plan *plan);
/* info[0..8] = n, nrays, nangles, #irregular rays, #degenerate angles, bp window,
 * bp angles/stage, bp kernel variant (window and stage are those of the variant), and the
 * angle groups a one-slice full-range backprojection divides into (1: not divided; a
 * small image's launch is divided so that it fills the GPU, and the groups' partial
 * images are added in group order) */
int raft_ss_plan_info(const raft_ss_plan *plan, int *info, int ninfo);
/* copy the per-ray sample ranges [A][R] x (lo, hi) to host memory */
int raft_ss_plan_bounds(const raft_ss_plan *plan, int *host_bounds);
/* copy the per-(angle, 64-ray group) union/intersection of the ranges, [A][ceil(R/64)] x
 * (union lo, union hi, intersection lo, intersection hi); (0,-1,0,-1) for an empty group */
int raft_ss_plan_wave_bounds(const raft_ss_plan *plan, int *host_wbounds);
/* the same ranges (and per-angle Q32.32 alpha0/gamma/beta, branch flags) computed on
 * the host, for checking index ranges without a GPU */
int raft_ss_plan_host_bounds(int n, int nrays, int nangles, const double *img_corners, const double *sino_corners,
                             int *host_bounds, long long *host_fix, int *host_branch);

/* Plan on arbitrary tables: ray k at detector coordinate t[k], angle j at theta[j]
 * (the geometry of raft_radon_slantstack_withmesh).  Only the Radon (raft_ss_radon*)
 * accepts such a plan; the other operators need the uniform grid.  nrays >= 2. */
int raft_ss_plan_create_mesh(raft_ss_plan **plan, int n, const double *img_corners, int nrays, const double *t,
                             int nangles, const double *theta, void *stream);
/* the mesh plan's ranges ([A][R] x (lo, hi)) and Q32.32 trajectory offsets ([A][R])
 * computed on the host */
int raft_ss_plan_host_bounds_mesh(int n, const double *img_corners, int nrays, const double *t, int nangles,
                                  const double *theta, int *host_bounds, long long *host_alpha);

/* plan summary computed on the host (no device needed): info[0..4] = #degenerate
 * angles, bp window, bp angles per stage, 64-ray groups, 0 */
int raft_ss_plan_host_info(int n, int nrays, int nangles, const double *img_corners, const double *sino_corners,
                           int *info);

/* backprojection staging of every kernel variant computed on the host:
 * windows[v] = rays staged per angle, per_stage[v] = angles per LDS stage (0 when the
 * window does not fit variant v); arrays of RAFT_SS_BP_NVARIANTS ints */
int raft_ss_plan_host_bp_staging(int n, int nrays, int nangles, const double *img_corners,
                                 const double *sino_corners, int *windows, int *per_stage);
/* pixel-driven exact transpose staging computed on the host: windows[v], per_stage[v] as
 * above (0 angles when some angle's rays are closer than one cell: the line form runs),
 * and per angle (table, nangles x 5 doubles) its transpose constants rt_ig, rt_a0, rt_b,
 * branch (1: samples walk image columns) and ray spacing |gamma| in cells */
int raft_ss_plan_host_rt_staging(int n, int nrays, int nangles, const double *img_corners,
                                 const double *sino_corners, int *windows, int *per_stage, double *table);

/* ================================================================ device-pointer operators */

/* Slant-stack Radon of nslices images (angle columns [j0, j1) of the sinogram are
 * written; the rest is untouched). */
int raft_ss_radon(const raft_ss_plan *plan, const float *d_image, float *d_sino, int nslices, void *stream);
int raft_ss_radon_range(const raft_ss_plan *plan, const float *d_image, float *d_out, int angle_major, int j0, int j1,
                        int nslices, void *stream);

/* Slant-stack backprojection (sum over angles [j0, j1) times dtheta). */
int raft_ss_backprojection(const raft_ss_plan *plan, const float *d_sino, float *d_image, int nslices, void *stream);
int raft_ss_backprojection_range(const raft_ss_plan *plan, const float *d_sino, float *d_image, int angle_major,
                                 int j0, int j1, int nslices, void *stream);

/* Exact matched transpose of raft_ss_radon (no reference counterpart): the adjoint
 * that the autograd backward needs, <R x, y> == <x, R^T y> to rounding. */
int raft_ss_radon_transpose(const raft_ss_plan *plan, const float *d_sino, float *d_image, int nslices, void *stream);
int raft_ss_radon_transpose_range(const raft_ss_plan *plan, const float *d_sino, float *d_image, int angle_major,
                                  int j0, int j1, int nslices, void *stream);

/* FBP row filter of pyraft.lowpassino_fbp (pyraft/filters/filters.py:36-68): each angle
 * column is circularly convolved along t with g = Re(ifft(fftshift(|w| cos(pi w dt)))). */
int raft_ss_fbp_filter(raft_ss_plan *plan, const float *d_sino, float *d_out, int angle_major, int nslices,
                       void *stream);
/* The FP32 convolution kernel g (length R) of lowpassino_fbp for a detector grid
 * t = linspace(tmin, tmax, R), computed in FP64 on the host. */
int raft_ss_ramp_kernel(int R, double tmin, double tmax, float *host_g);
/* Kernel of either numpy FBP filter: RAFT_SS_FILTER_COS_RAMP, h = |w| cos(pi w dt)
 * (pyraft.lowpassino_fbp); RAFT_SS_FILTER_TIKHONOV_RAMP, h = |w| / (1 + param |w|^2)
 * (pyraft.lowpassino_reg_fbp, pyraft/filters/filters.py:73-111, param = L). */
#define RAFT_SS_FILTER_COS_RAMP 0
#define RAFT_SS_FILTER_TIKHONOV_RAMP 1
int raft_ss_filter_kernel(int R, double tmin, double tmax, int kind, double param, float *host_g);
/* Circular convolution of nrows contiguous length-R rows with the device kernel d_g;
 * an FFT convolution up to R = 2048, a direct sum up to R = 6826 (larger R: error).
 * Not in place. */
int raft_ss_filter_rows(const float *d_in, float *d_out, const float *d_g, int R, int nrows, void *stream);
/* filter + backprojection (pyraft/inversion/inversion.py:48-56) */
int raft_ss_fbp(raft_ss_plan *plan, const float *d_sino, float *d_image, int nslices, void *stream);

/* Transmission EM of raft_tr_em_gpu on device buffers: d_raw and d_flat [s][R][A], d_x the
 * starting image on entry and the estimate on exit. */
int raft_ss_tr_em(raft_ss_plan *plan, const float *d_raw, const float *d_flat, float *d_x, int niter, int nslices,
                  void *stream);

/* Ordered-subsets EM (nsubsets = 1: the EM of raft/raft_em_gpu.cu:41-69).  d_x is the
 * starting image on entry and the estimate on exit.  Per subset: forward projection
 * with the guarded ratio s/p and FP64 log-likelihood partials fused in its epilogue,
 * then backprojection with x <- x * BP_b(ratio) / BP_b(1) fused in its epilogue.
 * ll_host (optional) receives niter values of sum(s log p - p) over rays with p > 0.
 * The subsets' backprojection variant is the plan's when set by
 * raft_ss_plan_set_bp_variant; otherwise the first call per subset size times the
 * variants on one subset's angle range (not under stream capture) and keeps the fastest.
 * Variants differ by ~1e-7 relative in FP32 rounding, so a timed choice can change the
 * low bits between processes; setting the variant makes them reproducible.
 * Stream capture: with ll_host NULL the call enqueues kernels and its stream-ordered
 * workspace allocation and release only (no memset, copy or wait), so it can be
 * captured into a graph; with ll_host set it copies the log-likelihood back and waits
 * for the stream, and returns RAFT_SS_EINVAL while the stream is being captured. */
int raft_ss_em(raft_ss_plan *plan, const float *d_sino, float *d_x, int niter, int nsubsets, int nslices,
               double *ll_host, void *stream);

/* Rasterise ellipses (rows of 6 doubles: value, a, b, x0, y0, rotation) into a
 * device FP32 image of lines x cols with corners {tl_x, tl_y, br_x, br_y}
 * (row 0 at tl_y, as raft/raft_image_phantom_fromdesc.c:18-41). */
int raft_ss_phantom(const double *desc_rows, int nell, int lines, int cols, const double *corners, float *d_out,
                    void *stream);

/* ================================================================ kernel variants and the autotuner
 *
 * No reference counterpart (the reference has one fixed CUDA launch shape,
 * raft/raft_backprojection_gpu_function.cu:76-113).  The backprojection has
 * RAFT_SS_BP_NVARIANTS tile shapes: variant v covers 32 columns x 8*rows[v] pixel rows
 * per 256-thread block and stages up to stage[v] (value, difference) pairs of
 * sinogram windows in LDS per block.  Every variant sums the same terms in the same
 * per-pixel angle order; they differ only in the FP32 rounding of the detector
 * coordinate, which is taken relative to the tile origin (~1e-7 relative). */
#define RAFT_SS_BP_NVARIANTS 6
int raft_ss_bp_variant_shape(int variant, int *rows, int *stage);
/* select a variant for a plan (RAFT_SS_EINVAL if the window does not fit its stage); it is
 * then also the variant of raft_ss_em's subset backprojections */
int raft_ss_plan_set_bp_variant(raft_ss_plan *plan, int variant);
/* time every runnable variant on this plan's full angle range (warm-up + best of
 * `reps` launches, HIP events on `stream`), keep the fastest and return its index.
 * times_us (optional, RAFT_SS_BP_NVARIANTS doubles) receives the times, -1 for
 * variants that do not fit.  Results are cached by device architecture and geometry,
 * per process and in a JSON file that later processes read ($PYRAFT_AMD_TUNE_CACHE, an
 * empty value for none; else $XDG_CACHE_HOME/pyraft_amd/tune.json or
 * $HOME/.cache/pyraft_amd/tune.json); raft_ss_em's subset variants are cached the same
 * way.  Flag RAFT_SS_TUNE_FORCE measures again.  Not to be called while another thread
 * uses the plan. */
#define RAFT_SS_TUNE_FORCE 1
int raft_ss_plan_tune(raft_ss_plan *plan, int reps, int flags, double *times_us, void *stream);
/* Each Radon launch runs in one of two block shapes (staged window of 4096 records in
 * chunks of up to 32 rows, or 4864 records in chunks of up to 40), and launches of few
 * blocks (few angles, small images) divide each ray's samples into slabs of rows added by
 * a finishing pass.  By default the first launch per (angle count, slice count) times
 * both shapes undivided and, where the default rule divides the launch, both shapes at
 * the default slab size and shape 0 at twice and half of it, keeps the fastest and caches
 * it with the backprojection choices.  rows > 0 pins the slab size (and shape 0) for
 * every Radon of the plan (rows >= n: no division); rows = 0 returns to the tuned
 * choice.  Shapes and slab sizes change the FP32 summation order (~1e-7 relative);
 * pinning makes the bits reproducible. */
int raft_ss_plan_set_radon_slab_rows(raft_ss_plan *plan, int rows);
/* The exact transpose runs pixel tiles of the backprojection's shapes (variant v: 32
 * columns x 8*rows[v] rows, stage[v] staged pairs) when every angle's rays are at least
 * one cell apart, else an image-line form.  This times the runnable tile variants on
 * `nslices` slices over the full angle range, keeps the fastest for the plan and returns
 * it (-1 and no timing when the line form applies); times_us as raft_ss_plan_tune.
 * Cached with the other choices. */
int raft_ss_plan_tune_transpose(raft_ss_plan *plan, int nslices, int reps, int flags, double *times_us,
                                void *stream);

/* ================================================================ batches of uneven slices
 *
 * No reference counterpart: the reference runs one slice per call
 * (raft/raft_radon_gpu.cu:16-35, raft/raft_em_gpu.cu:20-88, and pyraft.osem loops
 * over subsets in Python, pyraft/iterative/em.py:265-320).  A batch holds nslices
 * slices with their own n_i, R_i, A_i and corners.  Images and sinograms are packed
 * back to back: slice i's image starts at img_off[i] floats (n_i x n_i) and its
 * sinogram at sino_off[i] floats (R_i x A_i, or A_i x R_i angle-major).  Each
 * operator runs every slice in one launch per kernel: the Radon and backprojection
 * kernels as persistent launches (a fixed grid of resident blocks walking the work
 * items of all slices, heaviest first), the layout kernels (gather layouts, sinogram
 * transposes) with one block per tile of every slice; and with
 * RAFT_SS_BATCH_GRAPH the whole call is captured once into a HIP graph and
 * replayed by a single graph launch on later calls with the same arguments. */
typedef struct raft_ss_batch raft_ss_batch;
#define RAFT_SS_BATCH_GRAPH 1
/* img_corners / sino_corners: nslices x 4 doubles, or NULL for the defaults (image
 * [-1,1]^2; theta in [0, pi), t in [-1, 1]).  Synchronises `stream`. */
int raft_ss_batch_create(raft_ss_batch **batch, int nslices, const int *n, const int *nrays, const int *nangles,
                         const double *img_corners, const double *sino_corners, void *stream);
int raft_ss_batch_destroy(raft_ss_batch *batch);
/* img_off / sino_off (nslices each, optional), totals[2] = total image / sinogram floats */
int raft_ss_batch_layout(const raft_ss_batch *batch, long long *img_off, long long *sino_off, long long *totals);
/* info[0..5] of one slice: n, nrays, nangles, #irregular rays, bp variant, bp angles/stage */
int raft_ss_batch_info(const raft_ss_batch *batch, int slice, int *info, int ninfo);
/* bp variant of one slice (slice = -1: all slices) */
int raft_ss_batch_set_bp_variant(raft_ss_batch *batch, int slice, int variant);
/* autotune: per distinct image side n, time the variants on that class's median-A
 * slice with the batched kernel and give the class the fastest */
int raft_ss_batch_tune(raft_ss_batch *batch, int reps, int flags, void *stream);
/* Radon of all slices: packed images -> packed sinograms ([R][A], or [A][R] if
 * angle_major) */
int raft_ss_batch_radon(raft_ss_batch *batch, const float *d_images, float *d_sinos, int angle_major, int flags,
                        void *stream);
/* backprojection of all slices over their full angle ranges */
int raft_ss_batch_backprojection(raft_ss_batch *batch, const float *d_sinos, float *d_images, int angle_major,
                                 int flags, void *stream);
/* OSEM of all slices (raft_ss_em per slice, subset b of slice i = angles
 * [b*A_i/S, (b+1)*A_i/S)); d_sinos in [R][A] layout, d_x the packed starting images
 * (estimates on exit).  ll_host (optional): niter x nslices log-likelihoods. */
int raft_ss_batch_em(raft_ss_batch *batch, const float *d_sinos, float *d_x, int niter, int nsubsets,
                     double *ll_host, int flags, void *stream);

#if defined(__GNUC__)
#pragma GCC visibility pop
#endif

#ifdef __cplusplus
}
#endif

#endif /* RAFT_SLANTSTACK_H */
