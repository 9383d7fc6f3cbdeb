// ss_kernels.h -- launch wrappers for the slant-stack HIP kernels (ss_kernels.hip).
// All device pointers; every wrapper enqueues on `stream` and returns hipSuccess or
// the launch error.  Layout names:
//   image      [slice][n][n]        row-major, row 0 = ymin (bottom)
//   sino       [slice][R][A]        the reference layout (angle contiguous)
//   sinoT      [slice][A][R]        angle-major (detector contiguous), internal
//   gatherA/B  [slice][n][pitch]    Radon gather layouts (pitch = n+2 rounded up to 16) of float2 (lower tap, difference to the
//                                   upper tap) per interpolation cell; A: image columns, B: rows
#ifndef PYRAFT_AMD_SS_KERNELS_H
#define PYRAFT_AMD_SS_KERNELS_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "ss_geometry.h"

// Radon output modes; SS_OUT_PARTIAL: a row slab's partial ray sums (launches of few
// blocks), which a finishing pass stores in one of the other modes
enum { SS_OUT_SINO = 0, SS_OUT_SINOT = 1, SS_OUT_EM_RATIO = 2, SS_OUT_PARTIAL = 3 };
// Backprojection output modes
enum { SS_BP_PLAIN = 0, SS_BP_EM_UPDATE = 1, SS_BP_PARTIAL = 2 };

// Backprojection kernel variants (chosen per plan, by the autotuner or by hand).
// A block is 256 threads over a tile of 32 columns x 8*rows pixel rows; each thread
// owns `rows` pixels of one column.  `stage` is the LDS capacity for the staged
// sinogram windows in (value, difference) pairs of 8 bytes; a stage holds
// min(64, stage / window) angles.
#define SS_BP_NVARIANTS 6
#define SS_BP_DEFAULT_VARIANT 0
#define SS_RT_DEFAULT_VARIANT 3  // pixel-driven transpose tile (4 rows per thread, 2048 staged pairs)
static const int kSSBpRows[SS_BP_NVARIANTS] = {4, 2, 8, 4, 2, 8};
static const int kSSBpStage[SS_BP_NVARIANTS] = {4096, 2048, 4096, 2048, 4096, 2048};

struct SSDevPlan {
  SSGeom g;
  const SSAngle* angles;   // [A]
  const int2* bounds;      // [A][R]   exact sample range per ray; (0,-1) empty, (0,-2) irregular
  const int4* wbounds;     // [A][ceil(R/64)] (union lo, union hi, inter lo, inter hi)
  const int* irregular;    // list of j*R + k
  const double* ray_t;       // mesh plans: detector coordinate of each ray [R]; nullptr: uniform grid
  const int64_t* ray_alpha;  // mesh plans: Q32.32 trajectory offset per ray [A][R]; nullptr: alpha0 + k*gamma
  int nirregular;
  int nwaves;              // ceil(R/64)
  int bp_variant;          // index into kSSBpRows / kSSBpStage
  int bp_window;           // rays staged per angle per pixel tile (for bp_variant)
  int bp_angles;           // angles per LDS stage (for bp_variant)
  int radon_slab_rows;     // Radon: sample rows per slab (n: no division; 0: the default rule)
  int radon_shape;         // Radon: block shape (0: 4096 staged records, 32-row chunks; 1: 4864, 40)
  // exact transpose, pixel-driven form: tile variant (the backprojection's shapes), rays
  // staged per angle and angles per stage; rt_angles = 0: not available (rays closer than
  // one cell on some angle), the line form runs instead
  int rt_variant, rt_window, rt_angles;
};

// ---- batched execution over slices of different geometry
// Packed buffers hold every slice at its own offset.  One launch covers a list of
// entries (slice, angle range); its work items are numbered consecutively entry by
// entry (item0 = first item of the entry) and a fixed grid of resident blocks walks
// them with stride gridDim.x.
struct SSBatchSlice {
  SSDevPlan p;
  int64_t img_off;   // floats: image [n][n]
  int64_t sino_off;  // floats: sinogram of R*A values (either layout)
  int64_t lay_off;   // float2: one gather layout [n][pitch]
};
struct SSBatchEntry {
  int slice, j0, j1, item0;
  // Radon launches of few blocks and their finishing pass: the entry's sample rows in
  // nslab slabs of slab_rows, its partial ray sums at part[part_off ...] ([z][j - j0][k])
  int nslab, slab_rows;
  int64_t part_off;
  int which;  // gather launches: the layouts to build (SS_GATHER_*), those angles j0..j1-1 read
};
struct SSBatchLaunch {
  const SSBatchSlice* slices;  // device
  const SSBatchEntry* entries; // device
  int nentries;
  int total;                   // work items
  int grid;                    // blocks launched (<= total)
};

// ---- plan construction
hipError_t ss_launch_plan_bounds(const SSGeom& g, const SSAngle* angles, int2* bounds, int* irregular,
                                 int* nirregular, int capacity, const double* ray_t, int64_t* ray_alpha,
                                 hipStream_t stream);
hipError_t ss_launch_plan_wave_bounds(const SSGeom& g, const int2* bounds, int4* wbounds, hipStream_t stream);

// ---- layouts
// the gather layouts are built in square tiles of SS_GATHER_TILE records per direction;
// records q = 0..n+1 of each row
#define SS_GATHER_TILE 64
static inline __host__ __device__ int ss_gather_tiles(int n) { return (n + 2 + SS_GATHER_TILE - 1) / SS_GATHER_TILE; }
// which layouts to build: the Radon reads angle j's samples from gatherA when
// SSAngle::branch_a, else from gatherB, so a range of angles on one branch needs only one
#define SS_GATHER_A 1
#define SS_GATHER_B 2
#define SS_GATHER_BOTH 3
hipError_t ss_launch_gather_layouts(const SSDevPlan& p, const float* image, float2* gatherA, float2* gatherB,
                                    int nslices, hipStream_t stream, int which = SS_GATHER_BOTH);
hipError_t ss_launch_sino_to_t(const float* sino, float* sinoT, int R, int A, int nslices, hipStream_t stream);
hipError_t ss_launch_t_to_sino(const float* sinoT, float* sino, int R, int A, int nslices, hipStream_t stream);

// ---- forward
struct SSEmRatioArgs {
  const float* meas;   // measured sinogram, angle-major [slice][A][R]
  double* ll_partials; // one FP64 partial per block (+ one per irregular ray)
};
// A Radon launch of fewer than SS_RADON_SLAB_BLOCKS blocks divides the sample rows of
// every (angle block, ray group) into slabs taken by separate blocks, which write partial
// ray sums; a finishing pass adds them.  ss_radon_slab_plan: slab count (1: no division)
// and rows per slab of an n x n image in a launch of `blocks` blocks.
#ifndef RADON_WAVES
#define RADON_WAVES 8  // angles (one per wave) of a Radon block
#endif
#define SS_RADON_SLAB_BLOCKS (8192 / RADON_WAVES)
#define RADON_SLAB_ALIGN 32  // rows per row slab: a multiple of this
int ss_radon_slab_plan(int n, long long blocks, int* rows);
// blocks of the Radon of angles [j0, j1) on an image with `nwaves` ray groups
static inline long long ss_radon_blocks(int nwaves, int j0, int j1) {
  return (long long)nwaves * ((j1 - j0 + RADON_WAVES - 1) / RADON_WAVES);
}
// Slabs of a Radon launch over angles [j0, j1) of nslices slices: p.radon_slab_rows when
// set (the plan's tuned choice), else the default rule; returns the slab count (1: not
// divided) and sets *rows.
int ss_radon_slabs(const SSDevPlan& p, int j0, int j1, int nslices, int* rows);
// scratch: ss_radon_scratch(p, j0, j1, nslices) bytes (the slabs' partial sums; 0 bytes,
// and scratch may be null, when the launch is not divided)
size_t ss_radon_scratch(const SSDevPlan& p, int j0, int j1, int nslices);
hipError_t ss_launch_radon(const SSDevPlan& p, const float2* gatherA, const float2* gatherB, int j0, int j1,
                           int nslices, int mode, float* out, SSEmRatioArgs em, float* scratch, hipStream_t stream);
// FP64 log-likelihood partials an EM_RATIO launch writes (ll_partials[0 .. count))
int ss_radon_partials(const SSDevPlan& p, int j0, int j1, int nslices);

// ---- backprojection (reads angle-major sinogram)
struct SSEmUpdateArgs {
  float* x;            // image updated in place
  const float* weight; // BP_b(1)
};
// A launch of fewer than SS_BP_SPLIT_TILES pixel tiles (a small image) divides its angles
// into groups taken by separate blocks, at least SS_BP_GROUP_ANGLES angles each, so that
// the launch has about SS_BP_SPLIT_BLOCKS blocks; each group writes a partial image
// (unscaled sums) to scratch and a finishing pass adds the groups in group order, scales by
// dtheta and stores or applies the EM update.  ss_bp_groups: the group count (1: no
// division); ss_bp_scratch: the bytes of the partial images (0 when not divided).
#define SS_BP_SPLIT_TILES 512
#define SS_BP_SPLIT_BLOCKS 1024
#define SS_BP_GROUP_ANGLES 8
int ss_bp_groups(const SSDevPlan& p, int j0, int j1, int nslices);
size_t ss_bp_scratch(const SSDevPlan& p, int j0, int j1, int nslices);
// scratch: ss_bp_scratch(p, j0, j1, nslices) bytes, or nullptr (then the launch is not divided)
hipError_t ss_launch_backprojection(const SSDevPlan& p, const float* sinoT, int j0, int j1, int nslices, int mode,
                                    float* out, SSEmUpdateArgs em, float* scratch, hipStream_t stream);

// ---- exact transpose of the Radon (reads angle-major sinogram)
// scratch bytes the transpose needs: one transposed n x n FP32 image per slice, where the
// lines of the angles that walk image columns accumulate before k_transpose_add
size_t ss_radon_transpose_scratch(const SSDevPlan& p, int nslices);
hipError_t ss_launch_radon_transpose(const SSDevPlan& p, const float* sinoT, int j0, int j1, int nslices,
                                     float* out, float* scratch, hipStream_t stream);

// ---- batched forms (items: gather = SS_GATHER_TILE^2 tiles; radon = 64-ray groups x blocks of
//      8 angles x row slabs; finish = 256 rays of an entry; fixup = irregular rays; bp = pixel
//      tiles of the variant; transpose = 32x32 tiles)
int ss_bp_tiles(int n, int variant);
// gather tile side for a batch of images with sides n[0..count): SS_GATHER_TILE when the
// batch has at least 1024 such tiles, else 32 (items are tiles of that side)
int ss_batch_gather_tile(const int* n, int count);
hipError_t ss_launch_batch_gather(const SSBatchLaunch& L, int tile, const float* images, float2* gatherA,
                                  float2* gatherB, hipStream_t stream);
hipError_t ss_launch_batch_transpose(const SSBatchLaunch& L, const float* in, float* out, int to_angle_major,
                                     hipStream_t stream);
// part != nullptr: the entries' partial sums go to part (launch in slab form, then
// ss_launch_batch_radon_finish with `mode`); otherwise the Radon stores in `mode` itself
hipError_t ss_launch_batch_radon(const SSBatchLaunch& L, const float2* gatherA, const float2* gatherB, float* part,
                                 int mode, float* out, const float* meas, double* ll_partials, hipStream_t stream);
hipError_t ss_launch_batch_radon_finish(const SSBatchLaunch& L, const float* part, int mode, float* out,
                                        const float* meas, double* ll_partials, hipStream_t stream);
hipError_t ss_launch_batch_radon_fixup(const SSBatchLaunch& L, const float2* gatherA, const float2* gatherB, int mode,
                                       float* out, const float* meas, double* ll_partials, hipStream_t stream);
hipError_t ss_launch_batch_backprojection(const SSBatchLaunch& L, int variant, const float* sinoT, int mode,
                                          float* out, float* x, const float* weight, hipStream_t stream);
// per-slice sums of LL partials: ll_out[i] = sum of the `nranges` (start, count) ranges of slice i
hipError_t ss_launch_batch_ll(const int2* ranges, int nranges, int nslices, const double* partials, double* ll_out,
                              hipStream_t stream);
hipError_t ss_launch_fill_f64(double* p, double v, size_t n, hipStream_t stream);

// grid of a batched kernel: the occupancy-based resident grid (blocks per CU x CUs) of the
// Radon and backprojection kernels; one block per item for the layout kernels
int ss_batch_grid(int kind, int variant, int mode, int total);
enum { SS_KIND_GATHER = 0, SS_KIND_TRANSPOSE, SS_KIND_RADON, SS_KIND_FIXUP, SS_KIND_BP, SS_KIND_RADON_FINISH };

// ---- FBP row filter (rows of length R, circular convolution with g)
// FFT form of the row filter (ss_kernels.hip): the spectrum of a tap set, built on the
// device, and the filter launch.  Sizes up to R = SS_FFT_MAX / 2 (two M-point FP64
// complex buffers in LDS); ss_launch_ramp_filter takes this form where it fits.
#define SS_FFT_MAX 4096
struct SSFftTaps {
  const double2* G;  // [M] FFT of the zero-padded taps
  const double2* W;  // [M] exp(-2 pi i k / M)
  int R, M, lgM;
};
int ss_fft_size(int R, int* lg);     // M = the power of two >= 2R (0 if above SS_FFT_MAX)
size_t ss_fft_taps_bytes(int R);     // device bytes of W and G (0 if the FFT form does not fit)
hipError_t ss_fft_taps_build(const float* g, int R, bool toeplitz, void* mem, SSFftTaps* t, hipStream_t stream);
// in_cols / out_cols: 0 for angle-major rows, A for the (R, A) layout (row = slice * A + a)
hipError_t ss_launch_fft_filter(const float* in, float* out, const SSFftTaps& t, int nrows, hipStream_t stream,
                                int in_cols = 0, int out_cols = 0);
hipError_t ss_launch_ramp_filter(const float* in, float* out, const float* g, int R, int nrows, hipStream_t stream,
                                 bool toeplitz = false);

// ---- FP64 reductions / EM helpers
hipError_t ss_launch_sum_partials(const double* partials, int n, double* dst, hipStream_t stream);
hipError_t ss_launch_fill(float* p, float v, size_t n, hipStream_t stream);
// transmission EM elementwise steps
hipError_t ss_launch_tr_expected(float* p, const float* flat, size_t n, hipStream_t stream);
hipError_t ss_launch_tr_update(float* x, const float* b, const float* w, size_t n, hipStream_t stream);

// ---- phantoms (FP64)
struct SSEllipse {  // value, a, b, x0, y0, cos(rot), sin(rot)
  double val, a, b, x0, y0, c, s;
};
hipError_t ss_launch_phantom(const SSEllipse* ell, int nell, int lines, int cols, double tl_x, double tl_y,
                             double hs, double vs, double* out64, float* out32, hipStream_t stream);
// per (ellipse, column): c, s, m, mab, delt  (host-built, see raft/raft_radon_fromdesc.c)
struct SSEllipseView {
  double m, mab, delt;
};
hipError_t ss_launch_radon_fromdesc(const SSEllipseView* views, int nell, int lines, int cols, double* out64,
                                    hipStream_t stream);

#endif
