// ss_kernels.hip -- gfx950 kernels for the slant-stack projector pair, its exact
// transpose, the FBP row filter, the fused EM epilogues and the ellipse phantoms.
//
// Arithmetic is FP32 with FP32 accumulation.  FP64 appears only in the plan (angle
// tables, exact sample ranges), in the knife-edge predicate fallback of the
// backprojection and in the EM log-likelihood sums.
#include <hip/hip_runtime.h>

#include <type_traits>

#include <algorithm>

#include "ss_kernels.h"
#include "ss_plan_logic.h"

#define SS_WAVE 64
#define SS_INV_FIX 2.3283064365386963e-10f  // 2^-32
#define SS_FIX_ONE_I64 ((int64_t)1 << 32)    // one cell in Q32.32

// =====================================================================================
// plan: exact per-ray sample ranges
// =====================================================================================

// ray_t == nullptr: the uniform detector grid.  Otherwise ray k sits at ray_t[k] (the
// arbitrary-mesh form) and its Q32.32 trajectory offset goes to ray_alpha[j][k].
__global__ __launch_bounds__(256) void k_plan_bounds(SSGeom g, const SSAngle* __restrict__ angles,
                                                     int2* __restrict__ bounds, int* __restrict__ irregular,
                                                     int* __restrict__ nirregular, int capacity,
                                                     const double* __restrict__ ray_t, int64_t* __restrict__ ray_alpha) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y;
  if (k >= g.nrays) return;
  const SSAngle a = angles[j];
  const SSTraj tr = ss_trajectory(g, a);
  int lo, hi, irr;
  if (ray_t) {
    const double t = ray_t[k];
    const double u0 = ss_ray_alpha(g, a, t);
    ray_alpha[(size_t)j * g.nrays + k] = __double2ll_rn(u0 * SS_FIX_ONE);
    irr = ss_ray_range_t(g, a, tr.beta, t, u0, &lo, &hi);
  } else {
    irr = ss_ray_range(g, a, tr, k, &lo, &hi);
  }
  if (irr) {
    lo = 0;
    hi = -2;  // empty for the main kernels; the fixup kernels evaluate it sample by sample
    const int slot = atomicAdd(nirregular, 1);
    if (slot < capacity) irregular[slot] = j * g.nrays + k;
  }
  bounds[(size_t)j * g.nrays + k] = make_int2(lo, hi);
}

// Per (angle, group of 64 rays): union and intersection of the non-empty ranges.
__global__ __launch_bounds__(64) void k_plan_wave_bounds(SSGeom g, const int2* __restrict__ bounds,
                                                         int4* __restrict__ wbounds) {
  const int j = blockIdx.y;
  const int grp = blockIdx.x;
  const int k = grp * SS_WAVE + threadIdx.x;
  int lo = 0x7fffffff, hi = -1, ilo = -0x7fffffff, ihi = 0x7fffffff;
  bool any = false;
  if (k < g.nrays) {
    const int2 b = bounds[(size_t)j * g.nrays + k];
    if (b.y >= b.x) {
      lo = b.x;
      hi = b.y;
      ilo = b.x;
      ihi = b.y;
      any = true;
    }
  }
  // every lane takes part in every shuffle (no short-circuit around them)
  for (int off = 32; off > 0; off >>= 1) {
    lo = min(lo, __shfl_xor(lo, off));
    hi = max(hi, __shfl_xor(hi, off));
    ilo = max(ilo, __shfl_xor(ilo, off));
    ihi = min(ihi, __shfl_xor(ihi, off));
  }
  const bool any_lane = __ballot(any ? 1 : 0) != 0ull;
  if (threadIdx.x == 0) {
    int4 r;
    if (!any_lane) {
      r = make_int4(0, -1, 0, -1);
    } else {
      r = make_int4(lo, hi, ilo, ihi);
    }
    wbounds[(size_t)j * gridDim.x + grp] = r;
  }
}

hipError_t ss_launch_plan_bounds(const SSGeom& g, const SSAngle* angles, int2* bounds, int* irregular,
                                 int* nirregular, int capacity, const double* ray_t, int64_t* ray_alpha,
                                 hipStream_t stream) {
  dim3 grid((g.nrays + 255) / 256, g.nangles);
  hipLaunchKernelGGL(k_plan_bounds, grid, dim3(256), 0, stream, g, angles, bounds, irregular, nirregular, capacity,
                     ray_t, ray_alpha);
  return hipGetLastError();
}

hipError_t ss_launch_plan_wave_bounds(const SSGeom& g, const int2* bounds, int4* wbounds, hipStream_t stream) {
  dim3 grid((g.nrays + SS_WAVE - 1) / SS_WAVE, g.nangles);
  hipLaunchKernelGGL(k_plan_wave_bounds, grid, dim3(SS_WAVE), 0, stream, g, bounds, wbounds);
  return hipGetLastError();
}

// =====================================================================================
// layouts
// =====================================================================================

// image [n][n] -> pair layouts [n][pitch] of float2, q in [0, n+1] (pitch >= n+2):
//   gatherB[r][q] = (F(r, q-1), F(r, q) - F(r, q-1))      image rows
//   gatherA[c][q] = (F(q-1, c), F(q, c) - F(q-1, c))      image columns
// with F = the image inside [0, n)^2 and 0 outside.  Element m+1 of a row holds the
// lower tap of interpolation cell m and the difference to the upper tap, so a Radon
// sample is one aligned 8-byte load and a + w*d.
//
// One tile = GT values of q in both directions (GT = 32 or 64) for a block of NW waves:
// the image square with a one-pixel halo ((GT+1)^2) goes to LDS, every halo load issued
// before the first LDS write; then the waves write rows of GT records (GT * 8
// contiguous bytes), first of layout B, then of layout A, 64 / GT rows per store.  64 x 64
// tiles with 8 waves serve launches of many tiles; 32 x 32 tiles with 4 waves serve small
// ones (one 1024^2 image), where the larger tiles leave the CUs short of waves.
template <int GT, int NW>
__device__ __forceinline__ void ss_gather_tile(const float* __restrict__ src, float2* __restrict__ dA,
                                               float2* __restrict__ dB, int n, int pitch, int bx, int by,
                                               float (*T)[GT + 1], int which = SS_GATHER_BOTH) {
  constexpr int H = GT + 1;                       // halo side
  constexpr int NT = NW * SS_WAVE;                // threads
  constexpr int LOADS = (H * H + NT - 1) / NT;    // halo values per thread
  constexpr int SUB = SS_WAVE / GT;               // rows per wave store
  float v[LOADS];
#pragma unroll
  for (int i = 0; i < LOADS; ++i) {
    const int idx = (int)threadIdx.x + NT * i;
    const int a = idx / H, b = idx - a * H;
    const int r = by - 1 + a, c = bx - 1 + b;
    const bool in = idx < H * H && r >= 0 && r < n && c >= 0 && c < n;
    v[i] = in ? src[(size_t)r * n + c] : 0.f;
  }
#pragma unroll
  for (int i = 0; i < LOADS; ++i) {
    const int idx = (int)threadIdx.x + NT * i;
    if (idx < H * H) T[idx / H][idx % H] = v[i];
  }
  __syncthreads();
  const int lane = threadIdx.x & (SS_WAVE - 1);
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x / SS_WAVE);
  const int x = lane % GT, sub = lane / GT;
  // B: record (r, q) = (F(r, q-1), F(r, q) - F(r, q-1)) = T[r - by + 1][q - bx (+1)]
  const int qb = bx + x;
  for (int i = wv * SUB + sub; (which & SS_GATHER_B) && i < GT; i += NW * SUB) {
    const int r = by + i;
    if (r < n && qb <= n + 1) {
      const float lo = T[i + 1][x];
      dB[(size_t)r * pitch + qb] = make_float2(lo, T[i + 1][x + 1] - lo);
    }
  }
  // A: record (c, q) = (F(q-1, c), F(q, c) - F(q-1, c)) = T[q - by (+1)][c - bx + 1]
  const int qa = by + x;
  for (int j = wv * SUB + sub; (which & SS_GATHER_A) && j < GT; j += NW * SUB) {
    const int cc = bx + j;
    if (cc < n && qa <= n + 1) {
      const float lo = T[x][j + 1];
      dA[(size_t)cc * pitch + qa] = make_float2(lo, T[x + 1][j + 1] - lo);
    }
  }
}

template <int GT, int NW>
__global__ __launch_bounds__(NW * SS_WAVE) void k_gather_pairs(const float* __restrict__ img, float2* __restrict__ gA,
                                                               float2* __restrict__ gB, int n, int pitch, int which) {
  __shared__ float T[GT + 1][GT + 1];  // F(by-1+a, bx-1+b)
  const size_t slice = blockIdx.z;
  ss_gather_tile<GT, NW>(img + slice * n * n, gA + slice * n * pitch, gB + slice * n * pitch, n, pitch,
                         blockIdx.x * GT, blockIdx.y * GT, T, which);
}

hipError_t ss_launch_gather_layouts(const SSDevPlan& p, const float* image, float2* gatherA, float2* gatherB,
                                    int nslices, hipStream_t stream, int which) {
  const int n = p.g.n;
  const int tiles = ss_gather_tiles(n);
  if (!(which & SS_GATHER_BOTH)) return hipSuccess;
  if ((long long)tiles * tiles * nslices >= 1024) {  // at least 4 large tiles per CU
    hipLaunchKernelGGL((k_gather_pairs<SS_GATHER_TILE, 8>), dim3(tiles, tiles, nslices), dim3(8 * SS_WAVE), 0, stream,
                       image, gatherA, gatherB, n, p.g.pitch, which);
  } else {
    const int t32 = (n + 2 + 31) / 32;
    hipLaunchKernelGGL((k_gather_pairs<32, 4>), dim3(t32, t32, nslices), dim3(4 * SS_WAVE), 0, stream, image,
                       gatherA, gatherB, n, p.g.pitch, which);
  }
  return hipGetLastError();
}

// [rows][cols] -> [cols][rows], one 32x32 tile
__device__ __forceinline__ void ss_transpose_tile(const float* __restrict__ src, float* __restrict__ dst, int rows,
                                                  int cols, int bx, int by, float (*tile)[33]) {
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  for (int yy = ty; yy < 32; yy += 8) {
    const int r = by + yy, c = bx + tx;
    if (r < rows && c < cols) tile[yy][tx] = src[(size_t)r * cols + c];
  }
  __syncthreads();
  for (int xx = ty; xx < 32; xx += 8) {
    const int c = bx + xx, r = by + tx;
    if (r < rows && c < cols) dst[(size_t)c * rows + r] = tile[tx][xx];
  }
}

// [s][rows][cols] -> [s][cols][rows]
__global__ __launch_bounds__(256) void k_transpose2d(const float* __restrict__ in, float* __restrict__ out, int rows,
                                                     int cols) {
  __shared__ float tile[32][33];
  const size_t off = (size_t)blockIdx.z * rows * cols;
  ss_transpose_tile(in + off, out + off, rows, cols, blockIdx.x * 32, blockIdx.y * 32, tile);
}

hipError_t ss_launch_sino_to_t(const float* sino, float* sinoT, int R, int A, int nslices, hipStream_t stream) {
  dim3 grid((A + 31) / 32, (R + 31) / 32, nslices);
  hipLaunchKernelGGL(k_transpose2d, grid, dim3(256), 0, stream, sino, sinoT, R, A);
  return hipGetLastError();
}

hipError_t ss_launch_t_to_sino(const float* sinoT, float* sino, int R, int A, int nslices, hipStream_t stream) {
  dim3 grid((R + 31) / 32, (A + 31) / 32, nslices);
  hipLaunchKernelGGL(k_transpose2d, grid, dim3(256), 0, stream, sinoT, sino, A, R);
  return hipGetLastError();
}

// =====================================================================================
// forward: ray-driven slant-stack Radon
//
// One wave = 64 consecutive rays of one angle.  Sample s walks
// gather rows (image rows for branch B, image columns for branch A, stored
// transposed), so at each s the 64 lanes read a contiguous stretch of one padded
// row.  Loop bounds come from the plan's exact FP64 ranges: predicated head and
// tail, unpredicated middle.
// =====================================================================================

__device__ __forceinline__ double ss_block_sum_f64(double v, double* scratch) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  const int wv = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) scratch[wv] = v;
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x == 0)
    for (int i = 0; i < (int)(blockDim.x >> 6); ++i) r += scratch[i];
  return r;  // valid in thread 0
}

// The plan fields the forward kernel reads (kept small: in the batched kernels they
// are loaded per item into scalar registers).
struct SSRadonView {
  const SSAngle* angles;
  const int2* bounds;
  const int4* wbounds;
  const int64_t* ray_alpha;  // per-ray trajectory offsets of a mesh plan, else nullptr
  int n, R, A, pitch, nwaves;
};

__device__ __forceinline__ SSRadonView ss_radon_view(const SSDevPlan& p) {
  return SSRadonView{p.angles, p.bounds, p.wbounds, p.ray_alpha, p.g.n, p.g.nrays, p.g.nangles, p.g.pitch, p.nwaves};
}

// out / meas point at this slice's sinogram
template <int MODE>
__device__ __forceinline__ void ss_radon_store(const SSRadonView& p, int j, int k, float val, float* out,
                                               const float* meas, double* ll) {
  const int R = p.R, A = p.A;
  if (MODE == SS_OUT_SINO) {
    out[(size_t)k * A + j] = val;
  } else if (MODE == SS_OUT_SINOT) {
    out[(size_t)j * R + k] = val;
  } else {
    const size_t idx = (size_t)j * R + k;
    const float s = meas[idx];
    float rho = 0.f;
    if (val > 0.f) {
      rho = s / val;
      *ll += (double)s * log((double)val) - (double)val;
    }
    out[idx] = rho;
  }
}

// A block is 8 waves: 8 consecutive angles over the same 64-ray group, one angle per wave.
// Launches too small to fill the chip (few angles, e.g. one OSEM subset) divide the sample
// rows into slabs taken by separate blocks instead (ss_radon_slab_plan).

typedef float ss_f32x2 __attribute__((ext_vector_type(2)));
#define SS_RSRC_DWORD3 0x00020000  // gfx9 buffer word 3: 32-bit data format, no swizzle

// Structured buffer load of one 8-byte record: address = base + soffset + vindex * stride.
// The record index comes straight from the high word of the Q32.32 trajectory, so a
// sample costs no address arithmetic.
__device__ ss_f32x2 ss_struct_load_f32x2(__amdgpu_buffer_rsrc_t rsrc, int vindex, int voffset, int soffset,
                                         int aux) __asm("llvm.amdgcn.struct.ptr.buffer.load.v2f32");

// tap pair (lower tap, difference) at record `rec` of the gather row at byte offset
// `row_off` (per-slice offsets stay below 2^32 for n <= 16384; unsigned on the hardware)
__device__ __forceinline__ ss_f32x2 ss_pair(__amdgpu_buffer_rsrc_t rs, uint32_t row_off, int rec) {
  return ss_struct_load_f32x2(rs, rec, 0, (int)row_off, 0);
}

// Radon block shapes, picked per launch by the autotuner: staged (tap, difference)
// records per block and sample rows per chunk.  Shape 0: 4096 records (32 KiB), chunks
// of up to 32 rows; shape 1: 4864 records (38 KiB; four blocks still fit a CU's
// 160 KiB), chunks of up to 40 rows.
template <int SH>
struct SSRadonShape {
  static constexpr int CAP = SH == 0 ? 4096 : 4864;
  static constexpr int CHUNK = SH == 0 ? 32 : 40;
};
#ifndef RADON_ROW_ROUND
#define RADON_ROW_ROUND 8  // chunk rows rounded down to a multiple of this where the window allows
#endif

#define RADON_TABLE SS_WAVE  // chunks per window table: one (chunk, wave) pair per thread

template <int CAP>
struct SSRadonLds {
  double red[RADON_WAVES];
  int64_t wpar[RADON_WAVES][3];  // per wave: lowest and highest trajectory offset of its rays, slope
  int wmeta[RADON_WAVES][4];     // per wave: branch (-1: no samples), first row, last row
  int wred[2][RADON_WAVES];      // per wave: widest window of a table pass (two passes in turn)
  int2 table[RADON_TABLE];       // per chunk of the super-chunk: first staged record, records per row
  __attribute__((aligned(16))) float2 stage[CAP + 2];  // chunk window; record CAP stays (0, 0)
};

typedef __attribute__((address_space(1))) const void* ss_global_cptr;
typedef __attribute__((address_space(3))) void* ss_lds_ptr;

// wave-uniform copies (scalar registers) of values the compiler cannot prove uniform
__device__ __forceinline__ int ss_rfl(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ int64_t ss_rfl64(int64_t v) {
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)v);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)((uint64_t)v >> 32));
  return (int64_t)(((uint64_t)hi << 32) | lo);
}
__device__ __forceinline__ float ss_rflf(float v) { return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v))); }

// Direct sample reader: (row s, trajectory u) -> (lower tap, difference), where u is
// the Q32.32 trajectory plus one cell, its high word the record index.
struct SSGatherReader {
  __amdgpu_buffer_rsrc_t rs;
  uint32_t row_bytes;
  __device__ __forceinline__ float2 operator()(int s, int64_t u) const {
    const ss_f32x2 t = ss_pair(rs, (uint32_t)s * row_bytes, (int)(u >> 32));
    return make_float2(t.x, t.y);
  }
};

// G consecutive samples s, s+1, ... of one lane's ray from the staged window, all G
// reads issued before the first is used.  With CHECK, a sample at or beyond `end` or
// outside the lane's own range [lo, lo + span] reads record CAP (past the window), which holds
// (0, 0), instead of branching: it adds exactly zero to both sums.  u advances G steps.
template <int G, bool CHECK, int CAP>
__device__ __forceinline__ void ss_radon_group(const float2* __restrict__ st, int s, int end, int64_t& u,
                                               int64_t step, int lo, unsigned span, float& acc_a, float& acc_d) {
  float2 t[G];
  uint32_t w[G];
  int idx[G];
#pragma unroll
  for (int g = 0; g < G; ++g) {
    idx[g] = (int)(u >> 32);
    if (CHECK) idx[g] = (s + g < end && (unsigned)(s + g - lo) <= span) ? idx[g] : CAP;
    w[g] = (uint32_t)u;
    u += step;
    __asm__ volatile("" : "+v"(u));  // one 64-bit add per sample, not offsets re-derived per sample
  }
#pragma unroll
  for (int g = 0; g < G; ++g) t[g] = st[idx[g]];
  float sa = t[0].x;
  if (G == 4) sa = (t[0].x + t[1].x) + (t[2].x + t[3].x);
  if (G == 8) sa = ((t[0].x + t[1].x) + (t[2].x + t[3].x)) + ((t[4].x + t[5].x) + (t[6].x + t[7].x));
  acc_a += sa;
#pragma unroll
  for (int g = 0; g < G; ++g) acc_d = fmaf((float)w[g], t[g].y, acc_d);
}

// Samples [sa, sb) of one lane's ray from the staged window (u carries the row offset
// in its high word): groups of 4 with the range test before zlo (outside the group's
// common range), unchecked groups of 8 in [zlo, zhi1), and the rest (the last group
// of the common range and the samples from zhi1 on) again in checked groups of 4.
template <int CAP>
__device__ __forceinline__ void ss_radon_rows_staged(const float2* __restrict__ st, int sa, int sb, int zlo,
                                                     int zhi1, int64_t u, int64_t step, int lo, unsigned span,
                                                     float& acc_a, float& acc_d) {
  int s = sa;
  const int h_end = min(sb, zlo);
  if (s < h_end) {
    const int64_t u0 = u;
    for (; s < h_end; s += 4) ss_radon_group<4, true, CAP>(st, s, h_end, u, step, lo, span, acc_a, acc_d);
    s = h_end;
    u = u0 + (int64_t)(h_end - sa) * step;
  }
  const int m_end = min(sb, zhi1);
  for (; s + 8 <= m_end; s += 8) ss_radon_group<8, false, CAP>(st, s, 0, u, step, lo, span, acc_a, acc_d);
  for (; s < sb; s += 4) ss_radon_group<4, true, CAP>(st, s, sb, u, step, lo, span, acc_a, acc_d);
}

// Samples [sa, sb) of one lane's ray: predicated before zlo and from zhi1 on (outside
// the group's common range), unpredicated in between with 4 samples in flight, each
// on its own trajectory register (one 64-bit add per sample; the empty asm keeps the
// compiler from re-deriving the four from one base, which costs extra adds).
template <class Rd>
__device__ __forceinline__ void ss_radon_rows(const Rd& rd, int sa, int sb, int zlo, int zhi1, int64_t u,
                                              int64_t beta, int lo, unsigned span, float& acc_a, float& acc_d) {
  int s = sa;
  const int h_end = min(sb, zlo);
  for (; s < h_end; ++s) {
    if ((unsigned)(s - lo) <= span) {
      const float2 t = rd(s, u);
      acc_a += t.x;
      acc_d = fmaf((float)(uint32_t)u, t.y, acc_d);
    }
    u += beta;
  }
  const int m_end = min(sb, zhi1);
  if (s + 4 <= m_end) {
    int64_t u0 = u, u1 = u + beta, u2 = u + 2 * beta, u3 = u + 3 * beta;
    const int64_t beta4 = 4 * beta;
    for (; s + 4 <= m_end; s += 4) {
      const float2 t0 = rd(s, u0), t1 = rd(s + 1, u1), t2 = rd(s + 2, u2), t3 = rd(s + 3, u3);
      acc_a += (t0.x + t1.x) + (t2.x + t3.x);
      acc_d = fmaf((float)(uint32_t)u0, t0.y, acc_d);
      acc_d = fmaf((float)(uint32_t)u1, t1.y, acc_d);
      acc_d = fmaf((float)(uint32_t)u2, t2.y, acc_d);
      acc_d = fmaf((float)(uint32_t)u3, t3.y, acc_d);
      u0 += beta4;
      u1 += beta4;
      u2 += beta4;
      u3 += beta4;
      __asm__ volatile("" : "+v"(u0), "+v"(u1), "+v"(u2), "+v"(u3));
    }
    u = u0;
  }
  for (; s < m_end; ++s) {
    const float2 t = rd(s, u);
    acc_a += t.x;
    acc_d = fmaf((float)(uint32_t)u, t.y, acc_d);
    u += beta;
  }
  for (; s < sb; ++s) {
    if ((unsigned)(s - lo) <= span) {
      const float2 t = rd(s, u);
      acc_a += t.x;
      acc_d = fmaf((float)(uint32_t)u, t.y, acc_d);
    }
    u += beta;
  }
}

// Window table of `nch` chunks of `rows` sample rows from row sc (the last chunk ends
// at s_hi at the latest): for each chunk, the gather records [rec0, rec0 + Wc) of each
// row that the block's waves read in its rows (Wc = 0: none).  Each thread takes one
// (chunk, wave) pair, bounds the cells of that wave's rays over the chunk's rows at the
// corners (the trajectory is affine in ray and row; valid samples have cells in
// [-1, n-1]), and the RADON_WAVES lanes of a chunk take the union.  Returns the widest
// Wc of the table (block-uniform); `pass` selects the reduction slots, so that a second
// pass right after the first needs no barrier in between.
template <int CAP>
__device__ __forceinline__ int ss_radon_windows(SSRadonLds<CAP>& lds, int pass, int sc, int rows, int nch, int s_hi,
                                                int n) {
  const int t = threadIdx.x;
  const int c = t / RADON_WAVES, w = t % RADON_WAVES;
  int cmin = 0x7fffffff, cmax = -0x7fffffff;
  if (c < nch) {
    const int s0 = sc + c * rows, s1 = min(s0 + rows, s_hi + 1);
    const int ra = max(s0, lds.wmeta[w][1]), rb = min(s1 - 1, lds.wmeta[w][2]);
    if (lds.wmeta[w][0] >= 0 && ra <= rb) {
      const int64_t beta = lds.wpar[w][2];
      const int64_t c00 = lds.wpar[w][0] + (int64_t)ra * beta;
      const int64_t c10 = lds.wpar[w][1] + (int64_t)ra * beta;
      const int64_t ds = (int64_t)(rb - ra) * beta;
      const int m0 = (int)(c00 >> 32), m1 = (int)(c10 >> 32);
      const int m2 = (int)((c00 + ds) >> 32), m3 = (int)((c10 + ds) >> 32);
      cmin = max(min(min(m0, m1), min(m2, m3)), -1);
      cmax = min(max(max(m0, m1), max(m2, m3)), n - 1);
    }
  }
  for (int off = 1; off < RADON_WAVES; off <<= 1) {
    cmin = min(cmin, __shfl_xor(cmin, off));
    cmax = max(cmax, __shfl_xor(cmax, off));
  }
  // cells c_lo..c_hi need records c_lo+1..c_hi+1; rec0 is even so that rows of the
  // window start 16-byte aligned in the layout (even pitch) and in LDS (even Wc)
  const int rec0 = (cmin + 1) & ~1;
  int wc = cmax >= cmin ? ((cmax + 2 - rec0) + 1) & ~1 : 0;
  if (c >= nch) wc = 0;
  if (c < nch && w == 0) lds.table[c] = make_int2(rec0, wc);
  for (int off = 32; off > 0; off >>= 1) wc = max(wc, __shfl_xor(wc, off));
  if ((t & (SS_WAVE - 1)) == 0) lds.wred[pass][t / SS_WAVE] = wc;
  __syncthreads();
  int wmax = 0;
  for (int i = 0; i < RADON_WAVES; ++i) wmax = max(wmax, lds.wred[pass][i]);
  return ss_rfl(wmax);
}

// One block's work: rays [grp*64, grp*64+64) of the RADON_WAVES angles starting at
// j0 + ablk*RADON_WAVES, clipped to [j0, j1), over the sample rows
// [row_lo, row_hi).  gA/gB/out/meas point at this slice's data; in EM mode the block's
// FP64 log-likelihood partial goes to *ll_slot.  In SS_OUT_PARTIAL mode the block writes
// its partial ray sums to part_out ([j - j0][k], delta not applied) for a finishing pass.
//
// All angles of the block on one branch (the usual case): the block walks its sample
// rows in chunks of up to CHUNK rows.  For each chunk it bounds the cells every
// wave can touch (the trajectory is affine in ray and row, so the extremes are at the
// corners), copies that window of the gather rows into LDS with contiguous loads, and
// the waves interpolate from LDS.  A window too wide for LDS falls back to direct
// gathers for that chunk.  A block straddling the branch switch gathers directly from
// the layouts.
template <int MODE, int SH>
__device__ __forceinline__ void ss_radon_item(const SSRadonView& p, const float2* __restrict__ gA,
                                              const float2* __restrict__ gB, int j0, int j1, int grp, int ablk,
                                              int row_lo, int row_hi, float* __restrict__ out,
                                              const float* __restrict__ meas, double* ll_slot,
                                              float* __restrict__ part_out,
                                              SSRadonLds<SSRadonShape<SH>::CAP>& lds) {
  constexpr int CAP = SSRadonShape<SH>::CAP, CHUNK = SSRadonShape<SH>::CHUNK;
  const int lane = threadIdx.x & 63;
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int j = j0 + ablk * RADON_WAVES + wv;
  const int R = p.R, n = p.n, pitch = p.pitch;
  const int k = grp * SS_WAVE + lane;
  const bool active = j < j1;
  float acc_a = 0.f, acc_d = 0.f;
  float delta = 0.f;
  int branch = -1;
  int4 wb = make_int4(0, -1, 0, -1);
  int lo = 0, hi = -1;
  int64_t alpha = 0, gamma = 0, beta = 0;
  if (active) {
    const SSAngle& a = p.angles[j];
    delta = ss_rflf(a.delta);
    alpha = ss_rfl64(a.alpha0);
    gamma = ss_rfl64(a.gamma);
    beta = ss_rfl64(a.beta);
    const int4 w4 = p.wbounds[(size_t)j * p.nwaves + grp];
    wb = make_int4(ss_rfl(w4.x), ss_rfl(w4.y), ss_rfl(w4.z), ss_rfl(w4.w));
    if (wb.y >= wb.x) branch = ss_rfl(a.branch_a);
    if (k < R) {
      const int2 b = p.bounds[(size_t)j * R + k];
      lo = b.x;
      hi = b.y;
    }
  }
  const bool empty = hi < lo;
  const unsigned span = (unsigned)(hi - lo);
  const uint32_t row_bytes = (uint32_t)pitch * 8u;
  // this lane's trajectory offset: u = a_k + s*beta
  int64_t a_k = alpha + (int64_t)k * gamma;
  if (p.ray_alpha && active && k < R) a_k = p.ray_alpha[(size_t)j * R + k];

  bool staged = false;
  {  // staged or direct: decided per block
    // the wave's extreme trajectory offsets (over its rays on the detector)
    int64_t amin = (k < R) ? a_k : INT64_MAX, amax = (k < R) ? a_k : INT64_MIN;
    for (int off = 32; off > 0; off >>= 1) {
      amin = min(amin, (int64_t)__shfl_xor((long long)amin, off));
      amax = max(amax, (int64_t)__shfl_xor((long long)amax, off));
    }
    if (lane == 0) {
      lds.wmeta[wv][0] = branch;
      lds.wmeta[wv][1] = wb.x;
      lds.wmeta[wv][2] = wb.y;
      lds.wpar[wv][0] = amin;
      lds.wpar[wv][1] = amax;
      lds.wpar[wv][2] = beta;
    }
    if (threadIdx.x == 0) lds.stage[CAP] = make_float2(0.f, 0.f);
    __syncthreads();
    int bmin = 2, bmax = -1, s_lo = 0x7fffffff, s_hi = -1;
    for (int w = 0; w < RADON_WAVES; ++w) {
      const int b = lds.wmeta[w][0];
      if (b >= 0) {
        bmin = min(bmin, b);
        bmax = max(bmax, b);
        s_lo = min(s_lo, lds.wmeta[w][1]);
        s_hi = max(s_hi, lds.wmeta[w][2]);
      }
    }
    bmin = ss_rfl(bmin);
    bmax = ss_rfl(bmax);
    s_lo = max(ss_rfl(s_lo), row_lo);
    s_hi = min(ss_rfl(s_hi), row_hi - 1);
    staged = bmax < 0 || bmin == bmax;  // block-uniform
    if (staged && bmax >= 0) {
      const float2* lay = bmax ? gA : gB;
      const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
          (void*)lay, (short)sizeof(float2), (int)((uint32_t)n * (uint32_t)pitch * 8u), SS_RSRC_DWORD3);
      const SSGatherReader grd{rs, row_bytes};
      // the pure trajectory of this lane's ray at the first row of the current chunk
      int64_t t0 = a_k + (int64_t)s_lo * beta;
      // rows in super-chunks of up to RADON_TABLE chunks: one pass over the table gives
      // every chunk's window and the widest; chunks of CHUNK rows unless that
      // window would not fit, then of the rows it allows (a second pass)
      for (int sc = s_lo; sc <= s_hi;) {
        int rows = CHUNK;
        int nch = min(RADON_TABLE, (s_hi - sc + rows) / rows);
        const int wmax = ss_radon_windows(lds, 0, sc, rows, nch, s_hi, n);
        if (wmax * rows > CAP) {
          int r2 = CAP / wmax;
          if (r2 >= RADON_ROW_ROUND) r2 = r2 / RADON_ROW_ROUND * RADON_ROW_ROUND;
          rows = max(r2, 1);
          nch = min(RADON_TABLE, (s_hi - sc + rows) / rows);
          (void)ss_radon_windows(lds, 1, sc, rows, nch, s_hi, n);
        }
        const int64_t chunk_step = (int64_t)rows * beta;
        for (int c = 0; c < nch; ++c) {
          const int s0 = sc + c * rows, s1 = min(s0 + rows, s_hi + 1);
          const int2 tw = lds.table[c];
          const int rec0 = ss_rfl(tw.x), Wc = ss_rfl(tw.y);
          const bool fits = Wc > 0 && (s1 - s0) * Wc <= CAP;  // block-uniform
          if (c > 0) __syncthreads();  // every wave is done with the previous chunk
          if (fits) {
            // direct global -> LDS copies, 16 bytes (two records) per lane: wave wv copies
            // rows wv, wv+8, ...; the last record read is rec0 + Wc - 1 <= c_hi + 2 <= n + 1
            const float2* src0 = lay + (size_t)(s0 + wv) * pitch + rec0 + 2 * lane;
            for (int seg = 0; seg < Wc; seg += 2 * SS_WAVE) {
              if (seg + 2 * lane < Wc) {
                const float2* src = src0 + seg;
                float2* dst = lds.stage + wv * Wc + seg;
                for (int r = wv; r < s1 - s0; r += RADON_WAVES) {
                  __builtin_amdgcn_global_load_lds((ss_global_cptr)src, (ss_lds_ptr)dst, 16, 0, 0);
                  src += (size_t)RADON_WAVES * pitch;
                  dst += RADON_WAVES * Wc;
                }
              }
            }
          }
          __syncthreads();  // window staged
          const int ra2 = max(s0, wb.x), rb2 = min(s1 - 1, wb.y);
          if (branch >= 0 && ra2 <= rb2 && Wc > 0) {
            const int64_t t = t0 + (int64_t)(ra2 - s0) * beta;
            if (fits) {
              // high word: staged record index (cell + 1 - rec0) + (row - s0) * Wc
              int64_t u = t + ((int64_t)(1 - rec0 + (ra2 - s0) * Wc) << 32);
              int64_t step = beta + ((int64_t)Wc << 32);
              if (empty) {
                u = (int64_t)CAP << 32;
                step = 0;
              }
              ss_radon_rows_staged<CAP>(lds.stage, ra2, rb2 + 1, wb.z, wb.w + 1, u, step, lo, span, acc_a, acc_d);
            } else {
              int64_t u = t + ((int64_t)1 << 32);
              int64_t step = beta;
              if (empty) {
                u = 0;
                step = 0;
              }
              ss_radon_rows(grd, ra2, rb2 + 1, wb.z, wb.w + 1, u, step, lo, span, acc_a, acc_d);
            }
          }
          t0 += chunk_step;
        }
        sc += nch * rows;
      }
    }
  }
  if (!staged && active && branch >= 0) {
    // the group's sample range [wb.x, wb.y] within the rows
    const int sa = max(wb.x, row_lo), sb = min(wb.y + 1, row_hi);
    // u carries the trajectory plus one cell: its high word is the record index m + 1
    // of the lower tap in the gather row, its low word the interpolation weight
    int64_t u = a_k + (int64_t)sa * beta + ((int64_t)1 << 32);
    int64_t step = beta;
    if (empty) {
      // an empty ray reads record 0 (lower tap 0) with weight 0: exactly 0 in the unpredicated loop
      u = 0;
      step = 0;
    }
    const float2* lay = branch ? gA : gB;
    const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
        (void*)lay, (short)sizeof(float2), (int)((uint32_t)n * (uint32_t)pitch * 8u), SS_RSRC_DWORD3);
    ss_radon_rows(SSGatherReader{rs, row_bytes}, sa, sb, wb.z, wb.w + 1, u, step, lo, span, acc_a, acc_d);
  }
  double ll = 0.0;
  if (active && k < R) {
    if (MODE == SS_OUT_PARTIAL) {
      part_out[(size_t)(j - j0) * R + k] = fmaf(acc_d, SS_INV_FIX, acc_a);
    } else {
      const float val = fmaf(acc_d, SS_INV_FIX, acc_a) * delta;
      ss_radon_store<MODE>(p, j, k, val, out, meas, &ll);
    }
  }
  if (MODE == SS_OUT_EM_RATIO) {
    const double tot = ss_block_sum_f64(ll, lds.red);
    if (threadIdx.x == 0) *ll_slot = tot;
  }
}

// The slabs' partial sums of one ray, added in slab order (the per-plan and batched
// finishing passes give the same bits); unrolled so that the loads issue together.
__device__ __forceinline__ float ss_slab_sum(const float* __restrict__ src, size_t per, int nslab) {
  float sum = 0.f;
#pragma unroll 8
  for (int z = 0; z < nslab; ++z) sum += src[(size_t)z * per];
  return sum;
}

// blockIdx.z = slice * nslab + z: slab z takes sample rows [z * slab_rows, (z + 1) * slab_rows).
// In SS_OUT_PARTIAL mode the blocks write partial ray sums to part[slice][z][j - j0][k] for
// k_radon_slab_finish; in the other modes (nslab = 1) they store the output themselves.
template <int MODE, int SH>
__global__ __launch_bounds__(RADON_WAVES * SS_WAVE) void k_radon(SSDevPlan p, const float2* __restrict__ gA,
                                               const float2* __restrict__ gB, int j0, int j1, int nslab,
                                               int slab_rows, float* __restrict__ out, float* __restrict__ part,
                                               SSEmRatioArgs em) {
  __shared__ SSRadonLds<SSRadonShape<SH>::CAP> lds;
  const int z = (int)(blockIdx.z % (unsigned)nslab);
  const size_t slice = blockIdx.z / (unsigned)nslab;
  const size_t lay = slice * p.g.n * p.g.pitch;
  const size_t so = slice * p.g.nangles * p.g.nrays;
  double* slot = nullptr;
  if (MODE == SS_OUT_EM_RATIO) slot = em.ll_partials + (slice * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x;
  float* po = MODE == SS_OUT_PARTIAL ? part + (slice * nslab + z) * (size_t)(j1 - j0) * p.g.nrays : nullptr;
  ss_radon_item<MODE, SH>(ss_radon_view(p), gA + lay, gB + lay, j0, j1, blockIdx.x, blockIdx.y, z * slab_rows,
                      min(p.g.n, (z + 1) * slab_rows), out + so, MODE == SS_OUT_EM_RATIO ? em.meas + so : nullptr,
                      slot, po, lds);
}

// Sums the slabs' partial sums of rays [j0, j1) x [0, R) of each slice and stores them
// as k_radon does (delta applied; in EM mode the ratio, and one FP64 log-likelihood
// partial per block at ll_partials[slice * gridDim.x + blockIdx.x]).
template <int MODE>
__global__ __launch_bounds__(256) void k_radon_slab_finish(SSDevPlan p, const float* __restrict__ part, int nslab,
                                                           int j0, int j1, float* __restrict__ out, SSEmRatioArgs em) {
  __shared__ double red[4];
  const size_t slice = blockIdx.y;
  const int R = p.g.nrays;
  const size_t per = (size_t)(j1 - j0) * R;
  const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t so = slice * p.g.nangles * (size_t)R;
  double ll = 0.0;
  if (idx < per) {
    const int jj = (int)(idx / (unsigned)R), k = (int)(idx - (size_t)jj * R);
    const float* src = part + slice * nslab * per + idx;
    float sum = 0.f;
    sum = ss_slab_sum(src, per, nslab);
    ss_radon_store<MODE>(ss_radon_view(p), j0 + jj, k, sum * p.angles[j0 + jj].delta, out + so,
                         MODE == SS_OUT_EM_RATIO ? em.meas + so : nullptr, &ll);
  }
  if (MODE == SS_OUT_EM_RATIO) {
    const double tot = ss_block_sum_f64(ll, red);
    if (threadIdx.x == 0) em.ll_partials[slice * gridDim.x + blockIdx.x] = tot;
  }
}

// Irregular rays (knife-edge angles): one block per ray, samples across threads,
// validity evaluated per sample with the exact FP64 predicate.  `idx` indexes the
// plan's irregular list; rays outside [j0, j1) write a zero partial only.
template <int MODE>
__device__ __forceinline__ void ss_radon_fixup_item(const SSDevPlan& p, const float2* __restrict__ gA,
                                                    const float2* __restrict__ gB, int j0, int j1, int idx,
                                                    float* __restrict__ out, const float* __restrict__ meas,
                                                    double* ll_slot, float* redf) {
  const int id = p.irregular[idx];
  const int R = p.g.nrays, n = p.g.n, pitch = p.g.pitch;
  const int j = id / R, k = id % R;
  if (j < j0 || j >= j1) {
    if (MODE == SS_OUT_EM_RATIO && threadIdx.x == 0) *ll_slot = 0.0;
    return;
  }
  const SSAngle& a = p.angles[j];
  const float2* img = (a.branch_a ? gA : gB) + 1;
  const double t = p.ray_t ? p.ray_t[k] : ss_ray_t(p.g, k);
  const int64_t a_k = p.ray_alpha ? p.ray_alpha[(size_t)j * R + k] : a.alpha0 + (int64_t)k * a.gamma;
  float acc = 0.f;
  for (int s = threadIdx.x; s < n; s += blockDim.x) {
    if (ss_radon_sample_valid_t(p.g, a, t, s)) {
      const int64_t u = a_k + (int64_t)s * a.beta;
      const int m = (int)(u >> 32);
      const float w = (float)(uint32_t)u * SS_INV_FIX;
      const float2 rec = img[(size_t)s * pitch + m];
      acc += fmaf(w, rec.y, rec.x);
    }
  }
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off);
  if ((threadIdx.x & 63) == 0) redf[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    float tot = 0.f;
    for (int i = 0; i < (int)(blockDim.x >> 6); ++i) tot += redf[i];
    double ll = 0.0;
    ss_radon_store<MODE>(ss_radon_view(p), j, k, tot * a.delta, out, meas, &ll);
    if (MODE == SS_OUT_EM_RATIO) *ll_slot = ll;
  }
}

template <int MODE>
__global__ __launch_bounds__(256) void k_radon_fixup(SSDevPlan p, const float2* __restrict__ gA,
                                                     const float2* __restrict__ gB, int j0, int j1,
                                                     float* __restrict__ out, SSEmRatioArgs em, int partial_base) {
  __shared__ float redf[4];
  const size_t slice = blockIdx.y;
  const size_t lay = slice * p.g.n * p.g.pitch;
  const size_t so = slice * p.g.nangles * p.g.nrays;
  double* slot = nullptr;
  if (MODE == SS_OUT_EM_RATIO) slot = em.ll_partials + partial_base + slice * gridDim.x + blockIdx.x;
  ss_radon_fixup_item<MODE>(p, gA + lay, gB + lay, j0, j1, blockIdx.x, out + so,
                            MODE == SS_OUT_EM_RATIO ? em.meas + so : nullptr, slot, redf);
}

// A launch of few blocks (a few angles: one OSEM subset) divides the sample rows of every
// (angle block, ray group) into slabs handled by separate blocks, so that every CU gets
// blocks while each block keeps the LDS-staged path; a finishing pass then adds the
// slabs' partial sums.  Rows per slab are a multiple of RADON_SLAB_ALIGN, and the divided
// launch stays within SS_RADON_SLAB_BLOCKS blocks (4 per CU, one round: a second round
// for a few blocks would double the launch time).
int ss_radon_slab_plan(int n, long long blocks, int* rows) {
  *rows = n;
  if (blocks >= SS_RADON_SLAB_BLOCKS) return 1;
  const long long want = SS_RADON_SLAB_BLOCKS / std::max(blocks, 1LL);
  if (want <= 1) return 1;
  int L = (int)((n + want - 1) / want);
  L = (L + RADON_SLAB_ALIGN - 1) / RADON_SLAB_ALIGN * RADON_SLAB_ALIGN;
  const int count = (n + L - 1) / L;
  if (count <= 1) return 1;
  *rows = L;
  return count;
}

int ss_radon_slabs(const SSDevPlan& p, int j0, int j1, int nslices, int* rows) {
  const int n = p.g.n;
  *rows = n;
  if (j1 <= j0) return 1;
  if (p.radon_slab_rows > 0) {
    const int L = (std::min(p.radon_slab_rows, n) + RADON_SLAB_ALIGN - 1) / RADON_SLAB_ALIGN * RADON_SLAB_ALIGN;
    const int count = (n + L - 1) / L;
    if (count <= 1) return 1;
    *rows = L;
    return count;
  }
  return ss_radon_slab_plan(n, ss_radon_blocks(p.nwaves, j0, j1) * nslices, rows);
}

static int ss_radon_finish_blocks(int nj, int R) { return (int)(((long long)nj * R + 255) / 256); }

size_t ss_radon_scratch(const SSDevPlan& p, int j0, int j1, int nslices) {
  int rows;
  const int nslab = ss_radon_slabs(p, j0, j1, nslices, &rows);
  return nslab > 1 ? sizeof(float) * (size_t)nslab * nslices * (size_t)(j1 - j0) * p.g.nrays : 0;
}

int ss_radon_partials(const SSDevPlan& p, int j0, int j1, int nslices) {
  int rows;
  const int nslab = ss_radon_slabs(p, j0, j1, nslices, &rows);
  const int main = nslab > 1 ? ss_radon_finish_blocks(j1 - j0, p.g.nrays)
                             : p.nwaves * ((j1 - j0 + RADON_WAVES - 1) / RADON_WAVES);
  return (main + p.nirregular) * nslices;
}

hipError_t ss_launch_radon(const SSDevPlan& p, const float2* gatherA, const float2* gatherB, int j0, int j1,
                           int nslices, int mode, float* out, SSEmRatioArgs em, float* scratch, hipStream_t stream) {
  if (j1 <= j0) return hipSuccess;
  int rows;
  const int nslab = ss_radon_slabs(p, j0, j1, nslices, &rows);
  if (nslab > 1 && !scratch) return hipErrorInvalidValue;
  dim3 grid(p.nwaves, (j1 - j0 + RADON_WAVES - 1) / RADON_WAVES, nslices * nslab);
  dim3 block(RADON_WAVES * SS_WAVE);
  int partial_base;
  // the main pass in the plan's block shape (0 or 1)
#define SS_RADON_LAUNCH(MODE, OUTP, PARTP, NSLAB, ROWS, EM)                                                        \
  do {                                                                                                            \
    if (p.radon_shape == 1)                                                                                       \
      hipLaunchKernelGGL((k_radon<MODE, 1>), grid, block, 0, stream, p, gatherA, gatherB, j0, j1, NSLAB, ROWS,    \
                         OUTP, PARTP, EM);                                                                        \
    else                                                                                                          \
      hipLaunchKernelGGL((k_radon<MODE, 0>), grid, block, 0, stream, p, gatherA, gatherB, j0, j1, NSLAB, ROWS,    \
                         OUTP, PARTP, EM);                                                                        \
  } while (0)
  if (nslab > 1) {
    const SSEmRatioArgs none{nullptr, nullptr};
    SS_RADON_LAUNCH(SS_OUT_PARTIAL, nullptr, scratch, nslab, rows, none);
    const int fb = ss_radon_finish_blocks(j1 - j0, p.g.nrays);
    dim3 fin(fb, nslices);
    switch (mode) {
      case SS_OUT_SINO:
        hipLaunchKernelGGL(k_radon_slab_finish<SS_OUT_SINO>, fin, dim3(256), 0, stream, p, scratch, nslab, j0, j1, out,
                           em);
        break;
      case SS_OUT_SINOT:
        hipLaunchKernelGGL(k_radon_slab_finish<SS_OUT_SINOT>, fin, dim3(256), 0, stream, p, scratch, nslab, j0, j1,
                           out, em);
        break;
      default:
        hipLaunchKernelGGL(k_radon_slab_finish<SS_OUT_EM_RATIO>, fin, dim3(256), 0, stream, p, scratch, nslab, j0, j1,
                           out, em);
        break;
    }
    partial_base = fb * nslices;
  } else {
    switch (mode) {
      case SS_OUT_SINO:
        SS_RADON_LAUNCH(SS_OUT_SINO, out, nullptr, 1, p.g.n, em);
        break;
      case SS_OUT_SINOT:
        SS_RADON_LAUNCH(SS_OUT_SINOT, out, nullptr, 1, p.g.n, em);
        break;
      default:
        SS_RADON_LAUNCH(SS_OUT_EM_RATIO, out, nullptr, 1, p.g.n, em);
        break;
    }
#undef SS_RADON_LAUNCH
    partial_base = (int)(grid.x * grid.y) * nslices;
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess || p.nirregular == 0) return e;
  dim3 fgrid(p.nirregular, nslices);
  switch (mode) {
    case SS_OUT_SINO:
      hipLaunchKernelGGL(k_radon_fixup<SS_OUT_SINO>, fgrid, dim3(256), 0, stream, p, gatherA, gatherB, j0, j1, out,
                         em, partial_base);
      break;
    case SS_OUT_SINOT:
      hipLaunchKernelGGL(k_radon_fixup<SS_OUT_SINOT>, fgrid, dim3(256), 0, stream, p, gatherA, gatherB, j0, j1, out,
                         em, partial_base);
      break;
    default:
      hipLaunchKernelGGL(k_radon_fixup<SS_OUT_EM_RATIO>, fgrid, dim3(256), 0, stream, p, gatherA, gatherB, j0, j1,
                         out, em, partial_base);
      break;
  }
  return hipGetLastError();
}

// =====================================================================================
// backprojection: pixel-driven, 32 x (8*ROWS) pixel tile per block, angle blocks
// staged in LDS
//
// For each angle of a stage the block computes (FP64) the ray coordinate U of the
// tile origin and the tile's U range, then stages the W consecutive rays
// [base, base+W) that the tile can touch.  Each thread owns ROWS pixels (one column,
// rows ty, ty+8, ...) and interpolates from LDS with FP32 offsets relative to `base`
// (|u| < ~W, so FP32 resolves ~1e-5 of a detector bin).
//
// Exactness of the 0 <= m <= R-2 predicate: a tile whose U range lies inside
// [eps, R-1-eps] needs no test; one entirely outside contributes nothing; the rest
// test each pixel, and pixels within 1e-3 bins of an edge re-evaluate the
// reference's FP64 expression.
// =====================================================================================

#define BP_TILE_W 32

struct BPAngle {
  float f0, b, c;
  int flag;  // 0 outside, 1 interior, 2 boundary
  int base;
  int pad0, pad1, pad2;
};

__device__ __forceinline__ int ss_bp_exact_valid(const SSDevPlan& p, int j, int row, int col) {
  const SSAngle& a = p.angles[j];
  return ss_bp_sample_valid(p.g, a.cost, a.sint, row, col) ? 1 : 0;
}

template <int STAGE>
struct SSBpLds {
  float2 win[STAGE];  // per staged ray w: (value - w * difference, difference to the next ray)
  BPAngle par[64];
};

// One pixel tile [c0, c0+32) x [r0, r0+8*ROWS) over angles [j0, j1).  sl points at this
// slice's angle-major sinogram; out (plain) or x/weight (EM update) at its image.
template <int MODE, int ROWS, int STAGE>
__device__ __forceinline__ void ss_bp_item(const SSDevPlan& p, const float* __restrict__ sl, int j0, int j1, int c0,
                                           int r0, float* __restrict__ out, float* __restrict__ xs,
                                           const float* __restrict__ ws, SSBpLds<STAGE>& lds) {
  const int n = p.g.n, R = p.g.nrays;
  const int W = p.bp_window;
  const int J = p.bp_angles;
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int tx = tid & 31, ty = tid >> 5;
  const int col = c0 + tx;
  constexpr int TH = 8 * ROWS;
  float acc[ROWS];
  float fty[ROWS];
#pragma unroll
  for (int i = 0; i < ROWS; ++i) {
    acc[i] = 0.f;
    fty[i] = (float)(ty + 8 * i);
  }
  const float ftx = (float)tx;
  const float lo_rel_eps = 1e-3f;

  for (int jb = j0; jb < j1; jb += J) {
    const int nj = min(J, j1 - jb);
    // every wave: lane l holds the parameters of angle jb + l (J <= 64)
    BPAngle q{0.f, 0.f, 0.f, 0, 0, 0, 0, 0};
    if (lane < nj) {
      const SSAngle& a = p.angles[jb + lane];
      const double U = a.bp_a + (double)c0 * a.bp_b + (double)r0 * a.bp_c;
      const double ec = (double)(BP_TILE_W - 1) * a.bp_b, er = (double)(TH - 1) * a.bp_c;
      const double umin = U + fmin(0.0, ec) + fmin(0.0, er);
      const double umax = U + fmax(0.0, ec) + fmax(0.0, er);
      const double base = floor(umin) - 1.0;
      q.f0 = (float)(U - base);
      q.b = (float)a.bp_b;
      q.c = (float)a.bp_c;
      const double eps = 1e-6;
      if (umax < -eps || umin >= (double)(R - 1) + eps)
        q.flag = 0;
      else if (umin >= eps && umax < (double)(R - 1) - eps)
        q.flag = 1;
      else
        q.flag = 2;
      q.base = (int)base;
    }
    // a stage whose angles are all interior for this tile runs without predicates
    const bool mixed = __ballot(lane < nj && q.flag != 1) != 0ull;
    __syncthreads();  // the previous stage (or item) is done with win / par
    if (wv == 0) {  // the used fields only (whole-struct copies kept the padding in scratch)
      lds.par[lane].f0 = q.f0;
      lds.par[lane].b = q.b;
      lds.par[lane].c = q.c;
      lds.par[lane].flag = q.flag;
      lds.par[lane].base = q.base;
    }
    // staging: wave wv fills the windows of angles wv, wv+4, ...; base is at least one
    // ray below the tile's lowest U, so every pixel has u >= 1 and (int)u < W - 1
    for (int jj = wv; jj < nj; jj += 4) {
      const int fl = __builtin_amdgcn_readlane(q.flag, jj);
      const int bs = __builtin_amdgcn_readlane(q.base, jj);
      const float* row = sl + (size_t)(jb + jj) * R;
      for (int w = lane; w < W; w += 64) {
        const int ray = bs + w;
        float v0 = 0.f, v1 = 0.f;
        if (fl != 0) {
          if ((unsigned)ray < (unsigned)R) v0 = row[ray];
          if ((unsigned)(ray + 1) < (unsigned)R) v1 = row[ray + 1];
        }
        // (v0 - w d, d): the value at u in [w, w + 1) is then fma(u, d, first), so the
        // interpolation needs no fraction of u
        const float d = v1 - v0;
        lds.win[jj * W + w] = make_float2(fmaf(-(float)w, d, v0), d);
      }
    }
    __syncthreads();
    if (!mixed) {
#pragma unroll 2
      for (int jj = 0; jj < nj; ++jj) {
        const float f0 = lds.par[jj].f0, pb = lds.par[jj].b, pc = lds.par[jj].c;
        const float2* wrow = lds.win + jj * W;
        const float t0 = fmaf(ftx, pb, f0);
        // u >= 1 here, so truncation is the floor
#pragma unroll
        for (int i = 0; i < ROWS; ++i) {
          const float u = fmaf(fty[i], pc, t0);
          const float2 t = wrow[(int)u];
          acc[i] += fmaf(u, t.y, t.x);
        }
      }
    } else {
      for (int jj = 0; jj < nj; ++jj) {
        BPAngle qa;
        qa.f0 = lds.par[jj].f0;
        qa.b = lds.par[jj].b;
        qa.c = lds.par[jj].c;
        qa.flag = lds.par[jj].flag;
        qa.base = lds.par[jj].base;
        const int flag = __builtin_amdgcn_readfirstlane(qa.flag);
        if (flag == 0) continue;
        const float2* wrow = lds.win + jj * W;
        const float t0 = fmaf(ftx, qa.b, qa.f0);
        float us[ROWS], vals[ROWS];
#pragma unroll
        for (int i = 0; i < ROWS; ++i) {
          us[i] = fmaf(fty[i], qa.c, t0);
          const float2 t = wrow[(int)us[i]];
          vals[i] = fmaf(us[i], t.y, t.x);
        }
        if (flag == 2) {
          // boundary tile: per-pixel 0 <= m <= R-2 test; pixels within 1e-3 bins of an
          // edge take the reference's FP64 expression.  The result is an integer mask
          // applied by select, keeping the divergent region free of value phis.
          const float lo_rel = (float)(-qa.base), hi_rel = (float)(R - 1 - qa.base);
          int keep = 0;
#pragma unroll
          for (int i = 0; i < ROWS; ++i) {
            const float u = us[i];
            const bool inside = (u > lo_rel + lo_rel_eps) && (u < hi_rel - lo_rel_eps);
            const bool outside = (u < lo_rel - lo_rel_eps) || (u > hi_rel + lo_rel_eps);
            int ok = inside ? 1 : 0;
            if (!inside && !outside) ok = ss_bp_exact_valid(p, jb + jj, r0 + ty + 8 * i, col);
            keep |= ok << i;
          }
#pragma unroll
          for (int i = 0; i < ROWS; ++i) vals[i] = ((keep >> i) & 1) ? vals[i] : 0.f;
        }
#pragma unroll
        for (int i = 0; i < ROWS; ++i) acc[i] += vals[i];
      }
    }
  }
  if (col >= n) return;
  const float dth = (float)p.g.dth;
#pragma unroll
  for (int i = 0; i < ROWS; ++i) {
    const int row = r0 + ty + 8 * i;
    if (row >= n) continue;
    const size_t idx = (size_t)row * n + col;
    const float bp = acc[i] * dth;
    if (MODE == SS_BP_PLAIN) {
      out[idx] = bp;
    } else if (MODE == SS_BP_PARTIAL) {
      out[idx] = acc[i];
    } else {
      const float wgt = ws[idx];
      const float x = xs[idx];
      xs[idx] = wgt > 0.f ? x * bp / wgt : x;
    }
  }
}

// blockIdx.z = slice * ngroups + g.  SS_BP_PARTIAL: group g takes angles
// [j0 + g*nj/ngroups, j0 + (g+1)*nj/ngroups) and writes its unscaled sums to
// out[slice][g][n][n]; the other modes have ngroups = 1.
template <int MODE, int ROWS, int STAGE>
__global__ __launch_bounds__(256) void k_backprojection(SSDevPlan p, const float* __restrict__ sinoT, int j0, int j1,
                                                        int ngroups, float* __restrict__ out, SSEmUpdateArgs em) {
  __shared__ SSBpLds<STAGE> lds;
  const size_t slice = blockIdx.z / (unsigned)ngroups;
  const int g = (int)(blockIdx.z % (unsigned)ngroups);
  const size_t so = slice * p.g.nangles * p.g.nrays;
  const size_t per = (size_t)p.g.n * p.g.n;
  const size_t io = slice * per;
  int ja = j0, jb = j1;
  if (MODE == SS_BP_PARTIAL) {
    const long long nj = j1 - j0;
    ja = j0 + (int)(g * nj / ngroups);
    jb = j0 + (int)((g + 1) * nj / ngroups);
  }
  float* dst = MODE == SS_BP_PLAIN ? out + io : (MODE == SS_BP_PARTIAL ? out + (slice * ngroups + g) * per : nullptr);
  ss_bp_item<MODE, ROWS, STAGE>(p, sinoT + so, ja, jb, blockIdx.x * BP_TILE_W, blockIdx.y * 8 * ROWS, dst,
                                MODE == SS_BP_EM_UPDATE ? em.x + io : nullptr,
                                MODE == SS_BP_EM_UPDATE ? em.weight + io : nullptr, lds);
}

// the groups' partial sums of each pixel added in group order, times dtheta, stored
// (SS_BP_PLAIN) or applied as x <- x * bp / w (SS_BP_EM_UPDATE), as ss_bp_item stores them
template <int MODE>
__global__ __launch_bounds__(256) void k_bp_groups_finish(const float* __restrict__ part, int ngroups, int n, float dth,
                                                          float* __restrict__ out, SSEmUpdateArgs em) {
  const size_t per = (size_t)n * n;
  const size_t idx = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (idx >= per) return;
  const size_t slice = blockIdx.y;
  const float* src = part + slice * ngroups * per + idx;
  float sum = 0.f;
  for (int g = 0; g < ngroups; ++g) sum += src[(size_t)g * per];
  const float bp = sum * dth;
  const size_t o = slice * per + idx;
  if (MODE == SS_BP_PLAIN) {
    out[o] = bp;
  } else {
    const float wgt = em.weight[o];
    const float x = em.x[o];
    em.x[o] = wgt > 0.f ? x * bp / wgt : x;
  }
}

int ss_bp_tiles(int n, int variant) {
  const int th = 8 * kSSBpRows[variant];
  return ((n + BP_TILE_W - 1) / BP_TILE_W) * ((n + th - 1) / th);
}

int ss_bp_groups(const SSDevPlan& p, int j0, int j1, int nslices) {
  const long long tiles = (long long)ss_bp_tiles(p.g.n, p.bp_variant) * std::max(nslices, 1);
  const int nj = j1 - j0;
  if (tiles >= SS_BP_SPLIT_TILES || nj < 2 * SS_BP_GROUP_ANGLES) return 1;
  const long long want = (SS_BP_SPLIT_BLOCKS + tiles - 1) / tiles;
  return (int)std::max(1LL, std::min(want, (long long)(nj / SS_BP_GROUP_ANGLES)));
}

size_t ss_bp_scratch(const SSDevPlan& p, int j0, int j1, int nslices) {
  const int G = ss_bp_groups(p, j0, j1, nslices);
  return G > 1 ? sizeof(float) * (size_t)G * nslices * p.g.n * p.g.n : 0;
}

// variant id -> (ROWS, STAGE); must match kSSBpRows / kSSBpStage
#define SS_BP_FOR_VARIANTS(X) X(0, 4, 4096) X(1, 2, 2048) X(2, 8, 4096) X(3, 4, 2048) X(4, 2, 4096) X(5, 8, 2048)

hipError_t ss_launch_backprojection(const SSDevPlan& p, const float* sinoT, int j0, int j1, int nslices, int mode,
                                    float* out, SSEmUpdateArgs em, float* scratch, hipStream_t stream) {
  if (j1 <= j0 && mode == SS_BP_EM_UPDATE) return hipSuccess;
  const int n = p.g.n;
  const int th = 8 * kSSBpRows[p.bp_variant];
  const int G = scratch ? ss_bp_groups(p, j0, j1, nslices) : 1;
  const int kmode = G > 1 ? SS_BP_PARTIAL : mode;
  float* kout = G > 1 ? scratch : out;
  dim3 grid((n + BP_TILE_W - 1) / BP_TILE_W, (n + th - 1) / th, nslices * G);
#define SS_BP_LAUNCH(V, ROWS, STAGE)                                                                               \
  case V:                                                                                                          \
    if (kmode == SS_BP_PLAIN)                                                                                      \
      hipLaunchKernelGGL((k_backprojection<SS_BP_PLAIN, ROWS, STAGE>), grid, dim3(256), 0, stream, p, sinoT, j0,  \
                         j1, 1, kout, em);                                                                         \
    else if (kmode == SS_BP_PARTIAL)                                                                               \
      hipLaunchKernelGGL((k_backprojection<SS_BP_PARTIAL, ROWS, STAGE>), grid, dim3(256), 0, stream, p, sinoT, j0, \
                         j1, G, kout, em);                                                                         \
    else                                                                                                           \
      hipLaunchKernelGGL((k_backprojection<SS_BP_EM_UPDATE, ROWS, STAGE>), grid, dim3(256), 0, stream, p, sinoT,  \
                         j0, j1, 1, kout, em);                                                                     \
    break;
  switch (p.bp_variant) {
    SS_BP_FOR_VARIANTS(SS_BP_LAUNCH)
    default:
      return hipErrorInvalidValue;
  }
#undef SS_BP_LAUNCH
  hipError_t e = hipGetLastError();
  if (e != hipSuccess || G == 1) return e;
  const size_t per = (size_t)n * n;
  dim3 fin((unsigned)((per + 255) / 256), nslices);
  if (mode == SS_BP_PLAIN)
    hipLaunchKernelGGL(k_bp_groups_finish<SS_BP_PLAIN>, fin, dim3(256), 0, stream, scratch, G, n, (float)p.g.dth, out,
                       em);
  else
    hipLaunchKernelGGL(k_bp_groups_finish<SS_BP_EM_UPDATE>, fin, dim3(256), 0, stream, scratch, G, n,
                       (float)p.g.dth, out, em);
  return hipGetLastError();
}

// =====================================================================================
// exact transpose of the Radon: image lines accumulated in LDS from the forward's own
// sample map
//
// The forward adds (1-w)*F[m] + w*F[m+1] for every valid sample (ray k, sample s), with
// m and w from the Q32.32 trajectory u = alpha0 + k*gamma + s*beta.  Its transpose
// gives pixel m the amount (1-w)*delta*q_k and pixel m+1 the amount w*delta*q_k.
//
// k_radon_transpose_lines: one wave owns one image line along the forward's sample
// axis -- image row s for branch-B angles, image column s for branch-A angles -- as
// an LDS row of n+2 cells (pixel c at index c+1, the two pads absorb m = -1 and
// m+1 = n).  For each angle of its branch, the lanes take the rays that can cross the
// line 64 at a time, step u by 64*gamma between passes, check the sample's validity
// and add the two contributions.  Each line belongs to one wave, so no global atomics
// or block barriers are needed; the LDS holds only the lines, so up to 16384-pixel
// lines fit and many waves share a CU.  The rays' exact ranges and sinogram values are
// read from global memory (the angle's rows are shared by every line, so they stay in
// cache).  Branch-B lines are image rows and are written to the output; branch-A lines
// are image columns and go to a transposed scratch image that k_transpose_add folds
// in.  Irregular rays are added by the fixup kernel afterwards.
// =====================================================================================

#define RT_LINES 8         // image lines (waves) per block
#define RT_PHASES_MAX 8    // dense rays: lane phases before falling back to LDS atomics

// LDS of k_radon_transpose_lines: one line of n+2 cells per wave, plus two spare cells
// that lanes without a sample update with zero
__host__ __device__ inline int ss_rt_line_stride(int n) { return n + 4; }
__host__ __device__ inline size_t ss_rt_line_lds(int n, int nl) {
  return sizeof(float) * (size_t)nl * ss_rt_line_stride(n);
}

// The angle table read through the constant address space: the plan's angles are not
// written while a kernel runs, so the reads are scalar loads into wave-uniform registers.
typedef __attribute__((address_space(4))) const SSAngle* ss_const_angles;

// next angle >= j of `branch` (or j1)
__device__ __forceinline__ int ss_rt_next(ss_const_angles angles, int j, int j1, int branch) {
  while (j < j1 && angles[j].branch_a != branch) ++j;
  return j;
}

// structured buffer loads of one ray's record (address = base + vindex * stride)
typedef int ss_i32x2 __attribute__((ext_vector_type(2)));
__device__ ss_i32x2 ss_struct_load_i32x2(__amdgpu_buffer_rsrc_t rsrc, int vindex, int voffset, int soffset,
                                         int aux) __asm("llvm.amdgcn.struct.ptr.buffer.load.v2i32");
__device__ float ss_struct_load_f32(__amdgpu_buffer_rsrc_t rsrc, int vindex, int voffset, int soffset,
                                    int aux) __asm("llvm.amdgcn.struct.ptr.buffer.load.f32");

// How the lanes of a pass add their two terms to the LDS line:
//   RT_DISTINCT  rays at least one cell apart (|gamma| >= 1): the lanes hit distinct cells,
//                so plain read-modify-writes, first cell m, then cell m+1, in program order
//                (the empty asm statements keep the compiler from moving an update across
//                the previous one); lanes without a sample add zero to the spare cells;
//   RT_PHASED    denser rays: the same, one phase of lanes `phases` apart at a time;
//   RT_ATOMIC    rays denser than 1/RT_PHASES_MAX cell: LDS atomics.
// In the first two every cell takes its terms in a fixed order (deterministic sums).
enum { RT_DISTINCT = 0, RT_PHASED = 1, RT_ATOMIC = 2 };

// A wave owns image line s of SL slices (SL = 2 shares the geometry work of a pass
// between two slices of a batch; every slice has its own LDS line and sinogram row).
#define RT_SL_MAX 4  // slices per wave, at most
struct SSRtLine {
  float* line[RT_SL_MAX];    // LDS lines: cell c + 1 holds pixel c; cells n+2, n+3 are spare
  int n, s;                  // line length, line index (the forward's sample)
  int k_lo, k_hi;            // candidate rays
  bool all_edge;             // every sample reads its ray's exact range
  int p_first, p_last;       // passes p_first .. p_last - 1 hold only valid samples
  int phases, my_phase;      // RT_PHASED
  __amdgpu_buffer_rsrc_t rb;     // the angle's range row (int2)
  __amdgpu_buffer_rsrc_t rq[RT_SL_MAX];  // the angle's sinogram row (float) of each slice
  int64_t gamma;             // trajectory step per ray
  int64_t u0, step;          // trajectory (+1 cell) of ray k_lo, and its step per pass
  float delta;
};

// RT_PHASED / RT_ATOMIC update of one pass on one line: trajectory u, term qd = q * delta
// (0 if the sample is not valid)
template <int MODE>
__device__ __forceinline__ void ss_rt_update(const SSRtLine& L, float* line, bool valid, int64_t u, float qd) {
  const int idx = (int)(u >> 32);
  const float wq = (float)(uint32_t)u * SS_INV_FIX * qd;
  if (MODE == RT_PHASED) {
    float* cell = line + idx;
    for (int ph = 0; ph < L.phases; ++ph) {
      const bool mine = valid && L.my_phase == ph;
      if (mine) cell[0] += qd - wq;
      __asm__ volatile("" ::: "memory");
      if (mine) cell[1] += wq;
      __asm__ volatile("" ::: "memory");
    }
  } else if (valid) {
    atomicAdd(line + idx, qd - wq);
    atomicAdd(line + idx + 1, wq);
  }
}

// One iteration over two passes (pass, pass + 1) of the candidate rays: lane 0's ray kb and
// trajectory ub (+1 cell), the passes' sinogram values q0[], q1[] per slice; loads the
// values of the next iteration's passes into n0[], n1[] (indices clamped to the last
// candidate: the loads past the last pass stay in bounds and go unused).
//
// Passes p_first .. p_last - 1 hold only valid samples (see the kernel); a pair of them
// takes a path without range checks.  Other samples near the line's ends (every sample
// of other angles) read the ray's exact range.
template <int MODE, int SL>
__device__ __forceinline__ void ss_rt_pair(const SSRtLine& L, int lane, int pass, int npass, int kb, int64_t ub,
                                           int64_t lane_u, const float* q0, const float* q1, float* n0, float* n1) {
  const int n = L.n, s = L.s, k_hi = L.k_hi, last = k_hi - 1;
  const int64_t ub1 = ub + L.step;
  const int64_t u = ub + lane_u, u1 = ub1 + lane_u;
  const int idx0 = (int)(u >> 32), idx1 = (int)(u1 >> 32);  // m + 1, in [0, n] if valid
  const int k = kb + lane;
  if (MODE == RT_DISTINCT && pass >= L.p_first && pass + 1 < L.p_last) {
    // every sample of both passes is valid
#pragma unroll
    for (int i = 0; i < SL; ++i) {
      n0[i] = ss_struct_load_f32(L.rq[i], min(k + 2 * SS_WAVE, last), 0, 0, 0);
      n1[i] = ss_struct_load_f32(L.rq[i], min(k + 3 * SS_WAVE, last), 0, 0, 0);
    }
    const float ds = L.delta * SS_INV_FIX;
    const float f0 = (float)(uint32_t)u, f1 = (float)(uint32_t)u1;
    float w0[SL], w1[SL], x0[SL], x1[SL];
    float *c0[SL], *c1[SL];
#pragma unroll
    for (int i = 0; i < SL; ++i) {
      w0[i] = f0 * (q0[i] * ds);
      w1[i] = f1 * (q1[i] * ds);
      c0[i] = L.line[i] + idx0;
      c1[i] = L.line[i] + idx1;
      x0[i] = c0[i][0];
      x1[i] = c1[i][0];
    }
#pragma unroll
    for (int i = 0; i < SL; ++i) {
      c0[i][0] = fmaf(q0[i], L.delta, x0[i]) - w0[i];
      c1[i][0] = fmaf(q1[i], L.delta, x1[i]) - w1[i];
    }
    __asm__ volatile("" ::: "memory");
#pragma unroll
    for (int i = 0; i < SL; ++i) {
      x0[i] = c0[i][1];
      x1[i] = c1[i][1];
    }
#pragma unroll
    for (int i = 0; i < SL; ++i) {
      c0[i][1] = x0[i] + w0[i];
      c1[i][1] = x1[i] + w1[i];
    }
    __asm__ volatile("" ::: "memory");
    return;
  }
  const bool two = pass + 1 < npass;  // a second pass this iteration (uniform)
  const bool e0 = L.all_edge || (unsigned)(idx0 - 2) > (unsigned)(n - 4);
  const bool e1 = two && (L.all_edge || (unsigned)(idx1 - 2) > (unsigned)(n - 4));
  ss_i32x2 b0 = {0, -1}, b1 = {0, -1};
  if (e0) b0 = ss_struct_load_i32x2(L.rb, min(k, last), 0, 0, 0);
  if (e1) b1 = ss_struct_load_i32x2(L.rb, min(k + SS_WAVE, last), 0, 0, 0);
  // issued after the ranges, so that waiting for the ranges does not wait for these
#pragma unroll
  for (int i = 0; i < SL; ++i) {
    n0[i] = ss_struct_load_f32(L.rq[i], min(k + 2 * SS_WAVE, last), 0, 0, 0);
    n1[i] = ss_struct_load_f32(L.rq[i], min(k + 3 * SS_WAVE, last), 0, 0, 0);
  }
  bool v0 = k < k_hi, v1 = k + SS_WAVE < k_hi;
  if (e0) v0 = v0 && s >= b0.x && s <= b0.y;
  if (e1) v1 = v1 && s >= b1.x && s <= b1.y;
  if (MODE == RT_DISTINCT) {
    // the first cells of both passes are distinct, and so are their second cells: read
    // both, then write both, one cell phase at a time; lanes without a sample add zero to
    // the spare cells
    const float f0 = (float)(uint32_t)u * SS_INV_FIX, f1 = (float)(uint32_t)u1 * SS_INV_FIX;
    const int i0 = v0 ? idx0 : n + 2, i1 = v1 ? idx1 : n + 2;
    float qd0[SL], qd1[SL], w0[SL], w1[SL], x0[SL], x1[SL];
    float *c0[SL], *c1[SL];
#pragma unroll
    for (int i = 0; i < SL; ++i) {
      qd0[i] = v0 ? q0[i] * L.delta : 0.f;
      qd1[i] = v1 ? q1[i] * L.delta : 0.f;
      w0[i] = f0 * qd0[i];
      w1[i] = f1 * qd1[i];
      c0[i] = L.line[i] + i0;
      c1[i] = L.line[i] + i1;
      x0[i] = c0[i][0];
      x1[i] = c1[i][0];
    }
#pragma unroll
    for (int i = 0; i < SL; ++i) {
      c0[i][0] = x0[i] + (qd0[i] - w0[i]);
      c1[i][0] = x1[i] + (qd1[i] - w1[i]);
    }
    __asm__ volatile("" ::: "memory");
#pragma unroll
    for (int i = 0; i < SL; ++i) {
      x0[i] = c0[i][1];
      x1[i] = c1[i][1];
    }
#pragma unroll
    for (int i = 0; i < SL; ++i) {
      c0[i][1] = x0[i] + w0[i];
      c1[i][1] = x1[i] + w1[i];
    }
    __asm__ volatile("" ::: "memory");
  } else {
#pragma unroll
    for (int i = 0; i < SL; ++i) {
      ss_rt_update<MODE>(L, L.line[i], v0, u, v0 ? q0[i] * L.delta : 0.f);
      if (two) ss_rt_update<MODE>(L, L.line[i], v1, u1, v1 ? q1[i] * L.delta : 0.f);
    }
  }
}

// the passes over the candidate rays, two per ss_rt_pair; the sinogram values alternate
// between two register sets (each half loads the other half's values)
template <int MODE, int SL>
__device__ __forceinline__ void ss_rt_passes(const SSRtLine& L, int lane) {
  const int npass = (L.k_hi - L.k_lo + SS_WAVE - 1) / SS_WAVE;
  const int last = L.k_hi - 1;
  const int64_t lane_u = (int64_t)lane * L.gamma;
  int64_t ub = L.u0;  // lane 0's trajectory (+1 cell)
  int kb = L.k_lo;    // lane 0's ray
  float qa[SL], qb[SL], qc[SL], qd[SL];
#pragma unroll
  for (int i = 0; i < SL; ++i) {
    qa[i] = ss_struct_load_f32(L.rq[i], min(kb + lane, last), 0, 0, 0);
    qb[i] = ss_struct_load_f32(L.rq[i], min(kb + SS_WAVE + lane, last), 0, 0, 0);
  }
  for (int pass = 0;;) {
    ss_rt_pair<MODE, SL>(L, lane, pass, npass, kb, ub, lane_u, qa, qb, qc, qd);
    pass += 2;
    kb += 2 * SS_WAVE;
    ub += 2 * L.step;
    if (pass >= npass) break;
    ss_rt_pair<MODE, SL>(L, lane, pass, npass, kb, ub, lane_u, qc, qd, qa, qb);
    pass += 2;
    kb += 2 * SS_WAVE;
    ub += 2 * L.step;
    if (pass >= npass) break;
  }
}

// slices slice0 + blockIdx.y * SL .. + SL - 1
template <int SL>
__global__ __launch_bounds__(RT_LINES * SS_WAVE) void k_radon_transpose_lines(SSDevPlan p, const float* __restrict__ qT,
                                                                             int j0, int j1, int slice0,
                                                                             float* __restrict__ outB,
                                                                             float* __restrict__ outA) {
  extern __shared__ float sh[];
  const int n = p.g.n, R = p.g.nrays, A = p.g.nangles;
  const int lane = threadIdx.x & 63;
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int nw = (int)(blockDim.x >> 6);
  const int groups = (n + nw - 1) / nw;
  const int branch = blockIdx.x >= (unsigned)groups ? 1 : 0;
  const int s = ((int)blockIdx.x - branch * groups) * nw + wv;  // image row (B) or column (A)
  if (s >= n) return;  // waves work independently: no block barrier follows
  const size_t slice = (size_t)slice0 + (size_t)blockIdx.y * SL;
  const size_t sino_size = (size_t)A * R;
  const float* q = qT + slice * sino_size;
  SSRtLine L;
  // LDS is addressed as sh + offset, so that every access compiles to an LDS instruction
#pragma unroll
  for (int i = 0; i < SL; ++i) {
    L.line[i] = sh + (wv * SL + i) * ss_rt_line_stride(n);
    for (int c = lane; c < ss_rt_line_stride(n); c += SS_WAVE) L.line[i][c] = 0.f;
  }
  L.n = n;
  L.s = s;
  const ss_const_angles angles = (ss_const_angles)p.angles;
  for (int j = ss_rt_next(angles, j0, j1, branch); j < j1; j = ss_rt_next(angles, j + 1, j1, branch)) {
    const __attribute__((address_space(4))) SSAngle& a = angles[j];
    const int64_t gamma = a.gamma;
    // candidate rays [k_lo, k_hi) of line s: ray k sits at cell x = (k + a0(s)) / rt_ig on it,
    // and a valid sample has x in [-1, n]; two rays of margin against rounding (every
    // candidate is still checked)
    const double a0 = a.rt_a0 + (double)s * a.rt_b;
    int k_lo = 0, k_hi = R;
    if (gamma != 0) {
      const double kd = floor(a.rt_w - a0) - 2.0;
      k_lo = ss_rfl((int)fmin(fmax(kd, 0.0), (double)R));
      k_hi = ss_rfl((int)fmax(fmin(kd + (double)(a.rt_wn + 5), (double)R), 0.0));
    }
    if (k_lo >= k_hi) continue;
    L.k_lo = k_lo;
    L.k_hi = k_hi;
    L.all_edge = !(a.regular && n >= 4);
    // On a regular angle a sample whose Q32.32 cell m lies in [1, n-3] is valid: the FP64
    // coordinate of the exact predicate is within ~(k + s) 2^-33 cells of the trajectory,
    // so its floor is in [0, n-2].  Those samples are the rays reaching cells [1, n-2]; the
    // passes whose 64 rays lie at least one ray inside that interval (and below k_hi) hold
    // only valid samples.
    L.p_first = 0;
    L.p_last = 0;
    if (!L.all_edge && gamma != 0) {
      const int kf = ss_rfl((int)fmin(fmax(floor(a.rt_i0 - a0) + 2.0, -1.0), (double)R));  // first fast ray
      const int kl = min(ss_rfl((int)fmin(fmax(ceil(a.rt_i1 - a0) - 2.0, -1.0), (double)R)), k_hi - 1);
      // pass p covers rays k_lo + 64 p .. k_lo + 64 p + 63
      const int pf = kf <= k_lo ? 0 : (kf - k_lo + SS_WAVE - 1) / SS_WAVE;
      const int span = kl - k_lo - (SS_WAVE - 1);
      if (span >= 0) {
        L.p_first = pf;
        L.p_last = span / SS_WAVE + 1;  // one past the last fast pass
      }
    }
    L.delta = a.delta;
    // record k of the angle's range row (8 bytes) and sinogram rows (4 bytes); every index
    // used is clamped to [0, R-1]
    L.rb = __builtin_amdgcn_make_buffer_rsrc((void*)(p.bounds + (size_t)j * R), (short)sizeof(int2),
                                             (int)((uint32_t)R * 8u), SS_RSRC_DWORD3);
#pragma unroll
    for (int i = 0; i < SL; ++i)
      L.rq[i] = __builtin_amdgcn_make_buffer_rsrc((void*)(q + i * sino_size + (size_t)j * R), (short)sizeof(float),
                                                  (int)((uint32_t)R * 4u), SS_RSRC_DWORD3);
    L.gamma = gamma;
    L.u0 = a.alpha0 + (int64_t)s * a.beta + (int64_t)k_lo * gamma + ((int64_t)1 << 32);
    L.step = (int64_t)SS_WAVE * gamma;
    const uint64_t ag = (uint64_t)(gamma < 0 ? -gamma : gamma);
    if (ag >= (uint64_t)SS_FIX_ONE_I64) {
      ss_rt_passes<RT_DISTINCT, SL>(L, lane);
    } else if (ag * RT_PHASES_MAX >= (uint64_t)SS_FIX_ONE_I64) {
      // lanes `phases` apart are at least one cell apart
      L.phases = ss_rfl((int)(((uint64_t)SS_FIX_ONE_I64 + ag - 1) / ag));
      L.my_phase = lane % L.phases;
      ss_rt_passes<RT_PHASED, SL>(L, lane);
    } else {
      ss_rt_passes<RT_ATOMIC, SL>(L, lane);
    }
  }
#pragma unroll
  for (int i = 0; i < SL; ++i) {
    float* dst = (branch ? outA : outB) + ((slice + i) * n + s) * n;
    for (int c = lane; c < n; c += SS_WAVE) dst[c] = L.line[i][c + 1];
  }
}

// out[s][r][c] += t[s][c][r], 32x32 tiles
__global__ __launch_bounds__(256) void k_transpose_add(const float* __restrict__ t, float* __restrict__ out, int n) {
  __shared__ float tile[32][33];
  const size_t off = (size_t)blockIdx.z * n * n;
  const int bx = blockIdx.x * 32, by = blockIdx.y * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  for (int yy = ty; yy < 32; yy += 8) {
    const int r = by + yy, c = bx + tx;  // t row r (image column), column c (image row)
    if (r < n && c < n) tile[yy][tx] = t[off + (size_t)r * n + c];
  }
  __syncthreads();
  for (int xx = ty; xx < 32; xx += 8) {
    const int r = bx + xx, c = by + tx;  // image row r, column c
    if (r < n && c < n) out[off + (size_t)r * n + c] += tile[tx][xx];
  }
}

// irregular rays: scatter their exact contributions
__global__ __launch_bounds__(256) void k_radon_transpose_fixup(SSDevPlan p, const float* __restrict__ qT, int j0,
                                                               int j1, float* __restrict__ out) {
  const int id = p.irregular[blockIdx.x];
  const int slice = blockIdx.y;
  const int n = p.g.n, R = p.g.nrays, A = p.g.nangles;
  const int j = id / R, k = id % R;
  if (j < j0 || j >= j1) return;
  const SSAngle& a = p.angles[j];
  const float v = qT[((size_t)slice * A + j) * R + k] * a.delta;
  float* img = out + (size_t)slice * n * n;
  for (int s = threadIdx.x; s < n; s += blockDim.x) {
    if (!ss_radon_sample_valid(p.g, a, k, s)) continue;
    const int64_t u = a.alpha0 + (int64_t)k * a.gamma + (int64_t)s * a.beta;
    const int m = (int)(u >> 32);
    const float w = (float)(uint32_t)u * SS_INV_FIX;
    // pixel index along the interpolation axis m (weight 1-w) and m+1 (weight w)
    if (m >= 0 && m < n) {
      const size_t idx = a.branch_a ? (size_t)m * n + s : (size_t)s * n + m;
      atomicAdd(img + idx, (1.f - w) * v);
    }
    if (m + 1 >= 0 && m + 1 < n) {
      const size_t idx = a.branch_a ? (size_t)(m + 1) * n + s : (size_t)s * n + m + 1;
      atomicAdd(img + idx, w * v);
    }
  }
}

// =====================================================================================
// exact transpose, pixel-driven (rays at least one cell apart on every line)
//
// The forward adds (1-w) F[m] + w F[m+1] for a valid sample whose Q32.32 position on its
// line is u = m + w, so pixel p of line s receives Delta q_k max(0, 1 - |u_k(s) - p|)
// from ray k: a hat one cell wide.  On line s the rays sit at u_k = (alpha0 + k gamma +
// s beta) 2^-32, so the pixel lies at the real ray index
//   kappa = (p 2^32 - alpha0 - s beta) / gamma = p rt_ig - rt_a0 - s rt_b
// and ray k contributes Delta q_k max(0, 1 - g |k - kappa|), g = |gamma| 2^-32 the ray
// spacing in cells.  With g >= 1 only k0 = floor(kappa) and k0 + 1 can contribute:
//   w0 = clamp(1 - g f),  w1 = clamp(1 - g + g f),  f = kappa - k0.
// kappa is affine in the pixel's column and row (branch B: p = column, s = row; branch A:
// p = row, s = column), so a pixel tile evaluates it as the backprojection evaluates its
// detector coordinate: per angle of an LDS stage the tile origin's kappa in FP64 relative
// to a staged base ray, FP32 offsets per pixel, and staged (Delta q_k, Delta q_{k+1})
// pairs.  Both branches add into the same pixels in one pass.
//
// Validity.  A sample is valid when its FP64 predicate coordinate x lies in [0, n-1); on
// a regular angle x is within ~(k + s) 2^-33 cells of the Q32.32 trajectory, so every
// sample whose cell lies in [1, n-3] is valid and pixels p in [2, n-3] need no test.  In
// tiles holding pixels p < 2 or p > n-3 a regular angle tests, at those pixels, each
// candidate's position on its line, u = p - f gs or p + (1 - f) gs (gs the signed
// spacing), which the FP32
// evaluation gives within ~2e-5 cells: u in [eps, n-1-eps) is valid, u < -eps or
// u >= n-1+eps is not, and only u within eps = 1e-3 of 0 or n-1 reads the candidate's
// exact range (s in [lo_k, hi_k]).  An angle that is not regular (degenerate, or with
// irregular rays) reads the exact range of every candidate; irregular rays have empty
// ranges here and the fixup kernel adds them.  The weights are FP32 evaluations of the
// forward's (~1e-5 cells).
// =====================================================================================

#ifndef RTP_UNROLL
#define RTP_UNROLL 2  // angles per unrolled step of the pixel-driven transpose
#endif

struct RTPAngle {
  float f0, b, c, g;
  int flag;  // 0: no candidate ray on the detector, 1: no validity test, 2: test by position
             // (regular angle, tile at a line end), 3: exact ranges (angle not regular)
  int base;
  int branch_a;
  float gs;  // signed ray spacing in cells (gamma 2^-32)
};

// SL = 2: the block takes two slices of a launch and shares each pixel's geometry (kappa,
// floor, fraction, weights) between them.  win holds, per staged angle, the window of
// Delta q_k for rays base .. base + W (rt_window W; windows WS = W + 4 rays apart),
// copied straight from the pre-scaled sinogram into LDS: one float per ray, or with two
// slices the pair (slice 0, slice 1) per ray, as the pre-scaled copy interleaves them.  A
// pixel reads its two candidates with one two-word (two-pair) LDS read.
template <int STAGE, int SL>
struct SSRtpLds {
  alignas(16) float win[2 * STAGE];  // STAGE pairs' worth: rt_angles x WS floats
  RTPAngle par[64];
};

// sl: the sinogram [A][R] pre-scaled by each angle's Delta; with SL = 2 the two slices
// interleaved, [A][R][2]
template <int ROWS, int STAGE, int SL>
__device__ __forceinline__ void ss_rtp_item(const SSDevPlan& p, const float* __restrict__ sl, int j0, int j1,
                                            int c0, int r0, float* __restrict__ out0, float* __restrict__ out1,
                                            SSRtpLds<STAGE, SL>& lds) {
  const int n = p.g.n, R = p.g.nrays;
  const int W = p.rt_window;
  const int WS = W + 4;
  const int J = SL == 2 ? p.rt_angles / 2 : p.rt_angles;  // two windows per angle with two slices
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int tx = tid & 31, ty = tid >> 5;
  const int col = c0 + tx;
  constexpr int TH = 8 * ROWS;
  // two accumulators per pixel and slice (the two candidates' terms) keep the FMA chains short
  float acc0[SL][ROWS], acc1[SL][ROWS];
  float fty[ROWS];
#pragma unroll
  for (int i = 0; i < ROWS; ++i) {
    fty[i] = (float)(ty + 8 * i);
#pragma unroll
    for (int z = 0; z < SL; ++z) acc0[z][i] = acc1[z][i] = 0.f;
  }
  const float ftx = (float)tx;
  // tile pixels along each branch's cell axis: columns (branch B), rows (branch A)
  const bool inner_b = c0 >= 2 && c0 + BP_TILE_W - 1 <= n - 3;
  const bool inner_a = r0 >= 2 && r0 + TH - 1 <= n - 3;

  for (int jb = j0; jb < j1; jb += J) {
    const int nj = min(J, j1 - jb);
    RTPAngle q{0.f, 0.f, 0.f, 0.f, 0, 0, 0, 0.f};
    if (lane < nj) {
      const SSAngle& a = p.angles[jb + lane];
      const double kb = a.branch_a ? -a.rt_b : a.rt_ig;  // kappa per column
      const double kc = a.branch_a ? a.rt_ig : -a.rt_b;  // kappa per row
      const double K = -a.rt_a0 + (double)c0 * kb + (double)r0 * kc;
      const double eb = (double)(BP_TILE_W - 1) * kb, ec = (double)(TH - 1) * kc;
      const double kmin = K + fmin(0.0, eb) + fmin(0.0, ec);
      const double kmax = K + fmax(0.0, eb) + fmax(0.0, ec);
      const double base = floor(kmin) - 1.0;
      q.f0 = (float)(K - base);
      q.b = (float)kb;
      q.c = (float)kc;
      q.g = (float)(fabs((double)a.gamma) * (1.0 / SS_FIX_ONE));
      q.gs = (float)((double)a.gamma * (1.0 / SS_FIX_ONE));
      q.base = (int)base;
      q.branch_a = a.branch_a;
      if (kmax < -1.0 || kmin >= (double)R)
        q.flag = 0;
      else if (!a.regular)
        q.flag = 3;
      else if (a.branch_a ? inner_a : inner_b)
        q.flag = 1;
      else
        q.flag = 2;
    }
    const bool checked = __ballot(lane < nj && q.flag >= 2) != 0ull;
    __syncthreads();  // the previous stage is done with win / par
    if (wv == 0) {
      lds.par[lane].f0 = q.f0;
      lds.par[lane].b = q.b;
      lds.par[lane].c = q.c;
      lds.par[lane].g = q.g;
      lds.par[lane].flag = q.flag;
      lds.par[lane].base = q.base;
      lds.par[lane].branch_a = q.branch_a;
      lds.par[lane].gs = q.gs;
    }
    // staging: wave wv fills the windows of angles wv, wv+4, ... with direct buffer -> LDS
    // copies, one ray per lane; rays outside [0, R) (and every ray of an angle without
    // candidates) read 0 through the buffer's range check.  kappa - base >= 1 for every
    // pixel, so the truncated offset is the floor and stays below W.
    for (int jj = wv; jj < nj; jj += 4) {
      const int fl = __builtin_amdgcn_readlane(q.flag, jj);
      const int bs = __builtin_amdgcn_readlane(q.base, jj);
      const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
          (void*)(sl + (size_t)(jb + jj) * R * SL), (short)0, R * SL * (int)sizeof(float), 0x00020000);
      float* dst = lds.win + jj * SL * WS;
      for (int w = 0; w < SL * (W + 1); w += 64) {
        // the whole offset in the register (none in the instruction's offset field), so
        // the range check sees a ray below 0 as out of range
        int off = fl != 0 ? (SL * bs + w + lane) * (int)sizeof(float) : 0x7ffffff0;
        __asm__ volatile("" : "+v"(off));
        if (w + lane < SL * (W + 1))  // active lanes write dst[w + lane]: stay in this window
          __builtin_amdgcn_raw_ptr_buffer_load_lds(rs, (ss_lds_ptr)(dst + w), 4, off, 0, 0, 0);
      }
    }
    __syncthreads();
    if (!checked) {
#pragma unroll RTP_UNROLL
      for (int jj = 0; jj < nj; ++jj) {
        const float f0 = lds.par[jj].f0, pb = lds.par[jj].b, pc = lds.par[jj].c, g = lds.par[jj].g;
        const float g1 = 1.f - g;
        const float* wrow = lds.win + jj * SL * WS;
        const float t0 = fmaf(ftx, pb, f0);
#pragma unroll
        for (int i = 0; i < ROWS; ++i) {
          const float kap = fmaf(fty[i], pc, t0);
          const int kr = (int)kap;
          const float f = __builtin_amdgcn_fractf(kap);
          const float w0 = __builtin_amdgcn_fmed3f(fmaf(-g, f, 1.f), 0.f, 1.f);
          const float w1 = __builtin_amdgcn_fmed3f(fmaf(g, f, g1), 0.f, 1.f);
          if constexpr (SL == 2) {
            const float2 t0 = reinterpret_cast<const float2*>(wrow)[kr];
            const float2 t1 = reinterpret_cast<const float2*>(wrow)[kr + 1];
            acc0[0][i] = fmaf(w0, t0.x, acc0[0][i]);
            acc1[0][i] = fmaf(w1, t1.x, acc1[0][i]);
            acc0[1][i] = fmaf(w0, t0.y, acc0[1][i]);
            acc1[1][i] = fmaf(w1, t1.y, acc1[1][i]);
          } else {
            acc0[0][i] = fmaf(w0, wrow[kr], acc0[0][i]);
            acc1[0][i] = fmaf(w1, wrow[kr + 1], acc1[0][i]);
          }
        }
      }
    } else {
      for (int jj = 0; jj < nj; ++jj) {
        const int flag = __builtin_amdgcn_readfirstlane(lds.par[jj].flag);
        if (flag == 0) continue;
        const float f0 = lds.par[jj].f0, pb = lds.par[jj].b, pc = lds.par[jj].c, g = lds.par[jj].g;
        const float g1 = 1.f - g;
        const float* wrow = lds.win + jj * SL * WS;
        const float t0 = fmaf(ftx, pb, f0);
        const int base = lds.par[jj].base;
        const int branch_a = __builtin_amdgcn_readfirstlane(lds.par[jj].branch_a);
        const float gs = lds.par[jj].gs;
        const int2* rng = p.bounds + (size_t)(jb + jj) * R;
        const float top = (float)(n - 1);
        constexpr float eps = 1e-3f;
#pragma unroll
        for (int i = 0; i < ROWS; ++i) {
          const float kap = fmaf(fty[i], pc, t0);
          const int kr = (int)kap;
          const float f = __builtin_amdgcn_fractf(kap);
          float w0 = __builtin_amdgcn_fmed3f(fmaf(-g, f, 1.f), 0.f, 1.f);
          float w1 = __builtin_amdgcn_fmed3f(fmaf(g, f, g1), 0.f, 1.f);
          const int row = r0 + ty + 8 * i;
          const int pix = branch_a ? row : col;  // the pixel's cell on its sample line
          if (flag == 3 || (flag == 2 && (pix <= 1 || pix >= n - 2))) {
            const int sline = branch_a ? col : row;  // the candidates' sample line
            const float pcell = (float)pix;
            const int k0 = base + kr;
            int keep;
            if (flag == 2) {
              // position on the line decides, except within eps of either end
              const float u0 = fmaf(-f, gs, pcell), u1 = u0 + gs;
              keep = ((u0 >= eps && u0 < top - eps) ? 1 : 0) | ((u1 >= eps && u1 < top - eps) ? 2 : 0);
              const bool amb0 = fabsf(u0) < eps || fabsf(u0 - top) < eps;
              const bool amb1 = fabsf(u1) < eps || fabsf(u1 - top) < eps;
              if (amb0 && (unsigned)k0 < (unsigned)R) {
                const int2 b0 = rng[k0];
                keep |= (sline >= b0.x && sline <= b0.y) ? 1 : 0;
              }
              if (amb1 && (unsigned)(k0 + 1) < (unsigned)R) {
                const int2 b1 = rng[k0 + 1];
                keep |= (sline >= b1.x && sline <= b1.y) ? 2 : 0;
              }
            } else {
              // the candidates' exact ranges
              int2 b0 = make_int2(0, -1), b1 = make_int2(0, -1);
              if ((unsigned)k0 < (unsigned)R) b0 = rng[k0];
              if ((unsigned)(k0 + 1) < (unsigned)R) b1 = rng[k0 + 1];
              keep = ((sline >= b0.x && sline <= b0.y) ? 1 : 0) | ((sline >= b1.x && sline <= b1.y) ? 2 : 0);
            }
            w0 = (keep & 1) ? w0 : 0.f;
            w1 = (keep & 2) ? w1 : 0.f;
          }
          if constexpr (SL == 2) {
            const float2 t0 = reinterpret_cast<const float2*>(wrow)[kr];
            const float2 t1 = reinterpret_cast<const float2*>(wrow)[kr + 1];
            acc0[0][i] = fmaf(w0, t0.x, acc0[0][i]);
            acc1[0][i] = fmaf(w1, t1.x, acc1[0][i]);
            acc0[1][i] = fmaf(w0, t0.y, acc0[1][i]);
            acc1[1][i] = fmaf(w1, t1.y, acc1[1][i]);
          } else {
            acc0[0][i] = fmaf(w0, wrow[kr], acc0[0][i]);
            acc1[0][i] = fmaf(w1, wrow[kr + 1], acc1[0][i]);
          }
        }
      }
    }
  }
  if (col >= n) return;
#pragma unroll
  for (int i = 0; i < ROWS; ++i) {
    const int row = r0 + ty + 8 * i;
    if (row >= n) continue;
    out0[(size_t)row * n + col] = acc0[0][i] + acc1[0][i];
    if constexpr (SL == 2) out1[(size_t)row * n + col] = acc0[1][i] + acc1[1][i];
  }
}

// Blocks per CU the compiler must fit (registers): eight 256-thread blocks (8 waves per
// SIMD, 64 VGPRs) for single-slice tiles of up to 4 rows per thread, whose register need
// is near 64 (2048^2 x 1500: 2.66 -> 2.57 ms); slice pairs would spill under that bound.
#ifndef RTP_BLOCKS_SMALL
#define RTP_BLOCKS_SMALL 8
#endif
#define RTP_MIN_BLOCKS(ROWS, SL) ((ROWS) <= 4 && (SL) == 1 ? RTP_BLOCKS_SMALL : 1)

// slices slice0 + blockIdx.z * SL .. + SL - 1
template <int ROWS, int STAGE, int SL>
__global__ __launch_bounds__(256, RTP_MIN_BLOCKS(ROWS, SL)) void k_radon_transpose_px(SSDevPlan p, const float* __restrict__ qT, int j0, int j1,
                                                            int slice0, float* __restrict__ out) {
  __shared__ SSRtpLds<STAGE, SL> lds;
  const size_t slice = (size_t)slice0 + (size_t)blockIdx.z * SL;
  const size_t sino = (size_t)p.g.nangles * p.g.nrays, img = (size_t)p.g.n * p.g.n;
  // the pre-scaled copy holds slice pairs interleaved in the pairs' own space
  ss_rtp_item<ROWS, STAGE, SL>(p, qT + slice * sino, j0, j1, blockIdx.x * BP_TILE_W, blockIdx.y * 8 * ROWS,
                               out + slice * img, out + (slice + SL - 1) * img, lds);
}

// LDS a block may use on the current device (gfx950: 160 KiB, also assumed if the query
// fails)
static size_t ss_lds_per_block() {
  int dev = 0, v = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess || v <= 0)
    return 160 * 1024;
  return (size_t)v;
}

// lines (waves) per block for the line kernel holding `sl` slices per wave: the most whose
// lines fit one block's LDS (one line of the largest image, 16384 pixels, takes 64 KiB)
static int ss_rt_lines(int n, int sl) {
  const size_t budget = ss_lds_per_block();
  for (int nl = RT_LINES; nl >= 1; nl >>= 1)
    if (ss_rt_line_lds(n, nl * sl) <= budget) return nl;
  return 0;
}

// The pixel-driven transpose's pre-scaled rays: in[slice][a][r] * Delta_a for the angles
// j0 .. j1 - 1; slices 2i, 2i + 1 (i < pairs) interleaved in their own space,
// out[2i][a][r][2], the others as they are
__global__ __launch_bounds__(256) void k_rtp_scale(SSDevPlan p, const float* __restrict__ in, int j0, int j1,
                                                   int pairs, float* __restrict__ out) {
  const int R = p.g.nrays;
  const size_t slice = blockIdx.z;
  const int a = j0 + blockIdx.y;
  if (a >= j1) return;
  const size_t sino = (size_t)p.g.nangles * R;
  const size_t row = slice * sino + (size_t)a * R;
  const float d = p.angles[a].delta;
  const bool paired = slice < 2 * (size_t)pairs;
  float* __restrict__ dst = paired ? out + (slice & ~(size_t)1) * sino + (size_t)a * R * 2 + (slice & 1) : out + row;
  const int step = paired ? 2 : 1;
  for (int r = blockIdx.x * 256 + threadIdx.x; r < R; r += gridDim.x * 256) dst[(size_t)r * step] = in[row + r] * d;
}

size_t ss_radon_transpose_scratch(const SSDevPlan& p, int nslices) {
  return p.rt_angles > 0 ? sizeof(float) * (size_t)nslices * p.g.nangles * p.g.nrays
                         : sizeof(float) * (size_t)nslices * p.g.n * p.g.n;
}

hipError_t ss_launch_radon_transpose(const SSDevPlan& p, const float* sinoT, int j0, int j1, int nslices,
                                     float* out, float* scratch, hipStream_t stream) {
  const int n = p.g.n;
  if (p.rt_angles > 0) {  // the pixel-driven form (every angle's rays at least one cell apart)
    if (!scratch) return hipErrorInvalidValue;
    const int th = 8 * kSSBpRows[p.rt_variant];
    // slices in pairs (a pair shares each pixel's geometry) when a pair's stage holds at
    // least 8 angles, the last odd slice alone
    const int pairs = (p.rt_angles / 2 >= 8) ? nslices / 2 : 0;
    hipLaunchKernelGGL(k_rtp_scale, dim3((p.g.nrays + 255) / 256, j1 - j0, nslices), dim3(256), 0, stream, p, sinoT,
                       j0, j1, pairs, scratch);
    const hipError_t es = hipGetLastError();
    if (es != hipSuccess) return es;
    const float* scaled = scratch;
    const int singles = nslices - 2 * pairs;
    for (int part = 0; part < 2; ++part) {
      const int count = part == 0 ? pairs : singles;
      if (count == 0) continue;
      dim3 grid((n + BP_TILE_W - 1) / BP_TILE_W, (n + th - 1) / th, count);
      const int slice0 = part == 0 ? 0 : 2 * pairs;
#define SS_RTP_LAUNCH(V, ROWS, STAGE)                                                                            \
  case V:                                                                                                        \
    if (part == 0)                                                                                               \
      hipLaunchKernelGGL((k_radon_transpose_px<ROWS, STAGE, 2>), grid, dim3(256), 0, stream, p, scaled, j0, j1, \
                         slice0, out);                                                                           \
    else                                                                                                         \
      hipLaunchKernelGGL((k_radon_transpose_px<ROWS, STAGE, 1>), grid, dim3(256), 0, stream, p, scaled, j0, j1, \
                         slice0, out);                                                                           \
    break;
      switch (p.rt_variant) {
        SS_BP_FOR_VARIANTS(SS_RTP_LAUNCH)
        default:
          return hipErrorInvalidValue;
      }
#undef SS_RTP_LAUNCH
      const hipError_t e = hipGetLastError();
      if (e != hipSuccess) return e;
    }
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess || p.nirregular == 0) return e;
    hipLaunchKernelGGL(k_radon_transpose_fixup, dim3(p.nirregular, nslices), dim3(256), 0, stream, p, sinoT, j0, j1,
                       out);
    return hipGetLastError();
  }
  // several slices per wave where the image is small enough for their lines to leave
  // room for many waves per CU: four for n <= 512, two for n <= 1024 (8 KiB per wave);
  // the remaining slices go in pairs, then alone
  const int quads = (n <= 512) ? nslices / 4 : 0;
  const int pairs = (n <= 1024) ? (nslices - 4 * quads) / 2 : 0;
  const int singles = nslices - 4 * quads - 2 * pairs;
  if (quads > 0) {
    const int nl = ss_rt_lines(n, 4);
    if (nl == 0) return hipErrorInvalidConfiguration;
    const int groups = (n + nl - 1) / nl;
    hipLaunchKernelGGL(k_radon_transpose_lines<4>, dim3(2 * groups, quads), dim3(nl * SS_WAVE),
                       ss_rt_line_lds(n, 4 * nl), stream, p, sinoT, j0, j1, 0, out, scratch);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  if (pairs > 0) {
    const int nl = ss_rt_lines(n, 2);
    if (nl == 0) return hipErrorInvalidConfiguration;
    const int groups = (n + nl - 1) / nl;
    hipLaunchKernelGGL(k_radon_transpose_lines<2>, dim3(2 * groups, pairs), dim3(nl * SS_WAVE),
                       ss_rt_line_lds(n, 2 * nl), stream, p, sinoT, j0, j1, 4 * quads, out, scratch);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  if (singles > 0) {
    const int nl = ss_rt_lines(n, 1);
    if (nl == 0) return hipErrorInvalidConfiguration;  // a line does not fit the LDS
    const int groups = (n + nl - 1) / nl;
    hipLaunchKernelGGL(k_radon_transpose_lines<1>, dim3(2 * groups, singles), dim3(nl * SS_WAVE),
                       ss_rt_line_lds(n, nl), stream, p, sinoT, j0, j1, 4 * quads + 2 * pairs, out, scratch);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  const int t = (n + 31) / 32;
  hipLaunchKernelGGL(k_transpose_add, dim3(t, t, nslices), dim3(256), 0, stream, scratch, out, n);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess || p.nirregular == 0) return e;
  hipLaunchKernelGGL(k_radon_transpose_fixup, dim3(p.nirregular, nslices), dim3(256), 0, stream, p, sinoT, j0, j1,
                     out);
  return hipGetLastError();
}

// =====================================================================================
// FBP row filter: circular convolution of each length-R row with g
// (g = Re(ifft(fftshift(h))), the real kernel equivalent to pyraft's lowpassino_fbp)
// 4 rows per block; g (duplicated, length 2R) and the rows in LDS.
// =====================================================================================

#define RF_ROWS 4
// out[row][k] = sum_m in[row][m] * taps[k - m + R].  toeplitz = 0: g holds R taps of a
// circular convolution (taps[j] = g[j mod R]); toeplitz = 1: g holds the 2R Toeplitz
// taps themselves (entry 0 unused).
__global__ __launch_bounds__(256) void k_ramp_filter(const float* __restrict__ in, float* __restrict__ out,
                                                     const float* __restrict__ g, int R, int nrows, int toeplitz) {
  extern __shared__ float lds[];
  float* g2 = lds;                          // 2R
  float* rows = lds + ((2 * R + 3) & ~3);   // R x RF_ROWS, interleaved [m][row], 16-byte aligned
  const int row0 = blockIdx.x * RF_ROWS;
  if (toeplitz) {
    for (int i = threadIdx.x; i < 2 * R; i += 256) g2[i] = g[i];
  } else {
    for (int i = threadIdx.x; i < R; i += 256) {
      const float v = g[i];
      g2[i] = v;
      g2[i + R] = v;
    }
  }
  for (int i = threadIdx.x; i < R * RF_ROWS; i += 256) {
    const int rr = i / R, m = i - rr * R;
    const int row = row0 + rr;
    rows[m * RF_ROWS + rr] = row < nrows ? in[(size_t)row * R + m] : 0.f;
  }
  __syncthreads();
  for (int k = threadIdx.x; k < R; k += 256) {
    float a0 = 0.f, a1 = 0.f, a2 = 0.f, a3 = 0.f;
    const float* gk = g2 + k + R;
    for (int m = 0; m < R; ++m) {
      const float gv = gk[-m];
      const float4 s = *reinterpret_cast<const float4*>(rows + m * RF_ROWS);
      a0 = fmaf(gv, s.x, a0);
      a1 = fmaf(gv, s.y, a1);
      a2 = fmaf(gv, s.z, a2);
      a3 = fmaf(gv, s.w, a3);
    }
    if (row0 + 0 < nrows) out[(size_t)(row0 + 0) * R + k] = a0;
    if (row0 + 1 < nrows) out[(size_t)(row0 + 1) * R + k] = a1;
    if (row0 + 2 < nrows) out[(size_t)(row0 + 2) * R + k] = a2;
    if (row0 + 3 < nrows) out[(size_t)(row0 + 3) * R + k] = a3;
  }
}

hipError_t ss_launch_ramp_filter(const float* in, float* out, const float* g, int R, int nrows, hipStream_t stream,
                                 bool toeplitz) {
  if (nrows <= 0) return hipSuccess;
  if (const size_t bytes = ss_fft_taps_bytes(R)) {
    // FFT form, its spectrum in stream-ordered scratch
    void* mem = nullptr;
    hipError_t e = hipMallocAsync(&mem, bytes, stream);
    if (e != hipSuccess) return e;
    SSFftTaps t;
    e = ss_fft_taps_build(g, R, toeplitz, mem, &t, stream);
    if (e == hipSuccess) e = ss_launch_fft_filter(in, out, t, nrows, stream);
    const hipError_t f = hipFreeAsync(mem, stream);
    return e != hipSuccess ? e : f;
  }
  const size_t lds = (size_t)(((2 * R + 3) & ~3) + RF_ROWS * R) * sizeof(float);
  if (lds > 160 * 1024) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_ramp_filter, dim3((nrows + RF_ROWS - 1) / RF_ROWS), dim3(256), lds, stream, in, out, g, R,
                     nrows, toeplitz ? 1 : 0);
  return hipGetLastError();
}

// =====================================================================================
// FFT form of the row filter.  With taps2[j] (j in [0, 2R)) the filter is
//   y[n] = sum_m x[m] taps2[n - m + R],   n, m in [0, R),
// where n - m + R stays in [1, 2R - 1].  Zero-padded to a power of two M >= 2R, the
// M-point circular convolution c = x (*) taps2 holds y[n] = c[n + R]: no term wraps.
// Two rows go into one complex sequence (row 2i real, row 2i + 1 imaginary); the taps
// are real, so the real and imaginary parts of c are the two rows' convolutions.
// A block computes c = IFFT(FFT(z) * G) in LDS with G = FFT(taps2), both transforms
// Stockham autosort passes (radix 4, one radix-2 pass first when log2 M is odd) over
// a table W[k] = exp(-2 pi i k / M).  The transforms run in FP64: FP32 rounding in the
// spectrum of a smooth row (large low-frequency bins, which the ramp suppresses) would
// reach 1e-5 of the filtered row at R = 2048.
// =====================================================================================

#define FF_THREADS 512

__device__ __forceinline__ double2 ss_zadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 ss_zsub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ double2 ss_zmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// The M-point DFT (INV: with exp(+2 pi i k / M), unnormalised) of the sequence in `a`,
// with `b` as the second buffer; returns the buffer holding the result.  Every thread
// of the block takes part (the passes end with block barriers).
template <bool INV>
__device__ double2* ss_fft_lds(double2* a, double2* b, const double2* __restrict__ W, int M, int lgM) {
  const int tid = threadIdx.x;
  int Ns = 1;
  if (lgM & 1) {
    const int h = M >> 1;
    for (int j = tid; j < h; j += FF_THREADS) {
      const double2 v0 = a[j], v1 = a[j + h];
      b[2 * j] = ss_zadd(v0, v1);
      b[2 * j + 1] = ss_zsub(v0, v1);
    }
    __syncthreads();
    double2* t = a;
    a = b;
    b = t;
    Ns = 2;
  }
  const int q = M >> 2;
  for (; Ns < M; Ns <<= 2) {
    const int ts = q / Ns;  // twiddle stride: exp(-2 pi i r k / (4 Ns)) = W[r k ts]
    for (int j = tid; j < q; j += FF_THREADS) {
      const int k = j & (Ns - 1);
      double2 w1 = W[k * ts], w2 = W[2 * k * ts], w3 = W[3 * k * ts];
      if (INV) {
        w1.y = -w1.y;
        w2.y = -w2.y;
        w3.y = -w3.y;
      }
      const double2 v0 = a[j], v1 = ss_zmul(a[j + q], w1), v2 = ss_zmul(a[j + 2 * q], w2),
                    v3 = ss_zmul(a[j + 3 * q], w3);
      const double2 a0 = ss_zadd(v0, v2), a1 = ss_zsub(v0, v2), a2 = ss_zadd(v1, v3), d = ss_zsub(v1, v3);
      const double2 a3 = INV ? make_double2(-d.y, d.x) : make_double2(d.y, -d.x);  // d * (+-i)
      const int o = (j - k) * 4 + k;
      b[o] = ss_zadd(a0, a2);
      b[o + Ns] = ss_zadd(a1, a3);
      b[o + 2 * Ns] = ss_zsub(a0, a2);
      b[o + 3 * Ns] = ss_zsub(a1, a3);
    }
    __syncthreads();
    double2* t = a;
    a = b;
    b = t;
  }
  return a;
}

__global__ __launch_bounds__(256) void k_fft_twiddles(double2* __restrict__ W, int M) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= M) return;
  double sv, cv;
  sincospi(-2.0 * (double)k / (double)M, &sv, &cv);
  W[k] = make_double2(cv, sv);
}

// G = FFT(taps2 zero-padded to M); taps2[j] = g[j] (toeplitz) or g[j mod R]
__global__ __launch_bounds__(FF_THREADS) void k_fft_taps(const float* __restrict__ g, int R, int toeplitz,
                                                         const double2* __restrict__ W, int M, int lgM,
                                                         double2* __restrict__ G) {
  extern __shared__ double2 zb[];
  double2* a = zb;
  for (int m = threadIdx.x; m < M; m += FF_THREADS) {
    double v = 0.0;
    if (m < 2 * R) v = toeplitz ? (m > 0 ? g[m] : 0.f) : g[m < R ? m : m - R];
    a[m] = make_double2(v, 0.0);
  }
  __syncthreads();
  const double2* f = ss_fft_lds<false>(a, zb + M, W, M, lgM);
  for (int m = threadIdx.x; m < M; m += FF_THREADS) G[m] = f[m];
}

// Row r of the filter's input / output: angle-major (cols = 0: row r at r * R, unit
// stride) or the (R, A) layout (cols = A: row r = slice * A + a at slice * R * A + a,
// stride A), which spares a small sinogram the layout copies around the filter.
__device__ __forceinline__ size_t ss_ff_row(int r, int R, int cols) {
  return cols > 0 ? (size_t)(r / cols) * R * cols + (r % cols) : (size_t)r * R;
}

__global__ __launch_bounds__(FF_THREADS) void k_fft_filter(const float* __restrict__ in, float* __restrict__ out,
                                                           const double2* __restrict__ G,
                                                           const double2* __restrict__ W, int R, int M, int lgM,
                                                           int nrows, int in_cols, int out_cols) {
  extern __shared__ double2 zb[];
  double2* a = zb;
  double2* b = zb + M;
  const int r0 = 2 * blockIdx.x;
  const bool two = r0 + 1 < nrows;
  const size_t is = in_cols > 0 ? (size_t)in_cols : 1;
  const float* x0 = in + ss_ff_row(r0, R, in_cols);
  const float* x1 = in + ss_ff_row(two ? r0 + 1 : r0, R, in_cols);
  for (int m = threadIdx.x; m < M; m += FF_THREADS) {
    double2 z = make_double2(0.0, 0.0);
    if (m < R) {
      z.x = x0[m * is];
      if (two) z.y = x1[m * is];
    }
    a[m] = z;
  }
  __syncthreads();
  double2* f = ss_fft_lds<false>(a, b, W, M, lgM);
  for (int m = threadIdx.x; m < M; m += FF_THREADS) f[m] = ss_zmul(f[m], G[m]);
  __syncthreads();
  const double2* c = ss_fft_lds<true>(f, f == a ? b : a, W, M, lgM);
  const double scale = 1.0 / (double)M;
  const size_t os = out_cols > 0 ? (size_t)out_cols : 1;
  float* y0 = out + ss_ff_row(r0, R, out_cols);
  float* y1 = out + ss_ff_row(two ? r0 + 1 : r0, R, out_cols);
  for (int n = threadIdx.x; n < R; n += FF_THREADS) {
    const double2 v = c[n + R];
    y0[n * os] = (float)(v.x * scale);
    if (two) y1[n * os] = (float)(v.y * scale);
  }
}

int ss_fft_size(int R, int* lg) {
  int M = 4, l = 2;
  while (M < 2 * R) {
    M <<= 1;
    ++l;
  }
  if (M > SS_FFT_MAX) return 0;
  if (lg) *lg = l;
  return M;
}

size_t ss_fft_taps_bytes(int R) {
  const int M = ss_fft_size(R, nullptr);
  return M ? 2 * sizeof(double2) * (size_t)M : 0;
}

hipError_t ss_fft_taps_build(const float* g, int R, bool toeplitz, void* mem, SSFftTaps* t, hipStream_t stream) {
  int lg = 0;
  const int M = ss_fft_size(R, &lg);
  if (!M) return hipErrorInvalidValue;
  double2* W = reinterpret_cast<double2*>(mem);
  double2* G = W + M;
  hipLaunchKernelGGL(k_fft_twiddles, dim3((M + 255) / 256), dim3(256), 0, stream, W, M);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_fft_taps, dim3(1), dim3(FF_THREADS), 2 * sizeof(double2) * (size_t)M, stream, g, R,
                     toeplitz ? 1 : 0, W, M, lg, G);
  e = hipGetLastError();
  if (e != hipSuccess) return e;
  t->W = W;
  t->G = G;
  t->R = R;
  t->M = M;
  t->lgM = lg;
  return hipSuccess;
}

hipError_t ss_launch_fft_filter(const float* in, float* out, const SSFftTaps& t, int nrows, hipStream_t stream,
                                int in_cols, int out_cols) {
  if (nrows <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_fft_filter, dim3((nrows + 1) / 2), dim3(FF_THREADS), 2 * sizeof(double2) * (size_t)t.M,
                     stream, in, out, t.G, t.W, t.R, t.M, t.lgM, nrows, in_cols, out_cols);
  return hipGetLastError();
}

// =====================================================================================
// transmission EM (raft/raft_transmission_em.cu:52-61)
// =====================================================================================

// expected counts: p <- flat * exp(-p) (raft/raft_radon_gpu_function.cu:31-39 with expo = -1,
// then the product with the flat field)
__global__ __launch_bounds__(256) void k_tr_expected(float* __restrict__ p, const float* __restrict__ flat, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    p[i] = flat[i] * expf(-p[i]);
}

// x <- (x * r) * r, r = b / w: the reference applies its multiplicative update twice;
// pixels with w = 0 keep their value (the reference divides unguarded)
__global__ __launch_bounds__(256) void k_tr_update(float* __restrict__ x, const float* __restrict__ b,
                                                   const float* __restrict__ w, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const float wi = w[i];
    if (wi != 0.f) {
      const float r = b[i] / wi;
      x[i] = (x[i] * r) * r;
    }
  }
}

hipError_t ss_launch_tr_expected(float* p, const float* flat, size_t n, hipStream_t stream) {
  hipLaunchKernelGGL(k_tr_expected, dim3(2048), dim3(256), 0, stream, p, flat, n);
  return hipGetLastError();
}

hipError_t ss_launch_tr_update(float* x, const float* b, const float* w, size_t n, hipStream_t stream) {
  hipLaunchKernelGGL(k_tr_update, dim3(2048), dim3(256), 0, stream, x, b, w, n);
  return hipGetLastError();
}

// =====================================================================================
// small helpers
// =====================================================================================

// dst = sum(partials[0..n)) in a fixed order (one block); a store, so the caller clears nothing
__global__ __launch_bounds__(256) void k_sum_partials(const double* __restrict__ partials, int n,
                                                      double* __restrict__ dst) {
  __shared__ double red[4];
  double v = 0.0;
  for (int i = threadIdx.x; i < n; i += 256) v += partials[i];
  const double tot = ss_block_sum_f64(v, red);
  if (threadIdx.x == 0) *dst = tot;
}

hipError_t ss_launch_sum_partials(const double* partials, int n, double* dst, hipStream_t stream) {
  hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(256), 0, stream, partials, n, dst);
  return hipGetLastError();
}

__global__ __launch_bounds__(256) void k_fill(float* __restrict__ p, float v, size_t n) {
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) p[i] = v;
}

hipError_t ss_launch_fill(float* p, float v, size_t n, hipStream_t stream) {
  size_t blocks = (n + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  if (blocks == 0) return hipSuccess;
  hipLaunchKernelGGL(k_fill, dim3((unsigned)blocks), dim3(256), 0, stream, p, v, n);
  return hipGetLastError();
}

// =====================================================================================
// batched execution: persistent kernels over the work items of many slices
//
// A launch covers entries (slice, angle range) of a batch whose slices differ in
// n, R and A.  Items are numbered entry by entry; a fixed grid of resident blocks
// takes items blockIdx.x, blockIdx.x + gridDim.x, ...  The host orders entries by
// decreasing item cost, so the last items to run are the cheapest.  Each item runs
// the same device function as the per-plan kernels, on that slice's plan and
// buffers, so a slice's result does not depend on what it is batched with.
// =====================================================================================

// largest entry e with entries[e].item0 <= item (entries with no items are skipped
// because a later entry shares their item0)
__device__ __forceinline__ int ss_find_entry(const SSBatchEntry* __restrict__ e, int n, int item) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (e[mid].item0 <= item)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

// A copy of *src held in scalar registers: the tables live in ordinary global memory,
// which the kernels also write, so the compiler would otherwise load them per lane.
template <class T>
__device__ __forceinline__ T ss_uniform_copy(const T* src) {
  static_assert(sizeof(T) % 4 == 0, "dword-sized type");
  int w[sizeof(T) / 4];
  const int* p = reinterpret_cast<const int*>(src);
#pragma unroll
  for (int i = 0; i < (int)(sizeof(T) / 4); ++i) w[i] = __builtin_amdgcn_readfirstlane(p[i]);
  T out;
  __builtin_memcpy(&out, w, sizeof(T));
  return out;
}

// The entry of `item` with every field wave-uniform: the compiler then keeps the
// slice's plan in scalar registers instead of per-lane copies.
__device__ __forceinline__ SSBatchEntry ss_entry_of(const SSBatchLaunch& L, int item) {
  const int idx = __builtin_amdgcn_readfirstlane(ss_find_entry(L.entries, L.nentries, item));
  SSBatchEntry e = L.entries[idx];
  e.slice = __builtin_amdgcn_readfirstlane(e.slice);
  e.j0 = __builtin_amdgcn_readfirstlane(e.j0);
  e.j1 = __builtin_amdgcn_readfirstlane(e.j1);
  e.item0 = __builtin_amdgcn_readfirstlane(e.item0);
  e.nslab = __builtin_amdgcn_readfirstlane(e.nslab);
  e.slab_rows = __builtin_amdgcn_readfirstlane(e.slab_rows);
  e.part_off = ss_rfl64(e.part_off);
  e.which = __builtin_amdgcn_readfirstlane(e.which);
  return e;
}

// GT x GT tiles with NW waves (ss_batch_gather_tile picks the shape per launch)
template <int GT, int NW>
__global__ __launch_bounds__(NW * SS_WAVE) void k_batch_gather(SSBatchLaunch L, const float* __restrict__ images,
                                                               float2* __restrict__ gA, float2* __restrict__ gB) {
  __shared__ float T[GT + 1][GT + 1];
  for (int item = blockIdx.x; item < L.total; item += gridDim.x) {
    const SSBatchEntry e = ss_entry_of(L, item);
    const SSBatchSlice s = ss_uniform_copy(L.slices + e.slice);
    const int n = s.p.g.n;
    const int tiles = (n + 2 + GT - 1) / GT;
    const int local = item - e.item0;
    const int ty = local / tiles, tx = local - ty * tiles;
    ss_gather_tile<GT, NW>(images + s.img_off, gA + s.lay_off, gB + s.lay_off, n, s.p.g.pitch, tx * GT, ty * GT, T,
                           e.which);
    __syncthreads();
  }
}

// to_t: [R][A] -> [A][R] per slice, else [A][R] -> [R][A]
__global__ __launch_bounds__(256) void k_batch_transpose(SSBatchLaunch L, const float* __restrict__ in,
                                                         float* __restrict__ out, int to_t) {
  __shared__ float tile[32][33];
  for (int item = blockIdx.x; item < L.total; item += gridDim.x) {
    const SSBatchEntry e = ss_entry_of(L, item);
    const SSBatchSlice s = ss_uniform_copy(L.slices + e.slice);
    const int rows = to_t ? s.p.g.nrays : s.p.g.nangles;
    const int cols = to_t ? s.p.g.nangles : s.p.g.nrays;
    const int tcols = (cols + 31) / 32;
    const int local = item - e.item0;
    const int ty = local / tcols, tx = local - ty * tcols;
    ss_transpose_tile(in + s.sino_off, out + s.sino_off, rows, cols, tx * 32, ty * 32, tile);
    __syncthreads();
  }
}

// Eight waves per SIMD (four 512-thread blocks per CU, as the per-plan k_radon gets): the
// slice's plan values held in scalar registers would otherwise push the kernel to 87
// VGPRs, two blocks per CU, and a second round of blocks for a 1024^2 OSEM subset.  The
// scalar values that do not fit go to VGPR lanes (no scratch memory).
//
// Item = (row slab, ray group, block of 8 angles) of an entry, slab fastest.  In
// SS_OUT_PARTIAL mode the items write their partial ray sums for k_batch_radon_finish;
// launches in the other modes are never divided into slabs.
template <int MODE>
__global__ __launch_bounds__(RADON_WAVES * SS_WAVE) __attribute__((amdgpu_waves_per_eu(8, 8))) void k_batch_radon(
    SSBatchLaunch L, const float2* __restrict__ gA, const float2* __restrict__ gB, float* __restrict__ part,
    float* __restrict__ out, const float* __restrict__ meas, double* __restrict__ ll) {
  __shared__ SSRadonLds<SSRadonShape<0>::CAP> lds;
  for (int item = blockIdx.x; item < L.total; item += gridDim.x) {
    const SSBatchEntry e = ss_entry_of(L, item);
    const SSBatchSlice* sp = L.slices + e.slice;
    SSRadonView v;
    v.angles = ss_uniform_copy(&sp->p.angles);
    v.bounds = ss_uniform_copy(&sp->p.bounds);
    v.wbounds = ss_uniform_copy(&sp->p.wbounds);
    v.ray_alpha = ss_uniform_copy(&sp->p.ray_alpha);
    v.n = ss_uniform_copy(&sp->p.g.n);
    v.R = ss_uniform_copy(&sp->p.g.nrays);
    v.A = ss_uniform_copy(&sp->p.g.nangles);
    v.pitch = ss_uniform_copy(&sp->p.g.pitch);
    v.nwaves = ss_uniform_copy(&sp->p.nwaves);
    const int64_t lay_off = ss_uniform_copy(&sp->lay_off), sino_off = ss_uniform_copy(&sp->sino_off);
    const int local = item - e.item0;
    constexpr bool divisible = MODE == SS_OUT_PARTIAL;
    const int ns = divisible ? max(e.nslab, 1) : 1;
    const int z = local % ns, r = local / ns;
    const int ablk = r / v.nwaves, grp = r - ablk * v.nwaves;
    const int rows = ns > 1 ? e.slab_rows : v.n;
    float* po = divisible ? part + e.part_off + (size_t)z * (e.j1 - e.j0) * v.R : nullptr;
    ss_radon_item<MODE, 0>(v, gA + lay_off, gB + lay_off, e.j0, e.j1, grp, ablk, z * rows, min(v.n, (z + 1) * rows),
                        out + sino_off, MODE == SS_OUT_EM_RATIO ? meas + sino_off : nullptr,
                        MODE == SS_OUT_EM_RATIO ? ll + item : nullptr, po, lds);
    __syncthreads();
  }
}

template <int MODE>
__global__ __launch_bounds__(256) void k_batch_radon_fixup(SSBatchLaunch L, const float2* __restrict__ gA,
                                                           const float2* __restrict__ gB, float* __restrict__ out,
                                                           const float* __restrict__ meas, double* __restrict__ ll) {
  __shared__ float redf[4];
  for (int item = blockIdx.x; item < L.total; item += gridDim.x) {
    const SSBatchEntry e = ss_entry_of(L, item);
    const SSBatchSlice s = ss_uniform_copy(L.slices + e.slice);
    ss_radon_fixup_item<MODE>(s.p, gA + s.lay_off, gB + s.lay_off, e.j0, e.j1, item - e.item0, out + s.sino_off,
                              MODE == SS_OUT_EM_RATIO ? meas + s.sino_off : nullptr,
                              MODE == SS_OUT_EM_RATIO ? ll + item : nullptr, redf);
    __syncthreads();
  }
}

// Item = 256 consecutive rays [j - j0][k] of an entry: the sum of its nslab partial sums,
// stored as k_batch_radon stores (EM mode: one FP64 log-likelihood partial per item).
template <int MODE>
__global__ __launch_bounds__(256) void k_batch_radon_finish(SSBatchLaunch L, const float* __restrict__ part,
                                                            float* __restrict__ out, const float* __restrict__ meas,
                                                            double* __restrict__ ll) {
  __shared__ double red[4];
  for (int item = blockIdx.x; item < L.total; item += gridDim.x) {
    const SSBatchEntry e = ss_entry_of(L, item);
    const SSBatchSlice* sp = L.slices + e.slice;
    SSRadonView v;
    v.angles = ss_uniform_copy(&sp->p.angles);
    v.R = ss_uniform_copy(&sp->p.g.nrays);
    v.A = ss_uniform_copy(&sp->p.g.nangles);
    const int64_t sino_off = ss_uniform_copy(&sp->sino_off);
    const size_t per = (size_t)(e.j1 - e.j0) * v.R;
    const size_t idx = (size_t)(item - e.item0) * 256 + threadIdx.x;
    double l = 0.0;
    if (idx < per) {
      const int jj = (int)(idx / (unsigned)v.R), k = (int)(idx - (size_t)jj * v.R);
      const float* src = part + e.part_off + idx;
      float sum = 0.f;
      sum = ss_slab_sum(src, per, max(e.nslab, 1));
      ss_radon_store<MODE>(v, e.j0 + jj, k, sum * v.angles[e.j0 + jj].delta, out + sino_off,
                           MODE == SS_OUT_EM_RATIO ? meas + sino_off : nullptr, &l);
    }
    if (MODE == SS_OUT_EM_RATIO) {
      const double tot = ss_block_sum_f64(l, red);
      if (threadIdx.x == 0) ll[item] = tot;
    }
    __syncthreads();
  }
}

template <int MODE, int ROWS, int STAGE>
__global__ __launch_bounds__(256) void k_batch_bp(SSBatchLaunch L, const float* __restrict__ sinoT,
                                                  float* __restrict__ out, float* __restrict__ x,
                                                  const float* __restrict__ weight) {
  __shared__ SSBpLds<STAGE> lds;
  constexpr int TH = 8 * ROWS;
  for (int item = blockIdx.x; item < L.total; item += gridDim.x) {
    const SSBatchEntry e = ss_entry_of(L, item);
    const SSBatchSlice s = ss_uniform_copy(L.slices + e.slice);
    const int tiles_x = (s.p.g.n + BP_TILE_W - 1) / BP_TILE_W;
    const int local = item - e.item0;
    const int ty = local / tiles_x, tx = local - ty * tiles_x;
    ss_bp_item<MODE, ROWS, STAGE>(s.p, sinoT + s.sino_off, e.j0, e.j1, tx * BP_TILE_W, ty * TH,
                                  MODE == SS_BP_PLAIN ? out + s.img_off : nullptr,
                                  MODE == SS_BP_PLAIN ? nullptr : x + s.img_off,
                                  MODE == SS_BP_PLAIN ? nullptr : weight + s.img_off, lds);
    __syncthreads();
  }
}

// resident blocks per CU of each batched kernel, from the runtime's occupancy query
namespace {
template <class K>
int ss_resident_blocks(K kernel, int block) {
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, block, 0) != hipSuccess || per_cu < 1) per_cu = 1;
  int dev = 0, cus = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1)
    cus = 256;
  return per_cu * cus;
}
}  // namespace

// Blocks take items blockIdx.x + i*gridDim.x.  Item costs repeat with the ray-group
// or tile-column count of a slice (often a power of two), so a grid that shares a
// factor with it would hand some blocks only the long central rays; a prime grid
// larger than any such period gives every block the full mix.
static int ss_prime_at_most(int g) {
  for (; g > 2; --g) {
    bool prime = (g % 2) != 0;
    for (int d = 3; prime && d * d <= g; d += 2) prime = (g % d) != 0;
    if (prime) return g;
  }
  return g;
}

int ss_batch_grid(int kind, int variant, int mode, int total) {
  if (total <= 0) return 0;
  // the layout kernels' items are single tiles: one block per item lets the dispatcher
  // overlap their loads and stores across blocks, where a resident grid walks its items
  // one after another (4% of a batched OSEM iteration at 256 slices)
  if (kind == SS_KIND_GATHER || kind == SS_KIND_TRANSPOSE || kind == SS_KIND_RADON_FINISH) return total;
  int g = 0;
  switch (kind) {
    case SS_KIND_RADON:
      g = mode == SS_OUT_EM_RATIO  ? ss_resident_blocks(k_batch_radon<SS_OUT_EM_RATIO>, RADON_WAVES * SS_WAVE)
          : mode == SS_OUT_PARTIAL ? ss_resident_blocks(k_batch_radon<SS_OUT_PARTIAL>, RADON_WAVES * SS_WAVE)
                                   : ss_resident_blocks(k_batch_radon<SS_OUT_SINOT>, RADON_WAVES * SS_WAVE);
      break;
    case SS_KIND_FIXUP:
      g = ss_resident_blocks(k_batch_radon_fixup<SS_OUT_EM_RATIO>, 256);
      break;
    default: {
#define SS_BP_OCC(V, ROWS, STAGE)                                                               \
  case V:                                                                                       \
    g = mode == SS_BP_PLAIN ? ss_resident_blocks(k_batch_bp<SS_BP_PLAIN, ROWS, STAGE>, 256)     \
                            : ss_resident_blocks(k_batch_bp<SS_BP_EM_UPDATE, ROWS, STAGE>, 256); \
    break;
      switch (variant) {
        SS_BP_FOR_VARIANTS(SS_BP_OCC)
        default:
          g = 256;
      }
#undef SS_BP_OCC
    }
  }
  return g < total ? ss_prime_at_most(g) : total;
}

int ss_batch_gather_tile(const int* n, int count) {
  long long big = 0;
  for (int i = 0; i < count; ++i) big += (long long)ss_gather_tiles(n[i]) * ss_gather_tiles(n[i]);
  return big >= 1024 ? SS_GATHER_TILE : 32;  // the per-plan rule: at least 4 large tiles per CU
}

hipError_t ss_launch_batch_gather(const SSBatchLaunch& L, int tile, const float* images, float2* gatherA,
                                  float2* gatherB, hipStream_t stream) {
  if (L.total <= 0) return hipSuccess;
  if (tile == SS_GATHER_TILE)
    hipLaunchKernelGGL((k_batch_gather<SS_GATHER_TILE, 8>), dim3(L.grid), dim3(8 * SS_WAVE), 0, stream, L, images,
                       gatherA, gatherB);
  else
    hipLaunchKernelGGL((k_batch_gather<32, 4>), dim3(L.grid), dim3(4 * SS_WAVE), 0, stream, L, images, gatherA,
                       gatherB);
  return hipGetLastError();
}

hipError_t ss_launch_batch_transpose(const SSBatchLaunch& L, const float* in, float* out, int to_angle_major,
                                     hipStream_t stream) {
  if (L.total <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_batch_transpose, dim3(L.grid), dim3(256), 0, stream, L, in, out, to_angle_major);
  return hipGetLastError();
}

hipError_t ss_launch_batch_radon(const SSBatchLaunch& L, const float2* gatherA, const float2* gatherB, float* part,
                                 int mode, float* out, const float* meas, double* ll_partials, hipStream_t stream) {
  if (L.total <= 0) return hipSuccess;
  const dim3 grid(L.grid), block(RADON_WAVES * SS_WAVE);
  if (part) mode = SS_OUT_PARTIAL;  // the output mode is the finishing pass's
  switch (mode) {
    case SS_OUT_PARTIAL:
      hipLaunchKernelGGL(k_batch_radon<SS_OUT_PARTIAL>, grid, block, 0, stream, L, gatherA, gatherB, part, out, meas,
                         ll_partials);
      break;
    case SS_OUT_SINO:
      hipLaunchKernelGGL(k_batch_radon<SS_OUT_SINO>, grid, block, 0, stream, L, gatherA, gatherB, part, out, meas,
                         ll_partials);
      break;
    case SS_OUT_SINOT:
      hipLaunchKernelGGL(k_batch_radon<SS_OUT_SINOT>, grid, block, 0, stream, L, gatherA, gatherB, part, out, meas,
                         ll_partials);
      break;
    default:
      hipLaunchKernelGGL(k_batch_radon<SS_OUT_EM_RATIO>, grid, block, 0, stream, L, gatherA, gatherB, part, out, meas,
                         ll_partials);
  }
  return hipGetLastError();
}

hipError_t ss_launch_batch_radon_finish(const SSBatchLaunch& L, const float* part, int mode, float* out,
                                        const float* meas, double* ll_partials, hipStream_t stream) {
  if (L.total <= 0) return hipSuccess;
  const dim3 grid(L.grid), block(256);
  switch (mode) {
    case SS_OUT_SINO:
      hipLaunchKernelGGL(k_batch_radon_finish<SS_OUT_SINO>, grid, block, 0, stream, L, part, out, meas, ll_partials);
      break;
    case SS_OUT_SINOT:
      hipLaunchKernelGGL(k_batch_radon_finish<SS_OUT_SINOT>, grid, block, 0, stream, L, part, out, meas, ll_partials);
      break;
    default:
      hipLaunchKernelGGL(k_batch_radon_finish<SS_OUT_EM_RATIO>, grid, block, 0, stream, L, part, out, meas,
                         ll_partials);
  }
  return hipGetLastError();
}

hipError_t ss_launch_batch_radon_fixup(const SSBatchLaunch& L, const float2* gatherA, const float2* gatherB, int mode,
                                       float* out, const float* meas, double* ll_partials, hipStream_t stream) {
  if (L.total <= 0) return hipSuccess;
  const dim3 grid(L.grid), block(256);
  switch (mode) {
    case SS_OUT_SINO:
      hipLaunchKernelGGL(k_batch_radon_fixup<SS_OUT_SINO>, grid, block, 0, stream, L, gatherA, gatherB, out, meas,
                         ll_partials);
      break;
    case SS_OUT_SINOT:
      hipLaunchKernelGGL(k_batch_radon_fixup<SS_OUT_SINOT>, grid, block, 0, stream, L, gatherA, gatherB, out, meas,
                         ll_partials);
      break;
    default:
      hipLaunchKernelGGL(k_batch_radon_fixup<SS_OUT_EM_RATIO>, grid, block, 0, stream, L, gatherA, gatherB, out,
                         meas, ll_partials);
  }
  return hipGetLastError();
}

hipError_t ss_launch_batch_backprojection(const SSBatchLaunch& L, int variant, const float* sinoT, int mode,
                                          float* out, float* x, const float* weight, hipStream_t stream) {
  if (L.total <= 0) return hipSuccess;
  const dim3 grid(L.grid), block(256);
#define SS_BP_BATCH(V, ROWS, STAGE)                                                                            \
  case V:                                                                                                      \
    if (mode == SS_BP_PLAIN)                                                                                   \
      hipLaunchKernelGGL((k_batch_bp<SS_BP_PLAIN, ROWS, STAGE>), grid, block, 0, stream, L, sinoT, out, x,     \
                         weight);                                                                              \
    else                                                                                                       \
      hipLaunchKernelGGL((k_batch_bp<SS_BP_EM_UPDATE, ROWS, STAGE>), grid, block, 0, stream, L, sinoT, out, x, \
                         weight);                                                                              \
    break;
  switch (variant) {
    SS_BP_FOR_VARIANTS(SS_BP_BATCH)
    default:
      return hipErrorInvalidValue;
  }
#undef SS_BP_BATCH
  return hipGetLastError();
}

// one block per slice: the slice's partial ranges summed in a fixed order
__global__ __launch_bounds__(256) void k_batch_ll(const int2* __restrict__ ranges, int nranges,
                                                  const double* __restrict__ partials, double* __restrict__ ll_out) {
  __shared__ double red[4];
  const size_t slice = blockIdx.x;
  double v = 0.0;
  for (int r = 0; r < nranges; ++r) {
    const int2 rg = ranges[slice * nranges + r];
    for (int i = threadIdx.x; i < rg.y; i += 256) v += partials[(size_t)rg.x + i];
  }
  const double tot = ss_block_sum_f64(v, red);
  if (threadIdx.x == 0) ll_out[slice] = tot;
}

hipError_t ss_launch_batch_ll(const int2* ranges, int nranges, int nslices, const double* partials, double* ll_out,
                              hipStream_t stream) {
  if (nslices <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_batch_ll, dim3(nslices), dim3(256), 0, stream, ranges, nranges, partials, ll_out);
  return hipGetLastError();
}

__global__ __launch_bounds__(256) void k_fill_f64(double* __restrict__ p, double v, size_t n) {
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) p[i] = v;
}

hipError_t ss_launch_fill_f64(double* p, double v, size_t n, hipStream_t stream) {
  size_t blocks = (n + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  if (blocks == 0) return hipSuccess;
  hipLaunchKernelGGL(k_fill_f64, dim3((unsigned)blocks), dim3(256), 0, stream, p, v, n);
  return hipGetLastError();
}

// =====================================================================================
// ellipse phantoms, FP64 (raft/raft_image_phantom_fromdesc.c:4-59, raft/raft_radon_fromdesc.c:4-67)
// =====================================================================================

__global__ __launch_bounds__(256) void k_phantom(const SSEllipse* __restrict__ ell, int nell, int lines, int cols,
                                                 double tl_x, double tl_y, double hs, double vs,
                                                 double* __restrict__ out64, float* __restrict__ out32) {
  SS_NOCONTRACT
  const size_t idx = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (idx >= (size_t)lines * cols) return;
  const int i = (int)(idx / cols), j = (int)(idx % cols);
  double acc = out64 ? out64[idx] : 0.0;
  for (int e = 0; e < nell; ++e) {
    const SSEllipse E = ell[e];
    const double x = tl_x + j * hs - E.x0;
    const double y = tl_y + i * vs - E.y0;
    double pv = (y * E.s) + (x * E.c);
    pv /= E.a;
    pv *= pv;
    double qv = (y * E.c) - (x * E.s);
    qv /= E.b;
    qv *= qv;
    pv += qv;
    pv = sqrt(pv);
    if (pv <= 1) acc += E.val;
  }
  if (out64) out64[idx] = acc;
  if (out32) out32[idx] = (float)acc;
}

hipError_t ss_launch_phantom(const SSEllipse* ell, int nell, int lines, int cols, double tl_x, double tl_y,
                             double hs, double vs, double* out64, float* out32, hipStream_t stream) {
  const size_t total = (size_t)lines * cols;
  hipLaunchKernelGGL(k_phantom, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, stream, ell, nell, lines, cols,
                     tl_x, tl_y, hs, vs, out64, out32);
  return hipGetLastError();
}

__global__ __launch_bounds__(256) void k_radon_fromdesc(const SSEllipseView* __restrict__ views, int nell,
                                                        int lines, int cols, double* __restrict__ out64) {
  SS_NOCONTRACT
  const size_t idx = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (idx >= (size_t)lines * cols) return;
  const int i = (int)(idx / cols), j = (int)(idx % cols);
  double acc = out64[idx];
  for (int e = 0; e < nell; ++e) {
    const SSEllipseView v = views[(size_t)e * cols + j];
    double tp = v.mab * i - v.delt;
    tp *= tp;
    if (tp <= 1.0) acc += v.m * sqrt(1.0 - tp);
  }
  out64[idx] = acc;
}

hipError_t ss_launch_radon_fromdesc(const SSEllipseView* views, int nell, int lines, int cols, double* out64,
                                    hipStream_t stream) {
  const size_t total = (size_t)lines * cols;
  hipLaunchKernelGGL(k_radon_fromdesc, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, stream, views, nell,
                     lines, cols, out64);
  return hipGetLastError();
}
