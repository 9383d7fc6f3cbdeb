// dda_kernels.hip -- line-traversal (Bresenham / DDA) projector pair on gfx950.
//
// One lane per ray; the 64 lanes of a wave are 64 consecutive rays of one angle, so at
// each fast move they touch one line of pixels of the (fast, slow) layout chosen for
// the angle.  The move sequence is decided in FP64 with the reference's operation
// order (raft/raft_ddaiterator.c), so every ray visits the reference's pixels; taps
// are FP32, the Radon's ray sums FP64.
//
//   k_dda_radon      gather: sino[r][a] = sum over the ray's pixels of f(pixel) * len
//   k_dda_scatter    scatter: acc(pixel) += q * len with atomics (FP32 for R^T, FP64 for
//                    the normalised backprojection and its weights R^T 1)
//   k_dda_fold       out = acc + accT^T (optionally / weights) through an LDS tile
//   k_dda_transpose_image  the column-major copy of the image the Radon reads for
//                          angles whose fast axis is the column
#include "dda_kernels.h"

#include <hip/hip_runtime.h>

#define DDA_TILE 32

__device__ __forceinline__ bool dda_inside(int f, int s, int nf, int ns) {
  return (unsigned)f < (unsigned)nf && (unsigned)s < (unsigned)ns;
}

__device__ __forceinline__ double dda_ray_t(const SSDdaGeom& g, int r) {
  DDA_NOCONTRACT
  return g.t0 + (r * g.dt);
}

// ---------------------------------------------------------------- Radon (gather)

// The wave's fast index at which its first ray enters the image (all lanes share the
// angle, so the direction dfast): lanes wait until the wave's index reaches theirs and
// then advance with it, so at every step the active lanes sit on one line of the
// layout and their 64 taps fall in a few cache lines.
__device__ __forceinline__ int dda_wave_start(bool alive, int f, int dfast) {
  int key = alive ? (dfast > 0 ? f : -f) : 0x7fffffff;
  for (int off = 32; off > 0; off >>= 1) key = min(key, __shfl_xor(key, off));
  return dfast > 0 ? key : -key;
}

#ifndef DDA_RADON_U
#define DDA_RADON_U 8
#endif
#define DDA_RSRC_DWORD3 0x00020000  // gfx9 buffer word 3: 32-bit data format, no swizzle

// Exact-geometry angles in trapezoid form (SSDdaAngle::tr_*): the ray's length in the
// pixel at fast index k, slow index j is lmax z(|sigma_k - j|), and at most two slow
// indices per k have z > 0.  Each lane carries tau = sigma - half_s + DDA_PADW + 1 in Q32
// (integer part, 32-bit fraction), stepped by tr_d per fast index with an add and a
// carry; per k it adds q0 z0 + q1 z1 of its candidates (padded slow indices floor(tau),
// floor(tau) + 1), z from the FP32 fraction.  The wave runs the union of its lanes' fast
// ranges where a candidate can be inside the image; rows are summed in FP32 per block of
// rows, the blocks in FP64.  src is the slice's padded image (dda_pad_pitch lines).
//
// Wave-staged form (tr_wstage): per chunk of DDA_WS_ROWS fast indices the wave copies,
// for each, the DDA_WS_SEG padded pixels from one chunk-wide start sb into its own LDS
// rows, two fast indices per 16-byte buffer-to-LDS load (lanes 0-31 one, 32-63 the
// next), and then reads each lane's candidates from there with LDS reads: half the
// loads of the direct form, whose one 8-byte load per ray and fast index keeps the
// texture addresser busy.  sb is the lowest ray's floor(tau) at the chunk's first index,
// less 1 and the chunk's possible fall (tr_wgap), aligned down to 4: one scalar per
// chunk, stepped in Q32 by tr_qchunk, so the copies take a fixed lane offset and a
// scalar row offset.  Every lane's candidates on the chunk's rows then lie in
// [sb, sb + spread + (DDA_WS_ROWS - 1) |tr_d| + 9), inside the segment, and the DDA_PADW
// padding keeps the segment inside its line for every fast index of the wave's range.
// Two chunk buffers per wave: the next chunk's copies are in flight while the wave sums
// the current one.  Otherwise the direct form: one 8-byte buffer load per ray and fast
// index, the index clamped to the padded line (its padding makes the clamp exact).
//
// s_waitcnt immediates (gfx9 encoding: vmcnt bits 3:0 and 15:14, expcnt 6:4, lgkmcnt
// 11:8; a counter at its maximum is not waited on)
#define DDA_WAIT_VM0 0x0F70
#define DDA_WAIT_LGKM0 0xC07F
#define DDA_WAIT_VM_CHUNK (0x0F70 | (DDA_WS_ROWS / 2))  // vmcnt(DDA_WS_ROWS / 2), below 16
typedef float dda_f2v __attribute__((ext_vector_type(2)));
typedef __attribute__((address_space(3))) void* dda_lds_ptr;
typedef __attribute__((address_space(3))) const float* dda_lds_fptr;
__device__ dda_f2v dda_raw_load_f2(__amdgpu_buffer_rsrc_t rsrc, int voffset, int soffset,
                                   int aux) __asm("llvm.amdgcn.raw.ptr.buffer.load.v2f32");

struct SSDdaTrapState {
  int hi;
  unsigned lo;
};

__device__ __forceinline__ void dda_trap_step(const SSDdaAngle& a, SSDdaTrapState& st) {
  const unsigned lo = st.lo + a.tr_dlo;
  st.hi += a.tr_dhi + (lo < st.lo ? 1 : 0);
  st.lo = lo;
}

__device__ __forceinline__ float dda_trap_term(const SSDdaAngle& a, float x0, float x1, unsigned lo, float acc) {
  const float fr = (float)lo;
  const float z0 = __builtin_amdgcn_fmed3f(fmaf(-fabsf(a.tr_c1 - fr), a.tr_kz, a.tr_pz), 0.f, 1.f);
  const float z1 = __builtin_amdgcn_fmed3f(fmaf(fr, a.tr_kz, a.tr_pz1), 0.f, 1.f);
  return fmaf(x1, z1, fmaf(x0, z0, acc));
}

__device__ __forceinline__ SSDdaTrapState dda_trap_start(const SSDdaAngle& a, double tau0, int k) {
  const double v = fmin(fmax(fma((double)k, a.tr_d, tau0), -1073741824.0), 1073741824.0);
  const double fv = floor(v);
  SSDdaTrapState st;
  st.hi = (int)fv;
  st.lo = (unsigned)((v - fv) * 4294967296.0);
  return st;
}

// the wave's fast range: indices where floor(tau) can be in [DDA_PADW - 1,
// ns + DDA_PADW - 1] (a candidate inside the image) for some lane, one index of margin
__device__ __forceinline__ int2 dda_trap_range(const SSDdaAngle& a, double tau0, bool lane_ok, int nf, int ns) {
  const double ka = (DDA_PADW - 1.0 - tau0) / a.tr_d, kb = (ns + DDA_PADW - tau0) / a.tr_d;
  const double kl = fmax(fmin(ka, kb), -1.0), kh = fmin(fmax(ka, kb), (double)nf);
  int klo = nf, khi = -1;
  if (lane_ok && kl <= kh) {  // false for NaN
    klo = (int)fmax(0.0, floor(kl) - 1.0);
    khi = (int)fmin(nf - 1.0, ceil(kh) + 1.0);
  }
  for (int off = 32; off > 0; off >>= 1) {
    klo = min(klo, __shfl_xor(klo, off));
    khi = max(khi, __shfl_xor(khi, off));
  }
  // wave-uniform: rows in scalar registers
  return make_int2(__builtin_amdgcn_readfirstlane(klo), __builtin_amdgcn_readfirstlane(khi));
}

template <int N>
__device__ __forceinline__ float dda_trap_rows(const SSDdaAngle& a, __amdgpu_buffer_rsrc_t rs, int row_bytes,
                                               int pitch_bytes, int smax, SSDdaTrapState& st) {
  float blk = 0.f;
#pragma unroll
  for (int u = 0; u < N; ++u) {
    // below 0 wraps to a large unsigned value and clamps to smax: zeros there too
    const unsigned idx = min((unsigned)st.hi, (unsigned)smax);
    const dda_f2v q = dda_raw_load_f2(rs, (int)(idx * sizeof(float)), row_bytes + u * pitch_bytes, 0);
    blk = dda_trap_term(a, q[0], q[1], st.lo, blk);
    dda_trap_step(a, st);
  }
  return blk;
}

__device__ __forceinline__ double dda_radon_trap(const SSDdaAngle& a, const float* __restrict__ src, double t,
                                                 bool lane_ok, int nf, int ns) {
  const double tau0 = fma(t, a.tr_inv, a.tr_off);
  const int2 kr = dda_trap_range(a, tau0, lane_ok, nf, ns);
  if (kr.x > kr.y) return 0.0;
  SSDdaTrapState st = dda_trap_start(a, tau0, kr.x);
  const int pitch = dda_pad_pitch(ns);
  const int pitch_bytes = pitch * (int)sizeof(float);
  const __amdgpu_buffer_rsrc_t rs =
      __builtin_amdgcn_make_buffer_rsrc((void*)src, (short)0, nf * pitch_bytes, DDA_RSRC_DWORD3);
  double tot = 0.0;
  int k = kr.x;
  for (; k + DDA_RADON_U - 1 <= kr.y; k += DDA_RADON_U)
    tot += (double)dda_trap_rows<DDA_RADON_U>(a, rs, k * pitch_bytes, pitch_bytes, pitch - 2, st);
  for (; k <= kr.y; ++k) tot += (double)dda_trap_rows<1>(a, rs, k * pitch_bytes, pitch_bytes, pitch - 2, st);
  return lane_ok ? tot * a.tr_lmax : 0.0;
}

// copies chunk rows k0 .. k0 + nrow - 1 of the segments from sb into buf (DDA_WS_ROWS
// rows of DDA_WS_SEG floats), DDA_WS_ROWS / 2 loads
__device__ __forceinline__ void dda_ws_copy(__amdgpu_buffer_rsrc_t rs, float* buf, int k0, int sb, int nrow,
                                            int pitch_bytes, int l16, int upper) {
  const int chunk_bytes = k0 * pitch_bytes + sb * (int)sizeof(float);
  if (nrow == DDA_WS_ROWS) {
    // a fixed lane offset, the row pair as a scalar offset
    const int lane_off = l16 + upper * pitch_bytes;
#pragma unroll
    for (int i = 0; i < DDA_WS_ROWS / 2; ++i) {
      int off = lane_off;
      __asm__ volatile("" : "+v"(off));
      // the scalar offset through readfirstlane: a uniform value that instruction
      // selection placed in a vector register would otherwise become a waterfall loop
      __builtin_amdgcn_raw_ptr_buffer_load_lds(rs, (dda_lds_ptr)(buf + 2 * i * DDA_WS_SEG), 16, off,
                                               __builtin_amdgcn_readfirstlane(chunk_bytes + 2 * i * pitch_bytes), 0,
                                               0);
    }
  } else {
#pragma unroll
    for (int i = 0; i < DDA_WS_ROWS / 2; ++i) {
      // rows past the range repeat the last one (inside the image, never read); the whole
      // offset per lane, so that no scalar offset is built from lane values
      int off = chunk_bytes + l16 + min(2 * i + upper, nrow - 1) * pitch_bytes;
      __asm__ volatile("" : "+v"(off));
      __builtin_amdgcn_raw_ptr_buffer_load_lds(rs, (dda_lds_ptr)(buf + 2 * i * DDA_WS_SEG), 16, off, 0, 0, 0);
    }
  }
}

// the chunk's segment start from the lowest ray's Q32 state ql at its first index
__device__ __forceinline__ int dda_ws_start(const SSDdaAngle& a, long long ql) {
  return ((int)(ql >> 32) - 1 - a.tr_wgap) & ~3;
}

// wlds: this wave's two chunk buffers in LDS, DDA_WS_ROWS x DDA_WS_SEG floats each (the
// next chunk's copies are in flight while the wave sums the current one).  Every lane,
// also one past the last ray, reads inside its buffer: the lanes' tau is linear in the
// ray index, so each lies between the lowest lane's and that plus the wave's spread.
__device__ __forceinline__ double dda_radon_trap_ws(const SSDdaAngle& a, const float* __restrict__ src, double t,
                                                    bool lane_ok, int nf, int ns, float* wlds) {
  const double tau0 = fma(t, a.tr_inv, a.tr_off);
  const int2 kr = dda_trap_range(a, tau0, lane_ok, nf, ns);
  if (kr.x > kr.y) return 0.0;
  SSDdaTrapState st = dda_trap_start(a, tau0, kr.x);
  const int lane = threadIdx.x & 63;
  const int pitch_bytes = dda_pad_pitch(ns) * (int)sizeof(float);
  const __amdgpu_buffer_rsrc_t rs =
      __builtin_amdgcn_make_buffer_rsrc((void*)src, (short)0, nf * pitch_bytes, DDA_RSRC_DWORD3);
  const int l16 = (lane & 31) * 16;  // this lane's 16 bytes of a staged segment
  const int upper = lane >> 5;       // 1: the second fast index of a load
  // the lowest ray's Q32 state (scalar) at the first index of the next chunk to copy
  long long ql = (long long)(((unsigned long long)(unsigned)__builtin_amdgcn_readlane(st.hi, a.tr_lowlane) << 32) |
                             (unsigned)__builtin_amdgcn_readlane((int)st.lo, a.tr_lowlane));
  const int nchunk = (kr.y - kr.x) / DDA_WS_ROWS + 1;
  int sb = dda_ws_start(a, ql);
  dda_ws_copy(rs, wlds, kr.x, sb, min(DDA_WS_ROWS, kr.y - kr.x + 1), pitch_bytes, l16, upper);
  double tot = 0.0;
  for (int c = 0; c < nchunk; ++c) {
    const int k0 = kr.x + c * DDA_WS_ROWS;
    const int nrow = min(DDA_WS_ROWS, kr.y - k0 + 1);
    const float* __restrict__ cur = wlds + (c & 1) * (DDA_WS_ROWS * DDA_WS_SEG);
    int sb_next = 0;
    if (c + 1 < nchunk) {
      ql += a.tr_qchunk;
      sb_next = dda_ws_start(a, ql);
      // the other buffer's LDS reads (chunk c - 1) have returned before it is overwritten
      __builtin_amdgcn_s_waitcnt(DDA_WAIT_LGKM0);
      dda_ws_copy(rs, wlds + ((c + 1) & 1) * (DDA_WS_ROWS * DDA_WS_SEG), k0 + DDA_WS_ROWS, sb_next,
                  min(DDA_WS_ROWS, kr.y - k0 - DDA_WS_ROWS + 1), pitch_bytes, l16, upper);
      // buffer-to-LDS copies count on vmcnt, in order, and no barrier orders them here:
      // this chunk's have landed once at most the next chunk's are outstanding
      __builtin_amdgcn_s_waitcnt(DDA_WAIT_VM_CHUNK);
    } else {
      __builtin_amdgcn_s_waitcnt(DDA_WAIT_VM0);
    }
    // LDS byte address of a lane's first candidate on row u: a scalar row base plus 4 hi
    const int base = __builtin_amdgcn_readfirstlane((int)(size_t)(dda_lds_fptr)cur) - 4 * sb;
    float blk = 0.f;
    if (nrow == DDA_WS_ROWS) {
      // wave-uniform: a full chunk steps its rows, issues its reads together, then sums
      unsigned lo[DDA_WS_ROWS];
      float x0[DDA_WS_ROWS], x1[DDA_WS_ROWS];
#pragma unroll
      for (int u = 0; u < DDA_WS_ROWS; ++u) {
        const dda_lds_fptr q = (dda_lds_fptr)(size_t)(unsigned)(__builtin_amdgcn_readfirstlane(base + u * DDA_WS_SEG * 4) +
                                                                 4 * st.hi);
        x0[u] = q[0];
        x1[u] = q[1];
        lo[u] = st.lo;
        dda_trap_step(a, st);
      }
#pragma unroll
      for (int u = 0; u < DDA_WS_ROWS; ++u) blk = dda_trap_term(a, x0[u], x1[u], lo[u], blk);
    } else {
      for (int u = 0; u < nrow; ++u) {
        const dda_lds_fptr q = (dda_lds_fptr)(size_t)(unsigned)(base + u * DDA_WS_SEG * 4 + 4 * st.hi);
        blk = dda_trap_term(a, q[0], q[1], st.lo, blk);
        dda_trap_step(a, st);
      }
    }
    tot += (double)blk;
    sb = sb_next;
  }
  return lane_ok ? tot * a.tr_lmax : 0.0;
}

__global__ __launch_bounds__(256) void k_dda_radon(SSDdaGeom g, const SSDdaAngle* __restrict__ angles,
                                                   const float* __restrict__ img, const float* __restrict__ imgT,
                                                   float* __restrict__ sino) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  const int ai = blockIdx.y;
  const size_t slice = blockIdx.z;
  const bool lane_ok = r < g.nrays;
  const SSDdaAngle a = angles[ai];
  const int nf = a.fast_i ? g.lines : g.cols;
  const int ns = a.fast_i ? g.cols : g.lines;
  // padded layouts: DDA_PADW zero columns either side of each line (dda_pad_pitch)
  const int pitch = dda_pad_pitch(ns);
  const float* __restrict__ src = (a.fast_i ? img : imgT) + slice * (size_t)nf * pitch;
  if (a.fixed && (long long)nf * pitch < (1LL << 29)) {
    __shared__ float wlds[4][2 * DDA_WS_ROWS * DDA_WS_SEG];  // two chunk buffers per wave
    // every lane's own ray (past R too): the staged form takes the wave's lowest ray from
    // its lanes
    const double t = dda_ray_t(g, r);
    const double acc = a.tr_wstage ? dda_radon_trap_ws(a, src, t, lane_ok, nf, ns, wlds[threadIdx.x >> 6])
                                   : dda_radon_trap(a, src, t, lane_ok, nf, ns);
    if (lane_ok) sino[(slice * g.nrays + r) * g.nangles + ai] = (float)acc;
    return;
  }
  SSDdaRay ray = dda_ray_start(g, a, dda_ray_t(g, lane_ok ? r : 0));
  bool alive = lane_ok && dda_inside(ray.f, ray.s, nf, ns);
  // FP64 sum: the lengths are FP64 already, so a tap costs one convert and one FMA as
  // in FP32, without FP32's N*eps growth over a ray of N equal terms
  double acc = 0.0;
  int F = dda_wave_start(alive, ray.f, a.dfast);
  // the pixel's offset in the (fast, slow) image, stepped with the ray (an image side is
  // at most 16384, so a slice's offsets fit an int); meaningful while the lane is alive
  int pix = alive ? ray.f * pitch + ray.s + DDA_PADW : 0;
  {
    const int fast_step = a.dfast * pitch;
    // lanes start at most nf moves apart and then take at most nf + 1 moves each
    for (int it = 0; it <= 2 * nf + 1; ++it) {
      if (__ballot(alive) == 0) break;
      if (alive && ray.f == F) {
        acc = fma((double)src[pix], ray.len, acc);
        if (ray.slow_next) {
          ray.s = dda_step(ray.s, a.dslow);
          pix += a.dslow;
          ray.len = a.len_full - ray.len;
          ray.slow_next = 0;
          alive = dda_inside(ray.f, ray.s, nf, ns);
          if (alive) acc = fma((double)src[pix], ray.len, acc);
        }
        if (alive) {
          dda_fast_move(a, ray);
          pix += fast_step;
          alive = dda_inside(ray.f, ray.s, nf, ns);
        }
      }
      F += a.dfast;
    }
  }
  if (lane_ok) sino[(slice * g.nrays + r) * g.nangles + ai] = (float)acc;
}

// ---------------------------------------------------------------- transpose (scatter)

// acc: pixels of angles whose fast axis is the line ([lines][cols]); accT: the others
// ([cols][lines]).  ones: every ray carries 1 (the normalisation R^T 1).  T = float for
// the transpose, double for the backprojection, whose normalisation can cancel (the
// reference's non-square-pixel traversal yields negative lengths).
template <typename T>
__global__ __launch_bounds__(256) void k_dda_scatter(SSDdaGeom g, const SSDdaAngle* __restrict__ angles,
                                                     const int* __restrict__ walk, const float* __restrict__ sino,
                                                     int ones, T* __restrict__ acc, T* __restrict__ accT) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  const int ai = walk[blockIdx.y];
  const size_t slice = blockIdx.z;
  if (r >= g.nrays) return;
  const T q = ones ? T(1) : T(sino[(slice * g.nrays + r) * g.nangles + ai]);
  if (q == T(0)) return;  // adds nothing
  const SSDdaAngle a = angles[ai];
  const int nf = a.fast_i ? g.lines : g.cols;
  const int ns = a.fast_i ? g.cols : g.lines;
  T* dst = (a.fast_i ? acc : accT) + slice * (size_t)g.lines * g.cols;
  SSDdaRay ray = dda_ray_start(g, a, dda_ray_t(g, r));
  for (int it = 0; it <= nf; ++it) {
    if (!dda_inside(ray.f, ray.s, nf, ns)) break;
    unsafeAtomicAdd(dst + (size_t)ray.f * ns + ray.s, q * T(ray.len));
    if (ray.slow_next) {
      ray.s = dda_step(ray.s, a.dslow);
      ray.len = a.len_full - ray.len;
      ray.slow_next = 0;
      if (!dda_inside(ray.f, ray.s, nf, ns)) break;
      unsafeAtomicAdd(dst + (size_t)ray.f * ns + ray.s, q * T(ray.len));
    }
    dda_fast_move(a, ray);
  }
}

// ---------------------------------------------------------------- transpose (gather)

#define DDA_GX 64     // image columns per block (one per lane)
#define DDA_GROWS 4   // image lines per thread
#define DDA_GY (4 * DDA_GROWS)

// acc[i][j] = sum over the table's angles and the rays crossing pixel (i, j) of
// q(r, a) * L(t_r - t_c) (trapezoid lengths, FP64 geometry and sums).  Overwrites acc.
// rows: the angle-major sinogram, each row scaled by its angle's lmax and padded with
// two zeros either side (row a starts at rows + a * (R + 4) + 2), so the K = 2 path reads
// its candidates r0, r0 + 1 in [-2, R + 1] without bounds tests.  K = 0: any number of
// candidates (A.k), with bounds tests.  ones: q = 1 (the normalisation R^T 1).
template <typename T, int K>
__global__ __launch_bounds__(256) void k_dda_gather(SSDdaGeom g, const SSDdaGatherAngle* __restrict__ ga, int nga,
                                                    const float* __restrict__ rows, int ones, T* __restrict__ acc) {
  const int j = blockIdx.x * DDA_GX + (threadIdx.x & (DDA_GX - 1));
  const int i0 = blockIdx.y * DDA_GY + (threadIdx.x / DDA_GX) * DDA_GROWS;
  const size_t slice = blockIdx.z;
  if (j >= g.cols || i0 >= g.lines) return;
  const double x = g.ox + j * g.hx;
  const double y0 = g.oy + i0 * g.hy;
  const double inv_dt = 1.0 / g.dt;
  const int pitch = g.nrays + 4;
  const float* __restrict__ srows = rows + slice * (size_t)g.nangles * pitch + 2;
  double sum[DDA_GROWS];
#pragma unroll
  for (int rr = 0; rr < DDA_GROWS; ++rr) sum[rr] = 0.0;
  for (int k = 0; k < nga; ++k) {
    const SSDdaGatherAngle A = ga[k];
    const float* __restrict__ q = srows + (size_t)A.a * pitch;
    double tc = fma(x, A.c, y0 * A.s);
    const double p_inv = A.p * A.inv_pq;
#pragma unroll
    for (int rr = 0; rr < DDA_GROWS; ++rr) {
      const double rc = (tc - g.t0) * inv_dt;
      const double lo = fmin(fmax(floor(rc - A.half) + 1.0, -2.0), (double)g.nrays);
      const int r0 = (int)lo;
      if (K == 2) {
        // d = t_r - t_c for r0 and r0 + 1; z = clamp((p - |d|) / (p - q))
        const double d0 = fma(lo, g.dt, g.t0) - tc;
        const double d1 = d0 + g.dt;
        const double z0 = fmin(1.0, fmax(0.0, fma(-fabs(d0), A.inv_pq, p_inv)));
        const double z1 = fmin(1.0, fmax(0.0, fma(-fabs(d1), A.inv_pq, p_inv)));
        if (ones) {  // rays that exist only
          const double w0 = (unsigned)r0 < (unsigned)g.nrays ? z0 : 0.0;
          const double w1 = (unsigned)(r0 + 1) < (unsigned)g.nrays ? z1 : 0.0;
          sum[rr] += A.lmax * (w0 + w1);
        } else {
          sum[rr] = fma((double)q[r0], z0, sum[rr]);
          sum[rr] = fma((double)q[r0 + 1], z1, sum[rr]);
        }
      } else {
        for (int m = 0; m < A.k; ++m) {
          const int r = r0 + m;
          if ((unsigned)r < (unsigned)g.nrays) {
            const double d = fma((double)r, g.dt, g.t0) - tc;
            const double z = fmin(1.0, fmax(0.0, fma(-fabs(d), A.inv_pq, p_inv)));
            sum[rr] = fma(ones ? A.lmax : (double)q[r], z, sum[rr]);
          }
        }
      }
      tc += A.hys;
    }
  }
  T* __restrict__ out = acc + slice * (size_t)g.lines * g.cols;
#pragma unroll
  for (int rr = 0; rr < DDA_GROWS; ++rr)
    if (i0 + rr < g.lines) out[(size_t)(i0 + rr) * g.cols + j] = T(sum[rr]);
}

// The two-candidate gather of a sinogram (k <= 2 for every angle) in FP32 ray units.
// A block covers DDA_G2_COLS columns x 4 G lines of pixels and takes the
// angles in groups of DDA_G2_ANGLES: each wave stages, for a quarter of the group's
// angles, the segment of the angle's (lmax-scaled, zero-padded) row that the block's
// candidate rays span into LDS (rays rb .. rb + ws - 1 with rb = floor(min over the block
// of rc - half) - 2; the resource spans the padded row, so rays off the row read 0),
// or marks the angle as missing the block.  Then per (thread, angle) FP64 gives
// rc - half - rb at the thread's first line and splits it into an integer base and an
// FP32 fraction g0 in [0, 1); each of the thread's G lines takes
// fl = floor(g), fr = fract(g) of g = g0 + rr ky, reads q[r0], q[r0 + 1]
// (r0 = rb + base + fl + 1) from the segment with one two-word LDS read and adds
// q0 z0 + q1 z1 with the angle's hat weights.  The segment index base + fl + 1 lies in
// [1, 6 + span] for span = (cols - 1)|kx| + (lines - 1)|ky| of the block (the FP32
// fraction is within 1e-5 ray units of the FP64 value, so floor is off by at most one),
// and ws >= span + 8.  Terms are summed in FP32 over a group, those sums in FP64.
// s_waitcnt immediates (gfx9 encoding: vmcnt bits 3:0 and 15:14, expcnt 6:4, lgkmcnt
// 11:8; a counter at its maximum is not waited on)
#define DDA_WAIT_VM0 0x0F70
#define DDA_WAIT_LGKM0 0xC07F
#define DDA_WAIT_VM_CHUNK (0x0F70 | (DDA_WS_ROWS / 2))  // vmcnt(DDA_WS_ROWS / 2), below 16
typedef float dda_f2v __attribute__((ext_vector_type(2)));
__device__ float dda_raw_load_f32(__amdgpu_buffer_rsrc_t rsrc, int voffset, int soffset,
                                  int aux) __asm("llvm.amdgcn.raw.ptr.buffer.load.f32");

__device__ __forceinline__ int dda_floor_i32(float v) {
  int r;
  __asm__("v_cvt_flr_i32_f32 %0, %1" : "=v"(r) : "v"(v));
  return r;
}

#define DDA_G2_MISS (-0x7fffffff - 1)
typedef __attribute__((address_space(3))) float dda_lds_float;

template <typename T, int G>
__global__ __launch_bounds__(256) void k_dda_gather2(SSDdaGeom g, const SSDdaGatherAngle* __restrict__ ga, int nga,
                                                     const float* __restrict__ rows, int ws, T* __restrict__ acc) {
  constexpr int NA = DDA_G2_ANGLES;
  __shared__ float seg[NA * DDA_G2_WSMAX];
  __shared__ int seg_rb[NA];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int jb = blockIdx.x * DDA_G2_COLS, ib = blockIdx.y * (4 * G);
  const int j = jb + lane;
  const int i0 = ib + wave * G;
  const size_t slice = blockIdx.z;
  const double jd = j, id = i0;
  const int pitch = g.nrays + 4;
  const float* __restrict__ srows = rows + slice * (size_t)g.nangles * pitch;
  double tot[G];
#pragma unroll
  for (int rr = 0; rr < G; ++rr) tot[rr] = 0.0;
  for (int k0 = 0; k0 < nga; k0 += NA) {
    const int na = min(NA, nga - k0);
    __syncthreads();  // the previous group's segments are read
    for (int t = wave; t < na; t += 4) {
      const SSDdaGatherAngle& A = ga[k0 + t];
      const double lo = fma((double)jb, A.kx, fma((double)ib, A.ky, A.k0h)) + fmin(0.0, (DDA_G2_COLS - 1) * A.kx) +
                        fmin(0.0, (4 * G - 1) * A.ky);
      const double hi = lo + (DDA_G2_COLS - 1) * fabs(A.kx) + (4 * G - 1) * fabs(A.ky);
      // candidates floor(lo) + 1 .. floor(hi) + 2 meet the rays 0 .. R - 1
      const bool meets = hi >= -4.0 && lo <= g.nrays + 4.0;
      const int rb = meets ? (int)floor(lo) - 2 : DDA_G2_MISS;
      if (lane == 0) seg_rb[t] = rb;
      if (meets) {
        const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
            (void*)(srows + (size_t)A.a * pitch), (short)0, pitch * (int)sizeof(float), DDA_RSRC_DWORD3);
        // byte offset of padded element rb + 2 (ray rb), opaque so that no constant part
        // moves into the load's instruction offset: the range check sees the whole
        // offset, and one below 0 reads as out of range
        int off = (rb + 2) * (int)sizeof(float);
        __asm__ volatile("" : "+v"(off));
        for (int m = lane; m < ws; m += 64)
          seg[t * DDA_G2_WSMAX + m] = dda_raw_load_f32(rs, off + m * (int)sizeof(float), 0, 0);
      }
    }
    __syncthreads();
    float sum[G];
#pragma unroll
    for (int rr = 0; rr < G; ++rr) sum[rr] = 0.f;
    for (int t = 0; t < na; ++t) {
      const int rb = __builtin_amdgcn_readfirstlane(seg_rb[t]);
      if (rb == DDA_G2_MISS) continue;
      const SSDdaGatherAngle& A = ga[k0 + t];
      const double v = fma(jd, A.kx, fma(id, A.ky, A.k0h)) - (double)rb;
      const double b = floor(v);
      const float g0 = (float)(v - b);
      // the angle's segment from the pixel's base, opaque so that each line's address is
      // one shift-add of its floor
      dda_lds_float* sg = (dda_lds_float*)(seg + t * DDA_G2_WSMAX + ((int)b + 1));
      __asm__ volatile("" : "+v"(sg));
      float fr[G];
      dda_f2v q[G];
#pragma unroll
      for (int rr = 0; rr < G; ++rr) {
        const float gr = rr == 0 ? g0 : fmaf((float)rr, A.kyf, g0);
        const int fl = dda_floor_i32(gr);
        fr[rr] = __builtin_amdgcn_fractf(gr);
        q[rr][0] = sg[fl];
        q[rr][1] = sg[fl + 1];
      }
#pragma unroll
      for (int rr = 0; rr < G; ++rr) {
        const float z0 = __builtin_amdgcn_fmed3f(fmaf(-fabsf(A.c1 - fr[rr]), A.kz, A.pz), 0.f, 1.f);
        const float z1 = __builtin_amdgcn_fmed3f(fmaf(fr[rr], A.kz, A.pz1), 0.f, 1.f);
        sum[rr] = fmaf(q[rr][0], z0, sum[rr]);
        sum[rr] = fmaf(q[rr][1], z1, sum[rr]);
      }
    }
#pragma unroll
    for (int rr = 0; rr < G; ++rr) tot[rr] += (double)sum[rr];
  }
  if (j >= g.cols) return;
  T* __restrict__ out = acc + slice * (size_t)g.lines * g.cols;
#pragma unroll
  for (int rr = 0; rr < G; ++rr)
    if (i0 + rr < g.lines) out[(size_t)(i0 + rr) * g.cols + j] = T(tot[rr]);
}

// rows[slice][a][2 + r] = sino[slice][r][a] * scale[a], two zeros either side of each row
__global__ __launch_bounds__(256) void k_dda_sino_rows(int R, int A, const float* __restrict__ sino,
                                                       const float* __restrict__ scale, float* __restrict__ rows) {
  __shared__ float tile[DDA_TILE][DDA_TILE + 1];
  const size_t slice = blockIdx.z;
  const int r0 = blockIdx.y * DDA_TILE, a0 = blockIdx.x * DDA_TILE;
  const int pitch = R + 4;
  const float* __restrict__ s = sino + slice * (size_t)R * A;
  float* __restrict__ o = rows + slice * (size_t)A * pitch;
  for (int rr = threadIdx.y; rr < DDA_TILE; rr += blockDim.y) {
    const int r = r0 + rr, a = a0 + threadIdx.x;
    tile[rr][threadIdx.x] = (r < R && a < A) ? s[(size_t)r * A + a] : 0.f;
  }
  __syncthreads();
  for (int aa = threadIdx.y; aa < DDA_TILE; aa += blockDim.y) {
    const int a = a0 + aa, r = r0 + threadIdx.x;
    if (a < A && r < R) o[(size_t)a * pitch + 2 + r] = tile[threadIdx.x][aa] * scale[a];
    if (a < A && blockIdx.y == 0 && threadIdx.x < 4) {
      const int pad = threadIdx.x < 2 ? threadIdx.x : R + threadIdx.x;  // 0, 1, R + 2, R + 3
      o[(size_t)a * pitch + pad] = 0.f;
    }
  }
}

// ---------------------------------------------------------------- layouts

template <typename T>
__global__ __launch_bounds__(256) void k_dda_fill(T* __restrict__ p, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = T(0);
}

// Straight angles: vcol[slice][j] (vline[slice][i]) = sum over the straight angles whose
// slow index is the column (line) of len * (sum of q over the rays of that index), q = 1
// for `ones`; FP64 sums.  One thread per column, then per line.
template <typename T>
__global__ __launch_bounds__(256) void k_dda_straight(SSDdaGeom g, const SSDdaStraight* __restrict__ st, int nst,
                                                      const int* __restrict__ ranges, const float* __restrict__ sino,
                                                      int ones, T* __restrict__ vcol, T* __restrict__ vline) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  const size_t slice = blockIdx.y;
  if (idx >= g.cols + g.lines) return;
  const int col = idx < g.cols;
  const int s = col ? idx : idx - g.cols;
  const float* __restrict__ q = sino + slice * (size_t)g.nrays * g.nangles;
  double sum = 0.0;
  for (int k = 0; k < nst; ++k) {
    const SSDdaStraight S = st[k];
    if (S.slow_col != col) continue;
    const int lo = ranges[S.off + 2 * s], hi = ranges[S.off + 2 * s + 1];
    double qs = 0.0;
    if (ones)
      qs = hi - lo;
    else
      for (int r = lo; r < hi; ++r) qs += (double)q[(size_t)r * g.nangles + S.a];
    sum = fma(qs, S.len, sum);
  }
  if (col)
    vcol[slice * g.cols + s] = T(sum);
  else
    vline[slice * g.lines + s] = T(sum);
}

// out[lines][cols] = acc[lines][cols] + accT[cols][lines]^T + vcol[cols] + vline[lines]
// (accT, vcol, vline may be null), divided by w when given.
// Block (32, 8) handles a 32 x 32 tile; accT is read along its lines through LDS.
template <typename T, typename Out>
__global__ __launch_bounds__(256) void k_dda_fold(int lines, int cols, const T* __restrict__ acc,
                                                  const T* __restrict__ accT, const T* __restrict__ vcol,
                                                  const T* __restrict__ vline, const double* __restrict__ w,
                                                  Out* __restrict__ out) {
  __shared__ T tile[DDA_TILE][DDA_TILE + 1];
  const size_t slice = blockIdx.z;
  const size_t npix = (size_t)lines * cols;
  const int i0 = blockIdx.y * DDA_TILE, j0 = blockIdx.x * DDA_TILE;
  const T* aT = accT ? accT + slice * npix : nullptr;
  for (int jj = threadIdx.y; jj < DDA_TILE; jj += blockDim.y) {
    const int j = j0 + jj, i = i0 + threadIdx.x;
    tile[jj][threadIdx.x] = (accT && j < cols && i < lines) ? aT[(size_t)j * lines + i] : T(0);
  }
  __syncthreads();
  for (int ii = threadIdx.y; ii < DDA_TILE; ii += blockDim.y) {
    const int i = i0 + ii, j = j0 + threadIdx.x;
    if (i < lines && j < cols) {
      const size_t p = (size_t)i * cols + j;
      T v = acc[slice * npix + p] + tile[threadIdx.x][ii];
      if (vcol) v += vcol[slice * cols + j] + vline[slice * lines + i];
      out[slice * npix + p] = w ? Out((double)v / w[p]) : Out(v);
    }
  }
}

// The Radon's padded layouts of img[slice][lines][cols]: pad[slice][lines][pc] and
// padT[slice][cols][pl] (the transpose), pc = dda_pad_pitch(cols), pl =
// dda_pad_pitch(lines): each line DDA_PADW zeros, its pixels, then zeros to the pitch.
// Block (32, 8) copies a 32 x 32 tile, the transposed copy through LDS; the first and
// last tile of a line also write its padding.
__global__ __launch_bounds__(256) void k_dda_pad_images(int lines, int cols, const float* __restrict__ src,
                                                        float* __restrict__ pad, float* __restrict__ padT) {
  __shared__ float tile[DDA_TILE][DDA_TILE + 1];
  const size_t slice = blockIdx.z;
  const int i0 = blockIdx.y * DDA_TILE, j0 = blockIdx.x * DDA_TILE;
  const int pc = dda_pad_pitch(cols), pl = dda_pad_pitch(lines);
  const float* __restrict__ s = src + slice * (size_t)lines * cols;
  float* __restrict__ o = pad + slice * (size_t)lines * pc;
  float* __restrict__ oT = padT + slice * (size_t)cols * pl;
  const int x = threadIdx.x;
  for (int ii = threadIdx.y; ii < DDA_TILE; ii += blockDim.y) {
    const int i = i0 + ii, j = j0 + x;
    const float v = (i < lines && j < cols) ? s[(size_t)i * cols + j] : 0.f;
    tile[ii][x] = v;
    if (i < lines) {
      float* __restrict__ line = o + (size_t)i * pc;
      if (j < cols) line[j + DDA_PADW] = v;
      if (blockIdx.x == 0)
        for (int c = x; c < DDA_PADW; c += DDA_TILE) line[c] = 0.f;
      if (j0 + DDA_TILE >= cols)
        for (int c = cols + DDA_PADW + x; c < pc; c += DDA_TILE) line[c] = 0.f;
    }
  }
  __syncthreads();
  for (int jj = threadIdx.y; jj < DDA_TILE; jj += blockDim.y) {
    const int j = j0 + jj, i = i0 + x;
    if (j < cols) {
      float* __restrict__ line = oT + (size_t)j * pl;
      if (i < lines) line[i + DDA_PADW] = tile[x][jj];
      if (blockIdx.y == 0)
        for (int c = x; c < DDA_PADW; c += DDA_TILE) line[c] = 0.f;
      if (i0 + DDA_TILE >= lines)
        for (int c = lines + DDA_PADW + x; c < pl; c += DDA_TILE) line[c] = 0.f;
    }
  }
}

// ---------------------------------------------------------------- launchers

static dim3 dda_tiles(int lines, int cols, int nslices) {
  return dim3((cols + DDA_TILE - 1) / DDA_TILE, (lines + DDA_TILE - 1) / DDA_TILE, nslices);
}

hipError_t dda_launch_radon(const SSDdaGeom& g, const SSDdaAngle* angles, const float* img, float* pad, float* sino,
                            int nslices, hipStream_t stream) {
  float* img_pad = pad;
  float* imgT_pad = pad + (size_t)nslices * g.lines * dda_pad_pitch(g.cols);
  hipLaunchKernelGGL(k_dda_pad_images, dda_tiles(g.lines, g.cols, nslices), dim3(DDA_TILE, 8), 0, stream, g.lines,
                     g.cols, img, img_pad, imgT_pad);
  hipLaunchKernelGGL(k_dda_radon, dim3((g.nrays + 255) / 256, g.nangles, nslices), dim3(256), 0, stream, g, angles,
                     (const float*)img_pad, (const float*)imgT_pad, sino);
  return hipGetLastError();
}

template <typename T, typename Out>
hipError_t dda_launch_transpose(const SSDdaGeom& g, const SSDdaTransposePlan& tp, const float* sino, bool ones,
                                float* sinoT, T* acc, T* accT, const double* w, Out* out, int nslices,
                                hipStream_t stream) {
  const size_t n = (size_t)g.lines * g.cols * nslices;
  if (tp.ngather > 0) {
    if (!ones)  // angle-major, lmax-scaled, zero-padded copy of the sinogram
      hipLaunchKernelGGL(k_dda_sino_rows, dda_tiles(g.nrays, g.nangles, nslices), dim3(DDA_TILE, 8), 0, stream,
                         g.nrays, g.nangles, sino, tp.row_scale, sinoT);
    const dim3 grid((g.cols + DDA_GX - 1) / DDA_GX, (g.lines + DDA_GY - 1) / DDA_GY, nslices);
    if (tp.kmax <= 2 && !ones && tp.gather_ws > 0) {
      // the taller tiles halve the per-angle work per pixel; they need a block per CU
      const int gx = (g.cols + DDA_G2_COLS - 1) / DDA_G2_COLS;
      const int gy = (g.lines + 4 * DDA_G2_LINES - 1) / (4 * DDA_G2_LINES);
      if ((long long)gx * gy * nslices >= 256)
        hipLaunchKernelGGL((k_dda_gather2<T, DDA_G2_LINES>), dim3(gx, gy, nslices), dim3(256), 0, stream, g,
                           tp.gather, tp.ngather, (const float*)sinoT, tp.gather_ws, acc);
      else
        hipLaunchKernelGGL((k_dda_gather2<T, DDA_G2_LINES_SMALL>), dim3(gx, (g.lines + 4 * DDA_G2_LINES_SMALL - 1) /
                                                                          (4 * DDA_G2_LINES_SMALL), nslices),
                           dim3(256), 0, stream, g, tp.gather, tp.ngather, (const float*)sinoT, tp.gather_ws, acc);
    }
    else if (tp.kmax <= 2)
      hipLaunchKernelGGL((k_dda_gather<T, 2>), grid, dim3(256), 0, stream, g, tp.gather, tp.ngather,
                         (const float*)sinoT, ones ? 1 : 0, acc);
    else
      hipLaunchKernelGGL((k_dda_gather<T, 0>), grid, dim3(256), 0, stream, g, tp.gather, tp.ngather,
                         (const float*)sinoT, ones ? 1 : 0, acc);
  } else {
    hipLaunchKernelGGL(k_dda_fill<T>, dim3(1024), dim3(256), 0, stream, acc, n);
  }
  if (tp.nwalk > 0) {
    hipLaunchKernelGGL(k_dda_fill<T>, dim3(1024), dim3(256), 0, stream, accT, n);
    hipLaunchKernelGGL(k_dda_scatter<T>, dim3((g.nrays + 255) / 256, tp.nwalk, nslices), dim3(256), 0, stream, g,
                       tp.angles, tp.walk, sino, ones ? 1 : 0, acc, accT);
  }
  // straight angles (present only without traced ones): column and line vectors in accT,
  // [nslices][cols] then [nslices][lines] (lines * cols >= lines + cols for sides >= 2)
  T* vcol = nullptr;
  T* vline = nullptr;
  if (tp.nstraight > 0 && tp.nwalk == 0) {
    vcol = accT;
    vline = accT + (size_t)nslices * g.cols;
    hipLaunchKernelGGL(k_dda_straight<T>, dim3((g.cols + g.lines + 255) / 256, nslices), dim3(256), 0, stream, g,
                       tp.straight, tp.nstraight, tp.ranges, sino, ones ? 1 : 0, vcol, vline);
  }
  hipLaunchKernelGGL((k_dda_fold<T, Out>), dda_tiles(g.lines, g.cols, nslices), dim3(DDA_TILE, 8), 0, stream, g.lines,
                     g.cols, (const T*)acc, (const T*)(tp.nwalk > 0 ? accT : nullptr), (const T*)vcol,
                     (const T*)vline, w, out);
  return hipGetLastError();
}

__global__ __launch_bounds__(256) void k_dda_to_f32(const double* __restrict__ src, float* __restrict__ dst, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    dst[i] = (float)src[i];
}

hipError_t dda_launch_to_f32(const double* src, float* dst, size_t n, hipStream_t stream) {
  hipLaunchKernelGGL(k_dda_to_f32, dim3(1024), dim3(256), 0, stream, src, dst, n);
  return hipGetLastError();
}

template hipError_t dda_launch_transpose<float, float>(const SSDdaGeom&, const SSDdaTransposePlan&, const float*,
                                                      bool, float*, float*, float*, const double*, float*, int,
                                                      hipStream_t);
template hipError_t dda_launch_transpose<double, float>(const SSDdaGeom&, const SSDdaTransposePlan&, const float*,
                                                       bool, float*, double*, double*, const double*, float*, int,
                                                       hipStream_t);
template hipError_t dda_launch_transpose<double, double>(const SSDdaGeom&, const SSDdaTransposePlan&, const float*,
                                                        bool, float*, double*, double*, const double*, double*, int,
                                                        hipStream_t);
