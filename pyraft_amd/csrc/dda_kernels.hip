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

__global__ __launch_bounds__(256) void k_dda_radon(SSDdaGeom g, const SSDdaAngle* __restrict__ angles,
                                                   const float* __restrict__ img, const float* __restrict__ imgT,
                                                   float* __restrict__ sino) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  const int ai = blockIdx.y;
  const size_t slice = blockIdx.z;
  if (r >= g.nrays) return;
  const SSDdaAngle a = angles[ai];
  const int nf = a.fast_i ? g.lines : g.cols;
  const int ns = a.fast_i ? g.cols : g.lines;
  const float* __restrict__ src = (a.fast_i ? img : imgT) + slice * (size_t)g.lines * g.cols;
  SSDdaRay ray = dda_ray_start(g, a, dda_ray_t(g, r));
  // FP64 sum: the lengths are FP64 already, so a tap costs one convert and one FMA as
  // in FP32, without FP32's N*eps growth over a ray of N equal terms
  double acc = 0.0;
  // every iteration advances the fast index by one: at most nf + 1 of them see the image
  for (int it = 0; it <= nf; ++it) {
    if (!dda_inside(ray.f, ray.s, nf, ns)) break;
    acc = fma((double)src[(size_t)ray.f * ns + ray.s], ray.len, acc);
    if (ray.slow_next) {
      ray.s = dda_step(ray.s, a.dslow);
      ray.len = a.len_full - ray.len;
      ray.slow_next = 0;
      if (!dda_inside(ray.f, ray.s, nf, ns)) break;
      acc = fma((double)src[(size_t)ray.f * ns + ray.s], ray.len, acc);
    }
    dda_fast_move(a, ray);
  }
  sino[(slice * g.nrays + r) * g.nangles + ai] = (float)acc;
}

// ---------------------------------------------------------------- transpose (scatter)

// acc: pixels of angles whose fast axis is the line ([lines][cols]); accT: the others
// ([cols][lines]).  ones: every ray carries 1 (the normalisation R^T 1).  T = float for
// the transpose, double for the backprojection, whose normalisation can cancel (the
// reference's non-square-pixel traversal yields negative lengths).
template <typename T>
__global__ __launch_bounds__(256) void k_dda_scatter(SSDdaGeom g, const SSDdaAngle* __restrict__ angles,
                                                     const float* __restrict__ sino, int ones, T* __restrict__ acc,
                                                     T* __restrict__ accT) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  const int ai = blockIdx.y;
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

// ---------------------------------------------------------------- layouts

template <typename T>
__global__ __launch_bounds__(256) void k_dda_fill(T* __restrict__ p, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = T(0);
}

// out[lines][cols] = acc[lines][cols] + accT[cols][lines]^T, divided by w when given.
// Block (32, 8) handles a 32 x 32 tile; accT is read along its lines through LDS.
template <typename T, typename Out>
__global__ __launch_bounds__(256) void k_dda_fold(int lines, int cols, const T* __restrict__ acc,
                                                  const T* __restrict__ accT, const double* __restrict__ w,
                                                  Out* __restrict__ out) {
  __shared__ T tile[DDA_TILE][DDA_TILE + 1];
  const size_t slice = blockIdx.z;
  const size_t npix = (size_t)lines * cols;
  const int i0 = blockIdx.y * DDA_TILE, j0 = blockIdx.x * DDA_TILE;
  const T* aT = accT + slice * npix;
  for (int jj = threadIdx.y; jj < DDA_TILE; jj += blockDim.y) {
    const int j = j0 + jj, i = i0 + threadIdx.x;
    tile[jj][threadIdx.x] = (j < cols && i < lines) ? aT[(size_t)j * lines + i] : T(0);
  }
  __syncthreads();
  for (int ii = threadIdx.y; ii < DDA_TILE; ii += blockDim.y) {
    const int i = i0 + ii, j = j0 + threadIdx.x;
    if (i < lines && j < cols) {
      const size_t p = (size_t)i * cols + j;
      const T v = acc[slice * npix + p] + tile[threadIdx.x][ii];
      out[slice * npix + p] = w ? Out((double)v / w[p]) : Out(v);
    }
  }
}

// dst[cols][lines] = src[lines][cols]^T
__global__ __launch_bounds__(256) void k_dda_transpose_image(int lines, int cols, const float* __restrict__ src,
                                                             float* __restrict__ dst) {
  __shared__ float tile[DDA_TILE][DDA_TILE + 1];
  const size_t slice = blockIdx.z;
  const size_t npix = (size_t)lines * cols;
  const int i0 = blockIdx.y * DDA_TILE, j0 = blockIdx.x * DDA_TILE;
  for (int ii = threadIdx.y; ii < DDA_TILE; ii += blockDim.y) {
    const int i = i0 + ii, j = j0 + threadIdx.x;
    tile[ii][threadIdx.x] = (i < lines && j < cols) ? src[slice * npix + (size_t)i * cols + j] : 0.f;
  }
  __syncthreads();
  for (int jj = threadIdx.y; jj < DDA_TILE; jj += blockDim.y) {
    const int j = j0 + jj, i = i0 + threadIdx.x;
    if (i < lines && j < cols) dst[slice * npix + (size_t)j * lines + i] = tile[threadIdx.x][jj];
  }
}

// ---------------------------------------------------------------- launchers

static dim3 dda_tiles(int lines, int cols, int nslices) {
  return dim3((cols + DDA_TILE - 1) / DDA_TILE, (lines + DDA_TILE - 1) / DDA_TILE, nslices);
}

hipError_t dda_launch_radon(const SSDdaGeom& g, const SSDdaAngle* angles, const float* img, float* imgT, float* sino,
                            int nslices, hipStream_t stream) {
  hipLaunchKernelGGL(k_dda_transpose_image, dda_tiles(g.lines, g.cols, nslices), dim3(DDA_TILE, 8), 0, stream, g.lines,
                     g.cols, img, imgT);
  hipLaunchKernelGGL(k_dda_radon, dim3((g.nrays + 255) / 256, g.nangles, nslices), dim3(256), 0, stream, g, angles,
                     img, (const float*)imgT, sino);
  return hipGetLastError();
}

template <typename T>
hipError_t dda_launch_scatter(const SSDdaGeom& g, const SSDdaAngle* angles, const float* sino, bool ones, T* acc,
                              T* accT, int nslices, hipStream_t stream) {
  const size_t n = (size_t)g.lines * g.cols * nslices;
  hipLaunchKernelGGL(k_dda_fill<T>, dim3(1024), dim3(256), 0, stream, acc, n);
  hipLaunchKernelGGL(k_dda_fill<T>, dim3(1024), dim3(256), 0, stream, accT, n);
  hipLaunchKernelGGL(k_dda_scatter<T>, dim3((g.nrays + 255) / 256, g.nangles, nslices), dim3(256), 0, stream, g,
                     angles, sino, ones ? 1 : 0, acc, accT);
  return hipGetLastError();
}

template <typename T, typename Out>
hipError_t dda_launch_fold(const SSDdaGeom& g, const T* acc, const T* accT, const double* w, Out* out, int nslices,
                           hipStream_t stream) {
  hipLaunchKernelGGL((k_dda_fold<T, Out>), dda_tiles(g.lines, g.cols, nslices), dim3(DDA_TILE, 8), 0, stream, g.lines,
                     g.cols, acc, accT, w, out);
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

template hipError_t dda_launch_scatter<float>(const SSDdaGeom&, const SSDdaAngle*, const float*, bool, float*, float*,
                                              int, hipStream_t);
template hipError_t dda_launch_scatter<double>(const SSDdaGeom&, const SSDdaAngle*, const float*, bool, double*,
                                               double*, int, hipStream_t);
template hipError_t dda_launch_fold<float, float>(const SSDdaGeom&, const float*, const float*, const double*, float*,
                                                  int, hipStream_t);
template hipError_t dda_launch_fold<double, float>(const SSDdaGeom&, const double*, const double*, const double*,
                                                   float*, int, hipStream_t);
template hipError_t dda_launch_fold<double, double>(const SSDdaGeom&, const double*, const double*, const double*,
                                                    double*, int, hipStream_t);
