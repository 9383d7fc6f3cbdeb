// filter_api.cpp -- the reference's C FBP filter (raft_filter1d_ramp) and raft_fbp.
//
// raft/raft_filter.cpp zero-pads each column x (length R) to Z = 2R - 1, takes the
// Z-point real FFT, multiplies bins 0..R/2 by ramp(k)/Z, leaves bins R/2+1..R-1 as they
// are, and keeps the first R samples of the unnormalised inverse.  The spectrum
// multiplier H is real and even, so the output is
//   y[n] = sum_m x[m] K[n - m],  K[d] = H[0] + 2 sum_{k=1}^{R-1} H[k] cos(2 pi k d / Z),
// a Toeplitz convolution with the 2R - 1 taps K[-(R-1) .. R-1].  The taps are built in
// FP64 here and the convolution runs on the device row filter.
#include <hip/hip_runtime.h>
#include <math.h>
#include <string.h>

#include <vector>

#include "../../include/raft_bresenham.h"
#include "../../include/raft_filter.h"
#include "ss_internal.h"

namespace {

// ramp of raft/raft_filter.cpp:189-206 over bins 0..F-1, F = R/2 + 1
std::vector<double> ramp_bins(int R, double cutoff, double dt) {
  const int F = R / 2 + 1;
  std::vector<double> f(F, 0.0);
  cutoff = (cutoff <= 1.0) ? cutoff : 1.0;
  double delta = 3.14159265358979323846264338327 / (F - 1);
  delta /= fabs(dt);
  const int length = (int)(F * cutoff);
  for (int i = 0; i < length && i < F; ++i) f[i] = delta * i;
  return f;
}

void toeplitz_taps(int R, double cutoff, double dt, float* taps) {
  const int Z = 2 * R - 1;
  const std::vector<double> ramp = ramp_bins(R, cutoff, dt);
  std::vector<double> H(R), ct(Z);
  for (int k = 0; k < R; ++k) H[k] = k < (int)ramp.size() ? ramp[k] / Z : 1.0;
  for (int j = 0; j < Z; ++j) ct[j] = cos(2.0 * M_PI * (double)j / Z);
  taps[0] = 0.f;
  for (int d = 0; d < R; ++d) {
    double acc = H[0];
    for (int k = 1; k < R; ++k) acc += 2.0 * H[k] * ct[((long long)k * d) % Z];
    taps[R + d] = (float)acc;  // K is even in d
    taps[R - d] = (float)acc;
  }
}

int check_taps_args(int R, double dt) {
  if (R < 2) return ss_set_error(RAFT_SS_EINVAL, "the filter needs R >= 2 (got %d)", R);
  if (!(fabs(dt) > 0.0) || !std::isfinite(dt)) return ss_set_error(RAFT_SS_EINVAL, "degenerate detector pitch");
  if ((size_t)(((2 * R + 3) & ~3) + 4 * R) * sizeof(float) > 160 * 1024)
    return ss_set_error(RAFT_SS_EINVAL, "the row filter holds at most R = %d rays", (160 * 1024 / 4 - 4) / 6);
  return 0;
}

// raft_filter1d_ramp on a host FP64 sinogram, in place
int host_filter1d_ramp(double cutoff, const raft_image& sino) {
  if (!sino.data.p_data) return ss_set_error(RAFT_SS_EINVAL, "empty image");
  const int R = sino.data.lines, A = sino.data.columns;
  const double dt = (sino.br_y - sino.tl_y) / (R - 1);  // raft_image_vsamplingdistance
  int rc = check_taps_args(R, dt);
  if (rc) return rc;
  if (A < 1) return 0;
  std::vector<float> taps(2 * R);
  toeplitz_taps(R, cutoff, dt, taps.data());
  const size_t n = (size_t)R * A;
  PinnedF32 h;
  HIP_TRY(h.alloc(n));
  // angle-major rows: row a = column a of the sinogram
  for (int a = 0; a < A; ++a)
    for (int r = 0; r < R; ++r) h.p[(size_t)a * R + r] = (float)RAFT_MATRIX_AT(sino.data, r, a);
  DevBuf din, dout, dtaps;
  HIP_TRY(din.alloc(sizeof(float) * n, nullptr));
  HIP_TRY(dout.alloc(sizeof(float) * n, nullptr));
  HIP_TRY(dtaps.alloc(sizeof(float) * 2 * R, nullptr));
  HIP_TRY(hipMemcpyAsync(din.p, h.p, sizeof(float) * n, hipMemcpyHostToDevice, nullptr));
  HIP_TRY(hipMemcpyAsync(dtaps.p, taps.data(), sizeof(float) * 2 * R, hipMemcpyHostToDevice, nullptr));
  HIP_TRY(ss_launch_ramp_filter(din.as<float>(), dout.as<float>(), dtaps.as<float>(), R, A, nullptr, true));
  HIP_TRY(hipMemcpyAsync(h.p, dout.p, sizeof(float) * n, hipMemcpyDeviceToHost, nullptr));
  HIP_TRY(hipStreamSynchronize(nullptr));
  for (int a = 0; a < A; ++a)
    for (int r = 0; r < R; ++r) RAFT_MATRIX_AT(sino.data, r, a) = (double)h.p[(size_t)a * R + r];
  return 0;
}

}  // namespace

extern "C" int raft_ss_filter1d_ramp_taps(int R, double cutoff, double dt, float* host_taps) {
  if (!host_taps) return ss_set_error(RAFT_SS_EINVAL, "null argument");
  int rc = check_taps_args(R, dt);
  if (rc) return rc;
  toeplitz_taps(R, cutoff, dt, host_taps);
  return 0;
}

extern "C" int raft_ss_filter_rows_toeplitz(const float* d_in, float* d_out, const float* d_taps, int R, int nrows,
                                            void* stream) {
  if (!d_in || !d_out || !d_taps || R < 2 || nrows < 0) return ss_set_error(RAFT_SS_EINVAL, "invalid filter arguments");
  if (d_in == d_out) return ss_set_error(RAFT_SS_EINVAL, "filter cannot run in place");
  if (nrows == 0) return 0;
  HIP_TRY(ss_launch_ramp_filter(d_in, d_out, d_taps, R, nrows, as_stream(stream), true));
  return 0;
}

extern "C" void raft_filter1d_ramp(double cutoff, raft_image sino, int nthreads) {
  (void)nthreads;
  ss_report_legacy("raft_filter1d_ramp", host_filter1d_ramp(cutoff, sino));
}

extern "C" void raft_fbp(double cutoff, raft_image sino, raft_image image, raft_radon_method method, int nthreads) {
  int rc = host_filter1d_ramp(cutoff, sino);
  if (rc) {
    ss_report_legacy("raft_fbp", rc);
    return;
  }
  switch (method) {
    case RAFT_BRESENHAM:
      raft_backprojection_bresenham(sino, image, nthreads);
      break;
    case RAFT_SLANTSTACK:
      raft_backprojection_slantstack(sino, image, nthreads);
      break;
    default:
      ss_report_legacy("raft_fbp", ss_set_error(RAFT_SS_EINVAL, "unknown method %d", (int)method));
  }
}
