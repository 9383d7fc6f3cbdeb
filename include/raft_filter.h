/*
 * raft_filter.h -- the reference's C FBP filter and FBP of pyraft_amd/lib/libraft<EXT_SUFFIX>.
 *
 * raft_filter1d_ramp filters every column (angle) of a sinogram in place exactly as
 * raft/raft_filter.cpp:118-213 does, half-spectrum quirk included: the column is zero
 * padded to Z = 2R - 1 samples, bins 0..R/2 of its Z-point spectrum are multiplied by
 * ramp(k)/Z with ramp(k) = k pi / ((R/2) |dt|) below int((R/2 + 1) cutoff) and 0 above,
 * bins R/2 + 1 .. R - 1 are left as they are, and the unnormalised inverse transform's
 * first R samples are kept (raft/raft_filter.cpp:70-92).  As a linear operator this is a
 * Toeplitz convolution with 2R - 1 taps, which is how it runs here.
 */
#ifndef RAFT_FILTER_H
#define RAFT_FILTER_H

#include "raft_types.h"

#ifdef __cplusplus
extern "C" {
#endif

#if defined(__GNUC__)
#pragma GCC visibility push(default)
#endif

/* raft/raft_enums.h:4-9 */
typedef enum { RAFT_BRESENHAM, RAFT_SLANTSTACK } raft_radon_method;

/* Replaces raft_filter1d_ramp (raft/raft_filter.h:17-20, raft/raft_filter.cpp:184-213), bound
 * at pyraft/raftypes.py:185-186.  In place on sino; cutoff above 1 (or NaN) counts as 1.
 * R >= 2.  nthreads is accepted and ignored. */
void raft_filter1d_ramp(double cutoff, raft_image sino, int nthreads);

/* Replaces raft_fbp (raft/raft_fbp.h:11-16, raft/raft_fbp.c:5-24): raft_filter1d_ramp on sino
 * (in place, as the reference does), then the backprojection of the method into image. */
void raft_fbp(double cutoff, raft_image sino, raft_image image, raft_radon_method method, int nthreads);

/* The 2R Toeplitz taps of raft_filter1d_ramp for R rays of pitch |dt| (taps[d + R] is the
 * weight of input sample m in output sample m + d, |d| < R; taps[0] = 0), computed in FP64
 * on the host.  No reference counterpart (the reference filters through FFTW). */
int raft_ss_filter1d_ramp_taps(int R, double cutoff, double dt, float *host_taps);
/* Toeplitz filter of nrows contiguous length-R rows with the device taps (2R floats);
 * an FFT convolution up to R = 2048, a direct sum up to R = 6826 (larger R: error).
 * Not in place. */
int raft_ss_filter_rows_toeplitz(const float *d_in, float *d_out, const float *d_taps, int R, int nrows,
                                 void *stream);

#if defined(__GNUC__)
#pragma GCC visibility pop
#endif

#ifdef __cplusplus
}
#endif

#endif /* RAFT_FILTER_H */
