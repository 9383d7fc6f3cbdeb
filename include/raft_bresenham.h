/*
 * raft_bresenham.h -- line-traversal (Bresenham / DDA) projector pair of
 * pyraft_amd/lib/libraft<EXT_SUFFIX>: pyraft's default Radon method.
 *
 * Geometry of the reference's raft_ddaiterator (raft/raft_ddaiterator.c:16-257) with
 * raft_image corners {tl_x, tl_y, br_x, br_y} (raft/raft_image.h:56-61):
 *   image    [nslices][lines][cols], line i <-> y = tl_y + i*(br_y - tl_y)/(lines - 1)
 *            (line 0 at tl_y: the top for the default corners), column j <-> x = tl_x
 *            + j*(br_x - tl_x)/(cols - 1); rectangular images are allowed
 *   sinogram [nslices][R][A], column a <-> theta = tl_x + a*(br_x - tl_x)/(A - 1)
 *            (both ends included), row r <-> t = tl_y + r*(br_y - tl_y)/(R - 1)
 * Ray {x cos(theta) + y sin(theta) = t} visits pixels of the image and weights each by
 * its intersection length.  lines, cols, R and A must be >= 2.
 *
 * Drop-in symbols first (void, by-value structs, reference signatures; status via
 * raft_ss_legacy_status()), then the device-pointer API (status codes RAFT_SS_E*,
 * message in raft_ss_last_error()).
 */
#ifndef RAFT_BRESENHAM_H
#define RAFT_BRESENHAM_H

#include "raft_types.h"

#ifdef __cplusplus
extern "C" {
#endif

#if defined(__GNUC__)
#pragma GCC visibility push(default)
#endif

/* Replaces raft_radon_bresenham (raft/raft_radon_bresenham.h:10-13,
 * raft/raft_radon_bresenham.cpp:60-105), bound at pyraft/raftypes.py:152-153 and called by
 * pyraft.radon(img, shape) with its default method (pyraft/sinogram/sino.py:199-200).
 * Every element of sino is overwritten.  nthreads is accepted and ignored. */
void raft_radon_bresenham(raft_image image, raft_image sino, int nthreads);

/* Replaces raft_radon_transpose_bresenham (raft/raft_radon_transpose_bresenham.h,
 * raft/raft_radon_transpose_bresenham.cpp), bound at pyraft/raftypes.py:156-157 and called
 * by pyraft.radon_transpose (pyraft/backprojection/backprojection.py:229-230) and
 * make_em (pyraft/iterative/em.py:109-142).  Adds R^T sino into image, as the
 * reference's column sums do (raft/raft_radon_transpose_bresenham.cpp:17-46). */
void raft_radon_transpose_bresenham(raft_image sino, raft_image image, int nthreads);

/* Replaces raft_backprojection_bresenham (raft/raft_backprojection_bresenham.h,
 * raft/raft_backprojection_bresenham.cpp:120-251), bound at pyraft/raftypes.py:154-155 and
 * called by pyraft.backprojection with a BRESENHAM plan
 * (pyraft/backprojection/backprojection.py:188-189): image = (image + R^T sino) / R^T 1,
 * unguarded as in the reference (0/0 where no ray passes a pixel). */
void raft_backprojection_bresenham(raft_image sino, raft_image image, int nthreads);

/* ---------------------------------------------------------------- device API */

typedef struct raft_dda_plan raft_dda_plan;

/* Geometry plan: the FP64 angle table (glibc cos/sin, as the reference evaluates
 * them).  img_corners / sino_corners are {tl_x, tl_y, br_x, br_y}.  No reference
 * counterpart (the reference traces from scratch on every call). */
int raft_dda_plan_create(raft_dda_plan **plan, int lines, int cols, int nrays, int nangles,
                         const double *img_corners, const double *sino_corners, void *stream);
int raft_dda_plan_destroy(raft_dda_plan *plan);
/* info[0..6] = lines, cols, nrays, nangles, #angles whose fast axis is the line,
 * #angles the transpose traces ray by ray, #angles whose rays all run straight along an
 * axis, which the transpose sums per column or line (the others it evaluates pixel by
 * pixel) */
int raft_dda_plan_info(const raft_dda_plan *plan, int *info, int ninfo);

/* sino[s] = R img[s] (overwritten), device FP32 buffers */
int raft_dda_radon(const raft_dda_plan *plan, const float *img, float *sino, int nslices, void *stream);
/* img[s] = R^T sino[s] (overwritten) */
int raft_dda_radon_transpose(const raft_dda_plan *plan, const float *sino, float *img, int nslices, void *stream);
/* img[s] = R^T sino[s] / R^T 1 (the reference's normalised backprojection; 0/0 where
 * no ray passes a pixel) */
int raft_dda_backprojection(raft_dda_plan *plan, const float *sino, float *img, int nslices, void *stream);
/* out = R^T 1 ([lines][cols]), the backprojection's normalisation (computed once per
 * plan and kept) */
int raft_dda_weights(raft_dda_plan *plan, float *out, void *stream);

/* The pixels the device Radon adds for ray `ray` of angle `angle`, in order, evaluated
 * on the host (no device needed; for checking index logic): on angles it traces move
 * by move (FP64 drift) the visited pixels; on exact-geometry angles the trapezoid form's
 * nonzero candidates per fast index (Q32 stepping as on the device, weights in FP64).
 * ij[2k], ij[2k+1] = line, column and len[k] = intersection length of the k-th pixel.
 * Where the device stages a wave's image segments in LDS, it also checks that both
 * candidates lie inside the segment the ray's wave copies (RAFT_SS_EINTERNAL if not).
 * Returns the number of pixels (at most cap are written) or a RAFT_SS_E* code. */
int raft_dda_host_ray(int lines, int cols, int nrays, int nangles, const double *img_corners,
                      const double *sino_corners, int angle, int ray, int *ij, double *len, int cap);

#if defined(__GNUC__)
#pragma GCC visibility pop
#endif

#ifdef __cplusplus
}
#endif

#endif /* RAFT_BRESENHAM_H */
