/*
 * ref_dda_harness.c -- TEST INFRASTRUCTURE ONLY (built into oracle/_ref/libraft_ref.so).
 *
 * Drives the reference's own line-traversal iterator (raft/raft_ddaiterator.c, compiled
 * from its source by oracle/Makefile) through the loops of the reference's
 * single-thread workers:
 *   harness_radon_bresenham            raft/raft_radon_bresenham.cpp:17-58
 *   harness_radon_transpose_bresenham  raft/raft_radon_transpose_bresenham.cpp:48-96, 100-
 *   harness_backprojection_bresenham   raft/raft_backprojection_bresenham.cpp:58-118, 120-251
 * The reference's own worker files need a Fortran BLAS (daxpy_, dcopy_) that this
 * image does not have; with one thread their BLAS calls only zero the partial image
 * and add it to the zero-initialised output, which these loops do directly.
 *
 * The iterator leaves the first intersection length of a ray with sin(theta) == 0 or
 * cos(theta) == 0 as the previous ray left it (uninitialised for a thread's first
 * ray).  The harness starts each such angle with the full pixel length, the value the
 * reference carries into every later ray of that angle.
 */
#include <string.h>

#include "raft_ddaiterator.h"
#include "raft_image.h"

static void set_angle(raft_ddaiterator *it, double theta) {
  raft_ddaiterator_set_theta(it, theta);
  if (it->sin_theta == 0.0 || it->cos_theta == 0.0) it->sigma = it->sigma_tilde;
}

void harness_radon_bresenham(raft_image image, raft_image sino) {
  raft_ddaiterator it;
  memset(&it, 0, sizeof it);
  raft_ddaiterator_set_image(&it, image);
  const double dth = raft_image_hsamplingdistance(sino);
  const double dt = raft_image_vsamplingdistance(sino);
  for (int j = 0; j < sino.data.columns; ++j) {
    set_angle(&it, sino.tl_x + (j * dth));
    for (int i = 0; i < sino.data.lines; ++i) {
      raft_image_sample(sino, i, j) = 0.0;
      raft_ddaiterator_set_t(&it, sino.tl_y + (i * dt));
      while (!raft_ddaiterator_end(&it)) {
        raft_image_sample(sino, i, j) += raft_image_sample(image, it.i, it.j) * it.sigma;
        raft_ddaiterator_update(&it);
      }
    }
  }
}

/* image += R^T sino; with scale != 0 also scale += R^T 1 */
static void transpose_into(raft_image sino, raft_image image, double *scale) {
  raft_ddaiterator it;
  memset(&it, 0, sizeof it);
  raft_ddaiterator_set_image(&it, image);
  const double dth = raft_image_hsamplingdistance(sino);
  const double dt = raft_image_vsamplingdistance(sino);
  for (int j = 0; j < sino.data.columns; ++j) {
    set_angle(&it, sino.tl_x + (j * dth));
    for (int i = 0; i < sino.data.lines; ++i) {
      raft_ddaiterator_set_t(&it, sino.tl_y + (i * dt));
      while (!raft_ddaiterator_end(&it)) {
        raft_image_sample(image, it.i, it.j) += raft_image_sample(sino, i, j) * it.sigma;
        if (scale) scale[(long)it.i * image.data.columns + it.j] += it.sigma;
        raft_ddaiterator_update(&it);
      }
    }
  }
}

void harness_radon_transpose_bresenham(raft_image sino, raft_image image) {
  transpose_into(sino, image, 0);
}

/* image = R^T sino / R^T 1 for a zero-initialised image; scale: lines*columns doubles,
   zero-initialised by the caller */
void harness_backprojection_bresenham(raft_image sino, raft_image image, double *scale) {
  transpose_into(sino, image, scale);
  for (int i = 0; i < image.data.lines; ++i)
    for (int j = 0; j < image.data.columns; ++j)
      raft_image_sample(image, i, j) /= scale[(long)i * image.data.columns + j];
}
