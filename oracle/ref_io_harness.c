/*
 * ref_io_harness.c -- TEST INFRASTRUCTURE ONLY (built into oracle/_ref/libraft_ref.so).
 *
 * Path-based wrappers around the reference's own raft_image_write / raft_image_read
 * (raft/raft_image.c:5-38, 135-151, over raft/raft_matrix.c:102-136, 261-287), so the
 * tests can exchange files between them and this library's binary I/O.
 */
#include <stdio.h>
#include <string.h>

#include "raft_image.h"
#include "raft_matrix.h"

int harness_image_write(const char *path, raft_image image) {
  FILE *f = fopen(path, "wb");
  if (!f) return -100;
  const int rc = raft_image_write(image, f);
  fclose(f);
  return rc;
}

/* reads an image; copies its values (row order) into out (capacity cap) and its shape and
   corners into shape[2], corners[4]; returns 0, 1 for an empty result, 2 when cap is short */
int harness_image_read(const char *path, double *out, long cap, int *shape, double *corners) {
  FILE *f = fopen(path, "rb");
  if (!f) return -100;
  raft_image im = raft_image_read(f);
  fclose(f);
  if (!im.data.p_data) return 1;
  shape[0] = im.data.lines;
  shape[1] = im.data.columns;
  corners[0] = im.tl_x;
  corners[1] = im.tl_y;
  corners[2] = im.br_x;
  corners[3] = im.br_y;
  if ((long)im.data.lines * im.data.columns > cap) {
    raft_image_destroy(&im);
    return 2;
  }
  for (int i = 0; i < im.data.lines; ++i)
    for (int j = 0; j < im.data.columns; ++j)
      out[(long)i * im.data.columns + j] = raft_matrix_element(im.data, i, j);
  raft_image_destroy(&im);
  return 0;
}
