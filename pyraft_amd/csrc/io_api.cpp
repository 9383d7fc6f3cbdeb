// io_api.cpp -- binary raft_matrix / raft_image files in the reference's layout
// (raft/raft_matrix.c:102-136, 261-287; raft/raft_image.c:5-38, 135-151).
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <vector>

#include "../../include/raft_io.h"

namespace {
const raft_matrix kEmptyMatrix = {nullptr, 0, 0, 0, 0};
}

extern "C" void raft_matrix_destroy(raft_matrix* p_mat) {
  if (!p_mat) return;
  free(p_mat->p_data);
  *p_mat = kEmptyMatrix;
}

extern "C" raft_matrix raft_matrix_read(FILE* f) {
  raft_matrix m = kEmptyMatrix;
  int lines = 0, columns = 0;
  if (!f || fread(&lines, sizeof(int), 1, f) != 1 || fread(&columns, sizeof(int), 1, f) != 1) return kEmptyMatrix;
  if (lines < 0 || columns < 0) return kEmptyMatrix;
  const size_t count = (size_t)lines * (size_t)columns;
  double* data = static_cast<double*>(malloc(count ? count * sizeof(double) : sizeof(double)));
  if (!data) return kEmptyMatrix;
  if (fread(data, sizeof(double), count, f) != count) {
    free(data);
    return kEmptyMatrix;
  }
  m.p_data = data;
  m.lines = lines;
  m.columns = columns;
  m.line_stride = 1;
  m.column_stride = columns;
  return m;
}

extern "C" int raft_matrix_write(raft_matrix mat, FILE* f) {
  if (!f || fwrite(&mat.lines, sizeof(int), 1, f) != 1) return 1;
  if (fwrite(&mat.columns, sizeof(int), 1, f) != 1) return 2;
  std::vector<double> row(mat.columns > 0 ? mat.columns : 1);
  for (int i = 0; i < mat.lines; ++i) {
    for (int j = 0; j < mat.columns; ++j) row[j] = RAFT_MATRIX_AT(mat, i, j);
    const size_t done = fwrite(row.data(), sizeof(double), (size_t)mat.columns, f);
    if (done != (size_t)mat.columns) return 3 + i * mat.columns + (int)done;
  }
  return 0;
}

extern "C" void raft_image_destroy(raft_image* p_image) {
  if (p_image) raft_matrix_destroy(&p_image->data);
}

extern "C" raft_image raft_image_read(FILE* f) {
  raft_image im;
  memset(&im, 0, sizeof(im));
  im.data = raft_matrix_read(f);
  if (!im.data.p_data) return im;
  double c[4];
  if (fread(c, sizeof(double), 4, f) != 4) {
    raft_image_destroy(&im);
    memset(&im, 0, sizeof(im));
    return im;
  }
  im.tl_x = c[0];
  im.tl_y = c[1];
  im.br_x = c[2];
  im.br_y = c[3];
  return im;
}

extern "C" int raft_image_write(raft_image image, FILE* f) {
  const int rc = raft_matrix_write(image.data, f);
  if (rc) return rc;
  const double c[4] = {image.tl_x, image.tl_y, image.br_x, image.br_y};
  for (int k = 0; k < 4; ++k)
    if (fwrite(&c[k], sizeof(double), 1, f) != 1) return -1 - k;
  return 0;
}
