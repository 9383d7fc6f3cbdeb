/*
 * raft_types.h -- the by-value C structs that cross the libraft boundary.
 *
 * Layout-compatible with the reference's strided FP64 views:
 *   raft_matrix : raft/raft_matrix.h:11-19   (element (i,j) at
 *                 p_data[i*column_stride + j*line_stride], raft/raft_matrix.h:67-68)
 *   raft_image  : raft/raft_image.h:11-19    (raft_matrix + field-of-view corners)
 *   raft_vector : raft/raft_vector.h:10-16
 * and with their ctypes mirrors in pyraft/raftypes.py:85-107.
 *
 * Sizes on x86-64: raft_matrix 24 bytes, raft_image 56 bytes, raft_vector 16 bytes.
 * All three are passed BY VALUE, exactly as the reference's exported functions do.
 */
#ifndef RAFT_TYPES_H
#define RAFT_TYPES_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  double *p_data;     /* first element                                    */
  int lines;          /* number of rows                                   */
  int line_stride;    /* distance (in doubles) between elements of a row  */
  int columns;        /* number of columns                                */
  int column_stride;  /* distance (in doubles) between elements of a column */
} raft_matrix;

typedef struct {
  raft_matrix data;
  double tl_x; /* x of the top-left corner     */
  double tl_y; /* y of the top-left corner     */
  double br_x; /* x of the bottom-right corner */
  double br_y; /* y of the bottom-right corner */
} raft_image;

typedef struct {
  double *p_data;
  int size;
  int stride;
} raft_vector;

#define RAFT_MATRIX_AT(mat, i, j) \
  ((mat).p_data[(long)(i) * (mat).column_stride + (long)(j) * (mat).line_stride])

#ifdef __cplusplus
}
#endif

#endif /* RAFT_TYPES_H */
