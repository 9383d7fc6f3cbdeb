/*
 * raft_io.h -- binary files of raft_matrix / raft_image, as the reference reads and writes
 * them (raft/raft_matrix.c:102-136, 261-287; raft/raft_image.c:5-38, 135-151):
 *   int32 lines, int32 columns, lines*columns FP64 values in row order, and for an image
 *   four more FP64 values tl_x, tl_y, br_x, br_y (host byte order).
 * Used for exchanging fixtures with the reference's C; no device work.
 */
#ifndef RAFT_IO_H
#define RAFT_IO_H

#include <stdio.h>

#include "raft_types.h"

#ifdef __cplusplus
extern "C" {
#endif

#if defined(__GNUC__)
#pragma GCC visibility push(default)
#endif

/* Replaces raft_matrix_read (raft/raft_matrix.h:35): a new row-major matrix (free with
 * raft_matrix_destroy), or an empty one (p_data NULL) when the file is short or the sizes
 * are invalid. */
raft_matrix raft_matrix_read(FILE *p_file);
/* Replaces raft_matrix_write (raft/raft_matrix.h:50): 0 on success, 1 / 2 when the line /
 * column count could not be written, 3 + i*columns + j when element (i, j) could not. */
int raft_matrix_write(raft_matrix mat, FILE *p_file);
/* Replaces raft_matrix_destroy (raft/raft_matrix.h:47): frees a matrix this library
 * allocated and leaves it empty. */
void raft_matrix_destroy(raft_matrix *p_mat);

/* Replaces raft_image_read (raft/raft_image.h:27): the matrix, then the four corners; an
 * empty image (corners 0) on any failure. */
raft_image raft_image_read(FILE *p_file);
/* Replaces raft_image_write (raft/raft_image.h:48): raft_matrix_write's codes, or -1..-4
 * when corner tl_x, tl_y, br_x or br_y could not be written. */
int raft_image_write(raft_image image, FILE *p_file);
/* Replaces raft_image_destroy: frees the image's matrix. */
void raft_image_destroy(raft_image *p_image);

#if defined(__GNUC__)
#pragma GCC visibility pop
#endif

#ifdef __cplusplus
}
#endif

#endif /* RAFT_IO_H */
