/*
 * oracle.c -- FP64 CPU restatement of the PyRAFT slant-stack path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the shipped library (pyraft_amd/) links,
 * loads or calls this file; only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py use it, and only as the checker / CPU baseline.
 *
 * Every function restates one reference routine with the same FP64 operation
 * order, so that the floor() indices and the 0 <= m <= N-2 predicates come out
 * identical (compile with -ffp-contract=off: no fused multiply-adds).
 *
 *   oracle_radon_slantstack            raft/raft_radon_slantstack.c:44-89, 122-214
 *   oracle_backprojection_slantstack   raft/raft_backprojection_slantstack.c:43-87, 121-181
 *   oracle_radon_transpose_slantstack  exact adjoint of the above forward (no reference
 *                                      counterpart; pinned by the dot-product identity)
 *   oracle_phantom_fromdesc            raft/raft_image_phantom_fromdesc.c:4-59
 *   oracle_radon_fromdesc              raft/raft_radon_fromdesc.c:4-67
 *   oracle_radon_withmesh              raft/raft_radon_slantstack_withmesh.c:49-224
 *   oracle_radon_view                  raft/raft_radon_slantstack_view.c:22-113
 *
 * Threading: the reference splits the outputs into nthreads+1 contiguous chunks
 * of floor(total/nthreads) (raft/raft_radon_slantstack.c:62-81), which leaves
 * trailing outputs uncomputed when total mod nthreads > floor(total/nthreads).
 * Here the outputs are split into nthreads chunks of ceil(total/nthreads), so every
 * output is computed; values do not depend on the partition.
 *
 * The angle range [j0, j1) restricts the work to a contiguous block of sinogram
 * columns while keeping the full-sinogram angle grid theta_j = tl_x + j*dth,
 * dth = (br_x - tl_x)/ncolumns.  j0 = 0, j1 = ncolumns is the reference call.
 */
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#include "../include/raft_types.h"

#define AT(mat, i, j) RAFT_MATRIX_AT(mat, i, j)

/* ------------------------------------------------------------------ */
/* thread fan-out                                                       */
/* ------------------------------------------------------------------ */

typedef void (*work_fn)(void *ctx, long begin, long end);

typedef struct {
  work_fn fn;
  void *ctx;
  long begin, end;
} chunk_t;

static void *chunk_main(void *arg) {
  chunk_t *c = (chunk_t *)arg;
  c->fn(c->ctx, c->begin, c->end);
  return NULL;
}

static void fan_out(work_fn fn, void *ctx, long total, int nthreads) {
  if (total <= 0) return;
  if (nthreads < 1) nthreads = 1;
  if (nthreads > total) nthreads = (int)total;
  if (nthreads == 1) {
    fn(ctx, 0, total);
    return;
  }
  long per = (total + nthreads - 1) / nthreads;
  pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * nthreads);
  chunk_t *ch = (chunk_t *)malloc(sizeof(chunk_t) * nthreads);
  int launched = 0;
  for (int n = 0; n < nthreads; ++n) {
    long b = per * n, e = b + per;
    if (e > total) e = total;
    ch[n].fn = fn; ch[n].ctx = ctx; ch[n].begin = b; ch[n].end = e;
    if (b >= e) continue;
    if (pthread_create(&th[n], NULL, chunk_main, &ch[n]) != 0) {
      fn(ctx, b, e); /* could not spawn: do the chunk here */
      ch[n].begin = ch[n].end = 0;
      continue;
    }
    ++launched;
  }
  for (int n = 0; n < nthreads; ++n)
    if (ch[n].begin < ch[n].end) pthread_join(th[n], NULL);
  (void)launched;
  free(th);
  free(ch);
}

/* ------------------------------------------------------------------ */
/* slant-stack Radon  (raft/raft_radon_slantstack.c:122-214)            */
/* ------------------------------------------------------------------ */

typedef struct {
  raft_image img, sino;
  int j0, j1;
} radon_ctx;

/* One ray integral at angle theta and detector coordinate t: the body shared by the
   uniform-grid routine (raft/raft_radon_slantstack.c:122-214) and the arbitrary-mesh
   ones (raft/raft_radon_slantstack_withmesh.c:147-224, raft/raft_radon_slantstack_view.c:40-113),
   which differ only in where theta and t come from. */
static double ray_sum_at(const raft_image *img, double theta, double t) {
  const int n = img->data.lines;            /* square image assumed (reference :140) */
  const double tol = 1.0 / sqrt(2);         /* :139 */

  const double xmin = img->tl_x, xmax = img->br_x;   /* :143-147 */
  const double ymin = img->br_y, ymax = img->tl_y;
  const double dx = (xmax - xmin) / (n - 1);          /* :150-151 */
  const double dy = (ymax - ymin) / (n - 1);

  const double c = cos(theta), s = sin(theta), tn = tan(theta); /* :160-164 */
  const double fs = fabs(s), fc = fabs(c);

  double acc = 0.0;
  if (fc < tol) {
    /* columns sampled, rows interpolated (:174-192) */
    const double scale = (dy / fs);
    for (int i = 0; i < n; ++i) {
      const double x = (double)(xmin + i * dx);
      const double y = (double)(t / s - x / tn);
      const int m = (int)floor((y - ymin) / dy);
      if (m > -1 && m < n - 1) {
        const double f0 = AT(img->data, m, i), f1 = AT(img->data, m + 1, i);
        acc += (f1 - f0) * (y - (ymin + m * dy)) / dy + f0;
      }
    }
    return acc * scale;
  } else {
    /* rows sampled, columns interpolated (:193-211); the row coordinate is built
       from xmin and dy exactly as the reference does (:200) */
    const double scale = (dx / fc);
    for (int i = 0; i < n; ++i) {
      const double y = (double)(xmin + i * dy);
      const double x = (double)(t / c - y * tn);
      const int m = (int)floor((x - xmin) / dx);
      if (m > -1 && m < n - 1) {
        const double f0 = AT(img->data, i, m), f1 = AT(img->data, i, m + 1);
        acc += (f1 - f0) * (x - (xmin + m * dx)) / dx + f0;
      }
    }
    return acc * scale;
  }
}

/* j = angle column, k = detector row of the uniform grid */
static double ray_sum(const raft_image *img, const raft_image *sino, int j, int k) {
  const int nviews = sino->data.columns;
  const int nrays = sino->data.lines;
  const double th0 = sino->tl_x, thf = sino->br_x;    /* :154-158 */
  const double dth = (thf - th0) / (nviews);
  const double theta = th0 + j * dth;
  const double tmin = sino->br_y, tmax = sino->tl_y;  /* :168-172 */
  const double dt = (tmax - tmin) / (nrays - 1);
  const double t = tmin + k * dt;
  return ray_sum_at(img, theta, t);
}

static void radon_work(void *vctx, long b, long e) {
  radon_ctx *ctx = (radon_ctx *)vctx;
  const int nrays = ctx->sino.data.lines;
  for (long w = b; w < e; ++w) {
    const int j = ctx->j0 + (int)(w / nrays);
    const int k = (int)(w % nrays);
    AT(ctx->sino.data, k, j) = ray_sum(&ctx->img, &ctx->sino, j, k);
  }
}

void oracle_radon_slantstack(raft_image phantom, raft_image radon, int j0, int j1, int nthreads) {
  if (j0 < 0) j0 = 0;
  if (j1 > radon.data.columns) j1 = radon.data.columns;
  if (j1 <= j0) return;
  radon_ctx ctx = {phantom, radon, j0, j1};
  fan_out(radon_work, &ctx, (long)(j1 - j0) * radon.data.lines, nthreads);
}

/* ------------------------------------------------------------------ */
/* arbitrary meshes (raft/raft_radon_slantstack_withmesh.c:49-145)      */
/* ------------------------------------------------------------------ */

typedef struct {
  raft_image img;
  raft_matrix out;
  raft_vector t, theta;
} mesh_ctx;

static void mesh_work(void *vctx, long b, long e) {
  mesh_ctx *ctx = (mesh_ctx *)vctx;
  const int nrays = ctx->out.lines;
  for (long w = b; w < e; ++w) {
    const int j = (int)(w / nrays);  /* :116-117 */
    const int k = (int)(w % nrays);
    const double th = ctx->theta.p_data[(long)j * ctx->theta.stride];
    const double tt = ctx->t.p_data[(long)k * ctx->t.stride];
    AT(ctx->out, k, j) = ray_sum_at(&ctx->img, th, tt);
  }
}

/* radon element (k, j) = ray sum at (theta[j], t[k]); every output is computed (the
   reference's split of nviews*nrays into nthreads chunks of floor(total/nthreads)
   leaves the remainder unwritten, :65-77) */
void oracle_radon_withmesh(raft_image phantom, raft_image radon, raft_vector t, raft_vector theta, int nthreads) {
  mesh_ctx ctx = {phantom, radon.data, t, theta};
  fan_out(mesh_work, &ctx, (long)radon.data.lines * radon.data.columns, nthreads);
}

/* one angle (raft/raft_radon_slantstack_view.c:22-38) */
void oracle_radon_view(raft_vector view, raft_image phantom, raft_vector t, double theta) {
  for (int k = 0; k < t.size; ++k)
    view.p_data[(long)k * view.stride] = ray_sum_at(&phantom, theta, t.p_data[(long)k * t.stride]);
}

/* ------------------------------------------------------------------ */
/* slant-stack backprojection (raft/raft_backprojection_slantstack.c:121-181) */
/* ------------------------------------------------------------------ */

typedef struct {
  raft_image sino, img;
  int j0, j1;
} bp_ctx;

static double angle_sum(const raft_image *sino, const raft_image *img, int row, int col, int j0, int j1) {
  const int nrays = sino->data.lines;      /* :135-136 */
  const int nviews = sino->data.columns;
  const int n = img->data.lines;

  const double tmin = sino->br_y, tmax = sino->tl_y;  /* :138-140 */
  const double dt = (tmax - tmin) / (nrays - 1);
  const double th0 = sino->tl_x, thf = sino->br_x;    /* :142-145 */
  const double dth = (thf - th0) / (nviews);

  const double xmin = img->tl_x, xmax = img->br_x;    /* :147-153 */
  const double dx = (xmax - xmin) / (n - 1);
  const double ymin = img->br_y, ymax = img->tl_y;
  const double dy = (ymax - ymin) / (n - 1);

  const double x = xmin + col * dx;                   /* :155-157 */
  const double y = ymin + row * dy;

  double acc = 0.0;
  for (int i = j0; i < j1; ++i) {                     /* :163-178 */
    const double th = th0 + i * dth;
    const double c = cos(th), s = sin(th);
    const double t = x * c + y * s;
    const int m = (int)floor((t - tmin) / dt);
    if (m > -1 && m < nrays - 1) {
      const double s0 = AT(sino->data, m, i), s1 = AT(sino->data, m + 1, i);
      acc += (s1 - s0) * (t - (tmin + m * dt)) / dt + s0;
    }
  }
  return acc * dth;                                   /* :180 */
}

static void bp_work(void *vctx, long b, long e) {
  bp_ctx *ctx = (bp_ctx *)vctx;
  const int n = ctx->img.data.lines;
  for (long w = b; w < e; ++w) {
    const int row = (int)(w / n), col = (int)(w % n);
    AT(ctx->img.data, row, col) = angle_sum(&ctx->sino, &ctx->img, row, col, ctx->j0, ctx->j1);
  }
}

void oracle_backprojection_slantstack(raft_image sinogram, raft_image backprojection, int j0, int j1,
                                      int nthreads) {
  if (j0 < 0) j0 = 0;
  if (j1 > sinogram.data.columns) j1 = sinogram.data.columns;
  bp_ctx ctx = {sinogram, backprojection, j0, j1};
  const long n = backprojection.data.lines;
  fan_out(bp_work, &ctx, n * n, nthreads);
}

/* The same pixels for a list of image rows only: out[i][col] = pixel (rows[i], col) of
   oracle_backprojection_slantstack on an n x n image with the corners of `geometry`
   (whose data is not read).  Lets a full-size backprojection be checked on a sample
   of its rows. */
typedef struct {
  raft_image sino, geom;
  int j0, j1;
  const int *rows;
  double *out;
} bp_rows_ctx;

static void bp_rows_work(void *vctx, long b, long e) {
  bp_rows_ctx *ctx = (bp_rows_ctx *)vctx;
  const int n = ctx->geom.data.lines;
  for (long w = b; w < e; ++w) {
    const long i = w / n;
    const int col = (int)(w % n);
    ctx->out[w] = angle_sum(&ctx->sino, &ctx->geom, ctx->rows[i], col, ctx->j0, ctx->j1);
  }
}

void oracle_backprojection_rows(raft_image sinogram, raft_image geometry, int j0, int j1, const int *rows,
                                int nrows, double *out, int nthreads) {
  if (j0 < 0) j0 = 0;
  if (j1 > sinogram.data.columns) j1 = sinogram.data.columns;
  bp_rows_ctx ctx = {sinogram, geometry, j0, j1, rows, out};
  fan_out(bp_rows_work, &ctx, (long)nrows * geometry.data.lines, nthreads);
}

/* ------------------------------------------------------------------ */
/* exact transpose of the slant-stack Radon (scatter form)              */
/* ------------------------------------------------------------------ */

typedef struct {
  raft_image sino, img;
  int j0, j1;
  double *partial; /* one n*n accumulator per chunk */
  long per;
} rt_ctx;

/* For ray (j,k): every sample that ray_sum() accumulates as f0 + w*(f1-f0) with
   w = (coord - grid)/step hands (1-w)*scale*q to f0's pixel and w*scale*q to
   f1's pixel.  Same expressions as ray_sum(), so the index pattern is identical. */
static void scatter_ray(const raft_image *img, const raft_image *sino, int j, int k, double q, double *acc) {
  const int n = img->data.lines;
  const int nviews = sino->data.columns;
  const int nrays = sino->data.lines;
  const double tol = 1.0 / sqrt(2);
  const double xmin = img->tl_x, xmax = img->br_x;
  const double ymin = img->br_y, ymax = img->tl_y;
  const double dx = (xmax - xmin) / (n - 1);
  const double dy = (ymax - ymin) / (n - 1);
  const double th0 = sino->tl_x, thf = sino->br_x;
  const double dth = (thf - th0) / (nviews);
  const double theta = th0 + j * dth;
  const double c = cos(theta), s = sin(theta), tn = tan(theta);
  const double fs = fabs(s), fc = fabs(c);
  const double tmin = sino->br_y, tmax = sino->tl_y;
  const double dt = (tmax - tmin) / (nrays - 1);
  const double t = tmin + k * dt;
  if (fc < tol) {
    const double v = q * (dy / fs);
    for (int i = 0; i < n; ++i) {
      const double x = (double)(xmin + i * dx);
      const double y = (double)(t / s - x / tn);
      const int m = (int)floor((y - ymin) / dy);
      if (m > -1 && m < n - 1) {
        const double w = (y - (ymin + m * dy)) / dy;
        acc[(long)m * n + i] += (1.0 - w) * v;
        acc[(long)(m + 1) * n + i] += w * v;
      }
    }
  } else {
    const double v = q * (dx / fc);
    for (int i = 0; i < n; ++i) {
      const double y = (double)(xmin + i * dy);
      const double x = (double)(t / c - y * tn);
      const int m = (int)floor((x - xmin) / dx);
      if (m > -1 && m < n - 1) {
        const double w = (x - (xmin + m * dx)) / dx;
        acc[(long)i * n + m] += (1.0 - w) * v;
        acc[(long)i * n + m + 1] += w * v;
      }
    }
  }
}

static void rt_work(void *vctx, long b, long e) {
  rt_ctx *ctx = (rt_ctx *)vctx;
  const int nrays = ctx->sino.data.lines;
  const long n = ctx->img.data.lines;
  double *acc = ctx->partial + (b / ctx->per) * n * n;
  for (long w = b; w < e; ++w) {
    const int j = ctx->j0 + (int)(w / nrays);
    const int k = (int)(w % nrays);
    const double q = AT(ctx->sino.data, k, j);
    if (q != 0.0) scatter_ray(&ctx->img, &ctx->sino, j, k, q, acc);
  }
}

void oracle_radon_transpose_slantstack(raft_image sinogram, raft_image image, int j0, int j1, int nthreads) {
  if (j0 < 0) j0 = 0;
  if (j1 > sinogram.data.columns) j1 = sinogram.data.columns;
  const long n = image.data.lines;
  const long total = (long)(j1 > j0 ? j1 - j0 : 0) * sinogram.data.lines;
  if (nthreads < 1) nthreads = 1;
  if (total > 0 && nthreads > total) nthreads = (int)total;
  const long per = total > 0 ? (total + nthreads - 1) / nthreads : 1;
  const long nchunks = total > 0 ? (total + per - 1) / per : 0;
  double *partial = (double *)calloc((size_t)(nchunks > 0 ? nchunks : 1) * n * n, sizeof(double));
  rt_ctx ctx = {sinogram, image, j0, j1, partial, per};
  fan_out(rt_work, &ctx, total, nthreads);
  for (long r = 0; r < n; ++r)
    for (long c = 0; c < n; ++c) {
      double s = 0.0;
      for (long ch = 0; ch < nchunks; ++ch) s += partial[ch * n * n + r * n + c];
      AT(image.data, r, c) = s;
    }
  free(partial);
}

/* ------------------------------------------------------------------ */
/* ellipse phantoms (raft/raft_image_phantom_fromdesc.c:4-59)           */
/* desc columns: value, a (horizontal semi-axis), b, x0, y0, rotation    */
/* ------------------------------------------------------------------ */

void oracle_phantom_fromdesc(raft_image image, raft_matrix desc) {
  if (!image.data.p_data || desc.columns < 6) return; /* the reference frees the caller's
                                                          buffer here (:11-15); not reproduced */
  const double hs = (image.br_x - image.tl_x) / (image.data.columns - 1);
  const double vs = (image.br_y - image.tl_y) / (image.data.lines - 1);
  for (int e = 0; e < desc.lines; ++e) {
    const double val = AT(desc, e, 0), a = AT(desc, e, 1), b = AT(desc, e, 2);
    const double x0 = AT(desc, e, 3), y0 = AT(desc, e, 4), rot = AT(desc, e, 5);
    const double c = cos(rot), s = sin(rot);
    for (int i = 0; i < image.data.lines; ++i)
      for (int j = 0; j < image.data.columns; ++j) {
        const double x = image.tl_x + j * hs - x0;
        const double y = image.tl_y + i * vs - y0;
        double p = (y * s) + (x * c);
        p /= a;
        p *= p;
        double q = (y * c) - (x * s);
        q /= b;
        q *= q;
        p += q;
        p = sqrt(p);
        if (p <= 1) AT(image.data, i, j) += val;
      }
  }
}

/* exact Radon transform of an ellipse set (raft/raft_radon_fromdesc.c:4-67) */
void oracle_radon_fromdesc(raft_image sino, raft_matrix desc) {
  const double hs = (sino.br_x - sino.tl_x) / (sino.data.columns - 1);
  const double vs = (sino.br_y - sino.tl_y) / (sino.data.lines - 1);
  for (int e = 0; e < desc.lines; ++e) {
    const double val = AT(desc, e, 0), a = AT(desc, e, 1), b = AT(desc, e, 2);
    const double rot = AT(desc, e, 5), ex = AT(desc, e, 3), ey = AT(desc, e, 4);
    for (int j = 0; j < sino.data.columns; ++j) {
      double th = sino.tl_x + j * hs;
      const double c = cos(th), s = sin(th);
      th -= rot;
      const double cp = cos(th), sp = sin(th);
      double m = 1.0 / sqrt(((cp * cp) / (b * b)) + ((sp * sp) / (a * a)));
      double mab = m / (a * b);
      m *= 2.0;
      m *= val;
      const double delt = mab * ((ex * c) + (ey * s) - sino.tl_y);
      mab *= vs;
      for (int i = 0; i < sino.data.lines; ++i) {
        double tp = mab * i - delt;
        tp *= tp;
        if (tp <= 1.0) AT(sino.data, i, j) += m * sqrt(1.0 - tp);
      }
    }
  }
}

/* ------------------------------------------------------------------ */
/* sample maps for checking index logic (not a reference routine)       */
/* ------------------------------------------------------------------ */

/* mask[(j*R + k)*N + s] = 1 iff sample s of the ray at t[k], theta[j] passes the
   reference's 0 <= m <= N-2 test (raft/raft_radon_slantstack_withmesh.c:184-186, 203-205);
   coord[...] = the reference's FP64 interpolation coordinate before floor(). */
void oracle_radon_sample_map_mesh(raft_image img, int nrays, const double *t, int nviews, const double *theta,
                                  unsigned char *mask, double *coord) {
  const int n = img.data.lines;
  const double tol = 1.0 / sqrt(2);
  const double xmin = img.tl_x, xmax = img.br_x, ymin = img.br_y, ymax = img.tl_y;
  const double dx = (xmax - xmin) / (n - 1), dy = (ymax - ymin) / (n - 1);
  for (int j = 0; j < nviews; ++j) {
    const double c = cos(theta[j]), s = sin(theta[j]), tn = tan(theta[j]);
    const int branch_a = fabs(c) < tol;
    for (int k = 0; k < nrays; ++k) {
      for (int i = 0; i < n; ++i) {
        double u;
        if (branch_a) {
          const double x = (double)(xmin + i * dx);
          const double y = (double)(t[k] / s - x / tn);
          u = (y - ymin) / dy;
        } else {
          const double y = (double)(xmin + i * dy);
          const double x = (double)(t[k] / c - y * tn);
          u = (x - xmin) / dx;
        }
        const int m = (int)floor(u);
        const size_t idx = ((size_t)j * nrays + k) * n + i;
        mask[idx] = (m > -1 && m < n - 1);
        if (coord) coord[idx] = u;
      }
    }
  }
}

/* the same map on the uniform grid theta = th0 + j*dth, t = tmin + k*dt
   (raft/raft_radon_slantstack.c:153-172) */
void oracle_radon_sample_map(raft_image img, raft_image sino, unsigned char *mask, double *coord) {
  const int nviews = sino.data.columns, nrays = sino.data.lines;
  const double th0 = sino.tl_x, thf = sino.br_x, dth = (thf - th0) / (nviews);
  const double tmin = sino.br_y, tmax = sino.tl_y, dt = (tmax - tmin) / (nrays - 1);
  double *t = (double *)malloc(sizeof(double) * nrays);
  double *theta = (double *)malloc(sizeof(double) * nviews);
  for (int k = 0; k < nrays; ++k) t[k] = tmin + k * dt;
  for (int j = 0; j < nviews; ++j) theta[j] = th0 + j * dth;
  oracle_radon_sample_map_mesh(img, nrays, t, nviews, theta, mask, coord);
  free(t);
  free(theta);
}

/* mask[(j*N + r)*N + c] = 1 iff pixel (r,c) takes angle j in the backprojection
   (raft/raft_backprojection_slantstack.c:170-174) */
void oracle_bp_sample_map(raft_image sino, raft_image img, unsigned char *mask) {
  const int n = img.data.lines, nviews = sino.data.columns, nrays = sino.data.lines;
  const double tmin = sino.br_y, tmax = sino.tl_y, dt = (tmax - tmin) / (nrays - 1);
  const double th0 = sino.tl_x, thf = sino.br_x, dth = (thf - th0) / (nviews);
  const double xmin = img.tl_x, xmax = img.br_x, dx = (xmax - xmin) / (n - 1);
  const double ymin = img.br_y, ymax = img.tl_y, dy = (ymax - ymin) / (n - 1);
  for (int j = 0; j < nviews; ++j) {
    const double th = th0 + j * dth;
    const double c = cos(th), s = sin(th);
    for (int r = 0; r < n; ++r)
      for (int col = 0; col < n; ++col) {
        const double x = xmin + col * dx, y = ymin + r * dy;
        const double t = x * c + y * s;
        const int m = (int)floor((t - tmin) / dt);
        mask[((size_t)j * n + r) * n + col] = (m > -1 && m < nrays - 1);
      }
  }
}

/* ------------------------------------------------------------------ */
/* line-traversal (Bresenham / DDA) projector pair                      */
/*   oracle_radon_bresenham            raft/raft_radon_bresenham.cpp:17-105  */
/*   oracle_radon_transpose_bresenham  raft/raft_radon_transpose_bresenham.cpp */
/*   oracle_backprojection_bresenham   raft/raft_backprojection_bresenham.cpp:17-251 */
/* with the traversal of raft/raft_ddaiterator.c:16-257.                 */
/*                                                                      */
/* Geometry (raft_image corners, signed pitches):                        */
/*   line i <-> y = tl_y + i*hy, hy = (br_y - tl_y)/(lines - 1)          */
/*   column j <-> x = tl_x + j*hx, hx = (br_x - tl_x)/(columns - 1)      */
/*   angle a <-> theta = stl_x + a*(sbr_x - stl_x)/(A - 1)               */
/*   ray r   <-> t = stl_y + r*(sbr_y - stl_y)/(R - 1)                   */
/* The ray {x cos + y sin = t} walks the pixel grid; each visited pixel  */
/* gets its intersection length.                                         */
/* ------------------------------------------------------------------ */

typedef struct {
  int lines, cols;
  double hx, hy;                    /* signed pixel pitches */
  double ox, oy;                    /* tl corner (pixel centre of line 0 / column 0) */
  double bx0, bx1, by0, by1;        /* pixel-edge bounding box: tl_x - hx/2, br_x + hx/2, tl_y - hy/2, br_y + hy/2 */
  /* per angle */
  double c, s;
  double xe, ye;                    /* edge lines a ray may enter through */
  int di, dj;                       /* index steps along the ray */
  int i0, j0;                       /* entry line / column on those edges */
  int fast_i;                       /* 1: the line index advances every step */
  double e_step, e_cap, len_full;   /* slow-axis drift per step, pixel width across, full length */
  int dfast, dslow;
  /* per ray */
  int i, j;
  double len, e;
  int slow_next;
} trace_t;

static double sgn(double v) { return v > 0.0 ? 1.0 : (v < 0.0 ? -1.0 : 0.0); }

/* double -> int as the reference's assignment compiles on x86-64 (cvttsd2si):
   NaN and out-of-range values give INT_MIN, i.e. an index outside the image */
static int to_index(double v) {
  if (!(v >= -2147483648.0 && v < 2147483648.0)) return (int)0x80000000u;
  return (int)v;
}

/* pixel index of coordinate v on an axis starting at v0 with pitch h
   (raft/raft_ddaiterator.c:10-14) */
static int axis_index(double v, double v0, double h, int dir) {
  if (dir >= 0) return to_index(floor(((v - v0) / h) - 0.5) + 1);
  return to_index(ceil(((v - v0) / h) - 0.5));
}

static void trace_grid(trace_t *tr, const raft_image *img) {
  tr->lines = img->data.lines;
  tr->cols = img->data.columns;
  tr->hx = (img->br_x - img->tl_x) / (tr->cols - 1);
  tr->hy = (img->br_y - img->tl_y) / (tr->lines - 1);
  tr->bx0 = img->tl_x - (tr->hx / 2.0);
  tr->bx1 = img->br_x + (tr->hx / 2.0);
  tr->by0 = img->tl_y - (tr->hy / 2.0);
  tr->by1 = img->br_y + (tr->hy / 2.0);
  tr->ox = img->tl_x;
  tr->oy = img->tl_y;
}

/* raft/raft_ddaiterator.c:38-106 */
static void trace_angle(trace_t *tr, double theta) {
  tr->c = cos(theta);
  tr->s = sin(theta);
  int sg = (int)sgn(tr->s);
  double p = sg * tr->bx1, q = sg * tr->bx0;
  tr->xe = sg * (p > q ? p : q);
  tr->dj = (int)(-sgn(tr->hx) * sg);
  sg = (int)sgn(tr->c);
  p = sg * tr->by1;
  q = sg * tr->by0;
  tr->ye = sg * (p < q ? p : q);
  tr->di = (int)(sgn(tr->hy) * sg);
  tr->i0 = tr->di < 0 ? tr->lines - 1 : 0;
  tr->j0 = tr->dj < 0 ? tr->cols - 1 : 0;
  tr->fast_i = fabs(tr->hx * tr->s) <= fabs(tr->hy * tr->c);
  if (tr->fast_i) {
    tr->e_step = fabs(tr->hy * tr->s / tr->c);
    tr->e_cap = fabs(tr->hx);
    tr->dslow = tr->dj;
    tr->dfast = tr->di;
    tr->len_full = fabs(tr->hy / tr->c);
  } else {
    tr->e_step = fabs(tr->hx * tr->c / tr->s);
    tr->e_cap = fabs(tr->hy);
    tr->dslow = tr->di;
    tr->dfast = tr->dj;
    tr->len_full = fabs(tr->hx / tr->s);
  }
}

/* one move along the ray (raft/raft_ddaiterator.c:196-248) */
static void trace_step(trace_t *tr) {
  if (tr->slow_next) {
    if (tr->fast_i) tr->j += tr->dslow; else tr->i += tr->dslow;
    tr->len = tr->len_full - tr->len;
    tr->slow_next = 0;
    return;
  }
  if (tr->fast_i) tr->i += tr->dfast; else tr->j += tr->dfast;
  const double e_prev = tr->e;
  tr->e += tr->e_step;
  tr->slow_next = tr->e > tr->e_cap;
  if (tr->slow_next) {
    double a;
    tr->e -= tr->e_cap;
    if (tr->e) {
      a = (tr->e_cap - e_prev) / tr->e;
      a /= (1.0 + a);
    } else {
      a = 1.0;
    }
    tr->len = a * tr->len_full;
  } else {
    tr->len = tr->len_full;
  }
}

static int trace_outside(const trace_t *tr) {
  return tr->i < 0 || tr->i >= tr->lines || tr->j < 0 || tr->j >= tr->cols;
}

/* entry pixel and first intersection of the ray at t (raft/raft_ddaiterator.c:108-194).
   Axis-parallel rays start with the full pixel length: the reference leaves the
   length of their first pixel as the previous ray left it (uninitialised for a
   thread's first ray) and carries len_full into every later ray of the angle. */
static void trace_ray(trace_t *tr, double t) {
  tr->slow_next = 0;
  double sx, sy, xy;
  if (tr->s == 0.0) {
    tr->j = axis_index(t, tr->ox, tr->hx, tr->dj);
    tr->i = tr->i0;
    tr->e = 0.0;
    tr->e_step = 0.0;
    tr->len = tr->len_full;
  } else if (tr->c == 0.0) {
    tr->i = axis_index(t, tr->oy, tr->hy, tr->di);
    tr->j = tr->j0;
    tr->e = 0.0;
    tr->e_step = 0.0;
    tr->len = tr->len_full;
  } else if ((sx = ((t * tr->c - tr->xe) / tr->s)) > (sy = ((tr->ye - t * tr->s) / tr->c))) {
    /* enters through the vertical edge x = xe */
    xy = (sx * tr->c) + (t * tr->s);
    tr->i = axis_index(xy, tr->oy, tr->hy, tr->di);
    tr->j = tr->j0;
    tr->e = fabs(xy - (tr->oy + (tr->i - (0.5 * tr->di)) * tr->hy));
    if (!tr->fast_i) {
      tr->j -= tr->dj;
      trace_step(tr);
    } else {
      tr->e -= fabs(tr->hy);
      tr->e /= tr->c;
      tr->len = fabs(tr->e);
      tr->e = fabs(tr->len * tr->s);
    }
  } else {
    /* enters through the horizontal edge y = ye */
    xy = (-sy * tr->s) + (t * tr->c);
    tr->j = axis_index(xy, tr->ox, tr->hx, tr->dj);
    tr->i = tr->i0;
    tr->e = fabs(xy - (tr->ox + (tr->j - (0.5 * tr->dj)) * tr->hx));
    if (tr->fast_i) {
      tr->i -= tr->di;
      trace_step(tr);
    } else {
      tr->e -= fabs(tr->hx);
      tr->e /= tr->s;
      tr->len = fabs(tr->e);
      tr->e = fabs(tr->len * tr->c);
    }
  }
}

/* a ray never visits more than lines + columns pixels (every move advances one
   index monotonically); the cap only matters for non-finite geometry */
#define TRACE_CAP(tr) ((long)(tr)->lines + (tr)->cols + 4)

typedef struct {
  raft_image img, sino;
  double *partial, *scale; /* per-chunk accumulators (transpose / backprojection) */
  int mode;                /* 0 radon, 1 transpose, 2 backprojection */
  long nchunks;
  int base, extra;         /* angle partition of the reference's threads */
} dda_ctx;

static void dda_angles(const dda_ctx *ctx, long chunk, int *a0, int *a1) {
  *a0 = (int)(chunk * ctx->base + (chunk < ctx->extra ? chunk : ctx->extra));
  *a1 = *a0 + ctx->base + (chunk < ctx->extra);
}

static void dda_work(void *vctx, long b, long e) {
  dda_ctx *ctx = (dda_ctx *)vctx;
  const raft_image *sino = &ctx->sino;
  const int R = sino->data.lines, A = sino->data.columns;
  const double dth = (sino->br_x - sino->tl_x) / (A - 1);
  const double dt = (sino->br_y - sino->tl_y) / (R - 1);
  trace_t tr;
  memset(&tr, 0, sizeof tr);
  trace_grid(&tr, &ctx->img);
  const long npix = (long)tr.lines * tr.cols;
  for (long ch = b; ch < e; ++ch) {
    int a0, a1;
    dda_angles(ctx, ch, &a0, &a1);
    double *acc = ctx->mode ? ctx->partial + ch * npix : NULL;
    double *sc = ctx->mode == 2 ? ctx->scale + ch * npix : NULL;
    for (int a = a0; a < a1; ++a) {
      trace_angle(&tr, sino->tl_x + (a * dth));
      for (int r = 0; r < R; ++r) {
        trace_ray(&tr, sino->tl_y + (r * dt));
        if (ctx->mode == 0) {
          double sum = 0.0;
          for (long it = 0; it < TRACE_CAP(&tr) && !trace_outside(&tr); ++it) {
            sum += AT(ctx->img.data, tr.i, tr.j) * tr.len;
            trace_step(&tr);
          }
          AT(sino->data, r, a) = sum;
        } else {
          const double q = AT(sino->data, r, a);
          for (long it = 0; it < TRACE_CAP(&tr) && !trace_outside(&tr); ++it) {
            acc[(long)tr.i * tr.cols + tr.j] += q * tr.len;
            if (sc) sc[(long)tr.i * tr.cols + tr.j] += tr.len;
            trace_step(&tr);
          }
        }
      }
    }
  }
}

/* nthreads chunks of consecutive angles, the first A mod nthreads one angle longer
   (raft/raft_radon_bresenham.cpp:68-101); the partial images are summed into the
   output in chunk order, as the reference's per-column daxpy does
   (raft/raft_backprojection_bresenham.cpp:17-56). */
static void dda_run(raft_image img, raft_image sino, int mode, int nthreads) {
  if (img.data.lines < 2 || img.data.columns < 2 || sino.data.lines < 2 || sino.data.columns < 2) return;
  if (nthreads < 1) nthreads = 1;
  dda_ctx ctx;
  memset(&ctx, 0, sizeof ctx);
  ctx.img = img;
  ctx.sino = sino;
  ctx.mode = mode;
  ctx.nchunks = nthreads;
  ctx.base = sino.data.columns / nthreads;
  ctx.extra = sino.data.columns % nthreads;
  const long npix = (long)img.data.lines * img.data.columns;
  if (mode) ctx.partial = (double *)calloc((size_t)nthreads * npix, sizeof(double));
  if (mode == 2) ctx.scale = (double *)calloc((size_t)nthreads * npix, sizeof(double));
  fan_out(dda_work, &ctx, nthreads, nthreads);
  if (mode) {
    for (long p = 0; p < npix; ++p) {
      const long i = p / img.data.columns, j = p % img.data.columns;
      double v = AT(img.data, i, j), w = 0.0;
      for (long ch = 0; ch < nthreads; ++ch) {
        v += ctx.partial[ch * npix + p];
        if (mode == 2) w += ctx.scale[ch * npix + p];
      }
      AT(img.data, i, j) = mode == 2 ? v / w : v;
    }
  }
  free(ctx.partial);
  free(ctx.scale);
}

/* sinogram (R x A) of the image; every element overwritten */
void oracle_radon_bresenham(raft_image image, raft_image sino, int nthreads) {
  dda_run(image, sino, 0, nthreads);
}

/* image += R^T sino (adds into the caller's image, as the reference does) */
void oracle_radon_transpose_bresenham(raft_image sino, raft_image image, int nthreads) {
  dda_run(image, sino, 1, nthreads);
}

/* image = (image + R^T sino) / R^T 1, unguarded: pixels no ray reaches get 0/0 */
void oracle_backprojection_bresenham(raft_image sino, raft_image image, int nthreads) {
  dda_run(image, sino, 2, nthreads);
}

/* the pixels ray r of angle a visits, in order: ij[2k], ij[2k+1] = line, column and
   len[k] = intersection length; returns the count (at most cap written) */
int oracle_dda_ray_path(raft_image img, raft_image sino, int a, int r, int *ij, double *len, int cap) {
  const int R = sino.data.lines, A = sino.data.columns;
  if (img.data.lines < 2 || img.data.columns < 2 || R < 2 || A < 2) return 0;
  trace_t tr;
  memset(&tr, 0, sizeof tr);
  trace_grid(&tr, &img);
  trace_angle(&tr, sino.tl_x + (a * ((sino.br_x - sino.tl_x) / (A - 1))));
  trace_ray(&tr, sino.tl_y + (r * ((sino.br_y - sino.tl_y) / (R - 1))));
  int count = 0;
  for (long it = 0; it < TRACE_CAP(&tr) && !trace_outside(&tr); ++it) {
    if (count < cap) {
      ij[2 * count] = tr.i;
      ij[2 * count + 1] = tr.j;
      len[count] = tr.len;
    }
    ++count;
    trace_step(&tr);
  }
  return count;
}
