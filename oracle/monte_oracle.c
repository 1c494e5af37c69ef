/*
 * monte_oracle.c -- FP64 CPU restatement of Tomato27/Monte's hot paths.
 *
 * TEST INFRASTRUCTURE ONLY (see monte_oracle.h).  The product never links it.
 *
 * Pinning (what checks each function against something other than itself) is
 * listed in tests/test_oracle_*.py; each function's reference lines are cited
 * at its definition.
 */
#define _GNU_SOURCE
#include "monte_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* =============================================================== threads */
typedef void (*orc_range_fn)(void *ctx, long begin, long end);
typedef struct {
  orc_range_fn fn;
  void *ctx;
  long n, chunk;
  long next;
  pthread_mutex_t mu;
} orc_pool;

static void *orc_worker(void *arg) {
  orc_pool *p = (orc_pool *)arg;
  for (;;) {
    pthread_mutex_lock(&p->mu);
    long b = p->next;
    p->next += p->chunk;
    pthread_mutex_unlock(&p->mu);
    if (b >= p->n) break;
    long e = b + p->chunk < p->n ? b + p->chunk : p->n;
    p->fn(p->ctx, b, e);
  }
  return NULL;
}

/* dynamic parallel-for over [0,n) in chunks */
static void orc_parallel_for(long n, long chunk, int nthreads, orc_range_fn fn, void *ctx) {
  if (nthreads < 1) nthreads = 1;
  if (chunk < 1) chunk = 1;
  orc_pool p = {fn, ctx, n, chunk, 0, PTHREAD_MUTEX_INITIALIZER};
  if (nthreads == 1) {
    orc_worker(&p);
    return;
  }
  pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * (size_t)nthreads);
  for (int i = 0; i < nthreads; ++i) pthread_create(&th[i], NULL, orc_worker, &p);
  for (int i = 0; i < nthreads; ++i) pthread_join(th[i], NULL);
  free(th);
}

/* =================================================================== RNG */
/* Philox4x32-10, Salmon, Moraes, Dror & Shaw, "Parallel random numbers: as
 * easy as 1, 2, 3", SC'11: ten rounds of a multiply-xor S-box over a 128-bit
 * counter with a Weyl-sequence key schedule. */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
  const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
  const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
  uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
  uint32_t k0 = key[0], k1 = key[1];
  for (int r = 0; r < 10; ++r) {
    uint64_t p0 = (uint64_t)M0 * c0;
    uint64_t p1 = (uint64_t)M1 * c2;
    uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
    uint32_t n1 = (uint32_t)p1;
    uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
    uint32_t n3 = (uint32_t)p0;
    c0 = n0; c1 = n1; c2 = n2; c3 = n3;
    k0 += W0; k1 += W1;
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

double orc_u01(uint32_t x) { return ((double)(x >> 8) + 0.5) * (1.0 / 16777216.0); }

/* Per-photon uniform stream: block b of photon n in view k of job j is
 * Philox4x32-10 at counter (b, n, k, j) with key (seed lo, seed hi).  Draw
 * order, shared with the product kernel:
 *   - the spectrum draw (if any) is the first uniform of block 0;
 *   - every Woodcock flight takes whole blocks, two steps per block as
 *     (distance, accept, distance, accept); a flight ending on the first
 *     step of a block leaves the other two unused;
 *   - every interaction starts on a fresh block and draws in sequence; the
 *     rest of its last block is unused (the next flight takes a new block). */
typedef struct {
  uint32_t ctr[4];
  uint32_t key[2];
  uint32_t buf[4];
  int idx;
} orc_stream;

static void orc_stream_init(orc_stream *s, uint32_t photon, uint32_t view, uint32_t job,
                            uint32_t seed_lo, uint32_t seed_hi) {
  s->ctr[0] = 0; s->ctr[1] = photon; s->ctr[2] = view; s->ctr[3] = job;
  s->key[0] = seed_lo; s->key[1] = seed_hi;
  s->idx = 4;
}

static double orc_next(orc_stream *s) {
  if (s->idx == 4) {
    orc_philox4x32_10(s->ctr, s->key, s->buf);
    s->ctr[0] += 1;
    s->idx = 0;
  }
  return orc_u01(s->buf[s->idx++]);
}

/* Whole next block as four uniforms; sequential draws after it start on a
 * fresh block. */
static void orc_block(orc_stream *s, double u[4]) {
  uint32_t b[4];
  orc_philox4x32_10(s->ctr, s->key, b);
  s->ctr[0] += 1;
  s->idx = 4;
  for (int i = 0; i < 4; ++i) u[i] = orc_u01(b[i]);
}

/* ================================================================ tables */
/* monte_cpp/CBCT_real2.cpp:633-668 reads the CSV with row k stored at index k
 * (k keV) and columns (coherent, incoherent, photoelectric, total).  The BOM
 * on row 1 made the original hard-code 1.372 (quirk Q1); here the BOM and CR
 * are stripped and every row is parsed. */
int orc_read_xcom_csv(const char *path, double tab[4][ORC_NROWS]) {
  FILE *f = fopen(path, "rb");
  if (!f) return -1;
  char line[512];
  int k = 0;
  memset(tab, 0, sizeof(double) * 4 * ORC_NROWS);
  while (fgets(line, sizeof line, f)) {
    char *p = line;
    if ((unsigned char)p[0] == 0xEF && (unsigned char)p[1] == 0xBB && (unsigned char)p[2] == 0xBF) p += 3;
    double v[4];
    int n = 0;
    char *q = p;
    while (n < 4) {
      char *end;
      v[n] = strtod(q, &end);
      if (end == q) break;
      ++n;
      q = end;
      while (*q == ',' || *q == ' ' || *q == '\t') ++q;
    }
    if (n != 4) continue;
    if (k + 1 >= ORC_NROWS) break;
    ++k;
    for (int c = 0; c < 4; ++c) tab[c][k] = v[c];
  }
  fclose(f);
  for (int c = 0; c < 4; ++c) tab[c][0] = tab[c][1];
  return k;
}

/* ============================================================ bp3d20 FDK */
/* recon/bp3d20_325.cpp:39-46 (and bp3d20.cpp:36-43):
 *   map_w = map * D / sqrt(D^2 + (H/2*pitch - r*pitch)^2 + (c*pitch - H/2*pitch)^2)
 * with D = 60 (the object-detector distance) and pixel EDGE coordinates. */
void orc_bp3d20_weight(const orc_geom *g, const float *map, double *map_w) {
  const int H = g->nu, W = g->nv;
  const double half = 0.5 * H * g->du, D = g->odd;
  for (int k = 0; k < g->n_views; ++k)
    for (int r = 0; r < H; ++r)
      for (int c = 0; c < W; ++c) {
        double a = -1.0 * (r * g->du) + half;
        double b = (c * g->dv) - 0.5 * W * g->dv;
        size_t i = ((size_t)k * H + r) * W + c;
        map_w[i] = map[i] * (D / sqrt(D * D + a * a + b * b));
      }
}

/* Ram-Lak kernel of recon/bp3d20_325.cpp:51-63: h[0]=1/4, h[even]=0,
 * h[odd n] = -1/(n pi)^2, |n| <= W-1 */
static double orc_ramlak_unit(int n) {
  if (n == 0) return 0.25;
  if (n % 2 == 0) return 0.0;
  return -1.0 / ((n * M_PI) * (n * M_PI));
}

/* Shepp-Logan kernel at unit spacing: h[n] = -2 / (pi^2 (4 n^2 - 1)) */
static double orc_shepp_unit(int n) { return -2.0 / (M_PI * M_PI * (4.0 * n * n - 1.0)); }

/* recon/bp3d20_325.cpp:68-81: out[k][d][b] = sum_c map_w[k][c][d] * 0.5 * h[c - b]
 * (filter along the first detector index, written transposed). */
void orc_bp3d20_filter(const orc_geom *g, const double *map_w, double *map_out) {
  const int H = g->nu, W = g->nv;
  double *h = (double *)malloc(sizeof(double) * (size_t)(2 * H - 1));
  for (int n = -(H - 1); n <= H - 1; ++n) h[n + H - 1] = orc_ramlak_unit(n < 0 ? -n : n);
  for (int k = 0; k < g->n_views; ++k)
    for (int d = 0; d < W; ++d)
      for (int b = 0; b < H; ++b) {
        double acc = 0.0;
        for (int c = 0; c < H; ++c) acc += map_w[((size_t)k * H + c) * W + d] * 0.5 * h[H - 1 - b + c];
        map_out[((size_t)k * W + d) * H + b] = acc;
      }
  free(h);
}

typedef struct {
  const orc_geom *g;
  const double *map_out;
  double *vol;
} orc_bp3d20_ctx;

/* recon/bp3d20_325.cpp:91-175 over z-slices [z0,z1) for all views */
static void orc_bp3d20_slices(void *vctx, long z0, long z1) {
  orc_bp3d20_ctx *c = (orc_bp3d20_ctx *)vctx;
  const orc_geom *g = c->g;
  const int H = g->nu, W = g->nv;
  const double pitch = g->du, half = 0.5 * H * pitch;
  const double vox = g->vx;
  const double x_start = g->ox - 0.5 * g->nx * vox;   /* -12.8 */
  const double y_start = g->oy + 0.5 * g->ny * g->vy; /* +12.8 */
  const double z_start = g->oz + 0.5 * g->nz * g->vz; /* +12.8 */
  for (int k = 0; k < g->n_views; ++k) {
    const int beta = k * g->step_deg; /* integer degrees */
    const double start_x = -g->sod, start_y = 0.0;
    const double primary_x = start_x * cos(M_PI * beta / 180) - start_y * sin(M_PI * beta / 180);
    const double primary_y = start_x * sin(M_PI * beta / 180) + start_y * cos(M_PI * beta / 180);
    const size_t kbase = (size_t)k * H * W, total = (size_t)g->n_views * H * W;
    for (long z = z0; z < z1; ++z)
      for (int t = 0; t < g->ny; ++t)
        for (int s = g->roi_x0; s < g->roi_x1; ++s) {
          double pv[3];
          pv[0] = (x_start + s * vox) - primary_x;
          pv[1] = (y_start - t * g->vy) - primary_y;
          pv[2] = (z_start - z * g->vz);
          double tmp_x = pv[0];
          pv[0] = pv[0] * cos(-1 * M_PI * beta / 180) - pv[1] * sin(-1 * M_PI * beta / 180);
          pv[1] = tmp_x * sin(-1 * M_PI * beta / 180) + pv[1] * cos(-1 * M_PI * beta / 180);
          double to_det = g->sdd / pv[0];
          for (int q = 0; q < 3; ++q) pv[q] *= to_det;
          if (fabs(pv[1]) > half || fabs(pv[2]) > half) continue;
          double y = -1 * (pv[1] / pitch - H / 2.0);
          double x = -1 * (pv[2] / pitch - W / 2.0);
          int yi = (int)y, xi = (int)x;
          double tmp_s = (x_start + s * vox) * cos(M_PI * beta / 180) - (y_start - t * g->vy) * sin(M_PI * beta / 180);
          double tmp_t = (x_start + s * vox) * sin(M_PI * beta / 180) + (y_start - t * g->vy) * cos(M_PI * beta / 180);
          double tb = tan((double)beta); /* integer degrees used as radians (quirk Q15) */
          double d = fabs(-1 * tb * tmp_t + 1 * tmp_s) / sqrt(1 + tb * tb);
          if (tmp_s < 0) d *= -1;
          if (g->mask_r >= 0) {
            long cz = g->nz / 2, cy = g->ny / 2, cx = g->nx / 2;
            long r2 = (z - cz) * (z - cz) + (long)(t - cy) * (t - cy) + (long)(s - cx) * (s - cx);
            if ((double)r2 > g->mask_r * g->mask_r) continue;
          }
          double val = 0.0;
          if (0 <= x && x <= W && 0 <= y && y <= H) {
            /* :163-166 index the flat array map_out[beta*H*W + xi*H + yi]; a
             * tap at yi+1 == H (or xi+1 == W) therefore reads the first
             * element of the next row (or the first row of the next view),
             * quirk Q18.  Restated literally; past the end of the whole stack
             * (the last view) the reference reads outside its allocation, which
             * is restated as zero. */
#define MO(xx, yy) ((kbase + (size_t)(xx) * H + (yy)) < total ? c->map_out[kbase + (size_t)(xx) * H + (yy)] : 0.0)
            val = (yi + 1 - y) * ((xi + 1 - x) * MO(xi, yi) + (x - xi) * MO(xi + 1, yi)) +
                  (y - yi) * ((xi + 1 - x) * MO(xi, yi + 1) + (x - xi) * MO(xi + 1, yi + 1));
#undef MO
          }
          double out = (g->odd * g->odd / ((g->odd - d) * (g->odd - d))) * val * g->step_deg * 2 * M_PI / 360;
          c->vol[((size_t)z * g->ny + t) * g->nx + s] += out * g->scale;
        }
  }
}

void orc_bp3d20_backproject(const orc_geom *g, const double *map_out, double *vol_xy, int nthreads) {
  orc_bp3d20_ctx c = {g, map_out, vol_xy};
  orc_parallel_for(g->nz, 4, nthreads, orc_bp3d20_slices, &c);
}

/* ========================================================= standard FDK */
/* Detector index convention shared by the standard preset: pixel iu covers
 * u_idx in [iu, iu+1) with u_idx = nu/2 + (y' - u_off)/du (y' the in-plane
 * detector coordinate in the rotating frame), likewise v with z. */
static double std_pix_u(const orc_geom *g, double idx) { return (idx - 0.5 * g->nu) * g->du + g->u_off; }
static double std_pix_v(const orc_geom *g, double idx) { return (idx - 0.5 * g->nv) * g->dv + g->v_off; }

void orc_fdk_std_weight(const orc_geom *g, const double *proj, double *proj_w) {
  for (int k = 0; k < g->n_views; ++k)
    for (int iu = 0; iu < g->nu; ++iu)
      for (int iv = 0; iv < g->nv; ++iv) {
        double a = std_pix_u(g, iu + 0.5), b = std_pix_v(g, iv + 0.5);
        size_t i = ((size_t)k * g->nu + iu) * g->nv + iv;
        proj_w[i] = proj[i] * g->sdd / sqrt(g->sdd * g->sdd + a * a + b * b);
      }
}
void orc_fdk_std_weight_T(const orc_geom *g, const double *proj_w, double *proj) {
  orc_fdk_std_weight(g, proj_w, proj); /* diagonal */
}

static double std_filter_tap(const orc_geom *g, int n) {
  n = n < 0 ? -n : n;
  return g->filter == ORC_FILTER_SHEPP_LOGAN ? orc_shepp_unit(n) : orc_ramlak_unit(n);
}
/* Kak & Slaney (3.?) equispaced fan/cone: Q = tau * sum R' g, g = h/2, h at
 * spacing tau = du*sod/sdd (detector scaled to the isocentre) */
static double std_filter_scale(const orc_geom *g) {
  double tau = g->du * g->sod / g->sdd;
  return 1.0 / (2.0 * tau);
}

void orc_fdk_std_filter(const orc_geom *g, const double *pw, double *filt) {
  const int nu = g->nu, nv = g->nv;
  const double cs = std_filter_scale(g);
  for (int k = 0; k < g->n_views; ++k)
    for (int iv = 0; iv < nv; ++iv)
      for (int b = 0; b < nu; ++b) {
        double acc = 0.0;
        for (int c = 0; c < nu; ++c) acc += pw[((size_t)k * nu + c) * nv + iv] * std_filter_tap(g, c - b);
        filt[((size_t)k * nv + iv) * nu + b] = cs * acc;
      }
}

void orc_fdk_std_filter_T(const orc_geom *g, const double *filt, double *pw) {
  const int nu = g->nu, nv = g->nv;
  const double cs = std_filter_scale(g);
  for (int k = 0; k < g->n_views; ++k)
    for (int iv = 0; iv < nv; ++iv)
      for (int c = 0; c < nu; ++c) {
        double acc = 0.0;
        for (int b = 0; b < nu; ++b) acc += filt[((size_t)k * nv + iv) * nu + b] * std_filter_tap(g, c - b);
        pw[((size_t)k * nu + c) * nv + iv] = cs * acc;
      }
}

typedef struct {
  const orc_geom *g;
  const double *filt;
  double *vol;
  int transpose;
  double *filt_out; /* per-thread accumulation for the transpose: guarded by views */
} orc_std_bp_ctx;

/* One view's contribution: for each voxel, bilinear sample of the filtered
 * view at the voxel's projection, weighted by (sod/U)^2 * dangle. */
static void std_bp_view(const orc_geom *g, int k, const double *filt_k, double *vol, double *filt_out_k,
                        long z0, long z1) {
  const int nu = g->nu, nv = g->nv;
  const double beta = g->angle0 + k * g->dangle;
  const double cb = cos(beta), sb = sin(beta);
  for (long z = z0; z < z1; ++z) {
    const double Z = (z - 0.5 * (g->nz - 1)) * g->vz + g->oz;
    for (int y = 0; y < g->ny; ++y) {
      const double Y = (y - 0.5 * (g->ny - 1)) * g->vy + g->oy;
      for (int x = 0; x < g->nx; ++x) {
        const double X = (x - 0.5 * (g->nx - 1)) * g->vx + g->ox;
        const double xr = X * cb + Y * sb, yr = -X * sb + Y * cb;
        const double U = g->sod + xr;
        const double yd = g->sdd * yr / U, zd = g->sdd * Z / U;
        const double su = 0.5 * nu + (yd - g->u_off) / g->du - 0.5;
        const double sv = 0.5 * nv + (zd - g->v_off) / g->dv - 0.5;
        const double w = (g->sod / U) * (g->sod / U) * g->dangle;
        const double fu = floor(su), fv = floor(sv);
        const int iu = (int)fu, iv = (int)fv;
        const double au = su - fu, av = sv - fv;
        const double wt[4] = {(1 - av) * (1 - au), (1 - av) * au, av * (1 - au), av * au};
        const int tu[4] = {iu, iu + 1, iu, iu + 1};
        const int tv[4] = {iv, iv, iv + 1, iv + 1};
        size_t vi = ((size_t)z * g->ny + y) * g->nx + x;
        for (int q = 0; q < 4; ++q) {
          if (tu[q] < 0 || tu[q] >= nu || tv[q] < 0 || tv[q] >= nv) continue;
          size_t pi = (size_t)tv[q] * nu + tu[q];
          if (filt_out_k)
            filt_out_k[pi] += w * wt[q] * vol[vi];
          else
            vol[vi] += w * wt[q] * filt_k[pi];
        }
      }
    }
  }
}

static void std_bp_slices(void *vctx, long z0, long z1) {
  orc_std_bp_ctx *c = (orc_std_bp_ctx *)vctx;
  for (int k = 0; k < c->g->n_views; ++k)
    std_bp_view(c->g, k, c->filt + (size_t)k * c->g->nu * c->g->nv, c->vol, NULL, z0, z1);
}

static void std_bpT_views(void *vctx, long k0, long k1) {
  orc_std_bp_ctx *c = (orc_std_bp_ctx *)vctx;
  for (long k = k0; k < k1; ++k)
    std_bp_view(c->g, (int)k, NULL, c->vol, c->filt_out + (size_t)k * c->g->nu * c->g->nv, 0, c->g->nz);
}

void orc_fdk_std_backproject(const orc_geom *g, const double *filt, double *vol, int nthreads) {
  orc_std_bp_ctx c = {g, filt, vol, 0, NULL};
  orc_parallel_for(g->nz, 1, nthreads, std_bp_slices, &c);
}

void orc_fdk_std_backproject_T(const orc_geom *g, const double *vol, double *filt, int nthreads) {
  memset(filt, 0, sizeof(double) * (size_t)g->n_views * g->nu * g->nv);
  orc_std_bp_ctx c = {g, NULL, (double *)vol, 1, filt};
  orc_parallel_for(g->n_views, 1, nthreads, std_bpT_views, &c);
}

void orc_fdk_std(const orc_geom *g, const double *proj, double *vol, int nthreads) {
  size_t np = (size_t)g->n_views * g->nu * g->nv;
  double *pw = (double *)malloc(sizeof(double) * np);
  double *f = (double *)malloc(sizeof(double) * np);
  orc_fdk_std_weight(g, proj, pw);
  orc_fdk_std_filter(g, pw, f);
  memset(vol, 0, sizeof(double) * (size_t)g->nx * g->ny * g->nz);
  orc_fdk_std_backproject(g, f, vol, nthreads);
  free(pw);
  free(f);
}

void orc_fdk_std_adjoint(const orc_geom *g, const double *dvol, double *dproj, int nthreads) {
  size_t np = (size_t)g->n_views * g->nu * g->nv;
  double *f = (double *)malloc(sizeof(double) * np);
  double *pw = (double *)malloc(sizeof(double) * np);
  orc_fdk_std_backproject_T(g, dvol, f, nthreads);
  orc_fdk_std_filter_T(g, f, pw);
  orc_fdk_std_weight_T(g, pw, dproj);
  free(f);
  free(pw);
}

/* ============================================================== phantom */
static void ph_bounds(const orc_phantom *ph, double b0[3], double b1[3]) {
  b0[0] = ph->ox - 0.5 * ph->nx * ph->vx; b1[0] = b0[0] + ph->nx * ph->vx;
  b0[1] = ph->oy - 0.5 * ph->ny * ph->vy; b1[1] = b0[1] + ph->ny * ph->vy;
  b0[2] = ph->oz - 0.5 * ph->nz * ph->vz; b1[2] = b0[2] + ph->nz * ph->vz;
}

/* slab test: parametric [t0,t1] of p + t d inside the box; returns 0 if missed */
static int ray_box(const double p[3], const double d[3], const double b0[3], const double b1[3],
                   double *t0, double *t1) {
  double lo = -INFINITY, hi = INFINITY;
  for (int a = 0; a < 3; ++a) {
    if (d[a] != 0.0) {
      double ta = (b0[a] - p[a]) / d[a], tb = (b1[a] - p[a]) / d[a];
      if (ta > tb) { double tt = ta; ta = tb; tb = tt; }
      if (ta > lo) lo = ta;
      if (tb < hi) hi = tb;
    } else if (p[a] < b0[a] || p[a] >= b1[a]) {
      return 0;
    }
  }
  if (hi <= lo) return 0;
  *t0 = lo; *t1 = hi;
  return 1;
}

/* Siddon (1985) exact radiological path, written as an index-stepping
 * traversal: each crossing parameter is recomputed from its plane index. */
void orc_siddon_lengths(const orc_phantom *ph, const double p0[3], const double p1[3], double *len) {
  double b0[3], b1[3], d[3], vox[3] = {ph->vx, ph->vy, ph->vz};
  int n[3] = {ph->nx, ph->ny, ph->nz};
  ph_bounds(ph, b0, b1);
  for (int a = 0; a < 3; ++a) d[a] = p1[a] - p0[a];
  const double L = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
  double t0, t1;
  if (!ray_box(p0, d, b0, b1, &t0, &t1)) return;
  if (t0 < 0) t0 = 0;
  if (t1 > 1) t1 = 1;
  if (t1 <= t0) return;
  int idx[3], step[3];
  double tnext[3];
  const double tm = 0.5 * (t0 + fmin(t1, t0 + 1e-9)); /* a point just inside */
  for (int a = 0; a < 3; ++a) {
    double pa = p0[a] + tm * d[a];
    int i = (int)floor((pa - b0[a]) / vox[a]);
    if (i < 0) i = 0;
    if (i >= n[a]) i = n[a] - 1;
    idx[a] = i;
    if (d[a] > 0) {
      step[a] = 1;
      tnext[a] = (b0[a] + (i + 1) * vox[a] - p0[a]) / d[a];
    } else if (d[a] < 0) {
      step[a] = -1;
      tnext[a] = (b0[a] + i * vox[a] - p0[a]) / d[a];
    } else {
      step[a] = 0;
      tnext[a] = INFINITY;
    }
  }
  double t = t0;
  for (;;) {
    int a = 0;
    if (tnext[1] < tnext[a]) a = 1;
    if (tnext[2] < tnext[a]) a = 2;
    double te = tnext[a] < t1 ? tnext[a] : t1;
    if (te > t) {
      /* labels past the table read as the last label, as in transport */
      int lab = ph->labels[((size_t)idx[2] * ph->ny + idx[1]) * ph->nx + idx[0]];
      if (lab > ORC_MAX_MAT - 1) lab = ORC_MAX_MAT - 1;
      len[lab] += (te - t) * L;
      t = te;
    }
    if (tnext[a] >= t1) break;
    idx[a] += step[a];
    if (idx[a] < 0 || idx[a] >= n[a]) break;
    tnext[a] = step[a] > 0 ? (b0[a] + (idx[a] + 1) * vox[a] - p0[a]) / d[a]
                           : (b0[a] + idx[a] * vox[a] - p0[a]) / d[a];
  }
}

/* Detector mapping shared by primary and MC: u_idx = un0 + us * y', v_idx = vn0 + vs * z' */
static void det_map(const orc_geom *g, double *un0, double *us, double *vn0, double *vs) {
  if (g->preset == ORC_PRESET_BP3D20) {
    /* CBCT_real325.cu:565-566: index = -(int)(10*d - 162.5) -> u_idx = nu/2 - y'/du */
    *un0 = 0.5 * g->nu; *us = -1.0 / g->du;
    *vn0 = 0.5 * g->nv; *vs = -1.0 / g->dv;
  } else {
    *un0 = 0.5 * g->nu - g->u_off / g->du; *us = 1.0 / g->du;
    *vn0 = 0.5 * g->nv - g->v_off / g->dv; *vs = 1.0 / g->dv;
  }
}

static double view_angle(const orc_geom *g, int k) {
  if (g->preset == ORC_PRESET_BP3D20) return M_PI * (k * g->step_deg) / 180.0;
  return g->angle0 + k * g->dangle;
}

typedef struct {
  const orc_geom *g;
  const orc_phantom *ph;
  const orc_tables *t;
  double e_kev;
  const double *spectrum;
  double *out;
} orc_prim_ctx;

static void prim_rows(void *vctx, long r0, long r1) {
  orc_prim_ctx *c = (orc_prim_ctx *)vctx;
  const orc_geom *g = c->g;
  double un0, us, vn0, vs;
  det_map(g, &un0, &us, &vn0, &vs);
  for (long r = r0; r < r1; ++r) {
    int k = (int)(r / g->nu), iu = (int)(r % g->nu);
    double beta = view_angle(g, k), cb = cos(beta), sb = sin(beta);
    for (int iv = 0; iv < g->nv; ++iv) {
      double yp = (iu + 0.5 - un0) / us, zp = (iv + 0.5 - vn0) / vs, xp = g->sdd - g->sod;
      double p0[3] = {-g->sod * cb, -g->sod * sb, 0.0};
      double p1[3] = {xp * cb - yp * sb, xp * sb + yp * cb, zp};
      double len[256] = {0};
      orc_siddon_lengths(c->ph, p0, p1, len);
      double L;
      if (!c->spectrum) {
        int row = (int)(c->e_kev + 0.5);
        L = 0;
        for (int l = 0; l < c->ph->n_labels; ++l) {
          int tb = c->ph->mat_table[l];
          if (tb < 0 || len[l] == 0) continue;
          L += len[l] * c->ph->density[l] * c->t->tab[((size_t)tb * 4 + 3) * ORC_NROWS + row];
        }
      } else {
        double num = 0, den = 0;
        for (int e = 1; e < ORC_NROWS; ++e) {
          double w = c->spectrum[e];
          if (w <= 0) continue;
          double Le = 0;
          for (int l = 0; l < c->ph->n_labels; ++l) {
            int tb = c->ph->mat_table[l];
            if (tb < 0 || len[l] == 0) continue;
            Le += len[l] * c->ph->density[l] * c->t->tab[((size_t)tb * 4 + 3) * ORC_NROWS + e];
          }
          num += w * exp(-Le);
          den += w;
        }
        L = -log(num / den);
      }
      c->out[((size_t)k * g->nu + iu) * g->nv + iv] = L;
    }
  }
}

void orc_primary_lineint(const orc_geom *g, const orc_phantom *ph, const orc_tables *t, double e_kev,
                         const double *spectrum, double *out, int nthreads) {
  orc_prim_ctx c = {g, ph, t, e_kev, spectrum, out};
  orc_parallel_for((long)g->n_views * g->nu, 8, nthreads, prim_rows, &c);
}

/* =================================================================== MC */
/* Kahn's rejection sampling of Klein-Nishina as in
 * monte_cu/CBCT_real325.cu:687-722 (lambda = 511/E, rho = lambda'/lambda). */
typedef double (*orc_uniform_fn)(void *);

static double kahn_sample(double e_kev, orc_uniform_fn next, void *st) {
  const double lambda = 511.0 / e_kev;
  for (;;) {
    double r1 = next(st);
    if (r1 < (lambda + 2.0) / (9.0 * lambda + 2.0)) {
      double r2 = next(st);
      double ro = 1.0 + (2.0 / lambda) * r2;
      double r3 = next(st);
      if (r3 <= 4.0 * ((1.0 / ro) - (1.0 / (ro * ro)))) return ro;
    } else {
      double r2 = next(st);
      double ro = (lambda + 2.0) / (lambda + 2.0 * (1.0 - r2));
      double r3 = next(st);
      double q = lambda - ro * lambda + 1.0;
      if (r3 <= 0.5 * (q * q + (1.0 / ro))) return ro;
    }
  }
}

typedef struct { const double *u; int i; } orc_arr_stream;
static double arr_next(void *s) { orc_arr_stream *a = (orc_arr_stream *)s; return a->u[a->i++]; }
static double stream_next(void *s) { return orc_next((orc_stream *)s); }

double orc_kahn_lambda_ratio(double e_kev, const double *u, int *n_used) {
  orc_arr_stream a = {u, 0};
  double r = kahn_sample(e_kev, arr_next, &a);
  if (n_used) *n_used = a.i;
  return r;
}

/* Thomson angular distribution p(mu) = 3/8 (1 + mu^2), inverted in closed
 * form (Cardano on mu^3 + 3 mu + 4 - 8u = 0). */
double orc_thomson_mu(double u) {
  double a = 4.0 * u - 2.0;
  double A = fabs(a) + sqrt(a * a + 1.0);
  double m = cbrt(A) - cbrt(1.0 / A);
  return a < 0 ? -m : m;
}

/* Rotate unit direction d by polar cosine c and azimuth phi (direction-cosine
 * form).  The new direction is c d + s (cos(phi) e1 + sin(phi) e2) with
 * s = sqrt(1 - c^2) and the orthonormal pair e1 = (dx dz, dy dz, -r^2)/r,
 * e2 = (-dy, dx, 0)/r, r = sqrt(dx^2 + dy^2) (accurate near the poles, where
 * 1 - dz^2 would cancel); only for d on the z axis is the pair the x and y
 * axes.  The reference updates (theta, phi) angles instead
 * (monte_cu/CBCT_real325.cu:725-766, without carrying sin_phi_a, quirk Q6). */
static void rotate_dir(double d[3], double c, double phi) {
  double s = sqrt(fmax(0.0, 1.0 - c * c));
  double cp = cos(phi), sp = sin(phi);
  double r = sqrt(d[0] * d[0] + d[1] * d[1]);
  double n[3];
  if (r > 1e-150) {
    n[0] = c * d[0] + s * (d[0] * d[2] * cp - d[1] * sp) / r;
    n[1] = c * d[1] + s * (d[1] * d[2] * cp + d[0] * sp) / r;
    n[2] = c * d[2] - s * r * cp;
  } else {
    n[0] = s * cp;
    n[1] = s * sp;
    n[2] = d[2] >= 0 ? c : -c;
  }
  double nn = 1.0 / sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
  d[0] = n[0] * nn; d[1] = n[1] * nn; d[2] = n[2] * nn;
}

void orc_rotate_dir(double d[3], double c, double phi) { rotate_dir(d, c, phi); }

/* Spectrum draw: the first keV bin e in [1, 200] whose cumulative weight
 * reaches u (a linear scan of the CDF as CBCT_real325.cu:479-486 scans its
 * table; bin 200 if none does). */
int orc_spectrum_bin(const double *cdf, double u) {
  int e = 1;
  while (e < ORC_NROWS - 1 && cdf[e] < u) ++e;
  return e;
}

/* ------------------------------------------------ first-flight majorant
 * A photon's first flight runs along the line from the source to the centre
 * of its pixel, so the labels that line crosses bound the attenuation it can
 * meet: the Woodcock majorant of that flight may be the largest mu among them
 * instead of among all labels (Woodcock tracking stays exact for any majorant
 * at least mu everywhere along the flight).  Label sets are bit masks: bit l
 * for labels 0..3, bit 4 for any label 4..255.
 *
 * orc_brick_masks: the label set of every brick of b^3 voxels, taken over the
 * brick widened by one voxel on every side, so a point whose voxel is at most
 * one voxel away from a brick (a sample point rounded across a brick face) is
 * covered by that brick's set.
 * orc_ray_masks: for view k and pixel pix (aim order iv * nu + iu), the union
 * of the sets of the bricks the segment source -> pixel centre crosses inside
 * the box [c0, c1] (an index-stepping traversal of the brick grid), or 16
 * (every label) when the segment misses the box.  The arithmetic is plain
 * IEEE FP64 without fused operations, in a fixed order, so the product
 * kernel, written with the same operations, yields the same masks bit for bit. */
static int label_bit(int lab) { return lab < 4 ? 1 << lab : 16; }

void orc_brick_masks(const orc_phantom *ph, int b, uint8_t *out) {
  const int nbx = (ph->nx + b - 1) / b, nby = (ph->ny + b - 1) / b, nbz = (ph->nz + b - 1) / b;
  for (int bz = 0; bz < nbz; ++bz)
    for (int by = 0; by < nby; ++by)
      for (int bx = 0; bx < nbx; ++bx) {
        int m = 0;
        const int z0 = bz * b - 1 < 0 ? 0 : bz * b - 1, z1 = bz * b + b < ph->nz ? bz * b + b : ph->nz - 1;
        const int y0 = by * b - 1 < 0 ? 0 : by * b - 1, y1 = by * b + b < ph->ny ? by * b + b : ph->ny - 1;
        const int x0 = bx * b - 1 < 0 ? 0 : bx * b - 1, x1 = bx * b + b < ph->nx ? bx * b + b : ph->nx - 1;
        for (int z = z0; z <= z1; ++z)
          for (int y = y0; y <= y1; ++y)
            for (int x = x0; x <= x1; ++x) m |= label_bit(ph->labels[((size_t)z * ph->ny + y) * ph->nx + x]);
        out[((size_t)bz * nby + by) * nbx + bx] = (uint8_t)m;
      }
}

/* label set of the segment S -> P (see above); org: grid corner, bw: brick size */
static int ray_set(const double S[3], const double P[3], const double c0[3], const double c1[3],
                   const double org[3], const double bw[3], const int nb[3], const uint8_t *bricks) {
  double d[3], inv[3];
  double lo = 0.0, hi = 1.0;
  for (int a = 0; a < 3; ++a) {
    d[a] = P[a] - S[a];
    if (d[a] != 0.0) {
      inv[a] = 1.0 / d[a];
      const double ta = (c0[a] - S[a]) * inv[a], tb = (c1[a] - S[a]) * inv[a];
      const double tl = ta < tb ? ta : tb, th = ta < tb ? tb : ta;
      if (tl > lo) lo = tl;
      if (th < hi) hi = th;
    } else {
      inv[a] = 0.0;
      if (S[a] < c0[a] || S[a] >= c1[a]) return 16;
    }
  }
  if (!(hi > lo)) return 16;
  const double span = hi - lo;
  const double tm = lo + (span * 0.5 < 1e-9 ? span * 0.5 : 1e-9); /* a point just inside */
  int idx[3];
  for (int a = 0; a < 3; ++a) {
    const double q = S[a] + tm * d[a];
    int i = (int)floor((q - org[a]) / bw[a]);
    idx[a] = i < 0 ? 0 : (i > nb[a] - 1 ? nb[a] - 1 : i);
  }
  int m = 0;
  for (;;) {
    m |= bricks[((size_t)idx[2] * nb[1] + idx[1]) * nb[0] + idx[0]];
    double tn[3];
    for (int a = 0; a < 3; ++a) {
      if (d[a] > 0.0) tn[a] = (org[a] + (double)(idx[a] + 1) * bw[a] - S[a]) * inv[a];
      else if (d[a] < 0.0) tn[a] = (org[a] + (double)idx[a] * bw[a] - S[a]) * inv[a];
      else tn[a] = INFINITY;
    }
    int a = 0;
    if (tn[1] < tn[a]) a = 1;
    if (tn[2] < tn[a]) a = 2;
    if (!(tn[a] < hi)) break;
    idx[a] += d[a] > 0.0 ? 1 : -1;
    if (idx[a] < 0 || idx[a] >= nb[a]) break;
  }
  return m;
}

void orc_ray_masks(const orc_geom *g, const orc_phantom *ph, const double c0[3], const double c1[3],
                   const uint8_t *bricks, int b, int view_begin, int n_views, uint8_t *out) {
  double un0, u_step, vn0, v_step;
  if (g->preset == ORC_PRESET_BP3D20) {
    un0 = 0.5 * g->nu; u_step = -g->du; vn0 = 0.5 * g->nv; v_step = -g->dv;
  } else {
    un0 = 0.5 * g->nu - g->u_off / g->du; u_step = g->du;
    vn0 = 0.5 * g->nv - g->v_off / g->dv; v_step = g->dv;
  }
  double org[3], b1[3];
  ph_bounds(ph, org, b1);
  const double bw[3] = {b * ph->vx, b * ph->vy, b * ph->vz};
  const int nb[3] = {(ph->nx + b - 1) / b, (ph->ny + b - 1) / b, (ph->nz + b - 1) / b};
  const size_t npix = (size_t)g->nu * g->nv;
  const double xd = g->sdd - g->sod;
  for (int kk = 0; kk < n_views; ++kk) {
    const double beta = view_angle(g, view_begin + kk), cb = cos(beta), sb = sin(beta);
    const double S[3] = {-g->sod * cb, -g->sod * sb, 0.0};
    for (size_t pix = 0; pix < npix; ++pix) {
      const int iv = (int)(pix / g->nu), iu = (int)(pix % g->nu);
      const double yp = ((double)iu + 0.5 - un0) * u_step, zp = ((double)iv + 0.5 - vn0) * v_step;
      const double P[3] = {xd * cb - yp * sb, xd * sb + yp * cb, zp};
      out[(size_t)kk * npix + pix] = (uint8_t)ray_set(S, P, c0, c1, org, bw, nb, bricks);
    }
  }
}

typedef struct {
  float mu[ORC_MAX_MAT][ORC_NROWS];  /* linear attenuation per label (cm^-1), FP32-rounded */
  float fa[ORC_MAX_MAT][ORC_NROWS];  /* photo / total */
  float fac[ORC_MAX_MAT][ORC_NROWS]; /* (photo + coherent) / total */
  float mumax[ORC_NROWS];
} orc_mc_tables;

/* Device tables are stored in FP32; the oracle rounds the same values so the
 * two sample identical distributions. */
static void build_mc_tables(const orc_phantom *ph, const orc_tables *t, orc_mc_tables *m) {
  memset(m, 0, sizeof *m);
  for (int e = 0; e < ORC_NROWS; ++e) {
    float mx = 0.f;
    for (int l = 0; l < ph->n_labels && l < ORC_MAX_MAT; ++l) {
      int tb = ph->mat_table[l];
      if (tb < 0) continue;
      const double *T = t->tab + (size_t)tb * 4 * ORC_NROWS;
      double tot = T[3 * ORC_NROWS + e];
      m->mu[l][e] = (float)(ph->density[l] * tot);
      m->fa[l][e] = (float)(T[2 * ORC_NROWS + e] / tot);
      m->fac[l][e] = (float)((T[2 * ORC_NROWS + e] + T[0 * ORC_NROWS + e]) / tot);
      if (m->mu[l][e] > mx) mx = m->mu[l][e];
    }
    m->mumax[e] = mx;
  }
}

static int ebin_of(double e) {
  int b = (int)(e + 0.5); /* CBCT_real325.cu:489: row (int)(Energy + 0.5) */
  if (b < 1) b = 1;
  if (b > ORC_NROWS - 1) b = ORC_NROWS - 1;
  return b;
}

typedef struct {
  const orc_geom *g;
  const orc_phantom *ph;
  const orc_mc_cfg *cfg;
  const orc_mc_tables *mt;
  double *primary, *scatter;
  orc_mc_stats *stats;
  pthread_mutex_t mu;
  long chunks_per_view;
  double c0[3], c1[3]; /* flight box: bounding box of the attenuating voxels */
  const uint8_t *ray_sets; /* [views][npix] first-flight label sets, or NULL (global majorant) */
} orc_mc_ctx;

/* Bounding box of the voxels whose label attenuates at some energy (labels
 * clamped to the table as in transport).  Outside it mu is 0, so clipping
 * flights to it changes no history's outcome, only the Woodcock steps spent
 * crossing the empty margin; the GPU transport clips to the same box.  With
 * no attenuating voxel the grid box is kept. */
static void support_box(const orc_phantom *ph, const orc_mc_tables *mt, double c0[3], double c1[3]) {
  double b0[3], b1[3];
  ph_bounds(ph, b0, b1);
  int solid[ORC_MAX_MAT];
  for (int l = 0; l < ORC_MAX_MAT; ++l) {
    solid[l] = 0;
    for (int e = 0; e < ORC_NROWS; ++e)
      if (mt->mu[l][e] > 0.f) { solid[l] = 1; break; }
  }
  int lo[3] = {ph->nx, ph->ny, ph->nz}, hi[3] = {-1, -1, -1};
  for (int z = 0; z < ph->nz; ++z)
    for (int y = 0; y < ph->ny; ++y) {
      const uint8_t *row = ph->labels + ((size_t)z * ph->ny + y) * ph->nx;
      for (int x = 0; x < ph->nx; ++x) {
        int lab = row[x] < ORC_MAX_MAT ? row[x] : ORC_MAX_MAT - 1;
        if (!solid[lab]) continue;
        const int idx[3] = {x, y, z};
        for (int a = 0; a < 3; ++a) {
          if (idx[a] < lo[a]) lo[a] = idx[a];
          if (idx[a] > hi[a]) hi[a] = idx[a];
        }
      }
    }
  const double vox[3] = {ph->vx, ph->vy, ph->vz};
  for (int a = 0; a < 3; ++a) {
    if (hi[0] < 0) { c0[a] = b0[a]; c1[a] = b1[a]; continue; }
    c0[a] = b0[a] + lo[a] * vox[a];
    c1[a] = b0[a] + (hi[a] + 1) * vox[a];
  }
}

static void mc_items(void *vctx, long i0, long i1) {
  orc_mc_ctx *c = (orc_mc_ctx *)vctx;
  const orc_geom *g = c->g;
  const orc_phantom *ph = c->ph;
  const orc_mc_cfg *cfg = c->cfg;
  const orc_mc_tables *mt = c->mt;
  const size_t npix = (size_t)g->nu * g->nv;
  double *bp = (double *)calloc(npix, sizeof(double));
  double *bs = (double *)calloc(npix, sizeof(double));
  double un0, us, vn0, vs, b0[3], b1[3];
  det_map(g, &un0, &us, &vn0, &vs);
  ph_bounds(ph, b0, b1);
  const double xdet = g->sdd - g->sod;
  for (long item = i0; item < i1; ++item) {
    orc_mc_stats st = {0};
    const int k = cfg->view_begin + (int)(item / c->chunks_per_view);
    const long ch = item % c->chunks_per_view;
    const uint64_t N = cfg->photons_per_view;
    const uint64_t p_begin = N * ch / c->chunks_per_view, p_end = N * (ch + 1) / c->chunks_per_view;
    const double beta = view_angle(g, k), cb = cos(beta), sb = sin(beta);
    memset(bp, 0, npix * sizeof(double));
    memset(bs, 0, npix * sizeof(double));
    for (uint64_t n = p_begin; n < p_end; ++n) {
      orc_stream rs;
      orc_stream_init(&rs, (uint32_t)n, (uint32_t)k, cfg->job, cfg->seed_lo, cfg->seed_hi);
      st.histories++;
      /* aim at the centre of pixel n mod npix (CBCT_real325.cu:464-466), in
       * the order iv * nu + iu (u fastest) */
      const size_t pix = (size_t)(n % npix);
      const int iv = (int)(pix / g->nu), iu = (int)(pix % g->nu);
      const double yp = (iu + 0.5 - un0) / us, zp = (iv + 0.5 - vn0) / vs;
      double dl[3] = {g->sdd, yp, zp};
      double dn = 1.0 / sqrt(dl[0] * dl[0] + dl[1] * dl[1] + dl[2] * dl[2]);
      double dir[3] = {(dl[0] * cb - dl[1] * sb) * dn, (dl[0] * sb + dl[1] * cb) * dn, dl[2] * dn};
      double pos[3] = {-g->sod * cb, -g->sod * sb, 0.0};
      double E = cfg->e_kev;
      if (cfg->spectrum_cdf) E = orc_spectrum_bin(cfg->spectrum_cdf, orc_next(&rs));
      int eb = ebin_of(E);
      int nint = 0;
      int first = 1;
      const int ray_set = c->ray_sets ? c->ray_sets[(size_t)(k - cfg->view_begin) * npix + pix] : 16;
      for (;;) {
        /* flight: Woodcock inside the box of attenuating voxels, straight
         * line outside */
        double t0, t1;
        int real = 0;
        /* majorant: the largest mu among the labels of the first flight's line,
         * else among all labels */
        const int set = first ? ray_set : 16;
        float mm = 0.f;
        if (set & 16) {
          mm = mt->mumax[eb];
        } else {
          for (int l = 0; l < 4; ++l)
            if ((set >> l) & 1 && mt->mu[l][eb] > mm) mm = mt->mu[l][eb];
        }
        first = 0;
        if (ray_box(pos, dir, c->c0, c->c1, &t0, &t1) && t1 > 0 && mm > 0.f) {
          /* advance to the box entry so flight parameters stay small */
          if (t0 > 0) {
            pos[0] += t0 * dir[0]; pos[1] += t0 * dir[1]; pos[2] += t0 * dir[2];
            t1 -= t0;
          }
          double t = 0;
          const double mumax = mm;
          /* two Woodcock steps per block: (distance, accept, distance, accept) */
          for (int ended = 0; !ended;) {
            double u[4];
            orc_block(&rs, u);
            for (int h = 0; h < 2 && !ended; ++h) {
              t += -log(u[2 * h]) / mumax;
              st.steps++;
              if (t >= t1) { ended = 1; break; }
              double P[3] = {pos[0] + t * dir[0], pos[1] + t * dir[1], pos[2] + t * dir[2]};
              int ix = (int)floor((P[0] - b0[0]) / ph->vx), iy = (int)floor((P[1] - b0[1]) / ph->vy),
                  iz = (int)floor((P[2] - b0[2]) / ph->vz);
              ix = ix < 0 ? 0 : (ix >= ph->nx ? ph->nx - 1 : ix);
              iy = iy < 0 ? 0 : (iy >= ph->ny ? ph->ny - 1 : iy);
              iz = iz < 0 ? 0 : (iz >= ph->nz ? ph->nz - 1 : iz);
              int lab = ph->labels[((size_t)iz * ph->ny + iy) * ph->nx + ix];
              if (lab > ORC_MAX_MAT - 1) lab = ORC_MAX_MAT - 1; /* past the table: the last label */
              double mu = mt->mu[lab][eb];
              if (u[2 * h + 1] * mumax < mu) {
                ended = 1;
                real = 1;
                pos[0] = P[0]; pos[1] = P[1]; pos[2] = P[2];
                /* interaction at this label */
                if (nint >= cfg->max_interactions) {
                  st.capped++;
                  goto done;
                }
                double u3 = orc_next(&rs); /* fresh block: orc_block left none */
                st.interactions++;
                if (u3 <= mt->fa[lab][eb]) {
                  st.photo++;
                  goto done;
                } else if (u3 <= mt->fac[lab][eb]) {
                  st.coherent++;
                  if (cfg->coherent_mode == ORC_COH_RAYLEIGH) {
                    double cth = orc_thomson_mu(orc_next(&rs));
                    double phi = 2.0 * M_PI * orc_next(&rs);
                    rotate_dir(dir, cth, phi);
                  }
                } else {
                  st.compton++;
                  double ro = kahn_sample(E, stream_next, &rs);
                  double lam = 511.0 / E, lamd = ro * lam;
                  double cth = 1.0 - (lamd - lam);
                  if (cth < -1) cth = -1;
                  E = 511.0 / lamd;
                  eb = ebin_of(E);
                  double phi = 2.0 * M_PI * orc_next(&rs);
                  rotate_dir(dir, cth, phi);
                }
                nint++;
              }
            }
          }
        }
        if (real) continue;
        /* escaped the phantom: straight line to the detector plane x' = sdd - sod */
        {
          double xq = pos[0] * cb + pos[1] * sb, yq = -pos[0] * sb + pos[1] * cb, zq = pos[2];
          double dx = dir[0] * cb + dir[1] * sb, dy = -dir[0] * sb + dir[1] * cb, dz = dir[2];
          if (dx <= 0) { st.missed++; goto done; }
          double tt = (xdet - xq) / dx;
          double yh = yq + tt * dy, zh = zq + tt * dz;
          double ui = un0 + us * yh, vi = vn0 + vs * zh;
          if (!(ui >= 0 && ui < g->nu && vi >= 0 && vi < g->nv)) { st.missed++; goto done; }
          size_t pi = (size_t)(int)floor(ui) * g->nv + (size_t)(int)floor(vi);
          if (nint == 0) { bp[pi] += 1.0; st.tallied_primary++; }
          else { bs[pi] += 1.0; st.tallied_scatter++; }
          goto done;
        }
      }
    done:;
    }
    pthread_mutex_lock(&c->mu);
    double *P = c->primary + (size_t)k * npix, *S = c->scatter + (size_t)k * npix;
    for (size_t i = 0; i < npix; ++i) { P[i] += bp[i]; S[i] += bs[i]; }
    orc_mc_stats *T = c->stats;
    T->histories += st.histories; T->steps += st.steps; T->interactions += st.interactions;
    T->photo += st.photo; T->coherent += st.coherent; T->compton += st.compton;
    T->tallied_primary += st.tallied_primary; T->tallied_scatter += st.tallied_scatter;
    T->missed += st.missed; T->capped += st.capped;
    pthread_mutex_unlock(&c->mu);
  }
  free(bp);
  free(bs);
}

void orc_mc_run(const orc_geom *g, const orc_phantom *ph, const orc_tables *t, const orc_mc_cfg *cfg,
                double *primary, double *scatter, orc_mc_stats *stats, int nthreads) {
  orc_mc_tables *mt = (orc_mc_tables *)malloc(sizeof(orc_mc_tables));
  build_mc_tables(ph, t, mt);
  orc_mc_stats local = {0};
  orc_mc_ctx c;
  c.g = g; c.ph = ph; c.cfg = cfg; c.mt = mt;
  support_box(ph, mt, c.c0, c.c1);
  uint8_t *bricks = NULL, *sets = NULL;
  c.ray_sets = NULL;
  if (cfg->ray_brick > 0 && cfg->view_end > cfg->view_begin) {
    const int b = cfg->ray_brick;
    const size_t nbr = (size_t)((ph->nx + b - 1) / b) * ((ph->ny + b - 1) / b) * ((ph->nz + b - 1) / b);
    bricks = (uint8_t *)malloc(nbr);
    sets = (uint8_t *)malloc((size_t)(cfg->view_end - cfg->view_begin) * g->nu * g->nv);
    orc_brick_masks(ph, b, bricks);
    orc_ray_masks(g, ph, c.c0, c.c1, bricks, b, cfg->view_begin, cfg->view_end - cfg->view_begin, sets);
    c.ray_sets = sets;
  }
  c.primary = primary; c.scatter = scatter; c.stats = stats ? stats : &local;
  pthread_mutex_init(&c.mu, NULL);
  int nv = cfg->view_end - cfg->view_begin;
  if (nv <= 0) { free(mt); return; }
  long cpv = (4L * (nthreads < 1 ? 1 : nthreads) + nv - 1) / nv;
  if ((uint64_t)cpv > cfg->photons_per_view) cpv = (long)(cfg->photons_per_view ? cfg->photons_per_view : 1);
  if (cpv < 1) cpv = 1;
  c.chunks_per_view = cpv;
  orc_parallel_for((long)nv * cpv, 1, nthreads, mc_items, &c);
  pthread_mutex_destroy(&c.mu);
  free(bricks);
  free(sets);
  free(mt);
}

void orc_flight_box(const orc_phantom *ph, const orc_tables *t, double c0[3], double c1[3]) {
  orc_mc_tables *mt = (orc_mc_tables *)malloc(sizeof(orc_mc_tables));
  build_mc_tables(ph, t, mt);
  support_box(ph, mt, c0, c1);
  free(mt);
}
