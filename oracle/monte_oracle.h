/*
 * monte_oracle.h -- FP64 CPU restatement of Tomato27/Monte's two hot paths.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (monte_amd/, include/,
 * libmonte.so) links, loads or calls this library.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg use it, and only as
 * the correctness reference or as the timed CPU baseline -- never as the
 * thing measured.
 *
 * What each part restates (reference paths are relative to the reference
 * repository root):
 *   orc_philox4x32_10      Philox4x32-10 (Salmon et al., SC'11).  No
 *                          reference counterpart: the reference uses cuRAND
 *                          MRG32k3a (monte_cu/CBCT_real325.cu:977-990).
 *   orc_read_xcom_csv      readcsv() semantics, monte_cpp/CBCT_real2.cpp:633-668
 *                          (row k -> k keV; BOM/CR stripped, quirk Q1 fixed).
 *   orc_bp3d20_*           recon/bp3d20_325.cpp:39-175 and recon/bp3d20.cpp:36-166
 *                          restated literally (FP64, including the tan(beta)
 *                          with beta in integer degrees used as radians).
 *   orc_fdk_std_*          textbook FDK (Feldkamp, Davis & Kress 1984; Kak &
 *                          Slaney ch. 3) in the same generic geometry as the
 *                          product's standard preset, plus its exact adjoint.
 *   orc_primary_lineint    exact Beer-Lambert line integrals by Siddon (1985)
 *                          traversal of a voxel label map.
 *   orc_mc_run             Woodcock photon transport following the physics of
 *                          monte_cu/CBCT_real325.cu:419-975 (aim at pixel
 *                          centres, Woodcock flights, photo/coherent/Compton with
 *                          Kahn sampling, detector-plane tally), in FP64, with
 *                          the same Philox stream layout and draw order as the
 *                          product kernel so individual histories can be compared.
 */
#ifndef MONTE_ORACLE_H
#define MONTE_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_NROWS 201 /* table rows: index = keV, 0..200 */

/* ------------------------------------------------------------------ RNG */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
/* 24-bit uniform in (0,1): ((x >> 8) + 0.5) * 2^-24, exact in FP32 and FP64 */
double orc_u01(uint32_t x);

/* --------------------------------------------------------------- tables */
/* Parse an XCOM-style CSV (coherent, incoherent, photoelectric, total; cm^2/g;
 * row k = k keV).  tab[c][k] for c in 0..3, k in 1..200; row 0 = row 1.
 * Returns number of rows parsed (200 for the reference files) or -1. */
int orc_read_xcom_csv(const char *path, double tab[4][ORC_NROWS]);

/* ------------------------------------------------------------- geometry */
enum { ORC_PRESET_STANDARD = 0, ORC_PRESET_BP3D20 = 1 };
enum { ORC_FILTER_RAMLAK = 0, ORC_FILTER_SHEPP_LOGAN = 1 };

typedef struct orc_geom {
  int preset;
  int n_views;
  int nu, nv;         /* detector: nu in-plane (u), nv axial (v); layout [view][u][v] */
  int nx, ny, nz;     /* volume; layout [z][y][x] */
  double sod, sdd;    /* source-isocentre, source-detector distance (cm) */
  double du, dv;      /* detector pitch (cm) */
  double u_off, v_off;/* detector centre offsets (standard preset) */
  double vx, vy, vz;  /* voxel size (cm) */
  double ox, oy, oz;  /* volume centre offset (cm) */
  double angle0, dangle; /* view k at angle0 + k*dangle (radians) */
  /* bp3d20 preset only */
  double odd;         /* distance used by bp3d20's weights (60) */
  double scale;       /* bp3d20 output scale (2.7*5 at 325^2, 2.7 at 65^2) */
  int step_deg;       /* bp3d20 view step in integer degrees */
  int roi_x0, roi_x1; /* bp3d20 reconstructs x columns [roi_x0, roi_x1) only */
  double mask_r;      /* bp3d20 65-variant sphere mask radius in voxels (<0: none) */
  int filter;         /* ORC_FILTER_* (standard preset) */
} orc_geom;

/* ----------------------------------------------------------- bp3d20 FDK */
/* Cosine weight, recon/bp3d20_325.cpp:39-46.  map, map_w: [view][nu][nv]. */
void orc_bp3d20_weight(const orc_geom *g, const float *map, double *map_w);
/* Ramp filter + transposed write, recon/bp3d20_325.cpp:51-81.
 * map_w [view][nu][nv] -> map_out [view][nv][nu]. */
void orc_bp3d20_filter(const orc_geom *g, const double *map_w, double *map_out);
/* Backprojection, recon/bp3d20_325.cpp:91-175.  map_out [view][nv][nu] ->
 * vol_xy [nz][ny][nx] (accumulated). */
void orc_bp3d20_backproject(const orc_geom *g, const double *map_out, double *vol_xy, int nthreads);

/* --------------------------------------------------------- standard FDK */
void orc_fdk_std_weight(const orc_geom *g, const double *proj, double *proj_w);
void orc_fdk_std_filter(const orc_geom *g, const double *proj_w, double *filt); /* -> [view][nv][nu] */
void orc_fdk_std_backproject(const orc_geom *g, const double *filt, double *vol, int nthreads);
/* exact adjoints (transposes) of the three stages */
void orc_fdk_std_backproject_T(const orc_geom *g, const double *vol, double *filt, int nthreads);
void orc_fdk_std_filter_T(const orc_geom *g, const double *filt, double *proj_w);
void orc_fdk_std_weight_T(const orc_geom *g, const double *proj_w, double *proj);
/* whole operator and its adjoint (FP64 in/out) */
void orc_fdk_std(const orc_geom *g, const double *proj, double *vol, int nthreads);
void orc_fdk_std_adjoint(const orc_geom *g, const double *dvol, double *dproj, int nthreads);

/* ------------------------------------------------------------- phantom */
#define ORC_MAX_MAT 16
typedef struct orc_phantom {
  const uint8_t *labels; /* [nz][ny][nx] */
  int nx, ny, nz;
  double vx, vy, vz;
  double ox, oy, oz;     /* centre offset */
  int n_labels;          /* labels 0..n_labels-1; others vacuum, labels > 15 read as 15 */
  int mat_table[ORC_MAX_MAT];   /* table index per label, -1 = vacuum */
  double density[ORC_MAX_MAT];  /* g/cm^3 per label */
} orc_phantom;

typedef struct orc_tables {
  int n_tables;
  const double *tab; /* [n_tables][4][ORC_NROWS]: coh, incoh, photo, total (cm^2/g) */
} orc_tables;

/* Exact line integrals (Siddon) from the source to each pixel centre.
 * If spectrum == NULL: mono energy e_kev, out = sum mu*l.
 * Else spectrum[k] (k = 0..200, weight of the k keV bin):
 *   out = -ln( sum_k w_k exp(-L_k) / sum_k w_k ).
 * out: [view][nu][nv]. */
void orc_primary_lineint(const orc_geom *g, const orc_phantom *ph, const orc_tables *t,
                         double e_kev, const double *spectrum, double *out, int nthreads);
/* Per-material path lengths of one ray p0->p1 (world cm); len[label] += l */
void orc_siddon_lengths(const orc_phantom *ph, const double p0[3], const double p1[3], double *len);

/* ----------------------------------------------------------------- MC */
enum { ORC_COH_FORWARD = 0, ORC_COH_RAYLEIGH = 1 };
typedef struct orc_mc_cfg {
  uint64_t photons_per_view;
  double e_kev;            /* mono energy if spectrum_cdf == NULL */
  const double *spectrum_cdf; /* cumulative weights over keV bins 0..200 (normalised), or NULL */
  int max_interactions;    /* reference: 5 (ScatterNUM) */
  int coherent_mode;       /* ORC_COH_* */
  uint32_t seed_lo, seed_hi;
  uint32_t job;
  int view_begin, view_end; /* subset of views to simulate */
  int ray_brick;            /* > 0: first flights take the majorant of the labels their line crosses,
                               from bricks of ray_brick^3 voxels (orc_ray_masks); 0: global majorant */
} orc_mc_cfg;

typedef struct orc_mc_stats {
  uint64_t histories, steps, interactions, photo, coherent, compton;
  uint64_t tallied_primary, tallied_scatter, missed, capped;
} orc_mc_stats;

/* Label sets (bit l: label l in 0..3, bit 4: any label >= 4) of the first
 * flight's line: bricks of b^3 voxels widened by one voxel
 * (out [ceil(nz/b)][ceil(ny/b)][ceil(nx/b)]), and per view of [view_begin,
 * view_begin + n_views) and pixel in aim order the union over the bricks the
 * source -> pixel-centre segment crosses inside the box [c0, c1]
 * (out [n_views][nv * nu]). */
void orc_brick_masks(const orc_phantom *ph, int b, uint8_t *out);
void orc_ray_masks(const orc_geom *g, const orc_phantom *ph, const double c0[3], const double c1[3],
                   const uint8_t *bricks, int b, int view_begin, int n_views, uint8_t *out);
/* the flight box of the transport: bounding box of the attenuating voxels */
void orc_flight_box(const orc_phantom *ph, const orc_tables *t, double c0[3], double c1[3]);

/* primary, scatter: [view][nu][nv] FP64 tallies (accumulated) */
void orc_mc_run(const orc_geom *g, const orc_phantom *ph, const orc_tables *t,
                const orc_mc_cfg *cfg, double *primary, double *scatter,
                orc_mc_stats *stats, int nthreads);

/* Single-history sampling primitives, exposed so tests can pin them against
 * closed forms (Klein-Nishina, Thomson, branching ratios). */
double orc_kahn_lambda_ratio(double e_kev, const double *u, int *n_used); /* returns lambda'/lambda */
double orc_thomson_mu(double u);  /* cos(theta) ~ (1+mu^2) */
/* d <- d turned by polar cosine c and azimuth phi about itself (unit vectors) */
void orc_rotate_dir(double d[3], double c, double phi);
/* keV bin of a spectrum draw u: first e in [1, 200] with cdf[e] >= u, else 200 */
int orc_spectrum_bin(const double *cdf, double u);

#ifdef __cplusplus
}
#endif
#endif
