/*
 * monte.h -- C ABI of libmonte.so, the MI355X (gfx950) engine for cone-beam CT
 * Monte Carlo projection and FDK reconstruction.
 *
 * This is the drop-in boundary for the two hot paths of Tomato27/Monte.  The
 * reference has no library API: each stage is a standalone program joined to
 * the next by raw files.  Every entry point below replaces one of those
 * programs (or one kernel inside one), with the same inputs and outputs, and
 * cites it.  Reference paths are relative to the reference repository root.
 *
 * Conventions
 *   - Return value: MONTE_OK (0) or a negative monte_status; never exit().
 *     monte_last_error() holds a message for the last failing call on a ctx.
 *   - Data pointers are DEVICE pointers (hipMalloc or torch tensors) unless the
 *     context was switched to host pointers with monte_set_pointer_mode(); in
 *     host mode every call stages through device buffers it allocates itself.
 *   - Calls are ordered on the context's stream (monte_set_stream; a new
 *     context starts on its own non-blocking stream).  Use one context per
 *     host thread.
 *   - Units: cm, keV, cm^-1.  All layouts are C order, slowest index first:
 *       projections / tallies / line integrals  [view][u][v]   (v fastest)
 *       filtered projections                    [view][v][u]   (u fastest)
 *       volumes                                 [z][y][x]      (x fastest)
 *       label phantoms                          [z][y][x]      uint8
 *     matching the reference's raw files (monte_cu/CBCT_real325.cu:296-300,
 *     recon/bp3d20_325.cpp:29-30, :68-81, :170).
 */
#ifndef MONTE_H
#define MONTE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MONTE_VERSION_MAJOR 0
#define MONTE_VERSION_MINOR 1
#define MONTE_TABLE_ROWS 201 /* attenuation table rows: index = keV, 0..200 */
#define MONTE_MAX_LABELS 16  /* labels 0..15 in a phantom */

typedef struct monte_ctx monte_ctx;

typedef enum monte_status {
  MONTE_OK = 0,
  MONTE_ERR_INVALID = -1,     /* bad argument: null pointer, size, shape */
  MONTE_ERR_HIP = -2,         /* HIP runtime failure */
  MONTE_ERR_NOMEM = -3,       /* device allocation failed */
  MONTE_ERR_STATE = -4,       /* prerequisite missing (e.g. tables not loaded) */
  MONTE_ERR_IO = -5,          /* file open / parse failure */
  MONTE_ERR_UNSUPPORTED = -6  /* valid request the library does not implement */
} monte_status;

typedef enum monte_preset {
  /* Textbook FDK / CBCT geometry: pixel centres, SDD cosine weight, (SOD/U)^2
   * distance weight, tau-correct filter scale (mu in cm^-1). */
  MONTE_PRESET_STANDARD = 0,
  /* Reproduces recon/bp3d20_325.cpp and recon/bp3d20.cpp exactly: D=60 cosine
   * weight at pixel edges, 0.5 filter scale, 60^2/(60-d)^2 distance weight with
   * tan(beta) of integer degrees taken as radians, x-column ROI, sphere mask,
   * and the MC detector index -(int)(10 d - 162.5) of CBCT_real325.cu:565. */
  MONTE_PRESET_BP3D20 = 1
} monte_preset;

typedef enum monte_filter { MONTE_FILTER_RAMLAK = 0, MONTE_FILTER_SHEPP_LOGAN = 1 } monte_filter;
typedef enum monte_coherent {
  MONTE_COH_FORWARD = 0, /* undeflected, as CBCT_real325.cu:642-681 */
  MONTE_COH_RAYLEIGH = 1 /* Thomson (1+cos^2) angular distribution */
} monte_coherent;

/* Scan and volume geometry.  The source sits at (-sod,0,0) and the detector
 * plane at x' = sdd - sod in a frame rotated by the view angle about z.
 * Detector pixel iu covers u_idx in [iu, iu+1) with
 *   standard: u_idx = nu/2 + (y' - u_off)/du      bp3d20: u_idx = nu/2 - y'/du
 * and likewise v with z'.  Voxel (x,y,z) centres:
 *   standard: X = ox + (x - (nx-1)/2) vx
 *   bp3d20:   X = ox - nx/2 vx + x vx,  Y = oy + ny/2 vy - y vy,  Z = oz + nz/2 vz - z vz
 *   (recon/bp3d20_325.cpp:107-109). */
typedef struct monte_geom {
  int32_t preset;           /* monte_preset */
  int32_t n_views;
  int32_t nu, nv;           /* detector size: u in-plane, v axial */
  int32_t nx, ny, nz;       /* volume size */
  double sod, sdd;          /* source-isocentre, source-detector distances */
  double du, dv;            /* detector pitch */
  double u_off, v_off;      /* detector offsets (standard) */
  double vx, vy, vz;        /* voxel size */
  double ox, oy, oz;        /* volume centre offset */
  double angle0, dangle;    /* view k at angle0 + k*dangle radians (standard) */
  double odd;               /* bp3d20: weight distance (60) */
  double scale;             /* bp3d20: output scale (2.7*5 or 2.7) */
  int32_t step_deg;         /* bp3d20: view k at k*step_deg integer degrees */
  int32_t roi_x0, roi_x1;   /* bp3d20: reconstruct x in [roi_x0, roi_x1) (others 0) */
  double mask_r;            /* bp3d20: sphere mask radius in voxels, < 0 = none */
  int32_t filter;           /* monte_filter (standard) */
} monte_geom;

/* Voxel label phantom ("phantom in").  Label l has attenuation table
 * mat_table[l] (an index into the loaded tables, -1 = vacuum) at density[l].
 * Labels at or above n_labels are vacuum; labels above 15 read as label 15
 * (transport, exact primary and flight box alike). */
typedef struct monte_phantom {
  const uint8_t *labels;    /* [nz][ny][nx] */
  int32_t nx, ny, nz;
  int32_t n_labels;
  double vx, vy, vz;        /* voxel size */
  double ox, oy, oz;        /* centre offset */
  int32_t mat_table[MONTE_MAX_LABELS];
  double density[MONTE_MAX_LABELS];
} monte_phantom;

typedef struct monte_mc_cfg {
  uint64_t photons_per_view; /* photon n aims at the centre of pixel n mod (nu nv) taken in the order
                                iv * nu + iu (CBCT_real325.cu:452-466 aims at pixel centres) */
  double e_kev;              /* mono-energetic source when spectrum == NULL */
  const double *spectrum;    /* HOST pointer: photon weight per keV bin [201], or NULL */
  int32_t max_interactions;  /* reference ScatterNUM = 5 (CBCT_real325.cu:9) */
  int32_t coherent_mode;     /* monte_coherent */
  uint64_t seed;             /* Philox key */
  uint32_t job;              /* stream id: counter word 3 */
  int32_t view_begin, view_end; /* simulate views [begin,end); 0,0 = all */
} monte_mc_cfg;

typedef struct monte_mc_stats {
  uint64_t histories, steps, interactions, photo, coherent, compton;
  uint64_t tallied_primary, tallied_scatter, missed, capped;
  double kernel_ms;          /* device time of the transport launch(es) */
} monte_mc_stats;

/* One job of a batched launch (monte_mc_batch). */
typedef struct monte_mc_job {
  const monte_geom *geom;
  const monte_phantom *phantom;
  const monte_mc_cfg *cfg;
  double *primary;           /* [views][nu][nv], accumulated */
  double *scatter;           /* [views][nu][nv], accumulated */
} monte_mc_job;

/* Launch parameters picked by monte_autotune (or set by hand). */
typedef struct monte_tuning {
  int32_t mc_block;          /* threads per block of the transport kernel */
  int32_t mc_batch;          /* photon cycles per work unit of one 8x8 detector tile (about 64 * mc_batch
                                histories per unit; views with fewer cycles group tiles).  monte_mc_batch
                                and monte_mc_project halve it, down to 16, for launches of fewer than 32
                                units per wave of the grid */
  int32_t mc_blocks_per_cu;  /* persistent grid = CUs * this */
  int32_t bp_zchunk;         /* voxels along z per backprojection thread */
  int32_t bp_tile_x, bp_tile_y; /* backprojection block tile (threads) */
  int32_t bp_view_chunk;     /* views per backprojection launch */
  int32_t pj_zchunk;         /* projector voxels per lane along z: 2, 4, 8, or 0 = the default (2) */
  int32_t mc_ray_brick;      /* > 0: a photon's first flight takes as Woodcock majorant the largest mu
                                among the labels its source -> pixel line crosses (found per view and
                                pixel on bricks of this many voxels a side); 0: the global majorant.
                                Either is exact; they draw different random numbers. */
} monte_tuning;

/* ------------------------------------------------------------ context */
int monte_create(monte_ctx **ctx, int device);
void monte_destroy(monte_ctx *ctx);
/* stream: a hipStream_t (e.g. torch.cuda.current_stream().cuda_stream); NULL
 * selects the device's default (null) stream.  A new context starts on a
 * non-blocking stream of its own. */
int monte_set_stream(monte_ctx *ctx, void *stream);
int monte_set_pointer_mode(monte_ctx *ctx, int host_pointers);
int monte_synchronize(monte_ctx *ctx);
const char *monte_status_str(int status);
const char *monte_last_error(const monte_ctx *ctx);
int monte_get_tuning(const monte_ctx *ctx, monte_tuning *out);
int monte_set_tuning(monte_ctx *ctx, const monte_tuning *in);

/* ------------------------------------------------------------ tables */
/* Replaces readcsv (monte_cpp/CBCT_real2.cpp:633-668) and readXcom
 * (monte_cu/CBCT_real325.cu:305-347): parse a 4-column XCOM CSV (coherent,
 * incoherent, photoelectric, total in cm^2/g; row k = k keV; BOM/CR tolerated)
 * into tab[4][201].  Host only; returns the row count or a monte_status. */
int monte_read_xcom_csv(const char *path, double *tab);
/* Upload n_tables tables, tab = [n_tables][4][201] host array. */
int monte_load_tables(monte_ctx *ctx, const double *tab, int n_tables);

/* ------------------------------------------------ FDK: projections -> volume */
/* Whole FDK of recon/bp3d20_325.cpp:main (:39-175): cosine weight, ramp filter,
 * voxel-driven backprojection.  proj [batch][views][nu][nv] -> vol [batch][nz][ny][nx]. */
int monte_fdk(monte_ctx *ctx, const monte_geom *g, const float *proj, float *vol, int batch);
/* Stage 1+2 (recon/bp3d20_325.cpp:39-81): weighted, filtered, transposed:
 * proj [batch][views][nu][nv] -> filt [batch][views][nv][nu]. */
int monte_fdk_filter(monte_ctx *ctx, const monte_geom *g, const float *proj, float *filt, int batch);
/* Stage 3 (recon/bp3d20_325.cpp:91-175): filt [batch][views][nv][nu] -> vol
 * (vol is overwritten unless accumulate != 0). */
int monte_fdk_backproject(monte_ctx *ctx, const monte_geom *g, const float *filt, float *vol, int batch,
                          int accumulate);

/* ------------------------------------- FDK adjoint: volume -> projections */
/* No reference counterpart (the reference has no adjoint).  Exact transposes:
 * backproject^T (a matched voxel-driven cone-beam projector), filter^T and
 * the diagonal weight, so <fdk(p), y> == <p, fdk_adjoint(y)>. */
int monte_fdk_project(monte_ctx *ctx, const monte_geom *g, const float *vol, float *filt, int batch);
int monte_fdk_filter_adjoint(monte_ctx *ctx, const monte_geom *g, const float *filt, float *proj, int batch);
int monte_fdk_adjoint(monte_ctx *ctx, const monte_geom *g, const float *dvol, float *dproj, int batch);

/* ------------------------------------------- MC: phantom -> projections */
/* Replaces the projection kernel of monte_cu/CBCT_real325.cu:419-975 (and its
 * voxel form CBCT_real325im.cu:425-962): Woodcock transport of
 * photons_per_view histories per view, FP64 tallies of primary and scatter
 * counts per pixel (accumulated into primary/scatter [views][nu][nv]). */
int monte_mc_project(monte_ctx *ctx, const monte_geom *g, const monte_phantom *ph, const monte_mc_cfg *cfg,
                     double *primary, double *scatter, monte_mc_stats *stats);
/* monte_mc_project without device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate during a call (graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Device
 * pointers only.  The launch's descriptors, attenuation tables, spectrum,
 * per-view table, work queue and counters live in a static 128 KiB device
 * array per GPU, which bounds the geometry to about 9400 views
 * (MONTE_ERR_INVALID beyond); every call on a GPU reuses that GPU's array, so
 * calls that overlap on one GPU race.  Every flight takes the global Woodcock
 * majorant (the first-flight label sets of tuning mc_ray_brick need memory in
 * proportion to views x pixels), so results equal monte_mc_project's run with
 * mc_ray_brick = 0. */
int monte_mc_project_noalloc(monte_ctx *ctx, const monte_geom *g, const monte_phantom *ph, const monte_mc_cfg *cfg,
                             double *primary, double *scatter, monte_mc_stats *stats);
/* The label sets that bound the first flights' Woodcock majorants (tuning
 * mc_ray_brick), for inspection and tests; no reference counterpart.  sets:
 * device uint8 [n_views][nv * nu], pixel iv * nu + iu (the aim order); bit l
 * for label l in 0..3, bit 4 (16) for any label >= 4 or a line that misses the
 * flight box.  A line's set is the union of the sets of the bricks of brick^3
 * voxels (each widened by one voxel) it crosses inside the flight box, found
 * in FP64 by the same operations as the CPU oracle. */
int monte_mc_first_flight_sets(monte_ctx *ctx, const monte_geom *g, const monte_phantom *ph, int brick,
                               uint8_t *sets);
/* Many independent jobs in one persistent launch fed by a device work queue
 * (no reference counterpart).  stats: n_jobs entries or NULL. */
int monte_mc_batch(monte_ctx *ctx, const monte_mc_job *jobs, int n_jobs, monte_mc_stats *stats);
/* Deterministic primary (scatter-off) estimator: exact line integrals of
 * mu along source->pixel-centre rays (Siddon).  Mono (spectrum == NULL) or
 * polyenergetic: -ln(sum_E w_E exp(-L_E) / sum_E w_E).  out [views][nu][nv]. */
int monte_primary_exact(monte_ctx *ctx, const monte_geom *g, const monte_phantom *ph, double e_kev,
                        const double *spectrum, float *lineint);
/* Counts -> line integrals, monte_cu/CBCT_real325.cu:264-294:
 * ln(N0) - ln(clamp(N, 1, N0)) with N0 the photons aimed at the pixel (see photons_per_view). */
int monte_counts_to_lineint(monte_ctx *ctx, const monte_geom *g, const double *counts,
                            uint64_t photons_per_view, float *lineint);
/* The same against a measured flat field (air scan): ln(F) - ln(clamp(N, 1, F)),
 * 0 where F <= 0.  flat is f64 [flat_views][nu][nv] with flat_views 1 (one air
 * scan for every view) or n_views.  The reference has no flat field: it divides
 * by the nominal photons per pixel (monte_cu/CBCT_real325.cu:264-294). */
int monte_counts_to_lineint_flat(monte_ctx *ctx, const monte_geom *g, const double *counts,
                                 const double *flat, int flat_views, float *lineint);

/* ------------------------------------------------ phantom generation */
/* Voxelise primitives into a label map (monte_cpp/make_image01.cpp:8-33 is the
 * reference's single-sphere generator).  Each primitive paints its label over
 * voxel centres inside it, in order. */
typedef enum monte_shape {
  MONTE_SHAPE_ELLIPSOID = 0,
  MONTE_SHAPE_CYLINDER_Z = 1,
  MONTE_SHAPE_BOX = 2,
  MONTE_SHAPE_CYLINDER_X = 3  /* axis along x, as the reference's water cylinder and Ca rods
                                 (monte_cu/CBCT_real325.cu:916-928) */
} monte_shape;
typedef struct monte_primitive {
  int32_t shape;            /* monte_shape */
  int32_t label;
  double cx, cy, cz;        /* centre (cm) */
  double rx, ry, rz;        /* semi-axes (cylinder along z: rx, ry radii, rz half-height;
                               along x: ry, rz radii, rx half-length) */
} monte_primitive;
int monte_phantom_paint(uint8_t *labels, int nx, int ny, int nz, double vx, double vy, double vz,
                        const monte_primitive *prims, int n_prims); /* host labels */

/* ------------------------------------------------------- tuning / graphs */
/* Time candidate launch shapes on this device for the given workload and keep
 * the fastest in the context (no reference counterpart).  what: bit 1 = MC
 * (block, batch, blocks per CU), bit 2 = backprojection (z chunk, tile), bit 4 =
 * matched projector (voxels per lane).  Writes the chosen values to *out if
 * non-NULL. */
int monte_autotune(monte_ctx *ctx, int what, const monte_geom *g, const monte_phantom *ph,
                   const monte_mc_cfg *cfg, monte_tuning *out);
/* Capture one full FDK (filter + backprojection over all view chunks) into a
 * hipGraph owned by the context, and replay it (no reference counterpart).
 * The graph is rebuilt when g, the pointers or batch change. */
int monte_fdk_graph(monte_ctx *ctx, const monte_geom *g, const float *proj, float *vol, int batch);

/* ------------------------------------------------------------ end to end */
/* Device buffers of one end-to-end step (caller-owned). */
typedef struct monte_e2e_buffers {
  double *primary, *scatter; /* [views][nu][nv] MC tallies, overwritten */
  float *lineint;            /* [views][nu][nv] line integrals of primary + scatter */
  float *vol;                /* [nz][ny][nx] FDK of lineint */
  const float *vol_ref;      /* [nz][ny][nx] target volume */
  float *residual;           /* [nz][ny][nx] vol - vol_ref */
  float *grad;               /* [views][nu][nv] gradient of the loss with respect to lineint */
  double *loss;              /* one double: 0.5 ||vol - vol_ref||^2 */
} monte_e2e_buffers;
/* One step of the differentiable scan, chaining what the reference runs as
 * separate programs joined by files (monte_cu/CBCT_real325.cu:241-300 then
 * recon/bp3d20_325.cpp:39-175) and adding what it lacks, the loss and its
 * gradient:
 *   MC projection of every view (tallies zeroed first), lineint = ln N0 -
 *   ln clamp(primary + scatter, 1, N0), vol = fdk(lineint), residual = vol -
 *   vol_ref, loss = 1/2 ||residual||^2 (FP64, fixed summation order),
 *   grad = fdk^T(residual).
 * use_graph != 0: the whole step is captured once into a hipGraph owned by the
 * context and replayed (rebuilt when any argument, the tables or the tuning
 * change); results are bitwise equal to use_graph == 0.  Device pointers only;
 * cfg must cover every view.  stats (nullable): the MC counters, with
 * kernel_ms the device time of the whole step. */
int monte_e2e_step(monte_ctx *ctx, const monte_geom *g, const monte_phantom *ph, const monte_mc_cfg *cfg,
                   const monte_e2e_buffers *buf, int use_graph, monte_mc_stats *stats);

#ifdef __cplusplus
}
#endif
#endif /* MONTE_H */
