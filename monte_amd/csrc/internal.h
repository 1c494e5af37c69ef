// Internal declarations shared by the host C++ and the gfx950 kernels of
// libmonte.so.  Not part of the public ABI (include/monte.h is).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/monte.h"

namespace monte {

// IEEE FP64 add, subtract and multiply that the compiler never fuses into an
// FMA, for expressions that must round exactly as a CPU evaluates them (the
// pragma marks the operations themselves, so it holds after inlining; the
// __dadd_rn family alone does not stop that fusion).
__host__ __device__ inline double xadd(double a, double b) {
#pragma clang fp contract(off)
  return a + b;
}
__host__ __device__ inline double xsub(double a, double b) {
#pragma clang fp contract(off)
  return a - b;
}
__host__ __device__ inline double xmul(double a, double b) {
#pragma clang fp contract(off)
  return a * b;
}

constexpr int kRows = MONTE_TABLE_ROWS;
constexpr int kMaxLabels = MONTE_MAX_LABELS;
// MC detector tile: one birth of a wave (64 lanes) covers one tile of one cycle.
constexpr int kTileU = 8, kTileV = 8;
constexpr int kPadLo = 1;  // zero guard cells before u / v in padded filtered projections
constexpr int kPadHi = 3;  // and after (covers the +1 bilinear tap at the window edge)

// ------------------------------------------------------------------ FDK
// Per-view backprojection coefficients (computed on the host in FP64).  For a
// voxel centre (X,Y,Z):
//   U  = cx X + cy Y + c0                      distance along the central ray
//   su = su0 + suk (ax X + ay Y) / U           sample coordinate along u
//   sv = sv0 + svk Z / U                       sample coordinate along v
//   Dd = d0 + dx X + dy Y - sgn(gx X + gy Y) |hx X + hy Y|
//   w  = fac (wn / Dd)^2                       distance weight x angular step
// Standard preset: Dd = U, wn = sod, fac = dangle.  bp3d20 preset: Dd = 60 - d
// with d the signed point-line distance of recon/bp3d20_325.cpp:145-153.
// The sample coordinates are formed per voxel column in FP64 (U, 1/U, su and
// the first sv of the column's z-chunk), then split into an integer pixel
// index and an FP32 fraction, so the per-voxel FP32 work only moves a small
// fraction along v.  The bp3d20 sign test sgn(gx X + gy Y) is evaluated in
// unfused FP64 exactly as the reference evaluates it: at columns where it is
// zero in exact arithmetic (e.g. X = -Y at 135 degrees) the reference's
// rounding decides the sign of d, and with it the weight.
struct BpView {
  double cx, cy, c0;
  double ax, ay;
  double su0, suk;
  double sv0, svk;
  double gx, gy;
  float wn, d0, dx, dy;
  float hx, hy, fac, pad0;
};
static_assert(sizeof(BpView) == 120, "BpView layout");

// Geometry-wide constants of the backprojector.
struct BpGrid {
  int nx, ny, nz, nu, nv;
  int pitch;          // padded row pitch of filtered projections (nu + kPadLo + kPadHi)
  int rows;           // padded rows (nv + kPadLo + kPadHi)
  double gx0, gsx;    // X = gx0 + x * gsx (unfused, as recon/bp3d20_325.cpp:107)
  double gy0, gsy;
  double gz0, gsz;
  double lo_u, hi_u;  // valid sample window along u (inclusive)
  double lo_v, hi_v;
  int x0, x1;         // x range reconstructed (bp3d20 ROI)
  int mask;           // 1: sphere mask in voxel index space
  float mask_r2;
  int mcx, mcy, mcz;  // mask centre (voxel index)
};

// Filter constants: weight(r,c) = D / sqrt(D^2 + (a_u + b_u r)^2 + (a_v + b_v c)^2),
// out = cs * sum_r weight * p * h[r - b].
struct FilterParams {
  int nu, nv, n_views;
  int pitch, rows;    // padded output layout
  float D, a_u, b_u, a_v, b_v;
  float cs;
};

void compute_bp_views(const monte_geom& g, std::vector<BpView>& out);
BpGrid compute_bp_grid(const monte_geom& g);
FilterParams compute_filter_params(const monte_geom& g);
void compute_filter_taps(const monte_geom& g, std::vector<float>& h);  // h[0..2nu-2], centre nu-1

// ------------------------------------------------------------------- MC
// Device attenuation tables of one job, per label and keV row:
//   m[l][e] = (mu, photo/total, (photo+coherent)/total, 0), mu in cm^-1
//   mumax[e] = max_l mu  (the Woodcock majorant)
struct McTables {
  float4 m[kMaxLabels][kRows];
  float mumax[kRows];
  float pad[3];
};

// Device job descriptor of the persistent transport kernel.
struct McJob {
  const uint8_t* labels;
  const McTables* tables;
  const float* spectrum_cdf;     // [kRows] or null (mono)
  const float2* view_cs;         // cos, sin per view
  double* primary;
  double* scatter;
  int nx, ny, nz, nu, nv;
  int view_begin, n_views_sim;   // simulated views (indices into the tally arrays)
  float vx, vy, vz;
  float bx0, by0, bz0, bx1, by1, bz1;  // phantom box
  float inv_vx, inv_vy, inv_vz;
  float ix0, iy0, iz0;           // -b0 / v: voxel index = q inv_v + i0
  float sod, sdd;
  float un0, us, vn0, vs;        // u_idx = un0 + us * y', v_idx = vn0 + vs * z'
  float u_step, v_step;          // 1/us, 1/vs (signed pitch)
  float e_kev;
  int max_int, coherent;
  uint32_t key0, key1, job_id;
  uint32_t photons_per_view;
  // Photon n of a view aims at pixel j = n mod (nu nv) in aim order (iv nu + iu),
  // in cycle c = n / (nu nv).  A work unit is a group of `tile_group` detector
  // tiles of kTileU x kTileV pixels (tiles numbered u fastest, `tiles_u` per row,
  // `tiles` in all) over `cycle_block` consecutive cycles, and a birth takes one
  // tile of one cycle: so a unit's primary tallies fall on a few lines that it
  // hits once per cycle, and the units in flight share a narrow detector band.
  uint32_t tiles_u, tiles, tile_group;
  uint32_t cycles, cycle_block, cycle_blocks;  // ceil(photons / (nu nv)); per unit; blocks
  uint32_t batches_per_view;     // ceil(tiles / tile_group) * cycle_blocks
  uint64_t unit_begin;           // first global work unit of this job
  uint64_t unit_count;
  const uint8_t* ray_set;        // [n_views_sim][nu nv] first-flight label sets (aim order), or null
};

// FP64 geometry of the first-flight label sets (ray_sets_kernel).  A photon's
// first flight runs from the source towards the centre of its pixel, so the
// labels that segment crosses bound the attenuation it meets, and the largest
// mu among them is a valid Woodcock majorant for that flight.  Sets are bit
// masks: bit l for labels 0..3, bit 4 for any label >= 4 (then the flight
// takes the global majorant).
struct RayGrid {
  double un0, u_step, vn0, v_step;  // pixel centre: y' = (iu + 1/2 - un0) u_step, z' likewise
  double sod, xd;                   // source distance, detector plane x' = xd
  double c0[3], c1[3];              // flight box (bounding box of the attenuating voxels)
  double org[3], bw[3];             // brick grid corner and brick size
  int nb[3];                        // bricks per axis
  int nu, nv;
};

// FP64 ray geometry of the exact primary estimator.
struct PrimaryGeom {
  double sod, sdd, un0, u_step, vn0, v_step;
  double b0[3], b1[3], vox[3];
  int n[3];
  int nu, nv, n_views;
};

struct McCounters {  // device-side statistics, one per job
  unsigned long long histories, steps, interactions, photo, coherent, compton;
  unsigned long long tallied_primary, tallied_scatter, missed, capped;
};

// Where a launch's job descriptors, tables, CDFs, view tables, queue, counters
// and flight boxes sit in its device buffer.
struct McLayout {
  size_t jobs_off = 0, tab_off = 0, q_off = 0, cnt_off = 0, box_off = 0, total = 0;
  std::vector<size_t> cdf_off, cs_off;
};

// A transport launch whose descriptors are in place in device memory;
// launching it is stream-ordered work only (capturable into a graph).
struct McPrepared {
  char* base = nullptr;
  McLayout L;
  int n = 0;
  uint64_t units = 0;
  uint32_t chunk = 1;  // units per queue chunk (the first job's simulated views; more with several jobs)
  unsigned int* queue() const { return (unsigned int*)(base + L.q_off); }
  McCounters* counters() const { return (McCounters*)(base + L.cnt_off); }
};

}  // namespace monte

// ------------------------------------------------------------------ context
struct monte_ctx {
  int device = 0;
  hipStream_t own_stream = nullptr;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;  // times transport launches (created once, with the context)
  bool host_ptrs = false;
  std::string err;
  int n_cu = 256;

  // attenuation tables (host copy, FP64): [n_tables][4][201]
  std::vector<double> tables;
  int n_tables = 0;

  monte_tuning tuning{};

  // growable device workspaces:
  //   0 forward filtered stack (padded)   1 adjoint filtered stack (padded)
  //   2 MC descriptors                    3 FDK constants (taps + per-view coefficients)
  //   4..7 staging for host-pointer mode and tuning scratch
  //   8 z-fastest copy of the volume for the matched projector
  //   9 tap quads of the forward filtered stack (backprojection input)
  //  10 column lists of the matched projector (per geometry)
  //  11 end-to-end step: MC descriptors and loss partial sums
  //  12 first-flight label sets of MC launches   13 those of the end-to-end step
  static constexpr int kSlots = 14;
  void* ws[kSlots] = {};
  size_t ws_size[kSlots] = {};

  // FDK constants cached by geometry
  std::vector<double> const_key;
  size_t taps_off = 0, views_off = 0;

  // projector column lists cached by geometry: offsets, then the lists
  std::vector<double> plan_key;
  size_t plan_list_off = 0;

  // hipGraph cache for monte_fdk_graph
  hipGraphExec_t graph_exec = nullptr;
  hipGraph_t graph = nullptr;
  std::vector<double> graph_key;

  // bumped by every monte_load_tables (keys of cached MC descriptors)
  uint64_t tables_gen = 0;

  // end-to-end step: prepared MC launch (keyed by job), and its graph
  monte::McPrepared e2e_mc;
  std::vector<double> e2e_mc_key;
  hipGraphExec_t e2e_exec = nullptr;
  hipGraph_t e2e_graph = nullptr;
  std::vector<double> e2e_graph_key;
};

namespace monte {

// error helpers
int fail(monte_ctx* c, int code, const std::string& msg);
#define MONTE_HIP(ctx, expr)                                                              \
  do {                                                                                    \
    hipError_t e_ = (expr);                                                               \
    if (e_ != hipSuccess)                                                                 \
      return ::monte::fail((ctx), MONTE_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

// workspace slot i grown to at least bytes; returns nullptr on failure
void* workspace(monte_ctx* c, int slot, size_t bytes);

// Device scratch of the allocation-free entry points (a static device array,
// one per GPU; noalloc.hip): its address on `device` and its size, or nullptr
// before noalloc_setup has run on that device.
constexpr size_t kNoAllocScratchBytes = 128 * 1024;
int noalloc_setup(int device);  // with `device` current; 0 or a hipError_t
void* noalloc_scratch(int device);

// ---------------------------------------------------------- kernel launches
// FDK (fdk.hip)
hipError_t launch_fill(float* p, float v, size_t n, hipStream_t s);
hipError_t launch_fdk_filter(const float* proj, float* filt_pad, const float* taps, const FilterParams& fp,
                             hipStream_t s);
hipError_t launch_pad_copy(const float* filt, float* filt_pad, int n_views, int nu, int nv, int pitch, int rows,
                           hipStream_t s);
hipError_t launch_unpad_copy(const float* filt_pad, float* filt, int n_views, int nu, int nv, int pitch, int rows,
                             hipStream_t s);
hipError_t launch_fill_pads(float* buf, int n_views, int nu, int nv, int pitch, int rows, int wrap, hipStream_t s);
hipError_t launch_fold_pads(float* buf, int n_views, int nu, int nv, int pitch, int rows, hipStream_t s);
// Tap quads of the padded filtered views: quads[i] holds the 2 x 2 bilinear
// neighbourhood whose top-left cell is padded cell i (n_views * rows * pitch),
// as (p00, p01 - p00, p10, p11 - p10).
hipError_t launch_tap_quads(const float* filt_pad, float4* quads, int n_views, int rows, int pitch, hipStream_t s);
hipError_t launch_backproject(const float4* quads, const BpView* views, int view0, int nviews_chunk,
                              const BpGrid& grid, float* vol, int accumulate, const monte_tuning& t, hipStream_t s);
// volT: the volume with z fastest ([y][x][z], see launch_transpose_zx); every
// cell of the padded views [view0, view0 + nviews_chunk) is written.  Padded
// rows [lds_r0, lds_r0 + lds_rows) must hold every row the views' voxels can
// reach (projector_rows); the others are written as zeros.
// Detector strips per view of the matched projector.
int project_strips(const BpGrid& grid);
// Column lists of the projector for views [0, n_views): with write false,
// offs[k * project_strips + st] receives each list's length; with write true,
// offs holds the exclusive prefix sums (plus the total at the end) and the
// lists are written to list.
hipError_t launch_project_plan(const BpView* views, int n_views, const BpGrid& grid, uint32_t* offs, uint32_t* list,
                               bool write, hipStream_t s);
hipError_t launch_project(const float* volT, const BpView* views, int view0, int nviews_chunk, const BpGrid& grid,
                          float* filt_pad, int lds_r0, int lds_rows, const uint32_t* plan_offs,
                          const uint32_t* plan_list, const monte_tuning& t, hipStream_t s);
// Padded detector rows reachable by the voxels of views [view0, view0 + n).
void projector_rows(const std::vector<BpView>& views, int view0, int n, const BpGrid& grid, int& lds_r0,
                    int& lds_rows);
hipError_t launch_transpose_zx(const float* vol, float* volT, int nx, int ny, int nz, hipStream_t s);
hipError_t launch_fdk_filter_T(const float* filt_pad, float* proj, const float* taps, const FilterParams& fp,
                               hipStream_t s);

// MC (mc.hip)
// The work queue is kMcQueues counters kMcQueueStride words apart, each
// handing out one chunk (`chunk` consecutive work units) per round, and block
// b draws from queue b mod kMcQueues first.  Blocks are dealt to the 8 XCDs
// round robin, so an XCD draws from 3 of the queues and its L2 serves their
// chunks.  With units of detector tiles, 8, 16, 24, 32 and 64 queues took
// 84.5, 84.2, 83.4, 88.0 and 105.8 ms at 360 views x 1e7 photons, and 24 took
// 815.8 against 824.7 ms for 32 at 1e8 photons (fewer queues keep an XCD's
// units on fewer tiles; with units of consecutive photons 8, 16 and 32 had
// taken 164, 161 and 158 ms, the counters' contention then mattering more).
constexpr int kMcQueues = 24;
constexpr int kMcQueueStride = 16;
constexpr size_t kMcQueueBytes = sizeof(unsigned int) * kMcQueues * kMcQueueStride;
hipError_t launch_mc_transport(const McJob* jobs_dev, int n_jobs, uint64_t total_units, uint32_t chunk,
                               unsigned int* queue, McCounters* counters, int block, int grid, hipStream_t s);
hipError_t launch_mc_reset(unsigned int* queue, McCounters* counters, int n_jobs, hipStream_t s);
// First-flight label sets: per brick of b^3 voxels widened by one voxel
// (bricks [nb2][nb1][nb0]), then per view of [view_begin, view_begin +
// n_views) and pixel in aim order (sets [n_views][nu nv]); cs64 holds FP64
// (cos, sin) of every view of the scan.
hipError_t launch_brick_sets(const uint8_t* labels, int nx, int ny, int nz, int b, const int nb[3], uint8_t* bricks,
                             hipStream_t s);
hipError_t launch_ray_sets(const RayGrid& R, const double* cs64, int view_begin, int n_views, const uint8_t* bricks,
                           uint8_t* sets, hipStream_t s);
// Bounding box (voxel indices, inclusive) of voxels whose clamped label has bit set in `solid`:
// box[0..2] = minima, box[3..5] = maxima; the caller initialises it to {INT_MAX x 3, -1 x 3}.
hipError_t launch_support_box(const uint8_t* labels, int nx, int ny, int nz, uint32_t solid, int* box, hipStream_t s);
hipError_t launch_primary_exact(const PrimaryGeom& P, const double* cs64, const uint8_t* labels,
                                const McTables* tables, float e_kev, const float* spec_w, int use_spec, float* out,
                                hipStream_t s);
// counts2 (nullable) is added to counts before the conversion
hipError_t launch_counts_to_lineint(const double* counts, const double* counts2, float* out, int n_views, int nu,
                                    int nv, uint64_t photons_per_view, hipStream_t s);
// out = a - b; partial[blk] = FP64 sum of out^2 over block blk's fixed share
// (grid-stride, reduced in a fixed order), then loss = 0.5 * sum(partial) in
// one more fixed-order pass: deterministic, so graph replays are bitwise equal.
constexpr int kLossBlocks = 1024;
hipError_t launch_residual_loss(const float* a, const float* b, float* out, size_t n, double* partial, double* loss,
                                hipStream_t s);
hipError_t launch_counts_to_lineint_flat(const double* counts, const double* flat, int flat_views, float* out,
                                         int n_views, int nu, int nv, hipStream_t s);

}  // namespace monte
