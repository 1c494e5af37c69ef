// C ABI of libmonte.so (declared in include/monte.h): argument checking,
// workspaces, constant caching, host-pointer staging, MC job construction,
// autotuning and hipGraph capture.  All device work is in fdk.hip and mc.hip.
#include <climits>
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "internal.h"

using namespace monte;

namespace monte {

int fail(monte_ctx* c, int code, const std::string& msg) {
  if (c) c->err = msg;
  return code;
}

void* workspace(monte_ctx* c, int slot, size_t bytes) {
  if (bytes == 0) bytes = 256;
  if (c->ws_size[slot] >= bytes) return c->ws[slot];
  if (c->ws[slot]) {
    (void)hipStreamSynchronize(c->stream);
    (void)hipFree(c->ws[slot]);
    c->ws[slot] = nullptr;
    c->ws_size[slot] = 0;
  }
  void* p = nullptr;
  if (hipMalloc(&p, bytes) != hipSuccess) return nullptr;
  c->ws[slot] = p;
  c->ws_size[slot] = bytes;
  return p;
}

}  // namespace monte

namespace {

constexpr int kSlotFwd = 0, kSlotAdj = 1, kSlotMc = 2, kSlotConst = 3, kSlotStageA = 4, kSlotStageB = 5,
              kSlotTuneA = 6, kSlotTuneB = 7, kSlotVolT = 8, kSlotQuad = 9,
              kSlotPlan = 10, kSlotE2E = 11, kSlotRay = 12, kSlotE2ERay = 13;

size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

int check_geom(monte_ctx* c, const monte_geom* g, bool need_volume) {
  if (!g) return fail(c, MONTE_ERR_INVALID, "geometry is null");
  if (g->preset != MONTE_PRESET_STANDARD && g->preset != MONTE_PRESET_BP3D20)
    return fail(c, MONTE_ERR_INVALID, "unknown preset");
  if (g->n_views < 1 || g->nu < 1 || g->nv < 1) return fail(c, MONTE_ERR_INVALID, "detector/view counts must be >= 1");
  if (g->nu > 1024) return fail(c, MONTE_ERR_UNSUPPORTED, "detector rows longer than 1024 pixels are not supported");
  if (g->nv > 1536) return fail(c, MONTE_ERR_UNSUPPORTED, "detectors taller than 1536 pixels are not supported");
  if (!(g->sod > 0) || !(g->sdd > g->sod)) return fail(c, MONTE_ERR_INVALID, "need 0 < sod < sdd");
  if (!(g->du > 0) || !(g->dv > 0)) return fail(c, MONTE_ERR_INVALID, "detector pitch must be > 0");
  if (g->preset == MONTE_PRESET_BP3D20 && g->step_deg < 1) return fail(c, MONTE_ERR_INVALID, "step_deg must be >= 1");
  if (g->preset == MONTE_PRESET_BP3D20 && g->nu != g->nv)
    return fail(c, MONTE_ERR_INVALID, "the bp3d20 preset indexes a square detector (nu == nv)");
  if (g->preset == MONTE_PRESET_STANDARD && g->filter != MONTE_FILTER_RAMLAK && g->filter != MONTE_FILTER_SHEPP_LOGAN)
    return fail(c, MONTE_ERR_INVALID, "unknown filter");
  if (need_volume) {
    if (g->nx < 1 || g->ny < 1 || g->nz < 1) return fail(c, MONTE_ERR_INVALID, "volume dims must be >= 1");
    if (!(g->vx > 0) || !(g->vy > 0) || !(g->vz > 0)) return fail(c, MONTE_ERR_INVALID, "voxel size must be > 0");
    if (g->preset == MONTE_PRESET_BP3D20 && (g->roi_x0 < 0 || g->roi_x1 > g->nx || g->roi_x0 >= g->roi_x1))
      return fail(c, MONTE_ERR_INVALID, "roi_x0/roi_x1 outside the volume");
    const double rx = 0.5 * g->nx * g->vx + std::fabs(g->ox), ry = 0.5 * g->ny * g->vy + std::fabs(g->oy);
    if (std::sqrt(rx * rx + ry * ry) >= g->sod) return fail(c, MONTE_ERR_INVALID, "volume reaches the source orbit");
  }
  return MONTE_OK;
}

std::vector<double> geom_key(const monte_geom& g) {
  return {(double)g.preset, (double)g.n_views, (double)g.nu, (double)g.nv, (double)g.nx, (double)g.ny,
          (double)g.nz, g.sod, g.sdd, g.du, g.dv, g.u_off, g.v_off, g.vx, g.vy, g.vz, g.ox, g.oy, g.oz,
          g.angle0, g.dangle, g.odd, g.scale, (double)g.step_deg, (double)g.roi_x0, (double)g.roi_x1, g.mask_r,
          (double)g.filter};
}

// Upload filter taps and per-view coefficients once per geometry.
int ensure_constants(monte_ctx* c, const monte_geom& g, const float** taps, const BpView** views) {
  std::vector<double> key = geom_key(g);
  if (key != c->const_key || !c->ws[kSlotConst]) {
    std::vector<float> h;
    compute_filter_taps(g, h);
    std::vector<BpView> v;
    compute_bp_views(g, v);
    const size_t toff = 0, voff = align_up(h.size() * sizeof(float), 256);
    const size_t bytes = voff + v.size() * sizeof(BpView);
    char* base = (char*)workspace(c, kSlotConst, bytes);
    if (!base) return fail(c, MONTE_ERR_NOMEM, "constants allocation failed");
    MONTE_HIP(c, hipMemcpyAsync(base + toff, h.data(), h.size() * sizeof(float), hipMemcpyHostToDevice, c->stream));
    MONTE_HIP(c, hipMemcpyAsync(base + voff, v.data(), v.size() * sizeof(BpView), hipMemcpyHostToDevice, c->stream));
    MONTE_HIP(c, hipStreamSynchronize(c->stream));
    c->const_key = key;
    c->taps_off = toff;
    c->views_off = voff;
  }
  *taps = (const float*)((char*)c->ws[kSlotConst] + c->taps_off);
  *views = (const BpView*)((char*)c->ws[kSlotConst] + c->views_off);
  return MONTE_OK;
}

// Build the projector's column lists once per geometry: a counting pass, the
// prefix sums on the host, then a writing pass.
int ensure_projector_plan(monte_ctx* c, const monte_geom& g, const BpView* views, const BpGrid& grid,
                          const uint32_t** offs, const uint32_t** list) {
  std::vector<double> key = geom_key(g);
  key.push_back((double)project_strips(grid));
  const size_t ncell = (size_t)g.n_views * project_strips(grid);
  const size_t obytes = align_up((ncell + 1) * sizeof(uint32_t), 256);
  if (key != c->plan_key || !c->ws[kSlotPlan]) {
    c->plan_key.clear();
    uint32_t* d = (uint32_t*)workspace(c, kSlotPlan, obytes);
    if (!d) return fail(c, MONTE_ERR_NOMEM, "projector plan allocation failed");
    MONTE_HIP(c, launch_project_plan(views, g.n_views, grid, d, nullptr, false, c->stream));
    std::vector<uint32_t> h(ncell + 1);
    MONTE_HIP(c, hipMemcpyAsync(h.data(), d, ncell * sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
    MONTE_HIP(c, hipStreamSynchronize(c->stream));
    uint64_t total = 0;
    for (size_t i = 0; i < ncell; ++i) {
      const uint32_t n = h[i];
      h[i] = (uint32_t)total;
      total += n;
      if (total > 0xffffffffull) return fail(c, MONTE_ERR_UNSUPPORTED, "projector plan exceeds 2^32 columns");
    }
    h[ncell] = (uint32_t)total;
    char* base = (char*)workspace(c, kSlotPlan, obytes + total * sizeof(uint32_t));
    if (!base) return fail(c, MONTE_ERR_NOMEM, "projector plan allocation failed");
    MONTE_HIP(c, hipMemcpyAsync(base, h.data(), h.size() * sizeof(uint32_t), hipMemcpyHostToDevice, c->stream));
    MONTE_HIP(c, launch_project_plan(views, g.n_views, grid, (uint32_t*)base, (uint32_t*)(base + obytes), true,
                                     c->stream));
    MONTE_HIP(c, hipStreamSynchronize(c->stream));
    c->plan_key = key;
    c->plan_list_off = obytes;
  }
  *offs = (const uint32_t*)c->ws[kSlotPlan];
  *list = (const uint32_t*)((char*)c->ws[kSlotPlan] + c->plan_list_off);
  return MONTE_OK;
}

size_t proj_elems(const monte_geom& g) { return (size_t)g.n_views * g.nu * g.nv; }
size_t vol_elems(const monte_geom& g) { return (size_t)g.nx * g.ny * g.nz; }
size_t padded_elems(const monte_geom& g) {
  return (size_t)g.n_views * (g.nu + kPadLo + kPadHi) * (g.nv + kPadLo + kPadHi);
}

// Host-pointer mode: copy an input to a staging slot / bring an output back.
struct Stage {
  monte_ctx* c;
  int code = MONTE_OK;
  const void* in(int slot, const void* host, size_t bytes) {
    if (!c->host_ptrs) return host;
    void* d = workspace(c, slot, bytes);
    if (!d) { code = fail(c, MONTE_ERR_NOMEM, "staging allocation failed"); return nullptr; }
    if (hipMemcpyAsync(d, host, bytes, hipMemcpyHostToDevice, c->stream) != hipSuccess) {
      code = fail(c, MONTE_ERR_HIP, "staging copy failed");
      return nullptr;
    }
    return d;
  }
  void* out(int slot, void* host, size_t bytes, bool copy_in) {
    if (!c->host_ptrs) return host;
    void* d = workspace(c, slot, bytes);
    if (!d) { code = fail(c, MONTE_ERR_NOMEM, "staging allocation failed"); return nullptr; }
    if (copy_in && hipMemcpyAsync(d, host, bytes, hipMemcpyHostToDevice, c->stream) != hipSuccess) {
      code = fail(c, MONTE_ERR_HIP, "staging copy failed");
      return nullptr;
    }
    return d;
  }
  int back(void* host, const void* dev, size_t bytes) {
    if (!c->host_ptrs) return MONTE_OK;
    MONTE_HIP(c, hipMemcpyAsync(host, dev, bytes, hipMemcpyDeviceToHost, c->stream));
    MONTE_HIP(c, hipStreamSynchronize(c->stream));
    return MONTE_OK;
  }
};

// ------------------------------------------------------------ FDK stages
int run_filter_padded(monte_ctx* c, const monte_geom& g, const float* proj, float* pad) {
  const float* taps;
  const BpView* views;
  int rc = ensure_constants(c, g, &taps, &views);
  if (rc) return rc;
  const FilterParams fp = compute_filter_params(g);
  MONTE_HIP(c, launch_fdk_filter(proj, pad, taps, fp, c->stream));
  MONTE_HIP(c, launch_fill_pads(pad, g.n_views, g.nu, g.nv, fp.pitch, fp.rows, g.preset == MONTE_PRESET_BP3D20,
                                c->stream));
  return MONTE_OK;
}

int run_backproject_padded(monte_ctx* c, const monte_geom& g, const float* pad, float* vol, int accumulate) {
  const float* taps;
  const BpView* views;
  int rc = ensure_constants(c, g, &taps, &views);
  if (rc) return rc;
  const BpGrid grid = compute_bp_grid(g);
  if (!accumulate && (grid.x0 > 0 || grid.x1 < g.nx)) {
    MONTE_HIP(c, launch_fill(vol, 0.f, vol_elems(g), c->stream));
    accumulate = 1;  // ROI columns added onto zeros
  }
  float4* quads = (float4*)workspace(c, kSlotQuad, padded_elems(g) * sizeof(float4));
  if (!quads) return fail(c, MONTE_ERR_NOMEM, "tap quad allocation failed");
  MONTE_HIP(c, launch_tap_quads(pad, quads, g.n_views, grid.rows, grid.pitch, c->stream));
  const int chunk = c->tuning.bp_view_chunk > 0 ? c->tuning.bp_view_chunk : g.n_views;
  for (int v0 = 0; v0 < g.n_views; v0 += chunk) {
    const int n = std::min(chunk, g.n_views - v0);
    MONTE_HIP(c, launch_backproject(quads, views, v0, n, grid, vol, (accumulate || v0 > 0) ? 1 : 0, c->tuning,
                                    c->stream));
  }
  return MONTE_OK;
}

int run_project_padded(monte_ctx* c, const monte_geom& g, const float* vol, float* pad) {
  const float* taps;
  const BpView* views;
  int rc = ensure_constants(c, g, &taps, &views);
  if (rc) return rc;
  const BpGrid grid = compute_bp_grid(g);
  const uint32_t *plan_offs, *plan_list;
  if ((rc = ensure_projector_plan(c, g, views, grid, &plan_offs, &plan_list))) return rc;
  float* volT = (float*)workspace(c, kSlotVolT, vol_elems(g) * sizeof(float));
  if (!volT) return fail(c, MONTE_ERR_NOMEM, "transposed volume allocation failed");
  MONTE_HIP(c, launch_transpose_zx(vol, volT, g.nx, g.ny, g.nz, c->stream));
  // strip edge columns are added atomically onto zero
  MONTE_HIP(c, launch_fill(pad, 0.f, padded_elems(g), c->stream));
  std::vector<BpView> host_views;
  compute_bp_views(g, host_views);
  const int chunk = c->tuning.bp_view_chunk > 0 ? c->tuning.bp_view_chunk : g.n_views;
  for (int v0 = 0; v0 < g.n_views; v0 += chunk) {
    const int n = std::min(chunk, g.n_views - v0);
    int lds_r0, lds_rows;
    projector_rows(host_views, v0, n, grid, lds_r0, lds_rows);
    MONTE_HIP(c, launch_project(volT, views, v0, n, grid, pad, lds_r0, lds_rows, plan_offs, plan_list, c->tuning,
                                c->stream));
  }
  if (g.preset == MONTE_PRESET_BP3D20)
    MONTE_HIP(c, launch_fold_pads(pad, g.n_views, g.nu, g.nv, grid.pitch, grid.rows, c->stream));
  return MONTE_OK;
}

int run_filter_T_padded(monte_ctx* c, const monte_geom& g, const float* pad, float* proj) {
  const float* taps;
  const BpView* views;
  int rc = ensure_constants(c, g, &taps, &views);
  if (rc) return rc;
  const FilterParams fp = compute_filter_params(g);
  MONTE_HIP(c, launch_fdk_filter_T(pad, proj, taps, fp, c->stream));
  return MONTE_OK;
}

// ------------------------------------------------------------------ MC
struct HostJob {
  McJob d{};
  double org[3] = {0, 0, 0}, vox[3] = {1, 1, 1};  // phantom grid origin (corner) and voxel size
  McTables tables{};
  std::vector<float> cdf;       // empty = mono
  std::vector<float2> view_cs;  // all views of the geometry
  std::vector<double> cs64;     // FP64 (cos, sin) of all views, for the first-flight label sets
  RayGrid ray{};                // first-flight label-set geometry (box and bricks filled in by prepare_mc)
};

// Work units: groups of detector tiles over blocks of photon cycles, about
// `batch` photons each (tiles are grouped when a view has fewer than batch / 64
// cycles); only tile rows that hold photons count when a view has fewer photons
// than pixels.
void set_units(McJob& J, int batch) {
  const uint64_t npix = (uint64_t)J.nu * (uint64_t)J.nv, ppv = J.photons_per_view;
  const uint64_t cycles = (ppv + npix - 1) / npix;
  const uint64_t rows = (std::min(ppv, npix) + J.nu - 1) / J.nu;
  const uint64_t per_unit = std::max<uint64_t>(1, (uint64_t)batch / 64);
  J.tiles_u = (uint32_t)((J.nu + kTileU - 1) / kTileU);
  J.tiles = (uint32_t)(J.tiles_u * ((rows + kTileV - 1) / kTileV));
  J.cycles = (uint32_t)cycles;
  J.cycle_block = (uint32_t)std::min(per_unit, cycles);
  J.tile_group = (uint32_t)std::max<uint64_t>(1, per_unit / J.cycle_block);
  J.cycle_blocks = (uint32_t)((cycles + J.cycle_block - 1) / J.cycle_block);
  J.batches_per_view = ((J.tiles + J.tile_group - 1) / J.tile_group) * J.cycle_blocks;
  J.unit_count = (uint64_t)J.batches_per_view * (uint64_t)J.n_views_sim;
}

int build_job(monte_ctx* c, const monte_geom* g, const monte_phantom* ph, const monte_mc_cfg* cfg, HostJob& hj,
              int batch) {
  int rc = check_geom(c, g, false);
  if (rc) return rc;
  if (!ph || !cfg) return fail(c, MONTE_ERR_INVALID, "phantom/config is null");
  if (!ph->labels) return fail(c, MONTE_ERR_INVALID, "phantom labels are null");
  if (ph->nx < 1 || ph->ny < 1 || ph->nz < 1 || !(ph->vx > 0) || !(ph->vy > 0) || !(ph->vz > 0))
    return fail(c, MONTE_ERR_INVALID, "bad phantom dimensions");
  if ((double)ph->nx * ph->ny * ph->nz > 4294967295.0)
    return fail(c, MONTE_ERR_INVALID, "phantom has more than 2^32 - 1 voxels");
  // the voxel index is two 24-bit multiply-adds: every operand, ny and the
  // partial index iz * ny + iy < ny * nz included, must fit in 24 bits
  if ((double)ph->ny * ph->nz >= 16777216.0 || ph->nx >= 16777216)
    return fail(c, MONTE_ERR_UNSUPPORTED, "phantom ny * nz and nx must be below 2^24 (24-bit voxel indexing)");
  if (ph->n_labels < 1 || ph->n_labels > kMaxLabels) return fail(c, MONTE_ERR_INVALID, "n_labels out of range");
  if (c->n_tables < 1) return fail(c, MONTE_ERR_STATE, "attenuation tables not loaded (monte_load_tables)");
  // photon indices are 32-bit (the kernel forms cycle * nu nv + pixel in 64 bits
  // and keeps the photons below photons_per_view)
  if (cfg->photons_per_view == 0 || cfg->photons_per_view > 0x100000000ull - (uint64_t)batch - 64)
    return fail(c, MONTE_ERR_INVALID, "photons_per_view must be in [1, 2^32 - batch - 64] (batch = 64 * mc_batch)");
  if (cfg->max_interactions < 0) return fail(c, MONTE_ERR_INVALID, "max_interactions < 0");
  if (!cfg->spectrum && !(cfg->e_kev >= 1.0 && cfg->e_kev <= kRows - 1))
    return fail(c, MONTE_ERR_INVALID, "e_kev must be in [1, 200]");
  const double half_x = 0.5 * ph->nx * ph->vx + std::fabs(ph->ox), half_y = 0.5 * ph->ny * ph->vy + std::fabs(ph->oy);
  if (std::sqrt(half_x * half_x + half_y * half_y) >= g->sod)
    return fail(c, MONTE_ERR_INVALID, "phantom reaches the source orbit");
  if (std::sqrt(half_x * half_x + half_y * half_y) >= g->sdd - g->sod)
    return fail(c, MONTE_ERR_INVALID, "phantom reaches the detector plane");
  int vb = cfg->view_begin, ve = cfg->view_end;
  if (vb == 0 && ve == 0) ve = g->n_views;
  if (vb < 0 || ve > g->n_views || vb >= ve) return fail(c, MONTE_ERR_INVALID, "view range outside the scan");

  // tables: mu = density * total; fractions from the mass coefficients
  McTables& T = hj.tables;
  std::memset(&T, 0, sizeof T);
  for (int e = 0; e < kRows; ++e) {
    float mx = 0.f;
    for (int l = 0; l < ph->n_labels; ++l) {
      const int tb = ph->mat_table[l];
      if (tb < 0) continue;
      if (tb >= c->n_tables) return fail(c, MONTE_ERR_INVALID, "mat_table index beyond the loaded tables");
      if (!(ph->density[l] >= 0)) return fail(c, MONTE_ERR_INVALID, "negative density");
      const double* t = c->tables.data() + (size_t)tb * 4 * kRows;
      const double tot = t[3 * kRows + e];
      float4 m;
      m.x = (float)(ph->density[l] * tot);
      m.y = tot > 0 ? (float)(t[2 * kRows + e] / tot) : 1.f;
      m.z = tot > 0 ? (float)((t[2 * kRows + e] + t[0 * kRows + e]) / tot) : 1.f;
      m.w = 0.f;
      T.m[l][e] = m;
      mx = std::max(mx, m.x);
    }
    T.mumax[e] = mx;
  }
  hj.cdf.clear();
  if (cfg->spectrum) {
    double tot = 0;
    for (int e = 1; e < kRows; ++e) {
      if (!(cfg->spectrum[e] >= 0)) return fail(c, MONTE_ERR_INVALID, "negative spectrum weight");
      tot += cfg->spectrum[e];
    }
    if (!(tot > 0)) return fail(c, MONTE_ERR_INVALID, "spectrum has no weight above 0 keV");
    hj.cdf.assign(kRows, 0.f);
    double run = 0;
    for (int e = 1; e < kRows; ++e) {
      run += cfg->spectrum[e];
      hj.cdf[e] = (float)(run / tot);
    }
    hj.cdf[kRows - 1] = 1.0f;
  }
  hj.view_cs.resize(g->n_views);
  hj.cs64.resize(2 * (size_t)g->n_views);
  for (int k = 0; k < g->n_views; ++k) {
    const double b = g->preset == MONTE_PRESET_BP3D20 ? 3.14159265358979323846 * (k * g->step_deg) / 180.0
                                                      : g->angle0 + k * g->dangle;
    hj.cs64[2 * k] = std::cos(b);
    hj.cs64[2 * k + 1] = std::sin(b);
    hj.view_cs[k] = make_float2((float)hj.cs64[2 * k], (float)hj.cs64[2 * k + 1]);
  }
  RayGrid& R = hj.ray;
  if (g->preset == MONTE_PRESET_BP3D20) {
    R.un0 = 0.5 * g->nu; R.u_step = -g->du; R.vn0 = 0.5 * g->nv; R.v_step = -g->dv;
  } else {
    R.un0 = 0.5 * g->nu - g->u_off / g->du; R.u_step = g->du;
    R.vn0 = 0.5 * g->nv - g->v_off / g->dv; R.v_step = g->dv;
  }
  R.sod = g->sod;
  R.xd = g->sdd - g->sod;
  R.nu = g->nu;
  R.nv = g->nv;
  McJob& J = hj.d;
  J.labels = ph->labels;
  J.nx = ph->nx; J.ny = ph->ny; J.nz = ph->nz;
  J.nu = g->nu; J.nv = g->nv;
  J.view_begin = vb;
  J.n_views_sim = ve - vb;
  J.vx = (float)ph->vx; J.vy = (float)ph->vy; J.vz = (float)ph->vz;
  hj.org[0] = ph->ox - 0.5 * ph->nx * ph->vx; hj.org[1] = ph->oy - 0.5 * ph->ny * ph->vy;
  hj.org[2] = ph->oz - 0.5 * ph->nz * ph->vz;
  hj.vox[0] = ph->vx; hj.vox[1] = ph->vy; hj.vox[2] = ph->vz;
  J.bx0 = (float)(ph->ox - 0.5 * ph->nx * ph->vx); J.bx1 = (float)(ph->ox + 0.5 * ph->nx * ph->vx);
  J.by0 = (float)(ph->oy - 0.5 * ph->ny * ph->vy); J.by1 = (float)(ph->oy + 0.5 * ph->ny * ph->vy);
  J.bz0 = (float)(ph->oz - 0.5 * ph->nz * ph->vz); J.bz1 = (float)(ph->oz + 0.5 * ph->nz * ph->vz);
  J.inv_vx = (float)(1.0 / ph->vx); J.inv_vy = (float)(1.0 / ph->vy); J.inv_vz = (float)(1.0 / ph->vz);
  J.ix0 = (float)(-(double)J.bx0 / ph->vx); J.iy0 = (float)(-(double)J.by0 / ph->vy);
  J.iz0 = (float)(-(double)J.bz0 / ph->vz);
  J.sod = (float)g->sod; J.sdd = (float)g->sdd;
  if (g->preset == MONTE_PRESET_BP3D20) {
    // CBCT_real325.cu:565-566: index -(int)(10 d - 162.5), i.e. u_idx = nu/2 - y'/du
    J.un0 = (float)(0.5 * g->nu); J.us = (float)(-1.0 / g->du); J.u_step = (float)(-g->du);
    J.vn0 = (float)(0.5 * g->nv); J.vs = (float)(-1.0 / g->dv); J.v_step = (float)(-g->dv);
  } else {
    J.un0 = (float)(0.5 * g->nu - g->u_off / g->du); J.us = (float)(1.0 / g->du); J.u_step = (float)g->du;
    J.vn0 = (float)(0.5 * g->nv - g->v_off / g->dv); J.vs = (float)(1.0 / g->dv); J.v_step = (float)g->dv;
  }
  J.e_kev = (float)cfg->e_kev;
  J.max_int = cfg->max_interactions;
  J.coherent = cfg->coherent_mode == MONTE_COH_RAYLEIGH ? 1 : 0;
  J.key0 = (uint32_t)(cfg->seed & 0xFFFFFFFFull);
  J.key1 = (uint32_t)(cfg->seed >> 32);
  J.job_id = cfg->job;
  J.photons_per_view = (uint32_t)cfg->photons_per_view;
  set_units(J, batch);
  return MONTE_OK;
}

McLayout mc_layout(const std::vector<HostJob>& hj) {
  const int n = (int)hj.size();
  McLayout L;
  size_t off = 0;
  L.jobs_off = off; off = align_up(off + sizeof(McJob) * n, 256);
  L.tab_off = off; off = align_up(off + sizeof(McTables) * n, 256);
  L.cdf_off.resize(n);
  L.cs_off.resize(n);
  for (int i = 0; i < n; ++i) {
    L.cdf_off[i] = off; off = align_up(off + sizeof(float) * hj[i].cdf.size(), 256);
    L.cs_off[i] = off; off = align_up(off + sizeof(float2) * hj[i].view_cs.size(), 256);
  }
  L.q_off = off; off = align_up(off + kMcQueueBytes, 256);
  L.cnt_off = off; off = align_up(off + sizeof(McCounters) * n, 256);
  L.box_off = off; off = align_up(off + sizeof(int) * 8 * n, 256);
  L.total = off;
  return L;
}

// First-flight label sets of every job (tuning mc_ray_brick > 0) in workspace
// slot `slot`: the FP64 view tables, the brick sets and the per-view,
// per-pixel sets; each job's descriptor then points at its sets.  The flight
// boxes must be final.
int prepare_ray_sets(monte_ctx* c, std::vector<HostJob>& hj, int slot) {
  const int b = c->tuning.mc_ray_brick;
  const int n = (int)hj.size();
  std::vector<size_t> cs_off(n), br_off(n), set_off(n);
  size_t off = 0;
  for (int i = 0; i < n; ++i) {
    HostJob& h = hj[i];
    const McJob& J = h.d;
    RayGrid& R = h.ray;
    const int dims[3] = {J.nx, J.ny, J.nz};
    for (int a = 0; a < 3; ++a) {
      R.org[a] = h.org[a];
      R.bw[a] = b * h.vox[a];
      R.nb[a] = (dims[a] + b - 1) / b;
    }
    cs_off[i] = off; off = align_up(off + sizeof(double) * h.cs64.size(), 256);
    br_off[i] = off; off = align_up(off + (size_t)R.nb[0] * R.nb[1] * R.nb[2], 256);
    set_off[i] = off; off = align_up(off + (size_t)J.n_views_sim * J.nu * J.nv, 256);
  }
  char* base = (char*)workspace(c, slot, off);
  if (!base) return fail(c, MONTE_ERR_NOMEM, "first-flight label set allocation failed");
  for (int i = 0; i < n; ++i) {
    HostJob& h = hj[i];
    McJob& J = h.d;
    MONTE_HIP(c, hipMemcpyAsync(base + cs_off[i], h.cs64.data(), sizeof(double) * h.cs64.size(),
                                hipMemcpyHostToDevice, c->stream));
    uint8_t* bricks = (uint8_t*)(base + br_off[i]);
    uint8_t* sets = (uint8_t*)(base + set_off[i]);
    MONTE_HIP(c, launch_brick_sets(J.labels, J.nx, J.ny, J.nz, b, h.ray.nb, bricks, c->stream));
    MONTE_HIP(c, launch_ray_sets(h.ray, (const double*)(base + cs_off[i]), J.view_begin, J.n_views_sim, bricks, sets,
                                 c->stream));
    J.ray_set = sets;
  }
  return MONTE_OK;
}

// Lays out the jobs in `base` (at least mc_layout(hj).total bytes; the
// context's MC workspace when null), computes each job's flight box and, with
// tuning mc_ray_brick > 0 and ray_slot >= 0, its first-flight label sets in
// workspace slot ray_slot, and uploads everything; returns once the uploads
// are complete.
int prepare_mc(monte_ctx* c, std::vector<HostJob>& hj, int block, int blocks_per_cu, char* base, McPrepared& P,
               int ray_slot) {
  const int n = (int)hj.size();
  const McLayout Lo = mc_layout(hj);
  if (!base) base = (char*)workspace(c, kSlotMc, Lo.total);
  if (!base) return fail(c, MONTE_ERR_NOMEM, "MC descriptor allocation failed");

  // Flights are clipped to the bounding box of the voxels that attenuate at
  // some energy: outside it mu is 0, so no history changes, and no Woodcock
  // steps are spent crossing the empty margin of the phantom grid.
  {
    std::vector<int> box(8 * (size_t)n);
    for (int i = 0; i < n; ++i) {
      int* b = box.data() + 8 * i;
      b[0] = b[1] = b[2] = INT_MAX;
      b[3] = b[4] = b[5] = -1;
    }
    int* dbox = (int*)(base + Lo.box_off);
    MONTE_HIP(c, hipMemcpyAsync(dbox, box.data(), sizeof(int) * box.size(), hipMemcpyHostToDevice, c->stream));
    for (int i = 0; i < n; ++i) {
      uint32_t solid = 0;
      for (int l = 0; l < kMaxLabels; ++l)
        for (int e = 0; e < kRows; ++e)
          if (hj[i].tables.m[l][e].x > 0.f) { solid |= 1u << l; break; }
      const McJob& J = hj[i].d;
      MONTE_HIP(c, launch_support_box(J.labels, J.nx, J.ny, J.nz, solid, dbox + 8 * i, c->stream));
    }
    MONTE_HIP(c, hipMemcpyAsync(box.data(), dbox, sizeof(int) * box.size(), hipMemcpyDeviceToHost, c->stream));
    MONTE_HIP(c, hipStreamSynchronize(c->stream));
    for (int i = 0; i < n; ++i) {
      const int* b = box.data() + 8 * i;
      McJob& J = hj[i].d;
      const double* o = hj[i].org;
      const double* v = hj[i].vox;
      RayGrid& R = hj[i].ray;
      const int dims[3] = {J.nx, J.ny, J.nz};
      if (b[3] < 0) {  // nothing attenuates: keep the grid box
        for (int a = 0; a < 3; ++a) {
          R.c0[a] = o[a];
          R.c1[a] = o[a] + dims[a] * v[a];
        }
        continue;
      }
      for (int a = 0; a < 3; ++a) {
        R.c0[a] = o[a] + b[a] * v[a];
        R.c1[a] = o[a] + (b[3 + a] + 1) * v[a];
      }
      J.bx0 = (float)R.c0[0]; J.bx1 = (float)R.c1[0];
      J.by0 = (float)R.c0[1]; J.by1 = (float)R.c1[1];
      J.bz0 = (float)R.c0[2]; J.bz1 = (float)R.c1[2];
    }
  }
  for (int i = 0; i < n; ++i) hj[i].d.ray_set = nullptr;
  if (ray_slot >= 0 && c->tuning.mc_ray_brick > 0) {
    const int rc = prepare_ray_sets(c, hj, ray_slot);
    if (rc) return rc;
  }

  uint64_t units = 0;
  std::vector<McJob> jobs(n);
  for (int i = 0; i < n; ++i) {
    McJob J = hj[i].d;
    J.tables = (const McTables*)(base + Lo.tab_off) + i;
    J.spectrum_cdf = hj[i].cdf.empty() ? nullptr : (const float*)(base + Lo.cdf_off[i]);
    J.view_cs = (const float2*)(base + Lo.cs_off[i]);
    J.unit_begin = units;
    units += J.unit_count;
    jobs[i] = J;
  }
  // the 32-bit queue counter passes the last unit by one fetch per wave
  const int grid_waves = std::max(1, c->n_cu * blocks_per_cu) * (block / 64);
  if (units > 0xFFFFFFFFull - (uint64_t)grid_waves)
    return fail(c, MONTE_ERR_INVALID, "too many work units in one launch (raise the batch)");
  MONTE_HIP(c, hipMemcpyAsync(base + Lo.jobs_off, jobs.data(), sizeof(McJob) * n, hipMemcpyHostToDevice, c->stream));
  for (int i = 0; i < n; ++i) {
    MONTE_HIP(c, hipMemcpyAsync(base + Lo.tab_off + sizeof(McTables) * i, &hj[i].tables, sizeof(McTables),
                                hipMemcpyHostToDevice, c->stream));
    if (!hj[i].cdf.empty())
      MONTE_HIP(c, hipMemcpyAsync(base + Lo.cdf_off[i], hj[i].cdf.data(), sizeof(float) * hj[i].cdf.size(),
                                  hipMemcpyHostToDevice, c->stream));
    MONTE_HIP(c, hipMemcpyAsync(base + Lo.cs_off[i], hj[i].view_cs.data(), sizeof(float2) * hj[i].view_cs.size(),
                                hipMemcpyHostToDevice, c->stream));
  }
  MONTE_HIP(c, hipStreamSynchronize(c->stream));  // the host copies above must outlive the uploads
  P.base = base;
  P.L = Lo;
  P.n = n;
  P.units = units;
  // A chunk is one tile group over all of the first job's views.  A wave changes
  // job only once its photons of the old job have all ended, so in a launch of
  // several jobs a chunk holds at least 8 units per wave drawing from its queue:
  // a wave then takes several units of one job in a row instead of a unit of
  // another job almost every time.  64 uneven jobs (144k units): 13.7 ms with
  // chunks of one job's views, 13.3, 13.0, 13.1, 12.7 and 13.2 ms with 1, 2, 4,
  // 8 and 16 units per wave, the tallies bitwise the same.
  P.chunk = (uint32_t)std::max(1, hj[0].d.n_views_sim);
  if (n > 1) P.chunk = std::max<uint32_t>(P.chunk, (uint32_t)(8 * grid_waves / kMcQueues));
  return MONTE_OK;
}

// Resets the work queue and counters, then launches the transport (with
// `start` recorded between the two when given, so it times the transport alone).
hipError_t launch_prepared(monte_ctx* c, const McPrepared& P, int block, int blocks_per_cu,
                           hipEvent_t start = nullptr) {
  hipError_t e = launch_mc_reset(P.queue(), P.counters(), P.n, c->stream);
  if (e != hipSuccess) return e;
  if (start && (e = hipEventRecord(start, c->stream)) != hipSuccess) return e;
  const int grid = std::max(1, c->n_cu * blocks_per_cu);
  return launch_mc_transport((const McJob*)(P.base + P.L.jobs_off), P.n, P.units, P.chunk, P.queue(), P.counters(),
                             block, grid, c->stream);
}

int read_counters(monte_ctx* c, const McPrepared& P, std::vector<McCounters>& cnt) {
  cnt.resize(P.n);
  MONTE_HIP(c, hipMemcpyAsync(cnt.data(), P.counters(), sizeof(McCounters) * P.n, hipMemcpyDeviceToHost, c->stream));
  MONTE_HIP(c, hipStreamSynchronize(c->stream));
  return MONTE_OK;
}

// Prepares and runs one launch in a device buffer: the context's MC workspace,
// or `base` (at least mc_layout(hj).total bytes) when the caller provides one.
int run_mc(monte_ctx* c, std::vector<HostJob>& hj, int block, int blocks_per_cu, float* ms_out,
           std::vector<McCounters>* counters_out, char* base = nullptr, int ray_slot = kSlotRay) {
  McPrepared P;
  int rc = prepare_mc(c, hj, block, blocks_per_cu, base, P, ray_slot);
  if (rc) return rc;
  const hipError_t le = launch_prepared(c, P, block, blocks_per_cu, c->ev0);
  MONTE_HIP(c, hipEventRecord(c->ev1, c->stream));
  std::vector<McCounters> cnt;
  if ((rc = read_counters(c, P, cnt))) return rc;
  float ms = 0.f;
  (void)hipEventElapsedTime(&ms, c->ev0, c->ev1);
  if (le != hipSuccess) return fail(c, MONTE_ERR_HIP, std::string("transport launch: ") + hipGetErrorString(le));
  if (ms_out) *ms_out = ms;
  if (counters_out) *counters_out = cnt;
  return MONTE_OK;
}

void fill_stats(monte_mc_stats* s, const McCounters& k, double ms) {
  s->histories = k.histories; s->steps = k.steps; s->interactions = k.interactions;
  s->photo = k.photo; s->coherent = k.coherent; s->compton = k.compton;
  s->tallied_primary = k.tallied_primary; s->tallied_scatter = k.tallied_scatter;
  s->missed = k.missed; s->capped = k.capped;
  s->kernel_ms = ms;
}

// Key of an MC job's prepared descriptors: everything build_job reads.
std::vector<double> mc_job_key(const monte_ctx* c, const monte_geom& g, const monte_phantom& ph,
                               const monte_mc_cfg& cfg) {
  std::vector<double> k = geom_key(g);
  const double vals[] = {(double)(uintptr_t)ph.labels, (double)ph.nx, (double)ph.ny, (double)ph.nz,
                         (double)ph.n_labels, ph.vx, ph.vy, ph.vz, ph.ox, ph.oy, ph.oz,
                         (double)cfg.photons_per_view, cfg.e_kev, (double)cfg.max_interactions,
                         (double)cfg.coherent_mode, (double)(cfg.seed & 0xFFFFFFFFull), (double)(cfg.seed >> 32),
                         (double)cfg.job, (double)cfg.view_begin, (double)cfg.view_end, (double)c->tables_gen,
                         (double)c->tuning.mc_batch, (double)c->tuning.mc_block, (double)c->tuning.mc_blocks_per_cu,
                         (double)c->tuning.mc_ray_brick};
  k.insert(k.end(), std::begin(vals), std::end(vals));
  for (int l = 0; l < kMaxLabels; ++l) {
    k.push_back((double)ph.mat_table[l]);
    k.push_back(ph.density[l]);
  }
  k.push_back(cfg.spectrum ? 1.0 : 0.0);
  if (cfg.spectrum) k.insert(k.end(), cfg.spectrum, cfg.spectrum + kRows);
  return k;
}

// The launches of one end-to-end step, in stream order (no host syncs, no
// allocation once the workspaces have their sizes): zero the tallies, MC
// transport, counts -> line integrals of primary + scatter, FDK, residual and
// loss, adjoint of the FDK applied to the residual.
int e2e_launches(monte_ctx* c, const monte_geom& g, const monte_e2e_buffers& b, uint64_t ppv, double* partial) {
  const size_t np = proj_elems(g), nvx = vol_elems(g);
  MONTE_HIP(c, launch_fill((float*)b.primary, 0.f, 2 * np, c->stream));
  MONTE_HIP(c, launch_fill((float*)b.scatter, 0.f, 2 * np, c->stream));
  const monte_tuning& t = c->tuning;
  MONTE_HIP(c, launch_prepared(c, c->e2e_mc, t.mc_block, t.mc_blocks_per_cu));
  MONTE_HIP(c, launch_counts_to_lineint(b.primary, b.scatter, b.lineint, g.n_views, g.nu, g.nv, ppv, c->stream));
  float* pad = (float*)c->ws[kSlotFwd];
  int rc = run_filter_padded(c, g, b.lineint, pad);
  if (!rc) rc = run_backproject_padded(c, g, pad, b.vol, 0);
  if (rc) return rc;
  MONTE_HIP(c, launch_residual_loss(b.vol, b.vol_ref, b.residual, nvx, partial, b.loss, c->stream));
  float* apad = (float*)c->ws[kSlotAdj];
  if ((rc = run_project_padded(c, g, b.residual, apad))) return rc;
  return run_filter_T_padded(c, g, apad, b.grad);
}

monte_tuning default_tuning() {
  monte_tuning t{};
  t.mc_block = 256;
  // 6 blocks of 256 per CU is the occupancy of the transport kernel (6 waves per
  // SIMD: its 6.6 KB LDS slot per wave and its 77 VGPRs both allow 6).  A unit of 64 cycles of
  // an 8 x 8 detector tile was fastest at 325^3 / 512^2 / 360 views x 1e7 photons
  // (102.1 ms at 4 waves per SIMD; 102.6, 103.7 and 108.9 ms for 32, 128 and 256
  // cycles): fewer, larger units contend less on the queue counters, smaller ones
  // keep each XCD's units on fewer detector tiles.
  t.mc_batch = 64;
  t.mc_blocks_per_cu = 6;
  t.bp_zchunk = 16;
  t.bp_tile_x = 8;   // fastest at 325^3 / 512^2 / 360 views
  t.bp_tile_y = 32;
  t.bp_view_chunk = 0;
  t.pj_zchunk = 0;
  t.mc_ray_brick = 4;  // 2.45 Woodcock steps per history at 325^3 / 512^2 (7.56 with the global majorant)
  return t;
}

}  // namespace

// ====================================================================== C ABI
extern "C" {

const char* monte_status_str(int s) {
  switch (s) {
    case MONTE_OK: return "ok";
    case MONTE_ERR_INVALID: return "invalid argument";
    case MONTE_ERR_HIP: return "HIP runtime error";
    case MONTE_ERR_NOMEM: return "out of device memory";
    case MONTE_ERR_STATE: return "missing prerequisite";
    case MONTE_ERR_IO: return "file error";
    case MONTE_ERR_UNSUPPORTED: return "unsupported";
    default: return "unknown status";
  }
}

const char* monte_last_error(const monte_ctx* c) { return c ? c->err.c_str() : "null context"; }

int monte_create(monte_ctx** out, int device) {
  if (!out) return MONTE_ERR_INVALID;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return MONTE_ERR_HIP;
  if (device < 0 || device >= n) return MONTE_ERR_INVALID;
  if (hipSetDevice(device) != hipSuccess) return MONTE_ERR_HIP;
  monte_ctx* c = new monte_ctx();
  c->device = device;
  if (hipStreamCreateWithFlags(&c->own_stream, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    return MONTE_ERR_HIP;
  }
  c->stream = c->own_stream;
  if (hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess) {
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    (void)hipStreamDestroy(c->own_stream);
    delete c;
    return MONTE_ERR_HIP;
  }
  // the allocation-free entry points' scratch: located once per device
  if (noalloc_setup(device) != 0) {
    (void)hipEventDestroy(c->ev0);
    (void)hipEventDestroy(c->ev1);
    (void)hipStreamDestroy(c->own_stream);
    delete c;
    return MONTE_ERR_HIP;
  }
  int cu = 0;
  if (hipDeviceGetAttribute(&cu, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess && cu > 0) c->n_cu = cu;
  c->tuning = default_tuning();
  *out = c;
  return MONTE_OK;
}

void monte_destroy(monte_ctx* c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->stream);
  if (c->graph_exec) (void)hipGraphExecDestroy(c->graph_exec);
  if (c->graph) (void)hipGraphDestroy(c->graph);
  if (c->e2e_exec) (void)hipGraphExecDestroy(c->e2e_exec);
  if (c->e2e_graph) (void)hipGraphDestroy(c->e2e_graph);
  for (int i = 0; i < monte_ctx::kSlots; ++i)
    if (c->ws[i]) (void)hipFree(c->ws[i]);
  if (c->ev0) (void)hipEventDestroy(c->ev0);
  if (c->ev1) (void)hipEventDestroy(c->ev1);
  if (c->own_stream) (void)hipStreamDestroy(c->own_stream);
  delete c;
}

int monte_set_stream(monte_ctx* c, void* stream) {
  if (!c) return MONTE_ERR_INVALID;
  c->stream = (hipStream_t)stream;  // NULL: the device's default (null) stream
  return MONTE_OK;
}

int monte_set_pointer_mode(monte_ctx* c, int host_pointers) {
  if (!c) return MONTE_ERR_INVALID;
  c->host_ptrs = host_pointers != 0;
  return MONTE_OK;
}

int monte_synchronize(monte_ctx* c) {
  if (!c) return MONTE_ERR_INVALID;
  MONTE_HIP(c, hipStreamSynchronize(c->stream));
  return MONTE_OK;
}

int monte_get_tuning(const monte_ctx* c, monte_tuning* out) {
  if (!c || !out) return MONTE_ERR_INVALID;
  *out = c->tuning;
  return MONTE_OK;
}

int monte_set_tuning(monte_ctx* c, const monte_tuning* in) {
  if (!c || !in) return MONTE_ERR_INVALID;
  monte_tuning t = *in;
  if (t.mc_block <= 0 || t.mc_block > 256 || t.mc_block % 64) return fail(c, MONTE_ERR_INVALID, "mc_block");
  if (t.mc_batch <= 0 || t.mc_batch > 4096) return fail(c, MONTE_ERR_INVALID, "mc_batch");
  if (t.mc_blocks_per_cu <= 0 || t.mc_blocks_per_cu > 64) return fail(c, MONTE_ERR_INVALID, "mc_blocks_per_cu");
  if (t.bp_zchunk != 4 && t.bp_zchunk != 8 && t.bp_zchunk != 16) return fail(c, MONTE_ERR_INVALID, "bp_zchunk");
  if (t.bp_tile_x <= 0 || 256 % t.bp_tile_x) return fail(c, MONTE_ERR_INVALID, "bp_tile_x");
  t.bp_tile_y = 256 / t.bp_tile_x;
  if (t.bp_view_chunk < 0) return fail(c, MONTE_ERR_INVALID, "bp_view_chunk");
  if (t.pj_zchunk != 0 && t.pj_zchunk != 2 && t.pj_zchunk != 4 && t.pj_zchunk != 8)
    return fail(c, MONTE_ERR_INVALID, "pj_zchunk");
  if (t.mc_ray_brick < 0 || t.mc_ray_brick > 64) return fail(c, MONTE_ERR_INVALID, "mc_ray_brick");
  c->tuning = t;
  return MONTE_OK;
}

int monte_read_xcom_csv(const char* path, double* tab) {
  if (!path || !tab) return MONTE_ERR_INVALID;
  FILE* f = std::fopen(path, "rb");
  if (!f) return MONTE_ERR_IO;
  std::memset(tab, 0, sizeof(double) * 4 * kRows);
  char line[512];
  int k = 0;
  while (std::fgets(line, sizeof line, f)) {
    char* p = line;
    if ((unsigned char)p[0] == 0xEF && (unsigned char)p[1] == 0xBB && (unsigned char)p[2] == 0xBF) p += 3;
    double v[4];
    int n = 0;
    char* q = p;
    while (n < 4) {
      char* e;
      v[n] = std::strtod(q, &e);
      if (e == q) break;
      ++n;
      q = e;
      while (*q == ',' || *q == ' ' || *q == '\t') ++q;
    }
    if (n != 4) continue;  // blank or trailing CR-only line
    if (k + 1 >= kRows) break;
    ++k;
    for (int col = 0; col < 4; ++col) tab[col * kRows + k] = v[col];
  }
  std::fclose(f);
  if (k == 0) return MONTE_ERR_IO;
  for (int col = 0; col < 4; ++col) tab[col * kRows] = tab[col * kRows + 1];
  return k;
}

int monte_load_tables(monte_ctx* c, const double* tab, int n_tables) {
  if (!c || !tab || n_tables < 1) return MONTE_ERR_INVALID;
  for (size_t i = 0; i < (size_t)n_tables * 4 * kRows; ++i)
    if (!(tab[i] >= 0)) return fail(c, MONTE_ERR_INVALID, "attenuation tables must be >= 0");
  c->tables.assign(tab, tab + (size_t)n_tables * 4 * kRows);
  c->n_tables = n_tables;
  c->tables_gen += 1;
  return MONTE_OK;
}

// --------------------------------------------------------------------- FDK
int monte_fdk_filter(monte_ctx* c, const monte_geom* g, const float* proj, float* filt, int batch) {
  if (!c) return MONTE_ERR_INVALID;
  int rc = check_geom(c, g, false);
  if (rc) return rc;
  if (!proj || !filt || batch < 1) return fail(c, MONTE_ERR_INVALID, "null buffer or batch < 1");
  MONTE_HIP(c, hipSetDevice(c->device));
  const size_t np = proj_elems(*g);
  float* pad = (float*)workspace(c, kSlotFwd, padded_elems(*g) * sizeof(float));
  if (!pad) return fail(c, MONTE_ERR_NOMEM, "filtered stack allocation failed");
  Stage st{c};
  const float* din = (const float*)st.in(kSlotStageA, proj, np * batch * sizeof(float));
  float* dout = (float*)st.out(kSlotStageB, filt, np * batch * sizeof(float), false);
  if (st.code) return st.code;
  const FilterParams fp = compute_filter_params(*g);
  for (int b = 0; b < batch; ++b) {
    if ((rc = run_filter_padded(c, *g, din + b * np, pad))) return rc;
    MONTE_HIP(c, launch_unpad_copy(pad, dout + b * np, g->n_views, g->nu, g->nv, fp.pitch, fp.rows, c->stream));
  }
  return st.back(filt, dout, np * batch * sizeof(float));
}

int monte_fdk_backproject(monte_ctx* c, const monte_geom* g, const float* filt, float* vol, int batch,
                          int accumulate) {
  if (!c) return MONTE_ERR_INVALID;
  int rc = check_geom(c, g, true);
  if (rc) return rc;
  if (!filt || !vol || batch < 1) return fail(c, MONTE_ERR_INVALID, "null buffer or batch < 1");
  MONTE_HIP(c, hipSetDevice(c->device));
  const size_t np = proj_elems(*g), nvx = vol_elems(*g);
  float* pad = (float*)workspace(c, kSlotFwd, padded_elems(*g) * sizeof(float));
  if (!pad) return fail(c, MONTE_ERR_NOMEM, "filtered stack allocation failed");
  Stage st{c};
  const float* din = (const float*)st.in(kSlotStageA, filt, np * batch * sizeof(float));
  float* dout = (float*)st.out(kSlotStageB, vol, nvx * batch * sizeof(float), accumulate != 0);
  if (st.code) return st.code;
  const FilterParams fp = compute_filter_params(*g);
  for (int b = 0; b < batch; ++b) {
    MONTE_HIP(c, launch_pad_copy(din + b * np, pad, g->n_views, g->nu, g->nv, fp.pitch, fp.rows, c->stream));
    if (g->preset == MONTE_PRESET_BP3D20)
      MONTE_HIP(c, launch_fill_pads(pad, g->n_views, g->nu, g->nv, fp.pitch, fp.rows, 1, c->stream));
    if ((rc = run_backproject_padded(c, *g, pad, dout + b * nvx, accumulate))) return rc;
  }
  return st.back(vol, dout, nvx * batch * sizeof(float));
}

int monte_fdk(monte_ctx* c, const monte_geom* g, const float* proj, float* vol, int batch) {
  if (!c) return MONTE_ERR_INVALID;
  int rc = check_geom(c, g, true);
  if (rc) return rc;
  if (!proj || !vol || batch < 1) return fail(c, MONTE_ERR_INVALID, "null buffer or batch < 1");
  MONTE_HIP(c, hipSetDevice(c->device));
  const size_t np = proj_elems(*g), nvx = vol_elems(*g);
  float* pad = (float*)workspace(c, kSlotFwd, padded_elems(*g) * sizeof(float));
  if (!pad) return fail(c, MONTE_ERR_NOMEM, "filtered stack allocation failed");
  Stage st{c};
  const float* din = (const float*)st.in(kSlotStageA, proj, np * batch * sizeof(float));
  float* dout = (float*)st.out(kSlotStageB, vol, nvx * batch * sizeof(float), false);
  if (st.code) return st.code;
  for (int b = 0; b < batch; ++b) {
    if ((rc = run_filter_padded(c, *g, din + b * np, pad))) return rc;
    if ((rc = run_backproject_padded(c, *g, pad, dout + b * nvx, 0))) return rc;
  }
  return st.back(vol, dout, nvx * batch * sizeof(float));
}

int monte_fdk_project(monte_ctx* c, const monte_geom* g, const float* vol, float* filt, int batch) {
  if (!c) return MONTE_ERR_INVALID;
  int rc = check_geom(c, g, true);
  if (rc) return rc;
  if (!vol || !filt || batch < 1) return fail(c, MONTE_ERR_INVALID, "null buffer or batch < 1");
  MONTE_HIP(c, hipSetDevice(c->device));
  const size_t np = proj_elems(*g), nvx = vol_elems(*g);
  float* pad = (float*)workspace(c, kSlotAdj, padded_elems(*g) * sizeof(float));
  if (!pad) return fail(c, MONTE_ERR_NOMEM, "adjoint stack allocation failed");
  Stage st{c};
  const float* din = (const float*)st.in(kSlotStageA, vol, nvx * batch * sizeof(float));
  float* dout = (float*)st.out(kSlotStageB, filt, np * batch * sizeof(float), false);
  if (st.code) return st.code;
  const FilterParams fp = compute_filter_params(*g);
  for (int b = 0; b < batch; ++b) {
    if ((rc = run_project_padded(c, *g, din + b * nvx, pad))) return rc;
    MONTE_HIP(c, launch_unpad_copy(pad, dout + b * np, g->n_views, g->nu, g->nv, fp.pitch, fp.rows, c->stream));
  }
  return st.back(filt, dout, np * batch * sizeof(float));
}

int monte_fdk_filter_adjoint(monte_ctx* c, const monte_geom* g, const float* filt, float* proj, int batch) {
  if (!c) return MONTE_ERR_INVALID;
  int rc = check_geom(c, g, false);
  if (rc) return rc;
  if (!filt || !proj || batch < 1) return fail(c, MONTE_ERR_INVALID, "null buffer or batch < 1");
  MONTE_HIP(c, hipSetDevice(c->device));
  const size_t np = proj_elems(*g);
  float* pad = (float*)workspace(c, kSlotAdj, padded_elems(*g) * sizeof(float));
  if (!pad) return fail(c, MONTE_ERR_NOMEM, "adjoint stack allocation failed");
  Stage st{c};
  const float* din = (const float*)st.in(kSlotStageA, filt, np * batch * sizeof(float));
  float* dout = (float*)st.out(kSlotStageB, proj, np * batch * sizeof(float), false);
  if (st.code) return st.code;
  const FilterParams fp = compute_filter_params(*g);
  for (int b = 0; b < batch; ++b) {
    MONTE_HIP(c, launch_pad_copy(din + b * np, pad, g->n_views, g->nu, g->nv, fp.pitch, fp.rows, c->stream));
    if ((rc = run_filter_T_padded(c, *g, pad, dout + b * np))) return rc;
  }
  return st.back(proj, dout, np * batch * sizeof(float));
}

int monte_fdk_adjoint(monte_ctx* c, const monte_geom* g, const float* dvol, float* dproj, int batch) {
  if (!c) return MONTE_ERR_INVALID;
  int rc = check_geom(c, g, true);
  if (rc) return rc;
  if (!dvol || !dproj || batch < 1) return fail(c, MONTE_ERR_INVALID, "null buffer or batch < 1");
  MONTE_HIP(c, hipSetDevice(c->device));
  const size_t np = proj_elems(*g), nvx = vol_elems(*g);
  float* pad = (float*)workspace(c, kSlotAdj, padded_elems(*g) * sizeof(float));
  if (!pad) return fail(c, MONTE_ERR_NOMEM, "adjoint stack allocation failed");
  Stage st{c};
  const float* din = (const float*)st.in(kSlotStageA, dvol, nvx * batch * sizeof(float));
  float* dout = (float*)st.out(kSlotStageB, dproj, np * batch * sizeof(float), false);
  if (st.code) return st.code;
  for (int b = 0; b < batch; ++b) {
    if ((rc = run_project_padded(c, *g, din + b * nvx, pad))) return rc;
    if ((rc = run_filter_T_padded(c, *g, pad, dout + b * np))) return rc;
  }
  return st.back(dproj, dout, np * batch * sizeof(float));
}

// ---------------------------------------------------------------------- MC
int monte_mc_batch(monte_ctx* c, const monte_mc_job* jobs, int n_jobs, monte_mc_stats* stats) {
  if (!c) return MONTE_ERR_INVALID;
  if (!jobs || n_jobs < 1) return fail(c, MONTE_ERR_INVALID, "no jobs");
  if (c->host_ptrs) return fail(c, MONTE_ERR_UNSUPPORTED, "monte_mc_batch takes device pointers only");
  MONTE_HIP(c, hipSetDevice(c->device));
  const int batch = 64 * c->tuning.mc_batch;
  std::vector<HostJob> hj(n_jobs);
  for (int i = 0; i < n_jobs; ++i) {
    int rc = build_job(c, jobs[i].geom, jobs[i].phantom, jobs[i].cfg, hj[i], batch);
    if (rc) return fail(c, rc, "job " + std::to_string(i) + ": " + c->err);
    if (!jobs[i].primary || !jobs[i].scatter) return fail(c, MONTE_ERR_INVALID, "job tallies are null");
    hj[i].d.primary = jobs[i].primary;
    hj[i].d.scatter = jobs[i].scatter;
  }
  // A launch ends with a tail in which the last units run on a few waves, so a
  // launch holding fewer than 32 units of `batch` photons per wave of the grid
  // halves its units, down to 16 cycles of a tile.  The tallies do not depend on
  // the unit size.  64 uneven jobs (4.6e8 histories): 12.7 ms with units of 64
  // cycles, 12.2-12.4 with 32, 12.4-12.5 with 16 and 13.5 with 8.
  {
    uint64_t hist = 0;
    for (const HostJob& h : hj) hist += (uint64_t)h.d.photons_per_view * (uint64_t)h.d.n_views_sim;
    const uint64_t waves =
        (uint64_t)std::max(1, c->n_cu * c->tuning.mc_blocks_per_cu) * (uint64_t)(c->tuning.mc_block / 64);
    int b = batch;
    while (b > 64 * 16 && hist < (uint64_t)b * 32 * waves) b /= 2;
    if (b != batch)
      for (HostJob& h : hj) set_units(h.d, b);
  }
  float ms = 0.f;
  std::vector<McCounters> cnt;
  int rc = run_mc(c, hj, c->tuning.mc_block, c->tuning.mc_blocks_per_cu, &ms, &cnt);
  if (rc) return rc;
  if (stats)
    for (int i = 0; i < n_jobs; ++i) fill_stats(&stats[i], cnt[i], ms);
  return MONTE_OK;
}

int monte_mc_project(monte_ctx* c, const monte_geom* g, const monte_phantom* ph, const monte_mc_cfg* cfg,
                     double* primary, double* scatter, monte_mc_stats* stats) {
  if (!c) return MONTE_ERR_INVALID;
  if (!g || !ph || !cfg || !primary || !scatter) return fail(c, MONTE_ERR_INVALID, "null argument");
  MONTE_HIP(c, hipSetDevice(c->device));
  const size_t np = proj_elems(*g);
  Stage st{c};
  monte_phantom dph = *ph;
  dph.labels = (const uint8_t*)st.in(kSlotStageA, ph->labels, (size_t)ph->nx * ph->ny * ph->nz);
  double* dp = (double*)st.out(kSlotStageB, primary, np * sizeof(double), true);
  double* ds = (double*)st.out(kSlotTuneA, scatter, np * sizeof(double), true);
  if (st.code) return st.code;
  const bool host = c->host_ptrs;
  c->host_ptrs = false;
  monte_mc_job job{g, &dph, cfg, dp, ds};
  int rc = monte_mc_batch(c, &job, 1, stats);
  c->host_ptrs = host;
  if (rc) return rc;
  if ((rc = st.back(primary, dp, np * sizeof(double)))) return rc;
  return st.back(scatter, ds, np * sizeof(double));
}

int monte_mc_project_noalloc(monte_ctx* c, const monte_geom* g, const monte_phantom* ph, const monte_mc_cfg* cfg,
                             double* primary, double* scatter, monte_mc_stats* stats) {
  if (!c) return MONTE_ERR_INVALID;
  if (!g || !ph || !cfg || !primary || !scatter) return fail(c, MONTE_ERR_INVALID, "null argument");
  if (c->host_ptrs) return fail(c, MONTE_ERR_UNSUPPORTED, "monte_mc_project_noalloc takes device pointers only");
  MONTE_HIP(c, hipSetDevice(c->device));
  char* scratch = (char*)noalloc_scratch(c->device);
  if (!scratch) return fail(c, MONTE_ERR_STATE, "no scratch on this device (context not created on it)");
  const int batch = 64 * c->tuning.mc_batch;
  std::vector<HostJob> hj(1);
  int rc = build_job(c, g, ph, cfg, hj[0], batch);
  if (rc) return rc;
  const size_t need = mc_layout(hj).total;
  if (need > kNoAllocScratchBytes)
    return fail(c, MONTE_ERR_INVALID, "the job needs " + std::to_string(need) + " bytes of scratch, more than the " +
                                          std::to_string(kNoAllocScratchBytes) + " available (fewer views)");
  hj[0].d.primary = primary;
  hj[0].d.scatter = scatter;
  float ms = 0.f;
  std::vector<McCounters> cnt;
  if ((rc = run_mc(c, hj, c->tuning.mc_block, c->tuning.mc_blocks_per_cu, &ms, &cnt, scratch, -1))) return rc;
  if (stats) fill_stats(stats, cnt[0], ms);
  return MONTE_OK;
}

int monte_mc_first_flight_sets(monte_ctx* c, const monte_geom* g, const monte_phantom* ph, int brick,
                               uint8_t* sets) {
  if (!c) return MONTE_ERR_INVALID;
  if (!g || !ph || !sets) return fail(c, MONTE_ERR_INVALID, "null argument");
  if (brick < 1 || brick > 64) return fail(c, MONTE_ERR_INVALID, "brick must be in [1, 64]");
  if (c->host_ptrs) return fail(c, MONTE_ERR_UNSUPPORTED, "monte_mc_first_flight_sets takes device pointers only");
  MONTE_HIP(c, hipSetDevice(c->device));
  monte_mc_cfg cfg{};
  cfg.photons_per_view = 1;
  cfg.e_kev = 60.0;
  std::vector<HostJob> hj(1);
  int rc = build_job(c, g, ph, &cfg, hj[0], 64);
  if (rc) return rc;
  // the flight box as the transport computes it, then the sets with `brick`
  const monte_tuning saved = c->tuning;
  c->tuning.mc_ray_brick = brick;
  McPrepared P;
  rc = prepare_mc(c, hj, saved.mc_block, saved.mc_blocks_per_cu, nullptr, P, kSlotRay);
  c->tuning = saved;
  if (rc) return rc;
  MONTE_HIP(c, hipMemcpyAsync(sets, hj[0].d.ray_set, (size_t)g->n_views * g->nu * g->nv, hipMemcpyDeviceToDevice,
                              c->stream));
  MONTE_HIP(c, hipStreamSynchronize(c->stream));
  return MONTE_OK;
}

int monte_primary_exact(monte_ctx* c, const monte_geom* g, const monte_phantom* ph, double e_kev,
                        const double* spectrum, float* lineint) {
  if (!c) return MONTE_ERR_INVALID;
  if (!g || !ph || !lineint) return fail(c, MONTE_ERR_INVALID, "null argument");
  MONTE_HIP(c, hipSetDevice(c->device));
  monte_mc_cfg cfg{};
  cfg.photons_per_view = 1;
  cfg.e_kev = spectrum ? 60.0 : e_kev;
  cfg.spectrum = nullptr;
  HostJob hj;
  Stage st{c};
  monte_phantom dph = *ph;
  dph.labels = (const uint8_t*)st.in(kSlotStageA, ph->labels, (size_t)ph->nx * ph->ny * ph->nz);
  if (st.code) return st.code;
  int rc = build_job(c, g, &dph, &cfg, hj, 64);
  if (rc) return rc;
  const size_t np = proj_elems(*g);
  float* dout = (float*)st.out(kSlotStageB, lineint, np * sizeof(float), false);
  if (st.code) return st.code;
  // tables, FP64 view table and spectrum weights in the MC slot
  size_t off = 0;
  const size_t tab_off = off; off = align_up(off + sizeof(McTables), 256);
  const size_t cs_off = off; off = align_up(off + sizeof(double) * 2 * g->n_views, 256);
  const size_t w_off = off; off = align_up(off + sizeof(float) * kRows, 256);
  char* base = (char*)workspace(c, kSlotMc, off);
  if (!base) return fail(c, MONTE_ERR_NOMEM, "primary descriptor allocation failed");
  std::vector<float> w(kRows, 0.f);
  if (spectrum) {
    double tot = 0;
    for (int e = 1; e < kRows; ++e) {
      if (!(spectrum[e] >= 0)) return fail(c, MONTE_ERR_INVALID, "negative spectrum weight");
      tot += spectrum[e];
    }
    if (!(tot > 0)) return fail(c, MONTE_ERR_INVALID, "empty spectrum");
    for (int e = 1; e < kRows; ++e) w[e] = (float)(spectrum[e] / tot);
  }
  std::vector<double> cs(2 * g->n_views);
  for (int k = 0; k < g->n_views; ++k) {
    const double b = g->preset == MONTE_PRESET_BP3D20 ? 3.14159265358979323846 * (k * g->step_deg) / 180.0
                                                      : g->angle0 + k * g->dangle;
    cs[2 * k] = std::cos(b);
    cs[2 * k + 1] = std::sin(b);
  }
  PrimaryGeom P{};
  P.sod = g->sod; P.sdd = g->sdd;
  if (g->preset == MONTE_PRESET_BP3D20) {
    P.un0 = 0.5 * g->nu; P.u_step = -g->du; P.vn0 = 0.5 * g->nv; P.v_step = -g->dv;
  } else {
    P.un0 = 0.5 * g->nu - g->u_off / g->du; P.u_step = g->du;
    P.vn0 = 0.5 * g->nv - g->v_off / g->dv; P.v_step = g->dv;
  }
  const double vox[3] = {ph->vx, ph->vy, ph->vz}, org[3] = {ph->ox, ph->oy, ph->oz};
  const int n[3] = {ph->nx, ph->ny, ph->nz};
  for (int a = 0; a < 3; ++a) {
    P.vox[a] = vox[a];
    P.n[a] = n[a];
    P.b0[a] = org[a] - 0.5 * n[a] * vox[a];
    P.b1[a] = org[a] + 0.5 * n[a] * vox[a];
  }
  P.nu = g->nu; P.nv = g->nv; P.n_views = g->n_views;
  MONTE_HIP(c, hipMemcpyAsync(base + tab_off, &hj.tables, sizeof(McTables), hipMemcpyHostToDevice, c->stream));
  MONTE_HIP(c, hipMemcpyAsync(base + cs_off, cs.data(), sizeof(double) * cs.size(), hipMemcpyHostToDevice, c->stream));
  MONTE_HIP(c, hipMemcpyAsync(base + w_off, w.data(), sizeof(float) * kRows, hipMemcpyHostToDevice, c->stream));
  MONTE_HIP(c, launch_primary_exact(P, (const double*)(base + cs_off), dph.labels, (const McTables*)(base + tab_off),
                                    (float)cfg.e_kev, (const float*)(base + w_off), spectrum ? 1 : 0, dout,
                                    c->stream));
  MONTE_HIP(c, hipStreamSynchronize(c->stream));  // host tables above must outlive the copies
  return st.back(lineint, dout, np * sizeof(float));
}

int monte_counts_to_lineint(monte_ctx* c, const monte_geom* g, const double* counts, uint64_t photons_per_view,
                            float* lineint) {
  if (!c) return MONTE_ERR_INVALID;
  int rc = check_geom(c, g, false);
  if (rc) return rc;
  if (!counts || !lineint || photons_per_view == 0) return fail(c, MONTE_ERR_INVALID, "null buffer or no photons");
  MONTE_HIP(c, hipSetDevice(c->device));
  const size_t np = proj_elems(*g);
  Stage st{c};
  const double* din = (const double*)st.in(kSlotStageA, counts, np * sizeof(double));
  float* dout = (float*)st.out(kSlotStageB, lineint, np * sizeof(float), false);
  if (st.code) return st.code;
  MONTE_HIP(c, launch_counts_to_lineint(din, nullptr, dout, g->n_views, g->nu, g->nv, photons_per_view, c->stream));
  return st.back(lineint, dout, np * sizeof(float));
}

int monte_counts_to_lineint_flat(monte_ctx* c, const monte_geom* g, const double* counts, const double* flat,
                                 int flat_views, float* lineint) {
  if (!c) return MONTE_ERR_INVALID;
  int rc = check_geom(c, g, false);
  if (rc) return rc;
  if (!counts || !flat || !lineint) return fail(c, MONTE_ERR_INVALID, "null buffer");
  if (flat_views != 1 && flat_views != g->n_views)
    return fail(c, MONTE_ERR_INVALID, "flat_views must be 1 or the number of views");
  MONTE_HIP(c, hipSetDevice(c->device));
  const size_t np = proj_elems(*g), nf = (size_t)flat_views * g->nu * g->nv;
  Stage st{c};
  const double* din = (const double*)st.in(kSlotStageA, counts, np * sizeof(double));
  const double* dflat = (const double*)st.in(kSlotTuneA, flat, nf * sizeof(double));
  float* dout = (float*)st.out(kSlotStageB, lineint, np * sizeof(float), false);
  if (st.code) return st.code;
  MONTE_HIP(c, launch_counts_to_lineint_flat(din, dflat, flat_views, dout, g->n_views, g->nu, g->nv, c->stream));
  return st.back(lineint, dout, np * sizeof(float));
}

// ----------------------------------------------------------------- phantom
int monte_phantom_paint(uint8_t* labels, int nx, int ny, int nz, double vx, double vy, double vz,
                        const monte_primitive* prims, int n_prims) {
  if (!labels || nx < 1 || ny < 1 || nz < 1 || !(vx > 0) || !(vy > 0) || !(vz > 0) || n_prims < 0 ||
      (n_prims > 0 && !prims))
    return MONTE_ERR_INVALID;
  for (int i = 0; i < n_prims; ++i)
    if (prims[i].label < 0 || prims[i].label > 255 || prims[i].shape < 0 || prims[i].shape > 3 ||
        !(prims[i].rx > 0) || !(prims[i].ry > 0) || !(prims[i].rz > 0))
      return MONTE_ERR_INVALID;
  unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  auto work = [&](int z0, int z1) {
    for (int z = z0; z < z1; ++z) {
      const double Z = (z - 0.5 * (nz - 1)) * vz;
      for (int y = 0; y < ny; ++y) {
        const double Y = (y - 0.5 * (ny - 1)) * vy;
        uint8_t* row = labels + ((size_t)z * ny + y) * nx;
        for (int x = 0; x < nx; ++x) {
          const double X = (x - 0.5 * (nx - 1)) * vx;
          for (int i = 0; i < n_prims; ++i) {
            const monte_primitive& p = prims[i];
            const double a = (X - p.cx) / p.rx, b = (Y - p.cy) / p.ry, cc = (Z - p.cz) / p.rz;
            bool in;
            if (p.shape == MONTE_SHAPE_ELLIPSOID) in = a * a + b * b + cc * cc <= 1.0;
            else if (p.shape == MONTE_SHAPE_CYLINDER_Z) in = a * a + b * b <= 1.0 && std::fabs(cc) <= 1.0;
            else if (p.shape == MONTE_SHAPE_CYLINDER_X) in = b * b + cc * cc <= 1.0 && std::fabs(a) <= 1.0;
            else in = std::fabs(a) <= 1.0 && std::fabs(b) <= 1.0 && std::fabs(cc) <= 1.0;
            if (in) row[x] = (uint8_t)p.label;
          }
        }
      }
    }
  };
  std::vector<std::thread> th;
  const int per = (nz + (int)nt - 1) / (int)nt;
  for (unsigned t = 0; t < nt; ++t) {
    const int z0 = (int)t * per, z1 = std::min(nz, z0 + per);
    if (z0 < z1) th.emplace_back(work, z0, z1);
  }
  for (auto& t : th) t.join();
  return MONTE_OK;
}

// ------------------------------------------------------------------ tuning
int monte_autotune(monte_ctx* c, int what, const monte_geom* g, const monte_phantom* ph, const monte_mc_cfg* cfg,
                   monte_tuning* out) {
  if (!c) return MONTE_ERR_INVALID;
  if (c->host_ptrs) return fail(c, MONTE_ERR_UNSUPPORTED, "autotune takes device pointers only");
  MONTE_HIP(c, hipSetDevice(c->device));
  monte_tuning best = c->tuning;
  if (what & 1) {
    if (!g || !ph || !cfg) return fail(c, MONTE_ERR_INVALID, "MC tuning needs geometry, phantom and config");
    // a run over all the job's views (units run views fastest, so the view
    // count shapes the memory access pattern) with 64 photon cycles per view
    // where the job has them (so the units are tiles over 64 cycles, as in a
    // long launch; 3 cycles of 2^28 photons over 360 views made units of 21
    // tiles, and a slower choice) and at least 2^28 photons in all, into
    // scratch tallies: the runs timed here must be long enough that the last
    // units' tail hardly counts (at 2^23 photons, launches of one wave per SIMD
    // won and were 2.5x slower on the full workload)
    monte_mc_cfg tc = *cfg;
    int vb = cfg->view_begin, ve = cfg->view_end;
    if (vb == 0 && ve == 0) ve = g->n_views;
    const uint64_t nvw = (uint64_t)std::max(ve - vb, 1), npix = (uint64_t)g->nu * (uint64_t)g->nv;
    const uint64_t per_view = std::max<uint64_t>(64 * npix, (1ull << 28) / nvw);
    tc.photons_per_view = std::min<uint64_t>(cfg->photons_per_view, per_view);
    tc.view_begin = vb;
    tc.view_end = ve;
    const size_t np = proj_elems(*g);
    double* tally = (double*)workspace(c, kSlotTuneA, 2 * np * sizeof(double));
    if (!tally) return fail(c, MONTE_ERR_NOMEM, "tuning tallies");
    float best_ms = 1e30f;
    const int blocks[] = {64, 128, 256};
    const int batches[] = {32, 64, 128};
    const int waves_per_cu[] = {16, 20, 24};  // 24 is the kernel's occupancy (6 waves per SIMD)
    for (int blk : blocks)
      for (int bt : batches)
        for (int wpc : waves_per_cu) {
          const int bpc = wpc / (blk / 64);
          if (bpc < 1 || blk * bpc > 2048) continue;
          HostJob hj;
          int rc = build_job(c, g, ph, &tc, hj, 64 * bt);
          if (rc) return rc;
          hj.d.primary = tally;
          hj.d.scatter = tally + np;
          std::vector<HostJob> v(1, hj);
          float ms = 0.f;
          if ((rc = run_mc(c, v, blk, bpc, &ms, nullptr))) return rc;
          if (ms < best_ms) {
            best_ms = ms;
            best.mc_block = blk;
            best.mc_batch = bt;
            best.mc_blocks_per_cu = bpc;
          }
        }
  }
  if (what & 2) {
    int rc = check_geom(c, g, true);
    if (rc) return rc;
    // time a slice of views on scratch buffers
    monte_geom tg = *g;
    const float* taps;
    const BpView* views;
    if ((rc = ensure_constants(c, tg, &taps, &views))) return rc;
    const int nviews = std::min(g->n_views, 360);  // L2 reuse across views shapes the best tile
    float4* quads = (float4*)workspace(c, kSlotQuad, padded_elems(tg) * sizeof(float4));
    float* vol = (float*)workspace(c, kSlotTuneB, vol_elems(tg) * sizeof(float));
    if (!quads || !vol) return fail(c, MONTE_ERR_NOMEM, "tuning buffers");
    MONTE_HIP(c, launch_fill((float*)quads, 0.f, 4 * padded_elems(tg), c->stream));
    const BpGrid grid = compute_bp_grid(tg);
    hipEvent_t e0, e1;
    MONTE_HIP(c, hipEventCreate(&e0));
    MONTE_HIP(c, hipEventCreate(&e1));
    float best_ms = 1e30f;
    const int zcs[] = {4, 8, 16};
    const int txs[] = {8, 16, 32, 64};
    for (int zc : zcs)
      for (int tx : txs) {
        monte_tuning t = c->tuning;
        t.bp_zchunk = zc;
        t.bp_tile_x = tx;
        t.bp_tile_y = 256 / tx;
        MONTE_HIP(c, launch_backproject(quads, views, 0, nviews, grid, vol, 0, t, c->stream));  // warm
        MONTE_HIP(c, hipEventRecord(e0, c->stream));
        MONTE_HIP(c, launch_backproject(quads, views, 0, nviews, grid, vol, 0, t, c->stream));
        MONTE_HIP(c, hipEventRecord(e1, c->stream));
        MONTE_HIP(c, hipEventSynchronize(e1));
        float ms = 0.f;
        MONTE_HIP(c, hipEventElapsedTime(&ms, e0, e1));
        if (ms < best_ms) {
          best_ms = ms;
          best.bp_zchunk = zc;
          best.bp_tile_x = tx;
          best.bp_tile_y = 256 / tx;
        }
      }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
  }
  if (what & 4) {
    // the matched projector: voxels per lane along z (the automatic choice, 0,
    // and each of 2, 4, 8) on a scratch volume, every view, after a warm-up
    int rc = check_geom(c, g, true);
    if (rc) return rc;
    const size_t np = padded_elems(*g);
    float* vol = (float*)workspace(c, kSlotTuneB, vol_elems(*g) * sizeof(float));
    float* pad = (float*)workspace(c, kSlotAdj, np * sizeof(float));
    if (!vol || !pad) return fail(c, MONTE_ERR_NOMEM, "tuning buffers");
    MONTE_HIP(c, launch_fill(vol, 1.f, vol_elems(*g), c->stream));
    hipEvent_t e0, e1;
    MONTE_HIP(c, hipEventCreate(&e0));
    MONTE_HIP(c, hipEventCreate(&e1));
    const monte_tuning keep = c->tuning;
    float best_ms = 1e30f;
    for (int zc : {0, 2, 4, 8}) {
      c->tuning = keep;
      c->tuning.pj_zchunk = zc;
      rc = run_project_padded(c, *g, vol, pad);  // warm (builds the column lists once)
      if (!rc) {
        (void)hipEventRecord(e0, c->stream);
        rc = run_project_padded(c, *g, vol, pad);
        (void)hipEventRecord(e1, c->stream);
      }
      if (rc) break;
      MONTE_HIP(c, hipEventSynchronize(e1));
      float ms = 0.f;
      MONTE_HIP(c, hipEventElapsedTime(&ms, e0, e1));
      if (ms < best_ms) {
        best_ms = ms;
        best.pj_zchunk = zc;
      }
    }
    c->tuning = keep;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc;
  }
  c->tuning = best;
  if (out) *out = best;
  return MONTE_OK;
}

// ------------------------------------------------------------------ graphs
int monte_fdk_graph(monte_ctx* c, const monte_geom* g, const float* proj, float* vol, int batch) {
  if (!c) return MONTE_ERR_INVALID;
  int rc = check_geom(c, g, true);
  if (rc) return rc;
  if (!proj || !vol || batch < 1) return fail(c, MONTE_ERR_INVALID, "null buffer or batch < 1");
  if (c->host_ptrs) return fail(c, MONTE_ERR_UNSUPPORTED, "graphs take device pointers only");
  if (c->stream == nullptr) return fail(c, MONTE_ERR_INVALID, "graph capture needs a created stream, not the null stream");
  MONTE_HIP(c, hipSetDevice(c->device));
  std::vector<double> key = geom_key(*g);
  key.push_back((double)(uintptr_t)proj);
  key.push_back((double)(uintptr_t)vol);
  key.push_back((double)batch);
  const monte_tuning& t = c->tuning;
  for (double v : {(double)t.bp_zchunk, (double)t.bp_tile_x, (double)t.bp_view_chunk}) key.push_back(v);
  // everything the captured launches touch must exist before capture
  const float* taps;
  const BpView* views;
  if ((rc = ensure_constants(c, *g, &taps, &views))) return rc;
  float* pad = (float*)workspace(c, kSlotFwd, padded_elems(*g) * sizeof(float));
  if (!pad) return fail(c, MONTE_ERR_NOMEM, "filtered stack allocation failed");
  if (!workspace(c, kSlotQuad, padded_elems(*g) * sizeof(float4))) return fail(c, MONTE_ERR_NOMEM, "tap quad allocation failed");
  key.push_back((double)(uintptr_t)pad);
  key.push_back((double)(uintptr_t)c->ws[kSlotQuad]);
  key.push_back((double)(uintptr_t)c->ws[kSlotConst]);
  if (key != c->graph_key || !c->graph_exec) {
    if (c->graph_exec) { (void)hipGraphExecDestroy(c->graph_exec); c->graph_exec = nullptr; }
    if (c->graph) { (void)hipGraphDestroy(c->graph); c->graph = nullptr; }
    c->graph_key.clear();
    MONTE_HIP(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    const size_t np = proj_elems(*g), nvx = vol_elems(*g);
    int crc = MONTE_OK;
    for (int b = 0; b < batch && !crc; ++b) {
      crc = run_filter_padded(c, *g, proj + b * np, pad);
      if (!crc) crc = run_backproject_padded(c, *g, pad, vol + b * nvx, 0);
    }
    hipGraph_t graph = nullptr;
    hipError_t ce = hipStreamEndCapture(c->stream, &graph);
    if (crc) { if (graph) (void)hipGraphDestroy(graph); return crc; }
    if (ce != hipSuccess) return fail(c, MONTE_ERR_HIP, std::string("end capture: ") + hipGetErrorString(ce));
    c->graph = graph;
    MONTE_HIP(c, hipGraphInstantiate(&c->graph_exec, c->graph, nullptr, nullptr, 0));
    c->graph_key = key;
  }
  MONTE_HIP(c, hipGraphLaunch(c->graph_exec, c->stream));
  return MONTE_OK;
}

// ------------------------------------------------------------ end to end
int monte_e2e_step(monte_ctx* c, const monte_geom* g, const monte_phantom* ph, const monte_mc_cfg* cfg,
                   const monte_e2e_buffers* buf, int use_graph, monte_mc_stats* stats) {
  if (!c) return MONTE_ERR_INVALID;
  int rc = check_geom(c, g, true);
  if (rc) return rc;
  if (!ph || !cfg || !buf) return fail(c, MONTE_ERR_INVALID, "null argument");
  const monte_e2e_buffers& b = *buf;
  if (!b.primary || !b.scatter || !b.lineint || !b.vol || !b.vol_ref || !b.residual || !b.grad || !b.loss)
    return fail(c, MONTE_ERR_INVALID, "null buffer");
  if (c->host_ptrs) return fail(c, MONTE_ERR_UNSUPPORTED, "monte_e2e_step takes device pointers only");
  if (use_graph && c->stream == nullptr)
    return fail(c, MONTE_ERR_INVALID, "graph capture needs a created stream, not the null stream");
  if (!((cfg->view_begin == 0 && (cfg->view_end == 0 || cfg->view_end == g->n_views))))
    return fail(c, MONTE_ERR_INVALID, "the end-to-end step simulates every view of the scan");
  MONTE_HIP(c, hipSetDevice(c->device));
  const int batch = 64 * c->tuning.mc_batch;

  // MC descriptors, prepared once per job in their own slot
  std::vector<double> mkey = mc_job_key(c, *g, *ph, *cfg);
  mkey.push_back((double)(uintptr_t)b.primary);  // the descriptors hold the tally pointers
  mkey.push_back((double)(uintptr_t)b.scatter);
  if (mkey != c->e2e_mc_key || !c->e2e_mc.base) {
    c->e2e_mc_key.clear();
    c->e2e_graph_key.clear();
    std::vector<HostJob> hj(1);
    if ((rc = build_job(c, g, ph, cfg, hj[0], batch))) return rc;
    hj[0].d.primary = b.primary;
    hj[0].d.scatter = b.scatter;
    const size_t need = mc_layout(hj).total + kLossBlocks * sizeof(double);
    char* base = (char*)workspace(c, kSlotE2E, need);
    if (!base) return fail(c, MONTE_ERR_NOMEM, "end-to-end descriptor allocation failed");
    if ((rc = prepare_mc(c, hj, c->tuning.mc_block, c->tuning.mc_blocks_per_cu, base, c->e2e_mc, kSlotE2ERay)))
      return rc;
    c->e2e_mc_key = mkey;
  }
  double* partial = (double*)(c->e2e_mc.base + c->e2e_mc.L.total);

  // every workspace the launches touch exists at its size before any capture
  const float* taps;
  const BpView* views;
  if ((rc = ensure_constants(c, *g, &taps, &views))) return rc;
  const BpGrid grid = compute_bp_grid(*g);
  const uint32_t *plan_offs, *plan_list;
  if ((rc = ensure_projector_plan(c, *g, views, grid, &plan_offs, &plan_list))) return rc;
  if (!workspace(c, kSlotFwd, padded_elems(*g) * sizeof(float)) ||
      !workspace(c, kSlotQuad, padded_elems(*g) * sizeof(float4)) ||
      !workspace(c, kSlotAdj, padded_elems(*g) * sizeof(float)) ||
      !workspace(c, kSlotVolT, vol_elems(*g) * sizeof(float)))
    return fail(c, MONTE_ERR_NOMEM, "end-to-end workspace allocation failed");

  MONTE_HIP(c, hipEventRecord(c->ev0, c->stream));
  if (!use_graph) {
    if ((rc = e2e_launches(c, *g, b, cfg->photons_per_view, partial))) return rc;
  } else {
    std::vector<double> key = mkey;
    const void* ptrs[] = {b.primary, b.scatter, b.lineint, b.vol, b.vol_ref, b.residual, b.grad, b.loss};
    for (const void* p : ptrs) key.push_back((double)(uintptr_t)p);
    for (int slot : {kSlotFwd, kSlotQuad, kSlotAdj, kSlotVolT, kSlotPlan, kSlotConst, kSlotE2E, kSlotE2ERay})
      key.push_back((double)(uintptr_t)c->ws[slot]);
    const monte_tuning& t = c->tuning;
    for (double v : {(double)t.bp_zchunk, (double)t.bp_tile_x, (double)t.bp_view_chunk, (double)t.pj_zchunk})
      key.push_back(v);
    if (key != c->e2e_graph_key || !c->e2e_exec) {
      if (c->e2e_exec) { (void)hipGraphExecDestroy(c->e2e_exec); c->e2e_exec = nullptr; }
      if (c->e2e_graph) { (void)hipGraphDestroy(c->e2e_graph); c->e2e_graph = nullptr; }
      c->e2e_graph_key.clear();
      MONTE_HIP(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
      const int crc = e2e_launches(c, *g, b, cfg->photons_per_view, partial);
      hipGraph_t graph = nullptr;
      const hipError_t ce = hipStreamEndCapture(c->stream, &graph);
      if (crc) { if (graph) (void)hipGraphDestroy(graph); return crc; }
      if (ce != hipSuccess) return fail(c, MONTE_ERR_HIP, std::string("end capture: ") + hipGetErrorString(ce));
      c->e2e_graph = graph;
      MONTE_HIP(c, hipGraphInstantiate(&c->e2e_exec, c->e2e_graph, nullptr, nullptr, 0));
      c->e2e_graph_key = key;
    }
    MONTE_HIP(c, hipGraphLaunch(c->e2e_exec, c->stream));
  }
  MONTE_HIP(c, hipEventRecord(c->ev1, c->stream));
  if (stats) {
    std::vector<McCounters> cnt;
    if ((rc = read_counters(c, c->e2e_mc, cnt))) return rc;
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, c->ev0, c->ev1);
    fill_stats(stats, cnt[0], ms);
  }
  return MONTE_OK;
}

}  // extern "C"
