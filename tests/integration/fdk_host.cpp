// The FDK call of INTEGRATION.md as a program: recon/bp3d20_325.cpp's weight,
// filter and backprojection loops (:39-175) replaced by one monte_fdk call on
// host arrays, keeping the reference's file I/O (input :29-30, image_xy :170).
//
//   fdk_host <map.raw, f32 [360][325][325]> <out image_xy.raw, f32 [256][256][256]>
#include <cstdio>
#include <vector>

#include "monte.h"

int main(int argc, char** argv) {
  if (argc != 3) {
    fprintf(stderr, "usage: %s map.raw image_xy.raw\n", argv[0]);
    return 64;
  }
  std::vector<float> map((size_t)360 * 325 * 325), vol_xy((size_t)256 * 256 * 256);
  FILE* f = fopen(argv[1], "rb");
  if (!f || fread(map.data(), sizeof(float), map.size(), f) != map.size()) { fprintf(stderr, "read failed\n"); return 3; }
  fclose(f);
  monte_ctx* ctx = nullptr;
  if (monte_create(&ctx, 0) != MONTE_OK) { fprintf(stderr, "no device\n"); return 1; }
  // [doc:fdk]
  monte_geom g = {};
  g.preset = MONTE_PRESET_BP3D20; g.n_views = 360; g.nu = g.nv = 325;
  g.nx = g.ny = g.nz = 256; g.vx = g.vy = g.vz = 0.1;
  g.sod = 160; g.sdd = 220; g.du = g.dv = 0.1; g.odd = 60; g.scale = 2.7 * 5; g.step_deg = 1;
  g.roi_x0 = 125; g.roi_x1 = 130; g.mask_r = -1;        // :104 x-column range
  monte_set_pointer_mode(ctx, 1);
  int rc = monte_fdk(ctx, &g, map.data() /* [360][325][325] from :29-30 */, vol_xy.data() /* [256][256][256] */, 1);
  // [/doc:fdk]
  if (rc != MONTE_OK) {
    fprintf(stderr, "%s\n", monte_last_error(ctx));
    monte_destroy(ctx);
    return 2;
  }
  monte_destroy(ctx);
  f = fopen(argv[2], "wb");
  if (!f || fwrite(vol_xy.data(), sizeof(float), vol_xy.size(), f) != vol_xy.size()) {
    fprintf(stderr, "write failed\n");
    return 3;
  }
  fclose(f);
  return 0;
}
