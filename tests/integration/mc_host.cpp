// The MC call of INTEGRATION.md as a program: how monte_cu/CBCT_real325.cu
// would call libmonte with host arrays in place of its RandStateGenerator +
// projection<<<>>> launch (:241-250), in the reference's geometry (SOD 160,
// SDD 220, 325 x 325 detector at 0.1 cm, index -(int)(10 d - 162.5)).
//
//   mc_host <dir with xcom2.csv, Ca.csv> <labels.raw, u8 [n][n][n] at 0.1 cm> <n>
//           <views> <step_deg> <photons per pixel> <out prefix>
// writes <prefix>primary.raw and <prefix>scatter.raw (FP64 [views][325][325])
// and prints the statistics as one line.
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "monte.h"

static int run(const std::string& dir, const uint8_t* geometry, int n, int n_views, int step_deg, uint64_t per,
               const std::string& prefix) {
  // [doc:mc]
  monte_ctx* ctx = nullptr;
  if (monte_create(&ctx, 0) != MONTE_OK) { fprintf(stderr, "no device\n"); return 1; }
  monte_set_pointer_mode(ctx, 1);                        // host arrays below

  double tab[2][4][MONTE_TABLE_ROWS];
  monte_read_xcom_csv((dir + "/xcom2.csv").c_str(), &tab[0][0][0]);  // water
  monte_read_xcom_csv((dir + "/Ca.csv").c_str(), &tab[1][0][0]);     // calcium
  monte_load_tables(ctx, &tab[0][0][0], 2);

  monte_geom g = {};                                     // bp3d20 preset = the reference's detector mapping
  g.preset = MONTE_PRESET_BP3D20; g.n_views = n_views; g.nu = g.nv = 325;
  g.sod = 160; g.sdd = 220; g.du = g.dv = 0.1; g.step_deg = step_deg;
  monte_phantom ph = {};                                 // label map read from the phantom .raw
  ph.labels = geometry; ph.nx = ph.ny = ph.nz = n; ph.vx = ph.vy = ph.vz = 0.1; ph.n_labels = 3;
  ph.mat_table[0] = -1; ph.mat_table[1] = 0; ph.mat_table[2] = 1;
  ph.density[1] = 1.0;  ph.density[2] = 1.55;
  monte_mc_cfg cfg = {};
  cfg.photons_per_view = per * 325 * 325;                // per photons per pixel
  cfg.e_kev = 140.0; cfg.max_interactions = 5;           // ScatterNUM (:9)
  cfg.coherent_mode = MONTE_COH_FORWARD;                 // undeflected, as :642-681
  cfg.seed = 20261202;

  std::vector<double> primary((size_t)n_views * 325 * 325), scatter(primary.size());
  monte_mc_stats st;
  int rc = monte_mc_project(ctx, &g, &ph, &cfg, primary.data(), scatter.data(), &st);
  if (rc != MONTE_OK) fprintf(stderr, "%s\n", monte_last_error(ctx));
  monte_destroy(ctx);
  // [/doc:mc]
  if (rc != MONTE_OK) return 2;
  for (const char* which : {"primary", "scatter"}) {
    const std::vector<double>& v = which[0] == 'p' ? primary : scatter;
    FILE* f = fopen((prefix + which + ".raw").c_str(), "wb");
    if (!f || fwrite(v.data(), sizeof(double), v.size(), f) != v.size()) { fprintf(stderr, "write failed\n"); return 3; }
    fclose(f);
  }
  printf("histories %llu steps %llu interactions %llu tallied_primary %llu tallied_scatter %llu missed %llu "
         "capped %llu\n", (unsigned long long)st.histories, (unsigned long long)st.steps,
         (unsigned long long)st.interactions, (unsigned long long)st.tallied_primary,
         (unsigned long long)st.tallied_scatter, (unsigned long long)st.missed, (unsigned long long)st.capped);
  return 0;
}

int main(int argc, char** argv) {
  if (argc != 8) {
    fprintf(stderr, "usage: %s data_dir labels.raw n views step_deg per out_prefix\n", argv[0]);
    return 64;
  }
  const int n = atoi(argv[3]);
  std::vector<uint8_t> labels((size_t)n * n * n);
  FILE* f = fopen(argv[2], "rb");
  if (!f || fread(labels.data(), 1, labels.size(), f) != labels.size()) { fprintf(stderr, "read failed\n"); return 3; }
  fclose(f);
  return run(argv[1], labels.data(), n, atoi(argv[4]), atoi(argv[5]), strtoull(argv[6], nullptr, 10), argv[7]);
}
