// Host-side geometry: per-view backprojection coefficients and filter
// constants for the two presets, computed in FP64.  The projective map of each
// view stays FP64 (the kernels form each column's U, 1/U and sample
// coordinates from it in FP64); the distance-weight terms and the filter
// constants are handed over in FP32.  One set of kernels serves both presets;
// the presets differ only in these numbers (see BpView in internal.h).
#include <algorithm>
#include <cmath>

#include "internal.h"

namespace monte {

namespace {
constexpr double kPi = 3.14159265358979323846;

double view_angle(const monte_geom& g, int k) {
  if (g.preset == MONTE_PRESET_BP3D20) return kPi * (double)(k * g.step_deg) / 180.0;
  return g.angle0 + k * g.dangle;
}
}  // namespace

void compute_bp_views(const monte_geom& g, std::vector<BpView>& out) {
  out.assign(g.n_views, BpView{});
  for (int k = 0; k < g.n_views; ++k) {
    const double b = view_angle(g, k), cb = std::cos(b), sb = std::sin(b);
    BpView& v = out[k];
    // U = X cos b + Y sin b + sod: distance from the source plane
    // (recon/bp3d20_325.cpp:111-118 rotates P - S by -beta and divides by it).
    v.cx = cb;
    v.cy = sb;
    v.c0 = g.sod;
    // in-plane numerator y' = -X sin b + Y cos b
    v.ax = -sb;
    v.ay = cb;
    if (g.preset == MONTE_PRESET_BP3D20) {
      // recon/bp3d20_325.cpp:134-135: y = -(w_y/pitch - H/2), w_y = y' sdd / U
      v.su0 = 0.5 * g.nu;
      v.suk = -g.sdd / g.du;
      v.sv0 = 0.5 * g.nv;
      v.svk = -g.sdd / g.dv;
      // :145-153: d = sgn(s') |s' - T t'| / sqrt(1 + T^2), s' = X cos - Y sin,
      // t' = X sin + Y cos, T = tan(beta) with beta the integer degree count
      // taken as radians; weight 60^2 / (60 - d)^2 (:169).  cos and sin here
      // are computed as the reference computes them, cos(M_PI*beta/180).
      const int beta = k * g.step_deg;
      const double cr = std::cos(kPi * beta / 180), sr = std::sin(kPi * beta / 180);
      const double T = std::tan((double)beta);
      const double kd = 1.0 / std::sqrt(1.0 + T * T);
      v.wn = (float)g.odd;
      v.d0 = (float)g.odd;
      v.dx = 0.f;
      v.dy = 0.f;
      v.gx = cr;
      v.gy = -sr;
      v.hx = (float)(kd * (cr - T * sr));
      v.hy = (float)(kd * (-sr - T * cr));
      // :169-170: * beta_span * 2 pi / 360 * 2.7 * 5
      v.fac = (float)(g.step_deg * 2.0 * kPi / 360.0 * g.scale);
    } else {
      // pixel centres: sample index = u_idx - 1/2
      v.su0 = 0.5 * g.nu - g.u_off / g.du - 0.5;
      v.suk = g.sdd / g.du;
      v.sv0 = 0.5 * g.nv - g.v_off / g.dv - 0.5;
      v.svk = g.sdd / g.dv;
      // FDK distance weight (sod / U)^2 and angular step
      v.wn = (float)g.sod;
      v.d0 = (float)g.sod;
      v.dx = (float)cb;
      v.dy = (float)sb;
      v.gx = v.gy = 0.0;
      v.hx = v.hy = 0.f;
      v.fac = (float)g.dangle;
    }
  }
}

BpGrid compute_bp_grid(const monte_geom& g) {
  BpGrid q{};
  q.nx = g.nx; q.ny = g.ny; q.nz = g.nz; q.nu = g.nu; q.nv = g.nv;
  q.pitch = g.nu + kPadLo + kPadHi;
  q.rows = g.nv + kPadLo + kPadHi;
  if (g.preset == MONTE_PRESET_BP3D20) {
    // recon/bp3d20_325.cpp:107-109: X = -12.8 + 0.1 s, Y = 12.8 - 0.1 t, Z = 12.8 - 0.1 z
    q.gx0 = g.ox - 0.5 * g.nx * g.vx; q.gsx = g.vx;
    q.gy0 = g.oy + 0.5 * g.ny * g.vy; q.gsy = -g.vy;
    q.gz0 = g.oz + 0.5 * g.nz * g.vz; q.gsz = -g.vz;
    // :127 and :162: |w| <= H/2 pitch and 0 <= x <= H, inclusive
    q.lo_u = 0.0; q.hi_u = g.nu;
    q.lo_v = 0.0; q.hi_v = g.nv;
    q.x0 = g.roi_x0; q.x1 = g.roi_x1;
    q.mask = g.mask_r >= 0 ? 1 : 0;
    q.mask_r2 = (float)(g.mask_r * g.mask_r);
    q.mcx = g.nx / 2; q.mcy = g.ny / 2; q.mcz = g.nz / 2;
  } else {
    q.gx0 = g.ox - 0.5 * (g.nx - 1) * g.vx; q.gsx = g.vx;
    q.gy0 = g.oy - 0.5 * (g.ny - 1) * g.vy; q.gsy = g.vy;
    q.gz0 = g.oz - 0.5 * (g.nz - 1) * g.vz; q.gsz = g.vz;
    // support of the zero-padded bilinear interpolant
    q.lo_u = -1.0; q.hi_u = g.nu;
    q.lo_v = -1.0; q.hi_v = g.nv;
    q.x0 = 0; q.x1 = g.nx;
    q.mask = 0;
  }
  return q;
}

FilterParams compute_filter_params(const monte_geom& g) {
  FilterParams f{};
  f.nu = g.nu; f.nv = g.nv; f.n_views = g.n_views;
  f.pitch = g.nu + kPadLo + kPadHi;
  f.rows = g.nv + kPadLo + kPadHi;
  if (g.preset == MONTE_PRESET_BP3D20) {
    // recon/bp3d20_325.cpp:43: 60 / sqrt(60^2 + (-0.1 r + 16.25)^2 + (0.1 c - 16.25)^2)
    f.D = (float)g.odd;
    f.a_u = (float)(0.5 * g.nu * g.du); f.b_u = (float)(-g.du);
    f.a_v = (float)(-0.5 * g.nv * g.dv); f.b_v = (float)g.dv;
    f.cs = 0.5f;  // :76
  } else {
    f.D = (float)g.sdd;
    f.a_u = (float)((0.5 - 0.5 * g.nu) * g.du + g.u_off); f.b_u = (float)g.du;
    f.a_v = (float)((0.5 - 0.5 * g.nv) * g.dv + g.v_off); f.b_v = (float)g.dv;
    const double tau = g.du * g.sod / g.sdd;  // detector pitch scaled to the isocentre
    f.cs = (float)(1.0 / (2.0 * tau));
  }
  return f;
}

void compute_filter_taps(const monte_geom& g, std::vector<float>& h) {
  const int n = g.nu;
  h.assign(2 * n - 1, 0.f);
  const bool shepp = g.preset == MONTE_PRESET_STANDARD && g.filter == MONTE_FILTER_SHEPP_LOGAN;
  for (int i = -(n - 1); i <= n - 1; ++i) {
    const int a = i < 0 ? -i : i;
    double v;
    if (shepp) {
      v = -2.0 / (kPi * kPi * (4.0 * a * a - 1.0));
    } else if (a == 0) {
      v = 0.25;  // recon/bp3d20_325.cpp:52
    } else if (a % 2 == 0) {
      v = 0.0;
    } else {
      v = -1.0 / ((a * kPi) * (a * kPi));  // :60-61
    }
    h[i + n - 1] = (float)v;
  }
}

// The v sample coordinate of a voxel is sv = sv0 + svk Z / U with U affine in
// (X, Y) and positive over the volume, so over the voxel-centre box its
// extremes lie at the box's 8 corners.  Rows written are floor(sv) and the one
// above, clamped to the v window; one row of margin each side covers rounding.
void projector_rows(const std::vector<BpView>& views, int view0, int n, const BpGrid& grid, int& lds_r0,
                    int& lds_rows) {
  double lo = 1e300, hi = -1e300;
  const double xs[2] = {grid.gx0 + grid.x0 * grid.gsx, grid.gx0 + (grid.x1 - 1) * grid.gsx};
  const double ys[2] = {grid.gy0, grid.gy0 + (grid.ny - 1) * grid.gsy};
  const double zs[2] = {grid.gz0, grid.gz0 + (grid.nz - 1) * grid.gsz};
  for (int k = view0; k < view0 + n; ++k) {
    const BpView& c = views[k];
    for (double X : xs)
      for (double Y : ys) {
        const double U = c.cx * X + c.cy * Y + c.c0;
        for (double Z : zs) {
          const double sv = c.sv0 + c.svk * Z / U;
          lo = std::min(lo, sv);
          hi = std::max(hi, sv);
        }
      }
  }
  const double wlo = std::floor(grid.lo_v), whi = std::floor(grid.hi_v);
  const double rlo = std::max(std::floor(lo) - 2.0, wlo), rhi = std::min(std::floor(hi) + 2.0, whi) + 1.0;
  int r0 = (int)rlo + kPadLo, r1 = (int)rhi + kPadLo;  // padded rows, inclusive
  r0 = std::max(r0, 0);
  r1 = std::min(r1, grid.rows - 1);
  if (r1 < r0) r1 = r0;
  lds_r0 = r0;
  lds_rows = r1 - r0 + 1;
}

}  // namespace monte
