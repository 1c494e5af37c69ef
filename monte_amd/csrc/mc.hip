// Monte Carlo photon transport for gfx950.
//
// Reference: the projection kernel of monte_cu/CBCT_real325.cu:419-846 and
// photon::delta_sampling :848-975 (voxel form: CBCT_real325im.cu:425-962).
// The reference gives each thread one detector pixel and loops serially over
// 360 views x 10^4 photons, so lanes idle whenever their photons' histories
// differ in length.  Here:
//   * one photon per lane, its random numbers drawn from a Philox4x32-10
//     stream keyed by (seed) with counter (block, photon, view, job), so any
//     history can be replayed and results do not depend on scheduling;
//   * the kernel is persistent: waves pull batches of photons (work units)
//     from a device queue shared by all jobs of a launch;
//   * the loop advances every live lane by one Woodcock step per iteration;
//     lanes whose photon has been absorbed, capped or tallied are refilled
//     at once from the wave's current batch (ballot + mbcnt prefix ranks), so
//     the wave stays full until the queue drains;
//   * flights are clipped to the phantom box (no Woodcock steps through air,
//     quirk Q9) and the first step starts at the box entry;
//   * primary and scatter counts are tallied with FP64 atomics.  Every tally
//     adds 1.0, so sums are exact integers and independent of order.
#include "internal.h"

namespace monte {

// ------------------------------------------------------------------ RNG
// Philox4x32-10 (Salmon et al., SC'11).
__device__ inline void philox10(uint32_t& c0, uint32_t& c1, uint32_t& c2, uint32_t& c3, uint32_t k0,
                                uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    const uint32_t lo0 = 0xD2511F53u * c0, hi0 = __umulhi(0xD2511F53u, c0);
    const uint32_t lo1 = 0xCD9E8D57u * c2, hi1 = __umulhi(0xCD9E8D57u, c2);
    const uint32_t n0 = hi1 ^ c1 ^ k0;
    const uint32_t n2 = hi0 ^ c3 ^ k1;
    c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
    k0 += 0x9E3779B9u; k1 += 0xBB67AE85u;
  }
}

struct Stream {
  uint32_t ctr, photon, view, job;  // counter words
  uint32_t b0, b1, b2, b3;          // unused outputs of the last block
  int left;
};

__device__ inline float next_u01(Stream& s, uint32_t k0, uint32_t k1) {
  if (s.left == 0) {
    uint32_t c0 = s.ctr, c1 = s.photon, c2 = s.view, c3 = s.job;
    philox10(c0, c1, c2, c3, k0, k1);
    s.b0 = c0; s.b1 = c1; s.b2 = c2; s.b3 = c3;
    s.ctr += 1;
    s.left = 4;
  }
  const uint32_t x = s.b0;
  s.b0 = s.b1; s.b1 = s.b2; s.b2 = s.b3;
  s.left -= 1;
  return ((float)(x >> 8) + 0.5f) * 5.9604644775390625e-08f;  // 2^-24
}

// --------------------------------------------------------------- physics
// Kahn's rejection sampling of the Klein-Nishina distribution as written in
// CBCT_real325.cu:687-722; returns rho = lambda'/lambda.
__device__ inline float kahn(float e_kev, Stream& s, uint32_t k0, uint32_t k1) {
  const float lambda = 511.0f / e_kev;
  const float p1 = (lambda + 2.0f) / (9.0f * lambda + 2.0f);
  for (int guard = 0; guard < 1000; ++guard) {
    const float r1 = next_u01(s, k0, k1);
    if (r1 < p1) {
      const float r2 = next_u01(s, k0, k1);
      const float ro = 1.0f + (2.0f / lambda) * r2;
      const float r3 = next_u01(s, k0, k1);
      if (r3 <= 4.0f * ((1.0f / ro) - (1.0f / (ro * ro)))) return ro;
    } else {
      const float r2 = next_u01(s, k0, k1);
      const float ro = (lambda + 2.0f) / (lambda + 2.0f * (1.0f - r2));
      const float r3 = next_u01(s, k0, k1);
      const float q = lambda - ro * lambda + 1.0f;
      if (r3 <= 0.5f * (q * q + (1.0f / ro))) return ro;
    }
  }
  return 1.0f;  // unreachable in practice (acceptance > 1/2 per trial)
}

// Thomson angular distribution (1 + mu^2): closed-form inverse CDF.
__device__ inline float thomson_mu(float u) {
  const float a = 4.0f * u - 2.0f;
  const float A = fabsf(a) + sqrtf(a * a + 1.0f);
  const float m = cbrtf(A) - cbrtf(1.0f / A);
  return a < 0.f ? -m : m;
}

// Rotate unit direction d by polar cosine c and azimuth phi.
__device__ inline void rotate_dir(float& dx, float& dy, float& dz, float c, float phi) {
  const float s = sqrtf(fmaxf(0.f, 1.f - c * c));
  float sp, cp;
  sincosf(phi, &sp, &cp);
  const float r = sqrtf(fmaxf(0.f, 1.f - dz * dz));
  float nx, ny, nz;
  if (r > 1e-5f) {
    const float ir = 1.0f / r;
    nx = c * dx + s * (dx * dz * cp - dy * sp) * ir;
    ny = c * dy + s * (dy * dz * cp + dx * sp) * ir;
    nz = c * dz - s * r * cp;
  } else {
    nx = s * cp;
    ny = s * sp;
    nz = dz >= 0.f ? c : -c;
  }
  const float nn = rsqrtf(nx * nx + ny * ny + nz * nz);
  dx = nx * nn; dy = ny * nn; dz = nz * nn;
}

__device__ inline int ebin_of(float e) {
  int b = (int)(e + 0.5f);  // CBCT_real325.cu:489: row (int)(Energy + 0.5)
  return b < 1 ? 1 : (b > kRows - 1 ? kRows - 1 : b);
}

// Slab test: parametric interval of p + t d inside the box.
__device__ inline bool ray_box(float px, float py, float pz, float dx, float dy, float dz, const McJob& J,
                               float& t0, float& t1) {
  float lo = -INFINITY, hi = INFINITY;
  const float p[3] = {px, py, pz}, d[3] = {dx, dy, dz};
  const float b0[3] = {J.bx0, J.by0, J.bz0}, b1[3] = {J.bx1, J.by1, J.bz1};
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    if (d[a] != 0.f) {
      const float inv = 1.0f / d[a];
      float ta = (b0[a] - p[a]) * inv, tb = (b1[a] - p[a]) * inv;
      if (ta > tb) { const float tt = ta; ta = tb; tb = tt; }
      lo = fmaxf(lo, ta);
      hi = fminf(hi, tb);
    } else if (p[a] < b0[a] || p[a] >= b1[a]) {
      return false;
    }
  }
  if (!(hi > lo)) return false;
  t0 = lo; t1 = hi;
  return true;
}

struct LaneCounters {
  uint32_t histories, steps, interactions, photo, coherent, compton, tp, ts, missed, capped;
};

__device__ inline unsigned long long wave_sum(uint32_t v) {
  unsigned long long x = v;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
  return x;
}

__device__ inline void flush_counters(LaneCounters& c, McCounters* dst) {
  const unsigned long long h = wave_sum(c.histories), st = wave_sum(c.steps), in = wave_sum(c.interactions),
                           ph = wave_sum(c.photo), co = wave_sum(c.coherent), cp = wave_sum(c.compton),
                           tp = wave_sum(c.tp), ts = wave_sum(c.ts), mi = wave_sum(c.missed),
                           ca = wave_sum(c.capped);
  if (__lane_id() == 0) {
    atomicAdd(&dst->histories, h); atomicAdd(&dst->steps, st); atomicAdd(&dst->interactions, in);
    atomicAdd(&dst->photo, ph); atomicAdd(&dst->coherent, co); atomicAdd(&dst->compton, cp);
    atomicAdd(&dst->tallied_primary, tp); atomicAdd(&dst->tallied_scatter, ts);
    atomicAdd(&dst->missed, mi); atomicAdd(&dst->capped, ca);
  }
  c = LaneCounters{};
}

__device__ inline int job_of_unit(const McJob* __restrict__ jobs, int n_jobs, uint64_t u) {
  int lo = 0, hi = n_jobs - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (jobs[mid].unit_begin <= u) lo = mid; else hi = mid - 1;
  }
  return lo;
}

// Photon state of one lane.
struct Photon {
  float px, py, pz;    // position (world)
  float dx, dy, dz;    // unit direction (world)
  float e;             // energy (keV)
  float t, t1;         // flight parameter and box-exit parameter
  float mumax, inv_mumax;
  float cb, sb;        // view rotation
  int eb;              // table row
  int nint;            // real interactions so far
  int view;            // tally view index (absolute)
};

// Straight line from the current position to the detector plane x' = sdd - sod.
__device__ inline void tally(const Photon& p, const McJob& J, LaneCounters& cnt) {
  const float xq = p.px * p.cb + p.py * p.sb, yq = -p.px * p.sb + p.py * p.cb;
  const float ddx = p.dx * p.cb + p.dy * p.sb, ddy = -p.dx * p.sb + p.dy * p.cb;
  if (!(ddx > 0.f)) { cnt.missed++; return; }
  const float tt = ((J.sdd - J.sod) - xq) / ddx;
  const float yh = fmaf(tt, ddy, yq), zh = fmaf(tt, p.dz, p.pz);
  const float ui = fmaf(J.us, yh, J.un0), vi = fmaf(J.vs, zh, J.vn0);
  if (!(ui >= 0.f && ui < (float)J.nu && vi >= 0.f && vi < (float)J.nv)) { cnt.missed++; return; }
  int iu = (int)floorf(ui), iv = (int)floorf(vi);
  iu = iu > J.nu - 1 ? J.nu - 1 : iu;
  iv = iv > J.nv - 1 ? J.nv - 1 : iv;
  const size_t idx = ((size_t)p.view * J.nu + iu) * J.nv + iv;
  if (p.nint == 0) {
    unsafeAtomicAdd(J.primary + idx, 1.0);
    cnt.tp++;
  } else {
    unsafeAtomicAdd(J.scatter + idx, 1.0);
    cnt.ts++;
  }
}

// Starts a flight from the current position; returns false if the photon
// leaves without entering the phantom box (caller tallies it).
__device__ inline bool start_flight(Photon& p, const McJob& J) {
  float t0, t1;
  if (!ray_box(p.px, p.py, p.pz, p.dx, p.dy, p.dz, J, t0, t1) || !(t1 > 0.f)) return false;
  const float mm = J.tables->mumax[p.eb];
  if (!(mm > 0.f)) return false;
  if (t0 > 0.f) {
    p.px = fmaf(t0, p.dx, p.px); p.py = fmaf(t0, p.dy, p.py); p.pz = fmaf(t0, p.dz, p.pz);
    t1 -= t0;
  }
  p.t = 0.f;
  p.t1 = t1;
  p.mumax = mm;
  p.inv_mumax = 1.0f / mm;
  return true;
}

__global__ __launch_bounds__(256) void mc_transport_kernel(const McJob* __restrict__ jobs, int n_jobs,
                                                           uint64_t total_units, unsigned int* __restrict__ queue,
                                                           McCounters* __restrict__ counters, int batch) {
  const uint64_t full = ~0ull;
  int job = -1;
  uint32_t next = 0, end = 0;  // photon range of the current unit (wave-uniform)
  int unit_view = 0;
  long long pending = -1;
  bool exhausted = false;

  bool alive = false;
  Photon p{};
  Stream rs{};
  LaneCounters cnt{};
  McJob J{};
  const int lane = __lane_id();

  for (;;) {
    const uint64_t dead = __ballot(!alive);
    // 1) secure a work unit when lanes are idle and the current one is spent
    if (dead != 0 && next >= end) {
      if (pending < 0 && !exhausted) {
        unsigned int u = 0;
        if (lane == 0) u = atomicAdd(queue, 1u);
        u = __shfl(u, 0, 64);
        if ((uint64_t)u >= total_units) exhausted = true; else pending = (long long)u;
      }
      if (pending >= 0) {
        const int pj = __builtin_amdgcn_readfirstlane(job_of_unit(jobs, n_jobs, (uint64_t)pending));
        if (pj == job || dead == full) {
          if (pj != job) {
            if (job >= 0) flush_counters(cnt, counters + job);
            job = pj;
            J = jobs[job];
          }
          const uint64_t local = (uint64_t)pending - J.unit_begin;
          const uint32_t vl = (uint32_t)(local / J.batches_per_view);
          const uint32_t bi = (uint32_t)(local - (uint64_t)vl * J.batches_per_view);
          unit_view = J.view_begin + (int)vl;
          next = bi * (uint32_t)batch;
          end = min(next + (uint32_t)batch, J.photons_per_view);
          pending = -1;
        }
      }
    }
    if (dead == full && next >= end && pending < 0 && exhausted) break;

    // 2) refill dead lanes from the current unit
    if (dead != 0 && next < end) {
      const uint32_t rank = __builtin_amdgcn_mbcnt_hi((uint32_t)(dead >> 32),
                                                      __builtin_amdgcn_mbcnt_lo((uint32_t)dead, 0u));
      const uint32_t ndead = (uint32_t)__popcll(dead);
      if (!alive) {
        const uint32_t id = next + rank;
        if (id < end) {
          cnt.histories++;
          rs.ctr = 0; rs.photon = id; rs.view = (uint32_t)unit_view; rs.job = J.job_id; rs.left = 0;
          // aim at the centre of pixel id mod (nu*nv) (CBCT_real325.cu:464-466)
          const uint32_t npix = (uint32_t)(J.nu * J.nv);
          const uint32_t pix = id % npix;
          const int iu = (int)(pix / (uint32_t)J.nv), iv = (int)(pix - (uint32_t)iu * (uint32_t)J.nv);
          const float yp = ((float)iu + 0.5f - J.un0) * J.u_step;
          const float zp = ((float)iv + 0.5f - J.vn0) * J.v_step;
          const float in = rsqrtf(J.sdd * J.sdd + yp * yp + zp * zp);
          const float2 cs = J.view_cs[unit_view];
          p.cb = cs.x; p.sb = cs.y;
          p.view = unit_view;
          p.dx = (J.sdd * cs.x - yp * cs.y) * in;
          p.dy = (J.sdd * cs.y + yp * cs.x) * in;
          p.dz = zp * in;
          p.px = -J.sod * cs.x; p.py = -J.sod * cs.y; p.pz = 0.f;
          float e = J.e_kev;
          if (J.spectrum_cdf) {
            const float u = next_u01(rs, J.key0, J.key1);
            int lo = 1, hi = kRows - 1;
            while (lo < hi) {
              const int mid = (lo + hi) >> 1;
              if (J.spectrum_cdf[mid] < u) lo = mid + 1; else hi = mid;
            }
            e = (float)lo;
          }
          p.e = e;
          p.eb = ebin_of(e);
          p.nint = 0;
          if (start_flight(p, J)) {
            alive = true;
          } else {
            tally(p, J, cnt);
          }
        }
      }
      next = min(end, next + ndead);
    }

    // 3) one Woodcock step for every live lane
    if (alive) {
      const float u = next_u01(rs, J.key0, J.key1);
      p.t = fmaf(-__logf(u), p.inv_mumax, p.t);
      cnt.steps++;
      if (p.t >= p.t1) {
        tally(p, J, cnt);
        alive = false;
      } else {
        const float qx = fmaf(p.t, p.dx, p.px), qy = fmaf(p.t, p.dy, p.py), qz = fmaf(p.t, p.dz, p.pz);
        int ix = (int)floorf((qx - J.bx0) * J.inv_vx);
        int iy = (int)floorf((qy - J.by0) * J.inv_vy);
        int iz = (int)floorf((qz - J.bz0) * J.inv_vz);
        ix = min(max(ix, 0), J.nx - 1);
        iy = min(max(iy, 0), J.ny - 1);
        iz = min(max(iz, 0), J.nz - 1);
        const int lab = J.labels[((size_t)iz * J.ny + iy) * J.nx + ix];
        const float4 m = J.tables->m[lab][p.eb];
        const float u2 = next_u01(rs, J.key0, J.key1);
        if (u2 * p.mumax < m.x) {
          // real collision
          p.px = qx; p.py = qy; p.pz = qz;
          if (p.nint >= J.max_int) {
            cnt.capped++;
            alive = false;
          } else {
            const float u3 = next_u01(rs, J.key0, J.key1);
            cnt.interactions++;
            if (u3 <= m.y) {
              cnt.photo++;
              alive = false;
            } else {
              if (u3 <= m.z) {
                cnt.coherent++;
                if (J.coherent) {
                  const float cth = thomson_mu(next_u01(rs, J.key0, J.key1));
                  const float phi = 6.283185307179586f * next_u01(rs, J.key0, J.key1);
                  rotate_dir(p.dx, p.dy, p.dz, cth, phi);
                }
              } else {
                cnt.compton++;
                const float ro = kahn(p.e, rs, J.key0, J.key1);
                const float lam = 511.0f / p.e, lamd = ro * lam;
                float cth = 1.0f - (lamd - lam);
                cth = cth < -1.f ? -1.f : cth;
                p.e = 511.0f / lamd;
                p.eb = ebin_of(p.e);
                const float phi = 6.283185307179586f * next_u01(rs, J.key0, J.key1);
                rotate_dir(p.dx, p.dy, p.dz, cth, phi);
              }
              p.nint += 1;
              if (!start_flight(p, J)) {
                tally(p, J, cnt);
                alive = false;
              }
            }
          }
        }
      }
    }
  }
  if (job >= 0) flush_counters(cnt, counters + job);
}

__global__ void mc_reset_kernel(unsigned int* queue, McCounters* counters, int n_jobs) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0) *queue = 0u;
  if (i < n_jobs) counters[i] = McCounters{};
}

hipError_t launch_mc_reset(unsigned int* queue, McCounters* counters, int n_jobs, hipStream_t s) {
  const int n = n_jobs > 1 ? n_jobs : 1;
  hipLaunchKernelGGL(mc_reset_kernel, dim3((n + 255) / 256), dim3(256), 0, s, queue, counters, n_jobs);
  return hipGetLastError();
}

hipError_t launch_mc_transport(const McJob* jobs_dev, int n_jobs, uint64_t total_units, unsigned int* queue,
                               McCounters* counters, int batch, int block, int grid, hipStream_t s) {
  if (block % 64 || block > 256 || block <= 0 || grid <= 0) return hipErrorInvalidValue;
  hipLaunchKernelGGL(mc_transport_kernel, dim3(grid), dim3(block), 0, s, jobs_dev, n_jobs, total_units, queue,
                     counters, batch);
  return hipGetLastError();
}

// ------------------------------------------------- exact primary (Siddon)
// One thread per (view, pixel): per-label path lengths along the source ->
// pixel-centre ray by index-stepping traversal (each crossing recomputed from
// its plane index, no running sums), kept in LDS, then combined with the
// attenuation of each spectral bin.
__global__ __launch_bounds__(256) void primary_exact_kernel(McJob J, const float* __restrict__ spec_w, int use_spec,
                                                            float* __restrict__ out, int n_views) {
  __shared__ float lens[kMaxLabels][256];
  const int tid = threadIdx.x;
  const long ray = (long)blockIdx.x * 256 + tid;
  const long nray = (long)n_views * J.nu * J.nv;
#pragma unroll
  for (int l = 0; l < kMaxLabels; ++l) lens[l][tid] = 0.f;
  if (ray >= nray) return;
  const int k = (int)(ray / ((long)J.nu * J.nv));
  const int pix = (int)(ray - (long)k * J.nu * J.nv);
  const int iu = pix / J.nv, iv = pix - iu * J.nv;
  const float2 cs = J.view_cs[J.view_begin + k];
  const float yp = ((float)iu + 0.5f - J.un0) * J.u_step;
  const float zp = ((float)iv + 0.5f - J.vn0) * J.v_step;
  const float xd = J.sdd - J.sod;
  const float p0x = -J.sod * cs.x, p0y = -J.sod * cs.y, p0z = 0.f;
  const float p1x = xd * cs.x - yp * cs.y, p1y = xd * cs.y + yp * cs.x, p1z = zp;
  float dx = p1x - p0x, dy = p1y - p0y, dz = p1z - p0z;
  const float L = sqrtf(dx * dx + dy * dy + dz * dz);
  const float il = 1.0f / L;
  dx *= il; dy *= il; dz *= il;
  float t0, t1;
  if (ray_box(p0x, p0y, p0z, dx, dy, dz, J, t0, t1)) {
    t0 = fmaxf(t0, 0.f);
    t1 = fminf(t1, L);
    if (t1 > t0) {
      const float ex = fmaf(t0, dx, p0x), ey = fmaf(t0, dy, p0y), ez = fmaf(t0, dz, p0z);
      const float total = t1 - t0;
      const float e[3] = {ex, ey, ez}, d[3] = {dx, dy, dz};
      const float b0[3] = {J.bx0, J.by0, J.bz0};
      const float vs[3] = {J.vx, J.vy, J.vz};
      const float ivs[3] = {J.inv_vx, J.inv_vy, J.inv_vz};
      const int n[3] = {J.nx, J.ny, J.nz};
      int idx[3], step[3];
      float sn[3], invd[3];
      const float sm = fminf(0.5f * total, 1e-4f);
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        int i = (int)floorf((fmaf(sm, d[a], e[a]) - b0[a]) * ivs[a]);
        i = min(max(i, 0), n[a] - 1);
        idx[a] = i;
        invd[a] = d[a] != 0.f ? 1.0f / d[a] : 0.f;
        if (d[a] > 0.f) {
          step[a] = 1;
          sn[a] = (fmaf((float)(i + 1), vs[a], b0[a]) - e[a]) * invd[a];
        } else if (d[a] < 0.f) {
          step[a] = -1;
          sn[a] = (fmaf((float)i, vs[a], b0[a]) - e[a]) * invd[a];
        } else {
          step[a] = 0;
          sn[a] = INFINITY;
        }
      }
      float s = 0.f;
      for (int guard = 0; guard < 4 * (J.nx + J.ny + J.nz) + 8; ++guard) {
        int a = 0;
        if (sn[1] < sn[a]) a = 1;
        if (sn[2] < sn[a]) a = 2;
        const float se = fminf(sn[a], total);
        if (se > s) {
          const int lab = J.labels[((size_t)idx[2] * J.ny + idx[1]) * J.nx + idx[0]];
          lens[lab][tid] += se - s;
          s = se;
        }
        if (sn[a] >= total) break;
        idx[a] += step[a];
        if (idx[a] < 0 || idx[a] >= n[a]) break;
        sn[a] = (fmaf((float)(step[a] > 0 ? idx[a] + 1 : idx[a]), vs[a], b0[a]) - e[a]) * invd[a];
      }
    }
  }
  float result;
  if (!use_spec) {
    const int eb = ebin_of(J.e_kev);
    float acc = 0.f;
    for (int l = 0; l < kMaxLabels; ++l) {
      const float len = lens[l][tid];
      if (len != 0.f) acc = fmaf(len, J.tables->m[l][eb].x, acc);
    }
    result = acc;
  } else {
    float num = 0.f, den = 0.f;
    for (int eb = 1; eb < kRows; ++eb) {
      const float w = spec_w[eb];
      if (w <= 0.f) continue;
      float le = 0.f;
      for (int l = 0; l < kMaxLabels; ++l) {
        const float len = lens[l][tid];
        if (len != 0.f) le = fmaf(len, J.tables->m[l][eb].x, le);
      }
      num = fmaf(w, expf(-le), num);
      den += w;
    }
    result = -logf(num / den);
  }
  out[ray] = result;
}

hipError_t launch_primary_exact(const McJob& job, const float* spec_w, int n_spec, float* out, hipStream_t s) {
  const long nray = (long)job.n_views_sim * job.nu * job.nv;
  if (nray <= 0) return hipSuccess;
  const long blocks = (nray + 255) / 256;
  hipLaunchKernelGGL(primary_exact_kernel, dim3((unsigned)blocks), dim3(256), 0, s, job, spec_w, n_spec > 0 ? 1 : 0,
                     out, job.n_views_sim);
  return hipGetLastError();
}

// ------------------------------------------------- counts -> line integrals
// CBCT_real325.cu:267-294: clamp N to [1, N0], L = ln N0 - ln N, with N0 the
// photons aimed at the pixel: floor(N/npix) + (pixel < N mod npix).
__global__ void counts_to_lineint_kernel(const double* __restrict__ counts, float* __restrict__ out, int n_views,
                                         int npix, uint64_t photons_per_view) {
  const size_t n = (size_t)n_views * npix;
  const uint64_t base = photons_per_view / (uint64_t)npix, rem = photons_per_view % (uint64_t)npix;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const int pix = (int)(i % (size_t)npix);
    const double n0 = (double)(base + ((uint64_t)pix < rem ? 1 : 0));
    double c = counts[i];
    c = c > n0 ? n0 : c;
    c = c < 1.0 ? 1.0 : c;
    out[i] = n0 > 0.0 ? (float)(log(n0) - log(c)) : 0.f;
  }
}

hipError_t launch_counts_to_lineint(const double* counts, float* out, int n_views, int nu, int nv,
                                    uint64_t photons_per_view, hipStream_t s) {
  const size_t n = (size_t)n_views * nu * nv;
  size_t blocks = (n + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  if (blocks == 0) return hipSuccess;
  hipLaunchKernelGGL(counts_to_lineint_kernel, dim3((unsigned)blocks), dim3(256), 0, s, counts, out, n_views,
                     nu * nv, photons_per_view);
  return hipGetLastError();
}

}  // namespace monte
