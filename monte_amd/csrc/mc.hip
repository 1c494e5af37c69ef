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
//   * the kernel is persistent: waves pull work units (a detector tile of 8 x 8
//     pixels over a block of photon cycles, in one view) from device work
//     queues shared by all jobs of a launch (24 counters, each block drawing
//     from its own first), so the units in flight share a thin slab of the
//     phantom and few tally lines;
//   * each loop iteration advances every flying lane by up to two Woodcock
//     steps (one Philox block).
//     Lanes that collide (real interaction pending), end their flight (tally
//     pending) or die (absorbed, capped) park; the wave runs the interaction
//     code, or tallies and refills (ballot + mbcnt prefix ranks), only once
//     enough lanes are parked or nothing else can run, so those branches
//     execute near full width instead of for one lane at a time;
//   * photons are born 64 at a time, one per lane with every lane active, into
//     a per-wave LDS buffer that dead lanes then take from;
//   * flights are clipped to the bounding box of the voxels that attenuate
//     (the reference clips to the phantom box, quirk Q9; mu is 0 in the
//     margin between the two, so no history changes) and the first step
//     starts at the box entry;
//   * primary and scatter counts are tallied with FP64 atomics.  Every tally
//     adds 1.0, so sums are exact integers and independent of order;
//   * statistics are counted per wave, from ballots of the lanes where events
//     happen, into 64-bit counters in LDS that go to the job's when it is
//     flushed.
#include <algorithm>
#include <climits>

#include "internal.h"

namespace monte {

// ------------------------------------------------------------------ RNG
// Philox4x32-10 (Salmon et al., SC'11).
__device__ inline void philox10(uint32_t& c0, uint32_t& c1, uint32_t& c2, uint32_t& c3, uint32_t k0,
                                uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    // one 32 x 32 -> 64-bit multiply gives both halves (v_mad_u64_u32)
    const uint64_t p0 = (uint64_t)0xD2511F53u * c0, p1 = (uint64_t)0xCD9E8D57u * c2;
    const uint32_t lo0 = (uint32_t)p0, hi0 = (uint32_t)(p0 >> 32);
    const uint32_t lo1 = (uint32_t)p1, hi1 = (uint32_t)(p1 >> 32);
    // hi ^ c ^ k as one three-input XOR (truth table 0x96)
    uint32_t n0, n2;
    asm("v_bitop3_b32 %0, %1, %2, %3 bitop3:0x96" : "=v"(n0) : "v"(hi1), "v"(c1), "s"(k0));
    asm("v_bitop3_b32 %0, %1, %2, %3 bitop3:0x96" : "=v"(n2) : "v"(hi0), "v"(c3), "s"(k1));
    c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
    k0 += 0x9E3779B9u; k1 += 0xBB67AE85u;
  }
}

// A photon's stream: counter (ctr, photon, view, job); the four outputs of the
// last block are consumed in order.
struct Stream {
  uint32_t ctr, photon, view;
  uint32_t b0, b1, b2, b3;
  int left;
};

// ((x >> 8) + 1/2) 2^-24 in one rounding
__device__ inline float u01(uint32_t x) { return fmaf((float)(x >> 8), 5.9604644775390625e-08f, 2.98023223876953125e-08f); }

__device__ inline float next_u01(Stream& s, uint32_t job, uint32_t k0, uint32_t k1) {
  if (s.left == 0) {
    uint32_t c0 = s.ctr, c1 = s.photon, c2 = s.view, c3 = job;
    philox10(c0, c1, c2, c3, k0, k1);
    s.b0 = c0; s.b1 = c1; s.b2 = c2; s.b3 = c3;
    s.ctr += 1;
    s.left = 4;
  }
  const uint32_t x = s.b0;
  s.b0 = s.b1; s.b1 = s.b2; s.b2 = s.b3;
  s.left -= 1;
  return u01(x);
}

// The stream's next three uniforms, in order, with at most one new block: the
// same numbers as three next_u01 calls, but with one Philox call site, so a
// wave whose lanes sit at different offsets in their blocks runs the rounds
// once per trial of a rejection loop rather than once per draw.
__device__ inline void next3_u01(Stream& s, uint32_t job, uint32_t k0, uint32_t k1, float& r1, float& r2, float& r3) {
  uint32_t n0 = 0, n1 = 0, n2 = 0, n3 = 0;
  const int left = s.left;
  if (left < 3) {
    n0 = s.ctr; n1 = s.photon; n2 = s.view; n3 = job;
    philox10(n0, n1, n2, n3, k0, k1);
    s.ctr += 1;
  }
  // the unconsumed words in order are b0 .. b(left-1), then n0 .. n3
  const uint32_t x0 = left == 0 ? n0 : s.b0;
  const uint32_t x1 = left == 0 ? n1 : (left == 1 ? n0 : s.b1);
  const uint32_t x2 = left == 0 ? n2 : (left == 1 ? n1 : (left == 2 ? n0 : s.b2));
  // and what stays: left + 1 words after a new block, left - 3 otherwise
  const uint32_t w0 = left == 0 ? n3 : (left == 1 ? n2 : (left == 2 ? n1 : s.b3));
  const uint32_t w1 = left == 1 ? n3 : n2;
  s.b0 = w0; s.b1 = w1; s.b2 = n3;
  s.left = left < 3 ? left + 1 : left - 3;
  r1 = u01(x0); r2 = u01(x1); r3 = u01(x2);
}

// Reciprocal to 1 ulp (v_rcp_f32).  The physics below draws angles and
// energies whose parity with the FP64 oracle is statistical, so it takes the
// hardware reciprocal, square root and sine instead of correctly rounded
// sequences.
__device__ inline float rcp(float x) { return __builtin_amdgcn_rcpf(x); }

// --------------------------------------------------------------- physics
// Kahn's rejection sampling of the Klein-Nishina distribution as written in
// CBCT_real325.cu:687-722 (lambda = 511/E); returns rho = lambda'/lambda.
// Each trial takes the stream's next three uniforms (r1, r2, r3) whichever
// branch r1 picks, as the reference's trials do.
__device__ inline float kahn(float lambda, Stream& s, uint32_t job, uint32_t k0, uint32_t k1) {
  const float il2 = 2.0f * rcp(lambda);
  const float p1 = (lambda + 2.0f) * rcp(9.0f * lambda + 2.0f);
  for (int guard = 0; guard < 1000; ++guard) {
    float r1, r2, r3;
    next3_u01(s, job, k0, k1, r1, r2, r3);
    if (r1 < p1) {
      const float ro = 1.0f + il2 * r2;
      const float iro = rcp(ro);
      if (r3 <= 4.0f * (iro - iro * iro)) return ro;
    } else {
      const float ro = (lambda + 2.0f) * rcp(lambda + 2.0f * (1.0f - r2));
      const float q = lambda - ro * lambda + 1.0f;
      if (r3 <= 0.5f * (q * q + rcp(ro))) return ro;
    }
  }
  return 1.0f;  // unreachable in practice (acceptance > 1/2 per trial)
}

// Thomson angular distribution (1 + mu^2): closed-form inverse CDF
// mu = cbrt(A) - cbrt(1/A), A = |a| + sqrt(a^2 + 1) >= 1, a = 4u - 2, with
// cbrt(A) = 2^(log2(A) / 3) on the hardware log and exp.
__device__ inline float thomson_mu(float u) {
  const float a = 4.0f * u - 2.0f;
  const float A = fabsf(a) + __builtin_amdgcn_sqrtf(a * a + 1.0f);
  const float c = __builtin_amdgcn_exp2f(__builtin_amdgcn_logf(A) * 0.333333343f);
  const float m = c - rcp(c);
  return a < 0.f ? -m : m;
}

// Rotate unit direction d by polar cosine c and azimuth phi = 2 pi w (w in
// revolutions, as the hardware sine takes it): c d + s (cos(phi) e1 +
// sin(phi) e2) with e1 = (dx dz, dy dz, -r^2)/r, e2 = (-dy, dx, 0)/r and
// r = sqrt(dx^2 + dy^2), which stays accurate near the poles (1 - dz^2 would
// cancel); only for d on the z axis is the pair the x and y axes.
__device__ inline void rotate_dir(float& dx, float& dy, float& dz, float c, float w) {
  const float s = __builtin_amdgcn_sqrtf(fmaxf(0.f, 1.f - c * c));
  const float sp = __builtin_amdgcn_sinf(w), cp = __builtin_amdgcn_cosf(w);
  const float r = __builtin_amdgcn_sqrtf(dx * dx + dy * dy);
  float nx, ny, nz;
  if (r > 1e-15f) {  // r^2 stays a normal FP32 number
    const float ir = rcp(r);
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

// Slab test along one axis, folded into [lo, hi]; false if the ray misses.
__device__ inline bool slab(float p, float d, float b0, float b1, float& lo, float& hi) {
  if (d != 0.f) {
    const float inv = __builtin_amdgcn_rcpf(d);  // 1 ulp: the box clip has voxel-scale slack
    float ta = (b0 - p) * inv, tb = (b1 - p) * inv;
    lo = fmaxf(lo, fminf(ta, tb));
    hi = fminf(hi, fmaxf(ta, tb));
    return true;
  }
  return p >= b0 && p < b1;
}

// Parametric interval of p + t d inside the phantom box.
template <class JT>
__device__ inline bool ray_box(float px, float py, float pz, float dx, float dy, float dz, const JT& J,
                               float& t0, float& t1) {
  float lo = -INFINITY, hi = INFINITY;
  const bool ok = slab(px, dx, J.bx0, J.bx1, lo, hi) && slab(py, dy, J.by0, J.by1, lo, hi) &&
                  slab(pz, dz, J.bz0, J.bz1, lo, hi);
  if (!ok || !(hi > lo)) return false;
  t0 = lo; t1 = hi;
  return true;
}

using CJob = const __attribute__((address_space(4))) McJob;
using CU32 = const __attribute__((address_space(4))) uint32_t;

__device__ inline int job_of_unit(const CJob* jobs, int n_jobs, uint64_t u) {
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
  float mumax;         // Woodcock majorant of the current table row
  float lscale;        // ln 2 / mumax: a free path is -log2(u) * lscale
  float mu0, mu1, mu2, mu3;  // mu of labels 0..3 at the current table row
  int eb;              // table row
  int lab;             // label at the pending collision
  int nint;            // real interactions so far
  int view;            // tally view index (absolute)
  float cv, sv;        // cos, sin of the view angle
};

// Per-wave LDS copy of the current job's per-energy tables, read at every
// birth and flight start (the spectrum search alone is 8 dependent reads).
constexpr int kGuide = 256;  // spectrum guide cells

// Photons born 64 at a time, one per lane with every lane active, wait here
// until dead lanes take them; all are of one view and one cycle of one
// detector tile, and slots are the lanes that made them, i.e. the tile's
// pixels.  The rays of the tile's pixels in that view (direction, flight-box
// entry t0 (0 when the source is not clipped) and exit t1 (-1: no flight),
// label set) are formed once per tile and view and serve every cycle of it;
// per photon there is only the table row eb.  slot[k] is the slot of the k-th
// photon that entered the flight box.  A taker forms the rest as the birth
// did: the photon index from the slot, the position from the source and t0,
// the energy from the table row, the majorant from the label set.
struct BirthBuffer {
  float4 d0[64];  // direction and box entry t0 of a slot's ray
  float t1[64];
  uint8_t set[64], eb[64], slot[64];
};

constexpr int kWaveStats = 10;
struct WaveTables {
  unsigned long long stats[kWaveStats];  // the wave's event counts for the current job
  float cdf[kRows];      // spectrum CDF (unused for mono jobs)
  float mumax[kRows];    // Woodcock majorant
  float mu[4][kRows];    // mu of labels 0..3
  // guide[g] = the energy row drawn for u = g / kGuide (rows are <= 200): a
  // draw u in [g, g + 1) / kGuide lies between guide[g] and guide[g + 1]
  uint8_t guide[kGuide + 4];
  BirthBuffer bb;
};
// 6 waves per SIMD (24 per CU) fit in a CU's 160 KB of LDS
static_assert(24 * sizeof(WaveTables) <= 160 * 1024, "WaveTables too large for 6 waves per SIMD");

// Inverse CDF of the spectrum (CBCT_real325.cu:479-486): the first row e in
// [1, kRows - 1] with cdf[e] >= u, or the last row, searched within [lo, hi].
__device__ inline int spectrum_row(const float* cdf, float u, int lo, int hi) {
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (cdf[mid] < u) lo = mid + 1; else hi = mid;
  }
  return lo;
}

// kTally: the flight ended; the photon waits to be tallied with the next refill
// batch (its position and direction do not change meanwhile)
enum { kDead = 0, kTally = 1, kFly = 2, kCollide = 3 };  // dead or tally pending: state <= kTally
// Statistics are counted per wave, not per lane: where every lane of the wave
// is active, a ballot of the lanes where an event happened is added to the
// wave's 64-bit counter in LDS, and the counters go to the job's when it is
// flushed.  (Ten per-lane counters, incremented in the branches, took 56 of the
// kernel's 121 VGPRs.)
enum { kStatHistories, kStatSteps, kStatInteractions, kStatPhoto, kStatCoherent, kStatCompton, kStatTp, kStatTs,
       kStatMissed, kStatCapped, kStats };
static_assert(kStats == kWaveStats, "one LDS counter per statistic");

__device__ inline unsigned popc(uint64_t m) { return (unsigned)__popcll(m); }

// Adds the lanes where `pred` holds to the wave's counter k.  Called where
// every lane of the wave is active; lane 0 adds for the wave.
__device__ inline void count_lanes(WaveTables& wt, int k, bool pred) {
  const uint64_t m = __ballot(pred);
  if (__lane_id() == 0 && m != 0) atomicAdd(&wt.stats[k], (unsigned long long)popc(m));
}

// Straight line from the current position to the detector plane x' = sdd - sod.
enum { kTallyMiss = 0, kTallyPrimary = 1, kTallyScatter = 2 };

// Counts the tallies of a section from each lane's tally code (-1: none).
// Misses are not counted here: every history ends once, tallied, missed,
// absorbed or capped, so the flush derives them.
__device__ inline void count_tallies(WaveTables& wt, int code) {
  const uint64_t mp = __ballot(code == kTallyPrimary), ms = __ballot(code == kTallyScatter);
  if (__lane_id() == 0 && (mp | ms) != 0) {
    atomicAdd(&wt.stats[kStatTp], (unsigned long long)popc(mp));
    atomicAdd(&wt.stats[kStatTs], (unsigned long long)popc(ms));
  }
}

template <class JT>
__device__ inline int tally(const Photon& p, const JT& J) {
  const float cx = p.cv, cy = p.sv;
  const float xq = p.px * cx + p.py * cy, yq = -p.px * cy + p.py * cx;
  const float ddx = p.dx * cx + p.dy * cy, ddy = -p.dx * cy + p.dy * cx;
  if (!(ddx > 0.f)) {
    return kTallyMiss;
  }
  const float tt = ((J.sdd - J.sod) - xq) * __builtin_amdgcn_rcpf(ddx);
  const float yh = fmaf(tt, ddy, yq), zh = fmaf(tt, p.dz, p.pz);
  const float ui = fmaf(J.us, yh, J.un0), vi = fmaf(J.vs, zh, J.vn0);
  if (!(ui >= 0.f && ui < (float)J.nu && vi >= 0.f && vi < (float)J.nv)) {
    return kTallyMiss;
  }
  int iu = (int)floorf(ui), iv = (int)floorf(vi);
  iu = iu > J.nu - 1 ? J.nu - 1 : iu;
  iv = iv > J.nv - 1 ? J.nv - 1 : iv;
  const size_t idx = ((size_t)p.view * J.nu + iu) * J.nv + iv;
  if (p.nint == 0) {
    unsafeAtomicAdd(J.primary + idx, 1.0);
    return kTallyPrimary;
  }
  unsafeAtomicAdd(J.scatter + idx, 1.0);
  return kTallyScatter;
}

// Label set meaning "any label": the flight takes the global majorant.
constexpr uint32_t kAnyLabel = 16u;

// The largest mu among the labels of `set` (bit l: label l in 0..3;
// kAnyLabel: every label) at the photon's table row, from its mu registers.
__device__ inline float set_majorant(const Photon& p, const WaveTables& wt, uint32_t set) {
  float mm;
  if (set & kAnyLabel) {
    mm = wt.mumax[p.eb];
  } else {
    mm = 0.f;
    if (set & 1u) mm = fmaxf(mm, p.mu0);
    if (set & 2u) mm = fmaxf(mm, p.mu1);
    if (set & 4u) mm = fmaxf(mm, p.mu2);
    if (set & 8u) mm = fmaxf(mm, p.mu3);
  }
  return mm;
}

// Starts a flight from the current position (after a scatter; a birth forms
// its tile's rays the same way, once per tile and view); returns false if the
// photon leaves without entering the flight box, or meets nothing that
// attenuates (the caller tallies it).  The Woodcock majorant is the largest mu
// among the labels of `set`.  t_in: the distance moved to the box entry, 0 if
// the photon starts inside.
template <class JT>
__device__ inline bool start_flight(Photon& p, const JT& J, const WaveTables& wt, uint32_t set, float& t_in) {
  p.mu0 = wt.mu[0][p.eb];
  p.mu1 = wt.mu[1][p.eb];
  p.mu2 = wt.mu[2][p.eb];
  p.mu3 = wt.mu[3][p.eb];
  const float mm = set_majorant(p, wt, set);
  if (!(mm > 0.f)) return false;
  float t0, t1;
  if (!ray_box(p.px, p.py, p.pz, p.dx, p.dy, p.dz, J, t0, t1) || !(t1 > 0.f)) return false;
  t_in = 0.f;
  if (t0 > 0.f) {
    p.px = fmaf(t0, p.dx, p.px); p.py = fmaf(t0, p.dy, p.py); p.pz = fmaf(t0, p.dz, p.pz);
    t1 -= t0;
    t_in = t0;
  }
  p.t = 0.f;
  p.t1 = t1;
  p.mumax = mm;
  p.lscale = 0.69314718055994531f * rcp(mm);
  return true;
}

// Copies a job's tables into the wave's LDS slot.
// The host builds the CDF from non-negative weights, so it is non-decreasing
// and the search within a guide cell returns the row a search over all rows
// would.
template <class JT>
__device__ inline void load_wave_tables(WaveTables& wt, const JT& J, int lane) {
  for (int e = lane; e < kRows; e += 64) {
    wt.cdf[e] = J.spectrum_cdf ? J.spectrum_cdf[e] : 0.f;
    wt.mumax[e] = J.tables->mumax[e];
#pragma unroll
    for (int l = 0; l < 4; ++l) wt.mu[l][e] = J.tables->m[l][e].x;
  }
  if (J.spectrum_cdf) {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");  // the CDF in LDS is complete
    __builtin_amdgcn_wave_barrier();
    for (int g = lane; g <= kGuide; g += 64)
      wt.guide[g] = (uint8_t)spectrum_row(wt.cdf, (float)g * (1.0f / kGuide), 1, kRows - 1);
  }
}

// Attenuation at a label: labels 0..3 from registers, others from the table.
template <class JT>
__device__ inline float mu_at(const Photon& p, int lab, const JT& J) {
  const bool odd = (lab & 1) != 0;
  const float lo = odd ? p.mu1 : p.mu0, hi = odd ? p.mu3 : p.mu2;
  float mu = (lab & 2) ? hi : lo;
  if (lab > 3) mu = J.tables->m[lab][p.eb].x;
  return mu;
}

// a * b + c with 24-bit a, b, at full rate (the compiler otherwise picks the
// quarter-rate 64-bit multiply-add when it cannot prove the operand widths)
__device__ inline uint32_t mad_u24(uint32_t a, uint32_t b, uint32_t c) {
  uint32_t r;
  asm("v_mad_u32_u24 %0, %1, %2, %3" : "=v"(r) : "v"(a), "s"(b), "v"(c));
  return r;
}

// min(max(x, 0), hi) for hi >= 0 as one median-of-three (the compiler forms
// it only when hi is a constant)
__device__ inline int clamp_index(int x, int hi) {
  int r;
  asm("v_med3_i32 %0, %1, 0, %2" : "=v"(r) : "v"(x), "s"(hi));
  return r;
}

// Position at flight parameter t and the index of its voxel.
template <class JT>
__device__ inline uint32_t voxel_at(const Photon& p, float t, const JT& J, float& qx, float& qy, float& qz) {
  qx = fmaf(t, p.dx, p.px); qy = fmaf(t, p.dy, p.py); qz = fmaf(t, p.dz, p.pz);
  // voxel index = (q - b0) / v as one FMA; truncation equals floor wherever
  // the clamp to [0, n - 1] does not already decide it
  const int ix = clamp_index((int)fmaf(qx, J.inv_vx, J.ix0), J.nx - 1);
  const int iy = clamp_index((int)fmaf(qy, J.inv_vy, J.iy0), J.ny - 1);
  const int iz = clamp_index((int)fmaf(qz, J.inv_vz, J.iz0), J.nz - 1);
  // 32-bit index from two 24-bit multiply-adds: the host checks
  // nz * ny < 2^24, nx < 2^24 and nx * ny * nz < 2^32
  return mad_u24(mad_u24((uint32_t)iz, (uint32_t)J.ny, (uint32_t)iy), (uint32_t)J.nx, (uint32_t)ix);
}

// Label byte of a voxel, by a global-address-space load; the caller clamps
// labels past the table (min(label, kMaxLabels - 1)) so no table read leaves it.
template <class JT>
__device__ inline uint32_t label_raw(uint32_t i, const JT& J) {
  return ((const __attribute__((address_space(1))) uint8_t*)J.labels)[i];
}

// A parked branch runs once at least this many lanes wait for it (or nothing
// else can run).  Interactions wait for 32 lanes: each batch costs a whole
// Klein-Nishina rejection loop and a flight set-up, so fuller batches pay
// (at 6 waves per SIMD, 16, 24 and 32 lanes took 96.7, 94.6 and 94.2 ms on the
// full workload at 1e7 photons per view; on the current kernel 40 and 48 lanes
// took 86.3 and 103.6 against 83.5 ms, the collided lanes then idling longer).
constexpr unsigned kInteractBatch = 32;
constexpr unsigned kRefillBatch = 16;

__global__ __launch_bounds__(256) void mc_transport_kernel(const McJob* __restrict__ jobs, int n_jobs,
                                                           uint64_t total_units, uint32_t chunk,
                                                           unsigned int* __restrict__ queue,
                                                           McCounters* __restrict__ counters) {
  int job = -1;
  const int q_first = (int)(blockIdx.x % kMcQueues);
  int q_tried = 0;  // queues found spent (the wave then draws from the next)
  // the current unit (wave-uniform): tiles [unit_t0, unit_t0 + unit_nt) over photon
  // cycles [unit_cyc, unit_cyc_end); the next birth takes tile unit_t0 + unit_g of
  // cycle unit_cyc
  uint32_t unit_t0 = 0, unit_nt = 0, unit_g = 0, unit_cyc = 0, unit_cyc_end = 0;
  int unit_view = 0;
  long long pending = -1;
  bool exhausted = false;
  uint32_t buf_head = 0, buf_n = 0;  // the birth buffer's next photon and photons left (wave-uniform)
  int buf_view = 0;                  // the view of the buffered photons, and its cos and sin
  float buf_cv = 0.f, buf_sv = 0.f;
  uint32_t buf_id0 = 0;              // photon index of the buffer's slot 0 (the tile's corner pixel)
  uint32_t ray_tile = ~0u;           // the tile and view whose rays the buffer holds (~0u: none)
  int ray_view = -1;

  int state = kDead;
  Photon p{};
  Stream rs{};
  const int lane = __lane_id();
  // The job descriptors do not change during the launch: read them through the
  // constant address space, so wave-uniform fields come in by scalar loads.
  const CJob* cjobs = (const CJob*)jobs;
  extern __shared__ WaveTables wave_tables[];  // one slot per wave of the block
  WaveTables& wt = wave_tables[threadIdx.x >> 6];
  if (lane < kStats) wt.stats[lane] = 0;
  // adds the wave's counters to the job's and clears them; called once all the
  // histories the wave began for the job have ended, so the misses and the
  // interactions follow from the other counts
  auto flush = [&](McCounters* dst) {
    if (__lane_id() == 0) {
      unsigned long long v[kStats];
#pragma unroll
      for (int k = 0; k < kStats; ++k) {
        v[k] = wt.stats[k];
        wt.stats[k] = 0;
      }
      v[kStatMissed] = v[kStatHistories] - v[kStatTp] - v[kStatTs] - v[kStatPhoto] - v[kStatCapped];
      v[kStatInteractions] = v[kStatPhoto] + v[kStatCoherent] + v[kStatCompton];
      atomicAdd(&dst->histories, v[kStatHistories]); atomicAdd(&dst->steps, v[kStatSteps]);
      atomicAdd(&dst->interactions, v[kStatInteractions]); atomicAdd(&dst->photo, v[kStatPhoto]);
      atomicAdd(&dst->coherent, v[kStatCoherent]); atomicAdd(&dst->compton, v[kStatCompton]);
      atomicAdd(&dst->tallied_primary, v[kStatTp]); atomicAdd(&dst->tallied_scatter, v[kStatTs]);
      atomicAdd(&dst->missed, v[kStatMissed]); atomicAdd(&dst->capped, v[kStatCapped]);
    }
  };

  for (;;) {
    // fields are read where they are used, by scalar loads
    CJob* Jp = cjobs + __builtin_amdgcn_readfirstlane(job < 0 ? 0 : job);
    uint64_t dead = __ballot(state <= kTally);
    const uint64_t fly = __ballot(state == kFly);
    const uint64_t coll = __ballot(state == kCollide);
    const uint64_t pend = __ballot(state == kTally);
    const bool all_dead = (fly | coll) == 0;

    // 1) secure a work unit when lanes are idle and the current one is spent
    if (dead != 0 && unit_cyc >= unit_cyc_end) {
      if (pending < 0 && !exhausted) {
        // in round r = k / chunk, queue q hands out chunk r kMcQueues + (q + r) mod kMcQueues,
        // unit k mod chunk of it: the chunks of a round are dealt to the queues rotated by
        // one per round, so no queue keeps the same detector rows when kMcQueues chunks
        // span a whole number of detector sweeps.  Units increase with k, so a queue is
        // spent once it hands out a unit past the end.
        while (q_tried < kMcQueues) {
          const int q = (q_first + q_tried) % kMcQueues;
          unsigned int k = 0;
          if (lane == 0) k = atomicAdd(queue + q * kMcQueueStride, 1u);
          k = (unsigned int)__builtin_amdgcn_readfirstlane((int)k);
          const uint32_t r = k / chunk;
          const uint64_t u = ((uint64_t)r * kMcQueues + (uint64_t)((q + r) % kMcQueues)) * chunk + (k - r * chunk);
          if (u < total_units) {
            pending = (long long)u;
            break;
          }
          ++q_tried;
        }
        if (pending < 0) exhausted = true;
      }
      if (pending >= 0) {
        const int pj = __builtin_amdgcn_readfirstlane(job_of_unit(cjobs, n_jobs, (uint64_t)pending));
        if (pj == job || (all_dead && buf_n == 0 && pend == 0)) {  // buffered and ended photons are of the current job
          const bool switched = pj != job;
          if (switched) {
            if (job >= 0) flush(counters + job);
            job = pj;
            ray_tile = ~0u;  // another job's geometry
          }
          // units run views fastest, then cycle blocks, then tile groups: the units in
          // flight share a few detector tiles, hence a thin z-slab of the phantom that
          // stays in L2, and each unit's primary tallies stay on a few lines
          const uint64_t local = (uint64_t)pending - cjobs[pj].unit_begin;
          const uint32_t nvs = (uint32_t)cjobs[pj].n_views_sim;
          const uint32_t bi = (uint32_t)(local / nvs);
          const uint32_t vl = (uint32_t)(local - (uint64_t)bi * nvs);
          unit_view = __builtin_amdgcn_readfirstlane(cjobs[pj].view_begin + (int)vl);
          {
            const uint32_t ncb = cjobs[pj].cycle_blocks, grp = cjobs[pj].tile_group;
            const uint32_t tg = bi / ncb, cb = bi - tg * ncb;
            unit_t0 = (uint32_t)__builtin_amdgcn_readfirstlane((int)(tg * grp));
            unit_nt = (uint32_t)__builtin_amdgcn_readfirstlane((int)min(grp, cjobs[pj].tiles - tg * grp));
            unit_g = 0;
            unit_cyc = (uint32_t)__builtin_amdgcn_readfirstlane((int)(cb * cjobs[pj].cycle_block));
            unit_cyc_end = (uint32_t)__builtin_amdgcn_readfirstlane(
                (int)min(unit_cyc + cjobs[pj].cycle_block, cjobs[pj].cycles));
          }
          pending = -1;
          if (switched) {  // J is re-read for the new job at the top of the loop
            load_wave_tables(wt, cjobs[pj], lane);
            __builtin_amdgcn_wave_barrier();
            continue;
          }
        }
      }
    }
    // 1b) tallies of the flights that ended, batched with the refills (all run when
    //     nothing flies, so none is left when the wave stops or changes job)
    if (pend != 0 && (popc(dead) >= kRefillBatch || fly == 0)) {
      int code = -1;
      if (state == kTally) {
        code = tally(p, *Jp);
        state = kDead;
      }
      count_tallies(wt, code);
    }
    if (all_dead && unit_cyc >= unit_cyc_end && pending < 0 && exhausted && buf_n == 0) break;

    // 2) refill dead lanes from the birth buffer, after filling it when it is empty
    if (dead != 0 && (buf_n != 0 || unit_cyc < unit_cyc_end) && (popc(dead) >= kRefillBatch || fly == 0)) {
      if (buf_n == 0) {
        // one tile of the unit in one cycle, one photon per lane with every lane
        // active: lane l aims at the centre of pixel (u_, v_) of the tile
        // (CBCT_real325.cu:464-466), and is photon n = cycle * nu nv + v_ nu + u_ of
        // the view, whose pixel in aim order is n mod (nu nv)
        const uint32_t t = unit_t0 + unit_g;
        const uint32_t tv = t / (uint32_t)Jp->tiles_u, tu = t - tv * (uint32_t)Jp->tiles_u;
        const uint32_t u_ = tu * kTileU + (uint32_t)lane % kTileU, v_ = tv * kTileV + (uint32_t)lane / kTileU;
        const uint32_t npix = (uint32_t)(Jp->nu * Jp->nv);
        const uint32_t pix = v_ * (uint32_t)Jp->nu + u_;
        const uint64_t n64 = (uint64_t)unit_cyc * npix + pix;
        const uint32_t id = (uint32_t)n64;
        CU32* vcs = (CU32*)Jp->view_cs + 2 * unit_view;  // read-only during the launch: scalar loads
        const float cv = __uint_as_float(vcs[0]), sv = __uint_as_float(vcs[1]);
        const bool on_det = u_ < (uint32_t)Jp->nu && v_ < (uint32_t)Jp->nv;
        if (t != ray_tile || unit_view != ray_view) {
          // the tile's rays in this view, one per lane: the first flight runs from the
          // source towards the pixel centre, its majorant comes from the labels the
          // line crosses, and it starts at the flight box (start_flight's clip)
          const uint32_t set =
              on_det && Jp->ray_set ? (uint32_t)Jp->ray_set[(size_t)(unit_view - Jp->view_begin) * npix + pix]
                                    : kAnyLabel;
          const float yp = ((float)u_ + 0.5f - Jp->un0) * Jp->u_step;
          const float zp = ((float)v_ + 0.5f - Jp->vn0) * Jp->v_step;
          const float in = rsqrtf(fmaf(zp, zp, fmaf(Jp->sdd, Jp->sdd, yp * yp)));
          // the fused products are spelled out, so the direction's rounding does not
          // depend on which product the compiler would fuse here
          const float dx = fmaf(Jp->sdd, cv, -(yp * sv)) * in;
          const float dy = fmaf(cv, yp, Jp->sdd * sv) * in;
          const float dz = zp * in;
          float t0, t1;
          const bool hit = ray_box(-Jp->sod * cv, -Jp->sod * sv, 0.f, dx, dy, dz, *Jp, t0, t1) && t1 > 0.f;
          float tin = 0.f;
          if (t0 > 0.f) {
            t1 -= t0;
            tin = t0;
          }
          wt.bb.d0[lane] = make_float4(dx, dy, dz, tin);
          wt.bb.t1[lane] = hit ? t1 : -1.f;
          wt.bb.set[lane] = (uint8_t)set;
          ray_tile = t;
          ray_view = unit_view;
        }
        bool born = false;
        int bcode = -1;  // tally code of a photon that misses the flight box
        const bool valid = on_det && n64 < Jp->photons_per_view;
        if (valid) {
          Photon q;
          q.view = unit_view;
          q.cv = cv;
          q.sv = sv;
          {
            const float4 d0 = wt.bb.d0[lane];
            q.dx = d0.x; q.dy = d0.y; q.dz = d0.z;
          }
          q.px = -Jp->sod * cv; q.py = -Jp->sod * sv; q.pz = 0.f;
          float e = Jp->e_kev;
          if (Jp->spectrum_cdf) {
            Stream bs{};
            bs.photon = id; bs.view = (uint32_t)unit_view;
            const float u = next_u01(bs, Jp->job_id, Jp->key0, Jp->key1);
            const int g = (int)(u * (float)kGuide);  // exact: u < 1 and kGuide is a power of 2
            e = (float)spectrum_row(wt.cdf, u, wt.guide[g], wt.guide[g + 1]);
          }
          q.e = e;
          q.eb = ebin_of(e);
          q.nint = 0;
          // start_flight's tests: something attenuates, and the line meets the box
          q.mu0 = wt.mu[0][q.eb];
          q.mu1 = wt.mu[1][q.eb];
          q.mu2 = wt.mu[2][q.eb];
          q.mu3 = wt.mu[3][q.eb];
          born = set_majorant(q, wt, wt.bb.set[lane]) > 0.f && wt.bb.t1[lane] >= 0.f;
          if (born) {
            wt.bb.eb[lane] = (uint8_t)q.eb;
          } else {
            bcode = tally(q, *Jp);
          }
        }
        count_lanes(wt, kStatHistories, valid);
        count_tallies(wt, bcode);
        const uint64_t bm = __ballot(born);
        if (born)
          wt.bb.slot[__builtin_amdgcn_mbcnt_hi((uint32_t)(bm >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)bm, 0u))] =
              (uint8_t)lane;
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");  // the buffer in LDS is complete
        __builtin_amdgcn_wave_barrier();
        buf_n = (uint32_t)popc(bm);
        buf_head = 0;
        buf_view = unit_view;
        buf_cv = cv;
        buf_sv = sv;
        buf_id0 = (uint32_t)((uint64_t)unit_cyc * npix + tv * kTileV * (uint32_t)Jp->nu + tu * kTileU);
        if (++unit_g == unit_nt) {  // the unit's tiles are done for this cycle
          unit_g = 0;
          ++unit_cyc;
        }
      }
      const uint32_t rank = __builtin_amdgcn_mbcnt_hi((uint32_t)(dead >> 32),
                                                      __builtin_amdgcn_mbcnt_lo((uint32_t)dead, 0u));
      if (state == kDead && rank < buf_n) {
        const int s = wt.bb.slot[buf_head + rank];
        // as the birth formed them (start_flight's box clip, majorant and scale)
        const float4 d0 = wt.bb.d0[s];
        p.dx = d0.x; p.dy = d0.y; p.dz = d0.z;
        const float tin = d0.w;
        p.px = fmaf(tin, p.dx, -Jp->sod * buf_cv);
        p.py = fmaf(tin, p.dy, -Jp->sod * buf_sv);
        p.pz = fmaf(tin, p.dz, 0.f);
        p.eb = wt.bb.eb[s];
        p.e = Jp->spectrum_cdf ? (float)p.eb : Jp->e_kev;  // a spectrum row is its energy in keV
        p.t = 0.f;
        p.t1 = wt.bb.t1[s];
        p.mu0 = wt.mu[0][p.eb];
        p.mu1 = wt.mu[1][p.eb];
        p.mu2 = wt.mu[2][p.eb];
        p.mu3 = wt.mu[3][p.eb];
        p.mumax = set_majorant(p, wt, wt.bb.set[s]);
        p.lscale = 0.69314718055994531f * rcp(p.mumax);
        p.nint = 0;
        p.view = buf_view;
        p.cv = buf_cv;
        p.sv = buf_sv;
        // the spectrum draw took block 0 of the photon's stream; flights take whole blocks
        rs.ctr = Jp->spectrum_cdf ? 1u : 0u;
        rs.photon = buf_id0 + (uint32_t)(s / kTileU) * (uint32_t)Jp->nu + (uint32_t)(s % kTileU);
        rs.view = (uint32_t)buf_view;
        rs.left = 0;
        state = kFly;
      }
      const uint32_t took = min((uint32_t)popc(dead), buf_n);
      buf_head += took;
      buf_n -= took;
    }
    // 3) interactions of the parked colliding lanes
    if (coll != 0 && (popc(coll) >= kInteractBatch || fly == 0)) {
      const bool cap = state == kCollide && p.nint >= Jp->max_int;
      count_lanes(wt, kStatCapped, cap);
      if (cap) state = kDead;
      int ev = 0;      // 1 photoelectric, 2 coherent, 3 Compton
      if (state == kCollide) {
        const float4 m = Jp->tables->m[p.lab][p.eb];
        rs.left = 0;  // interactions start on a fresh block
        const float u3 = next_u01(rs, Jp->job_id, Jp->key0, Jp->key1);
        asm volatile("" ::"v"(m.y), "v"(m.z));  // both cutoffs loaded and waited for together
        const bool photo = u3 <= m.y, coh = u3 <= m.z;
        if (photo) {
          ev = 1;
          state = kDead;
        } else {
          if (coh) {
            ev = 2;
            if (Jp->coherent) {
              const float cth = thomson_mu(next_u01(rs, Jp->job_id, Jp->key0, Jp->key1));
              const float w = next_u01(rs, Jp->job_id, Jp->key0, Jp->key1);  // azimuth / 2 pi
              rotate_dir(p.dx, p.dy, p.dz, cth, w);
            }
          } else {
            ev = 3;
            const float lam = 511.0f * rcp(p.e);
            const float ro = kahn(lam, rs, Jp->job_id, Jp->key0, Jp->key1);
            const float lamd = ro * lam;
            float cth = 1.0f - (lamd - lam);
            cth = cth < -1.f ? -1.f : cth;
            p.e = 511.0f * rcp(lamd);
            p.eb = ebin_of(p.e);
            const float w = next_u01(rs, Jp->job_id, Jp->key0, Jp->key1);  // azimuth / 2 pi
            rotate_dir(p.dx, p.dy, p.dz, cth, w);
          }
          p.nint += 1;
          float tin_;
          if (start_flight(p, *Jp, wt, kAnyLabel, tin_)) {
            state = kFly;
          } else {
            state = kTally;
          }
        }
      }
      count_lanes(wt, kStatPhoto, ev == 1);
      count_lanes(wt, kStatCoherent, ev == 2);
      count_lanes(wt, kStatCompton, ev == 3);
    }

    // 4) two Woodcock steps for every flying lane from one Philox block
    //    (distance, accept, distance, accept); both label loads are issued
    //    before either is used, and a flight ending on the first step leaves
    //    the block's other two uniforms unused.
    count_lanes(wt, kStatSteps, state == kFly);
    bool two = false;   // took the second step of the pair
    if (state == kFly) {
      uint32_t c0 = rs.ctr, c1 = rs.photon, c2 = rs.view, c3 = Jp->job_id;
      philox10(c0, c1, c2, c3, Jp->key0, Jp->key1);
      rs.ctr += 1;
      rs.left = 0;
      // u >= 2^-25, so the hardware log2 needs no denormal handling
      const float ta = fmaf(-__builtin_amdgcn_logf(u01(c0)), p.lscale, p.t);
      const float tb = fmaf(-__builtin_amdgcn_logf(u01(c2)), p.lscale, ta);
      float ax, ay, az, bx, by, bz;
      const uint32_t va = voxel_at(p, ta, *Jp, ax, ay, az);
      const uint32_t vb = voxel_at(p, tb, *Jp, bx, by, bz);
      // a step that leaves the flight box (most first flights end on their first
      // step) reads no label: its load would only add traffic to the line misses.
      // Both loads are issued before either value is used (the empty asm keeps
      // the clamps, and so the waits, after both).
      uint32_t ra = 0, rb = 0;
      if (ta < p.t1) ra = label_raw(va, *Jp);
      if (tb < p.t1) rb = label_raw(vb, *Jp);
      asm volatile("" : "+v"(ra), "+v"(rb));
      const int la = min((int)ra, kMaxLabels - 1), lb = min((int)rb, kMaxLabels - 1);
      const float mua = mu_at(p, la, *Jp), mub = mu_at(p, lb, *Jp);
      // Both attenuations are consumed here on every path (an empty asm taking
      // them as inputs): a table load for a label above 3 left pending on the
      // paths that end the flight made the compiler wait for every memory
      // operation, tally atomics included, at the start of the next step.
      asm volatile("" ::"v"(mua), "v"(mub));
      const bool real_a = u01(c1) * p.mumax < mua, real_b = u01(c3) * p.mumax < mub;
      if (ta >= p.t1) {
        state = kTally;
      } else if (real_a) {  // real collision: park until the wave runs interactions
        p.px = ax; p.py = ay; p.pz = az;
        p.lab = la;
        state = kCollide;
      } else {
        two = true;
        if (tb >= p.t1) {
          state = kTally;
        } else if (real_b) {
          p.px = bx; p.py = by; p.pz = bz;
          p.lab = lb;
          state = kCollide;
        } else {
          p.t = tb;
        }
      }
    }
    count_lanes(wt, kStatSteps, two);
  }
  if (job >= 0) flush(counters + job);
}

// --------------------------------------------- first-flight label sets
// (xadd, xsub and xmul are never fused into FMAs, so the label sets follow
// the oracle's operations bit for bit.)

__device__ inline uint32_t label_bit(int lab) { return lab < 4 ? 1u << lab : kAnyLabel; }

// One thread per brick: the labels of the brick's voxels and of their
// neighbours one voxel outside it, so a sample point rounded across a brick
// face into a voxel of the next brick is still covered.
__global__ __launch_bounds__(256) void brick_sets_kernel(const uint8_t* __restrict__ labels, int nx, int ny, int nz,
                                                         int b, int nb0, int nb1, int nb2,
                                                         uint8_t* __restrict__ bricks) {
  const long i = (long)blockIdx.x * 256 + threadIdx.x;
  if (i >= (long)nb0 * nb1 * nb2) return;
  const int bx = (int)(i % nb0), by = (int)((i / nb0) % nb1), bz = (int)(i / ((long)nb0 * nb1));
  const int x0 = max(bx * b - 1, 0), x1 = min(bx * b + b, nx - 1);
  const int y0 = max(by * b - 1, 0), y1 = min(by * b + b, ny - 1);
  const int z0 = max(bz * b - 1, 0), z1 = min(bz * b + b, nz - 1);
  uint32_t m = 0;
  for (int z = z0; z <= z1; ++z)
    for (int y = y0; y <= y1; ++y) {
      const uint8_t* row = labels + ((size_t)z * ny + y) * nx;
      for (int x = x0; x <= x1; ++x) m |= label_bit(row[x]);
    }
  bricks[i] = (uint8_t)m;
}

// One thread per (view, pixel): the union of the brick sets along the segment
// source -> pixel centre inside the flight box, by index-stepping traversal of
// the brick grid (every crossing recomputed from its plane index), or
// kAnyLabel if the segment misses the box.
__global__ __launch_bounds__(256) void ray_sets_kernel(RayGrid R, const double* __restrict__ cs64, int view_begin,
                                                       int n_views, const uint8_t* __restrict__ bricks,
                                                       uint8_t* __restrict__ sets) {
  const long npix = (long)R.nu * R.nv;
  const long i = (long)blockIdx.x * 256 + threadIdx.x;
  if (i >= npix * n_views) return;
  const int kk = (int)(i / npix);
  const int pix = (int)(i - (long)kk * npix);
  const int iv = pix / R.nu, iu = pix - iv * R.nu;
  const double cb = cs64[2 * (view_begin + kk)], sb = cs64[2 * (view_begin + kk) + 1];
  const double yp = xmul(xsub(xadd((double)iu, 0.5), R.un0), R.u_step);
  const double zp = xmul(xsub(xadd((double)iv, 0.5), R.vn0), R.v_step);
  const double S[3] = {xmul(-R.sod, cb), xmul(-R.sod, sb), 0.0};
  const double P[3] = {xsub(xmul(R.xd, cb), xmul(yp, sb)), xadd(xmul(R.xd, sb), xmul(yp, cb)), zp};
  double d[3], inv[3];
  double lo = 0.0, hi = 1.0;
  uint32_t m = 0;
  bool miss = false;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    d[a] = xsub(P[a], S[a]);
    if (d[a] != 0.0) {
      inv[a] = 1.0 / d[a];
      const double ta = xmul(xsub(R.c0[a], S[a]), inv[a]), tb = xmul(xsub(R.c1[a], S[a]), inv[a]);
      const double tl = ta < tb ? ta : tb, th = ta < tb ? tb : ta;
      if (tl > lo) lo = tl;
      if (th < hi) hi = th;
    } else {
      inv[a] = 0.0;
      if (S[a] < R.c0[a] || S[a] >= R.c1[a]) miss = true;
    }
  }
  if (miss || !(hi > lo)) {
    sets[i] = (uint8_t)kAnyLabel;
    return;
  }
  const double span = xsub(hi, lo);
  const double half = xmul(span, 0.5);
  const double tm = xadd(lo, half < 1e-9 ? half : 1e-9);  // a point just inside
  int idx[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const double q = xadd(S[a], xmul(tm, d[a]));
    const int j = (int)floor(xsub(q, R.org[a]) / R.bw[a]);
    idx[a] = j < 0 ? 0 : (j > R.nb[a] - 1 ? R.nb[a] - 1 : j);
  }
  for (int guard = 0; guard < 4 * (R.nb[0] + R.nb[1] + R.nb[2]) + 8; ++guard) {
    m |= bricks[((size_t)idx[2] * R.nb[1] + idx[1]) * R.nb[0] + idx[0]];
    double tn[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      if (d[a] > 0.0) tn[a] = xmul(xsub(xadd(R.org[a], xmul((double)(idx[a] + 1), R.bw[a])), S[a]), inv[a]);
      else if (d[a] < 0.0) tn[a] = xmul(xsub(xadd(R.org[a], xmul((double)idx[a], R.bw[a])), S[a]), inv[a]);
      else tn[a] = INFINITY;
    }
    int a = 0;
    if (tn[1] < tn[a]) a = 1;
    if (tn[2] < tn[a]) a = 2;
    if (!(tn[a] < hi)) break;
    idx[a] += d[a] > 0.0 ? 1 : -1;
    if (idx[a] < 0 || idx[a] >= R.nb[a]) break;
  }
  sets[i] = (uint8_t)m;
}

hipError_t launch_brick_sets(const uint8_t* labels, int nx, int ny, int nz, int b, const int nb[3], uint8_t* bricks,
                             hipStream_t s) {
  const long n = (long)nb[0] * nb[1] * nb[2];
  if (n <= 0 || b < 1) return hipErrorInvalidValue;
  hipLaunchKernelGGL(brick_sets_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, labels, nx, ny, nz, b,
                     nb[0], nb[1], nb[2], bricks);
  return hipGetLastError();
}

hipError_t launch_ray_sets(const RayGrid& R, const double* cs64, int view_begin, int n_views, const uint8_t* bricks,
                           uint8_t* sets, hipStream_t s) {
  const long n = (long)R.nu * R.nv * n_views;
  if (n <= 0) return hipSuccess;
  if ((n + 255) / 256 > 0x7fffffffL) return hipErrorInvalidValue;
  hipLaunchKernelGGL(ray_sets_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, R, cs64, view_begin,
                     n_views, bricks, sets);
  return hipGetLastError();
}

__global__ void mc_reset_kernel(unsigned int* queue, McCounters* counters, int n_jobs) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < kMcQueues) queue[i * kMcQueueStride] = 0u;
  if (i < n_jobs) counters[i] = McCounters{};
}

// Bounding box of the voxels whose label attenuates at some energy
// (`solid` has bit l set for such labels; labels are clamped as in transport):
// box = {x_min, y_min, z_min, x_max, y_max, z_max}, initialised by the caller
// to {INT_MAX x 3, -1 x 3}.  One block per z slice, reduced in the block, so
// the box takes six atomics per slice.
__global__ __launch_bounds__(256) void support_box_kernel(const uint8_t* __restrict__ labels, int nx, int ny, int nz,
                                                          uint32_t solid, int* __restrict__ box) {
  __shared__ int red[4][4];
  const int z = blockIdx.x;
  const uint8_t* slice = labels + (size_t)z * ny * nx;
  int xmin = INT_MAX, xmax = -1, ymin = INT_MAX, ymax = -1;
  for (int y = 0; y < ny; ++y) {
    for (int x = threadIdx.x; x < nx; x += 256) {
      const int lab = min((int)slice[(size_t)y * nx + x], kMaxLabels - 1);
      if ((solid >> lab) & 1u) {
        xmin = min(xmin, x);
        xmax = max(xmax, x);
        ymin = min(ymin, y);
        ymax = y;
      }
    }
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    xmin = min(xmin, __shfl_xor(xmin, o, 64));
    xmax = max(xmax, __shfl_xor(xmax, o, 64));
    ymin = min(ymin, __shfl_xor(ymin, o, 64));
    ymax = max(ymax, __shfl_xor(ymax, o, 64));
  }
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    red[w][0] = xmin; red[w][1] = xmax; red[w][2] = ymin; red[w][3] = ymax;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int i = 1; i < 4; ++i) {
      xmin = min(xmin, red[i][0]); xmax = max(xmax, red[i][1]);
      ymin = min(ymin, red[i][2]); ymax = max(ymax, red[i][3]);
    }
    if (xmax >= 0) {
      atomicMin(box + 0, xmin); atomicMin(box + 1, ymin); atomicMin(box + 2, z);
      atomicMax(box + 3, xmax); atomicMax(box + 4, ymax); atomicMax(box + 5, z);
    }
  }
}

hipError_t launch_support_box(const uint8_t* labels, int nx, int ny, int nz, uint32_t solid, int* box, hipStream_t s) {
  if (nx < 1 || ny < 1 || nz < 1) return hipErrorInvalidValue;
  hipLaunchKernelGGL(support_box_kernel, dim3(nz), dim3(256), 0, s, labels, nx, ny, nz, solid, box);
  return hipGetLastError();
}

hipError_t launch_mc_reset(unsigned int* queue, McCounters* counters, int n_jobs, hipStream_t s) {
  const int n = n_jobs > 1 ? n_jobs : 1;
  hipLaunchKernelGGL(mc_reset_kernel, dim3((n + 255) / 256), dim3(256), 0, s, queue, counters, n_jobs);
  return hipGetLastError();
}

hipError_t launch_mc_transport(const McJob* jobs_dev, int n_jobs, uint64_t total_units, uint32_t chunk,
                               unsigned int* queue, McCounters* counters, int block, int grid, hipStream_t s) {
  if (block % 64 || block > 256 || block <= 0 || grid <= 0 || chunk == 0) return hipErrorInvalidValue;
  hipLaunchKernelGGL(mc_transport_kernel, dim3(grid), dim3(block), sizeof(WaveTables) * (block / 64), s, jobs_dev,
                     n_jobs, total_units, chunk, queue, counters);
  return hipGetLastError();
}

// ------------------------------------------------- exact primary (Siddon)
// One thread per (view, pixel): per-label path lengths along the source ->
// pixel-centre ray by index-stepping traversal (each crossing recomputed from
// its plane index, no running sums), kept in LDS, then combined with the
// attenuation of each spectral bin.  This is a deterministic integral, not MC
// tracking, and it runs in FP64 end to end: in FP32 the ray set-up from a
// source tens of cm away moves grazing rays by ~1e-6 cm, which changes short
// chords by ~1e-4 relative.
__global__ __launch_bounds__(256) void primary_exact_kernel(PrimaryGeom P, const double* __restrict__ cs64,
                                                            const uint8_t* __restrict__ labels,
                                                            const McTables* __restrict__ tables, float e_kev,
                                                            const float* __restrict__ spec_w, int use_spec,
                                                            float* __restrict__ out) {
  __shared__ double lens[kMaxLabels][256];
  const int tid = threadIdx.x;
  const long ray = (long)blockIdx.x * 256 + tid;
  const long nray = (long)P.n_views * P.nu * P.nv;
#pragma unroll
  for (int l = 0; l < kMaxLabels; ++l) lens[l][tid] = 0.0;
  if (ray >= nray) return;
  const int k = (int)(ray / ((long)P.nu * P.nv));
  const int pix = (int)(ray - (long)k * P.nu * P.nv);
  const int iu = pix / P.nv, iv = pix - iu * P.nv;
  const double cb = cs64[2 * k], sb = cs64[2 * k + 1];
  const double yp = ((double)iu + 0.5 - P.un0) * P.u_step;
  const double zp = ((double)iv + 0.5 - P.vn0) * P.v_step;
  const double xd = P.sdd - P.sod;
  const double p0[3] = {-P.sod * cb, -P.sod * sb, 0.0};
  double d[3] = {xd * cb - yp * sb - p0[0], xd * sb + yp * cb - p0[1], zp};
  const double L = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
  for (int a = 0; a < 3; ++a) d[a] /= L;
  double t0 = -INFINITY, t1 = INFINITY;
  bool hit = true;
  for (int a = 0; a < 3; ++a) {
    if (d[a] != 0.0) {
      double ta = (P.b0[a] - p0[a]) / d[a], tb = (P.b1[a] - p0[a]) / d[a];
      if (ta > tb) { const double tt = ta; ta = tb; tb = tt; }
      t0 = fmax(t0, ta);
      t1 = fmin(t1, tb);
    } else if (p0[a] < P.b0[a] || p0[a] >= P.b1[a]) {
      hit = false;
    }
  }
  t0 = fmax(t0, 0.0);
  t1 = fmin(t1, L);
  if (hit && t1 > t0) {
    int idx[3], step[3];
    double sn[3];
    const double tm = t0 + fmin(0.5 * (t1 - t0), 1e-9);
    for (int a = 0; a < 3; ++a) {
      int i = (int)floor((p0[a] + tm * d[a] - P.b0[a]) / P.vox[a]);
      i = min(max(i, 0), P.n[a] - 1);
      idx[a] = i;
      if (d[a] > 0.0) {
        step[a] = 1;
        sn[a] = (P.b0[a] + (i + 1) * P.vox[a] - p0[a]) / d[a];
      } else if (d[a] < 0.0) {
        step[a] = -1;
        sn[a] = (P.b0[a] + i * P.vox[a] - p0[a]) / d[a];
      } else {
        step[a] = 0;
        sn[a] = INFINITY;
      }
    }
    double t = t0;
    for (int guard = 0; guard < 4 * (P.n[0] + P.n[1] + P.n[2]) + 8; ++guard) {
      int a = 0;
      if (sn[1] < sn[a]) a = 1;
      if (sn[2] < sn[a]) a = 2;
      const double te = fmin(sn[a], t1);
      if (te > t) {
        // labels past the table read as the last label, as in transport
        const int lab = min((int)labels[((size_t)idx[2] * P.n[1] + idx[1]) * P.n[0] + idx[0]], kMaxLabels - 1);
        lens[lab][tid] += te - t;
        t = te;
      }
      if (sn[a] >= t1) break;
      idx[a] += step[a];
      if (idx[a] < 0 || idx[a] >= P.n[a]) break;
      sn[a] = (P.b0[a] + (step[a] > 0 ? idx[a] + 1 : idx[a]) * P.vox[a] - p0[a]) / d[a];
    }
  }
  double result;
  if (!use_spec) {
    const int eb = ebin_of(e_kev);
    double acc = 0.0;
    for (int l = 0; l < kMaxLabels; ++l) {
      const double len = lens[l][tid];
      if (len != 0.0) acc += len * (double)tables->m[l][eb].x;
    }
    result = acc;
  } else {
    double num = 0.0, den = 0.0;
    for (int eb = 1; eb < kRows; ++eb) {
      const double w = spec_w[eb];
      if (w <= 0.0) continue;
      double le = 0.0;
      for (int l = 0; l < kMaxLabels; ++l) {
        const double len = lens[l][tid];
        if (len != 0.0) le += len * (double)tables->m[l][eb].x;
      }
      num += w * exp(-le);
      den += w;
    }
    result = -log(num / den);
  }
  out[ray] = (float)result;
}

hipError_t launch_primary_exact(const PrimaryGeom& P, const double* cs64, const uint8_t* labels,
                                const McTables* tables, float e_kev, const float* spec_w, int use_spec, float* out,
                                hipStream_t s) {
  const long nray = (long)P.n_views * P.nu * P.nv;
  if (nray <= 0) return hipSuccess;
  const long blocks = (nray + 255) / 256;
  hipLaunchKernelGGL(primary_exact_kernel, dim3((unsigned)blocks), dim3(256), 0, s, P, cs64, labels, tables, e_kev,
                     spec_w, use_spec, out);
  return hipGetLastError();
}

// ------------------------------------------------- counts -> line integrals
// CBCT_real325.cu:267-294: clamp N to [1, N0], L = ln N0 - ln N, with N0 the
// photons aimed at the pixel: floor(N/npix) + (a < N mod npix), a = iv nu + iu
// the pixel's place in the aim order.
// `counts2`, when given, is added to `counts` first (primary + scatter tallies
// in one pass).
__global__ void counts_to_lineint_kernel(const double* __restrict__ counts, const double* __restrict__ counts2,
                                         float* __restrict__ out, int n_views, int nu, int nv,
                                         uint64_t photons_per_view) {
  const int npix = nu * nv;
  const size_t n = (size_t)n_views * npix;
  const uint64_t base = photons_per_view / (uint64_t)npix, rem = photons_per_view % (uint64_t)npix;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const int pix = (int)(i % (size_t)npix);  // iu nv + iv
    const int iu = pix / nv, iv = pix - iu * nv;
    const int a = iv * nu + iu;
    const double n0 = (double)(base + ((uint64_t)a < rem ? 1 : 0));
    double c = counts2 ? counts[i] + counts2[i] : counts[i];
    c = c > n0 ? n0 : c;
    c = c < 1.0 ? 1.0 : c;
    out[i] = n0 > 0.0 ? (float)(log(n0) - log(c)) : 0.f;
  }
}

// Against a measured flat field (air scan) flat [flat_views][nu][nv], flat_views 1 or
// n_views: ln(F) - ln(clamp(N, 1, F)), 0 where F <= 0.
__global__ void counts_to_lineint_flat_kernel(const double* __restrict__ counts, const double* __restrict__ flat,
                                              float* __restrict__ out, size_t n, size_t flat_n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const double f = flat[i % flat_n];
    double c = counts[i];
    c = c > f ? f : c;
    c = c < 1.0 ? 1.0 : c;
    out[i] = f > 0.0 ? (float)(log(f) - log(c)) : 0.f;
  }
}

hipError_t launch_counts_to_lineint_flat(const double* counts, const double* flat, int flat_views, float* out,
                                         int n_views, int nu, int nv, hipStream_t s) {
  const size_t npix = (size_t)nu * nv, n = (size_t)n_views * npix;
  size_t blocks = (n + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  if (blocks == 0) return hipSuccess;
  hipLaunchKernelGGL(counts_to_lineint_flat_kernel, dim3((unsigned)blocks), dim3(256), 0, s, counts, flat, out, n,
                     (size_t)flat_views * npix);
  return hipGetLastError();
}

hipError_t launch_counts_to_lineint(const double* counts, const double* counts2, float* out, int n_views, int nu,
                                    int nv, uint64_t photons_per_view, hipStream_t s) {
  const size_t n = (size_t)n_views * nu * nv;
  size_t blocks = (n + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  if (blocks == 0) return hipSuccess;
  hipLaunchKernelGGL(counts_to_lineint_kernel, dim3((unsigned)blocks), dim3(256), 0, s, counts, counts2, out, n_views,
                     nu, nv, photons_per_view);
  return hipGetLastError();
}

}  // namespace monte
