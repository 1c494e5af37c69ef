// FDK kernels for gfx950: fused cosine weight + ramp filter + transpose,
// voxel-driven backprojection, and their exact transposes (the matched
// projector and the transposed filter) for the adjoint.
//
// Reference: recon/bp3d20_325.cpp:39-175 (weight :39-46, filter :51-81,
// backprojection :91-175).  The reference runs these as scalar FP64 loops; here
// each is one kernel:
//   filter:  one block per (view, 16 detector columns); the weighted input
//            tile is transposed through LDS and every thread keeps a 4 x NB
//            register tile of outputs, so each LDS tap feeds 4 FMAs.
//   backprojection: one thread per voxel column (x,y) and ZC voxels along z.
//            Along a column u, the distance weight and the v-slope are
//            constant, so per voxel-update only v moves (one FMA) before a
//            4-tap bilinear gather from the zero-padded filtered view.  Tiles
//            are ordered so each XCD owns a contiguous z-slab (its L2 then
//            holds only the detector rows that slab projects to).
#include "internal.h"

namespace monte {

// ------------------------------------------------------------------- fill
__global__ void fill_kernel(float* __restrict__ p, float v, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) p[i] = v;
}

hipError_t launch_fill(float* p, float v, size_t n, hipStream_t s) {
  if (n == 0) return hipSuccess;
  size_t blocks = (n + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL(fill_kernel, dim3((unsigned)blocks), dim3(256), 0, s, p, v, n);
  return hipGetLastError();
}

// ---------------------------------------------------------------- filter
__device__ inline float cos_weight(const FilterParams& fp, int r, int c) {
  const float a = fmaf(fp.b_u, (float)r, fp.a_u);
  const float b = fmaf(fp.b_v, (float)c, fp.a_v);
  return fp.D / sqrtf(fmaf(fp.D, fp.D, fmaf(a, a, b * b)));
}

// The ramp filter of one view is a product with a Toeplitz matrix:
//   out[v][b] = cs sum_c A[v][c] H[c][b],  H[c][b] = h[c - b + nu - 1],
// with A the cosine-weighted projection (forward) or the filtered stack
// (transpose; h is symmetric, so the same H serves).  Each block computes a
// 64 (v) x 64 kFnt (b) tile with v_mfma_f32_32x32x2_f32: wave w owns rows
// 32 (w & 1) and the kFnt 32-column tiles at 32 kFnt (w >> 1).  A is staged in LDS
// 32 columns of c at a time as As[c][v]; H is read straight from the taps in
// LDS, shifted by kFbn so every index of the tile stays inside the array.
// Operand layouts (one wave, 32 x 32 x 2): A from lane l is A[l % 32][l / 32],
// B from lane l is B[l / 32][l % 32], and accumulator i of lane l is
// D[8 (i / 4) + 4 (l / 32) + i % 4][l % 32].
constexpr int kFnt = 4;  // 32-column product tiles per wave
constexpr int kFbm = 64, kFbn = 64 * kFnt, kFkc = 32;
static_assert(kFnt % 2 == 0, "the transposed filter stages its result two tiles per wave at a time");
constexpr int kFas = kFbm + 1;  // LDS row pitch of As

typedef float floatx16 __attribute__((ext_vector_type(16)));

// Taps into LDS: Ht[i + kFbn] = h[i], zero elsewhere (2 nu + kFkc + 2 kFbn floats).
__device__ inline void stage_taps(float* __restrict__ Ht, const float* __restrict__ taps, int nu) {
  const int n = 2 * nu + kFkc + 2 * kFbn;
  for (int i = threadIdx.x; i < n; i += 256) {
    const int j = i - kFbn;
    Ht[i] = (j >= 0 && j < 2 * nu - 1) ? taps[j] : 0.f;
  }
}

// acc[t] += A[rows][c0 .. c0 + kFkc) x H[c0 .. c0 + kFkc)[cols of tile t]
__device__ inline void mfma_chunk(const float* __restrict__ As, const float* __restrict__ Ht, int c0, int b0, int nu,
                                  floatx16 (&acc)[kFnt]) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int r = (w & 1) * 32 + (lane & 31), kl = lane >> 5;
  const int col = b0 + (w >> 1) * (32 * kFnt) + (lane & 31);
  // H index of (c0 + kk + kl, col + 32 t) is c - b + nu - 1 + kFbn
  const int h0 = c0 + kl - col + nu - 1 + kFbn;
#pragma unroll
  for (int kk = 0; kk < kFkc; kk += 2) {
    const float a = As[(kk + kl) * kFas + r];
    float bv[kFnt];
#pragma unroll
    for (int t = 0; t < kFnt; ++t) bv[t] = Ht[h0 + kk - 32 * t];
#pragma unroll
    for (int t = 0; t < kFnt; ++t) acc[t] = __builtin_amdgcn_mfma_f32_32x32x2f32(a, bv[t], acc[t], 0, 0, 0);
  }
}

// proj [view][nu][nv] -> filt_pad [view][rows][pitch] (interior at +kPadLo).
// Thread (tc, tv) stages A[c0 + tc + 4 j][v0 + tv], j < 8; the next chunk's
// loads are issued before the current chunk's products, so their latency
// hides behind the MFMAs.
__global__ __launch_bounds__(256) void filter_fwd_mfma_kernel(const float* __restrict__ proj,
                                                              float* __restrict__ out,
                                                              const float* __restrict__ taps, FilterParams fp) {
  extern __shared__ float sm[];
  float* const A0 = sm;  // double-buffered [kFkc][kFas]
  float* const A1 = sm + kFkc * kFas;
  float* Ht = sm + 2 * kFkc * kFas;
  const int nu = fp.nu, nv = fp.nv;
  const int v0 = blockIdx.x * kFbm, b0 = blockIdx.y * kFbn, view = blockIdx.z;
  stage_taps(Ht, taps, nu);
  const float* src = proj + (size_t)view * nu * nv;
  const int tv = threadIdx.x & 63, tc = threadIdx.x >> 6;
  const int v = v0 + tv, vc = v < nv ? v : nv - 1;
  float pre[kFkc / 4];
  auto load = [&](int c0) {
#pragma unroll
    for (int j = 0; j < kFkc / 4; ++j) {
      const int c = c0 + tc + 4 * j;
      pre[j] = src[(size_t)(c < nu ? c : nu - 1) * nv + vc];
    }
  };
  auto stage = [&](float* A, int c0) {
#pragma unroll
    for (int j = 0; j < kFkc / 4; ++j) {
      const int kc = tc + 4 * j, c = c0 + kc;
      A[kc * kFas + tv] = (c < nu && v < nv) ? pre[j] * cos_weight(fp, c, v) : 0.f;
    }
  };
  load(0);
  stage(A0, 0);
  __syncthreads();
  floatx16 acc[kFnt] = {};
  // chunk k is multiplied from one buffer while chunk k + 1 is loaded and
  // staged into the other; one barrier per chunk
  int buf = 0;
  for (int c0 = 0; c0 < nu; c0 += kFkc) {
    const bool more = c0 + kFkc < nu;
    if (more) load(c0 + kFkc);
    mfma_chunk(buf ? A1 : A0, Ht, c0, b0, nu, acc);
    if (more) stage(buf ? A0 : A1, c0 + kFkc);
    __syncthreads();
    buf ^= 1;
  }
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  float* dst = out + (size_t)view * fp.rows * fp.pitch;
#pragma unroll
  for (int t = 0; t < kFnt; ++t) {
    const int b = b0 + (w >> 1) * (32 * kFnt) + 32 * t + (lane & 31);
#pragma unroll
    for (int i = 0; i < 16; ++i) {
      const int vo = v0 + (w & 1) * 32 + 8 * (i / 4) + 4 * (lane >> 5) + (i & 3);
      if (vo < nv && b < nu) dst[(size_t)(vo + kPadLo) * fp.pitch + b + kPadLo] = fp.cs * acc[t][i];
    }
  }
}

// filt_pad [view][rows][pitch] (interior) -> proj [view][nu][nv], weighted: the
// transpose of the forward filter.  Thread (tk, tr) stages A[v0 + tr + 8 j][c0 + tk],
// j < 8, pipelined like the forward kernel; results go through LDS so the
// [u][v] output is written in runs along v.
__global__ __launch_bounds__(256) void filter_T_mfma_kernel(const float* __restrict__ in, float* __restrict__ proj,
                                                            const float* __restrict__ taps, FilterParams fp) {
  extern __shared__ float sm[];
  float* const A0 = sm;  // double-buffered [kFkc][kFas]
  float* const A1 = sm + kFkc * kFas;
  float* Ht = sm + 2 * kFkc * kFas;
  const int nu = fp.nu, nv = fp.nv;
  const int v0 = blockIdx.x * kFbm, b0 = blockIdx.y * kFbn, view = blockIdx.z;
  stage_taps(Ht, taps, nu);
  const float* src = in + (size_t)view * fp.rows * fp.pitch;
  const int tk = threadIdx.x & 31, tr = threadIdx.x >> 5;
  float pre[kFbm / 8];
  auto load = [&](int c0) {
    const int c = c0 + tk, cc = c < nu ? c : nu - 1;
#pragma unroll
    for (int j = 0; j < kFbm / 8; ++j) {
      const int v = v0 + tr + 8 * j;
      pre[j] = src[(size_t)((v < nv ? v : nv - 1) + kPadLo) * fp.pitch + cc + kPadLo];
    }
  };
  auto stage = [&](float* A, int c0) {
    const bool cin = c0 + tk < nu;
#pragma unroll
    for (int j = 0; j < kFbm / 8; ++j) {
      const int vl = tr + 8 * j;
      A[tk * kFas + vl] = (cin && v0 + vl < nv) ? pre[j] : 0.f;
    }
  };
  load(0);
  stage(A0, 0);
  __syncthreads();
  floatx16 acc[kFnt] = {};
  int buf = 0;
  for (int c0 = 0; c0 < nu; c0 += kFkc) {
    const bool more = c0 + kFkc < nu;
    if (more) load(c0 + kFkc);
    mfma_chunk(buf ? A1 : A0, Ht, c0, b0, nu, acc);
    if (more) stage(buf ? A0 : A1, c0 + kFkc);
    __syncthreads();
    buf ^= 1;
  }
  // stage D as Ds[b_local][v_local] over the same LDS (the loop ended on a
  // barrier), two tiles per wave at a time, then write proj[b][v] along v.
  // In pass h, Ds row bl holds b = b0 + 32 kFnt (bl / 64) + 64 h + bl % 64.
  float* Ds = sm;  // [128][kFas]
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  float* dst = proj + (size_t)view * nu * nv;
#pragma unroll
  for (int h = 0; h < kFnt / 2; ++h) {
    if (h > 0) __syncthreads();
#pragma unroll
    for (int t = 0; t < 2; ++t) {
      const int bl = (w >> 1) * 64 + 32 * t + (lane & 31);
#pragma unroll
      for (int i = 0; i < 16; ++i) {
        const int vl = (w & 1) * 32 + 8 * (i / 4) + 4 * (lane >> 5) + (i & 3);
        Ds[bl * kFas + vl] = acc[2 * h + t][i];
      }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < 128 * kFbm; i += 256) {
      const int bl = i / kFbm, vl = i - bl * kFbm;
      const int b = b0 + (bl >> 6) * (32 * kFnt) + 64 * h + (bl & 63), v = v0 + vl;
      if (b < nu && v < nv) dst[(size_t)b * nv + v] = fp.cs * Ds[bl * kFas + vl] * cos_weight(fp, b, v);
    }
  }
}

// LDS bytes: two As buffers and the taps; the transposed filter also stages its
// result, 128 x 64 at a time, over the same memory.
static size_t filter_mfma_lds(int nu, bool transposed) {
  const size_t stage = 2 * (size_t)kFkc * kFas + 2 * (size_t)nu + kFkc + 2 * kFbn;
  const size_t out = transposed ? (size_t)128 * kFas : 0;
  return sizeof(float) * (stage > out ? stage : out);
}

hipError_t launch_fdk_filter(const float* proj, float* filt_pad, const float* taps, const FilterParams& fp,
                             hipStream_t s) {
  const size_t lds = filter_mfma_lds(fp.nu, false);
  if (lds > 64 * 1024) return hipErrorInvalidValue;
  dim3 grid((fp.nv + kFbm - 1) / kFbm, (fp.nu + kFbn - 1) / kFbn, fp.n_views);
  hipLaunchKernelGGL(filter_fwd_mfma_kernel, grid, dim3(256), lds, s, proj, filt_pad, taps, fp);
  return hipGetLastError();
}

hipError_t launch_fdk_filter_T(const float* filt_pad, float* proj, const float* taps, const FilterParams& fp,
                               hipStream_t s) {
  const size_t lds = filter_mfma_lds(fp.nu, true);
  if (lds > 64 * 1024) return hipErrorInvalidValue;
  dim3 grid((fp.nv + kFbm - 1) / kFbm, (fp.nu + kFbn - 1) / kFbn, fp.n_views);
  hipLaunchKernelGGL(filter_T_mfma_kernel, grid, dim3(256), lds, s, filt_pad, proj, taps, fp);
  return hipGetLastError();
}

// -------------------------------------------------------------- pad copies
__global__ void pad_copy_kernel(const float* __restrict__ src, float* __restrict__ dst, int n_views, int nu, int nv,
                                int pitch, int rows, int to_padded) {
  const size_t n = (size_t)n_views * rows * pitch;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const size_t k = i / ((size_t)rows * pitch);
    const int rem = (int)(i - k * rows * pitch);
    const int r = rem / pitch - kPadLo, c = rem % pitch - kPadLo;
    const bool inside = r >= 0 && r < nv && c >= 0 && c < nu;
    const size_t si = (k * nv + (inside ? r : 0)) * nu + (inside ? c : 0);
    if (to_padded)
      dst[i] = inside ? src[si] : 0.f;
    else if (inside)
      dst[si] = src[i];
  }
}

hipError_t launch_pad_copy(const float* filt, float* filt_pad, int n_views, int nu, int nv, int pitch, int rows,
                           hipStream_t s) {
  const size_t n = (size_t)n_views * rows * pitch;
  size_t blocks = (n + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL(pad_copy_kernel, dim3((unsigned)blocks), dim3(256), 0, s, filt, filt_pad, n_views, nu, nv,
                     pitch, rows, 1);
  return hipGetLastError();
}

// Guard cells of a padded stack.  Standard preset: zeros (the bilinear
// interpolant's zero extension).  bp3d20 preset (wrap != 0): the value the
// reference reads there.  recon/bp3d20_325.cpp:163-166 index the flat array
// map_out[view*H*W + row*H + col], so a tap at col == H lands on the next
// row's first elements and a tap at row == H on the next view's first row;
// cells past the last view read as zero.  Cells before the interior are never
// sampled in that preset and are zeroed.
__device__ inline bool wrap_source(int k, int r, int c, int n_views, int nu, int nv, int& kk, int& rr, int& cc) {
  if (r < 0 || c < 0) return false;
  const long flat = (long)r * nu + c;
  const long plane = (long)nu * nv;
  kk = k + (int)(flat / plane);
  const long w = flat % plane;
  rr = (int)(w / nu);
  cc = (int)(w % nu);
  return kk < n_views;
}

// One thread per guard cell (the interior is left alone): per view, the guard
// rows above the detector, the guard columns of its nv rows, the guard rows below.
__global__ void fill_pads_kernel(float* __restrict__ buf, int n_views, int nu, int nv, int pitch, int rows,
                                 int wrap) {
  const int top = kPadLo * pitch, side = pitch - nu, mid = nv * side;
  const int per_view = rows * pitch - nu * nv;
  const size_t n = (size_t)n_views * per_view;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const int k = (int)(i / per_view);
    int j = (int)(i - (size_t)k * per_view);
    int r, c;  // detector coordinates of the guard cell
    if (j < top) {
      r = j / pitch - kPadLo;
      c = j % pitch - kPadLo;
    } else if ((j -= top) < mid) {
      r = j / side;
      const int q = j - r * side;
      c = q < kPadLo ? q - kPadLo : nu + (q - kPadLo);
    } else {
      j -= mid;
      r = nv + j / pitch;
      c = j % pitch - kPadLo;
    }
    float v = 0.f;
    int kk, rr, cc;
    if (wrap && wrap_source(k, r, c, n_views, nu, nv, kk, rr, cc))
      v = buf[((size_t)kk * rows + rr + kPadLo) * pitch + cc + kPadLo];
    buf[((size_t)k * rows + r + kPadLo) * pitch + c + kPadLo] = v;
  }
}

hipError_t launch_fill_pads(float* buf, int n_views, int nu, int nv, int pitch, int rows, int wrap, hipStream_t s) {
  const size_t n = (size_t)n_views * ((size_t)rows * pitch - (size_t)nu * nv);
  size_t blocks = (n + 255) / 256;
  if (blocks > 16384) blocks = 16384;
  hipLaunchKernelGGL(fill_pads_kernel, dim3((unsigned)blocks), dim3(256), 0, s, buf, n_views, nu, nv, pitch, rows,
                     wrap);
  return hipGetLastError();
}

// Transpose of the bp3d20 guard-cell wrap: what the matched projector
// deposited on a guard cell belongs to the interior cell it aliases.  An
// interior cell (k, r, c) is aliased by at most three guard cells (for
// nu >= kPadHi and nv > kPadHi):
//   (k, r - 1, nu + c)        when c < kPadHi and r >= 1,
//   (k - 1, nv + r - 1, nu + c) when c < kPadHi and r <= kPadHi,
//   (k - 1, nv + r, c)        when r < kPadHi,
// the inverse of wrap_source.  Each interior cell gathers them in that fixed
// order, so the sum does not depend on scheduling.
__global__ void fold_pads_kernel(float* __restrict__ buf, int n_views, int nu, int nv, int pitch, int rows) {
  const size_t n = (size_t)n_views * nv * nu;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const int k = (int)(i / ((size_t)nv * nu));
    const int rem = (int)(i - (size_t)k * nv * nu);
    const int r = rem / nu, c = rem - r * nu;
    if (c >= kPadHi && r > kPadHi) continue;
    auto at = [&](int kk, int rr, int cc) -> float& {
      return buf[((size_t)kk * rows + rr + kPadLo) * pitch + cc + kPadLo];
    };
    float v = at(k, r, c);
    if (c < kPadHi && r >= 1) v += at(k, r - 1, nu + c);
    if (k >= 1 && c < kPadHi && r <= kPadHi) v += at(k - 1, nv + r - 1, nu + c);
    if (k >= 1 && r < kPadHi) v += at(k - 1, nv + r, c);
    at(k, r, c) = v;
  }
}

hipError_t launch_fold_pads(float* buf, int n_views, int nu, int nv, int pitch, int rows, hipStream_t s) {
  if (nu < kPadHi || nv <= kPadHi) return hipErrorInvalidValue;
  const size_t n = (size_t)n_views * nv * nu;
  size_t blocks = (n + 255) / 256;
  if (blocks > 16384) blocks = 16384;
  hipLaunchKernelGGL(fold_pads_kernel, dim3((unsigned)blocks), dim3(256), 0, s, buf, n_views, nu, nv, pitch, rows);
  return hipGetLastError();
}

hipError_t launch_unpad_copy(const float* filt_pad, float* filt, int n_views, int nu, int nv, int pitch, int rows,
                             hipStream_t s) {
  const size_t n = (size_t)n_views * rows * pitch;
  size_t blocks = (n + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL(pad_copy_kernel, dim3((unsigned)blocks), dim3(256), 0, s, filt_pad, filt, n_views, nu, nv,
                     pitch, rows, 0);
  return hipGetLastError();
}

// ------------------------------------------------------- backprojection
// Maps a hardware block index to a logical tile so that the blocks the
// dispatcher sends to XCD x (blocks x, x+8, x+16, ...) cover one contiguous
// range of logical tiles.  Tiles are numbered z-chunk slowest.
__device__ inline int xcd_tile(int b, int nblocks) {
  const int q = nblocks >> 3, r = nblocks & 7;
  const int xcd = b & 7, j = b >> 3;
  return xcd * q + (xcd < r ? xcd : r) + j;
}


// Per-column quantities of one view, shared by the backprojector and its
// transpose so both evaluate bit-identical weights.
struct ColumnCore {
  int iu;       // u pixel index of the left tap (sample window clamped)
  float fu;     // u fraction
  float w;      // distance weight x angular factor
  bool ok;      // su inside the window
  double kv;    // d(sv)/dZ = svk / U
  float dsv;    // sv increment per voxel step along z
};

__device__ inline ColumnCore column_core(const BpView& c, double X, double Y, float Xf, float Yf, const BpGrid& g) {
  ColumnCore r;
  const double U = fma(c.cx, X, fma(c.cy, Y, c.c0));
  const double inv = 1.0 / U;
  double su = fma(c.suk * fma(c.ax, X, c.ay * Y), inv, c.su0);
  r.ok = su >= g.lo_u && su <= g.hi_u;
  su = fmin(fmax(su, g.lo_u), g.hi_u);
  const double fu0 = floor(su);
  r.iu = (int)fu0;
  r.fu = (float)(su - fu0);
  r.kv = c.svk * inv;
  r.dsv = (float)(r.kv * g.gsz);
  // sign test in unfused FP64 (as recon/bp3d20_325.cpp:145,151 evaluates it)
  const double gp = xadd(xmul(X, c.gx), xmul(Y, c.gy));
  const float hp = fabsf(fmaf(c.hx, Xf, c.hy * Yf));
  const float dd = fmaf(c.dx, Xf, fmaf(c.dy, Yf, c.d0)) - (gp < 0.0 ? -hp : hp);
  const float q = c.wn / dd;
  r.w = c.fac * q * q;
  return r;
}

// v coordinate at a 16-aligned base voxel: integer part and FP32 fraction.
struct ChunkBase {
  int ivb;
  float fvb;
  float flo, fhi;  // v window relative to ivb
};

// The v coordinate of voxel z is formed from a base taken at the 16-aligned z
// at or below it (zb), whatever the thread's z-chunk, so every chunking and
// tile shape evaluates the same floating-point expression for each voxel and
// the volume is bitwise independent of the launch tuning.
constexpr int kZBase = 16;

__device__ inline double base_z(const BpGrid& g, int zb) {
  // recon/bp3d20_325.cpp:109 forms the coordinate as start - index * step, unfused
  return xadd(g.gz0, xmul((double)zb, g.gsz));
}

__device__ inline ChunkBase chunk_base(const ColumnCore& cc, const BpView& c, double Zb, const BpGrid& g) {
  ChunkBase b;
  double sv = fma(cc.kv, Zb, c.sv0);
  sv = fmin(fmax(sv, -1.0e6), 1.0e6);
  const double fv0 = floor(sv);
  b.ivb = (int)fv0;
  b.fvb = (float)(sv - fv0);
  b.flo = (float)(g.lo_v - fv0);
  b.fhi = (float)(g.hi_v - fv0);
  return b;
}

// Sample position of voxel zb + j, as the byte offset of its tap quad in the
// view (`off0` is that of row ivb, column iu; rows are `prow` bytes apart) and
// the row fraction; false when the sample lies outside the v window.  The
// clamped coordinate equals the unclamped one exactly when it is inside, so
// one compare gives the window test.  |floor(f)| < 2^23 (sv is clamped to
// +-1e6 and the window is a few thousand rows) and prow < 2^23, so a 24-bit
// multiply forms the row step; the sum is taken mod 2^32 and the true offset
// lies inside the view.
__device__ inline bool chunk_tap(const ColumnCore& cc, const ChunkBase& b, int j, uint32_t off0, uint32_t prow,
                                 uint32_t& off, float& fv) {
  const float f = fmaf((float)j, cc.dsv, b.fvb);
  const float fc = fminf(fmaxf(f, b.flo), b.fhi);
  const float fl = floorf(fc);
  fv = fc - fl;
  off = off0 + (uint32_t)__mul24((int)fl, (int)prow);
  return fc == f;
}

__device__ inline bool in_mask(const BpGrid& g, int x, int y, int z) {
  if (!g.mask) return true;
  const int dz = z - g.mcz, dy = y - g.mcy, dx = x - g.mcx;
  return (float)(dz * dz + dy * dy + dx * dx) <= g.mask_r2;
}

// recon/bp3d20_325.cpp:107-108 form X and Y as start + index * step, unfused
__device__ inline double voxel_x(const BpGrid& g, int x) { return xadd(g.gx0, xmul((double)x, g.gsx)); }
__device__ inline double voxel_y(const BpGrid& g, int y) { return xadd(g.gy0, xmul((double)y, g.gsy)); }

template <int ZC>
__global__ __launch_bounds__(256) void backproject_kernel(const float4* __restrict__ quads,
                                                          const BpView* __restrict__ views, int view0, int nviews,
                                                          BpGrid g, float* __restrict__ vol, int accumulate,
                                                          int tile_x, int tiles_x, int tiles_y) {
  const int t = xcd_tile(blockIdx.x, gridDim.x);
  const int per_z = tiles_x * tiles_y;
  const int zc = t / per_z;
  const int rem = t - zc * per_z;
  const int ty_t = rem / tiles_x, tx_t = rem - ty_t * tiles_x;
  const int tile_y = 256 / tile_x;
  const int x = g.x0 + tx_t * tile_x + (int)(threadIdx.x % tile_x);
  const int y = ty_t * tile_y + (int)(threadIdx.x / tile_x);
  const int z0 = zc * ZC;
  if (x >= g.x1 || y >= g.ny) return;
  const int zb = (z0 / kZBase) * kZBase, joff = z0 - zb;
  const double X = voxel_x(g, x), Y = voxel_y(g, y), Zb = base_z(g, zb);
  const float Xf = (float)X, Yf = (float)Y;
  bool live[ZC];
#pragma unroll
  for (int j = 0; j < ZC; ++j) live[j] = (z0 + j) < g.nz && in_mask(g, x, y, z0 + j);
  float acc[ZC];
#pragma unroll
  for (int j = 0; j < ZC; ++j) acc[j] = 0.f;

  // One 16-byte load per voxel: the vector memory pipeline costs about the
  // same per instruction whatever its width, so the four bilinear taps come
  // from one tap quad rather than two 8-byte row pairs.
  const size_t view_stride = (size_t)g.rows * g.pitch;
  const uint32_t prow = (uint32_t)g.pitch * 16u;
  for (int k = view0; k < view0 + nviews; ++k) {
    const BpView& c = views[k];
    const ColumnCore cc = column_core(c, X, Y, Xf, Yf, g);
    const ChunkBase cb = chunk_base(cc, c, Zb, g);
    const char* vb = (const char*)(quads + (size_t)k * view_stride);
    const uint32_t off0 = (uint32_t)(cb.ivb + kPadLo) * prow + (uint32_t)(cc.iu + kPadLo) * 16u;
    const float wu = cc.ok ? cc.w : 0.f;
#pragma unroll
    for (int j = 0; j < ZC; ++j) {
      uint32_t off;
      float fv;
      const bool okv = chunk_tap(cc, cb, joff + j, off0, prow, off, fv);
      const float4 t = *(const float4*)(vb + off);
      const float a = fmaf(cc.fu, t.y, t.x);
      const float b = fmaf(cc.fu, t.w, t.z);
      const float val = fmaf(fv, b - a, a);
      acc[j] = fmaf((okv && live[j]) ? wu : 0.f, val, acc[j]);
    }
  }
#pragma unroll
  for (int j = 0; j < ZC; ++j) {
    if (z0 + j >= g.nz) break;
    float* o = vol + ((size_t)(z0 + j) * g.ny + y) * g.nx + x;
    *o = accumulate ? *o + acc[j] : acc[j];
  }
}

static void bp_tiling(const BpGrid& grid, const monte_tuning& t, int zc, int& tile_x, int& tiles_x, int& tiles_y,
                      long& nblocks) {
  tile_x = t.bp_tile_x > 0 ? t.bp_tile_x : 16;
  const int tile_y = 256 / tile_x;
  const int ncols = grid.x1 - grid.x0;
  tiles_x = (ncols + tile_x - 1) / tile_x;
  tiles_y = (grid.ny + tile_y - 1) / tile_y;
  const int zchunks = (grid.nz + zc - 1) / zc;
  nblocks = (long)tiles_x * tiles_y * zchunks;
}

// quads [view][rows][pitch]: quad (r, c) = (p[r][c], p[r][c + 1] - p[r][c], p[r + 1][c],
// p[r + 1][c + 1] - p[r + 1][c]) of the padded filtered view p (zero past the last
// row or column): the two rows of a voxel's bilinear taps with their u
// differences, which the backprojector would otherwise form per voxel.
// One thread per padded cell: blocks of 64 columns x 4 rows, one grid layer
// per view, so the cell's view, row and column need no division.
constexpr int kQuadCols = 64, kQuadRows = 4;
__global__ __launch_bounds__(kQuadCols * kQuadRows) void tap_quads_kernel(const float* __restrict__ pad,
                                                                          float4* __restrict__ quads, int rows,
                                                                          int pitch) {
  const int cl = blockIdx.x * kQuadCols + (int)(threadIdx.x % kQuadCols);
  const int r = blockIdx.y * kQuadRows + (int)(threadIdx.x / kQuadCols);
  if (cl >= pitch || r >= rows) return;
  const size_t i = ((size_t)blockIdx.z * rows + r) * pitch + cl;
  const float* p = pad + i;
  const bool rr = r + 1 < rows, cr = cl + 1 < pitch;
  const float p00 = p[0], p01 = cr ? p[1] : 0.f;
  const float p10 = rr ? p[pitch] : 0.f, p11 = (rr && cr) ? p[pitch + 1] : 0.f;
  quads[i] = make_float4(p00, p01 - p00, p10, p11 - p10);
}

hipError_t launch_tap_quads(const float* filt_pad, float4* quads, int n_views, int rows, int pitch, hipStream_t s) {
  if (n_views <= 0 || rows <= 0 || pitch <= 0) return hipSuccess;
  if (n_views > 65535) return hipErrorInvalidValue;  // grid z limit
  const dim3 grid((unsigned)((pitch + kQuadCols - 1) / kQuadCols), (unsigned)((rows + kQuadRows - 1) / kQuadRows),
                  (unsigned)n_views);
  hipLaunchKernelGGL(tap_quads_kernel, grid, dim3(kQuadCols * kQuadRows), 0, s, filt_pad, quads, rows, pitch);
  return hipGetLastError();
}

hipError_t launch_backproject(const float4* quads, const BpView* views, int view0, int nviews_chunk,
                              const BpGrid& grid, float* vol, int accumulate, const monte_tuning& t,
                              hipStream_t s) {
  const int zc = t.bp_zchunk > 0 ? t.bp_zchunk : 8;
  int tile_x, tiles_x, tiles_y;
  long nblocks;
  if (t.bp_tile_x > 0 && 256 % t.bp_tile_x) return hipErrorInvalidValue;
  // quads are addressed by 32-bit byte offsets with a 24-bit row multiply
  if ((long long)grid.pitch * 16 >= (1LL << 23) || (long long)grid.rows * grid.pitch * 16 >= (1LL << 32))
    return hipErrorInvalidValue;
  bp_tiling(grid, t, zc, tile_x, tiles_x, tiles_y, nblocks);
  if (nblocks <= 0) return hipSuccess;
  if (nblocks > 0x7fffffffL) return hipErrorInvalidValue;
  dim3 gd((unsigned)nblocks), bd(256);
  switch (zc) {
    case 4: hipLaunchKernelGGL((backproject_kernel<4>), gd, bd, 0, s, quads, views, view0, nviews_chunk, grid, vol, accumulate, tile_x, tiles_x, tiles_y); break;
    case 8: hipLaunchKernelGGL((backproject_kernel<8>), gd, bd, 0, s, quads, views, view0, nviews_chunk, grid, vol, accumulate, tile_x, tiles_x, tiles_y); break;
    case 16: hipLaunchKernelGGL((backproject_kernel<16>), gd, bd, 0, s, quads, views, view0, nviews_chunk, grid, vol, accumulate, tile_x, tiles_x, tiles_y); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

// ------------------------------------------- transposed backprojection
// The matched projector, written as a gather so no atomics are needed.
// One wave owns a strip of kStripW padded detector columns (all rows) of one
// view, held in its own LDS.  It walks the voxel columns whose left tap
// lands in the strip (each voxel column of the view belongs to exactly one
// strip; its right tap may spill into the next strip's first column), from a
// list enumerated once per geometry: for each slab of the volume across the view's rays
// (x-slabs when |cos b| >= |sin b|, y-slabs otherwise) the projective map
// su(X,Y) is inverted at the strip's two edges, giving a range of columns on
// that slab (widened by one voxel each side).  The lanes evaluate the
// candidate columns of a range in parallel with the backprojector's own
// column arithmetic and keep those whose taps fall in the strip.  Each kept
// column is then walked along z by the whole wave: lane l takes the C
// consecutive voxels of chunk l, from a z-fastest copy of the volume.  The
// column core (u tap, fraction, distance weight) is the backprojector's own;
// the v coordinate sv(z) = A + z D of each voxel is formed in 32.32 fixed
// point from the column's FP64 A and D (integer row = high word, fraction =
// low word), within 2^-32 rows of the backprojector's value.
//
// Accumulation is a plain read-add-write of LDS, without atomics: in one
// instruction the active lanes' voxels are C apart along z, hence at least
// C |dsv| rows apart on the detector, and when C |dsv| > 1.01 no two lanes
// touch the same cell.  Columns with a smaller v-slope run in several phases
// over interleaved lane subsets that satisfy the same bound.  The upper tap
// row is updated by a second read-add-write after the first completes, and a
// wave's LDS operations execute in program order, so updates of a cell by
// neighbouring lanes, later voxels and later columns all see each other.
// LDS column kStripC0 + kStripW collects the right-tap spill into the next
// strip.  The kernel's speed follows the number of waves a CU holds, and a
// wave's LDS is its strip: with the column lists built once per geometry
// (below), one-column strips ran 13.1 ms at 325^3 / 512^2 / 360 views,
// against 14.7, 15.2 and 16.6 ms for strips of 2, 3 and 5 columns.  A row is
// then the pair (left tap, spill), read and written as one 8-byte access
// (11.8 ms, against 12.3 ms for pitch 3 and 13.0 ms for pitch 4).
constexpr int kStripW = 1;
constexpr int kStripP = 2;
constexpr int kStripC0 = 0;  // LDS column of the strip's first column

// A cell and its right neighbour (left and right tap of one row), as one
// 8-byte LDS access when the layout keeps them 8-byte aligned.
constexpr bool kPairAligned = kStripP % 2 == 0 && kStripC0 % 2 == 0;
__device__ inline float2 lds_get2(const float* strip, int i) {
  if constexpr (kPairAligned) return *(const float2*)(strip + i);
  return make_float2(strip[i], strip[i + 1]);
}
__device__ inline void lds_put2(float* strip, int i, float2 v) {
  if constexpr (kPairAligned) {
    *(float2*)(strip + i) = v;
  } else {
    strip[i] = v.x;
    strip[i + 1] = v.y;
  }
}
__device__ inline void lds_add2(float* strip, int i, float a, float b) {
  float2 x = lds_get2(strip, i);
  x.x += a;
  x.y += b;
  lds_put2(strip, i, x);
}

// Columns of one strip on each slab.  The strip's two edges su = s_lo and
// su = s_hi are lines through the source; on a slab (x fixed for x-slabs, y
// fixed for y-slabs) each crosses the other coordinate at
//   t = (q (c1 coord + c0) - a1 coord) / (a2 - q c2),  q = (s - su0) / suk,
// with (c1, c2, a1, a2) = (cx, cy, ax, ay) for x-slabs and (cy, cx, ay, ax)
// for y-slabs.  The denominators depend only on the view and the edge, so the
// per-slab work is two multiply-adds per edge.  The range is widened by one
// voxel each side; every column in it is then tested exactly.
struct SlabRanges {
  double q0, q1, id0, id1;  // edge parameters and reciprocal denominators
  double o0, ios;           // voxel index = (t - o0) * ios along the slab
  int n;                    // voxels along the slab
  bool whole;               // an edge runs parallel to the slabs: take every column

  __device__ void init(double s_lo, double s_hi, const BpView& c, bool x_slabs, const BpGrid& g) {
    q0 = (s_lo - c.su0) / c.suk;
    q1 = (s_hi - c.su0) / c.suk;
    const double a2 = x_slabs ? c.ay : c.ax, c2 = x_slabs ? c.cy : c.cx;
    const double d0 = a2 - q0 * c2, d1 = a2 - q1 * c2;
    whole = fabs(d0) < 1e-12 || fabs(d1) < 1e-12;
    id0 = 1.0 / d0;
    id1 = 1.0 / d1;
    o0 = x_slabs ? g.gy0 : g.gx0;
    ios = 1.0 / (x_slabs ? g.gsy : g.gsx);
    n = x_slabs ? g.ny : g.nx;
  }
  __device__ void range(const BpView& c, bool x_slabs, double coord, int& a, int& b) const {
    if (whole) { a = 0; b = n - 1; return; }
    const double c1 = x_slabs ? c.cx : c.cy, a1 = x_slabs ? c.ax : c.ay;
    const double base = c1 * coord + c.c0;
    const double t0 = (q0 * base - a1 * coord) * id0, t1 = (q1 * base - a1 * coord) * id1;
    const double i0 = (t0 - o0) * ios, i1 = (t1 - o0) * ios;
    a = (int)fmax(floor(fmin(i0, i1)) - 1.0, 0.0);
    b = (int)fmin(ceil(fmax(i0, i1)) + 1.0, (double)(n - 1));
  }
};

__device__ inline int bcast_i(int v, int src) { return __builtin_amdgcn_readlane(v, src); }
__device__ inline float bcast_f(float v, int src) { return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), src)); }
__device__ inline double bcast_d(double v, int src) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffLL), src);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), src);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// Keeps the compiler from moving LDS accesses across this point; the hardware
// already executes one wave's LDS operations in order.
__device__ inline void lds_order() { asm volatile("" ::: "memory"); }

// Passes (64 chunks of C voxels) whose volume values are fetched one column
// ahead; columns longer than kPrefetchZ / C / 64 passes load the rest in place.
constexpr int kPrefetchZ = 512;
constexpr int kAhead = 2;
// Views per group of the projector's block order (see project_kernel): 4, 8
// and 16 took 10.4, 10.0 and 10.4 ms at 325^3 / 360 views (51 -> 35, 26 and
// 29 GB of fabric reads), 76.5, 74.1 and 75.3 ms at 512^3 / 720 views.
constexpr int kProjViewGroup = 8;
// LDS floats after the strip that absorb the updates of lanes with nothing to
// add (outside the v window or past the column end), so the updates need no
// branches: two per lane, two row pitches and one cell of overlap.  (Branches
// around the updates, or handing the upper-row update to the next lane by a
// DPP shift, both measured slower.)
constexpr int kTrash = 2 * 64 + 2 * kStripP + 2;

template <int C>
struct ColumnVals {
  static constexpr int NP = (kPrefetchZ / C + 63) / 64;
  float v[NP][C];
};

template <bool MASK>
__device__ inline bool in_mask_t(const BpGrid& g, int x, int y, int z) {
  if (!MASK) return true;
  const int dz = z - g.mcz, dy = y - g.mcy, dx = x - g.mcx;
  return (float)(dz * dz + dy * dy + dx * dx) <= g.mask_r2;
}

// Volume values of chunk (p * 64 + lane) of a column, zero past the end and
// outside the mask; the load itself is clamped into the column, so it is
// unconditional.
template <int C, bool MASK>
__device__ inline void load_chunk(float (&v)[C], const float* __restrict__ col, int p, int x, int y,
                                  const BpGrid& g, int lane) {
#pragma unroll
  for (int j = 0; j < C; ++j) {
    const int z = (p * 64 + lane) * C + j;
    const float val = col[min(z, g.nz - 1)];
    v[j] = (z < g.nz && in_mask_t<MASK>(g, x, y, z)) ? val : 0.f;
  }
}

template <int C, bool MASK>
__device__ inline void load_column(ColumnVals<C>& cv, const float* __restrict__ col, int x, int y, const BpGrid& g,
                                   int lane, int npass) {
#pragma unroll
  for (int p = 0; p < ColumnVals<C>::NP; ++p)
    if (p < npass) load_chunk<C, MASK>(cv.v[p], col, p, x, y, g, lane);
}

// The v coordinate of a column in 32.32 fixed point: sv(z) = a + z d, and the
// v window [lo, hi] (the backprojector's, inclusive).
struct ColumnFx {
  long long a, d, lo, hi;
};

__device__ inline ColumnFx column_fx(const ColumnCore& cc, const BpView& c, const BpGrid& g) {
  ColumnFx f;
  const double A = fma(cc.kv, g.gz0, c.sv0), D = cc.kv * g.gsz;
  f.a = __double2ll_rn(fmin(fmax(A, -1.0e6), 1.0e6) * 4294967296.0);
  f.d = __double2ll_rn(D * 4294967296.0);
  f.lo = (long long)g.lo_v * 4294967296LL;
  f.hi = (long long)g.hi_v * 4294967296LL;
  return f;
}

// Row (floor) and fraction of a fixed-point coordinate.
__device__ inline int fx_row(long long s) { return (int)(s >> 32); }
__device__ inline float fx_frac(long long s) { return (float)(uint32_t)s * 2.3283064365386963e-10f; }  // 2^-32

// The C voxel updates of one lane; lanes with `act` false (or nothing to add)
// update their trash floats instead.
template <int C>
__device__ inline void pass_updates(float* __restrict__ strip, const float (&v)[C], const ColumnCore& cc,
                                    long long s0, long long d, long long lo, long long hi, int z0, bool act,
                                    int cell0, int trash, const BpGrid& g) {
#pragma unroll
  for (int j = 0; j < C; ++j) {
    const long long sj = s0 + j * d;
    const int row = fx_row(sj);
    const float fv = fx_frac(sj);
    const bool live = act && sj >= lo && sj <= hi && z0 + j < g.nz;
    const float sw = cc.w * v[j];
    const float bb = sw * fv, aa = sw - bb;
    const float a1 = aa * cc.fu, b1 = bb * cc.fu;
    const int i0 = live ? cell0 + row * kStripP : trash;
    lds_add2(strip, i0, aa - a1, a1);
    lds_order();
    lds_add2(strip, i0 + kStripP, bb - b1, b1);
    lds_order();
  }
}

// Both voxels of a 2-voxel chunk, when the v-slope is in (0, 1): the second
// voxel's row is the first's or the one above, so the lane's four row updates
// fold into three consecutive rows r, r + 1, r + 2, added in three rounds.  In
// round k every lane updates its row r + k; the lanes' rows r strictly
// increase, so no two lanes of a round touch the same cell.
struct PairRows {
  int i0;            // LDS index of (row r, left tap), or the trash floats
  float q0, q1, q2;  // totals for rows r, r + 1, r + 2
};

__device__ inline PairRows pair_rows(const float (&v)[2], const ColumnCore& cc, long long s0, long long d,
                                     long long lo, long long hi, int z0, int cell0, int trash, const BpGrid& g) {
  const long long s1 = s0 + d;
  const int r0 = fx_row(s0), r1 = fx_row(s1);
  const bool live0 = s0 >= lo && s0 <= hi && z0 < g.nz, live1 = s1 >= lo && s1 <= hi && z0 + 1 < g.nz;
  const float sw0 = live0 ? cc.w * v[0] : 0.f, sw1 = live1 ? cc.w * v[1] : 0.f;
  const float b0 = sw0 * fx_frac(s0), a0 = sw0 - b0;
  const float b1 = sw1 * fx_frac(s1), a1 = sw1 - b1;
  const bool same = r1 == r0;
  PairRows p;
  p.q0 = a0 + (same ? a1 : 0.f);
  p.q1 = b0 + (same ? b1 : a1);
  p.q2 = same ? 0.f : b1;
  p.i0 = (live0 || live1) ? cell0 + r0 * kStripP : trash;
  return p;
}

__device__ inline float pair_q(const PairRows& p, int k) { return k == 0 ? p.q0 : (k == 1 ? p.q1 : p.q2); }

__device__ inline void pass_updates_pair(float* __restrict__ strip, const float (&v)[2], const ColumnCore& cc,
                                         long long s0, long long d, long long lo, long long hi, int z0, int cell0,
                                         int trash, const BpGrid& g) {
  const PairRows p = pair_rows(v, cc, s0, d, lo, hi, z0, cell0, trash, g);
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const int i = p.i0 + k * kStripP;
    const float q = pair_q(p, k), t = q * cc.fu;
    lds_add2(strip, i, q - t, t);
    lds_order();
  }
}

// Two pair passes whose rows cannot meet, run round by round together so each
// round's two reads are in flight at once.  Passes p and p + 2 of a column
// start 256 voxels apart, so at a v-slope above 0.505 rows per voxel the first
// row of the later pass lies more than 64 rows above the last row (r + 2) of
// the earlier.
__device__ inline void pass_updates_pair2(float* __restrict__ strip, const float (&va)[2], const float (&vb)[2],
                                          const ColumnCore& cc, long long sa, long long sb, long long d,
                                          long long lo, long long hi, int za, int zb, int cell0, int trash,
                                          const BpGrid& g) {
  const PairRows a = pair_rows(va, cc, sa, d, lo, hi, za, cell0, trash, g);
  const PairRows b = pair_rows(vb, cc, sb, d, lo, hi, zb, cell0, trash, g);
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const int ia = a.i0 + k * kStripP, ib = b.i0 + k * kStripP;
    const float qa = pair_q(a, k), ta = qa * cc.fu;
    const float qb = pair_q(b, k), tb = qb * cc.fu;
    const float2 xa = lds_get2(strip, ia), xb = lds_get2(strip, ib);
    lds_put2(strip, ia, make_float2(xa.x + (qa - ta), xa.y + ta));
    lds_put2(strip, ib, make_float2(xb.x + (qb - tb), xb.y + tb));
    lds_order();
  }
}

// Every pass of an unfolded column at once (npass <= NP, one lane phase),
// round by round: per voxel j of a lane's chunk, every pass updates the lower
// then the upper row of its voxel.  The rows of all lanes of all passes in one
// round are distinct (consecutive lanes, and the last lane of a pass and the
// first of the next, are C voxels apart, so their rows differ by at least
// C |dsv| > 1.01), so a round's reads of all passes are issued before its
// writes and the passes' LDS latencies overlap.  Rounds follow each other in
// program order.
template <int C, int NP>
__device__ inline void column_interleaved(float* __restrict__ strip, const float (&v)[NP][C], const ColumnCore& cc,
                                          const ColumnFx& fx, int cell0, int trash, int npass, const BpGrid& g,
                                          int lane) {
#pragma unroll
  for (int j = 0; j < C; ++j) {
    int i0[NP];
    float aa[NP], a1[NP], bb[NP], b1[NP];
#pragma unroll
    for (int p = 0; p < NP; ++p) {
      const int z = (p * 64 + lane) * C + j;
      const long long sj = fx.a + (long long)z * fx.d;
      const bool live = p < npass && sj >= fx.lo && sj <= fx.hi && z < g.nz;
      const float sw = cc.w * v[p][j];
      bb[p] = sw * fx_frac(sj);
      aa[p] = sw - bb[p];
      a1[p] = aa[p] * cc.fu;
      b1[p] = bb[p] * cc.fu;
      i0[p] = live ? cell0 + fx_row(sj) * kStripP : trash;
    }
    float2 x[NP];
#pragma unroll
    for (int p = 0; p < NP; ++p)
      if (p < npass) x[p] = lds_get2(strip, i0[p]);
#pragma unroll
    for (int p = 0; p < NP; ++p)
      if (p < npass) lds_put2(strip, i0[p], make_float2(x[p].x + (aa[p] - a1[p]), x[p].y + a1[p]));
    lds_order();
#pragma unroll
    for (int p = 0; p < NP; ++p)
      if (p < npass) x[p] = lds_get2(strip, i0[p] + kStripP);
#pragma unroll
    for (int p = 0; p < NP; ++p)
      if (p < npass) lds_put2(strip, i0[p] + kStripP, make_float2(x[p].x + (bb[p] - b1[p]), x[p].y + b1[p]));
    lds_order();
  }
}

// One pass of a column: lane l takes the C voxels of chunk ch0 + l.  `cell0`
// is the LDS index of (row 0, left tap); `trash` the lane's trash floats;
// `phases` is the number of interleaved lane phases, or 0 for the folded
// two-voxel updates (C = 2 and a v-slope in (0.505, 0.999)).
template <int C>
__device__ inline void project_pass(float* __restrict__ strip, const float (&v)[C], const ColumnCore& cc,
                                    const ColumnFx& fx, int cell0, int trash, int ch0, int phases,
                                    const BpGrid& g, int lane) {
  const int z0 = (ch0 + lane) * C;
  const long long s0 = fx.a + (long long)z0 * fx.d;
  if constexpr (C == 2) {
    if (phases == 0) {
      pass_updates_pair(strip, v, cc, s0, fx.d, fx.lo, fx.hi, z0, cell0, trash, g);
      return;
    }
  }
  if (phases == 1) {
    pass_updates<C>(strip, v, cc, s0, fx.d, fx.lo, fx.hi, z0, true, cell0, trash, g);
  } else {
    const int ml = lane % phases;
    for (int ph = 0; ph < phases; ++ph)
      pass_updates<C>(strip, v, cc, s0, fx.d, fx.lo, fx.hi, z0, ml == ph, cell0, trash, g);
  }
}

// Up to 64 kept voxel columns of one strip, lane i holding the i-th.
struct ColumnBatch {
  int n = 0;
  int iu = 0, x = 0, y = 0;
  float fu = 0.f, w = 0.f, dsv = 0.f;
  double kv = 0.0;
};

// The projector's per-column constants, computed by lane i for the i-th
// listed column of a batch (all 64 columns at once), then read by the whole
// wave with readlane when the column's turn comes.
struct ColumnPrep {
  int x = 0, y = 0;
  float fu = 0.f, w = 0.f;
  long long a = 0, d = 0;  // sv(z) = a + z d in 32.32 fixed point
  int phases = 1;          // lane phases, 0 = folded two-voxel updates
  int cell0 = 0;           // LDS index of (padded row 0, left tap)
};

__device__ inline long long bcast_ll(long long v, int src) {
  const int lo = __builtin_amdgcn_readlane((int)(v & 0xffffffffLL), src);
  const int hi = __builtin_amdgcn_readlane((int)(v >> 32), src);
  return ((long long)hi << 32) | (unsigned int)lo;
}

__device__ inline int permute_i(int dst, int v) { return __builtin_amdgcn_ds_permute(dst << 2, v); }
__device__ inline float permute_f(int dst, float v) {
  return __int_as_float(__builtin_amdgcn_ds_permute(dst << 2, __float_as_int(v)));
}

// Walks the strip's slabs four at a time, 16 lanes per slab, and appends the
// kept columns to the batch (each kept lane pushes its column to lane n + its
// rank; the other lanes push to the lanes past the new end), until the batch
// is full or the slabs are exhausted.  Returns false when nothing is left.
struct Enumerator {
  SlabRanges sr;
  int s0 = 0;   // first slab of the current group of four
  int rnd = 0;  // 16-column round within the group

  __device__ bool fill(ColumnBatch& cb, const BpView& c, const BpGrid& g, bool x_slabs, int n_slabs, int u_lo,
                       int u_hi, int lane) {
    cb.n = 0;
    for (;;) {
      if (s0 >= n_slabs) return cb.n > 0;
      const int sl = s0 + (lane >> 4);
      int a = 0, b = -1;
      if (sl < n_slabs) {
        const int sc = x_slabs ? g.x0 + sl : sl;
        sr.range(c, x_slabs, x_slabs ? voxel_x(g, sc) : voxel_y(g, sc), a, b);
        if (!x_slabs) {
          a = max(a, g.x0);
          b = min(b, g.x1 - 1);
        }
      }
      const int w = max(b - a + 1, 0);
      const int need = max(max(bcast_i(w, 0), bcast_i(w, 16)), max(bcast_i(w, 32), bcast_i(w, 48)));
      const int off = a + rnd * 16 + (lane & 15);
      const int o = min(off, max(b, a));
      const int xo = x_slabs ? g.x0 + sl : o, yo = x_slabs ? o : sl;
      const int x = min(xo, g.x1 - 1), y = min(yo, g.ny - 1);
      const double X = voxel_x(g, x), Y = voxel_y(g, y);
      const ColumnCore mine = column_core(c, X, Y, (float)X, (float)Y, g);
      const bool keep = sl < n_slabs && off <= b && mine.ok && mine.iu >= u_lo && mine.iu <= u_hi;
      const uint64_t m = __ballot(keep);
      const int cnt = (int)__popcll(m);
      if (cb.n + cnt > 64) return true;  // full: this round is taken again by the next fill
      if (++rnd * 16 >= need) {
        rnd = 0;
        s0 += 4;
      }
      if (!cnt) continue;
      const uint32_t rk = __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
      const int dst = keep ? cb.n + (int)rk : (cb.n + cnt + (lane - (int)rk)) & 63;
      const bool fresh = lane >= cb.n && lane < cb.n + cnt;
      const long long kvb = __double_as_longlong(mine.kv);
      const int kv_lo = permute_i(dst, (int)(kvb & 0xffffffffLL)), kv_hi = permute_i(dst, (int)(kvb >> 32));
      const int piu = permute_i(dst, mine.iu), px = permute_i(dst, x), py = permute_i(dst, y);
      const float pfu = permute_f(dst, mine.fu), pw = permute_f(dst, mine.w), pdsv = permute_f(dst, mine.dsv);
      if (fresh) {
        cb.iu = piu; cb.x = px; cb.y = py;
        cb.fu = pfu; cb.w = pw; cb.dsv = pdsv;
        cb.kv = __longlong_as_double(((long long)kv_hi << 32) | (unsigned int)kv_lo);
      }
      cb.n += cnt;
      if (cb.n == 64) return true;
    }
  }
};

// The strip's geometry in a view: owned detector columns and slab setup.
struct StripSetup {
  int pu0, u_lo, u_hi, n_slabs;
  bool x_slabs;
  double s_lo, s_hi;
  __device__ StripSetup(const BpView& c, const BpGrid& g, int st) {
    pu0 = st * kStripW;   // first padded column of the strip
    u_lo = pu0 - kPadLo;  // first detector column index owned
    u_hi = u_lo + kStripW - 1;
    x_slabs = fabs(c.cx) >= fabs(c.cy);
    // columns whose left tap iu satisfies u_lo <= iu <= u_hi, i.e. su in [u_lo, u_hi + 1)
    s_lo = (double)u_lo - 0.5;
    s_hi = (double)(u_hi + 1) + 0.5;
    n_slabs = x_slabs ? (g.x1 - g.x0) : g.ny;
  }
};

// Column lists of the matched projector, built once per geometry: for view k
// and strip st, the voxel columns whose left tap lands in the strip, in the
// enumeration order above, packed as x | y << 16.  The first pass (WRITE
// false) stores each list's length at offs[k * n_strips + st]; the second
// writes the list at list + offs[k * n_strips + st] (offsets by then).
template <bool WRITE>
__global__ __launch_bounds__(64) void project_plan_kernel(const BpView* __restrict__ views, int n_strips, BpGrid g,
                                                          uint32_t* __restrict__ offs, uint32_t* __restrict__ list) {
  const int k = blockIdx.x / n_strips;
  const int st = blockIdx.x - k * n_strips;
  const int lane = threadIdx.x;
  const BpView& c = views[k];
  const StripSetup ss(c, g, st);
  Enumerator en;
  en.sr.init(ss.s_lo, ss.s_hi, c, ss.x_slabs, g);
  ColumnBatch cb;
  const uint32_t base = WRITE ? offs[blockIdx.x] : 0u;
  uint32_t n = 0;
  while (en.fill(cb, c, g, ss.x_slabs, ss.n_slabs, ss.u_lo, ss.u_hi, lane)) {
    if (WRITE && lane < cb.n) list[base + n + lane] = (uint32_t)cb.x | ((uint32_t)cb.y << 16);
    n += (uint32_t)cb.n;
  }
  if (!WRITE && lane == 0) offs[blockIdx.x] = n;
}

template <int C, bool MASK>
__global__ __launch_bounds__(64) void project_kernel(const float* __restrict__ volT,
                                                     const BpView* __restrict__ views, int view0, int n_strips,
                                                     int nviews, BpGrid g, float* __restrict__ filt, int lds_r0,
                                                     int lds_rows,
                                                     const uint32_t* __restrict__ offs,
                                                     const uint32_t* __restrict__ list) {
  // [lds_rows][kStripP] (padded rows lds_r0.., column q = padded column pu0 + q - kStripC0), then kTrash floats
  extern __shared__ float strip[];
  // Block b runs on XCD b mod 8 (blocks are dealt to the XCDs round robin):
  // XCD x takes the views [x V / 8, (x + 1) V / 8) of the launch, in groups of
  // kProjViewGroup views, strip by strip within a group.  A voxel column lands
  // in nearby strips in nearby views, so the group's waves that read it run
  // together on one XCD and its L2 serves the column to most of them.
  const int xcd = (int)(blockIdx.x % 8u), jx = (int)(blockIdx.x / 8u);
  const int va = xcd * nviews / 8, vx = (xcd + 1) * nviews / 8 - va;
  if (jx >= vx * n_strips) return;  // (the whole one-wave block)
  const int grp = jx / (kProjViewGroup * n_strips), rem = jx - grp * kProjViewGroup * n_strips;
  const int gsize = min(kProjViewGroup, vx - grp * kProjViewGroup);
  const int st = rem / gsize;
  const int k = view0 + va + grp * kProjViewGroup + (rem - st * gsize);
  const int lane = threadIdx.x;
  const int trash = lds_rows * kStripP;
  for (int i = lane; i < trash + kTrash; i += 64) strip[i] = 0.f;
  __syncthreads();

  const BpView& c = views[k];
  const int pu0 = st * kStripW;  // first padded column of the strip
  const int nch = (g.nz + C - 1) / C;
  const int npass = (nch + 63) / 64;
  constexpr int NP = ColumnVals<C>::NP;
  const size_t cell = (size_t)k * n_strips + st;
  const uint32_t beg = offs[cell], end = offs[cell + 1];

  ColumnPrep cp;
  int ncols = 0;
  const ColumnFx win = {0, 0, (long long)g.lo_v * 4294967296LL, (long long)g.hi_v * 4294967296LL};
  // The volume values of the next kAhead columns are in flight while one is
  // accumulated: a column's accumulation takes less time than a load from HBM.
  auto fetch = [&](ColumnVals<C>& dst, int i) {
    const int x = bcast_i(cp.x, i), y = bcast_i(cp.y, i);
    load_column<C, MASK>(dst, volT + ((size_t)y * g.nx + x) * g.nz, x, y, g, lane, npass);
  };
  auto accumulate = [&](const ColumnVals<C>& vals, int i) {
    ColumnCore cc;
    cc.fu = bcast_f(cp.fu, i);
    cc.w = bcast_f(cp.w, i);
    cc.ok = true;
    const int phases = bcast_i(cp.phases, i);
    const int cell0 = bcast_i(cp.cell0, i);
    ColumnFx fx = win;
    fx.a = bcast_ll(cp.a, i);
    fx.d = bcast_ll(cp.d, i);
    if constexpr (C == 2) {
      // three folded passes: the first and last together, then the middle one
      if (npass == 3 && phases == 0) {
        const int za = 2 * lane, zb = 2 * (128 + lane);
        pass_updates_pair2(strip, vals.v[0], vals.v[2], cc, fx.a + (long long)za * fx.d,
                           fx.a + (long long)zb * fx.d, fx.d, fx.lo, fx.hi, za, zb, cell0, trash + 2 * lane, g);
        project_pass<C>(strip, vals.v[1], cc, fx, cell0, trash + 2 * lane, 64, phases, g, lane);
        return;
      }
    }
    if (npass <= NP && phases == 1) {  // every pass of an unfolded column at once
      column_interleaved<C, NP>(strip, vals.v, cc, fx, cell0, trash + 2 * lane, npass, g, lane);
      return;
    }
#pragma unroll
    for (int p = 0; p < NP; ++p)
      if (p < npass) project_pass<C>(strip, vals.v[p], cc, fx, cell0, trash + 2 * lane, p * 64, phases, g, lane);
    if (npass > NP) {  // columns longer than the prefetch window load the rest in place
      const int x = bcast_i(cp.x, i), y = bcast_i(cp.y, i);
      const float* col = volT + ((size_t)y * g.nx + x) * g.nz;
      for (int p = NP; p < npass; ++p) {
        float v[C];
        load_chunk<C, MASK>(v, col, p, x, y, g, lane);
        project_pass<C>(strip, v, cc, fx, cell0, trash + 2 * lane, p * 64, phases, g, lane);
      }
    }
  };
  // Batches of 64 listed columns: lane i prepares the i-th with the same column
  // arithmetic the enumeration used to select it.
  for (uint32_t b0 = beg; b0 < end; b0 += 64) {
    ncols = (int)min(64u, end - b0);
    if (lane < ncols) {
      const uint32_t e = list[b0 + lane];
      cp.x = (int)(e & 0xffffu);
      cp.y = (int)(e >> 16);
      const double X = voxel_x(g, cp.x), Y = voxel_y(g, cp.y);
      const ColumnCore mine = column_core(c, X, Y, (float)X, (float)Y, g);
      const ColumnFx f = column_fx(mine, c, g);
      cp.fu = mine.fu;
      cp.w = mine.w;
      cp.a = f.a;
      cp.d = f.d;
      const float step = (float)C * fabsf(mine.dsv);
      const int ph = step > 1.01f ? 1 : min(64, (int)ceilf(1.01f / fmaxf(step, 1e-6f)));
      cp.phases = (C == 2 && ph == 1 && mine.dsv > 0.f && mine.dsv < 0.999f) ? 0 : ph;
      cp.cell0 = (kPadLo - lds_r0) * kStripP + (mine.iu + kPadLo - pu0) + kStripC0;
    }
    ColumnVals<C> ring[kAhead];
#pragma unroll
    for (int q = 0; q < kAhead; ++q)
      if (q < ncols) fetch(ring[q], q);
    for (int i0 = 0; i0 < ncols; i0 += kAhead) {
#pragma unroll
      for (int q = 0; q < kAhead; ++q) {
        const int i = i0 + q;
        if (i < ncols) {
          const ColumnVals<C> vals = ring[q];
          if (i + kAhead < ncols) fetch(ring[q], i + kAhead);
          accumulate(vals, i);
        }
      }
    }
  }
  __syncthreads();
  // Strip column t (LDS column t + 1) is padded column pu0 + t.  Columns 1 ..
  // kStripW - 1 belong to this strip alone; column 0 also takes the previous
  // strip's spill, and the spill column kStripW is the next strip's column 0.
  // Those two are added atomically into the zeroed output: two addends onto
  // zero, so the sum is the same in either order.
  float* out = filt + (size_t)k * g.rows * g.pitch;
  constexpr int kOut = kStripW + 1;
  for (int i = lane; i < g.rows * kOut; i += 64) {
    const int r = i / kOut, t = i - r * kOut;
    const int lr = r - lds_r0;
    if (pu0 + t >= g.pitch) continue;
    float* o = out + (size_t)r * g.pitch + pu0 + t;
    const bool held = lr >= 0 && lr < lds_rows;
    if (t == 0 || t == kStripW) {
      if (held) {
        const float val = strip[lr * kStripP + t + kStripC0];
        if (val != 0.f) unsafeAtomicAdd(o, val);
      }
    } else {
      *o = held ? strip[lr * kStripP + t + kStripC0] : 0.f;
    }
  }
}

// ------------------------------------------------------------ L2 loss
// out = a - b and the FP64 sum of out^2: each block sums a fixed grid-stride
// share in a fixed order and reduces it in a fixed tree, so the result does
// not depend on scheduling.
__global__ __launch_bounds__(256) void residual_partial_kernel(const float* __restrict__ a,
                                                               const float* __restrict__ b, float* __restrict__ out,
                                                               size_t n, double* __restrict__ partial) {
  __shared__ double red[4];
  double acc = 0.0;
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) {
    const float r = a[i] - b[i];
    out[i] = r;
    acc = fma((double)r, (double)r, acc);
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) partial[blockIdx.x] = (red[0] + red[1]) + (red[2] + red[3]);
}

__global__ __launch_bounds__(256) void loss_total_kernel(const double* __restrict__ partial, int n,
                                                         double* __restrict__ loss) {
  __shared__ double red[4];
  double acc = 0.0;
  for (int i = threadIdx.x; i < n; i += 256) acc += partial[i];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) *loss = 0.5 * ((red[0] + red[1]) + (red[2] + red[3]));
}

hipError_t launch_residual_loss(const float* a, const float* b, float* out, size_t n, double* partial, double* loss,
                                hipStream_t s) {
  hipLaunchKernelGGL(residual_partial_kernel, dim3(kLossBlocks), dim3(256), 0, s, a, b, out, n, partial);
  hipLaunchKernelGGL(loss_total_kernel, dim3(1), dim3(256), 0, s, partial, kLossBlocks, loss);
  return hipGetLastError();
}

// vol [z][y][x] -> volT [y][x][z] through 32 x 32 LDS tiles (per y).
__global__ __launch_bounds__(256) void transpose_zx_kernel(const float* __restrict__ vol, float* __restrict__ volT,
                                                           int nx, int ny, int nz) {
  __shared__ float t[32][33];
  const int x0 = blockIdx.x * 32, z0 = blockIdx.y * 32, y = blockIdx.z;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
  for (int r = ty; r < 32; r += 8) {
    const int z = z0 + r, x = x0 + tx;
    t[r][tx] = (z < nz && x < nx) ? vol[((size_t)z * ny + y) * nx + x] : 0.f;
  }
  __syncthreads();
  for (int r = ty; r < 32; r += 8) {
    const int x = x0 + r, z = z0 + tx;
    if (x < nx && z < nz) volT[((size_t)y * nx + x) * nz + z] = t[tx][r];
  }
}

hipError_t launch_transpose_zx(const float* vol, float* volT, int nx, int ny, int nz, hipStream_t s) {
  dim3 gd((nx + 31) / 32, (nz + 31) / 32, ny);
  hipLaunchKernelGGL(transpose_zx_kernel, gd, dim3(256), 0, s, vol, volT, nx, ny, nz);
  return hipGetLastError();
}

int project_strips(const BpGrid& grid) { return (grid.pitch + kStripW - 1) / kStripW; }

hipError_t launch_project_plan(const BpView* views, int n_views, const BpGrid& grid, uint32_t* offs, uint32_t* list,
                               bool write, hipStream_t s) {
  const long nblocks = (long)project_strips(grid) * n_views;
  if (nblocks <= 0) return hipSuccess;
  if (nblocks > 0x7fffffffL || grid.nx > 65536 || grid.ny > 65536) return hipErrorInvalidValue;
  if (write)
    hipLaunchKernelGGL(project_plan_kernel<true>, dim3((unsigned)nblocks), dim3(64), 0, s, views,
                       project_strips(grid), grid, offs, list);
  else
    hipLaunchKernelGGL(project_plan_kernel<false>, dim3((unsigned)nblocks), dim3(64), 0, s, views,
                       project_strips(grid), grid, offs, list);
  return hipGetLastError();
}

hipError_t launch_project(const float* volT, const BpView* views, int view0, int nviews_chunk, const BpGrid& grid,
                          float* filt_pad, int lds_r0, int lds_rows, const uint32_t* plan_offs,
                          const uint32_t* plan_list, const monte_tuning& t, hipStream_t s) {
  const int n_strips = project_strips(grid);
  // (8 XCDs x the largest view share x strips; blocks past an XCD's share return)
  const long nblocks = (long)n_strips * 8 * ((nviews_chunk + 7) / 8);
  if (nblocks <= 0) return hipSuccess;
  if (nblocks > 0x7fffffffL) return hipErrorInvalidValue;
  if (lds_r0 < 0 || lds_rows < 1 || lds_r0 + lds_rows > grid.rows) return hipErrorInvalidValue;
  const size_t lds = sizeof(float) * ((size_t)lds_rows * kStripP + kTrash);
  if (lds > 64 * 1024) return hipErrorInvalidValue;
  // voxels per lane: 2 unless tuned otherwise.  Two voxels per lane need the
  // fewest dependent LDS read-add-writes per pass: 2 and 4 voxels per lane
  // took 10.6 and 13.8 ms at 325^3 / 512^2 / 360 views, 86 and 96 ms at
  // 512^3 / 1024 x 768 / 720 views (and 8 voxels per lane 163 ms).
  const int best = t.pj_zchunk ? t.pj_zchunk : 2;
  const dim3 gd((unsigned)nblocks), bd(64);
#define MONTE_PJ(CZ, MK)                                                                                      \
  hipLaunchKernelGGL((project_kernel<CZ, MK>), gd, bd, lds, s, volT, views, view0, n_strips, nviews_chunk, grid,   \
                     filt_pad,                                                                                  \
                     lds_r0, lds_rows, plan_offs, plan_list)
  const bool mk = grid.mask != 0;
  switch (best) {
    case 2: if (mk) MONTE_PJ(2, true); else MONTE_PJ(2, false); break;
    case 4: if (mk) MONTE_PJ(4, true); else MONTE_PJ(4, false); break;
    case 8: if (mk) MONTE_PJ(8, true); else MONTE_PJ(8, false); break;
    default: return hipErrorInvalidValue;
  }
#undef MONTE_PJ
  return hipGetLastError();
}

}  // namespace monte
