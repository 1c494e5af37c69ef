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
// Convolution core shared by the forward and transposed filters.  LDS holds
// A[TV][nu+1] (input rows, +1 pad against bank conflicts) and the taps
// H[2nu-1] (centre nu-1).  Wave w computes rows [w*RW, (w+1)*RW), lane l the
// columns b = l + 64 m, m < NB:  out[j][b] = sum_c A[j][c] H[c - b + nu - 1].
template <int NB, int RW>
__device__ inline void conv_rows(const float* __restrict__ A, const float* __restrict__ H, int nu, int lda,
                                 float (&acc)[RW][NB]) {
  const int lane = threadIdx.x & 63;
  const int w = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < RW; ++q)
#pragma unroll
    for (int m = 0; m < NB; ++m) acc[q][m] = 0.f;
  int hb[NB];
#pragma unroll
  for (int m = 0; m < NB; ++m) {
    int b = lane + 64 * m;
    if (b > nu - 1) b = nu - 1;  // lanes past the row compute a discarded copy
    hb[m] = nu - 1 - b;
  }
  const float* Arow = A + (w * RW) * lda;
  for (int c = 0; c < nu; ++c) {
    float a[RW];
#pragma unroll
    for (int q = 0; q < RW; ++q) a[q] = Arow[q * lda + c];
#pragma unroll
    for (int m = 0; m < NB; ++m) {
      const float h = H[c + hb[m]];
#pragma unroll
      for (int q = 0; q < RW; ++q) acc[q][m] = fmaf(a[q], h, acc[q][m]);
    }
  }
}

__device__ inline float cos_weight(const FilterParams& fp, int r, int c) {
  const float a = fmaf(fp.b_u, (float)r, fp.a_u);
  const float b = fmaf(fp.b_v, (float)c, fp.a_v);
  return fp.D / sqrtf(fmaf(fp.D, fp.D, fmaf(a, a, b * b)));
}

// proj [view][nu][nv]  ->  filt_pad [view][rows][pitch] (interior at +kPadLo)
template <int NB, int TV>
__global__ __launch_bounds__(256) void filter_fwd_kernel(const float* __restrict__ proj, float* __restrict__ out,
                                                         const float* __restrict__ taps, FilterParams fp) {
  constexpr int RW = TV / 4;
  extern __shared__ float sm[];
  const int nu = fp.nu, nv = fp.nv;
  const int lda = nu + 1;
  float* A = sm;
  float* H = sm + TV * lda;
  const int view = blockIdx.y;
  const int iv0 = blockIdx.x * TV;
  const int tid = threadIdx.x;
  for (int i = tid; i < 2 * nu - 1; i += 256) H[i] = taps[i];
  const float* src = proj + (size_t)view * nu * nv;
  for (int i = tid; i < nu * TV; i += 256) {
    const int c = i / TV, j = i - c * TV;  // detector row c (u), column iv0 + j (v)
    const int iv = iv0 + j;
    float v = 0.f;
    if (iv < nv) v = src[(size_t)c * nv + iv] * cos_weight(fp, c, iv);
    A[j * lda + c] = v;
  }
  __syncthreads();
  float acc[RW][NB];
  conv_rows<NB, RW>(A, H, nu, lda, acc);
  const int lane = tid & 63, w = tid >> 6;
  float* dst = out + (size_t)view * fp.rows * fp.pitch;
#pragma unroll
  for (int q = 0; q < RW; ++q) {
    const int iv = iv0 + w * RW + q;
    if (iv >= nv) continue;
#pragma unroll
    for (int m = 0; m < NB; ++m) {
      const int b = lane + 64 * m;
      if (b < nu) dst[(size_t)(iv + kPadLo) * fp.pitch + b + kPadLo] = fp.cs * acc[q][m];
    }
  }
}

// filt_pad [view][rows][pitch] (interior)  ->  proj [view][nu][nv], weighted:
// the transpose of filter_fwd (the taps are symmetric).
template <int NB, int TV>
__global__ __launch_bounds__(256) void filter_T_kernel(const float* __restrict__ in, float* __restrict__ proj,
                                                       const float* __restrict__ taps, FilterParams fp) {
  constexpr int RW = TV / 4;
  extern __shared__ float sm[];
  const int nu = fp.nu, nv = fp.nv;
  const int lda = nu + 1;
  float* A = sm;
  float* H = sm + TV * lda;
  const int view = blockIdx.y;
  const int iv0 = blockIdx.x * TV;
  const int tid = threadIdx.x;
  for (int i = tid; i < 2 * nu - 1; i += 256) H[i] = taps[i];
  const float* src = in + (size_t)view * fp.rows * fp.pitch;
  for (int i = tid; i < nu * TV; i += 256) {
    const int j = i / nu, b = i - j * nu;
    const int iv = iv0 + j;
    A[j * lda + b] = iv < nv ? src[(size_t)(iv + kPadLo) * fp.pitch + b + kPadLo] : 0.f;
  }
  __syncthreads();
  float acc[RW][NB];
  conv_rows<NB, RW>(A, H, nu, lda, acc);
  __syncthreads();
  // stage results in LDS as [c][j] so the [u][v] output is written in runs along v
  const int lane = tid & 63, w = tid >> 6;
#pragma unroll
  for (int q = 0; q < RW; ++q)
#pragma unroll
    for (int m = 0; m < NB; ++m) {
      const int c = lane + 64 * m;
      if (c < nu) A[c * TV + w * RW + q] = acc[q][m];
    }
  __syncthreads();
  float* dst = proj + (size_t)view * nu * nv;
  for (int i = tid; i < nu * TV; i += 256) {
    const int c = i / TV, j = i - c * TV;
    const int iv = iv0 + j;
    if (iv < nv) dst[(size_t)c * nv + iv] = fp.cs * A[c * TV + j] * cos_weight(fp, c, iv);
  }
}

template <template <int, int> class K>
struct FilterLaunch;

static int filter_nb(int nu) {
  int nb = (nu + 63) / 64;
  if (nb <= 1) return 1;
  if (nb <= 2) return 2;
  if (nb <= 4) return 4;
  if (nb <= 8) return 8;
  if (nb <= 16) return 16;
  return -1;
}

#define MONTE_FILTER_DISPATCH(KERNEL, NBV, TVV, ...)                                                   \
  do {                                                                                                \
    if (TVV == 16) {                                                                                  \
      switch (NBV) {                                                                                  \
        case 1: hipLaunchKernelGGL((KERNEL<1, 16>), __VA_ARGS__); break;                              \
        case 2: hipLaunchKernelGGL((KERNEL<2, 16>), __VA_ARGS__); break;                              \
        case 4: hipLaunchKernelGGL((KERNEL<4, 16>), __VA_ARGS__); break;                              \
        case 8: hipLaunchKernelGGL((KERNEL<8, 16>), __VA_ARGS__); break;                              \
        default: return hipErrorInvalidValue;                                                         \
      }                                                                                               \
    } else {                                                                                          \
      switch (NBV) {                                                                                  \
        case 8: hipLaunchKernelGGL((KERNEL<8, 8>), __VA_ARGS__); break;                               \
        case 16: hipLaunchKernelGGL((KERNEL<16, 8>), __VA_ARGS__); break;                             \
        default: return hipErrorInvalidValue;                                                         \
      }                                                                                               \
    }                                                                                                 \
  } while (0)

hipError_t launch_fdk_filter(const float* proj, float* filt_pad, const float* taps, const FilterParams& fp,
                             hipStream_t s) {
  const int nb = filter_nb(fp.nu);
  if (nb < 0) return hipErrorInvalidValue;
  const int tv = fp.nu <= 512 ? 16 : 8;
  const size_t lds = sizeof(float) * ((size_t)tv * (fp.nu + 1) + 2 * fp.nu - 1);
  dim3 grid((fp.nv + tv - 1) / tv, fp.n_views);
  MONTE_FILTER_DISPATCH(filter_fwd_kernel, nb, tv, grid, dim3(256), lds, s, proj, filt_pad, taps, fp);
  return hipGetLastError();
}

hipError_t launch_fdk_filter_T(const float* filt_pad, float* proj, const float* taps, const FilterParams& fp,
                               hipStream_t s) {
  const int nb = filter_nb(fp.nu);
  if (nb < 0) return hipErrorInvalidValue;
  const int tv = fp.nu <= 512 ? 16 : 8;
  size_t lds = sizeof(float) * ((size_t)tv * (fp.nu + 1) + 2 * fp.nu - 1);
  const size_t stage = sizeof(float) * (size_t)fp.nu * tv;
  if (stage > lds) lds = stage;
  dim3 grid((fp.nv + tv - 1) / tv, fp.n_views);
  MONTE_FILTER_DISPATCH(filter_T_kernel, nb, tv, grid, dim3(256), lds, s, filt_pad, proj, taps, fp);
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

__global__ void fill_pads_kernel(float* __restrict__ buf, int n_views, int nu, int nv, int pitch, int rows,
                                 int wrap) {
  const size_t n = (size_t)n_views * rows * pitch;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const int k = (int)(i / ((size_t)rows * pitch));
    const int rem = (int)(i - (size_t)k * rows * pitch);
    const int r = rem / pitch - kPadLo, c = rem % pitch - kPadLo;
    if (r >= 0 && r < nv && c >= 0 && c < nu) continue;
    float v = 0.f;
    int kk, rr, cc;
    if (wrap && wrap_source(k, r, c, n_views, nu, nv, kk, rr, cc))
      v = buf[((size_t)kk * rows + rr + kPadLo) * pitch + cc + kPadLo];
    buf[i] = v;
  }
}

hipError_t launch_fill_pads(float* buf, int n_views, int nu, int nv, int pitch, int rows, int wrap, hipStream_t s) {
  const size_t n = (size_t)n_views * rows * pitch;
  size_t blocks = (n + 255) / 256;
  if (blocks > 16384) blocks = 16384;
  hipLaunchKernelGGL(fill_pads_kernel, dim3((unsigned)blocks), dim3(256), 0, s, buf, n_views, nu, nv, pitch, rows,
                     wrap);
  return hipGetLastError();
}

// Transpose of the bp3d20 guard-cell wrap: what the matched projector
// deposited on a guard cell belongs to the interior cell it aliases.
__global__ void fold_pads_kernel(float* __restrict__ buf, int n_views, int nu, int nv, int pitch, int rows) {
  const size_t n = (size_t)n_views * rows * pitch;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const int k = (int)(i / ((size_t)rows * pitch));
    const int rem = (int)(i - (size_t)k * rows * pitch);
    const int r = rem / pitch - kPadLo, c = rem % pitch - kPadLo;
    if (r >= 0 && r < nv && c >= 0 && c < nu) continue;
    int kk, rr, cc;
    const float v = buf[i];
    if (v != 0.f && wrap_source(k, r, c, n_views, nu, nv, kk, rr, cc))
      atomicAdd(buf + ((size_t)kk * rows + rr + kPadLo) * pitch + cc + kPadLo, v);
  }
}

hipError_t launch_fold_pads(float* buf, int n_views, int nu, int nv, int pitch, int rows, hipStream_t s) {
  const size_t n = (size_t)n_views * rows * pitch;
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
struct ColumnView {
  int iu;       // u pixel index of the left tap (sample window clamped)
  float fu;     // u fraction
  int ivb;      // integer part of sv at the base voxel zb
  float fvb;    // its fraction
  float dsv;    // sv increment per voxel step along z
  float flo, fhi;  // v window relative to ivb
  float w;      // distance weight x angular factor
  bool ok;      // su inside the window
};

__device__ inline ColumnView column_view(const BpView& c, double X, double Y, float Xf, float Yf, double Zb,
                                         const BpGrid& g) {
  ColumnView r;
  const double U = fma(c.cx, X, fma(c.cy, Y, c.c0));
  const double inv = 1.0 / U;
  double su = fma(c.suk * fma(c.ax, X, c.ay * Y), inv, c.su0);
  r.ok = su >= g.lo_u && su <= g.hi_u;
  su = fmin(fmax(su, g.lo_u), g.hi_u);
  const double fu0 = floor(su);
  r.iu = (int)fu0;
  r.fu = (float)(su - fu0);
  const double k = c.svk * inv;
  double sv = fma(k, Zb, c.sv0);
  sv = fmin(fmax(sv, -1.0e6), 1.0e6);
  const double fv0 = floor(sv);
  r.ivb = (int)fv0;
  r.fvb = (float)(sv - fv0);
  r.dsv = (float)(k * g.gsz);
  r.flo = (float)(g.lo_v - fv0);
  r.fhi = (float)(g.hi_v - fv0);
  // sign test in unfused FP64 (as recon/bp3d20_325.cpp:145,151 evaluates it)
  const double gp = __dadd_rn(__dmul_rn(X, c.gx), __dmul_rn(Y, c.gy));
  const float hp = fabsf(fmaf(c.hx, Xf, c.hy * Yf));
  const float dd = fmaf(c.dx, Xf, fmaf(c.dy, Yf, c.d0)) - (gp < 0.0 ? -hp : hp);
  const float q = c.wn / dd;
  r.w = c.fac * q * q;
  return r;
}

// Sample position of voxel j of the chunk: row index and fraction, and whether
// it lies in the v window.  (Identical in both directions.)
__device__ inline void column_row(const ColumnView& cv, int j, int& row, float& fv, bool& okv) {
  float f = fmaf((float)j, cv.dsv, cv.fvb);
  okv = f >= cv.flo && f <= cv.fhi;
  f = fminf(fmaxf(f, cv.flo), cv.fhi);
  const float fl = floorf(f);
  fv = f - fl;
  row = cv.ivb + (int)fl;
}

__device__ inline bool in_mask(const BpGrid& g, int x, int y, int z) {
  if (!g.mask) return true;
  const int dz = z - g.mcz, dy = y - g.mcy, dx = x - g.mcx;
  return (float)(dz * dz + dy * dy + dx * dx) <= g.mask_r2;
}

// The v coordinate of voxel z is formed from a base taken at the 16-aligned z
// at or below it (zb), whatever the thread's z-chunk, so every chunking and
// tile shape evaluates the same floating-point expression for each voxel and
// the volume is bitwise independent of the launch tuning.
constexpr int kZBase = 16;

struct ColumnPos {
  int x, y, z0;
  int joff;           // z0 - zb
  double X, Y, Zb;    // Zb: Z of the base voxel zb
  float Xf, Yf;
};

__device__ inline bool column_pos(const BpGrid& g, int tile_x, int tiles_x, int tiles_y, int zc_size,
                                  ColumnPos& p) {
  const int t = xcd_tile(blockIdx.x, gridDim.x);
  const int per_z = tiles_x * tiles_y;
  const int zc = t / per_z;
  const int rem = t - zc * per_z;
  const int ty_t = rem / tiles_x, tx_t = rem - ty_t * tiles_x;
  const int tile_y = 256 / tile_x;
  p.x = g.x0 + tx_t * tile_x + (int)(threadIdx.x % tile_x);
  p.y = ty_t * tile_y + (int)(threadIdx.x / tile_x);
  p.z0 = zc * zc_size;
  const int zb = (p.z0 / kZBase) * kZBase;
  p.joff = p.z0 - zb;
  // recon/bp3d20_325.cpp:107-109 form the coordinate as start + index * step, unfused
  p.X = __dadd_rn(g.gx0, __dmul_rn((double)p.x, g.gsx));
  p.Y = __dadd_rn(g.gy0, __dmul_rn((double)p.y, g.gsy));
  p.Zb = __dadd_rn(g.gz0, __dmul_rn((double)zb, g.gsz));
  p.Xf = (float)p.X;
  p.Yf = (float)p.Y;
  return p.x < g.x1 && p.y < g.ny;
}

template <int ZC>
__global__ __launch_bounds__(256) void backproject_kernel(const float* __restrict__ filt,
                                                          const BpView* __restrict__ views, int view0, int nviews,
                                                          BpGrid g, float* __restrict__ vol, int accumulate,
                                                          int tile_x, int tiles_x, int tiles_y) {
  ColumnPos cp;
  if (!column_pos(g, tile_x, tiles_x, tiles_y, ZC, cp)) return;
  bool live[ZC];
#pragma unroll
  for (int j = 0; j < ZC; ++j) live[j] = (cp.z0 + j) < g.nz && in_mask(g, cp.x, cp.y, cp.z0 + j);
  float acc[ZC];
#pragma unroll
  for (int j = 0; j < ZC; ++j) acc[j] = 0.f;

  const size_t view_stride = (size_t)g.rows * g.pitch;
  for (int k = view0; k < view0 + nviews; ++k) {
    const ColumnView cv = column_view(views[k], cp.X, cp.Y, cp.Xf, cp.Yf, cp.Zb, g);
    const float* base = filt + (size_t)k * view_stride + (cv.iu + kPadLo);
    const float wu = cv.ok ? cv.w : 0.f;
#pragma unroll
    for (int j = 0; j < ZC; ++j) {
      int row;
      float fv;
      bool okv;
      column_row(cv, cp.joff + j, row, fv, okv);
      const float* p = base + (size_t)(row + kPadLo) * g.pitch;
      const float p00 = p[0], p01 = p[1], p10 = p[g.pitch], p11 = p[g.pitch + 1];
      const float a = fmaf(cv.fu, p01 - p00, p00);
      const float b = fmaf(cv.fu, p11 - p10, p10);
      const float val = fmaf(fv, b - a, a);
      acc[j] = fmaf((okv && live[j]) ? wu : 0.f, val, acc[j]);
    }
  }
#pragma unroll
  for (int j = 0; j < ZC; ++j) {
    if (cp.z0 + j >= g.nz) break;
    float* o = vol + ((size_t)(cp.z0 + j) * g.ny + cp.y) * g.nx + cp.x;
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

hipError_t launch_backproject(const float* filt_pad, const BpView* views, int view0, int nviews_chunk,
                              const BpGrid& grid, float* vol, int accumulate, const monte_tuning& t,
                              hipStream_t s) {
  const int zc = t.bp_zchunk > 0 ? t.bp_zchunk : 8;
  int tile_x, tiles_x, tiles_y;
  long nblocks;
  if (t.bp_tile_x > 0 && 256 % t.bp_tile_x) return hipErrorInvalidValue;
  bp_tiling(grid, t, zc, tile_x, tiles_x, tiles_y, nblocks);
  if (nblocks <= 0) return hipSuccess;
  if (nblocks > 0x7fffffffL) return hipErrorInvalidValue;
  dim3 gd((unsigned)nblocks), bd(256);
  switch (zc) {
    case 4: hipLaunchKernelGGL((backproject_kernel<4>), gd, bd, 0, s, filt_pad, views, view0, nviews_chunk, grid, vol, accumulate, tile_x, tiles_x, tiles_y); break;
    case 8: hipLaunchKernelGGL((backproject_kernel<8>), gd, bd, 0, s, filt_pad, views, view0, nviews_chunk, grid, vol, accumulate, tile_x, tiles_x, tiles_y); break;
    case 16: hipLaunchKernelGGL((backproject_kernel<16>), gd, bd, 0, s, filt_pad, views, view0, nviews_chunk, grid, vol, accumulate, tile_x, tiles_x, tiles_y); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

// ------------------------------------------- transposed backprojection
// The matched projector.  Every (voxel, view) pair deposits vol * w *
// bilinear weight on the four taps the backprojector gathers from, with the
// same column_view / column_row arithmetic, so <BP q, x> == <q, BP^T x> up to
// the order of the sums.  A block owns a tile of voxel columns x ZT voxels
// along z (values kept in registers) and walks the views: for each view it
// finds the tile's detector footprint (block min/max of the taps touched),
// accumulates into a privatised copy of that footprint in LDS with LDS
// atomics, then adds the footprint to the global stack with one global atomic
// per touched pixel.  Footprints larger than the LDS patch fall back to global
// atomics for that view.  Taps on guard cells are dropped or folded by the
// callers.
constexpr int kPatch = 8192;  // LDS floats per block

__device__ inline int wave_min_i(int v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o, 64));
  return v;
}
__device__ inline int wave_max_i(int v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, 64));
  return v;
}

template <int ZT>
__global__ __launch_bounds__(256) void project_kernel(const float* __restrict__ vol,
                                                      const BpView* __restrict__ views, int view0, int nviews,
                                                      BpGrid g, float* __restrict__ filt, int tile_x, int tiles_x,
                                                      int tiles_y) {
  __shared__ float patch[kPatch];
  __shared__ int bounds[4];  // umin, umax, vmin, vmax over the block
  ColumnPos cp;
  const bool col_ok = column_pos(g, tile_x, tiles_x, tiles_y, ZT, cp);
  float val[ZT];
  bool any = false;
#pragma unroll
  for (int j = 0; j < ZT; ++j) {
    const int z = cp.z0 + j;
    val[j] = (col_ok && z < g.nz && in_mask(g, cp.x, cp.y, z)) ? vol[((size_t)z * g.ny + cp.y) * g.nx + cp.x] : 0.f;
    any = any || val[j] != 0.f;
  }
  const int tid = threadIdx.x;
  const size_t view_stride = (size_t)g.rows * g.pitch;
  for (int k = view0; k < view0 + nviews; ++k) {
    ColumnView cv{};
    int umin = 1 << 30, umax = -(1 << 30), vmin = 1 << 30, vmax = -(1 << 30);
    const bool act = col_ok && any;
    if (act) {
      cv = column_view(views[k], cp.X, cp.Y, cp.Xf, cp.Yf, cp.Zb, g);
      if (cv.ok) {
        umin = cv.iu;
        umax = cv.iu + 1;
        int r0, r1;
        float f;
        bool o;
        column_row(cv, cp.joff, r0, f, o);
        column_row(cv, cp.joff + ZT - 1, r1, f, o);
        vmin = min(r0, r1);
        vmax = max(r0, r1) + 1;
      }
    }
    umin = wave_min_i(umin); umax = wave_max_i(umax);
    vmin = wave_min_i(vmin); vmax = wave_max_i(vmax);
    if (tid == 0) { bounds[0] = 1 << 30; bounds[1] = -(1 << 30); bounds[2] = 1 << 30; bounds[3] = -(1 << 30); }
    __syncthreads();
    if ((tid & 63) == 0) {
      atomicMin(&bounds[0], umin); atomicMax(&bounds[1], umax);
      atomicMin(&bounds[2], vmin); atomicMax(&bounds[3], vmax);
    }
    __syncthreads();
    const int u0 = bounds[0], v0 = bounds[2];
    const int fu = bounds[1] - u0 + 1, fvn = bounds[3] - v0 + 1;
    if (fu <= 0 || fvn <= 0) {  // nothing of this tile lands on the detector
      __syncthreads();
      continue;
    }
    float* out = filt + (size_t)k * view_stride;
    const bool lds = (long)fu * fvn <= kPatch;
    if (lds)
      for (int i = tid; i < fu * fvn; i += 256) patch[i] = 0.f;
    __syncthreads();
    if (act && cv.ok) {
#pragma unroll
      for (int j = 0; j < ZT; ++j) {
        int row;
        float fv;
        bool okv;
        column_row(cv, cp.joff + j, row, fv, okv);
        if (!okv || val[j] == 0.f) continue;
        const float sw = cv.w * val[j];
        const float b = sw * fv, a = sw - b;
        const float a1 = a * cv.fu, b1 = b * cv.fu;
        if (lds) {
          float* p = patch + (row - v0) * fu + (cv.iu - u0);
          atomicAdd(p, a - a1);
          atomicAdd(p + 1, a1);
          atomicAdd(p + fu, b - b1);
          atomicAdd(p + fu + 1, b1);
        } else {
          float* p = out + (size_t)(row + kPadLo) * g.pitch + cv.iu + kPadLo;
          atomicAdd(p, a - a1);
          atomicAdd(p + 1, a1);
          atomicAdd(p + g.pitch, b - b1);
          atomicAdd(p + g.pitch + 1, b1);
        }
      }
    }
    __syncthreads();
    if (lds) {
      for (int i = tid; i < fu * fvn; i += 256) {
        const float v = patch[i];
        if (v != 0.f) {
          const int r = i / fu, c = i - r * fu;
          atomicAdd(out + (size_t)(v0 + r + kPadLo) * g.pitch + u0 + c + kPadLo, v);
        }
      }
    }
    __syncthreads();
  }
}

hipError_t launch_project(const float* vol, const BpView* views, int view0, int nviews_chunk, const BpGrid& grid,
                          float* filt_pad, const monte_tuning& t, hipStream_t s) {
  constexpr int ZT = 16;
  int tile_x, tiles_x, tiles_y;
  long nblocks;
  monte_tuning tt = t;
  tt.bp_tile_x = 16;
  bp_tiling(grid, tt, ZT, tile_x, tiles_x, tiles_y, nblocks);
  if (nblocks <= 0) return hipSuccess;
  if (nblocks > 0x7fffffffL) return hipErrorInvalidValue;
  hipLaunchKernelGGL((project_kernel<ZT>), dim3((unsigned)nblocks), dim3(256), 0, s, vol, views, view0,
                     nviews_chunk, grid, filt_pad, tile_x, tiles_x, tiles_y);
  return hipGetLastError();
}

}  // namespace monte
