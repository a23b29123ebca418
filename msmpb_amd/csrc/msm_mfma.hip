// msm_mfma.hip -- level stencil correlations as implicit GEMMs on the matrix
// cores of gfx950, and the prolongation pass that combines the levels.
//
// For level l < L the stencil pass of grid2grid (code/forcefield0.c:898-931
// of hkayabilisim/msmpb) is
//     c_l[x][y][z] = sum_{kx,ky,kz} K[kx][ky][kz] q[x+kx-cx][y+ky-cy][z+kz-cz]
// (periodic, c = s/2).  Along z it is a banded Toeplitz product: for one
// output plane x, input plane xi = x+kx-cx and row shift ky,
//     C(y, z) += sum_k A(y, k) B(k, z),   A(y, k) = q[xi][y+ky-cy][z0-cz+k],
//                                         B(k, z) = K[kx][ky][k-z] (0 <= k-z < s_z).
// One wavefront owns a 16 x 16 (y, z) output tile of one plane and a range
// of kx; it stages each input slab (16+s_y-1 rows x K columns) in LDS and
// issues one 16x16xK MFMA per (ky, K-chunk).  B depends only on
// (kx, ky, chunk), so it is prepared once per box on the host in the exact
// per-lane operand layout and streamed from L2.
//   FP64:  v_mfma_f64_16x16x4_f64, K chunks of 4.
//   mixed: v_mfma_f32_16x16x32_bf16, K chunks of 32, q and K split into
//          bf16 hi + lo, three products (hi*hi + hi*lo + lo*hi) accumulated
//          in FP32.
// Every level's tiles go into a single launch (the stencil passes of
// different levels are independent); each kx-group writes its own partial
// grid, and k_prolong_sum adds the partials and the prolongated coarser
// potential (forcefield1.c:530-567) level by level, deterministically.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "msm_kernels.h"

namespace msm {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef double f64x4 __attribute__((ext_vector_type(4)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ int wrapm(int v, int m) {
  int r = v % m;
  return r < 0 ? r + m : r;
}

// float -> bf16 bits, round to nearest even (inputs are finite)
__host__ __device__ inline uint16_t bf16_bits(float f) {
  uint32_t u;
  std::memcpy(&u, &f, 4);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__host__ __device__ inline float bf16_float(uint16_t h) {
  uint32_t u = (uint32_t)h << 16;
  float f;
  std::memcpy(&f, &u, 4);
  return f;
}

template <int PREC>
struct Cfg;
template <>
struct Cfg<MSM_G2G_FP64> {
  static constexpr int KC = 4;     // K per MFMA
  static constexpr int KPL = 1;    // K elements per lane
  static constexpr int PAD = 1;    // LDS row padding (elements)
};
template <>
struct Cfg<MSM_G2G_BF16X3> {
  static constexpr int KC = 32;
  static constexpr int KPL = 8;
  static constexpr int PAD = 8;
};

// B operands of one (kx, ky) row: NCH chunks (x2 for the bf16 hi/lo pair).
template <int PREC, int NCH>
struct BRow;
template <int NCH>
struct BRow<MSM_G2G_FP64, NCH> {
  double b[NCH];
  __device__ void load(const void *base, long f0, int lane) {
    const double *p = reinterpret_cast<const double *>(base) + f0 * 64 + lane;
#pragma unroll
    for (int c = 0; c < NCH; ++c) b[c] = p[c * 64];
  }
};
template <int NCH>
struct BRow<MSM_G2G_BF16X3, NCH> {
  u32x4 hi[NCH], lo[NCH];
  __device__ void load(const void *base, long f0, int lane) {
    const u32x4 *p = reinterpret_cast<const u32x4 *>(base) + f0 * 128 + lane;
#pragma unroll
    for (int c = 0; c < NCH; ++c) {
      hi[c] = p[c * 128];
      lo[c] = p[c * 128 + 64];
    }
  }
};

template <int PREC, int NCH>
__global__ __launch_bounds__(64) void k_g2g_tiles(G2GArgs args) {
  chain_issue_priority(args.urgent);
  extern __shared__ __align__(16) unsigned char lds_raw[];
  constexpr int KC = Cfg<PREC>::KC, KPL = Cfg<PREC>::KPL, PAD = Cfg<PREC>::PAD;
  long task = blockIdx.x;
  int li = 0;
  while (li + 1 < args.nlev && task >= args.lv[li + 1].task0) ++li;
  const G2GLevel &L = args.lv[li];
  long t = task - L.task0;
  int grp = t % L.G;
  t /= L.G;
  int zt = t % L.tz;
  t /= L.tz;
  int yt = t % L.ty;
  int x = (int)(t / L.ty);
  const GridDims g = L.g, s = L.s;
  int per = (s.x + L.G - 1) / L.G;
  int kx0 = grp * per, kx1 = min(s.x, kx0 + per);
  int y0 = yt * 16, z0 = zt * 16;
  const int R = 16 + s.y - 1, W = NCH * KC, Wp = W + PAD;
  const int lane = threadIdx.x;
  const int row = lane & 15, kg = lane >> 4;
  f32x4 acc32 = {0.f, 0.f, 0.f, 0.f};
  f64x4 acc64 = {0.0, 0.0, 0.0, 0.0};
  for (int kx = kx0; kx < kx1; ++kx) {
    int xi = wrapm(x + kx - s.x / 2, g.x);
    const double *plane = L.q + (long)xi * g.y * g.z;
    BRow<PREC, NCH> bnext;
    bnext.load(L.bfrag, (long)(kx * s.y) * NCH, lane);  // row ky = 0, in flight during the slab load
    __syncthreads();  // previous slab fully consumed
    // slab: lane = column (periodic z index computed once), rows uniform
    for (int cb = 0; cb < W; cb += 64) {
      int cc = cb + lane;
      if (cc < W) {
        int zz = wrapm(z0 - s.z / 2 + cc, g.z);
        int yy = wrapm(y0 - s.y / 2, g.y);
        for (int r = 0; r < R; ++r) {
          double v = plane[(long)yy * g.z + zz];
          if (++yy == g.y) yy = 0;
          if constexpr (PREC == MSM_G2G_FP64) {
            reinterpret_cast<double *>(lds_raw)[r * Wp + cc] = v;
          } else {
            float f = (float)v;
            uint16_t hi = bf16_bits(f);
            uint16_t lo = bf16_bits(f - bf16_float(hi));
            reinterpret_cast<uint16_t *>(lds_raw)[r * Wp + cc] = hi;
            reinterpret_cast<uint16_t *>(lds_raw)[R * Wp + r * Wp + cc] = lo;
          }
        }
      }
    }
    __syncthreads();
    for (int ky = 0; ky < s.y; ++ky) {
      BRow<PREC, NCH> bcur = bnext;
      if (ky + 1 < s.y) bnext.load(L.bfrag, (long)(kx * s.y + ky + 1) * NCH, lane);
#pragma unroll
      for (int ch = 0; ch < NCH; ++ch) {
        if constexpr (PREC == MSM_G2G_FP64) {
          const double *slab = reinterpret_cast<const double *>(lds_raw);
          double av = slab[(row + ky) * Wp + ch * KC + kg];
          acc64 = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bcur.b[ch], acc64, 0, 0, 0);
        } else {
          const uint16_t *sh = reinterpret_cast<const uint16_t *>(lds_raw);
          int off = (row + ky) * Wp + ch * KC + kg * KPL;
          bf16x8 ahi = __builtin_bit_cast(bf16x8, *reinterpret_cast<const u32x4 *>(sh + off));
          bf16x8 alo = __builtin_bit_cast(bf16x8, *reinterpret_cast<const u32x4 *>(sh + R * Wp + off));
          bf16x8 bhi = __builtin_bit_cast(bf16x8, bcur.hi[ch]);
          bf16x8 blo = __builtin_bit_cast(bf16x8, bcur.lo[ch]);
          acc32 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(alo, bhi, acc32, 0, 0, 0);
          acc32 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ahi, blo, acc32, 0, 0, 0);
          acc32 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ahi, bhi, acc32, 0, 0, 0);
        }
      }
    }
  }
  // C layout, column lane%16 in both cases: the f32 16x16x32 result holds
  // rows 4*(lane/16) + i in register i; the f64 16x16x4 result holds rows
  // 4*i + lane/16 (measured on gfx950 with one-hot operands).
  double *out = L.part + (long)grp * g.size() + (long)x * g.y * g.z;
  int z = z0 + row;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    int y = y0 + ((PREC == MSM_G2G_FP64) ? 4 * i + kg : 4 * kg + i);
    if (y < g.y && z < g.z) {
      double v = (PREC == MSM_G2G_FP64) ? acc64[i] : (double)acc32[i];
      out[(long)y * g.z + z] = v;
    }
  }
}

// BF16 hi/lo level stencils.  A workgroup of four wavefronts takes one
// (plane x, 16-row band, segment of T z tiles, kx group): per kx the four
// stage the input slab together (q^l as bf16_split pairs), then wavefront w
// runs the rows ky = w, w + 4, ... of the stencil over all T tiles -- the
// B operands of a (kx, ky) are loaded once and serve the T tiles, whose K
// windows (32 NCH columns starting nz apart; nz a multiple of 4 keeps the
// 8-byte LDS reads of the A operand aligned) lie in the slab.  Columns >= nz
// of a tile's 16 are dropped (nz + s_z - 1 fits the chunks).  The four
// partial tiles are summed through LDS in a fixed order.
constexpr int kRowWaves = 1;
constexpr int kSlabRows = 8;  // slab rows per batch: kSlabRows x kSlabCols loads in flight
// slab columns per thread: the widest slab, (kG2GMaxTpw - 1) 16 + 4 x 32
// columns (tile width <= 16, at most four K chunks), fits 64 kRowWaves
// kSlabCols
constexpr int kSlabCols = 3;
static_assert((kG2GMaxTpw - 1) * 16 + 4 * 32 <= 64 * kRowWaves * kSlabCols, "slab columns");

template <int NCH, int T>
__global__ __launch_bounds__(64 * kRowWaves) void k_g2g_rows(G2GArgs args) {
  chain_issue_priority(args.urgent);
  extern __shared__ __align__(16) unsigned char lds_raw[];
  constexpr int KC = 32, KPL = 8, PAD = 8, BS = 64 * kRowWaves;
  long task = blockIdx.x;
  int li = 0;
  while (li + 1 < args.nlev && task >= args.lv[li + 1].task0) ++li;
  const G2GLevel &L = args.lv[li];
  long t = task - L.task0;
  const int grp = t % L.G;
  t /= L.G;
  const int zs = t % L.nzseg;
  t /= L.nzseg;
  const int yt = t % L.ty;
  const int x = (int)(t / L.ty);
  const GridDims g = L.g, s = L.s;
  const int nz = L.nz;
  const int per = (s.x + L.G - 1) / L.G;
  const int kx0 = grp * per, kx1 = min(s.x, kx0 + per);
  const int y0 = yt * 16, z0 = zs * T * nz;
  const int R = 16 + s.y - 1, W = (T - 1) * nz + NCH * KC, Wp = W + PAD;
  const int tid = threadIdx.x, lane = tid & 63;
  const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);  // uniform over the wavefront
  const int row = lane & 15, kg = lane >> 4;
  uint16_t *const shi = reinterpret_cast<uint16_t *>(lds_raw);
  uint16_t *const slo = shi + R * Wp;
  const int yb = y0 - s.y / 2, zb = z0 - s.z / 2;
  const float igy = 1.0f / g.y, igz = 1.0f / g.z;
  auto wrapf = [](int v, int m, float im) { return v - m * (int)floorf((v + 0.5f) * im); };
  // slab columns: lane l takes columns l + 64 j (the same for every kx and
  // row), their wrapped z index computed once; rows go kSlabRows at a time,
  // the row's wrapped y offset uniform over the wavefront
  const int nj = (W + BS - 1) / BS;  // <= kSlabCols (checked at launch)
  int zo[kSlabCols];
#pragma unroll
  for (int j = 0; j < kSlabCols; ++j) zo[j] = wrapf(zb + min(tid + BS * j, W - 1), g.z, igz);
  f32x4 acc[T];
#pragma unroll
  for (int tt = 0; tt < T; ++tt) acc[tt] = f32x4{0.f, 0.f, 0.f, 0.f};
  for (int kx = kx0; kx < kx1; ++kx) {
    const int xi = wrapm(x + kx - s.x / 2, g.x);
    const unsigned *plane = L.qb + (long)xi * g.y * g.z;
    __syncthreads();  // previous slab consumed
    for (int r0 = 0; r0 < R; r0 += kSlabRows) {
      unsigned v[kSlabRows][kSlabCols];
#pragma unroll
      for (int u = 0; u < kSlabRows; ++u) {
        const int yo = wrapf(yb + min(r0 + u, R - 1), g.y, igy) * g.z;  // past the end: valid, not stored
#pragma unroll
        for (int j = 0; j < kSlabCols; ++j)
          if (j < nj) v[u][j] = plane[yo + zo[j]];
      }
#pragma unroll
      for (int u = 0; u < kSlabRows; ++u) {
#pragma unroll
        for (int j = 0; j < kSlabCols; ++j) {
          const int cc = tid + BS * j;
          if (j < nj && r0 + u < R && cc < W) {
            const int at = (r0 + u) * Wp + cc;
            shi[at] = (uint16_t)(v[u][j] & 0xFFFFu);
            slo[at] = (uint16_t)(v[u][j] >> 16);
          }
        }
      }
    }
    __syncthreads();
    BRow<MSM_G2G_BF16X3, NCH> bnext;
    if (wv < s.y) bnext.load(L.bfrag, (long)(kx * s.y + wv) * NCH, lane);
    for (int ky = wv; ky < s.y; ky += kRowWaves) {
      const BRow<MSM_G2G_BF16X3, NCH> bcur = bnext;
      if (ky + kRowWaves < s.y) bnext.load(L.bfrag, (long)(kx * s.y + ky + kRowWaves) * NCH, lane);
#pragma unroll
      for (int ch = 0; ch < NCH; ++ch) {
        const bf16x8 bhi = __builtin_bit_cast(bf16x8, bcur.hi[ch]);
        const bf16x8 blo = __builtin_bit_cast(bf16x8, bcur.lo[ch]);
        bf16x8 ahi[T], alo[T];
#pragma unroll
        for (int tt = 0; tt < T; ++tt) {
          const int off = (row + ky) * Wp + tt * nz + ch * KC + kg * KPL;  // multiple of 4
          const uint2 h0 = *reinterpret_cast<const uint2 *>(shi + off);
          const uint2 h1 = *reinterpret_cast<const uint2 *>(shi + off + 4);
          const uint2 l0 = *reinterpret_cast<const uint2 *>(slo + off);
          const uint2 l1 = *reinterpret_cast<const uint2 *>(slo + off + 4);
          ahi[tt] = __builtin_bit_cast(bf16x8, u32x4{h0.x, h0.y, h1.x, h1.y});
          alo[tt] = __builtin_bit_cast(bf16x8, u32x4{l0.x, l0.y, l1.x, l1.y});
        }
#pragma unroll
        for (int tt = 0; tt < T; ++tt)
          acc[tt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(alo[tt], bhi, acc[tt], 0, 0, 0);
#pragma unroll
        for (int tt = 0; tt < T; ++tt)
          acc[tt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ahi[tt], blo, acc[tt], 0, 0, 0);
#pragma unroll
        for (int tt = 0; tt < T; ++tt)
          acc[tt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ahi[tt], bhi, acc[tt], 0, 0, 0);
      }
    }
  }
  // the four wavefronts' tiles, summed in a fixed order by wavefront 0
  float *red = reinterpret_cast<float *>(lds_raw);
  __syncthreads();
  if (wv > 0) {
#pragma unroll
    for (int tt = 0; tt < T; ++tt)
#pragma unroll
      for (int i = 0; i < 4; ++i) red[(((wv - 1) * T + tt) * 4 + i) * 64 + lane] = acc[tt][i];
  }
  __syncthreads();
  if (wv != 0) return;
  // f32 16x16 result: register i of lane l holds row 4 (l / 16) + i, column l % 16
  double *out = L.part + (long)grp * g.size() + (long)x * g.y * g.z;
  if (row < nz) {
#pragma unroll
    for (int tt = 0; tt < T; ++tt) {
      const int z = z0 + tt * nz + row;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        float v = acc[tt][i];
        for (int w = 1; w < kRowWaves; ++w) v += red[(((w - 1) * T + tt) * 4 + i) * 64 + lane];
        const int y = y0 + 4 * kg + i;
        if (y < g.y && z < g.z) out[(long)y * g.z + z] = (double)v;
      }
    }
  }
}

// e^l = sum_g part[g] + P e^{l+1} (prolongate, forcefield1.c:530-567), as
// separable passes over an LDS tile.  Fine index v receives J_{v-2m+H} e_m
// from every coarse m with |v - 2m| <= H (H = NU/2), periodic.  A workgroup
// owns an 8^3 fine brick: it loads the (H+4)^3 coarse values that reach it,
// applies the stencil along z, y and x in turn, adds the kx-group partials
// of the stencil pass in a fixed order and stores every point once.  With
// q1 != null (the finest level) it also writes the brick's sum of q1 e1,
// the grid-energy partial.
// Prolongation at fine index v >= 0 (forcefield1.c:530-567): the sum over
// the coarse m with |v - 2m| <= H of J_{v-2m+H} c(m), in ascending m.  With
// m = v/2 + d the weight of offset d is J_{H-2d} (v even) or J_{H+1-2d} (v
// odd): compile-time entries of J chosen by the parity, where an index into
// J that varies from lane to lane costs a memory load per tap.  Offsets
// whose weight is zero for the parity of v add exact zeros.
template <int NU, typename F>
__device__ __forceinline__ double prolong_taps(const double *J, int v, F c) {
  constexpr int H = NU / 2;
  const int j = v >> 1;
  const bool odd = v & 1;
  double acc = 0.0;
#pragma unroll
  for (int d = -(H / 2); d <= (H + 1) / 2; ++d) {
    const int i0 = H - 2 * d, i1 = H + 1 - 2 * d;
    const double w0 = (i0 >= 0 && i0 <= NU) ? J[i0] : 0.0;
    const double w1 = (i1 >= 0 && i1 <= NU) ? J[i1] : 0.0;
    acc += (odd ? w1 : w0) * c(j + d);
  }
  return acc;
}

constexpr int kPB = 8;  // fine brick edge of the prolongation

template <int NU>
__device__ __forceinline__ void prolong_brick(const DevParams &P, GridDims g, GridDims gc, int G,
                                              const double *__restrict__ part,
                                              const double *ec, double *__restrict__ el,
                                              const double *__restrict__ q1,
                                              double *__restrict__ epart) {
  constexpr int H = NU / 2, CW = H + 4, B = kPB;
  __shared__ double cs[CW * CW * CW];
  __shared__ double t1[CW * CW * B];  // [cx][cy][fz]
  __shared__ double t2[CW * B * B];   // [cx][fy][fz]
  __shared__ double wsum[4];
  const int tid = threadIdx.x;
  const int nbz = (g.z + B - 1) / B, nby = (g.y + B - 1) / B;
  int b = blockIdx.x;
  const int fz0 = (b % nbz) * B;
  b /= nbz;
  const int fy0 = (b % nby) * B, fx0 = (b / nby) * B;
  // lowest coarse index reaching the brick: ceil((v0 - H) / 2) = v0/2 - H/2
  const int cx0 = fx0 / 2 - H / 2, cy0 = fy0 / 2 - H / 2, cz0 = fz0 / 2 - H / 2;
  for (int k = tid; k < CW * CW * CW; k += 256) {
    const int a = k / (CW * CW), bb = (k / CW) % CW, c = k % CW;
    cs[k] = ec[((long)wrapm(cx0 + a, gc.x) * gc.y + wrapm(cy0 + bb, gc.y)) * gc.z +
               wrapm(cz0 + c, gc.z)];
  }
  __syncthreads();
  for (int k = tid; k < CW * CW * B; k += 256) {  // along z
    const int fz = k % B, cxy = k / B, v = fz0 + fz;
    t1[k] = prolong_taps<NU>(P.J, v, [&](int m) { return cs[cxy * CW + (m - cz0)]; });
  }
  __syncthreads();
  for (int k = tid; k < CW * B * B; k += 256) {  // along y
    const int fz = k % B, fy = (k / B) % B, cx = k / (B * B), v = fy0 + fy;
    t2[k] = prolong_taps<NU>(P.J, v, [&](int m) { return t1[(cx * CW + (m - cy0)) * B + fz]; });
  }
  __syncthreads();
  const long n = g.size();
  double es = 0.0;
  for (int k = tid; k < B * B * B; k += 256) {  // along x, plus the stencil partials
    const int fz = k % B, fy = (k / B) % B, fx = k / (B * B);
    const int x = fx0 + fx, y = fy0 + fy, z = fz0 + fz;
    if (x >= g.x || y >= g.y || z >= g.z) continue;
    double acc = prolong_taps<NU>(P.J, x, [&](int m) { return t2[((m - cx0) * B + fy) * B + fz]; });
    const long id = ((long)x * g.y + y) * g.z + z;
    for (int gi = 0; gi < G; ++gi) acc += part[(long)gi * n + id];
    el[id] = acc;
    if (q1) es += q1[id] * acc;
  }
  if (q1) {  // fixed-order block sum of q1 e1
    es = wave_incl_scan(es);  // lane 63: the wavefront total
    if ((tid & 63) == 63) wsum[tid >> 6] = es;
    __syncthreads();
    if (tid == 0) epart[blockIdx.x] = ((wsum[0] + wsum[1]) + wsum[2]) + wsum[3];
  }
}

template <int NU>
__global__ __launch_bounds__(256) void k_prolong_sum(DevParams P, GridDims g, GridDims gc, int G,
                                                     const double *__restrict__ part,
                                                     const double *__restrict__ ec,
                                                     double *__restrict__ el,
                                                     const double *__restrict__ q1,
                                                     double *__restrict__ epart) {
  chain_issue_priority();
  prolong_brick<NU>(P, g, gc, G, part, ec, el, q1, epart);
}

// The coarsest prolongation with the top level folded in: on a small top
// grid whose next level has few prolongation bricks, every workgroup
// computes the whole top-level potential itself -- the DFT convolution of
// the one-workgroup top-level kernel, in its own LDS -- and prolongates its
// brick from it, so that the chain runs one dependent launch fewer;
// workgroup 0 also stores the top-level potential.
template <int NU>
__global__ __launch_bounds__(256) void k_prolong_top(DevParams P, GridDims g, GridDims gc, int G,
                                                     const double *__restrict__ part,
                                                     const double *__restrict__ qL,
                                                     const double *__restrict__ khat,
                                                     const double *__restrict__ tw,
                                                     double *__restrict__ eL,
                                                     double *__restrict__ el, int asis) {
  chain_issue_priority();
  extern __shared__ double top_lds[];  // 4 n doubles: two complex buffers; the twiddles
  const int n = (int)gc.size();
  double *re = top_lds, *im = top_lds + n, *re2 = top_lds + 2 * n, *im2 = top_lds + 3 * n;
  double *const twl = top_lds + 4 * n;
  for (int k = threadIdx.x; k < 2 * (gc.x + gc.y + gc.z); k += 256) twl[k] = tw[k];
  for (int e = threadIdx.x; e < n; e += 256) {
    const int z = e % gc.z, y = (e / gc.z) % gc.y, x = e / (gc.y * gc.z);
    re[e] = top_input(gc, qL, x, y, z, asis, P.nu);
    im[e] = 0.0;
  }
  __syncthreads();
  dft_pass_lds<256>(gc, 2, -1.0, re, im, re2, im2, twl);
  dft_pass_lds<256>(gc, 1, -1.0, re2, im2, re, im, twl);
  dft_pass_lds<256>(gc, 0, -1.0, re, im, re2, im2, twl);
  for (int e = threadIdx.x; e < n; e += 256) re2[e] *= khat[e], im2[e] *= khat[e];
  __syncthreads();
  dft_pass_lds<256>(gc, 0, 1.0, re2, im2, re, im, twl);
  dft_pass_lds<256>(gc, 1, 1.0, re, im, re2, im2, twl);
  dft_pass_lds<256>(gc, 2, 1.0, re2, im2, re, im, twl);
  for (int e = threadIdx.x; e < n; e += 256) {
    const double v = re[e] / (double)n;
    re[e] = v;
    if (blockIdx.x == 0) eL[e] = v;
  }
  __syncthreads();
  prolong_brick<NU>(P, g, gc, G, part, re, el, nullptr, nullptr);
}

}  // namespace

// ---------------------------------------------------------------- host side
int g2g_chunk_k(int prec) { return prec == MSM_G2G_FP64 ? Cfg<MSM_G2G_FP64>::KC : Cfg<MSM_G2G_BF16X3>::KC; }

// BF16 tile width for nch chunks: the widest multiple of 4 up to 16 whose K
// window nz + s_z - 1 fits 32 nch columns (0 if none).
static int bf16_nz(int nch, int sz) {
  const int w = std::min(16, 32 * nch - sz + 1);
  return w >= 4 ? w & ~3 : 0;
}

void plan_g2g_level(int prec, GridDims g, GridDims s, int target_waves, G2GLevel &L) {
  int KC = g2g_chunk_k(prec);
  L.g = g;
  L.s = s;
  L.ty = (g.y + 15) / 16;
  if (prec == MSM_G2G_FP64) {
    L.nz = 16;
    L.nch = (16 + s.z - 1 + KC - 1) / KC;
  } else {  // fewest chunks per output column
    L.nch = 0;
    for (int c = 1; c <= g2g_max_nch(prec); ++c) {
      const int w = bf16_nz(c, s.z);
      if (w > 0 && (L.nch == 0 || (double)c / w < (double)L.nch / L.nz - 1e-12)) L.nch = c, L.nz = w;
    }
    if (L.nch == 0) L.nch = g2g_max_nch(prec) + 1, L.nz = 16;  // too wide: refused by the caller
  }
  L.tz = (g.z + L.nz - 1) / L.nz;
  L.tpw = 1;
  L.nzseg = L.tz;
  long tiles = (long)g.x * L.ty * L.tz;
  long G = (target_waves + tiles - 1) / tiles;
  if (G < 1) G = 1;
  if (G > s.x) G = s.x;
  int per = (s.x + (int)G - 1) / (int)G;
  L.G = (s.x + per - 1) / per;  // no empty groups
}

long g2g_level_tasks(const G2GLevel &L) { return (long)L.g.x * L.ty * L.nzseg * L.G; }

long plan_g2g_launch(int prec, G2GLevel *lv, int n, int target_waves) {
  int nch = 0;
  for (int k = 0; k < n; ++k) nch = std::max(nch, lv[k].nch);
  int tpw = 1;
  long big = -1;  // z tiles per wavefront: fewest wasted tiles on the largest level
  for (int k = 0; k < n; ++k) {
    G2GLevel &L = lv[k];
    L.nch = nch;
    if (prec == MSM_G2G_BF16X3 && nch <= g2g_max_nch(prec)) {
      L.nz = bf16_nz(nch, L.s.z);
      L.tz = (L.g.z + L.nz - 1) / L.nz;
      if (L.g.size() > big) {
        big = L.g.size();
        int best = -1;
        for (int t = 1; t <= kG2GMaxTpw; ++t) {
          const int waste = (L.tz + t - 1) / t * t - L.tz;
          if (best < 0 || waste <= best) best = waste, tpw = t;
        }
      }
    }
  }
  long tasks = 0;
  for (int k = 0; k < n; ++k) {
    G2GLevel &L = lv[k];
    L.tpw = tpw;
    L.nzseg = (L.tz + tpw - 1) / tpw;
    const long units = (long)L.g.x * L.ty * L.nzseg;
    long G = (target_waves + units - 1) / units;
    if (G < 1) G = 1;
    if (G > L.s.x) G = L.s.x;
    const int per = (L.s.x + (int)G - 1) / (int)G;
    L.G = (L.s.x + per - 1) / per;
    L.task0 = tasks;
    tasks += g2g_level_tasks(L);
  }
  return tasks;
}

size_t g2g_bfrag_elems(int prec, const G2GLevel &L) {
  size_t nfrag = (size_t)L.s.x * L.s.y * L.nch;
  return prec == MSM_G2G_FP64 ? nfrag * 64 : nfrag * 2 * 64 * 8;  // doubles / bf16 values
}

size_t g2g_bfrag_bytes(int prec, const G2GLevel &L) {
  return g2g_bfrag_elems(prec, L) * (prec == MSM_G2G_FP64 ? 8 : 2);
}

// B(k, j) = K[kx][ky][k - j] for 0 <= k - j < s_z, in per-lane operand order:
// lane holds column j = lane % 16 and K rows (lane / 16) * KPL + e of the chunk.
void g2g_bfrag_fill(int prec, const G2GLevel &L, const double *khat, void *host_out) {
  const int KC = g2g_chunk_k(prec), KPL = prec == MSM_G2G_FP64 ? 1 : 8;
  const GridDims s = L.s;
  auto kval = [&](int kx, int ky, int k, int j) {
    int kz = k - j;
    return (kz >= 0 && kz < s.z) ? khat[((long)kx * s.y + ky) * s.z + kz] : 0.0;
  };
  for (int kx = 0; kx < s.x; ++kx)
    for (int ky = 0; ky < s.y; ++ky)
      for (int ch = 0; ch < L.nch; ++ch) {
        long f = ((long)kx * s.y + ky) * L.nch + ch;
        for (int lane = 0; lane < 64; ++lane) {
          int j = lane % 16;
          for (int e = 0; e < KPL; ++e) {
            int k = ch * KC + (lane / 16) * KPL + e;
            double v = kval(kx, ky, k, j);
            if (prec == MSM_G2G_FP64) {
              static_cast<double *>(host_out)[f * 64 + lane] = v;
            } else {
              uint16_t *o = static_cast<uint16_t *>(host_out);
              float fv = (float)v;
              uint16_t hi = bf16_bits(fv), lo = bf16_bits(fv - bf16_float(hi));
              o[(f * 128 + lane) * 8 + e] = hi;
              o[(f * 128 + 64 + lane) * 8 + e] = lo;
            }
          }
        }
      }
}

namespace {
// Device twin of g2g_bfrag_fill: one thread per (fragment, lane).
template <int PREC>
__global__ __launch_bounds__(256) void k_g2g_bfrag_fill(G2GLevel L, const double *__restrict__ khat,
                                                        void *__restrict__ out) {
  constexpr int KC = Cfg<PREC>::KC, KPL = Cfg<PREC>::KPL;
  const long t = (long)blockIdx.x * 256 + threadIdx.x;
  const long nfrag = (long)L.s.x * L.s.y * L.nch;
  if (t >= nfrag * 64) return;
  const long f = t / 64;
  const int lane = (int)(t % 64), j = lane % 16;
  const int ch = (int)(f % L.nch), ky = (int)((f / L.nch) % L.s.y), kx = (int)(f / ((long)L.nch * L.s.y));
  for (int e = 0; e < KPL; ++e) {
    const int k = ch * KC + (lane / 16) * KPL + e, kz = k - j;
    const double v = (kz >= 0 && kz < L.s.z) ? khat[((long)kx * L.s.y + ky) * L.s.z + kz] : 0.0;
    if constexpr (PREC == MSM_G2G_FP64) {
      static_cast<double *>(out)[f * 64 + lane] = v;
    } else {
      uint16_t *o = static_cast<uint16_t *>(out);
      const float fv = (float)v;
      const uint16_t hi = bf16_bits(fv), lo = bf16_bits(fv - bf16_float(hi));
      o[(f * 128 + lane) * 8 + e] = hi;
      o[(f * 128 + 64 + lane) * 8 + e] = lo;
    }
  }
}
}  // namespace

void launch_g2g_bfrag_fill(int prec, const G2GLevel &L, const double *khat, void *out,
                           hipStream_t s) {
  const long n = (long)L.s.x * L.s.y * L.nch * 64;
  const unsigned nb = (unsigned)((n + 255) / 256);
  if (prec == MSM_G2G_FP64)
    k_g2g_bfrag_fill<MSM_G2G_FP64><<<nb, 256, 0, s>>>(L, khat, out);
  else
    k_g2g_bfrag_fill<MSM_G2G_BF16X3><<<nb, 256, 0, s>>>(L, khat, out);
}

size_t g2g_lds_bytes(int prec, const G2GLevel &L) {
  int R = 16 + L.s.y - 1;
  if (prec == MSM_G2G_FP64) return sizeof(double) * R * (L.nch * 4 + Cfg<MSM_G2G_FP64>::PAD);
  const size_t slab = 2 * sizeof(uint16_t) * R *
                      ((L.tpw - 1) * L.nz + L.nch * 32 + Cfg<MSM_G2G_BF16X3>::PAD);
  const size_t red = sizeof(float) * (kRowWaves - 1) * L.tpw * 4 * 64;  // tile reduction
  return std::max(slab, red);
}

template <int PREC, int NCH>
static void g2g_launch(const G2GArgs &args, size_t lds, hipStream_t s, hipEvent_t done) {
  launch_done(k_g2g_tiles<PREC, NCH>, (unsigned)args.ntasks, 64, lds, s, done, args);
}

int g2g_max_nch(int prec) { return prec == MSM_G2G_FP64 ? 16 : 4; }

void launch_g2g_tiles(int prec, G2GArgs args, size_t lds, hipStream_t s, hipEvent_t done) {
  if (args.ntasks <= 0) return;
  // all levels of one launch share the chunk count (the plan equalises it)
  int nch = args.lv[0].nch;
  if (prec == MSM_G2G_FP64) {
    switch (nch) {
#define C(N) case N: g2g_launch<MSM_G2G_FP64, N>(args, lds, s, done); break;
      C(1) C(2) C(3) C(4) C(5) C(6) C(7) C(8) C(9) C(10) C(11) C(12) C(13) C(14) C(15) C(16)
#undef C
    }
  } else {
    const int T = args.lv[0].tpw;
#define R(N, TT) \
  if (nch == N && T == TT)                                                                      \
    launch_done(k_g2g_rows<N, TT>, (unsigned)args.ntasks, 64 * kRowWaves, lds, s, done, args);
#define RT(N) R(N, 1) R(N, 2) R(N, 3) R(N, 4) R(N, 5)
    RT(1) RT(2) RT(3) RT(4)
#undef RT
#undef R
  }
}

int prolong_blocks(GridDims g) {
  return ((g.x + kPB - 1) / kPB) * ((g.y + kPB - 1) / kPB) * ((g.z + kPB - 1) / kPB);
}

bool prolong_top_fits(GridDims g, GridDims gc) {
  return gc.size() <= kTopFuse && prolong_blocks(g) <= kTopFuseBricks;
}

void launch_prolong_top(const DevParams &P, GridDims g, GridDims gc, int G, const double *part,
                        const double *qL, const double *khat, const double *tw, double *eL,
                        double *el, int asis, hipStream_t s) {
  const unsigned nb = (unsigned)prolong_blocks(g);
  const size_t lds = (4 * gc.size() + 2 * (gc.x + gc.y + gc.z)) * sizeof(double);
#define CALL(NU) \
  k_prolong_top<NU><<<nb, 256, lds, s>>>(P, g, gc, G, part, qL, khat, tw, eL, el, asis)
  switch (P.nu) {
    case 4: CALL(4); break;
    case 6: CALL(6); break;
    case 8: CALL(8); break;
    default: CALL(10); break;
  }
#undef CALL
}

void launch_prolong_sum(const DevParams &P, GridDims g, GridDims gc, int G, const double *part,
                        const double *ec, double *el, const double *q1, double *epart,
                        hipStream_t s) {
  unsigned nb = (unsigned)prolong_blocks(g);
  switch (P.nu) {
    case 4: k_prolong_sum<4><<<nb, 256, 0, s>>>(P, g, gc, G, part, ec, el, q1, epart); break;
    case 6: k_prolong_sum<6><<<nb, 256, 0, s>>>(P, g, gc, G, part, ec, el, q1, epart); break;
    case 8: k_prolong_sum<8><<<nb, 256, 0, s>>>(P, g, gc, G, part, ec, el, q1, epart); break;
    default: k_prolong_sum<10><<<nb, 256, 0, s>>>(P, g, gc, G, part, ec, el, q1, epart); break;
  }
}

}  // namespace msm
