// msm_kernels.hip -- HIP kernels of one MSM energy/force evaluation on
// gfx950 (MI355X).
//
// The chain restates FF_energy of hkayabilisim/msmpb
// (code/forcefield1.c:101-235) for a GPU: atoms are binned into cells and
// sorted every call (instead of walking a stored neighbour list), the pair
// sum runs over a full neighbour-cell shell (no atomics on forces), the
// level grids are periodic and unpadded, and prolongation is fused into the
// level stencil pass.  Wave size is 64.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>

#include "msm_kernels.h"

namespace msm {
namespace {

__device__ __forceinline__ int wrapi(int v, int m) {
  int r = v % m;
  return r < 0 ? r + m : r;
}

__device__ __forceinline__ int floordiv(int v, int m) {
  return v >= 0 ? v / m : -((-v + m - 1) / m);
}

// Sum over a 64-wide wavefront, then over the waves of a block through LDS.
// Returns the block total in thread 0.
template <int BS>
__device__ __forceinline__ double block_sum(double v) {
  __shared__ double red[BS / 64];
  v = wave_incl_scan(v);  // lane 63: the wavefront total
  int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 63) red[w] = v;
  __syncthreads();
  double t = 0;
  if (threadIdx.x == 0)
    for (int k = 0; k < BS / 64; ++k) t += red[k];
  return t;
}

// Weights of the order-NU B-spline at the NU grid points around t in [0,1):
// w[i] = Q_NU(i + t) from the stored first NU/2 pieces, the other half by
// symmetry (forcefield1.c:609-637).  dw = d/dt.
template <int NU, bool DERIV>
__device__ __forceinline__ void bspline(const double *Q, double t, double *w, double *dw) {
#pragma unroll
  for (int i = 0; i < NU; ++i) {
    int p = i < NU / 2 ? i : NU - 1 - i;
    double u = i < NU / 2 ? t : 1.0 - t;
    const double *c = Q + p * NU;
    double v = c[NU - 1], d = 0.0;
#pragma unroll
    for (int k = NU - 2; k >= 0; --k) {
      if (DERIV) d = v + u * d;
      v = c[k] + u * v;
    }
    w[i] = v;
    if (DERIV) dw[i] = i < NU / 2 ? d : -d;
  }
}

// One weight Q_NU(i + t) (and its t-derivative) for a uniform index i.
template <int NU, bool DERIV>
__device__ __forceinline__ double bspline_one(const double *Q, int i, double t, double *dw) {
  int p = i < NU / 2 ? i : NU - 1 - i;
  double u = i < NU / 2 ? t : 1.0 - t;
  const double *c = Q + p * NU;
  double v = c[NU - 1], d = 0.0;
#pragma unroll
  for (int k = NU - 2; k >= 0; --k) {
    if (DERIV) d = v + u * d;
    v = c[k] + u * v;
  }
  if (DERIV) *dw = i < NU / 2 ? d : -d;
  return v;
}

// ---------------------------------------------------------------- utilities
__global__ void k_fill_int(int *p, long n, int v) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}
__global__ void k_fill_dbl(double *p, long n, double v) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// ---------------------------------------------------------------- binning
constexpr int kBinBS = 256;

// Fractional coordinates A^-1 r: the diagonal for orthorhombic boxes, the
// full inverse for triclinic ones (TRI: the box shape, fixed per kernel
// instance so that the orthorhombic kernels carry no triclinic code).
template <bool TRI>
__device__ __forceinline__ void to_frac(const DevParams &P, const double r[3], double s[3]) {
  if constexpr (TRI) {
#pragma unroll
    for (int a = 0; a < 3; ++a) s[a] = P.Ai[3 * a] * r[0] + P.Ai[3 * a + 1] * r[1] + P.Ai[3 * a + 2] * r[2];
  } else {
#pragma unroll
    for (int a = 0; a < 3; ++a) s[a] = r[a] * P.ibox[a];
  }
}

// Field from the fractional gradient: E = -A^-T (M * grad_t).
template <bool TRI>
__device__ __forceinline__ void frac_field(const DevParams &P, GridDims g, double ex, double ey,
                                           double ez, double &Ex, double &Ey, double &Ez) {
  const double fx = ex * g.x, fy = ey * g.y, fz = ez * g.z;
  if constexpr (TRI) {
    Ex = -(P.Ai[0] * fx + P.Ai[3] * fy + P.Ai[6] * fz);
    Ey = -(P.Ai[1] * fx + P.Ai[4] * fy + P.Ai[7] * fz);
    Ez = -(P.Ai[2] * fx + P.Ai[5] * fy + P.Ai[8] * fz);
  } else {
    Ex = -(fx * P.ibox[0]), Ey = -(fy * P.ibox[1]), Ez = -(fz * P.ibox[2]);
  }
}

// Lattice vector A k (Cartesian shift of the periodic image k).
template <bool TRI>
__device__ __forceinline__ void lattice(const DevParams &P, int kx, int ky, int kz, double &x,
                                        double &y, double &z) {
  if constexpr (TRI) {
    x = P.A[0] * kx + P.A[1] * ky + P.A[2] * kz;
    y = P.A[3] * kx + P.A[4] * ky + P.A[5] * kz;
    z = P.A[6] * kx + P.A[7] * ky + P.A[8] * kz;
  } else {
    x = kx * P.box[0], y = ky * P.box[1], z = kz * P.box[2];
  }
}

// The same decided at run time (binning and pair count: not hot).
__device__ __forceinline__ void to_frac_rt(const DevParams &P, const double r[3], double s[3]) {
  if (P.tri) to_frac<true>(P, r, s);
  else to_frac<false>(P, r, s);
}
__device__ __forceinline__ void lattice_rt(const DevParams &P, int kx, int ky, int kz, double &x,
                                           double &y, double &z) {
  if (P.tri) lattice<true>(P, kx, ky, kz, x, y, z);
  else lattice<false>(P, kx, ky, kz, x, y, z);
}

// Cell of every atom, its wrapped position, and the cell's members: the
// counting atomic gives the atom's arrival rank r in its cell, and the atom
// is recorded in the cell's bucket row (cap entries) at r, or, past the row,
// on the overflow list (ovfn[0] entries).  Also the per-workgroup charge sums.
// Atom i's cell and wrapped position with its charge (w); false for a
// non-finite position or charge (then cell 0 and a zero position: no index
// is computed from it).  The binning and the placement pass both evaluate
// it, with the same result.
__device__ __forceinline__ bool bin_atom(const DevParams &P, const double *__restrict__ pos,
                                         const double *__restrict__ q, int i, int &cell,
                                         double4 &w) {
  const double r[3] = {pos[3 * (long)i], pos[3 * (long)i + 1], pos[3 * (long)i + 2]};
  const double qi = q[i];
  int c[3];
  double rw[3], sf[3], fl[3];
  bool ok = isfinite(qi);
  to_frac_rt(P, r, sf);
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    double s = sf[a];
    double f = floor(s);
    double sw = s - f;
    fl[a] = f;
    ok = ok && isfinite(s);
    // s - floor(s) may round up to 1 for tiny negative s; keep in range
    int ci = (int)(sw * P.nc[a]);
    if (!(sw >= 0.0) || ci < 0) ci = 0;
    if (ci >= P.nc[a]) ci = P.nc[a] - 1;
    c[a] = ci;
  }
  if (P.tri) {  // r - A floor(A^-1 r)
    rw[0] = r[0] - (P.A[0] * fl[0] + P.A[1] * fl[1] + P.A[2] * fl[2]);
    rw[1] = r[1] - (P.A[3] * fl[0] + P.A[4] * fl[1] + P.A[5] * fl[2]);
    rw[2] = r[2] - (P.A[6] * fl[0] + P.A[7] * fl[1] + P.A[8] * fl[2]);
  } else {
#pragma unroll
    for (int a = 0; a < 3; ++a) rw[a] = r[a] - fl[a] * P.box[a];
  }
#pragma unroll
  for (int a = 0; a < 3; ++a)
    if (!isfinite(rw[a])) rw[a] = 0.0;
  cell = (c[0] * P.nc[1] + c[1]) * P.nc[2] + c[2];
  w = make_double4(rw[0], rw[1], rw[2], qi);
  return ok;
}

__global__ __launch_bounds__(kBinBS) void k_bin(DevParams P, const double *__restrict__ pos,
                                                const double *__restrict__ q,
                                                int *__restrict__ cell_of, int *__restrict__ bucket,
                                                int cap, int *__restrict__ ovfn,
                                                int *__restrict__ ovf,
                                                int *__restrict__ cnt,
                                                double *__restrict__ qpart, int *__restrict__ flags) {
  int i = blockIdx.x * kBinBS + threadIdx.x;
  double qi = 0.0;
  if (i < P.N) {
    int cell;
    double4 w;
    if (!bin_atom(P, pos, q, i, cell, w)) atomicAdd(flags, 1);
    qi = w.w;
    cell_of[i] = cell;
    const int rk = atomicAdd(cnt + cell, 1);  // arrival order; k_place fixes the order
    if (rk < cap) bucket[(long)cell * cap + rk] = i;
    else ovf[atomicAdd(ovfn, 1)] = i;
  }
  double s1 = block_sum<kBinBS>(qi);
  __syncthreads();
  double s2 = block_sum<kBinBS>(qi * qi);
  if (threadIdx.x == 0) {
    qpart[2 * blockIdx.x] = s1;
    qpart[2 * blockIdx.x + 1] = s2;
  }
}

constexpr int kScanBS = 1024;

// Exclusive scan of the cell counts (start[0..n]) by one workgroup, in
// passes of kScanTiles tiles of 4 x kScanBS counts: thread t loads counts
// 4t .. 4t + 3 of every tile of the pass (coalesced, all loads in flight
// together; cnt and start are padded to whole int4 past n, the padding of
// cnt zero), the per-thread sums are scanned per tile across each
// wavefront, every wavefront scans the wavefront totals of each tile
// itself (one barrier per pass), and the prefixes are written back four at
// a time.  The counts are cleared as they are read, ready for the next
// evaluation's binning, and so is the overflow count (moved to ovfn[1] for
// k_place).  Also sums the per-block charge partials of the binning pass,
// in a fixed order.
constexpr int kScanTiles = 8;
__global__ __launch_bounds__(kScanBS) void k_scan(int n, int *__restrict__ cnt,
                                                  int *__restrict__ start,
                                                  const double *__restrict__ qpart, int nq,
                                                  double *__restrict__ scal,
                                                  int *__restrict__ ovfn) {
  constexpr int NW = kScanBS / 64, TILE = 4 * kScanBS;
  __shared__ int wtot[kScanTiles][NW];
  __shared__ double wq[2][NW];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  double q1 = 0, q2 = 0;
  for (int k = t; k < nq; k += kScanBS) q1 += qpart[2 * k], q2 += qpart[2 * k + 1];
  int carry = 0;
  for (int base = 0; base < n; base += kScanTiles * TILE) {
    int4 v[kScanTiles];
    int sum[kScanTiles], incl[kScanTiles];
#pragma unroll
    for (int j = 0; j < kScanTiles; ++j) {
      const int i = base + j * TILE + 4 * t;
      v[j] = i < n ? *reinterpret_cast<const int4 *>(cnt + i) : make_int4(0, 0, 0, 0);
      sum[j] = v[j].x + v[j].y + v[j].z + v[j].w;
    }
#pragma unroll
    for (int j = 0; j < kScanTiles; ++j) {
      const int i = base + j * TILE + 4 * t;
      if (i < n) *reinterpret_cast<int4 *>(cnt + i) = make_int4(0, 0, 0, 0);
      incl[j] = wave_incl_scan(sum[j]);
    }
    if (lane == 63)
#pragma unroll
      for (int j = 0; j < kScanTiles; ++j) wtot[j][w] = incl[j];
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kScanTiles; ++j) {
      // this wavefront's offset in the tile and the tile's total
      const int wi = wave_incl_scan(lane < NW ? wtot[j][lane] : 0);
      const int woff = w > 0 ? __builtin_amdgcn_readlane(wi, w - 1) : 0;
      const int tile = __builtin_amdgcn_readlane(wi, NW - 1);
      const int p = carry + woff + incl[j] - sum[j];
      const int i = base + j * TILE + 4 * t;
      if (i < n)
        *reinterpret_cast<int4 *>(start + i) =
            make_int4(p, p + v[j].x, p + v[j].x + v[j].y, p + v[j].x + v[j].y + v[j].z);
      carry += tile;  // the same value in every thread
    }
    __syncthreads();  // wtot is rewritten by the next pass
  }
  // charge sums for the constant energy terms: wavefront scans, then the
  // wavefront totals in order
  q1 = wave_incl_scan(q1);
  q2 = wave_incl_scan(q2);
  if (lane == 63) wq[0][w] = q1, wq[1][w] = q2;
  __syncthreads();
  if (t == 0) {
    start[n] = carry;  // after the barriers: the padded int4 stores above may cover it
    ovfn[1] = ovfn[0];
    ovfn[0] = 0;
    double a = 0, c2 = 0;
    for (int k = 0; k < NW; ++k) a += wq[0][k], c2 += wq[1][k];
    scal[0] = a, scal[1] = c2;
  }
}

// Placement of a sorted atom on the level-1 grid for the spreading
// (anterpolate, forcefield1.c:350-368): fractional offsets t in [0, 1) and
// the index of the lowest support point m - NU/2 + 1 per axis, reduced
// modulo the grid and packed 10 bits per axis, with the charge next to t.
template <bool TRI>
__device__ __forceinline__ void place_atom(const DevParams &P, GridDims g, const double4 &at,
                                           unsigned &packed, double4 &tq) {
  const double r[3] = {at.x, at.y, at.z};
  double sf[3], t[3];
  to_frac<TRI>(P, r, sf);
  const int dims[3] = {g.x, g.y, g.z};
  packed = 0;
#pragma unroll
  for (int ax = 0; ax < 3; ++ax) {
    double sv = sf[ax];
    sv = sv - floor(sv);
    const double tt = dims[ax] * sv, m = floor(tt);
    t[ax] = tt - m;
    int v = ((int)m - P.nu / 2 + 1) % dims[ax];
    v += v < 0 ? dims[ax] : 0;
    packed |= (unsigned)v << (10 * ax);
  }
  tq = make_double4(t[0], t[1], t[2], at.w);
}

// The cell-sorted order, one thread per atom: atom i's place in its cell c
// is the number of c's atoms with a smaller original index (so the order,
// and every sum built on it, is deterministic), counted over c's bucket row
// and, for a cell that overflowed it, the overflow list (a scan of the whole
// list per such atom: rows hold several times the mean occupancy, so
// overflows are rare).  Writes the sorted position and charge, the original
// index and, with sbase != null, the spreading placement (place_atom): the
// spreading kernel then follows the binning on the same stream directly,
// and is dispatched before the pair kernel on its own stream can take the
// CUs.
constexpr int kPlaceAhead = 8;  // bucket entries in flight per thread
template <bool TRI>
__global__ __launch_bounds__(256) void k_place(DevParams P, const double *__restrict__ pos,
                                               const double *__restrict__ q,
                                               const int *__restrict__ cell_of,
                                               const int *__restrict__ start,
                                               const int *__restrict__ bucket, int cap,
                                               const int *__restrict__ ovfn,
                                               const int *__restrict__ ovf,
                                               double4 *__restrict__ spos, int *__restrict__ sidx,
                                               GridDims g1, unsigned *__restrict__ sbase,
                                               double4 *__restrict__ st) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= P.N) return;
  int c;
  double4 v;
  bin_atom(P, pos, q, i, c, v);  // as the binning found it
  const int *row = bucket + (long)c * cap;
  const int b = start[c], n = start[c + 1] - b;
  int rank = 0;
  for (int k0 = 0; k0 < min(n, cap); k0 += kPlaceAhead) {
    int e[kPlaceAhead];
#pragma unroll
    for (int u = 0; u < kPlaceAhead; ++u) e[u] = k0 + u < min(n, cap) ? row[k0 + u] : i;
#pragma unroll
    for (int u = 0; u < kPlaceAhead; ++u) rank += e[u] < i;
  }
  if (n > cap) {
    const int nv = ovfn[1];
    for (int k = 0; k < nv; ++k) {
      const int j = ovf[k];
      rank += j < i && cell_of[j] == c;
    }
  }
  spos[b + rank] = v;
  sidx[b + rank] = i;
  if (sbase) {
    unsigned pk;
    double4 tq;
    place_atom<TRI>(P, g1, v, pk, tq);
    sbase[b + rank] = pk;
    st[b + rank] = tq;
  }
}

// ---------------------------------------------------------------- pairs
// partcl2partcl (forcefield1.c:240-321) as a full-shell gather: each atom
// sums q_j g(r) over every neighbour j != i and every periodic image inside
// the cutoff, with g(r) = 1/r - tau(r^2/a0^2 - 1)/a0.
//
// One wavefront per home cell.  The home atoms are taken up to kPassAtoms at
// a time (a pass); their coordinates are wavefront-uniform.  The neighbour
// atoms that can come within the cutoff of the pass's bounding box are
// staged in LDS (coordinates relative to the home-cell origin), and lane l
// evaluates staged atoms l, l + 64, ... against every home atom of the pass,
// so no lane idles whatever the cell population.  Neighbours are found by
// cell columns: for each (dx, dy) offset the cells dz in [dzlo, dzhi] are
// contiguous in the sorted order, so a column is one index range (two where
// it wraps around the box); the z range is first trimmed to the cells the
// bounding box can reach.  A round's candidates are numbered across its
// ranges; a byte table gives the range of each candidate number (a walk over
// the range offsets where a round has more candidates than the table holds).
// Each lane's partial sums are combined across the
// wavefront in a fixed butterfly order: no atomics, so the result is
// deterministic.
//
// Periodic images of an atom itself (boxes narrower than the cutoff) are
// evaluated like any other neighbour, and their sum q_i * self_img, which
// depends on the box only, is taken off phi_i; their forces cancel pairwise.
constexpr int kPairBS = 256;    // atoms per pair-partial block (energy partial buffer sizing)
constexpr int kPassAtoms = 8;   // home atoms per pass
constexpr int kSegMax = 128;    // atom ranges per round: 64 columns x <= 2 pieces

template <bool MIXED>
struct PairPrec;
template <>
struct PairPrec<true> {
  using T = float4;
  using R = float;
  static constexpr int CAP = 384;  // staged neighbours per interaction run (6 KB of LDS)
  static constexpr int TAB = 2048; // range table: every round at water96k (1024: 40 %, 1.0 % slower)
};
template <>
struct PairPrec<false> {
  using T = double4;
  using R = double;
  static constexpr int CAP = 384;  // 12 KB: longer runs beat occupancy here (256: +1 %, 512: +14 %)
  static constexpr int TAB = 1024; // 2048 would leave 8 instead of 9 wavefronts per CU
};

__device__ __forceinline__ float wave_uniform(float v) {
  return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v)));
}
__device__ __forceinline__ double wave_uniform(double v) {
  const long long b = __double_as_longlong(v);
  const unsigned lo = (unsigned)__builtin_amdgcn_readfirstlane((int)(b & 0xffffffffll));
  const unsigned hi = (unsigned)__builtin_amdgcn_readfirstlane((int)(b >> 32));
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// 1/sqrt(r2) in FP64: hardware estimate refined by one Newton step (a
// second step left the water96k energy unchanged in all 17 printed digits;
// the FP64 parity tests hold every energy, force and potential to 1e-10).
__device__ __forceinline__ double rsqrt_f64(double r2) {
  double y = __builtin_amdgcn_rsq(r2);
  const double h = 0.5 * r2;
  y = y * fma(-h * y, y, 1.5);
  return y;
}

// One neighbour (aj: position relative to the home-cell origin, charge) and
// one home atom (ix, iy, iz): sa += q_j g(r), (sx, sy, sz) += q_j g'(r)/r d
// with d = r_j - r_i; the pre-scaled coefficients fold 1/a0 into tau and
// 2/a0^3 into tau'.
//
// FP32: r^2 is clamped to at most a0^2 (one med3) instead of masking the
// neighbours beyond the cutoff: g and g' both vanish at r = a0 (the smoothed
// kernel joins 1/r with a continuous first derivative), so such a neighbour
// adds only the rounding residue of g(a0) and g'(a0), about 1e-7 of 1/a0.
// The self pair (d = 0 exactly) is masked through its charge.
//
// NU = 0 selects the Ewald term of the PME comparator (pme1.c:160-170):
// g(r) = erfc(beta r)/r, g'(r)/r = -(g(r) + 2 beta/sqrt(pi) exp(-beta^2 r^2))/r^2.
// It does not vanish at the cutoff, so neighbours beyond it are masked.
template <int NU>
__device__ __forceinline__ void pair_term(const DevParams &P, float4 aj, float ix, float iy,
                                          float iz, float &sa, float &sx, float &sy, float &sz) {
  const float dx = aj.x - ix, dy = aj.y - iy, dz = aj.z - iz;
  const float r2 = dx * dx + dy * dy + dz * dz;
  if constexpr (NU == 0) {
    const bool ok = r2 < P.fa02 && r2 > 0.f;
    const float r2s = ok ? r2 : P.fa02;
    const float qj = ok ? aj.w : 0.f;
    const float rinv = __builtin_amdgcn_rsqf(r2s);
    const float v = erfcf(P.fbeta * (r2s * rinv)) * rinv;
    sa += qj * v;
    const float qc = -qj * (v + P.fewc * __expf(-P.fbeta2 * r2s)) * (rinv * rinv);
    sx += qc * dx, sy += qc * dy, sz += qc * dz;
  } else {
    const float r2s = __builtin_amdgcn_fmed3f(r2, 1e-20f, P.fa02);
    const float qj = r2 > 0.f ? aj.w : 0.f;
    const float sv = r2s * P.fia02 - 1.f;
    float ts = P.ftau[NU - 1], t1 = P.fdtau[NU - 2];
#pragma unroll
    for (int k = NU - 2; k >= 0; --k) ts = P.ftau[k] + sv * ts;
#pragma unroll
    for (int k = NU - 3; k >= 0; --k) t1 = P.fdtau[k] + sv * t1;
    const float rinv = __builtin_amdgcn_rsqf(r2s);
    sa += qj * (rinv - ts);
    const float qc = qj * (-rinv * rinv * rinv - t1);
    sx += qc * dx, sy += qc * dy, sz += qc * dz;
  }
}

// Cell of workgroup b.  Consecutive workgroups are dispatched to the eight
// XCDs in turn; this gives XCD k a contiguous range of cells (an x-slab), so
// the neighbour atoms its waves stage stay in its own L2.  A bijection of
// [0, n).
__device__ __forceinline__ int xcd_slab(int b, int n) {
  const int k = b & 7, j = b >> 3, q = n >> 3, r = n & 7;
  return k * q + min(k, r) + j;
}

template <int NU>
__device__ __forceinline__ void pair_term(const DevParams &P, double4 aj, double ix, double iy,
                                          double iz, double &sa, double &sx, double &sy,
                                          double &sz) {
  const double dx = aj.x - ix, dy = aj.y - iy, dz = aj.z - iz;
  const double r2 = dx * dx + dy * dy + dz * dz;
  const bool ok = r2 < P.a02 && r2 > 0.0;
  const double r2s = ok ? r2 : P.a02;
  const double qj = ok ? aj.w : 0.0;
  const double rinv = rsqrt_f64(r2s);
  if constexpr (NU == 0) {  // Ewald (PME)
    const double v = erfc(P.beta * (r2s * rinv)) * rinv;
    sa += qj * v;
    const double qc = -qj * (v + P.ewc * exp(-P.beta2 * r2s)) * (rinv * rinv);
    sx += qc * dx, sy += qc * dy, sz += qc * dz;
  } else {
    const double sv = r2s * P.inv_a02 - 1.0;
    double ts = P.tau_s[NU - 1], t1 = P.dtau_s[NU - 2];
#pragma unroll
    for (int k = NU - 2; k >= 0; --k) ts = P.tau_s[k] + sv * ts;
#pragma unroll
    for (int k = NU - 3; k >= 0; --k) t1 = P.dtau_s[k] + sv * t1;
    sa += qj * (rinv - ts);
    const double qc = qj * (-rinv * rinv * rinv - t1);
    sx += qc * dx, sy += qc * dy, sz += qc * dz;
  }
}

// The staged atoms [0, npad) against the pass's G home atoms: lane l takes
// staged atoms l, l + 64, ... and adds into s[g] = {phi, fx, fy, fz} (FP32
// for the mixed kernel: at most CAP / 64 terms per run).
template <bool MIXED, int NU, int G>
__device__ __forceinline__ void pair_block(const DevParams &P,
                                           const typename PairPrec<MIXED>::T *jbuf, int npad,
                                           const typename PairPrec<MIXED>::R *ix,
                                           const typename PairPrec<MIXED>::R *iy,
                                           const typename PairPrec<MIXED>::R *iz,
                                           typename PairPrec<MIXED>::R (*s)[4]) {
  for (int e = threadIdx.x; e < npad; e += 64) {
    const auto aj = jbuf[e];
#pragma unroll
    for (int g = 0; g < G; ++g)
      pair_term<NU>(P, aj, ix[g], iy[g], iz[g], s[g][0], s[g][1], s[g][2], s[g][3]);
  }
}

template <bool MIXED, int NU>
__device__ __forceinline__ void pair_run(const DevParams &P, const typename PairPrec<MIXED>::T *jbuf,
                                         int npad, int ni, const typename PairPrec<MIXED>::R *ix,
                                         const typename PairPrec<MIXED>::R *iy,
                                         const typename PairPrec<MIXED>::R *iz,
                                         typename PairPrec<MIXED>::R (*s)[4]) {
  switch (ni) {
    case 8: pair_block<MIXED, NU, 8>(P, jbuf, npad, ix, iy, iz, s); break;
    case 7: pair_block<MIXED, NU, 7>(P, jbuf, npad, ix, iy, iz, s); break;
    case 6: pair_block<MIXED, NU, 6>(P, jbuf, npad, ix, iy, iz, s); break;
    case 5: pair_block<MIXED, NU, 5>(P, jbuf, npad, ix, iy, iz, s); break;
    case 4: pair_block<MIXED, NU, 4>(P, jbuf, npad, ix, iy, iz, s); break;
    case 3: pair_block<MIXED, NU, 3>(P, jbuf, npad, ix, iy, iz, s); break;
    case 2: pair_block<MIXED, NU, 2>(P, jbuf, npad, ix, iy, iz, s); break;
    default: pair_block<MIXED, NU, 1>(P, jbuf, npad, ix, iy, iz, s); break;
  }
}

// Halving exchanges.  Each sums two values a, b over a pair of lanes: the
// lane of the pair whose selecting bit is clear receives a summed over both
// lanes, the other b summed over both.  Lanes 32 apart (v_permlane32_swap:
// lanes 32-63 of a change places with lanes 0-31 of b) and 16 apart
// (v_permlane16_swap, in each half) need no selects; the rest go through
// DPP row moves whose lane pairing is an involution: row_ror:8 (lane ^ 8),
// row_half_mirror (lane ^ 7), quad_perm (lane ^ 2, lane ^ 1).
constexpr int kDppRor8 = 0x128, kDppHalfMirror = 0x141, kDppXor2 = 0x4E, kDppXor1 = 0xB1;

template <bool S32>
__device__ __forceinline__ unsigned swap_lanes(unsigned &a, unsigned b) {
  const auto r = S32 ? __builtin_amdgcn_permlane32_swap(a, b, false, false)
                     : __builtin_amdgcn_permlane16_swap(a, b, false, false);
  a = r[0];
  return r[1];
}
template <bool S32>
__device__ __forceinline__ float pair_sum_swap(float a, float b) {
  unsigned x = __float_as_uint(a);
  const unsigned y = swap_lanes<S32>(x, __float_as_uint(b));
  return __uint_as_float(x) + __uint_as_float(y);
}
template <bool S32>
__device__ __forceinline__ double pair_sum_swap(double a, double b) {
  const unsigned long long ua = __double_as_longlong(a), ub = __double_as_longlong(b);
  unsigned alo = (unsigned)ua, ahi = (unsigned)(ua >> 32);
  const unsigned blo = swap_lanes<S32>(alo, (unsigned)ub);
  const unsigned bhi = swap_lanes<S32>(ahi, (unsigned)(ub >> 32));
  return __longlong_as_double((long long)(((unsigned long long)ahi << 32) | alo)) +
         __longlong_as_double((long long)(((unsigned long long)bhi << 32) | blo));
}
template <int CTRL>
__device__ __forceinline__ double dpp_move(double x) {
  const unsigned long long u = __double_as_longlong(x);
  const int lo = __builtin_amdgcn_mov_dpp((int)u, CTRL, 0xF, 0xF, true);
  const int hi = __builtin_amdgcn_mov_dpp((int)(u >> 32), CTRL, 0xF, 0xF, true);
  return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}
template <int CTRL>
__device__ __forceinline__ double pair_sum_dpp(double a, double b, bool up) {
  return (up ? b : a) + dpp_move<CTRL>(up ? a : b);
}

// Sum over the wavefront of the run's partials s[g][d], scattered: lane l
// receives value (l >> 1) & 31 (W: eight atoms) or (l >> 2) & 15 (four),
// value 4 g + d.  The first exchange is FP32 for the mixed kernel (sums of
// two lanes' partials), the rest FP64, in a fixed order.
template <bool MIXED, bool W>
__device__ __forceinline__ double wave_sum_scatter(typename PairPrec<MIXED>::R (&s)[kPassAtoms][4],
                                                   int lane) {
  constexpr int V = W ? 32 : 16;
  double d[V / 2];
#pragma unroll
  for (int k = 0; k < V / 2; ++k)  // lane bit 5 selects value k + V/2
    d[k] = (double)pair_sum_swap<true>(s[k >> 2][k & 3], s[(k + V / 2) >> 2][(k + V / 2) & 3]);
#pragma unroll
  for (int k = 0; k < V / 4; ++k) d[k] = pair_sum_swap<false>(d[k], d[k + V / 4]);  // bit 4
  const bool b3 = lane & 8, b2 = lane & 4, b1 = lane & 2;
#pragma unroll
  for (int k = 0; k < V / 8; ++k) d[k] = pair_sum_dpp<kDppRor8>(d[k], d[k + V / 8], b3);
#pragma unroll
  for (int k = 0; k < V / 16; ++k) d[k] = pair_sum_dpp<kDppHalfMirror>(d[k], d[k + V / 16], b2);
  double r;
  if constexpr (W) {
    r = pair_sum_dpp<kDppXor2>(d[0], d[1], b1);
  } else {
    r = d[0];
    r += dpp_move<kDppXor2>(r);
  }
  r += dpp_move<kDppXor1>(r);
  return r;
}

template <bool MIXED, int NU, bool TRI>
__global__ __launch_bounds__(64) void k_pair_cell(DevParams P, const double4 *__restrict__ spos,
                                                  const int *__restrict__ sidx,
                                                  const int *__restrict__ start,
                                                  const int4 *__restrict__ cols, int ncol,
                                                  double *__restrict__ F, double *__restrict__ phi,
                                                  double *__restrict__ epart) {
  using T = typename PairPrec<MIXED>::T;
  using R = typename PairPrec<MIXED>::R;
  constexpr int CAP = PairPrec<MIXED>::CAP;  // multiple of 64
  constexpr int LB = 2;                      // candidate loads in flight per lane
  __shared__ T jbuf[CAP];
  __shared__ int seg_off[kSegMax + 1], seg_beg[kSegMax];
  __shared__ double seg_sx[kSegMax], seg_sy[kSegMax], seg_sz[kSegMax];
  constexpr int TAB = PairPrec<MIXED>::TAB;
  __shared__ unsigned char seg_of[TAB];  // range of each candidate of a round
  const int c = xcd_slab(blockIdx.x, P.ncell), lane = threadIdx.x;
  const int b0 = start[c], n = start[c + 1] - b0;
  const int cz = c % P.nc[2], cy = (c / P.nc[2]) % P.nc[1], cx = c / (P.nc[1] * P.nc[2]);
  // home-cell origin: all coordinates below are relative to it
  double ox, oy, oz;
  if constexpr (TRI) {
    const double fx = (double)cx / P.nc[0], fy = (double)cy / P.nc[1], fz = (double)cz / P.nc[2];
    ox = P.A[0] * fx + P.A[1] * fy + P.A[2] * fz;
    oy = P.A[3] * fx + P.A[4] * fy + P.A[5] * fz;
    oz = P.A[6] * fx + P.A[7] * fy + P.A[8] * fz;
  } else {
    ox = cx * P.cw[0], oy = cy * P.cw[1], oz = cz * P.cw[2];
  }
  const R cut2 = MIXED ? (R)P.fcut2 : (R)P.cut2;
  const float fcw0 = (float)P.cw[0], fcw1 = (float)P.cw[1], ficw2 = (float)(1.0 / P.cw[2]);
  const float fcut2 = P.fcut2;
  // cell membership may be off by binning round-off; FP32 column trimming too
  const float mrg = 1e-3f * (float)fmin(P.cw[0], fmin(P.cw[1], P.cw[2]));
  const unsigned long long lanes_below = (1ull << lane) - 1ull;
  const R far = (R)1e18;  // padding atoms sit far from everything
  double esum = 0.0;
  for (int ip = 0; ip < n; ip += kPassAtoms) {
    const int ni = min(kPassAtoms, n - ip);
    R ix[kPassAtoms], iy[kPassAtoms], iz[kPassAtoms];
    R lo[3] = {far, far, far}, hi[3] = {-far, -far, -far};
#pragma unroll
    for (int k = 0; k < kPassAtoms; ++k) {
      const double4 a = spos[b0 + ip + min(k, ni - 1)];
      ix[k] = wave_uniform((R)(a.x - ox));
      iy[k] = wave_uniform((R)(a.y - oy));
      iz[k] = wave_uniform((R)(a.z - oz));
      lo[0] = fmin(lo[0], ix[k]), hi[0] = fmax(hi[0], ix[k]);
      lo[1] = fmin(lo[1], iy[k]), hi[1] = fmax(hi[1], iy[k]);
      lo[2] = fmin(lo[2], iz[k]), hi[2] = fmax(hi[2], iz[k]);
    }
    const float flo0 = (float)lo[0], flo1 = (float)lo[1], flo2 = (float)lo[2];
    const float fhi0 = (float)hi[0], fhi1 = (float)hi[1], fhi2 = (float)hi[2];
    R s[kPassAtoms][4];
#pragma unroll
    for (int k = 0; k < kPassAtoms; ++k) s[k][0] = s[k][1] = s[k][2] = s[k][3] = (R)0;
    double tot = 0.0;  // this lane's share of the pass totals
    const bool wide = ni > 4;
    // lane l will hold value vi = 4 g + d of home atom g: its writer's
    // charge and original index are fetched now, used after the last run
    const int vi = wide ? (lane >> 1) & 31 : (lane >> 2) & 15;
    const bool writer = (lane & (wide ? 1 : 3)) == 0 && (vi >> 2) < ni;
    const int si = b0 + ip + min(vi >> 2, ni - 1);
    const double qi = reinterpret_cast<const double *>(spos + si)[3];
    const int oi = sidx[si];
    int np = 0;        // staged atoms in jbuf
    int c0 = 0, g0 = 0, total = 0, sl = 0;
    bool fresh = true, done = false;
    while (!done) {
      // ---- fill jbuf: rounds of 64 columns, batches of 64 x LB candidates
      for (;;) {
        if (fresh) {
          if (c0 >= ncol) {
            done = true;
            break;
          }
          // lane k: column c0 + k -> at most two contiguous atom ranges
          int beg0 = 0, cnt0 = 0, beg1 = 0, cnt1 = 0;
          double sx = 0, sy = 0, sz0 = 0, sx1 = 0, sy1 = 0, sz1 = 0;
          if (c0 + lane < ncol) {
            const int4 col = cols[c0 + lane];  // dx, dy, dzlo, dzhi
            // orthorhombic cells: trim the z range to the pass's bounding box
            // (triclinic cells are not axis-aligned: the whole column is
            // staged and the per-candidate bounding-box test decides)
            int zl = col.z, zh = col.w;
            float rz2 = 1.f;
            if constexpr (!TRI) {
              const float gx = fmaxf(fmaxf(col.x * fcw0 - mrg - fhi0, flo0 - (col.x + 1) * fcw0 - mrg), 0.f);
              const float gy = fmaxf(fmaxf(col.y * fcw1 - mrg - fhi1, flo1 - (col.y + 1) * fcw1 - mrg), 0.f);
              rz2 = fcut2 - gx * gx - gy * gy;
              const float rz = sqrtf(fmaxf(rz2, 0.f)) + mrg;
              zl = max(col.z, (int)floorf((flo2 - rz) * ficw2));
              zh = min(col.w, (int)floorf((fhi2 + rz) * ficw2));
            }
            if (rz2 > 0.f) {
              if (zl <= zh) {
                int X = cx + col.x, Y = cy + col.y;
                const int kx = floordiv(X, P.nc[0]), ky = floordiv(Y, P.nc[1]);
                X -= kx * P.nc[0], Y -= ky * P.nc[1];
                const int base = (X * P.nc[1] + Y) * P.nc[2];
                const int Z0 = cz + zl, Z1 = cz + zh;  // Z1 - Z0 < nc[2]: one wrap at most
                const int kz0 = floordiv(Z0, P.nc[2]);
                const int e0 = min(Z1, (kz0 + 1) * P.nc[2] - 1);
                beg0 = start[base + Z0 - kz0 * P.nc[2]];
                cnt0 = start[base + e0 - kz0 * P.nc[2] + 1] - beg0;
                lattice<TRI>(P, kx, ky, kz0, sx, sy, sz0);
                sx -= ox, sy -= oy, sz0 -= oz;
                if (e0 < Z1) {  // the rest lies in the next periodic image along z
                  beg1 = start[base];
                  cnt1 = start[base + Z1 - (kz0 + 1) * P.nc[2] + 1] - beg1;
                  lattice<TRI>(P, kx, ky, kz0 + 1, sx1, sy1, sz1);
                  sx1 -= ox, sy1 -= oy, sz1 -= oz;
                }
              }
            }
          }
          // non-empty ranges only, in column order: (ranges << 24) + atoms
          const int mine = ((int)(cnt0 > 0) + (int)(cnt1 > 0)) * (1 << 24) + cnt0 + cnt1;
          const int incl = wave_incl_scan(mine);
          const int excl = incl - mine, last = __builtin_amdgcn_readlane(incl, 63);
          const int kfirst = excl >> 24;
          int k = kfirst;
          const int o = excl & 0xffffff;
          if (cnt0 > 0) {
            seg_off[k] = o, seg_beg[k] = beg0;
            seg_sx[k] = sx, seg_sy[k] = sy, seg_sz[k] = sz0;
            ++k;
          }
          if (cnt1 > 0) {
            seg_off[k] = o + cnt0, seg_beg[k] = beg1;
            seg_sx[k] = sx1, seg_sy[k] = sy1, seg_sz[k] = sz1;
          }
          total = last & 0xffffff;
          if (lane == 0) seg_off[last >> 24] = total;  // sentinel
          if (total <= TAB) {
            // the range index of every candidate: one table read replaces
            // the walk over the ranges in the candidate loop
            for (int t = 0; t < cnt0; ++t) seg_of[o + t] = (unsigned char)kfirst;
            const int k1 = kfirst + (cnt0 > 0);
            for (int t = 0; t < cnt1; ++t) seg_of[o + cnt0 + t] = (unsigned char)k1;
          }
          __syncthreads();
          g0 = 0, sl = 0, fresh = false;
        }
        if (g0 >= total) {
          __syncthreads();  // the tables are rewritten by the next round
          c0 += 64, fresh = true;
          continue;
        }
        if (np > CAP - 64 * LB) break;  // evaluate before the next batch
        double4 a[LB];
        int sg[LB];
#pragma unroll
        for (int u = 0; u < LB; ++u) {  // loads first (clamped index: always valid)
          const int g = min(g0 + 64 * u + lane, total - 1);
          if (total <= TAB)
            sl = seg_of[g];
          else
            while (seg_off[sl + 1] <= g) ++sl;  // the sentinel total > g ends it
          sg[u] = sl;
          a[u] = spos[(unsigned)(seg_beg[sl] + g - seg_off[sl])];
        }
#pragma unroll
        for (int u = 0; u < LB; ++u) {
          T v;
          v.x = (R)(a[u].x + seg_sx[sg[u]]);
          v.y = (R)(a[u].y + seg_sy[sg[u]]);
          v.z = (R)(a[u].z + seg_sz[sg[u]]);
          v.w = (R)a[u].w;
          const R bx = fmax(fmax(lo[0] - v.x, v.x - hi[0]), (R)0);
          const R by = fmax(fmax(lo[1] - v.y, v.y - hi[1]), (R)0);
          const R bz = fmax(fmax(lo[2] - v.z, v.z - hi[2]), (R)0);
          const bool keep = g0 + 64 * u + lane < total && bx * bx + by * by + bz * bz < cut2;
          const unsigned long long bm = __ballot(keep);
          if (keep) jbuf[np + __popcll(bm & lanes_below)] = v;
          np += __popcll(bm);
        }
        g0 += 64 * LB;
      }
      if (np > 0) {
        // ---- evaluate the staged atoms (padded to a multiple of 64)
        const int npad = (np + 63) & ~63;
        if (np + lane < npad) {
          T pv;
          pv.x = far, pv.y = far, pv.z = far, pv.w = 0;
          jbuf[np + lane] = pv;
        }
        __syncthreads();
        pair_run<MIXED, NU>(P, jbuf, npad, ni, ix, iy, iz, s);
        __syncthreads();
        tot += wide ? wave_sum_scatter<MIXED, true>(s, lane) : wave_sum_scatter<MIXED, false>(s, lane);
#pragma unroll
        for (int k = 0; k < kPassAtoms; ++k) s[k][0] = s[k][1] = s[k][2] = s[k][3] = (R)0;
        np = 0;
      }
    }
    // ---- lane l holds value vi = 4 g + d of home atom g
    if (writer) {
      const int d = vi & 3;
      if (d == 0) {
        const double pv = tot - qi * P.self_img;
        if (phi) phi[oi] = pv;
        esum += 0.5 * qi * pv;
      } else if (F) {
        F[3 * (long)oi + d - 1] = qi * tot;
      }
    }
  }
  esum = wave_incl_scan(esum);  // lane 63: the wavefront total
  if (lane == 63) epart[c] = esum;
}

// Pairs i < j (cell-sorted order) whose nearest image lies within radius
// sqrt(r2max): the reference's neighbour-list statistic (forcefield0.c:109-160;
// radius a0 + margin), counted by cell columns built for that radius.  One wavefront per home cell; a round's
// columns (up to 64) become at most two contiguous index ranges each, laid
// out by a prefix sum, and the lanes take the candidates in turn, each
// against every home atom.  A pair is counted at the image whose fractional
// separation has every component in [-1/2, 1/2), r - floor(r/L + 1/2) L as
// in the reference (forcefield0.c:140-148), so that of two images tied at
// exactly half a box only one is counted.  In an orthorhombic box whose
// edges exceed twice the radius that image is the only one inside it, and
// IMG = false skips the test.
constexpr int kCountAhead = 4;  // candidate loads in flight per lane

template <bool IMG>
__global__ __launch_bounds__(64) void k_pair_count(DevParams P, const double4 *__restrict__ spos,
                                                   const int *__restrict__ start,
                                                   const int4 *__restrict__ cols, int ncol,
                                                   double r2max,
                                                   unsigned *__restrict__ cell_count) {
  __shared__ int seg_off[kSegMax + 1], seg_beg[kSegMax];
  __shared__ double seg_sx[kSegMax], seg_sy[kSegMax], seg_sz[kSegMax];
  const int c = blockIdx.x, lane = threadIdx.x;
  const int b0 = start[c], n = start[c + 1] - b0;
  if (n == 0) {  // wavefront-uniform
    if (lane == 0) cell_count[c] = 0;
    return;
  }
  const int cz = c % P.nc[2], cy = (c / P.nc[2]) % P.nc[1], cx = c / (P.nc[1] * P.nc[2]);
  unsigned cnt = 0;
  // home atoms kPassAtoms at a time, held in registers (wavefront-uniform)
  for (int ip = 0; ip < n; ip += kPassAtoms) {
    const int ni = min(kPassAtoms, n - ip);
    double hx[kPassAtoms], hy[kPassAtoms], hz[kPassAtoms];
#pragma unroll
    for (int u = 0; u < kPassAtoms; ++u) {
      const double4 a = spos[b0 + ip + min(u, ni - 1)];
      hx[u] = wave_uniform(a.x), hy[u] = wave_uniform(a.y), hz[u] = wave_uniform(a.z);
    }
    const int i0 = b0 + ip;  // sorted index of home atom u: i0 + u
    for (int c0 = 0; c0 < ncol; c0 += 64) {
      int beg0 = 0, cnt0 = 0, beg1 = 0, cnt1 = 0;
      double sx = 0, sy = 0, sz0 = 0, sz1 = 0, sx1 = 0, sy1 = 0;
      if (c0 + lane < ncol) {
        const int4 col = cols[c0 + lane];  // dx, dy, dzlo, dzhi (dzhi - dzlo < nc[2])
        int X = cx + col.x, Y = cy + col.y;
        const int kx = floordiv(X, P.nc[0]), ky = floordiv(Y, P.nc[1]);
        X -= kx * P.nc[0], Y -= ky * P.nc[1];
        const int base = (X * P.nc[1] + Y) * P.nc[2];
        const int Z0 = cz + col.z, Z1 = cz + col.w;
        const int kz0 = floordiv(Z0, P.nc[2]);
        const int e0 = min(Z1, (kz0 + 1) * P.nc[2] - 1);
        beg0 = start[base + Z0 - kz0 * P.nc[2]];
        cnt0 = start[base + e0 - kz0 * P.nc[2] + 1] - beg0;
        lattice_rt(P, kx, ky, kz0, sx, sy, sz0);
        if (e0 < Z1) {  // the rest lies in the next periodic image along z
          beg1 = start[base];
          cnt1 = start[base + Z1 - (kz0 + 1) * P.nc[2] + 1] - beg1;
          lattice_rt(P, kx, ky, kz0 + 1, sx1, sy1, sz1);
        }
      }
      const int mine = ((int)(cnt0 > 0) + (int)(cnt1 > 0)) * (1 << 24) + cnt0 + cnt1;
      const int incl = wave_incl_scan(mine);
      const int excl = incl - mine, last = __builtin_amdgcn_readlane(incl, 63);
      int k = excl >> 24;
      const int o = excl & 0xffffff;
      if (cnt0 > 0) {
        seg_off[k] = o, seg_beg[k] = beg0;
        seg_sx[k] = sx, seg_sy[k] = sy, seg_sz[k] = sz0;
        ++k;
      }
      if (cnt1 > 0) {
        seg_off[k] = o + cnt0, seg_beg[k] = beg1;
        seg_sx[k] = sx1, seg_sy[k] = sy1, seg_sz[k] = sz1;
      }
      const int total = last & 0xffffff;
      if (lane == 0) seg_off[last >> 24] = total;  // sentinel
      __syncthreads();
      // kCountAhead candidates per lane in flight: their loads first
      int sl = 0;
      for (int g0 = lane; g0 - lane < total; g0 += 64 * kCountAhead) {
        double4 aj[kCountAhead];
        double lx[kCountAhead], ly[kCountAhead], lz[kCountAhead];
        int jj[kCountAhead];
#pragma unroll
        for (int v = 0; v < kCountAhead; ++v) {
          const int g = min(g0 + 64 * v, total - 1);  // clamped: always a valid index
          while (seg_off[sl + 1] <= g) ++sl;  // the sentinel total > g ends it
          jj[v] = g0 + 64 * v < total ? seg_beg[sl] + g - seg_off[sl] : -1;
          aj[v] = spos[seg_beg[sl] + g - seg_off[sl]];
          lx[v] = seg_sx[sl], ly[v] = seg_sy[sl], lz[v] = seg_sz[sl];
        }
#pragma unroll
        for (int v = 0; v < kCountAhead; ++v) {
          const int j = jj[v];  // -1 past the end: below every home index
#pragma unroll
          for (int u = 0; u < kPassAtoms; ++u) {
            if (u >= ni) break;  // wavefront-uniform
            const double d[3] = {aj[v].x + (lx[v] - hx[u]), aj[v].y + (ly[v] - hy[u]),
                                 aj[v].z + (lz[v] - hz[u])};
            bool in = j > i0 + u && d[0] * d[0] + d[1] * d[1] + d[2] * d[2] < r2max;
            if constexpr (IMG) {
              if (in) {
                double sd[3];
                to_frac_rt(P, d, sd);
                in = floor(sd[0] + 0.5) == 0.0 && floor(sd[1] + 0.5) == 0.0 &&
                     floor(sd[2] + 0.5) == 0.0;
              }
            }
            cnt += in;
          }
        }
      }
      __syncthreads();  // the tables are rewritten by the next round
    }
  }
  cnt = wave_incl_scan((int)cnt);  // lane 63: the wavefront total
  if (lane == 63) cell_count[c] = cnt;  // one count per cell: no contended atomic
}

// ---------------------------------------------------------------- spread
// anterpolate (forcefield1.c:324-368): q^1_{m-i+NU/2} += q Q(i+t) per axis,
// restated per grid brick as a matrix product on the matrix cores.  A brick
// is 16 x 4 x 4 points; with wx, wy, wz an atom's B-spline weights folded
// onto the brick's x, y and z indices,
//     Q[x][(y, z)] = sum_atoms wx[x] * (q wy[y] wz[z]),
// a 16 x 16 output tile accumulated by v_mfma_f64_16x16x4_f64 over the atoms
// four at a time.  The workgroup lists the pair-search cell columns whose
// atoms can reach the brick (atoms are sorted by cell, so a column is one or
// two contiguous index ranges); its four wavefronts split the candidates,
// queue each atom that reaches the brick and, 32 at a time, fold the queued
// atoms' weights into LDS records (two lanes per atom) and issue the MFMAs.
// Their tiles are summed in a fixed order and every grid point is stored
// once.  No atomics: the result is deterministic.  Records of 32 atoms per
// wavefront keep a workgroup at about 45 KB of LDS, so it fits beside the
// pair kernel's wavefronts.
constexpr int kSX = 16, kSY = 4, kSZ = 4;
constexpr int kSRanges = 512;       // index ranges per batch of cell columns
constexpr int kSJump = 8;           // ranges skipped per step of the candidate walk
constexpr int kWXs = 17, kWYs = 5;  // padded record strides (doubles)
constexpr int kSFlush = 32;         // atoms per flush (records per wavefront)
constexpr int kSQueue = kSFlush + 64;  // queued atoms: < kSFlush + one batch of 64
constexpr unsigned kSpreadFine = 2048;  // bricks from which one wavefront takes a brick

typedef double sp_f64x4 __attribute__((ext_vector_type(4)));

// Cells holding atoms whose grid index m lies in [g0 - half, g0 + B + half - 2],
// i.e. fractional coordinate in [(g0 - half) / M, (g0 + B + half - 1) / M); one
// grid spacing of margin on each side absorbs rounding in the binning.
__device__ __forceinline__ void cell_span(int g0, int B, int half, int M, int nc, int &lo,
                                          int &len) {
  lo = (int)floor((double)nc * (g0 - half - 1) / M);
  int hi = (int)floor((double)nc * (g0 + B + half) / M);
  len = hi - lo + 1;
  if (len >= nc) lo = 0, len = nc;
}

// Sorted index of candidate k in the range table (roff[lo] <= k < roff[lo + 1]),
// advancing from a previous position lo (k only grows); -1 past the end.
__device__ __forceinline__ int spread_index(const int *roff, const int *rbeg, int nr, int k,
                                            int total, int &lo) {
  if (k >= total) return -1;
  // roff[nr .. nr + kSJump] = total > k ends both walks by the last range
  while (roff[lo + kSJump] <= k) lo += kSJump;
  while (roff[lo + 1] <= k) ++lo;
  return rbeg[lo] + (k - roff[lo]);
}

// Fractional offsets t and the brick-relative index of the lowest support
// point m - NU/2 + 1 per axis, reduced into [0, M) (m lies in [0, M] and org
// in [0, M): one addition of M leaves it negative for NU >= 6 with org close
// to M, and a negative index would write below the atom's LDS record);
// reach: the support meets the brick.
template <int NU, bool TRI>
__device__ __forceinline__ bool spread_place(const DevParams &P, const double4 &at,
                                             const int dims[3], const int org[3],
                                             const int ext[3], double t[3], int rlo[3]) {
  const double r[3] = {at.x, at.y, at.z};
  double sf[3];
  to_frac<TRI>(P, r, sf);
  bool reach = true;
#pragma unroll
  for (int ax = 0; ax < 3; ++ax) {
    double sv = sf[ax];
    sv = sv - floor(sv);
    const double tt = dims[ax] * sv, m = floor(tt);
    t[ax] = tt - m;
    int v = ((int)m - NU / 2 + 1 - org[ax]) % dims[ax];
    v += v < 0 ? dims[ax] : 0;
    rlo[ax] = v;
    reach = reach && (v < ext[ax] || v > dims[ax] - NU);
  }
  return reach;
}

// One axis of an atom's record: weight i of the support sits at brick index
// rlo + NU - 1 - i (periodic), kept where it falls inside the brick (EXT).
template <int NU, int EXT>
__device__ __forceinline__ void spread_fold(const DevParams &P, double t, double sc, int rlo,
                                            int M, double *rec) {
  double w[NU];
  bspline<NU, false>(P.Q, t, w, nullptr);
  if (M >= NU) {  // the support points are distinct grid points
#pragma unroll
    for (int i = 0; i < NU; ++i) {
      int rr = rlo + (NU - 1 - i);
      rr -= rr >= M ? M : 0;
      if (rr < EXT) rec[rr] = sc * w[i];
    }
  } else {  // grids narrower than the support: points repeat, accumulate
    for (int i = 0; i < NU; ++i) {
      const int rr = (rlo + (NU - 1 - i)) % M;
      if (rr < EXT) rec[rr] += sc * w[i];
    }
  }
}

// The queued atoms (`at`: atom lane % kSFlush, valid when `ok`): folded
// weights into the records -- lanes < kSFlush build the x record of their
// atom, the other half the y (times q) and z records -- then
// Q += WX^T (WY x WZ) by natoms / 4 MFMAs.
template <int NU, bool TRI>
__device__ __forceinline__ void spread_flush(const DevParams &P, const double4 &at, bool ok,
                                             int natoms, const int dims[3], const int org[3],
                                             double *RX, double *RY, double *RZ, int lane,
                                             sp_f64x4 &acc) {
  constexpr int ext[3] = {kSX, kSY, kSZ};
  {
    const int a = lane % kSFlush;
    double *const rx = RX + a * kWXs, *const ry = RY + a * kWYs, *const rz = RZ + a * kWYs;
    if (lane < kSFlush) {
#pragma unroll
      for (int u = 0; u < kSX; ++u) rx[u] = 0.0;
    } else {
#pragma unroll
      for (int u = 0; u < kSY; ++u) ry[u] = 0.0, rz[u] = 0.0;
    }
    if (ok) {
      double t[3];
      int rlo[3];
      spread_place<NU, TRI>(P, at, dims, org, ext, t, rlo);
      if (lane < kSFlush) {
        spread_fold<NU, kSX>(P, t[0], 1.0, rlo[0], dims[0], rx);
      } else {
        spread_fold<NU, kSY>(P, t[1], at.w, rlo[1], dims[1], ry);
        spread_fold<NU, kSZ>(P, t[2], 1.0, rlo[2], dims[2], rz);
      }
    }
  }
  // LDS operations of one wavefront complete in order; the barrier keeps the
  // compiler from moving the reads above the writes
  __builtin_amdgcn_wave_barrier();
  const int row = lane & 15, kq = lane >> 4;
  for (int a0 = 0; a0 < natoms; a0 += 4) {
    const int a = a0 + kq;
    const double av = RX[a * kWXs + row];
    const double bv = RY[a * kWYs + (row >> 2)] * RZ[a * kWYs + (row & 3)];
    acc = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, acc, 0, 0, 0);
  }
  __builtin_amdgcn_wave_barrier();
}

// WPB wavefronts per brick: 4 where the grid has few bricks (they split the
// brick's candidates), 1 on fine grids, where the bricks alone fill the GPU
// and a wavefront needs no barrier partners and about 11 KB of LDS.
template <int NU, int WPB, bool TRI>
__global__ __launch_bounds__(64 * WPB) void k_spread_mfma(DevParams P, GridDims g,
                                                          const double4 *__restrict__ spos,
                                                          const int *__restrict__ start,
                                                          double *__restrict__ q1,
                                                          unsigned *__restrict__ q1b) {
  chain_issue_priority();
  constexpr int BS = 64 * WPB;
  // per wavefront: records of kSFlush atoms and a queue of sorted indices
  __shared__ double WX[WPB * kSFlush * kWXs];
  __shared__ double WY[WPB * kSFlush * kWYs], WZ[WPB * kSFlush * kWYs];
  __shared__ double4 queue[WPB * kSQueue];
  __shared__ int rbeg[kSRanges], roff[kSRanges + 1 + kSJump];  // total repeated past the end
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int nbz = (g.z + kSZ - 1) / kSZ, nby = (g.y + kSY - 1) / kSY;
  int b = blockIdx.x;
  const int gz0 = (b % nbz) * kSZ;
  b /= nbz;
  const int gy0 = (b % nby) * kSY, gx0 = (b / nby) * kSX;
  int lox, lenx, loy, leny, loz, lenz;
  cell_span(gx0, kSX, NU / 2, g.x, P.nc[0], lox, lenx);
  cell_span(gy0, kSY, NU / 2, g.y, P.nc[1], loy, leny);
  cell_span(gz0, kSZ, NU / 2, g.z, P.nc[2], loz, lenz);
  const int ncol = lenx * leny;
  const int dims[3] = {g.x, g.y, g.z}, org[3] = {gx0, gy0, gz0}, ext[3] = {kSX, kSY, kSZ};
  const unsigned long long below = (1ull << lane) - 1ull;
  double *const RX = WX + wv * kSFlush * kWXs, *const RY = WY + wv * kSFlush * kWYs,
               *const RZ = WZ + wv * kSFlush * kWYs;
  double4 *const Qw = queue + wv * kSQueue;
  int npend = 0;  // queued atoms of this wavefront (uniform)
  sp_f64x4 acc = {0.0, 0.0, 0.0, 0.0};
  for (int c0 = 0; c0 < ncol; c0 += kSRanges / 2) {
    const int nc_b = min(ncol - c0, kSRanges / 2);
    __syncthreads();  // previous batch done with the table
    for (int cc = tid; cc < nc_b; cc += BS) {
      int col = c0 + cc;
      int x = wrapi(lox + col / leny, P.nc[0]), y = wrapi(loy + col % leny, P.nc[1]);
      int base = (x * P.nc[1] + y) * P.nc[2];
      int z0 = wrapi(loz, P.nc[2]), z1 = z0 + lenz;
      int e1 = min(z1, P.nc[2]), e2 = z1 > P.nc[2] ? z1 - P.nc[2] : 0;
      rbeg[2 * cc] = start[base + z0];
      roff[2 * cc + 1] = start[base + e1] - start[base + z0];
      rbeg[2 * cc + 1] = start[base];
      roff[2 * cc + 2] = start[base + e2] - start[base];
    }
    __syncthreads();
    const int nr = 2 * nc_b;
    if (wv == 0) {  // inclusive prefix of roff[1..nr]: 8 entries per lane
      int loc[8], run = 0;
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        int k = 8 * lane + u;
        run += k < nr ? roff[k + 1] : 0;
        loc[u] = run;
      }
      const int incl = wave_incl_scan(run);
      const int excl = incl - run;
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        int k = 8 * lane + u;
        if (k < nr) roff[k + 1] = excl + loc[u];
      }
      if (lane == 0) roff[0] = 0;
      const int tot = __builtin_amdgcn_readlane(incl, 63);
      if (lane < kSJump) roff[nr + 1 + lane] = tot;  // the jumps of spread_index stop here
    }
    __syncthreads();
    const int total = roff[nr];
    // wavefront wv takes candidates 64 (WPB r + wv) + lane, r = 0, 1, ...;
    // the next candidate's atom is fetched while the current one is processed
    int k = wv * 64 + lane, lo = 0;
    int inext = spread_index(roff, rbeg, nr, k, total, lo);
    double4 nxt = inext >= 0 ? spos[inext] : make_double4(0.0, 0.0, 0.0, 0.0);
    for (; k - lane < total; k += BS) {
      const double4 at = nxt;
      const int icur = inext;
      inext = spread_index(roff, rbeg, nr, k + BS, total, lo);
      nxt = inext >= 0 ? spos[inext] : make_double4(0.0, 0.0, 0.0, 0.0);
      double t[3];
      int rlo[3];
      const bool reach = icur >= 0 && spread_place<NU, TRI>(P, at, dims, org, ext, t, rlo);
      const unsigned long long bm = __ballot(reach);
      if (reach) Qw[npend + __popcll(bm & below)] = at;
      npend += __popcll(bm);
      while (npend >= kSFlush) {
        __builtin_amdgcn_wave_barrier();
        const double4 mine = Qw[lane % kSFlush];
        const double4 rest = Qw[min(kSFlush + lane, kSQueue - 1)];
        __builtin_amdgcn_wave_barrier();
        spread_flush<NU, TRI>(P, mine, true, kSFlush, dims, org, RX, RY, RZ, lane, acc);
        // move the leftover queue entries [kSFlush, npend) to the front
        if (kSFlush + lane < npend) Qw[lane] = rest;
        npend -= kSFlush;
        __builtin_amdgcn_wave_barrier();
      }
    }
  }
  if (npend > 0) {  // the last, partial group of atoms
    __builtin_amdgcn_wave_barrier();
    const double4 mine = Qw[lane % kSFlush];
    spread_flush<NU, TRI>(P, mine, lane % kSFlush < npend, (npend + 3) & ~3, dims, org, RX, RY, RZ,
                     lane, acc);
  }
  // ---- sum the wavefronts' tiles in a fixed order and store
  if constexpr (WPB > 1) {
    __syncthreads();
#pragma unroll
    for (int i = 0; i < 4; ++i) WX[(wv * 4 + i) * 64 + lane] = acc[i];
    __syncthreads();
  }
  if (wv == 0) {
    // f64 16x16x4 result layout: register i of lane l holds row 4 i + l / 16,
    // column l % 16
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      double v = acc[i];
      for (int w = 1; w < WPB; ++w) v += WX[(w * 4 + i) * 64 + lane];
      const int x = gx0 + 4 * i + (lane >> 4), col = lane & 15;
      const int y = gy0 + (col >> 2), z = gz0 + (col & 3);
      if (x < g.x && y < g.y && z < g.z) {
        const long id = ((long)x * g.y + y) * g.z + z;
        q1[id] = v;
        if (q1b) q1b[id] = bf16_split(v);
      }
    }
  }
}

// ---- spreading from per-atom placements (grids with M_x >= 16 + NU and
// M_y, M_z >= 4 + NU, every axis below 1024 points)
//
// k_spread_place computes once per atom, in cell-sorted order, what every
// brick the atom reaches needs: the lowest support point m - NU/2 + 1 (mod M)
// per axis, packed 10 bits each, and the fractional offsets t with the
// charge.  k_spread_bricks then tests a candidate with three integer
// comparisons and builds the MFMA operands of a flush directly from each
// atom's NU weights per axis: the lane for row r (A) or column (y, z) (B)
// picks weight i = rel + NU - 1 - r of its atom, zero outside [0, NU), where
// rel in [1 - NU, ext - 1] is the support's first point relative to the
// brick (negative where the support wraps in from below).  On these grids
// one such offset per axis covers every point of the brick the support
// reaches, so no point needs two weights of one atom.  The arithmetic, the
// atoms of each MFMA and their order are those of k_spread_mfma.
constexpr int kSPFlush = 16;               // atoms per flush
constexpr int kSPAhead = 4;                // candidate placements in flight per lane
constexpr int kSPQueue = kSPFlush + 64;    // queued atoms: < kSPFlush + one batch of 64

template <bool TRI>
__global__ __launch_bounds__(256) void k_spread_place(DevParams P, GridDims g,
                                                      const double4 *__restrict__ spos,
                                                      unsigned *__restrict__ sbase,
                                                      double4 *__restrict__ st) {
  chain_issue_priority();
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= P.N) return;
  place_atom<TRI>(P, g, spos[i], sbase[i], st[i]);
}

template <int NU, int WPB>
__global__ __launch_bounds__(64 * WPB) void k_spread_bricks(DevParams P, GridDims g,
                                                            const unsigned *__restrict__ sbase,
                                                            const double4 *__restrict__ st,
                                                            const int *__restrict__ start,
                                                            double *__restrict__ q1,
                                                            unsigned *__restrict__ q1b) {
  chain_issue_priority();
  constexpr int BS = 64 * WPB, NW = 3 * NU;
  constexpr int WSZ = WPB * (kSPFlush * NW > 256 ? kSPFlush * NW : 256);
  __shared__ double W[WSZ];                  // weights of one flush's atoms; the tile sum
  __shared__ double4 queue[WPB * kSPQueue];  // (t, q) of the queued atoms
  __shared__ unsigned qrel[WPB * kSPQueue];  // their offsets rel + NU - 1, 8 bits per axis
  __shared__ int rbeg[kSRanges], roff[kSRanges + 1 + kSJump];  // total repeated past the end
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int nbz = (g.z + kSZ - 1) / kSZ, nby = (g.y + kSY - 1) / kSY;
  int b = blockIdx.x;
  const int gz0 = (b % nbz) * kSZ;
  b /= nbz;
  const int gy0 = (b % nby) * kSY, gx0 = (b / nby) * kSX;
  int lox, lenx, loy, leny, loz, lenz;
  cell_span(gx0, kSX, NU / 2, g.x, P.nc[0], lox, lenx);
  cell_span(gy0, kSY, NU / 2, g.y, P.nc[1], loy, leny);
  cell_span(gz0, kSZ, NU / 2, g.z, P.nc[2], loz, lenz);
  const int ncol = lenx * leny;
  const int dims[3] = {g.x, g.y, g.z}, org[3] = {gx0, gy0, gz0}, ext[3] = {kSX, kSY, kSZ};
  const unsigned long long below = (1ull << lane) - 1ull;
  double *const Wv = W + wv * (WSZ / WPB);
  double4 *const Qw = queue + wv * kSPQueue;
  unsigned *const Rw = qrel + wv * kSPQueue;
  const int row = lane & 15, kq = lane >> 4;
  int npend = 0;  // queued atoms of this wavefront (uniform)
  sp_f64x4 acc = {0.0, 0.0, 0.0, 0.0};
  // weights of queued atoms [0, n) into Wv, then Q += WX^T (WY x WZ) by
  // (n + 3) / 4 MFMAs (atoms n .. of the last group contribute zeros)
  auto flush = [&](int n) {
    __builtin_amdgcn_wave_barrier();
#pragma unroll
    for (int r0 = 0; r0 < 3 * kSPFlush; r0 += 64) {  // job: (axis, atom)
      const int j = r0 + lane, a = j % kSPFlush, ax = j / kSPFlush;
      if (j < 3 * kSPFlush && a < n) {
        const double4 v = Qw[a];
        const double t = ax == 0 ? v.x : ax == 1 ? v.y : v.z, sc = ax == 1 ? v.w : 1.0;
        double w[NU];
        bspline<NU, false>(P.Q, t, w, nullptr);
#pragma unroll
        for (int i = 0; i < NU; ++i) Wv[a * NW + ax * NU + i] = sc * w[i];
      }
    }
    __builtin_amdgcn_wave_barrier();
    // all four groups' operands first (their LDS reads overlap), then the MFMAs
    double av[kSPFlush / 4], bv[kSPFlush / 4];
#pragma unroll
    for (int gi = 0; gi < kSPFlush / 4; ++gi) {
      const int a = 4 * gi + kq;
      const unsigned pr = Rw[a];
      const int ix = (int)(pr & 255u) - row, iy = (int)((pr >> 8) & 255u) - (row >> 2),
                iz = (int)((pr >> 16) & 255u) - (row & 3);
      const double *wa = Wv + a * NW;
      const bool ok = a < n;
      const double wx = wa[min(max(ix, 0), NU - 1)];
      const double wy = wa[NU + min(max(iy, 0), NU - 1)];
      const double wz = wa[2 * NU + min(max(iz, 0), NU - 1)];
      // (LDS past the queued atoms holds stale values: both operands zero)
      av[gi] = ok && (unsigned)ix < (unsigned)NU ? wx : 0.0;
      bv[gi] = ok && (unsigned)iy < (unsigned)NU && (unsigned)iz < (unsigned)NU ? wy * wz : 0.0;
    }
#pragma unroll
    for (int gi = 0; gi < kSPFlush / 4; ++gi)
      if (4 * gi < n) acc = __builtin_amdgcn_mfma_f64_16x16x4f64(av[gi], bv[gi], acc, 0, 0, 0);
    __builtin_amdgcn_wave_barrier();
  };
  for (int c0 = 0; c0 < ncol; c0 += kSRanges / 2) {
    const int nc_b = min(ncol - c0, kSRanges / 2);
    __syncthreads();  // previous batch done with the table
    for (int cc = tid; cc < nc_b; cc += BS) {
      int col = c0 + cc;
      int x = wrapi(lox + col / leny, P.nc[0]), y = wrapi(loy + col % leny, P.nc[1]);
      int base = (x * P.nc[1] + y) * P.nc[2];
      int z0 = wrapi(loz, P.nc[2]), z1 = z0 + lenz;
      int e1 = min(z1, P.nc[2]), e2 = z1 > P.nc[2] ? z1 - P.nc[2] : 0;
      rbeg[2 * cc] = start[base + z0];
      roff[2 * cc + 1] = start[base + e1] - start[base + z0];
      rbeg[2 * cc + 1] = start[base];
      roff[2 * cc + 2] = start[base + e2] - start[base];
    }
    __syncthreads();
    const int nr = 2 * nc_b;
    if (wv == 0) {  // inclusive prefix of roff[1..nr]: 8 entries per lane
      int loc[8], run = 0;
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        int k = 8 * lane + u;
        run += k < nr ? roff[k + 1] : 0;
        loc[u] = run;
      }
      const int incl = wave_incl_scan(run);
      const int excl = incl - run;
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        int k = 8 * lane + u;
        if (k < nr) roff[k + 1] = excl + loc[u];
      }
      if (lane == 0) roff[0] = 0;
      const int tot = __builtin_amdgcn_readlane(incl, 63);
      if (lane < kSJump) roff[nr + 1 + lane] = tot;  // the jumps of spread_index stop here
    }
    __syncthreads();
    const int total = roff[nr];
    // wavefront wv takes candidates 64 (WPB r + wv) + lane, r = 0, 1, ...,
    // kSPAhead at a time: their placements are fetched together, then tested
    // in candidate order
    int lo = 0;
    for (int k0 = wv * 64 + lane; k0 - lane < total; k0 += kSPAhead * BS) {
      unsigned bc[kSPAhead];
      double4 tc[kSPAhead];
      bool vc[kSPAhead];
#pragma unroll
      for (int u = 0; u < kSPAhead; ++u) {
        const int ic = spread_index(roff, rbeg, nr, k0 + u * BS, total, lo);
        vc[u] = ic >= 0;
        bc[u] = vc[u] ? sbase[ic] : 0u;
        tc[u] = vc[u] ? st[ic] : make_double4(0.0, 0.0, 0.0, 0.0);
      }
#pragma unroll
      for (int u = 0; u < kSPAhead; ++u) {
        if (k0 - lane + u * BS >= total) break;  // wavefront-uniform
        bool reach = vc[u];
        unsigned rel = 0;
#pragma unroll
        for (int ax = 0; ax < 3; ++ax) {
          int d = (int)((bc[u] >> (10 * ax)) & 1023u) - org[ax];
          d += d < 0 ? dims[ax] : 0;
          reach = reach && (d < ext[ax] || d > dims[ax] - NU);
          rel |= (unsigned)((d < ext[ax] ? d : d - dims[ax]) + NU - 1) << (8 * ax);
        }
        const unsigned long long bm = __ballot(reach);
        if (reach) {
          const int slot = npend + __popcll(bm & below);
          Qw[slot] = tc[u];
          Rw[slot] = rel;
        }
        npend += __popcll(bm);
        while (npend >= kSPFlush) {
          flush(kSPFlush);
          // move the leftover queue entries [kSPFlush, npend) to the front
          const bool mv = kSPFlush + lane < npend;
          const double4 rest = Qw[min(kSPFlush + lane, kSPQueue - 1)];
          const unsigned rrel = Rw[min(kSPFlush + lane, kSPQueue - 1)];
          __builtin_amdgcn_wave_barrier();
          if (mv) Qw[lane] = rest, Rw[lane] = rrel;
          npend -= kSPFlush;
          __builtin_amdgcn_wave_barrier();
        }
      }
    }
  }
  if (npend > 0) flush(npend);  // the last, partial group of atoms
  // ---- sum the wavefronts' tiles in a fixed order and store
  if constexpr (WPB > 1) {
    __syncthreads();
#pragma unroll
    for (int i = 0; i < 4; ++i) W[(wv * 4 + i) * 64 + lane] = acc[i];
    __syncthreads();
  }
  if (wv == 0) {
    // f64 16x16x4 result layout: register i of lane l holds row 4 i + l / 16,
    // column l % 16
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      double v = acc[i];
      for (int w = 1; w < WPB; ++w) v += W[(w * 4 + i) * 64 + lane];
      const int x = gx0 + 4 * i + (lane >> 4), col = lane & 15;
      const int y = gy0 + (col >> 2), z = gz0 + (col & 3);
      if (x < g.x && y < g.y && z < g.z) {
        const long id = ((long)x * g.y + y) * g.z + z;
        q1[id] = v;
        if (q1b) q1b[id] = bf16_split(v);
      }
    }
  }
}

// ---------------------------------------------------------------- restrict
// restrict_ (forcefield1.c:401-477): q^l_m = sum_n J_nx J_ny J_nz
// q^{l-1}_{2m+n}, n in [-NU/2, NU/2]^3, periodic, as separable passes like
// the reference's z -> y -> x (:442-472).  A workgroup owns a 2^3 coarse
// brick: the z pass reads the fine values it needs straight from global
// memory (each thread's NU + 1 loads per value in flight together), then
// contracts y and x through LDS.  1.6 KB of LDS at NU = 6, so that several
// of these one-wavefront workgroups fit on a CU beside the pair kernel (a
// staged (NU+3)^3 tile, 7.4 KB, let one run per CU there: the 1728 bricks
// of the 24^3 level took seven rounds).
constexpr int kRB = 2;
constexpr int kRBS = 64;  // one wavefront per brick: the work is small, and few
                          // wavefronts find slots quickly beside the pair kernel

template <int NU>
__global__ __launch_bounds__(kRBS) void k_restrict(DevParams P, GridDims f, GridDims c,
                                                  const double *__restrict__ qf,
                                                  double *__restrict__ qc,
                                                  unsigned *__restrict__ qcb) {
  chain_issue_priority();
  // coarse brick of kRB^3 points; fine indices 2m + n - H, n in [0, NU]
  constexpr int H = NU / 2, FW = 2 * (kRB - 1) + NU + 1;
  constexpr int NZ = FW * FW * kRB;        // z-pass outputs
  constexpr int PER = (NZ + kRBS - 1) / kRBS;
  __shared__ double t1[FW * FW * kRB];    // [fx][fy][cz]
  __shared__ double t2[FW * kRB * kRB];   // [fx][cy][cz]
  const int tid = threadIdx.x;
  const int nbz = (c.z + kRB - 1) / kRB, nby = (c.y + kRB - 1) / kRB;
  int b = blockIdx.x;
  const int cz0 = (b % nbz) * kRB;
  b /= nbz;
  const int cy0 = (b % nby) * kRB, cx0 = (b / nby) * kRB;
  const int fx0 = 2 * cx0 - H, fy0 = 2 * cy0 - H, fz0 = 2 * cz0 - H;
  {  // along z, from global memory: all of this thread's loads first
    double v[PER][NU + 1];
#pragma unroll
    for (int u = 0; u < PER; ++u) {
      const int k = min(tid + u * kRBS, NZ - 1);  // past the end: a valid address, not stored
      const int cz = k % kRB, fxy = k / kRB, a = fxy / FW, bb = fxy % FW;
      const double *col = qf + ((long)wrapi(fx0 + a, f.x) * f.y + wrapi(fy0 + bb, f.y)) * f.z;
#pragma unroll
      for (int n = 0; n <= NU; ++n) v[u][n] = col[wrapi(fz0 + 2 * cz + n, f.z)];
    }
#pragma unroll
    for (int u = 0; u < PER; ++u) {
      double acc = 0.0;
#pragma unroll
      for (int n = 0; n <= NU; ++n) acc += P.J[n] * v[u][n];
      if (tid + u * kRBS < NZ) t1[tid + u * kRBS] = acc;
    }
  }
  __syncthreads();
  for (int k = tid; k < FW * kRB * kRB; k += kRBS) {  // along y
    const int cz = k % kRB, cy = (k / kRB) % kRB, fx = k / (kRB * kRB);
    double acc = 0.0;
#pragma unroll
    for (int n = 0; n <= NU; ++n) acc += P.J[n] * t1[(fx * FW + 2 * cy + n) * kRB + cz];
    t2[k] = acc;
  }
  __syncthreads();
  if (tid < kRB * kRB * kRB) {  // along x
    const int cz = tid % kRB, cy = (tid / kRB) % kRB, cx = tid / (kRB * kRB);
    const int x = cx0 + cx, y = cy0 + cy, z = cz0 + cz;
    if (x < c.x && y < c.y && z < c.z) {
      double acc = 0.0;
#pragma unroll
      for (int n = 0; n <= NU; ++n) acc += P.J[n] * t2[((2 * cx + n) * kRB + cy) * kRB + cz];
      const long id = ((long)x * c.y + y) * c.z + z;
      qc[id] = acc;
      if (qcb) qcb[id] = bf16_split(acc);
    }
  }
}

// ---------------------------------------------------------------- top level
// FFTg2g (forcefield1.c:499-528): separable DFT along z, y, x (sign -1),
// pointwise product with khat, inverse (sign +1), real part / M^3.
// tw holds cos/sin(2 pi j / M_axis) for the three axes back to back.
// One workgroup of 256 threads with LDS sized to the grid (32 bytes per
// point), so that it finds room beside the concurrently running pair kernel.
constexpr int kTopBS = 256;
constexpr int kTopSmall = 2048;  // largest M^3 done in one workgroup (64 KB of LDS)

__global__ __launch_bounds__(kTopBS) void k_top_small(DevParams P, GridDims g,
                                                      const double *__restrict__ qL,
                                                      const double *__restrict__ khat,
                                                      const double *__restrict__ tw,
                                                      double *__restrict__ eL, int asis) {
  chain_issue_priority();
  extern __shared__ double top_lds[];  // 4 n doubles: two complex buffers; the twiddles
  long n = g.size();
  double *re = top_lds, *im = top_lds + n, *re2 = top_lds + 2 * n, *im2 = top_lds + 3 * n;
  double *const twl = top_lds + 4 * n;  // read once per DFT term: kept in LDS
  for (int k = threadIdx.x; k < 2 * (g.x + g.y + g.z); k += kTopBS) twl[k] = tw[k];
  for (long e = threadIdx.x; e < n; e += kTopBS) {
    int z = e % g.z, y = (e / g.z) % g.y, x = e / ((long)g.y * g.z);
    re[e] = top_input(g, qL, x, y, z, asis, P.nu);
    im[e] = 0.0;
  }
  __syncthreads();
  dft_pass_lds<kTopBS>(g, 2, -1.0, re, im, re2, im2, twl);
  dft_pass_lds<kTopBS>(g, 1, -1.0, re2, im2, re, im, twl);
  dft_pass_lds<kTopBS>(g, 0, -1.0, re, im, re2, im2, twl);
  for (long e = threadIdx.x; e < n; e += kTopBS) re2[e] *= khat[e], im2[e] *= khat[e];
  __syncthreads();
  dft_pass_lds<kTopBS>(g, 0, 1.0, re2, im2, re, im, twl);
  dft_pass_lds<kTopBS>(g, 1, 1.0, re, im, re2, im2, twl);
  dft_pass_lds<kTopBS>(g, 2, 1.0, re2, im2, re, im, twl);
  for (long e = threadIdx.x; e < n; e += kTopBS) eL[e] = re[e] / (double)n;
}

// Large top grids (L = 1 with a fine grid): one launch per pass, global memory.
__global__ __launch_bounds__(256) void k_top_pack(DevParams P, GridDims g,
                                                  const double *__restrict__ qL,
                                                  double *__restrict__ re,
                                                  double *__restrict__ im, int asis) {
  chain_issue_priority();
  long e = (long)blockIdx.x * 256 + threadIdx.x;
  if (e >= g.size()) return;
  int z = e % g.z, y = (e / g.z) % g.y, x = e / ((long)g.y * g.z);
  re[e] = top_input(g, qL, x, y, z, asis, P.nu);
  im[e] = 0.0;
}

__global__ __launch_bounds__(256) void k_dft_axis(GridDims g, int axis, double sign,
                                                  const double *__restrict__ ire,
                                                  const double *__restrict__ iim,
                                                  double *__restrict__ ore,
                                                  double *__restrict__ oim,
                                                  const double *__restrict__ tw,
                                                  const double *__restrict__ mul) {
  chain_issue_priority();
  long e = (long)blockIdx.x * 256 + threadIdx.x;
  if (e >= g.size()) return;
  int M = axis == 0 ? g.x : axis == 1 ? g.y : g.z;
  const double *c = tw + 2 * (axis == 0 ? 0 : axis == 1 ? g.x : g.x + g.y);
  long stride = axis == 0 ? (long)g.y * g.z : axis == 1 ? g.z : 1;
  int k = axis == 0 ? e / ((long)g.y * g.z) : axis == 1 ? (e / g.z) % g.y : e % g.z;
  long base = e - k * stride;
  double sr = 0, si = 0;
  for (int m = 0; m < M; ++m) {
    int km = (int)(((long)k * m) % M);
    double cr = c[2 * km], ci = sign * c[2 * km + 1];
    double xr = ire[base + m * stride], xi = iim[base + m * stride];
    sr += xr * cr - xi * ci;
    si += xr * ci + xi * cr;
  }
  if (mul) sr *= mul[e], si *= mul[e];
  ore[e] = sr;
  oim[e] = si;
}

__global__ __launch_bounds__(256) void k_top_out(long n, const double *__restrict__ re,
                                                 double *__restrict__ eL) {
  chain_issue_priority();
  long e = (long)blockIdx.x * 256 + threadIdx.x;
  if (e < n) eL[e] = re[e] / (double)n;
}

// Grids above kTopSmall points whose axes factor into 2, 3, 4, 5 and 7 (every
// MSM top grid; PME grids in practice): per axis, each workgroup loads B
// whole lines into LDS and transforms them with a mixed-radix Stockham FFT
// (natural order in and out).  Stage s of radix R, with Ns the product of
// the radices before it: butterfly j of a line reads x[j + r M/R],
// multiplies by W^(r (j mod Ns) M/(Ns R)), takes the R-point DFT and writes
// y[(j - j mod Ns) R + j mod Ns + q Ns]; W^k = exp(i sign 2 pi k / M) comes
// from the axis' table of M roots.  A workgroup reads and writes only its
// own lines, so the passes run in place.  Five passes: forward z (reading
// the real grid), forward y, forward x + product with khat + inverse x in
// one LDS tile, inverse y, inverse z (writing Re / M^3).
constexpr int kFftBS = 256;
constexpr int kFftLds = 65536;  // LDS per workgroup: roots + two line buffers

template <int R>
__device__ __forceinline__ void fft_stage(const double2 *__restrict__ X, double2 *__restrict__ Y,
                                          const double2 *__restrict__ W, int M, int nb, int Ns,
                                          double sign) {
  const int mr = M / R, span = M / (Ns * R);
  double2 root[R];  // exp(i sign 2 pi k / R)
#pragma unroll
  for (int k = 0; k < R; ++k) root[k] = make_double2(W[k * mr].x, sign * W[k * mr].y);
  for (int t = threadIdx.x; t < nb * mr; t += kFftBS) {
    const int b = t / mr, j = t - b * mr, jm = j % Ns;
    const double2 *x = X + b * M;
    double2 v[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const double2 a = x[j + r * mr];
      const double2 w = W[r * jm * span];  // r jm span < R Ns span = M
      const double ws = sign * w.y;
      v[r] = make_double2(a.x * w.x - a.y * ws, a.x * ws + a.y * w.x);
    }
    double2 *y = Y + b * M + (j - jm) * R + jm;
#pragma unroll
    for (int q = 0; q < R; ++q) {
      double sr = 0.0, si = 0.0;
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const double2 w = root[(r * q) % R];
        sr += v[r].x * w.x - v[r].y * w.y;
        si += v[r].x * w.y + v[r].y * w.x;
      }
      y[q * Ns] = make_double2(sr, si);
    }
  }
}

// All stages of the nb lines in X (result left in X; Y is scratch).
__device__ void fft_lines(double2 *&X, double2 *&Y, const double2 *W, int M, int nb,
                          const FftAxis &A, double sign) {
  int Ns = 1;
  for (int s = 0; s < A.nst; ++s) {
    const int R = A.rad[s];
    switch (R) {
      case 2: fft_stage<2>(X, Y, W, M, nb, Ns, sign); break;
      case 3: fft_stage<3>(X, Y, W, M, nb, Ns, sign); break;
      case 4: fft_stage<4>(X, Y, W, M, nb, Ns, sign); break;
      case 5: fft_stage<5>(X, Y, W, M, nb, Ns, sign); break;
      default: fft_stage<7>(X, Y, W, M, nb, Ns, sign); break;
    }
    __syncthreads();
    double2 *t = X;
    X = Y;
    Y = t;
    Ns *= R;
  }
}

// Grid index of element k of line l along an axis with `inner` points after it.
__device__ __forceinline__ long fft_index(long l, int k, int M, long inner) {
  return (l / inner) * (M * inner) + (l % inner) + (long)k * inner;
}

// MODE 0: forward z from the real grid qL; 1: one direction of y (sign);
// 2: forward x, times khat, inverse x; 3: inverse z into eL (real part / n).
template <int MODE>
__global__ __launch_bounds__(kFftBS) void k_fft_pass(DevParams P, GridDims g, FftAxis A, int axis,
                                                     double sign, double2 *__restrict__ buf,
                                                     const double *__restrict__ tw,
                                                     const double *__restrict__ khat,
                                                     const double *__restrict__ qL,
                                                     double *__restrict__ eL, int asis) {
  chain_issue_priority();
  extern __shared__ double2 fft_lds[];
  const int M = A.M, B = A.B;
  const long inner = axis == 2 ? 1 : axis == 1 ? (long)g.z : (long)g.y * g.z;
  const long lines = g.size() / M;
  const long l0 = (long)blockIdx.x * B;
  const int nb = (int)min((long)B, lines - l0);
  double2 *W = fft_lds, *X = fft_lds + M, *Y = X + B * M;
  const double *c = tw + 2 * (axis == 0 ? 0 : axis == 1 ? g.x : g.x + g.y);
  for (int k = threadIdx.x; k < M; k += kFftBS) W[k] = make_double2(c[2 * k], c[2 * k + 1]);
  // consecutive threads take consecutive addresses: along the line where it
  // is contiguous (z), across lines otherwise
  for (int e = threadIdx.x; e < nb * M; e += kFftBS) {
    const int b = inner == 1 ? e / M : e % nb, k = inner == 1 ? e % M : e / nb;
    const long l = l0 + b;
    double2 v;
    if constexpr (MODE == 0)
      v = make_double2(top_input(g, qL, (int)(l / g.y), (int)(l % g.y), k, asis, P.nu), 0.0);
    else
      v = buf[fft_index(l, k, M, inner)];
    X[b * M + k] = v;
  }
  __syncthreads();
  fft_lines(X, Y, W, M, nb, A, sign);
  if constexpr (MODE == 2) {
    for (int e = threadIdx.x; e < nb * M; e += kFftBS) {
      const int b = e / M, k = e % M;
      const double kh = khat[fft_index(l0 + b, k, M, inner)];
      X[e] = make_double2(X[e].x * kh, X[e].y * kh);
    }
    __syncthreads();
    fft_lines(X, Y, W, M, nb, A, -sign);
  }
  const double inv_n = 1.0 / (double)g.size();
  for (int e = threadIdx.x; e < nb * M; e += kFftBS) {
    const int b = inner == 1 ? e / M : e % nb, k = inner == 1 ? e % M : e / nb;
    const long idx = fft_index(l0 + b, k, M, inner);
    if constexpr (MODE == 3)
      eL[idx] = X[b * M + k].x * inv_n;
    else
      buf[idx] = X[b * M + k];
  }
}

// ---------------------------------------------------------------- levels
// e^l = P e^{l+1} + K_l * q^l.  The stencil pass is grid2grid
// (forcefield0.c:898-931): e_m = sum_k khat_k q_{m+k-s/2}, periodic.
// One wavefront computes a (LY x LZ*ZB) tile of one x-plane: for each of
// the s_x input planes it stages a (LY+s_y-1) x (LZ*ZB+s_z-1) slab in LDS,
// and every lane slides ZB outputs along z through register windows of
// KC taps.  Prolongation (forcefield1.c:530-567) is gathered per output:
// e_{2m+n} += J_n e^{l+1}_m.
constexpr int kZB = 8;
constexpr int kKC = 8;

template <int NU>
__device__ __forceinline__ double prolong_at(const double *J, GridDims gc, const double *ec,
                                             int x, int y, int z) {
  // fine index v receives J_n from coarse m with 2m + n == v (mod 2 M_c),
  // i.e. n has the parity of v and m = (v - n)/2 mod M_c
  const int H = NU / 2;
  double acc = 0.0;
  for (int nx = -H; nx <= H; ++nx) {
    if (((x - nx) & 1) != 0) continue;
    int mx = wrapi((x - nx) / 2, gc.x);
    double sy = 0.0;
    for (int ny = -H; ny <= H; ++ny) {
      if (((y - ny) & 1) != 0) continue;
      int my = wrapi((y - ny) / 2, gc.y);
      const double *row = ec + ((long)mx * gc.y + my) * gc.z;
      double sz = 0.0;
      for (int nz = -H; nz <= H; ++nz) {
        if (((z - nz) & 1) != 0) continue;
        sz += J[nz + H] * row[wrapi((z - nz) / 2, gc.z)];
      }
      sy += J[ny + H] * sz;
    }
    acc += J[nx + H] * sy;
  }
  return acc;
}

template <int NU>
__global__ __launch_bounds__(64) void k_level64(DevParams P, GridDims g, GridDims gc, GridDims sd,
                                                int LZ, int LY, const double *__restrict__ khat,
                                                const double *__restrict__ ql,
                                                const double *__restrict__ ec,
                                                double *__restrict__ el) {
  extern __shared__ double slab[];
  int tid = threadIdx.x;
  int lz = tid % LZ, ly = tid / LZ;
  int tz = LZ * kZB;
  int ztiles = (g.z + tz - 1) / tz, ytiles = (g.y + LY - 1) / LY;
  int b = blockIdx.x;
  int zt = b % ztiles;
  b /= ztiles;
  int yt = b % ytiles;
  int x = b / ytiles;
  int z0 = zt * tz, y0 = yt * LY;
  int rows = LY + sd.y - 1, pitch = tz + sd.z - 1 + kKC;  // +KC: window over-read pad
  double acc[kZB];
#pragma unroll
  for (int t = 0; t < kZB; ++t) acc[t] = 0.0;
  for (int kx = 0; kx < sd.x; ++kx) {
    int xi = wrapi(x + kx - sd.x / 2, g.x);
    const double *plane = ql + (long)xi * g.y * g.z;
    for (int e = tid; e < rows * pitch; e += 64) {
      int r = e / pitch, cc = e % pitch;
      int yy = wrapi(y0 + r - sd.y / 2, g.y), zz = wrapi(z0 + cc - sd.z / 2, g.z);
      slab[e] = plane[(long)yy * g.z + zz];
    }
    __syncthreads();
    if (ly < LY) {
      const double *kp = khat + (long)kx * sd.y * sd.z;
      for (int ky = 0; ky < sd.y; ++ky) {
        const double *row = slab + (ly + ky) * pitch + lz * kZB;
        const double *kr = kp + ky * sd.z;
        for (int k0 = 0; k0 < sd.z; k0 += kKC) {
          double w[kZB + kKC - 1];
#pragma unroll
          for (int t = 0; t < kZB + kKC - 1; ++t) w[t] = row[k0 + t];
#pragma unroll
          for (int k = 0; k < kKC; ++k) {
            double kv = (k0 + k < sd.z) ? kr[k0 + k] : 0.0;
#pragma unroll
            for (int t = 0; t < kZB; ++t) acc[t] += kv * w[t + k];
          }
        }
      }
    }
    __syncthreads();
  }
  if (ly < LY) {
    int y = y0 + ly;
    if (y < g.y) {
#pragma unroll
      for (int t = 0; t < kZB; ++t) {
        int z = z0 + lz * kZB + t;
        if (z < g.z) {
          double pv = ec ? prolong_at<NU>(P.J, gc, ec, x, y, z) : 0.0;
          el[((long)x * g.y + y) * g.z + z] = pv + acc[t];
        }
      }
    }
  }
}

// ---------------------------------------------------------------- energy
constexpr int kGridBS = 256;

__global__ __launch_bounds__(kGridBS) void k_grid_energy(DevParams P, GridDims g,
                                                         const double *__restrict__ q1,
                                                         const double *__restrict__ e1,
                                                         double *__restrict__ epart, int asis) {
  chain_issue_priority();
  long n = g.size();
  double s = 0.0;
  for (long e = (long)blockIdx.x * kGridBS + threadIdx.x; e < n; e += (long)gridDim.x * kGridBS) {
    double qv;
    if (!asis) {
      qv = q1[e];
    } else {
      // reference HEAD at L = 1 (forcefield1.c:196-212): the padded array
      // (dims M+nu-1, interior at nu/2-1, zero ghosts) read at flat index e
      int nu = P.nu, o = nu / 2 - 1;
      long py = g.y + nu - 1, pz = g.z + nu - 1;
      long a = e / (py * pz), b = (e / pz) % py, c = e % pz;
      bool in = a >= o && a < o + g.x && b >= o && b < o + g.y && c >= o && c < o + g.z;
      qv = in ? q1[((a - o) * g.y + (b - o)) * g.z + (c - o)] : 0.0;
    }
    s += qv * e1[e];
  }
  double t = block_sum<kGridBS>(s);
  if (threadIdx.x == 0) epart[blockIdx.x] = t;
}

__device__ __forceinline__ void neumaier(double &s, double &c, double v) {
  double t = s + v;
  if (fabs(s) >= fabs(v)) c += (s - t) + v;
  else c += (v - t) + s;
  s = t;
}

// Adds two compensated partial sums (s, c): error-free sum of the heads,
// carries added.
__device__ __forceinline__ void csum_add(double &s, double &c, double s2, double c2) {
  double t = s + s2, bb = t - s;
  double err = (s - (t - bb)) + (s2 - bb);
  s = t;
  c += c2 + err;
}

// Compensated sums of the pair and the grid partials over the threads of a
// BS-thread workgroup: thread t takes p[t], p[t + BS], ... of each in that
// order.  The partials sit in memory written by other kernels, so every
// load of a round (UP pair and UG grid partials per thread) is issued
// before any is summed: one memory latency per round instead of one per
// value.
template <int BS, int UP, int UG>
__device__ __forceinline__ void strided_csum2(const double *__restrict__ ep, int np,
                                              const double *__restrict__ eg, int ng, double &sp,
                                              double &cp, double &sg, double &cg) {
  const int t = threadIdx.x;
  for (int r = 0; t + r * UP * BS < np || t + r * UG * BS < ng; ++r) {  // rounds
    const int kp = t + r * UP * BS, kg = t + r * UG * BS;
    double v[UP], w[UG];
#pragma unroll
    for (int u = 0; u < UP; ++u) v[u] = kp + u * BS < np ? ep[kp + u * BS] : 0.0;
#pragma unroll
    for (int u = 0; u < UG; ++u) w[u] = kg + u * BS < ng ? eg[kg + u * BS] : 0.0;
#pragma unroll
    for (int u = 0; u < UP; ++u) neumaier(sp, cp, v[u]);  // adding 0.0 changes neither
#pragma unroll
    for (int u = 0; u < UG; ++u) neumaier(sg, cg, w[u]);
  }
}

// One step of a compensated wavefront scan: (s, c) combined with the pair
// DPP control CTRL brings (rows ROWS only; elsewhere it adds (0, 0), which
// changes neither).
template <int CTRL, int ROWS>
__device__ __forceinline__ void csum_step(double &s1, double &c1, double &s2, double &c2) {
  const double a = dpp_take<CTRL, ROWS>(s1), b = dpp_take<CTRL, ROWS>(c1);
  const double d = dpp_take<CTRL, ROWS>(s2), e = dpp_take<CTRL, ROWS>(c2);
  csum_add(s1, c1, a, b);
  csum_add(s2, c2, d, e);
}

// Two compensated sums (s1, c1), (s2, c2) over a workgroup at once: an
// inclusive scan of each over the wavefront by DPP in a fixed order (lane
// 63 ends with the wavefront total), then the wavefront totals in order
// (one barrier); thread 0 receives the totals.
template <int BS>
__device__ __forceinline__ void block_csum2(double &s1, double &c1, double &s2, double &c2) {
  constexpr int NW = BS / 64;
  __shared__ double ws[4][NW];
  csum_step<0x111, 0xF>(s1, c1, s2, c2);  // row_shr:1
  csum_step<0x112, 0xF>(s1, c1, s2, c2);  // row_shr:2
  csum_step<0x114, 0xF>(s1, c1, s2, c2);  // row_shr:4
  csum_step<0x118, 0xF>(s1, c1, s2, c2);  // row_shr:8
  csum_step<0x142, 0xA>(s1, c1, s2, c2);  // row_bcast:15
  csum_step<0x143, 0xC>(s1, c1, s2, c2);  // row_bcast:31
  const int t = threadIdx.x;
  if ((t & 63) == 63) ws[0][t >> 6] = s1, ws[1][t >> 6] = c1, ws[2][t >> 6] = s2, ws[3][t >> 6] = c2;
  __syncthreads();
  if (t == 0) {
    s1 = ws[0][0], c1 = ws[1][0], s2 = ws[2][0], c2 = ws[3][0];
    for (int w = 1; w < NW; ++w) csum_add(s1, c1, ws[0][w], ws[1][w]), csum_add(s2, c2, ws[2][w], ws[3][w]);
  }
}

// The energy of an evaluation by one workgroup: compensated sums of the
// per-cell pair partials and the per-block grid partials, each thread over
// a fixed strided subset, then the wavefronts and the wavefront totals in a
// fixed order (the result does not depend on scheduling), and the constant
// terms.  out = {E, constant part, pair part, grid part, flags}; the
// non-finite-input counter is copied to out[4] and cleared for the next
// evaluation.  One workgroup, so no other workgroup's results need to be
// made visible to it (a last-workgroup-done reduction needs a device-scope
// release per workgroup, which on this GPU writes back and invalidates L2).
template <int BS>
__device__ void energy_block(const DevParams &P, const FinalizeArgs &f) {
  const int t = threadIdx.x;
  double qs = 0.0, q2 = 0.0;
  int flags = 0;
  if (t == 0) qs = f.scal[0], q2 = f.scal[1], flags = *f.flags;  // in flight with the rest
  double sv = 0, cv = 0, s = 0, c = 0;
  strided_csum2<BS, 32, 4>(f.ep, f.np, f.eg, f.ng, sv, cv, s, c);
  block_csum2<BS>(sv, cv, s, c);
  if (t != 0) return;
  double epair = sv + cv, egrid = 0.5 * (s + c);
  // constant part (forcefield1.c:114-130)
  double econst = 0.5 * q2 * P.coeff1 - 0.5 * (qs * qs) * P.coeff2 - 0.5 * q2 * P.gamma0 * P.inv_a0;
  s = 0, c = 0;
  neumaier(s, c, econst);
  neumaier(s, c, epair);
  neumaier(s, c, egrid);
  f.out[0] = s + c;
  f.out[1] = econst;
  f.out[2] = epair;
  f.out[3] = egrid;
  f.out[4] = (double)flags;
  if (f.user) {
    f.user[0] = f.out[0];
    f.user[1] = econst;
    f.user[2] = epair;
    f.user[3] = egrid;
  }
  *f.flags = 0;
}

constexpr int kFinBS = 1024;
__global__ __launch_bounds__(kFinBS) void k_finalize(DevParams P, FinalizeArgs f) {
  chain_issue_priority();
  energy_block<kFinBS>(P, f);
}

// interpolate (forcefield1.c:570-607) + F += q E (:228-231), and the
// potential phi_i = dE/dq_i: grid part plus the constant self terms.
template <int NU, bool ADD, bool TRI>
__device__ __forceinline__ void gather_atom(const DevParams &P, GridDims g,
                                            const double4 *__restrict__ spos,
                                            const int *__restrict__ sidx,
                                            const double *__restrict__ e1,
                                            const double *__restrict__ scal,
                                            double *__restrict__ F, double *__restrict__ phi,
                                            int i) {
  double4 a = spos[i];
  double r[3] = {a.x, a.y, a.z};
  int dims[3] = {g.x, g.y, g.z};
  double t[3], sf[3];
  int m[3];
  to_frac<TRI>(P, r, sf);
#pragma unroll
  for (int ax = 0; ax < 3; ++ax) {
    double tt = dims[ax] * sf[ax], mm = floor(tt);
    t[ax] = tt - mm;
    m[ax] = (int)mm;
  }
  double wz[NU], dwz[NU];
  int iz[NU];
  bspline<NU, true>(P.Q, t[2], wz, dwz);
#pragma unroll
  for (int k = 0; k < NU; ++k) iz[k] = wrapi(m[2] - k + NU / 2, g.z);
  double ph = 0, ex = 0, ey = 0, ez = 0;
#pragma unroll 1
  for (int x = 0; x < NU; ++x) {
    double dwx;
    double wx = bspline_one<NU, true>(P.Q, x, t[0], &dwx);
    long xo = (long)wrapi(m[0] - x + NU / 2, g.x) * g.y;
    double px = 0, py = 0, pz = 0;
#pragma unroll 1
    for (int y = 0; y < NU; ++y) {
      double dwy;
      double wy = bspline_one<NU, true>(P.Q, y, t[1], &dwy);
      const double *row = e1 + (xo + wrapi(m[1] - y + NU / 2, g.y)) * g.z;
      double s0 = 0, s1 = 0;
#pragma unroll
      for (int z = 0; z < NU; ++z) {
        double v = row[iz[z]];
        s0 += wz[z] * v;
        s1 += dwz[z] * v;
      }
      px += wy * s0;
      py += dwy * s0;
      pz += wy * s1;
    }
    ph += wx * px;
    ex += dwx * px;
    ey += wx * py;
    ez += wx * pz;
  }
  double Ex, Ey, Ez;  // E = -A^{-T} (M * grad_t)
  frac_field<TRI>(P, g, ex, ey, ez, Ex, Ey, Ez);
  int oi = sidx[i];
  const double pv = ph + a.w * P.coeff_self - scal[0] * P.coeff2;
  if (F) {
    double *f = F + 3 * (long)oi;
    f[0] = (ADD ? f[0] : 0.0) + a.w * Ex;
    f[1] = (ADD ? f[1] : 0.0) + a.w * Ey;
    f[2] = (ADD ? f[2] : 0.0) + a.w * Ez;
  }
  if (phi) phi[oi] = (ADD ? phi[oi] : 0.0) + pv;
}

template <int NU, bool ADD, bool TRI>
__global__ __launch_bounds__(256) void k_gather(DevParams P, GridDims g,
                                                const double4 *__restrict__ spos,
                                                const int *__restrict__ sidx,
                                                const double *__restrict__ e1,
                                                const double *__restrict__ scal,
                                                double *__restrict__ F, double *__restrict__ phi) {
  chain_issue_priority();
  int i = blockIdx.x * 256 + threadIdx.x;
  if (i < P.N) gather_atom<NU, ADD, TRI>(P, g, spos, sidx, e1, scal, F, phi, i);
}

// The same interpolation from an LDS copy of the level-1 grid around a tile
// of pair-search cells.  A workgroup takes tb x tb x tb cells, copies the
// grid points their atoms can reach -- the box [o, o + R) per axis, wrapped
// periodically -- into LDS (coalesced along z), then its threads take the
// tile's atoms (column by column, as sorted) and read their NU^3 grid values
// from LDS instead of global memory.
constexpr int kTileMax = 8;

struct TileGeom {
  int tb;         // cells per tile edge; 0: tiles do not fit, use k_gather
  int ntile[3];   // tiles per axis
  int R[3];       // LDS region extent per axis (grid points)
};

TileGeom gather_tiles(const DevParams &P, GridDims g) {
  TileGeom tg{};
  const int dims[3] = {g.x, g.y, g.z};
  const double per_cell = (double)P.N / P.ncell;
  for (int tb = kTileMax; tb >= 1; --tb) {
    int R[3];
    long vol = 1;
    for (int a = 0; a < 3; ++a) {
      R[a] = (P.tri ? (int)std::ceil((double)tb * dims[a] / P.nc[a])
                    : (int)std::ceil(tb * P.cw[a] * dims[a] / P.box[a])) + P.nu + 4;
      vol *= R[a];
    }
    if (vol * (long)sizeof(double) > 48 * 1024 || R[2] > 64) continue;
    if (tb > 1 && tb * tb * tb * per_cell > 768.0) continue;
    tg.tb = tb;
    for (int a = 0; a < 3; ++a) {
      tg.R[a] = R[a];
      tg.ntile[a] = (P.nc[a] + tb - 1) / tb;
    }
    return tg;
  }
  return tg;
}

template <int NU, bool ADD, bool TRI>
__device__ __forceinline__ void gather_atom_tile(const DevParams &P, GridDims g,
                                                 const double *reg, int Ry, int Rz,
                                                 const int (&o)[3],
                                                 const double4 *__restrict__ spos,
                                                 const int *__restrict__ sidx,
                                                 const double *__restrict__ scal,
                                                 double *__restrict__ F, double *__restrict__ phi,
                                                 int i) {
  const double4 a = spos[i];
  const double r[3] = {a.x, a.y, a.z};
  const int dims[3] = {g.x, g.y, g.z};
  double t[3], sf[3];
  int m[3];
  to_frac<TRI>(P, r, sf);
#pragma unroll
  for (int ax = 0; ax < 3; ++ax) {
    const double tt = dims[ax] * sf[ax], mm = floor(tt);
    t[ax] = tt - mm;
    m[ax] = (int)mm - o[ax] + NU / 2;  // local index of weight 0's grid point
  }
  // all weights up front (register arrays, fully unrolled loops)
  double wx[NU], dwx[NU], wy[NU], dwy[NU], wz[NU], dwz[NU];
  bspline<NU, true>(P.Q, t[0], wx, dwx);
  bspline<NU, true>(P.Q, t[1], wy, dwy);
  bspline<NU, true>(P.Q, t[2], wz, dwz);
  double ph = 0, ex = 0, ey = 0, ez = 0;
#pragma unroll 1
  for (int x = 0; x < NU; ++x) {
    const int xo = (m[0] - x) * Ry;
    double px = 0, py = 0, pz = 0;
#pragma unroll
    for (int y = 0; y < NU; ++y) {
      const double *row = reg + (xo + m[1] - y) * Rz + m[2];
      double s0 = 0, s1 = 0;
#pragma unroll
      for (int z = 0; z < NU; ++z) {
        const double v = row[-z];
        s0 += wz[z] * v;
        s1 += dwz[z] * v;
      }
      px += wy[y] * s0;
      py += dwy[y] * s0;
      pz += wy[y] * s1;
    }
    ph += wx[0] * px;
    ex += dwx[0] * px;
    ey += wx[0] * py;
    ez += wx[0] * pz;
#pragma unroll
    for (int k = 0; k + 1 < NU; ++k) wx[k] = wx[k + 1], dwx[k] = dwx[k + 1];  // next x weight first
  }
  double Ex, Ey, Ez;  // E = -A^{-T} (M * grad_t)
  frac_field<TRI>(P, g, ex, ey, ez, Ex, Ey, Ez);
  const int oi = sidx[i];
  const double pv = ph + a.w * P.coeff_self - scal[0] * P.coeff2;
  if (F) {
    double *f = F + 3 * (long)oi;
    f[0] = (ADD ? f[0] : 0.0) + a.w * Ex;
    f[1] = (ADD ? f[1] : 0.0) + a.w * Ey;
    f[2] = (ADD ? f[2] : 0.0) + a.w * Ez;
  }
  if (phi) phi[oi] = (ADD ? phi[oi] : 0.0) + pv;
}

template <int NU, bool ADD, bool TRI>
__global__ __launch_bounds__(256) void k_gather_tile(DevParams P, GridDims g, TileGeom tg,
                                                     const double4 *__restrict__ spos,
                                                     const int *__restrict__ sidx,
                                                     const int *__restrict__ start,
                                                     const double *__restrict__ e1,
                                                     const double *__restrict__ scal,
                                                     double *__restrict__ F,
                                                     double *__restrict__ phi) {
  chain_issue_priority();
  extern __shared__ double reg[];
  __shared__ int seg_off[kTileMax * kTileMax + 1], seg_beg[kTileMax * kTileMax];
  const int b = blockIdx.x, t = threadIdx.x;
  const int tz = b % tg.ntile[2], ty = (b / tg.ntile[2]) % tg.ntile[1],
            tx = b / (tg.ntile[1] * tg.ntile[2]);
  const int c0[3] = {tx * tg.tb, ty * tg.tb, tz * tg.tb};
  const int c1[3] = {min(P.nc[0], c0[0] + tg.tb), min(P.nc[1], c0[1] + tg.tb),
                     min(P.nc[2], c0[2] + tg.tb)};
  const int dims[3] = {g.x, g.y, g.z};
  int o[3];
#pragma unroll
  for (int a = 0; a < 3; ++a)
    o[a] = (TRI ? (int)floor((double)c0[a] * dims[a] / P.nc[a])
                : (int)floor(c0[a] * P.cw[a] * (dims[a] * P.ibox[a]))) - NU / 2 - 1;
  const int Ry = tg.R[1], Rz = tg.R[2], nreg = tg.R[0] * Ry * Rz;
  {  // eight loads in flight per thread (unconditional: clamped index); the
     // divisions and periodic wraps by FP32 reciprocals, exact for these
     // small integers
    const float iRz = 1.0f / Rz, iRy = 1.0f / Ry;
    const float igx = 1.0f / g.x, igy = 1.0f / g.y, igz = 1.0f / g.z;
    auto wrapf = [](int v, int m, float im) { return v - m * (int)floorf((v + 0.5f) * im); };
    for (int k0 = t; k0 < nreg; k0 += 256 * 8) {
      double v[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int k = min(k0 + 256 * u, nreg - 1);
        const int q = (int)((k + 0.5f) * iRz), lz = k - q * Rz;
        const int lx = (int)((q + 0.5f) * iRy), ly = q - lx * Ry;
        v[u] = e1[((long)wrapf(o[0] + lx, g.x, igx) * g.y + wrapf(o[1] + ly, g.y, igy)) * g.z +
                  wrapf(o[2] + lz, g.z, igz)];
      }
#pragma unroll
      for (int u = 0; u < 8; ++u)
        if (k0 + 256 * u < nreg) reg[k0 + 256 * u] = v[u];
    }
  }
  // the tile's cell columns (cx, cy, c0z .. c1z - 1) are contiguous atom ranges
  const int ncy = c1[1] - c0[1], ncol = (c1[0] - c0[0]) * ncy;
  if (t < 64) {
    int beg = 0, cnt = 0;
    if (t < ncol) {
      const int base = ((c0[0] + t / ncy) * P.nc[1] + c0[1] + t % ncy) * P.nc[2];
      beg = start[base + c0[2]];
      cnt = start[base + c1[2]] - beg;
    }
    const int incl = wave_incl_scan(cnt);
    if (t < ncol) {
      seg_off[t] = incl - cnt;
      seg_beg[t] = beg;
    }
    if (t == ncol - 1) seg_off[ncol] = incl;
  }
  __syncthreads();
  const int total = seg_off[ncol];
  int sl = 0;
  for (int k = t; k < total; k += 256) {
    while (seg_off[sl + 1] <= k) ++sl;
    gather_atom_tile<NU, ADD, TRI>(P, g, reg, Ry, Rz, o, spos, sidx, scal, F, phi,
                              seg_beg[sl] + k - seg_off[sl]);
  }
}

// F += Fp and phi += phip (the short-range parts, written by the pair kernel
// beside the interpolation) by workgroups 1 ..; workgroup 0 reduces the
// energy meanwhile (energy_block).
__global__ __launch_bounds__(kFinBS) void k_combine(DevParams P, const double *__restrict__ Fp,
                                                    const double *__restrict__ phip,
                                                    double *__restrict__ F,
                                                    double *__restrict__ phi, FinalizeArgs fin) {
  chain_issue_priority();
  if (blockIdx.x == 0) {
    energy_block<kFinBS>(P, fin);
    return;
  }
  const long i = (long)(blockIdx.x - 1) * kFinBS + threadIdx.x;
  if (i < P.N) {
    if (F)
      for (int d = 0; d < 3; ++d) F[3 * i + d] += Fp[3 * i + d];
    if (phi) phi[i] += phip[i];
  }
}

inline unsigned nblk(long n, int bs) { return (unsigned)((n + bs - 1) / bs); }

}  // namespace

// ---------------------------------------------------------------- launchers
void launch_fill_int(int *p, long n, int v, hipStream_t s) {
  if (n > 0) k_fill_int<<<nblk(n, 256), 256, 0, s>>>(p, n, v);
}
void launch_fill_dbl(double *p, long n, double v, hipStream_t s) {
  if (n > 0) k_fill_dbl<<<nblk(n, 256), 256, 0, s>>>(p, n, v);
}

int bin_blocks(int N) { return (int)nblk(N, kBinBS); }

void launch_bin(const DevParams &P, const double *pos, const double *q, int *cell_of, int *bucket,
                int cap, int *ovfn, int *ovf, int *cnt, double *qpart, int *flags, hipStream_t s) {
  k_bin<<<bin_blocks(P.N), kBinBS, 0, s>>>(P, pos, q, cell_of, bucket, cap, ovfn, ovf, cnt, qpart,
                                           flags);
}

void launch_scan(const DevParams &P, int *cnt, int *start, const double *qpart, int nqpart,
                 double *scal, int *ovfn, hipStream_t s) {
  k_scan<<<1, kScanBS, 0, s>>>(P.ncell, cnt, start, qpart, nqpart, scal, ovfn);
}

void launch_place(const DevParams &P, const double *pos, const double *q, const int *cell_of,
                  const int *start, const int *bucket, int cap, const int *ovfn, const int *ovf,
                  double4 *spos, int *sidx, hipStream_t s, hipEvent_t done, GridDims g1,
                  unsigned *sbase, double4 *st) {
#define A P, pos, q, cell_of, start, bucket, cap, ovfn, ovf, spos, sidx, g1, sbase, st
  if (P.tri) launch_done(k_place<true>, nblk(P.N, 256), 256, 0, s, done, A);
  else launch_done(k_place<false>, nblk(P.N, 256), 256, 0, s, done, A);
#undef A
}

int pair_blocks(int N) { return (int)nblk(N, kPairBS); }

#define MSM_NU_SWITCH(nu, CALL) \
  switch (nu) {                 \
    case 4: CALL(4); break;     \
    case 6: CALL(6); break;     \
    case 8: CALL(8); break;     \
    default: CALL(10); break;   \
  }


void launch_pair_fp64(const DevParams &P, const double4 *spos, const int *sidx, const int *start,
                      const int4 *cols, int ncol, double *F, double *phi, double *epart,
                      hipStream_t s, hipEvent_t done) {
#define CALL(NU)                                                                             \
  if (P.tri)                                                                                 \
    launch_done(k_pair_cell<false, NU, true>, P.ncell, 64, 0, s, done, P, spos, sidx, start, \
                cols, ncol, F, phi, epart);                                                  \
  else                                                                                       \
    launch_done(k_pair_cell<false, NU, false>, P.ncell, 64, 0, s, done, P, spos, sidx, start, \
                cols, ncol, F, phi, epart)
  if (P.ewald) {
    CALL(0);
  } else {
    MSM_NU_SWITCH(P.nu, CALL)
  }
#undef CALL
}

void launch_pair_fp32(const DevParams &P, const double4 *spos, const int *sidx, const int *start,
                      const int4 *cols, int ncol, double *F, double *phi, double *epart,
                      hipStream_t s, hipEvent_t done) {
#define CALL(NU)                                                                            \
  if (P.tri)                                                                                \
    launch_done(k_pair_cell<true, NU, true>, P.ncell, 64, 0, s, done, P, spos, sidx, start, \
                cols, ncol, F, phi, epart);                                                 \
  else                                                                                      \
    launch_done(k_pair_cell<true, NU, false>, P.ncell, 64, 0, s, done, P, spos, sidx, start, \
                cols, ncol, F, phi, epart)
  if (P.ewald) {
    CALL(0);
  } else {
    MSM_NU_SWITCH(P.nu, CALL)
  }
#undef CALL
}

void launch_pair_count(const DevParams &P, const double4 *spos, const int *start, const int4 *cols,
                       int ncol, double r2max, unsigned *cell_count, hipStream_t s) {
  const double w = 2.0 * std::sqrt(r2max) * (1.0 + 1e-12);  // rounding of d / L stays below 1/2
  if (!P.tri && w < P.box[0] && w < P.box[1] && w < P.box[2])
    k_pair_count<false><<<P.ncell, 64, 0, s>>>(P, spos, start, cols, ncol, r2max, cell_count);
  else
    k_pair_count<true><<<P.ncell, 64, 0, s>>>(P, spos, start, cols, ncol, r2max, cell_count);
}

bool spread_by_placements(const DevParams &P, GridDims g) {
  return g.x >= kSX + P.nu && g.y >= kSY + P.nu && g.z >= kSZ + P.nu && g.x < 1024 &&
         g.y < 1024 && g.z < 1024;
}

void launch_spread(const DevParams &P, GridDims g, const double4 *spos, const int *start,
                   double *q1, hipStream_t s, unsigned *q1b, unsigned *sbase, double4 *st,
                   hipEvent_t done, bool placed) {
  unsigned nb = (unsigned)(((g.x + kSX - 1) / kSX) * ((g.y + kSY - 1) / kSY) * ((g.z + kSZ - 1) / kSZ));
  if (sbase && st && spread_by_placements(P, g)) {
    if (!placed) {
      if (P.tri)
        k_spread_place<true><<<nblk(P.N, 256), 256, 0, s>>>(P, g, spos, sbase, st);
      else
        k_spread_place<false><<<nblk(P.N, 256), 256, 0, s>>>(P, g, spos, sbase, st);
    }
    if (nb >= kSpreadFine) {
#define CALL(NU) launch_done(k_spread_bricks<NU, 1>, nb, 64, 0, s, done, P, g, sbase, st, start, q1, q1b)
      MSM_NU_SWITCH(P.nu, CALL)
#undef CALL
    } else {
#define CALL(NU) launch_done(k_spread_bricks<NU, 4>, nb, 256, 0, s, done, P, g, sbase, st, start, q1, q1b)
      MSM_NU_SWITCH(P.nu, CALL)
#undef CALL
    }
    return;
  }
  if (nb >= kSpreadFine) {
#define CALL(NU)                                                    \
  if (P.tri) launch_done(k_spread_mfma<NU, 1, true>, nb, 64, 0, s, done, P, g, spos, start, q1, q1b); \
  else launch_done(k_spread_mfma<NU, 1, false>, nb, 64, 0, s, done, P, g, spos, start, q1, q1b)
    MSM_NU_SWITCH(P.nu, CALL)
#undef CALL
  } else {
#define CALL(NU)                                                      \
  if (P.tri) launch_done(k_spread_mfma<NU, 4, true>, nb, 256, 0, s, done, P, g, spos, start, q1, q1b); \
  else launch_done(k_spread_mfma<NU, 4, false>, nb, 256, 0, s, done, P, g, spos, start, q1, q1b)
    MSM_NU_SWITCH(P.nu, CALL)
#undef CALL
  }
}

void launch_restrict(const DevParams &P, GridDims fine, GridDims coarse, const double *qf,
                     double *qc, hipStream_t s, unsigned *qcb) {
  const unsigned nb = (unsigned)(((coarse.x + kRB - 1) / kRB) * ((coarse.y + kRB - 1) / kRB) *
                                 ((coarse.z + kRB - 1) / kRB));
#define CALL(NU) k_restrict<NU><<<nb, kRBS, 0, s>>>(P, fine, coarse, qf, qc, qcb)
  MSM_NU_SWITCH(P.nu, CALL)
#undef CALL
}

// Radices of an FFT of length M (4 first, then 2, 3, 5, 7) and the lines per
// workgroup that fit kFftLds; false if M has another prime factor or one
// line does not fit.
bool fft_axis_plan(int M, FftAxis &A) {
  A.M = M;
  A.nst = 0;
  int m = M;
  for (int p : {4, 2, 3, 5, 7})
    while (m % p == 0 && A.nst < kFftMaxStages) {
      A.rad[A.nst++] = p;
      m /= p;
    }
  const int per_line = 2 * M * (int)sizeof(double2);
  A.B = std::min(32, (kFftLds - M * (int)sizeof(double2)) / per_line);
  return m == 1 && A.B >= 1;
}

void launch_top(const DevParams &P, GridDims g, const double *qL, const double *khat,
                const double *tw, double *work, double *eL, int asis, hipStream_t s) {
  long n = g.size();
  const long small_lds = (4 * n + 2 * (g.x + g.y + g.z)) * (long)sizeof(double);
  if (n <= kTopSmall && small_lds <= 65536) {
    k_top_small<<<1, kTopBS, (size_t)small_lds, s>>>(P, g, qL, khat, tw, eL, asis);
    return;
  }
  FftAxis ax[3];
  if (fft_axis_plan(g.x, ax[0]) && fft_axis_plan(g.y, ax[1]) && fft_axis_plan(g.z, ax[2])) {
    double2 *buf = reinterpret_cast<double2 *>(work);
    const int dims[3] = {g.x, g.y, g.z};
    auto grid = [&](int a) { return (unsigned)((n / dims[a] + ax[a].B - 1) / ax[a].B); };
    auto lds = [&](int a) { return (size_t)(dims[a] + 2 * ax[a].B * dims[a]) * sizeof(double2); };
    k_fft_pass<0><<<grid(2), kFftBS, lds(2), s>>>(P, g, ax[2], 2, -1.0, buf, tw, khat, qL, eL, asis);
    k_fft_pass<1><<<grid(1), kFftBS, lds(1), s>>>(P, g, ax[1], 1, -1.0, buf, tw, khat, qL, eL, asis);
    k_fft_pass<2><<<grid(0), kFftBS, lds(0), s>>>(P, g, ax[0], 0, -1.0, buf, tw, khat, qL, eL, asis);
    k_fft_pass<1><<<grid(1), kFftBS, lds(1), s>>>(P, g, ax[1], 1, 1.0, buf, tw, khat, qL, eL, asis);
    k_fft_pass<3><<<grid(2), kFftBS, lds(2), s>>>(P, g, ax[2], 2, 1.0, buf, tw, khat, qL, eL, asis);
    return;
  }
  double *r0 = work, *i0 = work + n, *r1 = work + 2 * n, *i1 = work + 3 * n;
  unsigned nb = nblk(n, 256);
  k_top_pack<<<nb, 256, 0, s>>>(P, g, qL, r0, i0, asis);
  k_dft_axis<<<nb, 256, 0, s>>>(g, 2, -1.0, r0, i0, r1, i1, tw, nullptr);
  k_dft_axis<<<nb, 256, 0, s>>>(g, 1, -1.0, r1, i1, r0, i0, tw, nullptr);
  k_dft_axis<<<nb, 256, 0, s>>>(g, 0, -1.0, r0, i0, r1, i1, tw, khat);
  k_dft_axis<<<nb, 256, 0, s>>>(g, 0, 1.0, r1, i1, r0, i0, tw, nullptr);
  k_dft_axis<<<nb, 256, 0, s>>>(g, 1, 1.0, r0, i0, r1, i1, tw, nullptr);
  k_dft_axis<<<nb, 256, 0, s>>>(g, 2, 1.0, r1, i1, r0, i0, tw, nullptr);
  k_top_out<<<nb, 256, 0, s>>>(n, r0, eL);
}

void launch_level_fp64(const DevParams &P, GridDims g, GridDims gc, GridDims sd,
                       const double *khat, const double *ql, const double *ec, double *el,
                       hipStream_t s) {
  int LZ = (g.z + kZB - 1) / kZB;
  if (LZ > 16) LZ = 16;
  int LY = 64 / LZ;
  if (LY > g.y) LY = g.y;
  int tz = LZ * kZB;
  int ztiles = (g.z + tz - 1) / tz, ytiles = (g.y + LY - 1) / LY;
  unsigned nb = (unsigned)(ztiles * ytiles * g.x);
  size_t lds = sizeof(double) * (LY + sd.y - 1) * (tz + sd.z - 1 + kKC);
#define CALL(NU) \
  k_level64<NU><<<nb, 64, lds, s>>>(P, g, gc, sd, LZ, LY, khat, ql, ec, el)
  MSM_NU_SWITCH(P.nu, CALL)
#undef CALL
}

int grid_blocks(long n) {
  long b = (n + kGridBS - 1) / kGridBS;
  return (int)(b < 1024 ? b : 1024);
}

void launch_grid_energy(const DevParams &P, GridDims g, const double *q1, const double *e1,
                        double *epart, int asis, hipStream_t s) {
  k_grid_energy<<<grid_blocks(g.size()), kGridBS, 0, s>>>(P, g, q1, e1, epart, asis);
}

void launch_gather(const DevParams &P, GridDims g, const double4 *spos, const int *sidx,
                   const int *start, const double *e1, const double *scal, double *F,
                   double *phi, bool add, hipStream_t s) {
  const TileGeom tg = gather_tiles(P, g);
  if (tg.tb > 0) {
    const unsigned nb = (unsigned)(tg.ntile[0] * tg.ntile[1] * tg.ntile[2]);
    const size_t lds = sizeof(double) * tg.R[0] * tg.R[1] * tg.R[2];
#define GT(NU, ADD, TRI) \
  k_gather_tile<NU, ADD, TRI><<<nb, 256, lds, s>>>(P, g, tg, spos, sidx, start, e1, scal, F, phi)
#define CALL(NU)                                         \
  if (add) {                                             \
    if (P.tri) GT(NU, true, true); else GT(NU, true, false);   \
  } else {                                               \
    if (P.tri) GT(NU, false, true); else GT(NU, false, false); \
  }
    MSM_NU_SWITCH(P.nu, CALL)
#undef CALL
#undef GT
    return;
  }
#define GA(NU, ADD, TRI) \
  k_gather<NU, ADD, TRI><<<nblk(P.N, 256), 256, 0, s>>>(P, g, spos, sidx, e1, scal, F, phi)
#define CALL(NU)                                         \
  if (add) {                                             \
    if (P.tri) GA(NU, true, true); else GA(NU, true, false);   \
  } else {                                               \
    if (P.tri) GA(NU, false, true); else GA(NU, false, false); \
  }
  MSM_NU_SWITCH(P.nu, CALL)
#undef CALL
#undef GA
}



void launch_combine(const DevParams &P, const double *Fp, const double *phip, double *F,
                    double *phi, const FinalizeArgs &fin, hipStream_t s) {
  k_combine<<<1 + nblk(P.N, kFinBS), kFinBS, 0, s>>>(P, Fp, phip, F, phi, fin);
}

void launch_finalize(const DevParams &P, const FinalizeArgs &fin, hipStream_t s) {
  k_finalize<<<1, kFinBS, 0, s>>>(P, fin);
}

}  // namespace msm
