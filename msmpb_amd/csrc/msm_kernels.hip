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
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) red[w] = v;
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

__global__ __launch_bounds__(kBinBS) void k_bin(DevParams P, const double *__restrict__ pos,
                                                const double *__restrict__ q,
                                                int *__restrict__ cell_of, int *__restrict__ rank,
                                                int *__restrict__ cnt, double4 *__restrict__ wpos,
                                                double *__restrict__ qpart, int *__restrict__ flags) {
  int i = blockIdx.x * kBinBS + threadIdx.x;
  double qi = 0.0;
  if (i < P.N) {
    double r[3] = {pos[3 * (long)i], pos[3 * (long)i + 1], pos[3 * (long)i + 2]};
    qi = q[i];
    int c[3];
    double rw[3];
    bool ok = isfinite(qi);
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      double s = r[a] * P.ibox[a];
      double f = floor(s);
      double sw = s - f;
      ok = ok && isfinite(s);
      // s - floor(s) may round up to 1 for tiny negative s; keep in range
      int ci = (int)(sw * P.nc[a]);
      if (!(sw >= 0.0) || ci < 0) ci = 0;
      if (ci >= P.nc[a]) ci = P.nc[a] - 1;
      c[a] = ci;
      rw[a] = r[a] - f * P.box[a];
      if (!isfinite(rw[a])) rw[a] = 0.0;
    }
    if (!ok) atomicAdd(flags, 1);
    int cell = (c[0] * P.nc[1] + c[1]) * P.nc[2] + c[2];
    cell_of[i] = cell;
    rank[i] = atomicAdd(cnt + cell, 1);  // arrival order; the cell sort fixes the order
    wpos[i] = make_double4(rw[0], rw[1], rw[2], qi);
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

// Exclusive scan of the cell counts (start[0..n]) in tiles of 8 elements per
// thread: a per-thread prefix, a wavefront shuffle scan of the thread totals
// and a scan of the 16 wavefront totals; the counts are left in place (the
// cell sort clears them for the next evaluation).  Also reduces the per-block
// charge sums of the binning pass.
__global__ __launch_bounds__(kScanBS) void k_scan(int n, const int *__restrict__ cnt,
                                                  int *__restrict__ start,
                                                  const double *__restrict__ qpart, int nq,
                                                  double *__restrict__ scal) {
  constexpr int PER = 8, NW = kScanBS / 64;
  __shared__ int wtot[NW];
  __shared__ int carry_s;
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  if (t == 0) carry_s = 0;
  __syncthreads();
  for (int base = 0; base < n; base += kScanBS * PER) {
    const int k0 = base + t * PER;
    int v[PER], run = 0;
#pragma unroll
    for (int u = 0; u < PER; ++u) {
      v[u] = k0 + u < n ? cnt[k0 + u] : 0;
      run += v[u];
    }
    int incl = run;
    for (int off = 1; off < 64; off <<= 1) {
      int o = __shfl_up(incl, off, 64);
      if (lane >= off) incl += o;
    }
    if (lane == 63) wtot[w] = incl;
    __syncthreads();
    int woff = 0, tile = 0;
#pragma unroll
    for (int j = 0; j < NW; ++j) {
      woff += j < w ? wtot[j] : 0;
      tile += wtot[j];
    }
    int pre = carry_s + woff + incl - run;
#pragma unroll
    for (int u = 0; u < PER; ++u) {
      if (k0 + u < n) start[k0 + u] = pre;
      pre += v[u];
    }
    __syncthreads();  // everyone has read wtot and carry_s
    if (t == 0) carry_s += tile;
    __syncthreads();
  }
  if (t == 0) start[n] = carry_s;
  // charge sums for the constant energy terms
  double q1 = 0, q2 = 0;
  for (int k = t; k < nq; k += kScanBS) q1 += qpart[2 * k], q2 += qpart[2 * k + 1];
  __syncthreads();
  double a = block_sum<kScanBS>(q1);
  __syncthreads();
  double c2 = block_sum<kScanBS>(q2);
  if (t == 0) scal[0] = a, scal[1] = c2;
}

__global__ __launch_bounds__(256) void k_scatter(int N, const int *__restrict__ cell_of,
                                                 const int *__restrict__ rank,
                                                 const double4 *__restrict__ wpos,
                                                 const int *__restrict__ start,
                                                 double4 *__restrict__ spos,
                                                 int *__restrict__ sidx,
                                                 int *__restrict__ scell) {
  int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= N) return;
  int c = cell_of[i];
  int slot = start[c] + rank[i];
  spos[slot] = wpos[i];
  sidx[slot] = i;
  scell[slot] = c;
}

// One wavefront per cell: ranks the cell's atoms by original index (so the
// order, and every sum built on it, is deterministic), moves them in place
// and writes the cell-local FP32 coordinates used by the mixed pair kernel.
__global__ __launch_bounds__(256) void k_cell_sort(DevParams P, const int *__restrict__ start,
                                                   int *__restrict__ cnt,
                                                   double4 *__restrict__ spos,
                                                   int *__restrict__ sidx,
                                                   float4 *__restrict__ lpos) {
  int c = blockIdx.x * 4 + (threadIdx.x >> 6);
  int lane = threadIdx.x & 63;
  if (c >= P.ncell) return;
  if (lane == 0) cnt[c] = 0;  // ready for the next evaluation's binning
  int b = start[c], n = start[c + 1] - b;
  int cz = c % P.nc[2], cy = (c / P.nc[2]) % P.nc[1], cx = c / (P.nc[1] * P.nc[2]);
  double ox = cx * P.cw[0], oy = cy * P.cw[1], oz = cz * P.cw[2];
  if (n <= 64) {
    int key = lane < n ? sidx[b + lane] : 0x7fffffff;
    double4 v = lane < n ? spos[b + lane] : make_double4(0, 0, 0, 0);
    int rank = 0;
    for (int j = 0; j < n; ++j) rank += __shfl(key, j, 64) < key;
    if (lane < n) {
      sidx[b + rank] = key;
      spos[b + rank] = v;
      lpos[b + rank] = make_float4((float)(v.x - ox), (float)(v.y - oy), (float)(v.z - oz), (float)v.w);
    }
  } else if (lane == 0) {  // rare crowded cell: insertion sort by one lane
    for (int k = b + 1; k < b + n; ++k) {
      int key = sidx[k];
      double4 v = spos[k];
      int m = k - 1;
      while (m >= b && sidx[m] > key) {
        sidx[m + 1] = sidx[m];
        spos[m + 1] = spos[m];
        --m;
      }
      sidx[m + 1] = key;
      spos[m + 1] = v;
    }
    for (int k = b; k < b + n; ++k) {
      double4 v = spos[k];
      lpos[k] = make_float4((float)(v.x - ox), (float)(v.y - oy), (float)(v.z - oz), (float)v.w);
    }
  }
}

// ---------------------------------------------------------------- pairs
// partcl2partcl (forcefield1.c:240-321) as a full-shell gather: each atom
// sums q_j g(r) over every neighbour j != i and every periodic image inside
// the cutoff, with g(r) = 1/r - tau(r^2/a0^2 - 1)/a0.  Neighbour cells are
// visited through a stencil of cell offsets; an offset that leaves the box
// carries the image shift, so small boxes get every image exactly once.
//
// One wavefront per home cell: the cell's atoms (up to 64 per pass) occupy
// lane groups of S lanes (S = 4, 2 or 1 as the cell holds <= 16, <= 32 or
// more atoms), and for every stencil offset -- the same for the whole
// wavefront -- the S lanes of a group stride over that neighbour cell's
// atoms.  All lanes run the same loop, a j atom is fetched once for all
// groups, and a group's partial sums are combined with shuffles; no
// atomics, so the result is deterministic.
constexpr int kPairBS = 256;

// Neighbour atoms are staged through LDS in chunks: the 64 lanes take 64
// stencil offsets at a time, a wavefront scan of their cell populations
// decides how many cells fit the chunk, and the lanes copy those atoms --
// already shifted into the home cell's frame -- cooperatively.  The
// interaction loop then reads only LDS.
template <bool MIXED>
struct PairStage;
template <>
struct PairStage<true> {
  using T = float4;
  static constexpr int CAP = 512;
};
template <>
struct PairStage<false> {
  using T = double4;
  static constexpr int CAP = 256;
};

// Interactions of one home atom (position ai, sorted index ii) with the
// staged atoms e = bs, bs + S, ... < npad (npad a multiple of 16 S; the tail
// holds far-away padding).  Self pairs have r2 == 0; with CHK, periodic self
// images (narrow cell grids only) are excluded by index as well.
template <bool CHK, int NU>
__device__ __forceinline__ void pair_run_fp32(const DevParams &P, const float4 *jbuf,
                                              const int *jidx, int npad, int bs, int S, float4 ai,
                                              int ii, double &acc, double &fx, double &fy,
                                              double &fz) {
  const float a02 = (float)P.a02, ia02 = (float)P.inv_a02, ia0 = (float)P.inv_a0,
              t3 = (float)P.two_inv_a03;
  float tau[NU], dtau[NU - 1];  // tau(s) and tau'(s) coefficients
#pragma unroll
  for (int k = 0; k < NU; ++k) tau[k] = (float)P.tau[k];
#pragma unroll
  for (int k = 0; k + 1 < NU; ++k) dtau[k] = (float)((k + 1) * P.tau[k + 1]);
  for (int e0 = bs; e0 < npad; e0 += 16 * S) {
    // FP32 partial sums over 16 neighbours, then added in FP64
    float sa = 0.f, sx = 0.f, sy = 0.f, sz = 0.f;
#pragma unroll 8
    for (int u = 0; u < 16; ++u) {
      // branch-free: some lane of the wavefront nearly always accepts
      const int e = e0 + u * S;
      const float4 aj = jbuf[e];
      const float dx = aj.x - ai.x, dy = aj.y - ai.y, dz = aj.z - ai.z;
      const float r2 = dx * dx + dy * dy + dz * dz;
      bool ok = r2 < a02 && r2 > 0.f;
      if constexpr (CHK) ok = ok && jidx[e] != ii;
      const float r2s = ok ? r2 : 1.0f;
      const float qj = ok ? aj.w : 0.0f;
      const float sv = r2s * ia02 - 1.0f;
      float ts = tau[NU - 1], t1 = dtau[NU - 2];
#pragma unroll
      for (int k = NU - 2; k >= 0; --k) ts = tau[k] + sv * ts;
#pragma unroll
      for (int k = NU - 3; k >= 0; --k) t1 = dtau[k] + sv * t1;
      const float rinv = __builtin_amdgcn_rsqf(r2s);  // r2 >= 1e-2 for physical input
      sa += qj * (rinv - ts * ia0);
      const float qc = qj * (-rinv * rinv * rinv - t1 * t3);
      sx += qc * dx, sy += qc * dy, sz += qc * dz;
    }
    acc += (double)sa;
    fx += (double)sx, fy += (double)sy, fz += (double)sz;
  }
}

// 1/sqrt(r2) in FP64: hardware estimate refined by two Newton steps.
__device__ __forceinline__ double rsqrt_f64(double r2) {
  double y = __builtin_amdgcn_rsq(r2);
  const double h = 0.5 * r2;
  y = y * fma(-h * y, y, 1.5);
  y = y * fma(-h * y, y, 1.5);
  return y;
}

template <bool CHK, int NU>
__device__ __forceinline__ void pair_run_fp64(const DevParams &P, const double4 *jbuf,
                                              const int *jidx, int npad, int bs, int S, double4 ai,
                                              int ii, double &acc, double &fx, double &fy,
                                              double &fz) {
  const double a02 = P.a02;
  double dtau[NU - 1];
#pragma unroll
  for (int k = 0; k + 1 < NU; ++k) dtau[k] = (k + 1) * P.tau[k + 1];
  for (int e = bs; e < npad; e += S) {
    const double4 aj = jbuf[e];
    const double dx = aj.x - ai.x, dy = aj.y - ai.y, dz = aj.z - ai.z;
    const double r2 = dx * dx + dy * dy + dz * dz;
    bool ok = r2 < a02 && r2 > 0.0;
    if constexpr (CHK) ok = ok && jidx[e] != ii;
    if (ok) {
      const double sv = r2 * P.inv_a02 - 1.0;
      double ts = P.tau[NU - 1], t1 = dtau[NU - 2];
#pragma unroll
      for (int k = NU - 2; k >= 0; --k) ts = P.tau[k] + sv * ts;
#pragma unroll
      for (int k = NU - 3; k >= 0; --k) t1 = dtau[k] + sv * t1;
      const double rinv = rsqrt_f64(r2);
      acc += aj.w * (rinv - ts * P.inv_a0);
      const double qc = aj.w * (-rinv * rinv * rinv - t1 * P.two_inv_a03);
      fx += qc * dx, fy += qc * dy, fz += qc * dz;
    }
  }
}

template <bool MIXED, int NU>
__global__ __launch_bounds__(64) void k_pair_cell(
    DevParams P, const double4 *__restrict__ spos, const float4 *__restrict__ lpos,
    const int *__restrict__ sidx, const int *__restrict__ start, const int4 *__restrict__ sten,
    int nsten, double *__restrict__ F, double *__restrict__ phi, double *__restrict__ epart) {
  using T = typename PairStage<MIXED>::T;
  using S1 = decltype(T().x);
  constexpr int CAP = PairStage<MIXED>::CAP;  // multiple of 64
  __shared__ T jbuf[CAP];
  __shared__ int jidx[CAP];
  __shared__ int seg_off[65], seg_beg[64];
  __shared__ T seg_shift[64];
  const int c = blockIdx.x, lane = threadIdx.x;
  const int b0 = start[c], n = start[c + 1] - b0;
  const int cz = c % P.nc[2], cy = (c / P.nc[2]) % P.nc[1], cx = c / (P.nc[1] * P.nc[2]);
  // bounding box of the home atoms: neighbour atoms farther than the cutoff
  // from it interact with none of them and are not staged
  S1 lo[3] = {(S1)1e30, (S1)1e30, (S1)1e30}, hi[3] = {(S1)-1e30, (S1)-1e30, (S1)-1e30};
  for (int k = lane; k < n; k += 64) {
    T v;
    if constexpr (MIXED) v = lpos[b0 + k];
    else v = spos[b0 + k];
    lo[0] = fmin(lo[0], v.x), lo[1] = fmin(lo[1], v.y), lo[2] = fmin(lo[2], v.z);
    hi[0] = fmax(hi[0], v.x), hi[1] = fmax(hi[1], v.y), hi[2] = fmax(hi[2], v.z);
  }
#pragma unroll
  for (int off = 1; off < 64; off <<= 1)
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      lo[d] = fmin(lo[d], __shfl_xor(lo[d], off, 64));
      hi[d] = fmax(hi[d], __shfl_xor(hi[d], off, 64));
    }
  // slightly enlarged cutoff: the box test must never drop a pair that the
  // interaction test (different rounding) would keep
  const S1 cut2 = (S1)P.a02 * (MIXED ? (S1)1.0001 : (S1)(1.0 + 1e-12));
  const unsigned long long lanes_below = (1ull << lane) - 1ull;
  // padding atoms and idle home slots sit far apart and far from the box
  const S1 far = (S1)1e18;
  double esum = 0.0;
  for (int i0 = 0; i0 < n;) {
    const int rem = n - i0;
    const int S = rem <= 16 ? 4 : (rem <= 32 ? 2 : 1);
    const int a = lane / S, bs = lane - a * S;
    const bool act = a < rem;
    const int ii = b0 + i0 + a;
    T ai;
    if constexpr (MIXED) ai = act ? lpos[ii] : make_float4(-far, -far, -far, 0.f);
    else ai = act ? spos[ii] : make_double4(-far, -far, -far, 0.0);
    double acc = 0.0, fx = 0.0, fy = 0.0, fz = 0.0;
    for (int o0 = 0; o0 < nsten;) {
      // ---- stage: lane k takes stencil offset o0 + k
      const int o = o0 + lane;
      int cnt = 0, jb = 0;
      T sh;
      if (o < nsten) {
        int4 d = sten[o];
        int ox = cx + d.x, oy = cy + d.y, oz = cz + d.z;
        int kx = floordiv(ox, P.nc[0]), ky = floordiv(oy, P.nc[1]), kz = floordiv(oz, P.nc[2]);
        ox -= kx * P.nc[0], oy -= ky * P.nc[1], oz -= kz * P.nc[2];
        int cn = (ox * P.nc[1] + oy) * P.nc[2] + oz;
        jb = start[cn];
        cnt = start[cn + 1] - jb;
        if constexpr (MIXED)  // neighbour-cell origin relative to the home-cell origin
          sh = make_float4(d.x * (float)P.cw[0], d.y * (float)P.cw[1], d.z * (float)P.cw[2], 0.f);
        else  // periodic image of the neighbour cell
          sh = make_double4(kx * P.box[0], ky * P.box[1], kz * P.box[2], 0.0);
      }
      int incl = cnt;
      for (int off = 1; off < 64; off <<= 1) {
        int t = __shfl_up(incl, off, 64);
        if (lane >= off) incl += t;
      }
      unsigned long long fit = __ballot(o < nsten && incl <= CAP);
      int m = __popcll(fit);  // leading lanes whose cells fit (counts are >= 0)
      if (m == 0) m = 1;      // one cell larger than the chunk: read it in pieces below
      const int total = __shfl(incl, m - 1, 64);
      if (lane < m) {
        seg_off[lane] = incl - cnt;
        seg_beg[lane] = jb;
        seg_shift[lane] = sh;
      }
      if (lane == 0) seg_off[m] = total;
      __syncthreads();
      for (int p0 = 0; p0 < total; p0 += CAP) {
        const int ng = min(CAP, total - p0);
        int np = 0;  // staged atoms that pass the bounding-box test
        // the chunk's global loads are issued four at a time before any of
        // them is used; the segment of element e only moves forward as e grows
        constexpr int IT = CAP / 64, LB = 4;
        int sl = 0;
#pragma unroll
        for (int it0 = 0; it0 < IT; it0 += LB) {
          T v[LB];
          int jj[LB];
#pragma unroll
          for (int u = 0; u < LB; ++u) {
            const int e = 64 * (it0 + u) + lane;
            jj[u] = -1;
            if (e < ng) {
              const int g = p0 + e;
              while (sl + 1 < m && seg_off[sl + 1] <= g) ++sl;
              jj[u] = seg_beg[sl] + (g - seg_off[sl]);
              if constexpr (MIXED) v[u] = lpos[jj[u]];
              else v[u] = spos[jj[u]];
              const T s2 = seg_shift[sl];
              v[u].x += s2.x, v[u].y += s2.y, v[u].z += s2.z;
            }
          }
#pragma unroll
          for (int u = 0; u < LB; ++u) {
            bool keep = false;
            if (jj[u] >= 0) {
              S1 bx = fmax(fmax(lo[0] - v[u].x, v[u].x - hi[0]), (S1)0);
              S1 by = fmax(fmax(lo[1] - v[u].y, v[u].y - hi[1]), (S1)0);
              S1 bz = fmax(fmax(lo[2] - v[u].z, v[u].z - hi[2]), (S1)0);
              keep = bx * bx + by * by + bz * bz < cut2;
            }
            unsigned long long bm = __ballot(keep);
            if (keep) {
              int slot = np + __popcll(bm & lanes_below);
              jbuf[slot] = v[u];
              jidx[slot] = jj[u];
            }
            np += __popcll(bm);
          }
        }
        // pad to a multiple of 64 (np <= CAP, CAP a multiple of 64)
        const int npad = (np + 63) & ~63;
        if (np + lane < npad) {
          T v;
          v.x = far, v.y = far, v.z = far, v.w = 0;
          jbuf[np + lane] = v;
          jidx[np + lane] = -1;
        }
        __syncthreads();
        // ---- interactions of the staged atoms with this lane group's atom
        if constexpr (MIXED) {
          if (P.self_wrap) pair_run_fp32<true, NU>(P, jbuf, jidx, npad, bs, S, ai, ii, acc, fx, fy, fz);
          else pair_run_fp32<false, NU>(P, jbuf, jidx, npad, bs, S, ai, ii, acc, fx, fy, fz);
        } else {
          if (P.self_wrap) pair_run_fp64<true, NU>(P, jbuf, jidx, npad, bs, S, ai, ii, acc, fx, fy, fz);
          else pair_run_fp64<false, NU>(P, jbuf, jidx, npad, bs, S, ai, ii, acc, fx, fy, fz);
        }
        __syncthreads();
      }
      o0 += m;
    }
    // combine the S lanes of each group (consecutive lanes)
    for (int off = 1; off < S; off <<= 1) {
      acc += __shfl_xor(acc, off, 64);
      fx += __shfl_xor(fx, off, 64);
      fy += __shfl_xor(fy, off, 64);
      fz += __shfl_xor(fz, off, 64);
    }
    if (act && bs == 0) {
      double qi = ai.w;
      int oi = sidx[ii];
      if (F) {
        F[3 * (long)oi] = qi * fx;
        F[3 * (long)oi + 1] = qi * fy;
        F[3 * (long)oi + 2] = qi * fz;
      }
      if (phi) phi[oi] = acc;
      esum += 0.5 * qi * acc;
    }
    i0 += 64 / S;
  }
  for (int off = 32; off > 0; off >>= 1) esum += __shfl_down(esum, off, 64);
  if (lane == 0) epart[c] = esum;
}

__global__ __launch_bounds__(kPairBS) void k_pair_count(
    DevParams P, const double4 *__restrict__ spos, const int *__restrict__ scell,
    const int *__restrict__ start, const int4 *__restrict__ sten, int nsten,
    unsigned long long *__restrict__ count) {
  int i = blockIdx.x * kPairBS + threadIdx.x;
  double n = 0;
  if (i < P.N) {
    double4 ai = spos[i];
    int c = scell[i];
    int cz = c % P.nc[2], cy = (c / P.nc[2]) % P.nc[1], cx = c / (P.nc[1] * P.nc[2]);
    for (int o = 0; o < nsten; ++o) {
      int4 d = sten[o];
      int ox = cx + d.x, oy = cy + d.y, oz = cz + d.z;
      int kx = floordiv(ox, P.nc[0]), ky = floordiv(oy, P.nc[1]), kz = floordiv(oz, P.nc[2]);
      ox -= kx * P.nc[0], oy -= ky * P.nc[1], oz -= kz * P.nc[2];
      double bx = kx * P.box[0] - ai.x, by = ky * P.box[1] - ai.y, bz = kz * P.box[2] - ai.z;
      int cn = (ox * P.nc[1] + oy) * P.nc[2] + oz;
      for (int j = start[cn]; j < start[cn + 1]; ++j) {
        if (j <= i) continue;
        double4 aj = spos[j];
        double dx = aj.x + bx, dy = aj.y + by, dz = aj.z + bz;
        // count each pair once, at its nearest image
        if (rint(dx * P.ibox[0]) != 0.0 || rint(dy * P.ibox[1]) != 0.0 ||
            rint(dz * P.ibox[2]) != 0.0)
          continue;
        if (dx * dx + dy * dy + dz * dz < P.a02) n += 1.0;
      }
    }
  }
  double t = block_sum<kPairBS>(n);
  if (threadIdx.x == 0) atomicAdd(count, (unsigned long long)t);
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
// two contiguous index ranges), stages the folded weights of the reaching
// atoms in LDS 256 candidates at a time, and its four wavefronts split the
// atoms of each round; their tiles are summed in a fixed order and every grid
// point is stored once.  No atomics: the result is deterministic.
constexpr int kSX = 16, kSY = 4, kSZ = 4;
constexpr int kSRanges = 512;       // index ranges per batch of cell columns
constexpr int kWXs = 17, kWYs = 5;  // padded record strides (doubles)

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

// Locates candidate k in the range table (roff[lo] <= k < roff[lo + 1]),
// advancing from a previous position lo (k only grows) and loads its atom.
__device__ __forceinline__ double4 spread_fetch(const int *roff, const int *rbeg, int nr, int k,
                                                int total, int &lo,
                                                const double4 *__restrict__ spos) {
  if (k >= total) return make_double4(0.0, 0.0, 0.0, 0.0);
  while (lo + 1 < nr && roff[lo + 1] <= k) ++lo;
  return spos[rbeg[lo] + (k - roff[lo])];
}

// An atom that reaches the brick, queued until 64 of them fill a wavefront:
// fractional offsets t and charge, and the brick-relative index of the
// lowest support point per axis.
struct SpreadItem {
  double t[3];
  double q;
  short rlo[3];  // grid dims stay below 2^15
  short pad;
};

// Folded weights of one atom (the wavefront's 64 queued atoms, one per lane)
// into its records, then Q += WX^T (WY x WZ) by 16 MFMAs.  `valid` lanes
// hold an atom; the others leave zero records.
template <int NU>
__device__ __forceinline__ void spread_flush(const DevParams &P, const SpreadItem &it, bool valid,
                                             int natoms, const int dims[3], double *RX,
                                             double *RY, double *RZ, int lane,
                                             sp_f64x4 &acc) {
  constexpr int ext[3] = {kSX, kSY, kSZ};
  double *rec[3] = {RX + lane * kWXs, RY + lane * kWYs, RZ + lane * kWYs};
#pragma unroll
  for (int u = 0; u < kSX; ++u) rec[0][u] = 0.0;
#pragma unroll
  for (int u = 0; u < kSY; ++u) rec[1][u] = 0.0, rec[2][u] = 0.0;
  if (valid) {
    double w[NU];
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) {
      bspline<NU, false>(P.Q, it.t[ax], w, nullptr);
      const double sc = ax == 1 ? it.q : 1.0;
      const int M = dims[ax];
      if (M >= NU) {  // the support points are distinct grid points
#pragma unroll
        for (int i = 0; i < NU; ++i) {  // weight i sits at point m + NU/2 - i
          int rr = it.rlo[ax] + (NU - 1 - i);
          rr -= rr >= M ? M : 0;
          if (rr < ext[ax]) rec[ax][rr] = sc * w[i];
        }
      } else {  // grids narrower than the support: points repeat, accumulate
        for (int i = 0; i < NU; ++i) {
          int rr = (it.rlo[ax] + (NU - 1 - i)) % M;
          if (rr < ext[ax]) rec[ax][rr] += sc * w[i];
        }
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

template <int NU>
__global__ __launch_bounds__(256) void k_spread_mfma(DevParams P, GridDims g,
                                                     const double4 *__restrict__ spos,
                                                     const int *__restrict__ start,
                                                     double *__restrict__ q1) {
  // per wavefront: records of 64 atoms and a queue of up to 127 reaching atoms
  __shared__ double WX[256 * kWXs];
  __shared__ double WY[256 * kWYs], WZ[256 * kWYs];
  __shared__ SpreadItem queue[4 * 128];
  __shared__ int rbeg[kSRanges], roff[kSRanges + 1];
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
  double *const RX = WX + wv * 64 * kWXs, *const RY = WY + wv * 64 * kWYs,
               *const RZ = WZ + wv * 64 * kWYs;
  SpreadItem *const Qw = queue + wv * 128;
  int npend = 0;  // queued atoms of this wavefront (uniform)
  sp_f64x4 acc = {0.0, 0.0, 0.0, 0.0};
  for (int c0 = 0; c0 < ncol; c0 += kSRanges / 2) {
    const int nc_b = min(ncol - c0, kSRanges / 2);
    __syncthreads();  // previous batch done with the table
    for (int cc = tid; cc < nc_b; cc += 256) {
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
      int incl = run;
      for (int off = 1; off < 64; off <<= 1) {
        int t = __shfl_up(incl, off, 64);
        if (lane >= off) incl += t;
      }
      const int excl = incl - run;
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        int k = 8 * lane + u;
        if (k < nr) roff[k + 1] = excl + loc[u];
      }
      if (lane == 0) roff[0] = 0;
    }
    __syncthreads();
    const int total = roff[nr];
    // wavefront wv takes candidates 64 (4 r + wv) + lane, r = 0, 1, ...; the
    // next candidate's atom is fetched while the current one is processed
    int k = wv * 64 + lane, lo = 0;
    double4 nxt = spread_fetch(roff, rbeg, nr, k, total, lo, spos);
    for (; k - lane < total; k += 256) {
      const double4 at = nxt;
      nxt = spread_fetch(roff, rbeg, nr, k + 256, total, lo, spos);
      SpreadItem it;
      bool reach = k < total;
      const double r[3] = {at.x, at.y, at.z};
#pragma unroll
      for (int ax = 0; ax < 3; ++ax) {
        double sv = r[ax] * P.ibox[ax];
        sv = sv - floor(sv);
        const double t = dims[ax] * sv, m = floor(t);
        it.t[ax] = t - m;
        // brick-relative index of the lowest support point m - NU/2 + 1:
        // m lies in [0, M] and org in [0, M), so one wrap suffices (NU >= 4)
        int v = (int)m - NU / 2 + 1 - org[ax];
        v += v < 0 ? dims[ax] : 0;
        it.rlo[ax] = (short)v;
        reach = reach && (v < ext[ax] || v > dims[ax] - NU);
      }
      it.q = at.w;
      const unsigned long long bm = __ballot(reach);
      const int nrel = __popcll(bm);
      if (reach) Qw[npend + __popcll(bm & below)] = it;
      npend += nrel;
      if (npend >= 64) {
        __builtin_amdgcn_wave_barrier();
        const SpreadItem mine = Qw[lane];
        spread_flush<NU>(P, mine, true, 64, dims, RX, RY, RZ, lane, acc);
        // move the leftover queue entries [64, npend) to the front
        if (64 + lane < npend) {
          const SpreadItem rest = Qw[64 + lane];
          Qw[lane] = rest;
        }
        npend -= 64;
        __builtin_amdgcn_wave_barrier();
      }
    }
  }
  if (npend > 0) {  // the last, partial wavefront of atoms
    __builtin_amdgcn_wave_barrier();
    const SpreadItem mine = Qw[lane];
    spread_flush<NU>(P, mine, lane < npend, (npend + 3) & ~3, dims, RX, RY, RZ, lane, acc);
  }
  // ---- sum the four wavefronts' tiles in a fixed order and store
  __syncthreads();
#pragma unroll
  for (int i = 0; i < 4; ++i) WX[(wv * 4 + i) * 64 + lane] = acc[i];
  __syncthreads();
  if (wv == 0) {
    // f64 16x16x4 result layout: register i of lane l holds row 4 i + l / 16,
    // column l % 16
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      double v = WX[i * 64 + lane];
      for (int w = 1; w < 4; ++w) v += WX[(w * 4 + i) * 64 + lane];
      const int x = gx0 + 4 * i + (lane >> 4), col = lane & 15;
      const int y = gy0 + (col >> 2), z = gz0 + (col & 3);
      if (x < g.x && y < g.y && z < g.z) q1[((long)x * g.y + y) * g.z + z] = v;
    }
  }
}

// ---------------------------------------------------------------- restrict
// restrict_ (forcefield1.c:401-477): q^l_m = sum_n J_nx J_ny J_nz
// q^{l-1}_{2m+n}, n in [-NU/2, NU/2]^3, periodic, as separable passes like
// the reference's z -> y -> x (:442-472).  A workgroup owns a 2^3 coarse
// brick: it loads the (NU+3)^3 fine values the brick reads into LDS, then
// contracts z, y and x in turn.  The small brick keeps the LDS footprint
// (7.4 KB at NU = 6) low enough to run beside the pair kernel.
constexpr int kRB = 2;

template <int NU>
__global__ __launch_bounds__(256) void k_restrict(DevParams P, GridDims f, GridDims c,
                                                  const double *__restrict__ qf,
                                                  double *__restrict__ qc) {
  // coarse brick of kRB^3 points; fine indices 2m + n - H, n in [0, NU]
  constexpr int H = NU / 2, FW = 2 * (kRB - 1) + NU + 1;
  __shared__ double fs[FW * FW * FW];
  __shared__ double t1[FW * FW * kRB];    // [fx][fy][cz]
  __shared__ double t2[FW * kRB * kRB];   // [fx][cy][cz]
  const int tid = threadIdx.x;
  const int nbz = (c.z + kRB - 1) / kRB, nby = (c.y + kRB - 1) / kRB;
  int b = blockIdx.x;
  const int cz0 = (b % nbz) * kRB;
  b /= nbz;
  const int cy0 = (b % nby) * kRB, cx0 = (b / nby) * kRB;
  const int fx0 = 2 * cx0 - H, fy0 = 2 * cy0 - H, fz0 = 2 * cz0 - H;
  for (int k = tid; k < FW * FW * FW; k += 256) {
    const int a = k / (FW * FW), bb = (k / FW) % FW, cc = k % FW;
    fs[k] = qf[((long)wrapi(fx0 + a, f.x) * f.y + wrapi(fy0 + bb, f.y)) * f.z + wrapi(fz0 + cc, f.z)];
  }
  __syncthreads();
  for (int k = tid; k < FW * FW * kRB; k += 256) {  // along z
    const int cz = k % kRB, fxy = k / kRB;
    const double *row = fs + fxy * FW + 2 * cz;
    double acc = 0.0;
#pragma unroll
    for (int n = 0; n <= NU; ++n) acc += P.J[n] * row[n];
    t1[k] = acc;
  }
  __syncthreads();
  for (int k = tid; k < FW * kRB * kRB; k += 256) {  // along y
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
      qc[((long)x * c.y + y) * c.z + z] = acc;
    }
  }
}

// ---------------------------------------------------------------- top level
// FFTg2g (forcefield1.c:499-528): separable DFT along z, y, x (sign -1),
// pointwise product with khat, inverse (sign +1), real part / M^3.
// tw holds cos/sin(2 pi j / M_axis) for the three axes back to back.
__device__ __forceinline__ double top_input(GridDims g, const double *qL, int x, int y, int z,
                                            int asis, int nu) {
  if (!asis) return qL[((long)x * g.y + y) * g.z + z];
  // reference HEAD (forcefield1.c:504-515 vs :406-415): unpadded data read
  // with the padded stride M + nu - 1 and offset nu/2 - 1
  int o = nu / 2 - 1;
  long f = ((long)(x + o) * (g.y + nu - 1) + (y + o)) * (g.z + nu - 1) + (z + o);
  return f < g.size() ? qL[f] : 0.0;
}

// One workgroup of 256 threads with LDS sized to the grid (32 bytes per
// point), so that it finds room beside the concurrently running pair kernel.
constexpr int kTopBS = 256;
constexpr int kTopSmall = 2048;  // largest M^3 done in one workgroup (64 KB of LDS)

// One DFT pass along `axis` from (ire, iim) into (ore, oim), all in LDS.
__device__ void dft_pass_lds(GridDims g, int axis, double sign, const double *ire,
                             const double *iim, double *ore, double *oim, const double *tw) {
  // (kept lean in registers: it runs beside the pair kernel's wavefronts)
  const int M = axis == 0 ? g.x : axis == 1 ? g.y : g.z;
  const double *c = tw + 2 * (axis == 0 ? 0 : axis == 1 ? g.x : g.x + g.y);
  const int stride = axis == 0 ? g.y * g.z : axis == 1 ? g.z : 1;
  const int n = (int)g.size();
#pragma unroll 1
  for (int e = threadIdx.x; e < n; e += kTopBS) {
    const int k = axis == 0 ? e / (g.y * g.z) : axis == 1 ? (e / g.z) % g.y : e % g.z;
    const int base = e - k * stride;
    double sr = 0, si = 0;
    int km = 0;  // k m mod M
#pragma unroll 1
    for (int m = 0, o = base; m < M; ++m, o += stride) {
      const double cr = c[2 * km], ci = sign * c[2 * km + 1];
      const double xr = ire[o], xi = iim[o];
      sr += xr * cr - xi * ci;
      si += xr * ci + xi * cr;
      km += k;
      km -= km >= M ? M : 0;
    }
    ore[e] = sr, oim[e] = si;
  }
  __syncthreads();
}

__global__ __launch_bounds__(kTopBS) void k_top_small(DevParams P, GridDims g,
                                                      const double *__restrict__ qL,
                                                      const double *__restrict__ khat,
                                                      const double *__restrict__ tw,
                                                      double *__restrict__ eL, int asis) {
  extern __shared__ double top_lds[];  // 4 n doubles: two complex buffers
  long n = g.size();
  double *re = top_lds, *im = top_lds + n, *re2 = top_lds + 2 * n, *im2 = top_lds + 3 * n;
  for (long e = threadIdx.x; e < n; e += kTopBS) {
    int z = e % g.z, y = (e / g.z) % g.y, x = e / ((long)g.y * g.z);
    re[e] = top_input(g, qL, x, y, z, asis, P.nu);
    im[e] = 0.0;
  }
  __syncthreads();
  dft_pass_lds(g, 2, -1.0, re, im, re2, im2, tw);
  dft_pass_lds(g, 1, -1.0, re2, im2, re, im, tw);
  dft_pass_lds(g, 0, -1.0, re, im, re2, im2, tw);
  for (long e = threadIdx.x; e < n; e += kTopBS) re2[e] *= khat[e], im2[e] *= khat[e];
  __syncthreads();
  dft_pass_lds(g, 0, 1.0, re2, im2, re, im, tw);
  dft_pass_lds(g, 1, 1.0, re, im, re2, im2, tw);
  dft_pass_lds(g, 2, 1.0, re2, im2, re, im, tw);
  for (long e = threadIdx.x; e < n; e += kTopBS) eL[e] = re[e] / (double)n;
}

// Large top grids (L = 1 with a fine grid): one launch per pass, global memory.
__global__ __launch_bounds__(256) void k_top_pack(DevParams P, GridDims g,
                                                  const double *__restrict__ qL,
                                                  double *__restrict__ re,
                                                  double *__restrict__ im, int asis) {
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
  long e = (long)blockIdx.x * 256 + threadIdx.x;
  if (e < n) eL[e] = re[e] / (double)n;
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

// Compensated sums of the per-cell pair partials and per-block grid partials:
// each of 256 threads sums a fixed strided subset, then a fixed binary tree
// combines the (sum, carry) pairs, so the result does not depend on
// scheduling.  out = {E, constant part, pair part, grid part, flags}; the
// non-finite-input counter is copied to out[4] and cleared for the next
// evaluation.
// Tree-combines the 256 threads' compensated pair sums (sv, cv) with the
// grid partials and writes out.
__device__ void finalize_tail(const DevParams &P, const FinalizeArgs &f, double sv, double cv) {
  __shared__ double ss[2][256], cc[2][256];
  int t = threadIdx.x;
  ss[0][t] = sv, cc[0][t] = cv;
  double s = 0, c = 0;
  for (int k = t; k < f.ng; k += 256) neumaier(s, c, f.eg[k]);
  ss[1][t] = s, cc[1][t] = c;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if (t < h)
      for (int w = 0; w < 2; ++w) {
        double a = ss[w][t], b = cc[w][t];
        csum_add(a, b, ss[w][t + h], cc[w][t + h]);
        ss[w][t] = a, cc[w][t] = b;
      }
    __syncthreads();
  }
  if (t != 0) return;
  double epair = ss[0][0] + cc[0][0], egrid = 0.5 * (ss[1][0] + cc[1][0]);
  double qs = f.scal[0], q2 = f.scal[1];
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
  f.out[4] = (double)*f.flags;
  if (f.user) {
    f.user[0] = f.out[0];
    f.user[1] = econst;
    f.user[2] = epair;
    f.user[3] = egrid;
  }
  *f.flags = 0;
}

__device__ void finalize_body(const DevParams &P, const FinalizeArgs &f) {
  double s = 0, c = 0;
  for (int k = threadIdx.x; k < f.np; k += 256) neumaier(s, c, f.ep[k]);
  finalize_tail(P, f, s, c);
}

__global__ __launch_bounds__(256) void k_finalize(DevParams P, FinalizeArgs f) {
  finalize_body(P, f);
}

// interpolate (forcefield1.c:570-607) + F += q E (:228-231), and the
// potential phi_i = dE/dq_i: grid part plus the constant self terms.
template <int NU, bool ADD>
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
  double t[3];
  int m[3];
#pragma unroll
  for (int ax = 0; ax < 3; ++ax) {
    double tt = dims[ax] * (r[ax] * P.ibox[ax]), mm = floor(tt);
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
  // E = -A^{-T} (M * grad_t), orthorhombic A^{-T} = diag(1/box)
  double Ex = -ex * g.x * P.ibox[0], Ey = -ey * g.y * P.ibox[1], Ez = -ez * g.z * P.ibox[2];
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

template <int NU, bool ADD>
__global__ __launch_bounds__(256) void k_gather(DevParams P, GridDims g,
                                                const double4 *__restrict__ spos,
                                                const int *__restrict__ sidx,
                                                const double *__restrict__ e1,
                                                const double *__restrict__ scal,
                                                double *__restrict__ F, double *__restrict__ phi) {
  int i = blockIdx.x * 256 + threadIdx.x;
  if (i < P.N) gather_atom<NU, ADD>(P, g, spos, sidx, e1, scal, F, phi, i);
}

// F += Fp and phi += phip (the short-range parts, written by the pair kernel
// beside the interpolation), then the energy reduction: every workgroup
// sums a fixed slice of the per-cell pair partials into a compensated
// (sum, carry) pair, and the last workgroup to finish combines those and the
// grid partials in a fixed order.
__global__ __launch_bounds__(256) void k_combine(DevParams P, const double *__restrict__ Fp,
                                                 const double *__restrict__ phip,
                                                 double *__restrict__ F,
                                                 double *__restrict__ phi, double *__restrict__ ep2,
                                                 FinalizeArgs fin) {
  __shared__ double ss[256], cc[256];
  __shared__ int last;
  const int t = threadIdx.x, b = blockIdx.x;
  const long i = (long)b * 256 + t;
  if (i < P.N) {
    if (F)
      for (int d = 0; d < 3; ++d) F[3 * i + d] += Fp[3 * i + d];
    if (phi) phi[i] += phip[i];
  }
  const int chunk = (fin.np + gridDim.x - 1) / gridDim.x;
  const int k0 = b * chunk, k1 = min(fin.np, k0 + chunk);
  double sv = 0.0, cv = 0.0;
  for (int k = k0 + t; k < k1; k += 256) neumaier(sv, cv, fin.ep[k]);
  ss[t] = sv, cc[t] = cv;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if (t < h) {
      double x = ss[t], y = cc[t];
      csum_add(x, y, ss[t + h], cc[t + h]);
      ss[t] = x, cc[t] = y;
    }
    __syncthreads();
  }
  if (t == 0) {
    ep2[2 * b] = ss[0];
    ep2[2 * b + 1] = cc[0];
    __threadfence();
    last = atomicAdd(fin.done, 1u) == gridDim.x - 1;
  }
  __syncthreads();
  if (!last) return;
  __threadfence();
  if (t == 0) *fin.done = 0u;
  // the pair partials are now the gridDim.x (sum, carry) pairs of ep2
  sv = 0.0, cv = 0.0;
  for (int k = t; k < (int)gridDim.x; k += 256) csum_add(sv, cv, ep2[2 * k], ep2[2 * k + 1]);
  finalize_tail(P, fin, sv, cv);
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

void launch_bin(const DevParams &P, const double *pos, const double *q, int *cell_of, int *rank,
                int *cnt, double4 *wpos, double *qpart, int *flags, hipStream_t s) {
  k_bin<<<bin_blocks(P.N), kBinBS, 0, s>>>(P, pos, q, cell_of, rank, cnt, wpos, qpart, flags);
}

void launch_scan(const DevParams &P, const int *cnt, int *start, const double *qpart, int nqpart,
                 double *scal, hipStream_t s) {
  k_scan<<<1, kScanBS, 0, s>>>(P.ncell, cnt, start, qpart, nqpart, scal);
}

void launch_scatter(const DevParams &P, const int *cell_of, const int *rank, const double4 *wpos,
                    const int *start, double4 *spos, int *sidx, int *scell, hipStream_t s) {
  k_scatter<<<nblk(P.N, 256), 256, 0, s>>>(P.N, cell_of, rank, wpos, start, spos, sidx, scell);
}

void launch_cell_sort(const DevParams &P, const int *start, int *cnt, double4 *spos, int *sidx,
                      float4 *lpos, hipStream_t s) {
  k_cell_sort<<<nblk(P.ncell, 4), 256, 0, s>>>(P, start, cnt, spos, sidx, lpos);
}

int pair_blocks(int N) { return (int)nblk(N, kPairBS); }

#define MSM_NU_SWITCH(nu, CALL) \
  switch (nu) {                 \
    case 4: CALL(4); break;     \
    case 6: CALL(6); break;     \
    case 8: CALL(8); break;     \
    default: CALL(10); break;   \
  }


void launch_pair_fp64(const DevParams &P, const double4 *spos, const int *sidx, const int *scell,
                      const int *start, const int4 *sten, int nsten, double *F, double *phi,
                      double *epart, hipStream_t s) {
  (void)scell;
#define CALL(NU)                                                                             \
  k_pair_cell<false, NU><<<P.ncell, 64, 0, s>>>(P, spos, nullptr, sidx, start, sten, nsten, F, \
                                                 phi, epart)
  MSM_NU_SWITCH(P.nu, CALL)
#undef CALL
}

void launch_pair_fp32(const DevParams &P, const float4 *lpos, const int *sidx, const int *scell,
                      const int *start, const int4 *sten, int nsten, double *F, double *phi,
                      double *epart, hipStream_t s) {
  (void)scell;
#define CALL(NU)                                                                             \
  k_pair_cell<true, NU><<<P.ncell, 64, 0, s>>>(P, nullptr, lpos, sidx, start, sten, nsten, F,  \
                                                phi, epart)
  MSM_NU_SWITCH(P.nu, CALL)
#undef CALL
}

void launch_pair_count(const DevParams &P, const double4 *spos, const int *scell, const int *start,
                       const int4 *sten, int nsten, unsigned long long *count, hipStream_t s) {
  k_pair_count<<<pair_blocks(P.N), kPairBS, 0, s>>>(P, spos, scell, start, sten, nsten, count);
}

void launch_spread(const DevParams &P, GridDims g, const double4 *spos, const int *start,
                   double *q1, hipStream_t s) {
  unsigned nb = (unsigned)(((g.x + kSX - 1) / kSX) * ((g.y + kSY - 1) / kSY) * ((g.z + kSZ - 1) / kSZ));
#define CALL(NU) k_spread_mfma<NU><<<nb, 256, 0, s>>>(P, g, spos, start, q1)
  MSM_NU_SWITCH(P.nu, CALL)
#undef CALL
}

void launch_restrict(const DevParams &P, GridDims fine, GridDims coarse, const double *qf,
                     double *qc, hipStream_t s) {
  const unsigned nb = (unsigned)(((coarse.x + kRB - 1) / kRB) * ((coarse.y + kRB - 1) / kRB) *
                                 ((coarse.z + kRB - 1) / kRB));
#define CALL(NU) k_restrict<NU><<<nb, 256, 0, s>>>(P, fine, coarse, qf, qc)
  MSM_NU_SWITCH(P.nu, CALL)
#undef CALL
}

void launch_top(const DevParams &P, GridDims g, const double *qL, const double *khat,
                const double *tw, double *work, double *eL, int asis, hipStream_t s) {
  long n = g.size();
  if (n <= kTopSmall) {
    k_top_small<<<1, kTopBS, 4 * n * sizeof(double), s>>>(P, g, qL, khat, tw, eL, asis);
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
                   const double *e1, const double *scal, double *F, double *phi, bool add,
                   hipStream_t s) {
#define CALL(NU)                                                                        \
  if (add) k_gather<NU, true><<<nblk(P.N, 256), 256, 0, s>>>(P, g, spos, sidx, e1, scal, F, phi); \
  else k_gather<NU, false><<<nblk(P.N, 256), 256, 0, s>>>(P, g, spos, sidx, e1, scal, F, phi)
  MSM_NU_SWITCH(P.nu, CALL)
#undef CALL
}

int combine_blocks(int N) { return (int)nblk(N, 256); }

void launch_combine(const DevParams &P, const double *Fp, const double *phip, double *F,
                    double *phi, double *ep2, const FinalizeArgs &fin, hipStream_t s) {
  k_combine<<<combine_blocks(P.N), 256, 0, s>>>(P, Fp, phip, F, phi, ep2, fin);
}

void launch_finalize(const DevParams &P, const FinalizeArgs &fin, hipStream_t s) {
  k_finalize<<<1, 256, 0, s>>>(P, fin);
}

}  // namespace msm
