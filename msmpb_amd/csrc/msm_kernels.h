// msm_kernels.h -- device-side parameter block and launchers of the MSM
// evaluation chain (implemented in msm_kernels.hip).
#pragma once

#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <cstdint>

#include "msm_plan.h"  // kMaxLevel

namespace msm {

// ---- wavefront scans shared by the kernel files
// Inclusive prefix sums over the wavefront by DPP: row shifts by 1, 2, 4, 8
// lanes (within each row of 16), then lane 15 of each row broadcast into
// the next row and lane 31 into the upper half (no LDS traffic; a lane with
// no source adds 0).
__device__ __forceinline__ int wave_incl_scan(int v) {
  v += __builtin_amdgcn_update_dpp(0, v, 0x111, 0xF, 0xF, false);  // row_shr:1
  v += __builtin_amdgcn_update_dpp(0, v, 0x112, 0xF, 0xF, false);  // row_shr:2
  v += __builtin_amdgcn_update_dpp(0, v, 0x114, 0xF, 0xF, false);  // row_shr:4
  v += __builtin_amdgcn_update_dpp(0, v, 0x118, 0xF, 0xF, false);  // row_shr:8
  v += __builtin_amdgcn_update_dpp(0, v, 0x142, 0xA, 0xF, false);  // row_bcast:15
  v += __builtin_amdgcn_update_dpp(0, v, 0x143, 0xC, 0xF, false);  // row_bcast:31
  return v;
}
template <int CTRL, int ROWS>
__device__ __forceinline__ double dpp_take(double v) {
  const unsigned long long u = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)u, CTRL, ROWS, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(u >> 32), CTRL, ROWS, 0xF, false);
  return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}
__device__ __forceinline__ double wave_incl_scan(double v) {
  v += dpp_take<0x111, 0xF>(v);
  v += dpp_take<0x112, 0xF>(v);
  v += dpp_take<0x114, 0xF>(v);
  v += dpp_take<0x118, 0xF>(v);
  v += dpp_take<0x142, 0xA>(v);
  v += dpp_take<0x143, 0xC>(v);
  return v;
}


// The grid chain runs beside the pair kernel and bounds the evaluation: its
// wavefronts take instruction issue ahead of the pair kernel's on a shared
// SIMD (wave priority 3; the pair kernel keeps the default 0).
__device__ __forceinline__ void chain_issue_priority() { __builtin_amdgcn_s_setprio(3); }
__device__ __forceinline__ void chain_issue_priority(bool on) {
  if (on) __builtin_amdgcn_s_setprio(3);
}

constexpr int kMaxNu = 10;

// Passed by value to every kernel (lives in the kernarg segment, read with
// scalar loads).  Orthorhombic boxes only: A and Ai are diagonal.
struct DevParams {
  double box[3];       // edge lengths (diagonal of A)
  double ibox[3];      // 1 / edge lengths (diagonal of Ai)
  int nc[3];           // pair-search cells per axis
  int ncell;
  double self_img;     // sum of g(r) over the periodic images of an atom itself
                       // inside the cutoff (nonzero only for boxes narrower than a0)
  double cw[3];        // cell widths
  double a0, a02, inv_a0, inv_a02, two_inv_a03;
  double cut2;         // a0^2 enlarged by 1e-12: bounding-box tests never drop a pair
  double tau_s[kMaxNu], dtau_s[kMaxNu];  // tau_k / a0 and (k + 1) tau_{k+1} 2 / a0^3
  float ftau[kMaxNu], fdtau[kMaxNu];     // the same in FP32
  float fa02, fia02, fcut2;              // fcut2: a0^2 enlarged by 1e-4
  int nu;
  int N;
  double tau[kMaxNu];
  double Q[kMaxNu * kMaxNu / 2];
  double J[kMaxNu + 1];
  double coeff_self;   // coeff1 - gamma(0)/a0 (per-charge constant potential)
  double coeff1, coeff2, gamma0;
  // PME: the pair term is erfc(beta r)/r (ewald != 0); ewc = 2 beta/sqrt(pi)
  int ewald;
  double beta, beta2, ewc;
  float fbeta, fbeta2, fewc;
  // triclinic box (tri != 0): A and A^-1 row-major; box/ibox/cw above then
  // hold the face-to-face widths, and cells are slabs in fractional space
  int tri;
  double A[9], Ai[9];
};

struct GridDims {
  int x, y, z;
  __host__ __device__ long size() const { return (long)x * y * z; }
};

// One axis of the top-level FFT: length, radices of its Stockham stages and
// lines per workgroup.
constexpr int kFftMaxStages = 13;
struct FftAxis {
  int M, nst, B;
  int rad[kFftMaxStages];
};
// Plans an axis; false when M has a prime factor above 7 or is too long
// (the top level then falls back to the direct separable DFT).
bool fft_axis_plan(int M, FftAxis &A);

// ---- launchers (stream-ordered, no synchronisation) ----
// `done` (optional, where a launcher takes it): an event bound to the
// completion of the launcher's last kernel, for another stream to wait on.
// It rides on that kernel's dispatch (hipExtLaunchKernel), where a separate
// hipEventRecord adds a marker packet that delays the next kernel of the
// recording stream by several microseconds.
template <typename K, typename... A>
void launch_done(K kernel, dim3 grid, dim3 block, size_t lds, hipStream_t s, hipEvent_t done,
                 A... args) {
  if (done)
    hipExtLaunchKernelGGL(kernel, grid, block, (std::uint32_t)lds, s, nullptr, done, 0u, args...);
  else
    kernel<<<grid, block, lds, s>>>(args...);
}
void launch_fill_int(int *p, long n, int v, hipStream_t s);
void launch_fill_dbl(double *p, long n, double v, hipStream_t s);

// Binning: wrap positions into the box, count atoms per cell (cnt must be
// zero on entry) and record each atom in its cell's bucket row of cap
// entries (bucket: ncell x cap) at its arrival rank, or past the row on the
// overflow list ovf (N entries; ovfn[0] counts them, zero on entry); partial
// sums of q and q^2 per block (2 doubles per block into qpart), non-finite
// inputs counted in flags[0].
int bin_blocks(int N);
void launch_bin(const DevParams &P, const double *pos, const double *q, int *cell_of, int *bucket,
                int cap, int *ovfn, int *ovf, int *cnt, double *qpart, int *flags, hipStream_t s);
// Exclusive scan of cnt into start[0..ncell], cnt and ovfn[0] cleared for
// the next binning (the overflow count moved to ovfn[1]), and reduction of
// qpart into scal[0] = sum q, scal[1] = sum q^2.
void launch_scan(const DevParams &P, int *cnt, int *start, const double *qpart, int nqpart,
                 double *scal, int *ovfn, hipStream_t s);
// The cell-sorted order: sorted wrapped positions and charges (spos) and
// original indices (sidx), atoms ordered inside each cell by original index
// (deterministic results); pos, q are the binning's inputs (the wrap is
// evaluated again).  With sbase, st (N entries each) it also writes every
// atom's spreading placement on the grid g1 (as the placement pass of
// launch_spread does).
void launch_place(const DevParams &P, const double *pos, const double *q, const int *cell_of,
                  const int *start, const int *bucket, int cap, const int *ovfn, const int *ovf,
                  double4 *spos, int *sidx, hipStream_t s, hipEvent_t done = nullptr,
                  GridDims g1 = GridDims{0, 0, 0}, unsigned *sbase = nullptr,
                  double4 *st = nullptr);

// Short-range pairs, full neighbour shell, one wavefront per home cell;
// cols: (dx, dy, dzlo, dzhi) neighbour cell columns, dzhi - dzlo < nc[2].
// Writes force (if F) and potential (if phi) at the original atom index,
// and one energy partial per cell into epart (ncell doubles).  The mixed
// kernel evaluates in FP32 relative to the home cell and sums in FP64.
int pair_blocks(int N);
void launch_pair_fp64(const DevParams &P, const double4 *spos, const int *sidx, const int *start,
                      const int4 *cols, int ncol, double *F, double *phi, double *epart,
                      hipStream_t s, hipEvent_t done = nullptr);
void launch_pair_fp32(const DevParams &P, const double4 *spos, const int *sidx, const int *start,
                      const int4 *cols, int ncol, double *F, double *phi, double *epart,
                      hipStream_t s, hipEvent_t done = nullptr);
// Pairs i < j whose nearest image lies within sqrt(r2max), counted per home
// cell into cell_count (ncell entries); cols are cell columns built for
// that radius.
void launch_pair_count(const DevParams &P, const double4 *spos, const int *start, const int4 *cols,
                       int ncol, double r2max, unsigned *cell_count, hipStream_t s);

// B-spline spreading of charges onto the level-1 grid, one workgroup per
// grid brick (every grid point written exactly once).
// q1b (optional): q^1 also as bf16_split pairs.  sbase, st (optional, N
// entries each): per-atom placements, computed first (placed: already
// written by the placement pass of the binning) and used where
// spread_by_placements(P, g) holds.
void launch_spread(const DevParams &P, GridDims g, const double4 *spos, const int *start,
                   double *q1, hipStream_t s, unsigned *q1b = nullptr, unsigned *sbase = nullptr,
                   double4 *st = nullptr, hipEvent_t done = nullptr, bool placed = false);
bool spread_by_placements(const DevParams &P, GridDims g);
// q^l = R q^{l-1} (full weighting with the two-scale stencil).
// qcb (optional): q^l also as bf16_split pairs.
void launch_restrict(const DevParams &P, GridDims fine, GridDims coarse, const double *qf,
                     double *qc, hipStream_t s, unsigned *qcb = nullptr);
// Top level: e^L = IDFT(DFT(q^L) * khat) / M^3.  work: 4 * M^3 doubles.
// asis selects how q^L is read: directly (0), or by one of the reference
// HEAD's index maps -- the FFT build's padded stride and offset at L >= 2
// (kTopAsisStride) or the DFT build's flat read of the padded level-1 array
// at L = 1 (kTopAsisFlat).
constexpr int kTopAsisStride = 1, kTopAsisFlat = 2;

// ---- top level in one workgroup (shared by the kernel files)
// FFTg2g (forcefield1.c:499-528): separable DFT along z, y, x (sign -1),
// pointwise product with khat, inverse (sign +1), real part / M^3.
// tw holds cos/sin(2 pi j / M_axis) for the three axes back to back.
// The top-level input at (x, y, z), with the as-is read maps.
__device__ __forceinline__ double top_input(GridDims g, const double *qL, int x, int y, int z,
                                            int asis, int nu) {
  if (asis == kTopAsisStride) {
    // reference HEAD, FFT build (forcefield1.c:504-515 vs :406-415): unpadded
    // data read with the padded stride M + nu - 1 and offset nu/2 - 1
    int o = nu / 2 - 1;
    long f = ((long)(x + o) * (g.y + nu - 1) + (y + o)) * (g.z + nu - 1) + (z + o);
    return f < g.size() ? qL[f] : 0.0;
  }
  if (asis == kTopAsisFlat) {
    // reference HEAD, DFT build at L = 1 (forcefield1.c:483-486 on the array
    // of :338-346): the padded level-1 array (dims M + nu - 1, interior at
    // nu/2 - 1, zero ghosts) read at the flat unpadded index
    const long f = ((long)x * g.y + y) * g.z + z;
    const int P1 = g.y + nu - 1, P2 = g.z + nu - 1, o = nu / 2 - 1;
    const int a = (int)(f / ((long)P1 * P2)) - o, b = (int)((f / P2) % P1) - o,
              c = (int)(f % P2) - o;
    return a >= 0 && a < g.x && b >= 0 && b < g.y && c >= 0 && c < g.z
               ? qL[((long)a * g.y + b) * g.z + c]
               : 0.0;
  }
  return qL[((long)x * g.y + y) * g.z + z];
}

// One DFT pass along `axis` from (ire, iim) into (ore, oim), all in LDS.
template <int BS>
__device__ inline void dft_pass_lds(GridDims g, int axis, double sign, const double *ire,
                             const double *iim, double *ore, double *oim, const double *tw) {
  // (kept lean in registers: it runs beside the pair kernel's wavefronts)
  const int M = axis == 0 ? g.x : axis == 1 ? g.y : g.z;
  const double *c = tw + 2 * (axis == 0 ? 0 : axis == 1 ? g.x : g.x + g.y);
  const int stride = axis == 0 ? g.y * g.z : axis == 1 ? g.z : 1;
  const int n = (int)g.size();
#pragma unroll 1
  for (int e = threadIdx.x; e < n; e += BS) {
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

void launch_top(const DevParams &P, GridDims g, const double *qL, const double *khat,
                const double *tw, double *work, double *eL, int asis, hipStream_t s);
// e^l = P e^{l+1} + K_l * q^l (prolongation fused with the stencil pass);
// FP64 on the vector ALUs.
void launch_level_fp64(const DevParams &P, GridDims g, GridDims gc, GridDims sd,
                       const double *khat, const double *ql, const double *ec, double *el,
                       hipStream_t s);

// ---- level stencils on the matrix cores (msm_mfma.hip) ----
constexpr int MSM_G2G_FP64 = 0;    // v_mfma_f64_16x16x4_f64
constexpr int MSM_G2G_BF16X3 = 1;  // v_mfma_f32_16x16x32_bf16, hi/lo split, FP32 accumulation

constexpr int kG2GMaxTpw = 5;  // BF16: z tiles one wavefront computes per B operand

// A grid value for the BF16 level stencils: its FP32 rounding split into a
// bf16 high part (low 16 bits, round to nearest even) and the bf16 of the
// remainder (high 16 bits); values are finite.
__host__ __device__ inline unsigned bf16_split(double v) {
  auto bits = [](float f) {
    union { float f; unsigned u; } c{f};
    return (c.u + 0x7FFFu + ((c.u >> 16) & 1u)) >> 16;
  };
  const float f = (float)v;
  const unsigned hi = bits(f);
  union { unsigned u; float f; } h{hi << 16};
  return hi | (bits(f - h.f) << 16);
}

struct G2GLevel {
  GridDims g, s;          // level grid and stencil dims
  int nch;                // K chunks per (kx, ky)
  int G;                  // kx groups (one partial grid each)
  int ty, tz;             // tiles along y (16 rows) and z (nz columns)
  int nz;                 // output columns per tile (16; BF16: a multiple of 4)
  int nzseg;              // z segments (BF16: a wavefront takes tpw consecutive z tiles)
  int tpw;                // z tiles per wavefront (1 .. kG2GMaxTpw; FP64: 1)
  long task0;             // first wavefront of this level in the launch
  const double *q;        // q^l
  const unsigned *qb;     // q^l as bf16_split pairs (BF16 stencils)
  double *part;           // G x M^3 partial correlations
  const void *bfrag;      // Toeplitz B operands in per-lane layout
};
struct G2GArgs {
  G2GLevel lv[9];
  int nlev;
  long ntasks;
  int urgent;  // on the chain's critical path: wave priority raised
};
// Chooses tiles, K chunks and kx groups for about target_waves wavefronts.
void plan_g2g_level(int prec, GridDims g, GridDims s, int target_waves, G2GLevel &L);
// Levels launched together share the chunk count and (BF16) the z tiles per
// wavefront: equalises them over lv[0..n) and fixes G, nz, nzseg; returns
// the tasks of the launch.
long plan_g2g_launch(int prec, G2GLevel *lv, int n, int target_waves);
long g2g_level_tasks(const G2GLevel &L);
size_t g2g_bfrag_bytes(int prec, const G2GLevel &L);
void g2g_bfrag_fill(int prec, const G2GLevel &L, const double *khat, void *host_out);
size_t g2g_lds_bytes(int prec, const G2GLevel &L);
int g2g_max_nch(int prec);
// All levels' stencil correlations in one launch (one wavefront per task).
void launch_g2g_tiles(int prec, G2GArgs args, size_t lds, hipStream_t s, hipEvent_t done = nullptr);
// e^l = sum of the G partials + prolongation of e^{l+1}; with q1 != null
// also one partial of sum q1 e1 per workgroup into epart (prolong_blocks(g)).
int prolong_blocks(GridDims g);
// The top level folded into the prolongation onto the next level (gc: the
// top grid, at most kTopFuse points; g: the next level, at most
// kTopFuseBricks prolongation bricks): e_L (also stored into eL) from qL as
// launch_top computes it, then el as launch_prolong_sum computes it.
constexpr long kTopFuse = 512;
constexpr int kTopFuseBricks = 64;
bool prolong_top_fits(GridDims g, GridDims gc);
void launch_prolong_top(const DevParams &P, GridDims g, GridDims gc, int G, const double *part,
                        const double *qL, const double *khat, const double *tw, double *eL,
                        double *el, int asis, hipStream_t s);
void launch_prolong_sum(const DevParams &P, GridDims g, GridDims gc, int G, const double *part,
                        const double *ec, double *el, const double *q1, double *epart,
                        hipStream_t s);

// Grid energy partials: sum q1*e1 per block (asis != 0: the reference
// HEAD's L = 1 index map).
int grid_blocks(long n);
void launch_grid_energy(const DevParams &P, GridDims g, const double *q1, const double *e1,
                        double *epart, int asis, hipStream_t s);
// Final energy reduction: out = {E, constant part, pair part, grid part,
// non-finite input count}; flags[0] is cleared.
struct FinalizeArgs {
  const double *ep;
  int np;
  const double *eg;
  int ng;
  const double *scal;
  int *flags;
  double *out;  // >= 5 doubles; null: no reduction
  double *user;  // optional caller buffer that also receives out[0..3]
};
// Interpolation: F (=, or += with add) q E and phi (=, +=) grid potential
// + constant self terms.  Tiles of cells (start: the cell ranges) read the
// grid through LDS.
void launch_gather(const DevParams &P, GridDims g, const double4 *spos, const int *sidx,
                   const int *start, const double *e1, const double *scal, double *F,
                   double *phi, bool add, hipStream_t s);
void launch_finalize(const DevParams &P, const FinalizeArgs &fin, hipStream_t s);
// F += Fp, phi += phip (either may be null) and the energy reduction.
void launch_combine(const DevParams &P, const double *Fp, const double *phip, double *F,
                    double *phi, const FinalizeArgs &fin, hipStream_t s);


// ---- box-dependent stencils on the GPU (FF_rebuild) ----
struct StencilBox {
  double A[3][3], Ai[3][3];  // box matrix (columns = box vectors) and inverse
  double as[3];              // column norms of Ai: the image search extent
  double aCut[kMaxLevel + 1];
  double tau[kMaxNu];        // softening series tau_0 .. tau_{nu-1}
  double taylor[7];          // top-level kernel series coefficients
  double detA;
  int nu, L;
  int M[kMaxLevel + 1][3];   // level grids
  int S[kMaxLevel + 1][3];   // level stencil extents
  double beta;               // PME: Ewald splitting parameter
};
// Level l's stencil into khat (S_l^3): image sum, then omega' along z (w2),
// y (w1) and x (w0); tmp_a, tmp_b hold S_l^3 doubles each.
void launch_stencil_level(const StencilBox &b, int l, const double *w0, const double *w1,
                          const double *w2, double *tmp_a, double *tmp_b, double *khat,
                          hipStream_t s);
// Top-level Fourier kernel into khat (M_L^3); cL = cL_x, cL_y, cL_z packed.
void launch_stencil_top(const StencilBox &b, const double *cL, double *khat, hipStream_t s);
// PME's Gaussian Fourier kernel into khat (M_1^3), same cL packing.
void launch_stencil_top_pme(const StencilBox &b, const double *cL, double *khat, hipStream_t s);
// Toeplitz B operands of a level stencil (device khat) in MFMA lane order,
// the same values g2g_bfrag_fill writes on the host.
void launch_g2g_bfrag_fill(int prec, const G2GLevel &L, const double *khat, void *out,
                           hipStream_t s);

}  // namespace msm
