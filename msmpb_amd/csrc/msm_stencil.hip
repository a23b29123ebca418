// msm_stencil.hip -- box-dependent stencils generated on the GPU (FF_rebuild).
//
// When the box changes (an NPT step), hkayabilisim/msmpb recomputes the
// top-level Fourier kernel and every level's real-space stencil on the host
// (FF_rebuild, code/forcefield0.c:598-680, with dAL :769-896 and
// kaphatA/kappaA :965-1081).  Here the same FP64 arithmetic runs as
// kernels, one thread per output value:
//   k_stencil_kappa  kappa_l(s) = sum over the periodic images p of
//                    gamma_{a_{l-1}}(A(s-p)) - gamma_{a_l}(A(s-p)) on the
//                    S_l^3 offsets (forcefield0.c:1041-1081);
//   k_stencil_omega  one separable anti-blurring pass (omega' along z, y or x,
//                    periodic offset map (d + 3M/2) % M - M/2,
//                    forcefield0.c:995-1031);
//   k_stencil_top    the top-level kernel khat_L[k] = psi(pi |A^-T k| a_L)
//                    cL_x cL_y cL_z M^3 (forcefield0.c:806-884): a 7-term
//                    series for t <= pi, the sin/cos closed form above;
//   k_stencil_top_pme  the PME comparator's Gaussian kernel (pme0.c:473-503).
// The anti-blurring weights omega' and cL (a few dozen values per axis) and
// the scalar constants stay on the host.  msm_plan.cpp computes the same
// stencils on the host for FF_build; both paths agree to rounding.
#include <hip/hip_runtime.h>

#include <cmath>

#include "msm_kernels.h"

namespace msm {
namespace {

constexpr int kStBS = 256;

__device__ __forceinline__ double st_soft(const StencilBox &b, double s) {
  double g = b.tau[b.nu - 1];
  for (int k = b.nu - 2; k >= 0; --k) g = b.tau[k] + s * g;
  return g;
}

__global__ __launch_bounds__(kStBS) void k_stencil_kappa(StencilBox b, int l,
                                                         double *__restrict__ kap) {
  const int S0 = b.S[l][0], S1 = b.S[l][1], S2 = b.S[l][2];
  const int idx = blockIdx.x * kStBS + threadIdx.x;
  if (idx >= S0 * S1 * S2) return;
  const int i = idx / (S1 * S2), j = (idx / S2) % S1, k = idx % S2;
  const double s[3] = {(double)(i - S0 / 2) / b.M[l][0], (double)(j - S1 / 2) / b.M[l][1],
                       (double)(k - S2 / 2) / b.M[l][2]};
  const double al = b.aCut[l];
  int lo[3], hi[3];
  for (int a = 0; a < 3; ++a) {
    lo[a] = (int)floor(s[a] - al * b.as[a] + 1.0);
    hi[a] = (int)ceil(s[a] + al * b.as[a] - 1.0);
  }
  double sum = 0.0;
  for (int px = lo[0]; px <= hi[0]; ++px)
    for (int py = lo[1]; py <= hi[1]; ++py)
      for (int pz = lo[2]; pz <= hi[2]; ++pz) {
        const double f[3] = {s[0] - px, s[1] - py, s[2] - pz};
        double r2 = 0.0;
        for (int a = 0; a < 3; ++a) {
          const double v = b.A[a][0] * f[0] + b.A[a][1] * f[1] + b.A[a][2] * f[2];
          r2 += v * v;
        }
        const double rho = sqrt(r2) / al;
        if (rho >= 1.0) continue;
        const double g_l = st_soft(b, rho * rho - 1.0) / al;
        const double rho2 = 2.0 * rho;  // r / a_{l-1}
        const double g_lm1 = (rho2 >= 1.0 ? 1.0 / rho2 : st_soft(b, rho2 * rho2 - 1.0)) * (2.0 / al);
        sum += g_lm1 - g_l;
      }
  kap[idx] = sum;
}

// out = the anti-blurring pass of `in` along axis ax (0: x, 1: y, 2: z) of
// an S0 x S1 x S2 stencil: out[.., t, ..] = sum_u w[|fold(u - t)|] in[.., u, ..]
// on a periodic axis of M points; w holds omega'_0 .. omega'_{D-1}.
__global__ __launch_bounds__(kStBS) void k_stencil_omega(const double *__restrict__ in,
                                                         double *__restrict__ out,
                                                         const double *__restrict__ w, int S0,
                                                         int S1, int S2, int ax, int M) {
  const int idx = blockIdx.x * kStBS + threadIdx.x;
  if (idx >= S0 * S1 * S2) return;
  const int i = idx / (S1 * S2), j = (idx / S2) % S1, k = idx % S2;
  const int n = ax == 0 ? S0 : ax == 1 ? S1 : S2;
  const int t = ax == 0 ? i : ax == 1 ? j : k;
  const long stride = ax == 0 ? (long)S1 * S2 : ax == 1 ? S2 : 1;
  const double *base = in + (idx - (long)t * stride);
  double acc = 0.0;
  for (int u = 0; u < n; ++u) {
    const int d = (u - t + (3 * M) / 2) % M - M / 2;
    acc += w[d < 0 ? -d : d] * base[u * stride];
  }
  out[idx] = acc;
}

// t^2 psi(t) numerator and denominator factors of the top-level kernel in
// powers of t^2 (forcefield0.c:820-836).
__device__ __forceinline__ void st_top_poly(int nu, double *S, double *C, int &n) {
  n = nu / 2;
  switch (nu) {
    case 4: S[0] = 15.0, S[1] = -6.0, C[0] = 15.0, C[1] = -1.0; break;
    case 6:
      S[0] = 945.0, S[1] = -420.0, S[2] = 15.0;
      C[0] = 945.0, C[1] = -105.0, C[2] = 1.0;
      break;
    case 8:
      S[0] = 135135.0, S[1] = -62370.0, S[2] = 3150.0, S[3] = -28.0;
      C[0] = 135135.0, C[1] = -17325.0, C[2] = 378.0, C[3] = -1.0;
      break;
    default:
      S[0] = 34459425.0, S[1] = -16216200.0, S[2] = 945945.0, S[3] = -13860.0, S[4] = 45.0;
      C[0] = 34459425.0, C[1] = -4729725.0, C[2] = 135135.0, C[3] = -990.0, C[4] = 1.0;
      break;
  }
}

__device__ double st_top_psi(const StencilBox &b, int kx, int ky, int kz) {
  const double ax = b.Ai[0][0] * kx + b.Ai[1][0] * ky + b.Ai[2][0] * kz;
  const double ay = b.Ai[0][1] * kx + b.Ai[1][1] * ky + b.Ai[2][1] * kz;
  const double az = b.Ai[0][2] * kx + b.Ai[1][2] * ky + b.Ai[2][2] * kz;
  const double k2 = ax * ax + ay * ay + az * az;
  if (k2 == 0.0) return 0.0;
  const double pi = 3.14159265358979323846;
  const double aL = b.aCut[b.L];
  const double t = pi * sqrt(k2) * aL, t2 = t * t;
  double psi;
  if (t <= pi) {
    double s = 0.0, tp = 1.0;
    for (int i = 0; i < 7; ++i) s += tp * b.taylor[i], tp *= t2;
    psi = s / t2;
  } else {
    double S[5], C[5];
    int n;
    st_top_poly(b.nu, S, C, n);
    double sf = 0.0, cf = 0.0, tp = 1.0;
    for (int i = 0; i < n; ++i) sf += S[i] * tp, cf += C[i] * tp, tp *= t2;
    psi = (sf * (sin(t) / t) - cf * cos(t)) / pow(t, 2.0 * b.nu);
  }
  double dfac = 1.0;  // (2 nu - 1)!!
  for (int i = 3; i <= 2 * b.nu - 1; i += 2) dfac *= i;
  return psi * (dfac * pi * aL * aL / b.detA);
}

// cL holds the top grid's cL_x, cL_y, cL_z one after the other.
__global__ __launch_bounds__(kStBS) void k_stencil_top(StencilBox b, const double *__restrict__ cL,
                                                       double *__restrict__ kh) {
  const int g0 = b.M[b.L][0], g1 = b.M[b.L][1], g2 = b.M[b.L][2];
  const long idx = (long)blockIdx.x * kStBS + threadIdx.x;
  if (idx >= (long)g0 * g1 * g2) return;
  const int kx = (int)(idx / ((long)g1 * g2)), ky = (int)((idx / g2) % g1), kz = (int)(idx % g2);
  const int cx = kx < (g0 + 1) / 2 ? kx : kx - g0;
  // y folds with a real-valued midpoint (forcefield0.c:872)
  const int cy = ky < (g1 + 1) / 2.0 ? ky : ky - g1;
  const double c = cL[kx] * cL[g0 + ky] * cL[g0 + g1 + kz];
  const double psi = kz < (g2 + 1) / 2 ? st_top_psi(b, cx, cy, kz) : st_top_psi(b, -cx, -cy, g2 - kz);
  kh[idx] = psi * c * g0 * g1 * g2;
}

// PME's Fourier kernel (dALp1, pme0.c:473-503):
// khat(k) = exp(-pi^2 |m|^2 / beta^2) / (pi V |m|^2) cL_x cL_y cL_z M^3,
// m = Ai^T k over the centred frequencies, khat(0) = 0.
__global__ __launch_bounds__(kStBS) void k_stencil_top_pme(StencilBox b,
                                                           const double *__restrict__ cL,
                                                           double *__restrict__ kh) {
  const int g0 = b.M[1][0], g1 = b.M[1][1], g2 = b.M[1][2];
  const long idx = (long)blockIdx.x * kStBS + threadIdx.x;
  if (idx >= (long)g0 * g1 * g2) return;
  const int i = (int)(idx / ((long)g1 * g2)), j = (int)((idx / g2) % g1), k = (int)(idx % g2);
  const int kv[3] = {i <= (g0 - 1) / 2 ? i : i - g0, j <= (g1 - 1) / 2 ? j : j - g1,
                     k <= (g2 - 1) / 2 ? k : k - g2};
  double m2 = 0.0;
  for (int a = 0; a < 3; ++a) {
    const double v = b.Ai[0][a] * kv[0] + b.Ai[1][a] * kv[1] + b.Ai[2][a] * kv[2];
    m2 += v * v;
  }
  const double pi = 3.14159265358979323846;
  const double c = cL[i] * cL[g0 + j] * cL[g0 + g1 + k];
  kh[idx] = m2 == 0.0 ? 0.0
                      : exp(-pi * pi * m2 / (b.beta * b.beta)) / (pi * b.detA * m2) * c *
                            ((double)g0 * g1 * g2);
}

unsigned st_blocks(long n) { return (unsigned)((n + kStBS - 1) / kStBS); }

}  // namespace

void launch_stencil_level(const StencilBox &b, int l, const double *w0, const double *w1,
                          const double *w2, double *tmp_a, double *tmp_b, double *khat,
                          hipStream_t s) {
  const int S0 = b.S[l][0], S1 = b.S[l][1], S2 = b.S[l][2];
  const unsigned nb = st_blocks((long)S0 * S1 * S2);
  k_stencil_kappa<<<nb, kStBS, 0, s>>>(b, l, tmp_a);
  // z, y, x as in kappaA (forcefield0.c:995-1031)
  k_stencil_omega<<<nb, kStBS, 0, s>>>(tmp_a, tmp_b, w2, S0, S1, S2, 2, b.M[l][2]);
  k_stencil_omega<<<nb, kStBS, 0, s>>>(tmp_b, tmp_a, w1, S0, S1, S2, 1, b.M[l][1]);
  k_stencil_omega<<<nb, kStBS, 0, s>>>(tmp_a, khat, w0, S0, S1, S2, 0, b.M[l][0]);
}

void launch_stencil_top(const StencilBox &b, const double *cL, double *khat, hipStream_t s) {
  const long n = (long)b.M[b.L][0] * b.M[b.L][1] * b.M[b.L][2];
  k_stencil_top<<<st_blocks(n), kStBS, 0, s>>>(b, cL, khat);
}

void launch_stencil_top_pme(const StencilBox &b, const double *cL, double *khat, hipStream_t s) {
  const long n = (long)b.M[1][0] * b.M[1][1] * b.M[1][2];
  k_stencil_top_pme<<<st_blocks(n), kStBS, 0, s>>>(b, cL, khat);
}

}  // namespace msm
