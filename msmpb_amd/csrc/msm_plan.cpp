// msm_plan.cpp -- parameter resolution and FP64 build-time stencils.
//
// Host-side control plane of the MSM library: what hkayabilisim/msmpb does
// in FF_build / FF_rebuild (code/forcefield0.c:271-680) before any energy
// evaluation.  Everything here is FP64 and runs once per box; the per-call
// evaluation runs on the GPU (msm_kernels.hip).
#include "msm_plan.h"

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>

namespace msm {
namespace {

const double kPi = 3.14159265358979323846;

double row_norm(const double m[3][3], int r) {
  return std::sqrt(m[r][0] * m[r][0] + m[r][1] * m[r][1] + m[r][2] * m[r][2]);
}
double col_norm(const double m[3][3], int c) {
  return std::sqrt(m[0][c] * m[0][c] + m[1][c] * m[1][c] + m[2][c] * m[2][c]);
}

// Smallest M' >= M whose prime factors are all <= 7 (the FFT-size rule of
// forcefield0.c:219-245).
int next_7smooth(int M) {
  for (int c = M < 1 ? 1 : M;; ++c) {
    int r = c;
    for (int p : {2, 3, 5, 7})
      while (r % p == 0) r /= p;
    if (r == 1) return c;
  }
}

int64_t binom(int n, int k) {
  if (k < 0 || k > n) return 0;
  int64_t r = 1;
  for (int i = 1; i <= k; ++i) r = r * (n - k + i) / i;
  return r;
}

int64_t ipow(int64_t b, int e) {
  int64_t r = 1;
  while (e-- > 0) r *= b;
  return r;
}

// Cardinal B-spline of order nu on [0, nu]:
//   Q_nu(x) = 1/(nu-1)! sum_k (-1)^k C(nu,k) (x-k)_+^(nu-1).
// Piece j (x in [j, j+1]) in powers of u = x - j has the coefficient
//   C(nu-1,p)/(nu-1)! * sum_{k<=j} (-1)^k C(nu,k) (j-k)^(nu-1-p)
// of u^p; the integer sum is exact, so each coefficient is rounded once.
void bspline_pieces(int nu, double *Q) {
  double fact = 1;
  for (int i = 2; i < nu; ++i) fact *= i;
  for (int j = 0; j < nu / 2; ++j)
    for (int p = 0; p < nu; ++p) {
      int64_t s = 0;
      for (int k = 0; k <= j; ++k) s += (k % 2 ? -1 : 1) * binom(nu, k) * ipow(j - k, nu - 1 - p);
      Q[j * nu + p] = (double)(binom(nu - 1, p) * s) / fact;
    }
}

// Q_nu at an integer knot, exact sum then one rounding.
double bspline_knot(int nu, int x) {
  double fact = 1;
  for (int i = 2; i < nu; ++i) fact *= i;
  int64_t s = 0;
  for (int k = 0; k <= nu && k < x; ++k) s += (k % 2 ? -1 : 1) * binom(nu, k) * ipow(x - k, nu - 1);
  return (double)s / fact;
}

// gamma(rho) = tau(rho^2 - 1) for rho < 1 (Horner over tau).
double soft(const Plan &p, double s) {
  double g = p.tau[p.nu - 1];
  for (int k = p.nu - 2; k >= 0; --k) g = p.tau[k] + s * g;
  return g;
}

// Polynomial factors of the top-level Fourier kernel (forcefield0.c:820-836):
// psi(t) * t^(2 nu) = S(t^2) sin(t)/t - C(t^2) cos(t), coefficients in t^2.
void top_poly(int nu, double *S, double *C, int &n) {
  static const double s4[] = {15.0, -6.0}, c4[] = {15.0, -1.0};
  static const double s6[] = {945.0, -420.0, 15.0}, c6[] = {945.0, -105.0, 1.0};
  static const double s8[] = {135135.0, -62370.0, 3150.0, -28.0};
  static const double c8[] = {135135.0, -17325.0, 378.0, -1.0};
  static const double s10[] = {34459425.0, -16216200.0, 945945.0, -13860.0, 45.0};
  static const double c10[] = {34459425.0, -4729725.0, 135135.0, -990.0, 1.0};
  const double *s = nu == 4 ? s4 : nu == 6 ? s6 : nu == 8 ? s8 : s10;
  const double *c = nu == 4 ? c4 : nu == 6 ? c6 : nu == 8 ? c8 : c10;
  n = nu / 2;
  for (int i = 0; i < n; ++i) S[i] = s[i], C[i] = c[i];
}

// Taylor coefficients b_i of t^2 psi(t) = sum_i b_i t^(2i), i = 0..6, derived
// from the sin/cos series of the closed form in extended precision (the
// first nu orders of the numerator cancel exactly).  They replace the
// reference's table of reciprocals (forcefield0.c:795-800).
void top_taylor(int nu, double *b) {
  double S[5], C[5];
  int n;
  top_poly(nu, S, C, n);
  for (int i = 0; i < 7; ++i) {
    int j = nu - 1 + i;  // numerator order t^(2j); psi t^2 = N / t^(2nu-2)
    long double a = 0;
    for (int m = 0; m < n && m <= j; ++m) {
      int d = j - m;
      long double fs = 1, fc = 1;  // (2d+1)! and (2d)!
      for (int q = 2; q <= 2 * d + 1; ++q) fs *= q;
      for (int q = 2; q <= 2 * d; ++q) fc *= q;
      long double sg = (d % 2) ? -1.0L : 1.0L;
      a += (long double)S[m] * sg / fs - (long double)C[m] * sg / fc;
    }
    b[i] = (double)a;
  }
}

double top_psi(const Plan &p, const double *taylor, int kx, int ky, int kz) {
  // |A^{-T} k| for the integer wave vector k (forcefield0.c:806-809)
  double ax = p.Ai[0][0] * kx + p.Ai[1][0] * ky + p.Ai[2][0] * kz;
  double ay = p.Ai[0][1] * kx + p.Ai[1][1] * ky + p.Ai[2][1] * kz;
  double az = p.Ai[0][2] * kx + p.Ai[1][2] * ky + p.Ai[2][2] * kz;
  double k2 = ax * ax + ay * ay + az * az;
  if (k2 == 0) return 0.0;
  double aL = p.aCut[p.L];
  double t = kPi * std::sqrt(k2) * aL, t2 = t * t;
  double psi;
  if (t <= kPi) {  // series branch for t <= tmax*pi/4 with tmax = 4 and 7 terms
    double s = 0, tp = 1;
    for (int i = 0; i < 7; ++i) s += tp * taylor[i], tp *= t2;
    psi = s / t2;
  } else {
    double S[5], C[5];
    int n;
    top_poly(p.nu, S, C, n);
    double sf = 0, cf = 0, tp = 1;
    for (int i = 0; i < n; ++i) sf += S[i] * tp, cf += C[i] * tp, tp *= t2;
    psi = (sf * (std::sin(t) / t) - cf * std::cos(t)) / std::pow(t, 2.0 * p.nu);
  }
  double dfac = 1;  // (2 nu - 1)!!
  for (int i = 3; i <= 2 * p.nu - 1; i += 2) dfac *= i;
  return psi * (dfac * kPi * aL * aL / p.detA);
}

// Level kernel kappa_l at fractional offset s: images within a_l of
// gamma_{a_{l-1}} - gamma_{a_l} (forcefield0.c:1041-1081).
double kappa(const Plan &p, int l, const double s[3], const double as[3]) {
  double al = p.aCut[l], sum = 0;
  int lo[3], hi[3];
  for (int a = 0; a < 3; ++a) {
    lo[a] = (int)std::floor(s[a] - al * as[a] + 1.0);
    hi[a] = (int)std::ceil(s[a] + al * as[a] - 1.0);
  }
  for (int px = lo[0]; px <= hi[0]; ++px)
    for (int py = lo[1]; py <= hi[1]; ++py)
      for (int pz = lo[2]; pz <= hi[2]; ++pz) {
        double f[3] = {s[0] - px, s[1] - py, s[2] - pz};
        double r2 = 0;
        for (int a = 0; a < 3; ++a) {
          double v = p.A[a][0] * f[0] + p.A[a][1] * f[1] + p.A[a][2] * f[2];
          r2 += v * v;
        }
        double rho = std::sqrt(r2) / al;
        if (rho >= 1.0) continue;
        double g_l = soft(p, rho * rho - 1.0) / al;
        double rho2 = 2.0 * rho;  // r / a_{l-1}
        double g_lm1 = (rho2 >= 1.0 ? 1.0 / rho2 : soft(p, rho2 * rho2 - 1.0)) * (2.0 / al);
        sum += g_lm1 - g_l;
      }
  return sum;
}

// Offset map of the separable anti-blurring passes (forcefield0.c:1002):
// displacement d on a periodic axis of M points, folded to [-M/2, M/2).
inline int fold(int d, int M) { return (d + (3 * M) / 2) % M - M / 2; }

void build_omegap(Plan &p) {
  int nu = p.nu, h = nu / 2;
  double Phi[5];
  for (int n = 0; n < h; ++n) Phi[n] = bspline_knot(nu, h + n);
  int opLim = 2 * p.nLim;
  for (int l = p.L; l >= 1; --l)
    for (int a = 0; a < 3; ++a) {
      int Ma = p.M[l][a];
      std::vector<double> c(Ma);
      for (int k = 0; k < Ma; ++k) {
        double th = 2.0 * kPi * (double)k / (double)Ma, v = Phi[0];
        for (int n = 1; n < h; ++n) v += 2.0 * Phi[n] * std::cos((double)n * th);
        c[k] = 1.0 / (v * v);
      }
      int D = (Ma / 2 > opLim) ? opLim + 1 : Ma / 2 + 1;
      std::vector<double> w(D);
      for (int m = 0; m < D; ++m) {
        double s = c[0];
        for (int k = 1; k <= (Ma - 1) / 2; ++k)
          s += 2.0 * c[k] * std::cos(2.0 * kPi * (double)(k * m) / (double)Ma);
        if (Ma % 2 == 0) s += c[Ma / 2] * std::cos(kPi * (double)m);
        w[m] = s / (double)Ma;
      }
      p.omegap[l][a] = w;
      if (l == p.L) p.cL[a] = c;
    }
}

void build_top_kernel(Plan &p) {
  double taylor[7];
  top_taylor(p.nu, taylor);
  const int *g = p.M[p.L];
  std::vector<double> kh((size_t)g[0] * g[1] * g[2], 0.0);
  for (int kx = 0; kx < g[0]; ++kx) {
    int cx = kx < (g[0] + 1) / 2 ? kx : kx - g[0];
    for (int ky = 0; ky < g[1]; ++ky) {
      // the reference folds y with a real-valued midpoint (forcefield0.c:872)
      int cy = ky < (g[1] + 1) / 2.0 ? ky : ky - g[1];
      for (int kz = 0; kz < g[2]; ++kz) {
        double c = p.cL[0][kx] * p.cL[1][ky] * p.cL[2][kz];
        double psi = kz < (g[2] + 1) / 2 ? top_psi(p, taylor, cx, cy, kz)
                                         : top_psi(p, taylor, -cx, -cy, g[2] - kz);
        kh[((size_t)kx * g[1] + ky) * g[2] + kz] = psi * c * g[0] * g[1] * g[2];
      }
    }
  }
  p.khat[p.L] = std::move(kh);
}

void build_level_kernel(Plan &p, int l, const double as[3]) {
  const int *g = p.M[l], *S = p.S[l];
  int sx = S[0], sy = S[1], sz = S[2];
  size_t n = (size_t)sx * sy * sz;
  std::vector<double> kap(n), t1(n), t2(n), out(n);
  auto at = [&](int i, int j, int k) { return ((size_t)i * sy + j) * sz + k; };
  for (int i = 0; i < sx; ++i)
    for (int j = 0; j < sy; ++j)
      for (int k = 0; k < sz; ++k) {
        double s[3] = {(double)(i - sx / 2) / g[0], (double)(j - sy / 2) / g[1],
                       (double)(k - sz / 2) / g[2]};
        kap[at(i, j, k)] = kappa(p, l, s, as);
      }
  const std::vector<double> *w = p.omegap[l];
  for (int i = 0; i < sx; ++i)  // z pass
    for (int j = 0; j < sy; ++j)
      for (int k1 = 0; k1 < sz; ++k1) {
        double acc = 0;
        for (int k0 = 0; k0 < sz; ++k0) acc += w[2][std::abs(fold(k0 - k1, g[2]))] * kap[at(i, j, k0)];
        t1[at(i, j, k1)] = acc;
      }
  for (int i = 0; i < sx; ++i)  // y pass
    for (int j1 = 0; j1 < sy; ++j1)
      for (int k = 0; k < sz; ++k) {
        double acc = 0;
        for (int j0 = 0; j0 < sy; ++j0) acc += w[1][std::abs(fold(j0 - j1, g[1]))] * t1[at(i, j0, k)];
        t2[at(i, j1, k)] = acc;
      }
  for (int i1 = 0; i1 < sx; ++i1)  // x pass
    for (int j = 0; j < sy; ++j)
      for (int k = 0; k < sz; ++k) {
        double acc = 0;
        for (int i0 = 0; i0 < sx; ++i0) acc += w[0][std::abs(fold(i0 - i1, g[0]))] * t2[at(i0, j, k)];
        out[at(i1, j, k)] = acc;
      }
  p.khat[l] = std::move(out);
}

bool box_dependent(Plan &p, std::string &err, bool stencils = true) {
  // image-range vector: the reference's column norms of A^-1, or with the
  // complete image range max(row, column) norm (Plan::kas)
  kappa_image_bound(p);
  // constant-energy coefficients (forcefield0.c:612-643): own images of
  // each charge inside the cutoff, and the neutralising background.  A
  // lattice vector A k inside a0 has |k_a| < a0 |row a of A^-1|; the column
  // norms the reference bounds k with can miss some where the cutoff
  // reaches a face-to-face width of a triclinic box (none lie inside a0
  // below that width, so both ranges give the same sum there)
  double a0 = p.aCut[0];
  int lim[3];
  for (int a = 0; a < 3; ++a) lim[a] = (int)std::ceil(a0 * p.kas[a] - 1.0);
  p.coeff1 = 0;
  for (int x = -lim[0]; x <= lim[0]; ++x)
    for (int y = -lim[1]; y <= lim[1]; ++y)
      for (int z = -lim[2]; z <= lim[2]; ++z) {
        if (x == 0 && y == 0 && z == 0) continue;
        double r2 = 0;
        for (int a = 0; a < 3; ++a) {
          double v = p.A[a][0] * x + p.A[a][1] * y + p.A[a][2] * z;
          r2 += v * v;
        }
        if (r2 >= a0 * a0) continue;
        p.coeff1 += 1.0 / std::sqrt(r2) - soft(p, r2 / (a0 * a0) - 1.0) / a0;
      }
  double aL = p.aCut[p.L];
  p.coeff2 = kPi * aL * aL / ((4 * p.nu + 2) * p.detA);
  build_omegap(p);
  if (stencils) {
    build_top_kernel(p);
    for (int l = p.L - 1; l >= 1; --l) build_level_kernel(p, l, p.kas);
  }
  (void)err;
  return true;
}

// Box shape: orthorhombic (diagonal A) or triclinic.  Below half the
// smallest face-to-face width w = 1 / |row a of A^-1| the reference's pair
// rule (the image nearest in fractional coordinates, then the images
// p in [-plim, plim]^3 around it with plim from the column norms of A^-1;
// neighborlist and partcl2partcl of forcefield0.c and forcefield1.c) and
// the sum over every image inside the cutoff coincide: an image within
// a0 < w/2 has every |s_a| < 1/2.  At larger cutoffs in a triclinic box the
// two part: a pair whose fractionally nearest image lies beyond the cutoff
// is not listed although another image may lie inside it, and column norms
// below the row norms leave images outside the plim box.  The MSM method
// with the complete image range (complete_images, the default layout) sums
// every image inside the cutoff at any cutoff -- the pair kernel's cell
// columns, lattice shifts and self-image constant hold every image for any
// box -- so only the reference's own rules (the as-is layout) and the PME
// comparator keep the bound.
static double min_width(const Plan &p) {
  double wmin = INFINITY;
  for (int a = 0; a < 3; ++a) wmin = std::fmin(wmin, 1.0 / row_norm(p.Ai, a));
  return wmin;
}

bool check_box(Plan &p, std::string &err) {
  p.ortho = true;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      if (i != j && p.A[i][j] != 0.0) p.ortho = false;
  if (p.ortho || (p.method == kMsm && p.complete_images)) return true;
  const double wmin = min_width(p);
  if (!(p.aCut[0] < 0.5 * wmin)) {
    char buf[256];
    std::snprintf(buf, sizeof buf,
                  "triclinic box: the cutoff %g must be below half the smallest face-to-face "
                  "width (%g) for %s", p.aCut[0], wmin,
                  p.method == kMsm ? "the reference's image rules (as-is layout)" : "the PME method");
    err = buf;
    return false;
  }
  return true;
}

// ---- PME (code/pme0.c) ----------------------------------------------------
// Ewald splitting: erfc(b) = c for b = beta a0, by Newton's method from
// b = 0 until the residual stops shrinking (pme0.c:209-224).
double ewald_root(double c) {
  double b = 0.0, res = std::erfc(b) - c, old = 2.0 * res;
  while (std::fabs(res) < std::fabs(old)) {
    b -= res / (-2.0 / std::sqrt(kPi) * std::exp(-b * b));
    old = res;
    res = std::erfc(b) - c;
  }
  return b;
}

// Constants and the Fourier kernel for the current box (FF_rebuild and
// dALp1, pme0.c:365-412 and :473-503):
//   coeff1 = sum over the own images |Ap| < a0 of erfc(beta r)/r - 2 beta/sqrt(pi),
//   coeff2 = pi / (beta^2 V),
//   khat(k) = exp(-pi^2 |m|^2 / beta^2) / (pi V |m|^2) cL_x cL_y cL_z M^3
// with m = Ai^T k over the centred frequencies and khat(0) = 0.
void pme_box_dependent(Plan &p, bool stencils) {
  const double a0 = p.aCut[0], b = p.beta;
  int lim[3];
  for (int a = 0; a < 3; ++a) lim[a] = (int)std::ceil(a0 * col_norm(p.Ai, a) - 1.0);
  p.coeff1 = -2.0 * b / std::sqrt(kPi);
  for (int x = -lim[0]; x <= lim[0]; ++x)
    for (int y = -lim[1]; y <= lim[1]; ++y)
      for (int z = -lim[2]; z <= lim[2]; ++z) {
        if (!x && !y && !z) continue;
        double r2 = 0;
        for (int a = 0; a < 3; ++a) {
          double v = p.A[a][0] * x + p.A[a][1] * y + p.A[a][2] * z;
          r2 += v * v;
        }
        if (r2 >= a0 * a0) continue;
        p.coeff1 += std::erfc(b * std::sqrt(r2)) / std::sqrt(r2);
      }
  p.coeff2 = kPi / (b * b * p.detA);
  p.gamma0 = 0.0;
  build_omegap(p);  // cL of the single level
  if (!stencils) return;
  const int *g = p.M[1];
  const double n3 = (double)g[0] * g[1] * g[2];
  std::vector<double> kh((size_t)n3, 0.0);
  for (int i = 0; i < g[0]; ++i)
    for (int j = 0; j < g[1]; ++j)
      for (int k = 0; k < g[2]; ++k) {
        const int kv[3] = {i <= (g[0] - 1) / 2 ? i : i - g[0], j <= (g[1] - 1) / 2 ? j : j - g[1],
                           k <= (g[2] - 1) / 2 ? k : k - g[2]};
        double m2 = 0;
        for (int a = 0; a < 3; ++a) {
          const double v = p.Ai[0][a] * kv[0] + p.Ai[1][a] * kv[1] + p.Ai[2][a] * kv[2];
          m2 += v * v;
        }
        if (m2 == 0.0) continue;
        kh[((size_t)i * g[1] + j) * g[2] + k] = std::exp(-kPi * kPi * m2 / (b * b)) /
                                                (kPi * p.detA * m2) * p.cL[0][i] * p.cL[1][j] *
                                                p.cL[2][k] * n3;
      }
  p.khat[1] = std::move(kh);
}

// FF_build of the PME comparator (pme0.c:192-322): order 6, cutoff 8 A and
// direct tolerance 1e-5 by default; the grid from the reference's spacing
// table when none is given (pme0.c:226-263, orders 4 and 6, cutoffs 6-10 A).
bool build_pme(const BuildRequest &r, Plan &p, std::string &err) {
  char buf[256];
  p.method = kPme;
  p.L = 1;
  const int nu = r.nu > 0 ? r.nu : 6;
  if (nu != 4 && nu != 6 && nu != 8 && nu != 10) {
    std::snprintf(buf, sizeof buf, "orderAcc must be 4, 6, 8 or 10 (got %d)", nu);
    err = buf;
    return false;
  }
  p.nu = nu;
  const double a0 = r.a0 > 0 ? r.a0 : 8.0;
  if (!std::isfinite(a0)) { err = "cutoff must be finite"; return false; }
  p.aCut[0] = a0;
  p.aCut[1] = 2.0 * a0;
  p.tolDir = r.tolDir > 0 ? r.tolDir : 1e-5;
  if (!(p.tolDir * a0 < 1.0)) { err = "tolDir * cutoff must be below 1"; return false; }
  p.beta = ewald_root(p.tolDir * a0) / a0;
  if (r.Mtop[0] > 0) {
    for (int a = 0; a < 3; ++a) p.M[1][a] = r.Mtop[a];
  } else {
    static const double href[2][9] = {{0.70, 0.77, 0.83, 0.91, 1.00, 1.11, 1.21, 1.33, 1.48},
                                      {1.05, 1.18, 1.29, 1.43, 1.54, 1.67, 1.82, 2.00, 2.22}};
    const int i = (nu - 4) / 2, j = (int)std::lround((a0 - 6.0) / 0.5);
    if (i > 1 || j < 0 || j > 8) {
      std::snprintf(buf, sizeof buf,
                    "no default PME grid for order %d and cutoff %g (set topGridDim)", nu, a0);
      err = buf;
      return false;
    }
    const double hstar = std::cbrt(p.detA / r.N);
    const double beta_ref = ewald_root(1e-5 * a0) / a0;
    const double h = std::pow(p.tolDir / (1e-5 * hstar), 1.0 / nu) * (beta_ref / p.beta) * href[i][j];
    for (int a = 0; a < 3; ++a) p.M[1][a] = (int)std::ceil(row_norm(p.A, a) / h);
  }
  for (int a = 0; a < 3; ++a)
    if (p.M[1][a] < nu) {
      std::snprintf(buf, sizeof buf, "PME grid dimension %d is below the order %d", p.M[1][a], nu);
      err = buf;
      return false;
    }
  bspline_pieces(nu, p.Q);
  pme_box_dependent(p, true);
  return true;
}

}  // namespace

bool reference_images_hold(const Plan &p) { return p.ortho || p.aCut[0] < 0.5 * min_width(p); }

double invert3(const double a[3][3], double ai[3][3]) {
  double adj[3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      int i1 = (i + 1) % 3, i2 = (i + 2) % 3, j1 = (j + 1) % 3, j2 = (j + 2) % 3;
      adj[i][j] = a[j1][i1] * a[j2][i2] - a[j2][i1] * a[j1][i2];
    }
  double det = a[0][0] * adj[0][0] + a[0][1] * adj[1][0] + a[0][2] * adj[2][0];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) ai[i][j] = adj[i][j] / det;
  return std::fabs(det);
}

double err_estimate(const Plan &p, const double *q, int nu) {
  double ax = row_norm(p.A, 0), ay = row_norm(p.A, 1), az = row_norm(p.A, 2);
  double M1[3];
  for (int a = 0; a < 3; ++a) M1[a] = p.M[p.L][a] * std::pow(2.0, p.L - 1);
  double Q2 = 0;
  for (int i = 0; i < p.N; ++i) Q2 += q[i] * q[i];
  double Fref = (Q2 / (double)p.N) / std::pow(p.detA / (double)p.N, 5. / 6.);
  double h1 = std::fmax(std::fmax(ax / M1[0], ay / M1[1]), az / M1[2]);
  double dF = Q2 * (1 - std::pow(4.0, -p.L)) * std::pow(h1, nu - 1) * std::pow(p.detA, -1. / 3.) *
              std::pow(p.N, -2.0 / 3.0);
  return dF / Fref;
}

double delta_f_estimate(const Plan &p, const double *q) {
  static const double C[4] = {0.825169, 5.16701, 73.6508, 1977.23};
  double ax = row_norm(p.A, 0), ay = row_norm(p.A, 1), az = row_norm(p.A, 2);
  double M1[3];
  for (int a = 0; a < 3; ++a) M1[a] = p.M[p.L][a] * std::pow(2.0, p.L - 1);
  double Q2 = 0;
  for (int i = 0; i < p.N; ++i) Q2 += q[i] * q[i];
  double h1 = std::fmax(std::fmax(ax / M1[0], ay / M1[1]), az / M1[2]);
  return C[(p.nu - 4) / 2] * Q2 * (1 - std::pow(4.0, -p.L)) * std::pow(h1, p.nu - 1) *
         std::pow(p.detA, -1. / 3.) * std::pow(p.N, -2.0 / 3.0) / std::pow(p.aCut[0], p.nu);
}

void kappa_image_bound(Plan &p) {
  for (int a = 0; a < 3; ++a) {
    const double c = col_norm(p.Ai, a);
    p.kas[a] = p.complete_images ? std::fmax(c, row_norm(p.Ai, a)) : c;
  }
}

bool build_plan(const BuildRequest &r, Plan &p, std::string &err) {
  char buf[256];
  p = Plan();
  p.N = r.N;
  p.complete_images = r.complete_images;
  if (r.N <= 0) { err = "N must be positive"; return false; }
  std::memcpy(p.A, r.edges, sizeof p.A);
  p.detA = invert3(p.A, p.Ai);
  if (!(p.detA > 0) || !std::isfinite(p.detA)) { err = "box matrix is singular"; return false; }
  if (r.method == kPme) return build_pme(r, p, err) && check_box(p, err);
  p.estErr = r.estErr > 0 ? r.estErr : 1e-3;
  p.L = r.L > 0 ? r.L : 1;
  p.fft = r.fft;
  // determineM (forcefield0.c:247-269): about eta grid points per atom on the
  // finest level, equal spacing on every axis (eta 5.7 for one level and 0.83
  // for more with the FFT, 1 for the DFT build)
  auto determine_M = [&](int L, int Mt[3]) {
    const double eta = p.fft ? (L >= 2 ? 0.83 : 5.7) : 1.0;
    double ax = row_norm(p.A, 0), ay = row_norm(p.A, 1), az = row_norm(p.A, 2), av[3] = {ax, ay, az};
    for (int a = 0; a < 3; ++a)
      Mt[a] = (int)std::ceil(std::pow(2., 1. - L) * std::pow(eta * r.N, 1. / 3.) * av[a] *
                             std::pow(ax * ay * az, -1. / 3.));
  };
  int Mt[3];
  if (r.Mtop[0] > 0) {
    for (int a = 0; a < 3; ++a) Mt[a] = r.Mtop[a];
  } else if (!p.fft) {
    // DFT build (forcefield0.c:292-307): the number of levels becomes the
    // first whose top grid has at most sqrt(N) points
    for (int l = 1;; ++l) {
      determine_M(l, Mt);
      if ((double)Mt[0] * Mt[1] * Mt[2] <= std::sqrt((double)r.N)) {
        p.L = l;
        break;
      }
    }
  } else {
    determine_M(p.L, Mt);
  }
  if (p.L > kMaxLevel) {
    std::snprintf(buf, sizeof buf, "maxLevel %d exceeds %d", p.L, kMaxLevel);
    err = buf;
    return false;
  }
  for (int a = 0; a < 3; ++a) {
    if (Mt[a] < 1) { err = "topGridDim must be positive"; return false; }
    // the FFT build rounds up to 7-smooth sizes (forcefield0.c:309-316)
    p.M[p.L][a] = p.fft ? next_7smooth(Mt[a]) : Mt[a];
  }
  for (int l = p.L - 1; l >= 1; --l)
    for (int a = 0; a < 3; ++a) p.M[l][a] = 2 * p.M[l + 1][a];
  // cutoff and order from the a-priori error model (forcefield0.c:319-348)
  double a0 = r.a0;
  int nu = r.nu;
  if (a0 <= 0) {
    if (nu <= 0) {
      a0 = INFINITY;
      for (int t = 4; t <= 10; t += 2) a0 = std::fmin(a0, std::pow(err_estimate(p, r.q, t) / p.estErr, 1.0 / t));
    } else {
      a0 = std::pow(err_estimate(p, r.q, nu) / p.estErr, 1.0 / nu);
    }
  }
  if (nu <= 0) {
    double best = INFINITY;
    nu = 4;
    for (int t = 4; t <= 10; t += 2) {
      double e = err_estimate(p, r.q, t) / std::pow(a0, t);
      if (e < best) best = e, nu = t;
    }
  }
  if (nu != 4 && nu != 6 && nu != 8 && nu != 10) {
    std::snprintf(buf, sizeof buf, "orderAcc must be 4, 6, 8 or 10 (got %d)", nu);
    err = buf;
    return false;
  }
  if (!(a0 > 0) || !std::isfinite(a0)) { err = "cutoff must be positive"; return false; }
  p.nu = nu;
  p.aCut[0] = a0;
  for (int l = 1; l <= p.L; ++l) p.aCut[l] = 2.0 * p.aCut[l - 1];
  int nbar = 0;
  for (int a = 0; a < 3; ++a)
    nbar = std::max(nbar, (int)std::ceil(p.aCut[p.L] * p.M[p.L][a] * row_norm(p.Ai, a)));
  p.relCutoff = nbar;
  p.nLim = (int)std::ceil(p.relCutoff - 1.0);
  int dmax = 2 * p.nLim + 1;
  for (int l = 1; l < p.L; ++l)
    for (int a = 0; a < 3; ++a) {
      p.S[l][a] = std::min(dmax, p.M[l][a]);
      // the reference pads intermediate grids by max(s, nu-1) but fills and
      // restricts with max(s, nu): outside s >= nu it corrupts its heap
      if (p.S[l][a] < nu) {
        std::snprintf(buf, sizeof buf,
                      "level %d stencil extent %d is below the order %d (grid too coarse for "
                      "this cutoff); use fewer levels or a finer top grid",
                      l, p.S[l][a], nu);
        err = buf;
        return false;
      }
    }
  // tables (forcefield0.c:369-416)
  p.tau[0] = 1.0;
  for (int k = 1; k < nu; ++k) p.tau[k] = -(1.0 - 0.5 / (double)k) * p.tau[k - 1];
  bspline_pieces(nu, p.Q);
  for (int n = 0; n <= nu; ++n) p.J[n] = (double)binom(nu, n) * std::pow(2.0, 1 - nu);
  p.gamma0 = soft(p, -1.0);
  return check_box(p, err) && box_dependent(p, err);
}

bool rebuild_plan(Plan &p, const double edges[3][3], std::string &err, bool stencils) {
  std::memcpy(p.A, edges, sizeof p.A);
  p.detA = invert3(p.A, p.Ai);
  if (!(p.detA > 0)) { err = "box matrix is singular"; return false; }
  if (!check_box(p, err)) return false;
  for (int l = 1; l <= p.L; ++l) {
    for (int a = 0; a < 3; ++a) p.omegap[l][a].clear();
    p.khat[l].clear();
  }
  if (p.method == kPme) {
    pme_box_dependent(p, stencils);
    return true;
  }
  return box_dependent(p, err, stencils);
}

void top_series(int nu, double b[7]) { top_taylor(nu, b); }

}  // namespace msm
