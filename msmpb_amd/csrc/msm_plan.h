// msm_plan.h -- build-time state of one MSM evaluation plan (host, FP64).
//
// Holds what hkayabilisim/msmpb computes in FF_build / FF_rebuild
// (code/forcefield0.c:271-680): the resolved parameters, the B-spline,
// softening and two-scale tables, the anti-blurring weights and the level
// stencils.  Grids are described in periodic (unpadded) form: level l has
// dims M[l][0..2], x-major with z fastest.
#pragma once

#include <string>
#include <vector>

namespace msm {

constexpr int kMaxLevel = 9;   // the reference's time_* arrays hold 10 levels

// Summation method of a plan.  kPme is the reference's smooth particle-mesh
// Ewald comparator (code/pme0.c, code/pme1.c): one grid level (L = 1) with
// the Ewald real-space term erfc(beta r)/r and a Gaussian Fourier kernel.
enum Method { kMsm = 0, kPme = 1 };

struct Plan {
  int method = kMsm;
  double beta = 0, tolDir = 0; // PME: Ewald splitting, erfc(beta a0) = tolDir a0
  int N = 0;
  double A[3][3] = {}, Ai[3][3] = {}, detA = 0;
  bool ortho = true;           // box matrix diagonal
  int nu = 0, L = 0;
  bool fft = true;             // FF_set_FFT: top level by FFT (else the DFT build's rules)
  int M[kMaxLevel + 1][3] = {};  // level grid dims, l = 1..L
  int S[kMaxLevel + 1][3] = {};  // stencil dims, l = 1..L-1 (s_l = min(dmax, M_l))
  double aCut[kMaxLevel + 1] = {};
  double relCutoff = 0;
  int nLim = 0;
  double estErr = 1e-3;
  double tau[10] = {};         // softening Taylor coefficients tau_0..tau_{nu-1}
  double Q[50] = {};           // B-spline pieces: (nu/2) x nu, piece j in powers of (t - j)
  double J[11] = {};           // two-scale stencil J[n + nu/2], n = -nu/2..nu/2
  double gamma0 = 0;           // gamma(0) = tau(-1)
  double coeff1 = 0, coeff2 = 0;
  // Image range of the level kernels kappa_l: images p with
  // |s_a - p_a| <= a_l * kas[a] are summed.  Every image inside a_l
  // satisfies it for kas[a] >= |row a of A^-1|.  The reference passes the
  // column norms of A^-1 (forcefield0.c:618-620, 670), which equal the row
  // norms for orthorhombic boxes but can drop images of a triclinic one;
  // complete_images = false reproduces that exactly (the as-is layout),
  // true takes max(row, column) norm, so every box the reference's range
  // already covers gets its numbers bit for bit.
  bool complete_images = true;
  double kas[3] = {};
  std::vector<double> omegap[kMaxLevel + 1][3];
  std::vector<double> cL[3];
  std::vector<double> khat[kMaxLevel + 1];  // l < L: S^3 real-space; l = L: M_L^3 Fourier
};

struct BuildRequest {
  int N = 0;
  const double *q = nullptr;   // for the a-priori error model (sum q^2)
  double edges[3][3] = {};
  int nu = -1, L = -1, Mtop[3] = {-1, -1, -1};
  double a0 = -1, estErr = -1;
  int method = kMsm;
  double tolDir = -1;          // PME only
  bool complete_images = true; // level-kernel image range (Plan::kas)
  bool fft = true;             // FF_set_FFT (forcefield0.c:209-215, 247-316)
};

// Resolves parameters with the reference's rules (FF_build of
// forcefield0.c, or of pme0.c for method kPme) and builds every table and
// stencil.  Returns false with a message for inputs outside the domain the
// reference computes correctly.
bool build_plan(const BuildRequest &req, Plan &plan, std::string &err);

// Recomputes the box-dependent parts (constants and stencils) for a new box
// with the parameters of an existing plan (FF_rebuild).  With stencils =
// false the level and top-level stencils (khat) are left empty for the GPU
// to generate; omega', cL and the constants are still computed.
bool rebuild_plan(Plan &plan, const double edges[3][3], std::string &err, bool stencils = true);

// Sets plan.kas from the box and plan.complete_images.
void kappa_image_bound(Plan &plan);

// True where the reference's own image rules (the as-is layout) hold every
// image inside the cutoff: an orthorhombic box, or a cutoff below half the
// smallest face-to-face width.
bool reference_images_hold(const Plan &plan);

// Series coefficients b_0..b_6 of t^2 psi(t) of the top-level kernel.
void top_series(int nu, double b[7]);

// A-priori relative force error estimates (forcefield0.c:521-587).
double err_estimate(const Plan &plan, const double *q, int nu);
double delta_f_estimate(const Plan &plan, const double *q);
double invert3(const double a[3][3], double ai[3][3]);

}  // namespace msm
