// msm_engine.hip -- the C ABI of libmsmpb_hip.so: forcefield.h entry points
// (hkayabilisim/msmpb code/forcefield.h:70-112) and the device-pointer
// extensions of msmpb_hip.h, on top of the HIP kernel chain.
//
// One FF owns one Engine: the resolved plan and its device buffers
// (allocated once in FF_build, sized for N atoms); the two library streams
// of the device are shared.  FF_energy stages
// host positions and charges into HBM, runs the chain and copies forces
// back; the *_device entry points work on caller-owned device memory and
// are stream-ordered.
#include <hip/hip_runtime.h>

#include <chrono>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/msm_bridge.h"
#include "../../include/msmpb_hip.h"
// forcefield.h defines GNU max/min macros for C callers; not wanted here
#undef max
#undef min
#include "msm_kernels.h"
#include "msm_plan.h"

// The reference's callers define `struct O o`; keep programs that do not
// define it linkable.
__attribute__((weak)) struct O o = {(char *)"csCl"};

namespace {

thread_local std::string g_err;

int fail(int code, const std::string &msg) {
  g_err = msg;
  return code;
}

#define HIP_TRY(call)                                                                  \
  do {                                                                                 \
    hipError_t e_ = (call);                                                            \
    if (e_ != hipSuccess)                                                              \
      return fail(MSM_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_));     \
  } while (0)

// Named GPU intervals of one or more evaluations (start/stop events on the
// stream that runs the stage, so concurrent stages are timed separately).
struct Timing {
  bool on = false;
  struct Iv {
    std::string tag;
    hipEvent_t a, b;
  };
  std::vector<Iv> iv;
  std::vector<hipEvent_t> open_ev;
  int open(hipStream_t s) {
    if (!on) return -1;
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return -1;
    hipEventRecord(e, s);
    open_ev.push_back(e);
    return (int)open_ev.size() - 1;
  }
  void close(int k, const char *tag, hipStream_t s) {
    if (!on || k < 0) return;
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return;
    hipEventRecord(e, s);
    iv.push_back({tag, open_ev[k], e});
  }
  void clear() {
    for (auto &v : iv) hipEventDestroy(v.b);
    for (auto e : open_ev) hipEventDestroy(e);
    iv.clear();
    open_ev.clear();
  }
};

struct Engine {
  msm::Plan plan;
  msm::DevParams P{};
  msm::GridDims gd[msm::kMaxLevel + 1]{}, sd[msm::kMaxLevel + 1]{};
  int device = 0;
  hipStream_t stream = nullptr;    // the null stream: host-buffer entry points, FF_build
  hipStream_t stream2 = nullptr;   // library stream (shared, not owned): the pairs, beside the chain
  hipStream_t stream3 = nullptr;   // library stream: the level-1 stencil pass, beside the coarser levels
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  hipEvent_t ev_q1 = nullptr, ev_g1 = nullptr;  // q^1 ready; level-1 stencil pass done
  int nsten = 0;  // neighbour cell columns of the pair search
  // device buffers
  double *d_pos = nullptr, *d_q = nullptr, *d_F = nullptr, *d_phi = nullptr;
  double *d_out = nullptr, *d_qpart = nullptr, *d_scal = nullptr;
  double *d_epair = nullptr, *d_egrid = nullptr;
  int *d_cell_of = nullptr, *d_cnt = nullptr, *d_start = nullptr, *d_sidx = nullptr;
  // cell members by arrival (ncell rows of bcap), the overflow list and its
  // counts (binning -> scan -> placement)
  int *d_bucket = nullptr, *d_ovf = nullptr, *d_ovfn = nullptr;
  int bcap = 0;
  double *d_Fp = nullptr, *d_phip = nullptr;  // short-range parts of F and phi
  int *d_flags = nullptr;
  double4 *d_spos = nullptr;
  // per-atom spreading placements (packed support origin; t and q), sorted order
  unsigned *d_sbase = nullptr;
  double4 *d_st = nullptr;
  int4 *d_sten = nullptr;
  unsigned *d_count = nullptr;  // pair counts per pair-search cell
  double *d_ql[msm::kMaxLevel + 1] = {}, *d_el[msm::kMaxLevel + 1] = {};
  unsigned *d_qb[msm::kMaxLevel + 1] = {};  // q^l as bf16 hi/lo pairs (l < L)
  double *d_khat[msm::kMaxLevel + 1] = {};
  void *d_ksplit[msm::kMaxLevel + 1] = {};
  void *d_qsplit = nullptr;
  double *d_tw = nullptr, *d_topwork = nullptr;
  // FF_rebuild on the GPU: omega' and cL tables, stencil scratch (2 x S^3)
  double *d_stab = nullptr, *d_sscr = nullptr;
  size_t stab_cap = 0, sscr_cap = 0;
  // level stencils on MFMA, per precision (0: FP64, 1: BF16 hi/lo)
  msm::G2GArgs g2g[2]{};
  size_t g2g_lds[2] = {0, 0};
  double *d_part[2][msm::kMaxLevel + 1] = {};
  void *d_bfrag[2][msm::kMaxLevel + 1] = {};
  // per-stage events of the evaluations since the last collection
  Timing T;
  int timed_evals = 0;
  // forward state kept for a following adjoint
  bool kept = false;        // level-1 grid and atom ordering of the last forward
  bool kept_pairs = false;  // short-range F, phi of the last forward in d_Fp, d_phip
  const double *kept_pos = nullptr, *kept_q = nullptr;
  double last_terms[3] = {0, 0, 0};

  ~Engine() { release(); }
  void release() {
    T.clear();
    void *ptrs[] = {d_pos,  d_q,    d_F,    d_phi,   d_out,  d_qpart, d_scal,  d_epair,
                    d_egrid, d_cell_of, d_cnt, d_start, d_sidx, d_flags,
                    d_bucket, d_ovf, d_ovfn, d_Fp, d_phip,
                    d_spos, d_sten, d_count, d_qsplit, d_tw, d_topwork, d_stab, d_sscr,
                    d_sbase, d_st};
    for (void *p : ptrs)
      if (p) hipFree(p);
    for (int pr = 0; pr < 2; ++pr)
      for (int l = 0; l <= msm::kMaxLevel; ++l) {
        if (d_part[pr][l]) hipFree(d_part[pr][l]);
        if (d_bfrag[pr][l]) hipFree(d_bfrag[pr][l]);
        d_part[pr][l] = nullptr;
        d_bfrag[pr][l] = nullptr;
      }
    for (int l = 0; l <= msm::kMaxLevel; ++l) {
      if (d_ql[l]) hipFree(d_ql[l]);
      if (d_el[l]) hipFree(d_el[l]);
      if (d_khat[l]) hipFree(d_khat[l]);
      if (d_ksplit[l]) hipFree(d_ksplit[l]);
      if (d_qb[l]) hipFree(d_qb[l]);
      d_qb[l] = nullptr;
      d_ql[l] = d_el[l] = d_khat[l] = nullptr;
      d_ksplit[l] = nullptr;
    }
    d_pos = d_q = d_F = d_phi = d_out = d_qpart = d_scal = d_epair = d_egrid = nullptr;
    d_cell_of = d_cnt = d_start = d_sidx = d_flags = d_bucket = d_ovf = d_ovfn = nullptr;
    d_Fp = d_phip = nullptr;
    d_sbase = nullptr;
    d_st = nullptr;
    d_spos = nullptr;
    d_sten = nullptr;
    d_count = nullptr;
    d_qsplit = nullptr;
    d_tw = d_topwork = nullptr;
    d_stab = d_sscr = nullptr;
    stab_cap = sscr_cap = 0;
    if (ev_fork) hipEventDestroy(ev_fork);
    if (ev_join) hipEventDestroy(ev_join);
    if (ev_q1) hipEventDestroy(ev_q1);
    if (ev_g1) hipEventDestroy(ev_g1);
    ev_fork = ev_join = ev_q1 = ev_g1 = nullptr;
    stream = stream2 = stream3 = nullptr;
  }
};

struct FFX {
  FF ff;  // first member: FF* and FFX* alias
  Engine *eng = nullptr;
  int precision = MSM_PREC_FP64;
  int layout = MSM_LAYOUT_FIXED;
  int timing = 0;
  int method = MSM_METHOD_MSM;
  double tolDir = -1;  // PME direct-sum tolerance (<= 0: the default)
  double margin = 0;   // neighbour-list margin of the last build or rebuild
  bool built = false;
  // host copies behind the public FF pointers
  std::vector<double> h_tau, h_Q, h_J, h_aCut, h_cL[3];
  std::vector<double *> h_khat_ptr;
  std::vector<std::vector<double>> h_khat;
  double *(*h_omegap)[3] = nullptr;
  std::vector<std::vector<double>> h_omegap_store;
};

FFX *X(FF *ff) { return reinterpret_cast<FFX *>(ff); }

template <typename T>
int dalloc(T **p, size_t n) {
  if (n == 0) n = 1;
  hipError_t e = hipMalloc((void **)p, n * sizeof(T));
  if (e != hipSuccess)
    return fail(MSM_ERR_ALLOC, std::string("hipMalloc of ") + std::to_string(n * sizeof(T)) +
                                   " bytes: " + hipGetErrorString(e));
  return MSM_OK;
}

#define TRY(x)              \
  do {                      \
    int r_ = (x);           \
    if (r_ != MSM_OK) return r_; \
  } while (0)

// Neighbour cell columns of the pair-search cell grid for pairs out to
// radius R.  A column (dx, dy, dzlo, dzhi) lists the offsets dz in
// [dzlo, dzhi] of the cells (dx, dy, dz) whose nearest points lie inside R;
// longer ranges are split so that each spans fewer than nc[2] cells (one
// wrap at most).  Triclinic cells are parallelepipeds: offset d is kept where
// |A (d / nc)| minus the longest cell diagonal stays below R (the reference's
// neighbour-cell rule, forcefield0.c:61-107); for each (dx, dy) those dz form
// one interval (the distance is convex in dz).
void cell_columns(const msm::Plan &p, const msm::DevParams &P, double R, std::vector<int4> &cols) {
  cols.clear();
  if (!p.ortho) {
    auto vec = [&](double fx, double fy, double fz, double v[3]) {
      for (int a = 0; a < 3; ++a) v[a] = p.A[a][0] * fx + p.A[a][1] * fy + p.A[a][2] * fz;
    };
    double diam = 0.0, v[3];
    for (int sx = -1; sx <= 1; sx += 2)
      for (int sy = -1; sy <= 1; sy += 2) {
        vec((double)sx / P.nc[0], (double)sy / P.nc[1], 1.0 / P.nc[2], v);
        diam = std::max(diam, std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]));
      }
    int lim[3];
    for (int a = 0; a < 3; ++a) lim[a] = (int)std::ceil((R + diam) / P.cw[a]);
    for (int dx = -lim[0]; dx <= lim[0]; ++dx)
      for (int dy = -lim[1]; dy <= lim[1]; ++dy) {
        int zlo = 1, zhi = 0;
        for (int dz = -lim[2]; dz <= lim[2]; ++dz) {
          vec((double)dx / P.nc[0], (double)dy / P.nc[1], (double)dz / P.nc[2], v);
          if (std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]) - diam < R) {
            if (zlo > zhi) zlo = dz;
            zhi = dz;
          }
        }
        for (int z0 = zlo; z0 <= zhi; z0 += P.nc[2])
          cols.push_back(make_int4(dx, dy, z0, std::min(zhi, z0 + P.nc[2] - 1)));
      }
    return;
  }
  int lim[3];
  for (int a = 0; a < 3; ++a) lim[a] = (int)std::ceil(R / P.cw[a]);
  auto gap2 = [&](int d, int a) {
    double g = (std::abs(d) - 1) * P.cw[a];
    return g > 0 ? g * g : 0.0;
  };
  for (int dx = -lim[0]; dx <= lim[0]; ++dx)
    for (int dy = -lim[1]; dy <= lim[1]; ++dy) {
      const double gxy = gap2(dx, 0) + gap2(dy, 1);
      if (gxy >= R * R) continue;
      int zmax = 0;
      while (zmax + 1 <= lim[2] && gxy + gap2(zmax + 1, 2) < R * R) ++zmax;
      for (int z0 = -zmax; z0 <= zmax; z0 += P.nc[2])
        cols.push_back(make_int4(dx, dy, z0, std::min(zmax, z0 + P.nc[2] - 1)));
    }
}

// The pair-search cell grid, its neighbour cell columns for the cutoff and
// the self-image constant: sum of g(|A k|) over the images k != 0 of an atom
// itself inside the cutoff (forcefield1.c:240-321 skips i == j at every
// image).  Cells are at least a0/3 wide (a0/2 where that would leave fewer
// than about two atoms per cell).
void pair_geometry(const msm::Plan &p, msm::DevParams &P, std::vector<int4> &cols) {
  const double a0 = p.aCut[0];
  const bool tri = !p.ortho;
  // face-to-face widths of the box: the diagonal, or 1 / |row a of A^-1|
  double W[3];
  for (int a = 0; a < 3; ++a)
    W[a] = tri ? 1.0 / std::sqrt(p.Ai[a][0] * p.Ai[a][0] + p.Ai[a][1] * p.Ai[a][1] +
                                 p.Ai[a][2] * p.Ai[a][2])
               : p.A[a][a];
  const double vol = tri ? p.detA : p.A[0][0] * p.A[1][1] * p.A[2][2];
  const double dens = p.N / vol, w3 = a0 / 3.0;
  const double div = dens * w3 * w3 * w3 >= 2.0 ? 3.0 : 2.0;
  for (int a = 0; a < 3; ++a) {
    double L = W[a];
    int n = (int)std::floor(L / (a0 / div));
    if (n < 1) n = 1;
    if (n > 256) n = 256;
    P.nc[a] = n;
    P.cw[a] = L / n;
  }
  P.ncell = P.nc[0] * P.nc[1] * P.nc[2];
  cell_columns(p, P, a0, cols);
  double C = 0.0, v[3];
  int K[3];
  for (int a = 0; a < 3; ++a) K[a] = (int)std::ceil(a0 / W[a]);
  for (int kx = -K[0]; kx <= K[0]; ++kx)
    for (int ky = -K[1]; ky <= K[1]; ++ky)
      for (int kz = -K[2]; kz <= K[2]; ++kz) {
        if (!kx && !ky && !kz) continue;
        for (int a = 0; a < 3; ++a) v[a] = p.A[a][0] * kx + p.A[a][1] * ky + p.A[a][2] * kz;
        const double r2 = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
        if (r2 >= a0 * a0) continue;
        if (p.method == msm::kPme) {
          C += std::erfc(p.beta * std::sqrt(r2)) / std::sqrt(r2);
          continue;
        }
        const double sv = r2 / (a0 * a0) - 1.0;
        double ts = p.tau[p.nu - 1];
        for (int k = p.nu - 2; k >= 0; --k) ts = p.tau[k] + sv * ts;
        C += 1.0 / std::sqrt(r2) - ts / a0;
      }
  P.self_img = C;
}

void fill_params(const msm::Plan &p, int N, msm::DevParams &P) {
  for (int a = 0; a < 3; ++a) {
    P.box[a] = p.A[a][a];
    P.ibox[a] = p.Ai[a][a];
  }
  P.tri = !p.ortho;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) P.A[3 * i + j] = p.A[i][j], P.Ai[3 * i + j] = p.Ai[i][j];
  double a0 = p.aCut[0];
  P.a0 = a0;
  P.a02 = a0 * a0;
  P.inv_a0 = 1.0 / a0;
  P.inv_a02 = 1.0 / (a0 * a0);
  P.two_inv_a03 = 2.0 / (a0 * a0 * a0);
  P.nu = p.nu;
  P.N = N;
  std::memset(P.tau, 0, sizeof P.tau);
  std::memset(P.Q, 0, sizeof P.Q);
  std::memset(P.J, 0, sizeof P.J);
  for (int k = 0; k < p.nu; ++k) P.tau[k] = p.tau[k];
  for (int k = 0; k < p.nu / 2 * p.nu; ++k) P.Q[k] = p.Q[k];
  for (int k = 0; k <= p.nu; ++k) P.J[k] = p.J[k];
  P.coeff1 = p.coeff1;
  P.coeff2 = p.coeff2;
  P.gamma0 = p.gamma0;
  P.coeff_self = p.coeff1 - p.gamma0 / a0;
  P.cut2 = P.a02 * (1.0 + 1e-12);
  std::memset(P.tau_s, 0, sizeof P.tau_s);
  std::memset(P.dtau_s, 0, sizeof P.dtau_s);
  std::memset(P.ftau, 0, sizeof P.ftau);
  std::memset(P.fdtau, 0, sizeof P.fdtau);
  for (int k = 0; k < p.nu; ++k) {
    P.tau_s[k] = p.tau[k] / a0;
    P.ftau[k] = (float)P.tau_s[k];
  }
  for (int k = 0; k + 1 < p.nu; ++k) {
    P.dtau_s[k] = (k + 1) * p.tau[k + 1] * P.two_inv_a03;
    P.fdtau[k] = (float)P.dtau_s[k];
  }
  P.fa02 = (float)P.a02;
  P.fia02 = (float)P.inv_a02;
  P.fcut2 = (float)(P.a02 * 1.0001);
  P.ewald = p.method == msm::kPme;
  P.beta = p.beta;
  P.beta2 = p.beta * p.beta;
  P.ewc = 2.0 * p.beta / std::sqrt(M_PI);
  P.fbeta = (float)P.beta;
  P.fbeta2 = (float)P.beta2;
  P.fewc = (float)P.ewc;
}

// Uploads the box-dependent device data (stencils, twiddles, parameters).
int upload_box(Engine &E) {
  msm::Plan &p = E.plan;
  int ncell_alloc = E.P.ncell;
  fill_params(p, p.N, E.P);
  std::vector<int4> sten;
  pair_geometry(p, E.P, sten);
  if (ncell_alloc > 0 && E.P.ncell > ncell_alloc)
    return fail(MSM_ERR_ARG, "box change enlarges the pair-search cell grid beyond its buffers; "
                             "rebuild the FF with FF_build");
  E.nsten = (int)sten.size();
  if (E.d_sten) hipFree(E.d_sten), E.d_sten = nullptr;
  TRY(dalloc(&E.d_sten, sten.size()));
  HIP_TRY(hipMemcpy(E.d_sten, sten.data(), sten.size() * sizeof(int4), hipMemcpyHostToDevice));
  for (int l = 1; l <= p.L; ++l) {
    if (E.d_khat[l]) hipFree(E.d_khat[l]), E.d_khat[l] = nullptr;
    TRY(dalloc(&E.d_khat[l], p.khat[l].size()));
    HIP_TRY(hipMemcpy(E.d_khat[l], p.khat[l].data(), p.khat[l].size() * sizeof(double),
                      hipMemcpyHostToDevice));
  }
  const int *gt = p.M[p.L];
  std::vector<double> tw;
  for (int a = 0; a < 3; ++a)
    for (int j = 0; j < gt[a]; ++j) {
      double ang = 2.0 * M_PI * (double)j / (double)gt[a];
      tw.push_back(std::cos(ang));
      tw.push_back(std::sin(ang));
    }
  if (E.d_tw) hipFree(E.d_tw), E.d_tw = nullptr;
  TRY(dalloc(&E.d_tw, tw.size()));
  HIP_TRY(hipMemcpy(E.d_tw, tw.data(), tw.size() * sizeof(double), hipMemcpyHostToDevice));
  // Toeplitz B operands of the level stencils in MFMA lane order
  for (int pr = 0; pr < 2; ++pr)
    for (int l = 1; l < p.L; ++l) {
      msm::G2GLevel &lv = E.g2g[pr].lv[l - 1];
      size_t bytes = msm::g2g_bfrag_bytes(pr, lv);
      std::vector<unsigned char> host(bytes);
      msm::g2g_bfrag_fill(pr, lv, p.khat[l].data(), host.data());
      if (E.d_bfrag[pr][l]) hipFree(E.d_bfrag[pr][l]), E.d_bfrag[pr][l] = nullptr;
      TRY(dalloc((unsigned char **)&E.d_bfrag[pr][l], bytes));
      HIP_TRY(hipMemcpy(E.d_bfrag[pr][l], host.data(), bytes, hipMemcpyHostToDevice));
      lv.bfrag = E.d_bfrag[pr][l];
    }
  return MSM_OK;
}

// FF_rebuild's device path.  The host plan (rebuild_plan without stencils)
// supplies the constants, omega' and cL; the level stencils, the top-level
// kernel and the Toeplitz B operands of both precisions are generated on
// the GPU into the buffers FF_build allocated (a rebuild keeps nu, the grids
// and nLim, so every size is unchanged), then copied back into the plan's
// host khat, which FF::khat exposes.  Blocking, like the reference's
// FF_rebuild: it first waits for work in flight that may read the stencils.
int rebuild_box(Engine &E) {
  msm::Plan &p = E.plan;
  const hipStream_t s = E.stream;
  HIP_TRY(hipDeviceSynchronize());
  const int ncell_alloc = E.P.ncell;
  fill_params(p, p.N, E.P);
  std::vector<int4> sten;
  pair_geometry(p, E.P, sten);
  if (E.P.ncell > ncell_alloc)
    return fail(MSM_ERR_ARG, "box change enlarges the pair-search cell grid beyond its buffers; "
                             "rebuild the FF with FF_build");
  E.nsten = (int)sten.size();
  if (E.d_sten) hipFree(E.d_sten), E.d_sten = nullptr;
  TRY(dalloc(&E.d_sten, sten.size()));
  HIP_TRY(hipMemcpy(E.d_sten, sten.data(), sten.size() * sizeof(int4), hipMemcpyHostToDevice));
  // omega'_l along x, y, z for the levels below the top, then cL_x, cL_y, cL_z
  std::vector<double> tab;
  size_t woff[msm::kMaxLevel + 1][3] = {}, cloff = 0;
  size_t smax = 1;
  for (int l = 1; l < p.L; ++l) {
    for (int a = 0; a < 3; ++a) {
      woff[l][a] = tab.size();
      tab.insert(tab.end(), p.omegap[l][a].begin(), p.omegap[l][a].end());
    }
    smax = std::max(smax, (size_t)p.S[l][0] * p.S[l][1] * p.S[l][2]);
  }
  cloff = tab.size();
  for (int a = 0; a < 3; ++a) tab.insert(tab.end(), p.cL[a].begin(), p.cL[a].end());
  if (tab.size() > E.stab_cap) {
    if (E.d_stab) hipFree(E.d_stab), E.d_stab = nullptr;
    TRY(dalloc(&E.d_stab, tab.size()));
    E.stab_cap = tab.size();
  }
  if (2 * smax > E.sscr_cap) {
    if (E.d_sscr) hipFree(E.d_sscr), E.d_sscr = nullptr;
    TRY(dalloc(&E.d_sscr, 2 * smax));
    E.sscr_cap = 2 * smax;
  }
  HIP_TRY(hipMemcpy(E.d_stab, tab.data(), tab.size() * sizeof(double), hipMemcpyHostToDevice));
  msm::StencilBox b{};
  std::memcpy(b.A, p.A, sizeof b.A);
  std::memcpy(b.Ai, p.Ai, sizeof b.Ai);
  for (int a = 0; a < 3; ++a) b.as[a] = p.kas[a];
  for (int l = 0; l <= p.L; ++l) b.aCut[l] = p.aCut[l];
  for (int k = 0; k < p.nu; ++k) b.tau[k] = p.tau[k];
  msm::top_series(p.nu, b.taylor);
  b.detA = p.detA;
  b.nu = p.nu;
  b.L = p.L;
  std::memcpy(b.M, p.M, sizeof b.M);
  std::memcpy(b.S, p.S, sizeof b.S);
  for (int l = 1; l < p.L; ++l)
    msm::launch_stencil_level(b, l, E.d_stab + woff[l][0], E.d_stab + woff[l][1],
                              E.d_stab + woff[l][2], E.d_sscr, E.d_sscr + smax, E.d_khat[l], s);
  b.beta = p.beta;
  if (p.method == msm::kPme)
    msm::launch_stencil_top_pme(b, E.d_stab + cloff, E.d_khat[1], s);
  else
    msm::launch_stencil_top(b, E.d_stab + cloff, E.d_khat[p.L], s);
  for (int pr = 0; pr < 2; ++pr)
    for (int l = 1; l < p.L; ++l)
      msm::launch_g2g_bfrag_fill(pr, E.g2g[pr].lv[l - 1], E.d_khat[l], E.d_bfrag[pr][l], s);
  for (int l = 1; l <= p.L; ++l) {
    const size_t n = l < p.L ? (size_t)p.S[l][0] * p.S[l][1] * p.S[l][2]
                             : (size_t)p.M[l][0] * p.M[l][1] * p.M[l][2];
    p.khat[l].resize(n);
    HIP_TRY(hipMemcpyAsync(p.khat[l].data(), E.d_khat[l], n * sizeof(double),
                           hipMemcpyDeviceToHost, s));
  }
  HIP_TRY(hipStreamSynchronize(s));
  HIP_TRY(hipGetLastError());
  E.kept = E.kept_pairs = false;
  return MSM_OK;
}

// Two streams per device, created once and shared by every FF of the
// process: the concurrent pair kernel and the level-1 stencil pass.  The
// chain runs on the caller's stream (the device-pointer entry points) or on
// HIP's null stream (the host-buffer entry points and FF_build's pair
// count).  HIP maps streams onto a few hardware queues (four per process by
// default): the caller's stream, the null stream that PyTorch also uses and
// these two fit; a third library stream made a caller on its own stream
// share a queue with a concurrent stage (0.36 instead of 0.30 ms per
// water96k evaluation), and streams per FF would do the same.  FFs on one
// device are thereby ordered with each other.
int device_streams(int dev, hipStream_t (&out)[2]) {
  static std::mutex mu;
  struct Pair {
    hipStream_t s[2] = {nullptr, nullptr};
  };
  static std::vector<Pair> tab;
  std::lock_guard<std::mutex> lk(mu);
  if ((int)tab.size() <= dev) tab.resize(dev + 1);
  Pair &t = tab[dev];
  if (!t.s[0]) {
    Pair n;
    for (int k = 0; k < 2; ++k)
      if (hipStreamCreateWithFlags(&n.s[k], hipStreamNonBlocking) != hipSuccess) {
        for (int j = 0; j < k; ++j) hipStreamDestroy(n.s[j]);
        return fail(MSM_ERR_HIP, "hipStreamCreateWithFlags failed");
      }
    t = n;
  }
  for (int k = 0; k < 2; ++k) out[k] = t.s[k];
  return MSM_OK;
}

// Tiles, kx groups and partial buffers of the MFMA level-stencil launch.

int plan_levels(Engine &E) {
  const msm::Plan &p = E.plan;
  static const int target[2] = {4096, 1024};
  for (int pr = 0; pr < 2; ++pr) {
    msm::G2GArgs &A = E.g2g[pr];
    A.nlev = 0;
    A.ntasks = 0;
    E.g2g_lds[pr] = 0;
    for (int l = 1; l < p.L; ++l) {
      msm::plan_g2g_level(pr, E.gd[l], E.sd[l], target[pr], A.lv[l - 1]);
      A.nlev = l;
    }
    // level 1 has a launch of its own, the levels 2 .. L-1 share one: within
    // a launch the levels share the K-chunk count (extra chunks of a smaller
    // stencil hold zero Toeplitz columns) and the z tiles per wavefront
    if (A.nlev >= 1) A.ntasks = msm::plan_g2g_launch(pr, &A.lv[0], 1, target[pr]);
    if (A.nlev >= 2) A.ntasks += msm::plan_g2g_launch(pr, &A.lv[1], A.nlev - 1, target[pr]);
    for (int k = 0; k < A.nlev; ++k) {
      msm::G2GLevel &lv = A.lv[k];
      if (lv.nch > msm::g2g_max_nch(pr))
        return fail(MSM_ERR_ARG, "FF_build: level stencil too wide for the MFMA tiles");
      lv.q = E.d_ql[k + 1];
      lv.qb = E.d_qb[k + 1];
      TRY(dalloc(&E.d_part[pr][k + 1], (size_t)lv.G * E.gd[k + 1].size()));
      lv.part = E.d_part[pr][k + 1];
      E.g2g_lds[pr] = std::max(E.g2g_lds[pr], msm::g2g_lds_bytes(pr, lv));
    }
    if (E.g2g_lds[pr] > 65536)
      return fail(MSM_ERR_ARG, "FF_build: level stencil too wide for one workgroup's LDS");
  }
  return MSM_OK;
}

int create_engine(FFX *x, const msm::Plan &plan) {
  Engine *E = new Engine();
  E->plan = plan;
  x->eng = E;
  HIP_TRY(hipGetDevice(&E->device));
  {
    hipStream_t st[2];
    TRY(device_streams(E->device, st));
    E->stream = nullptr;  // HIP's null stream: host-buffer entry points, FF_build
    E->stream2 = st[0], E->stream3 = st[1];
  }
  // fork/join events order kernels of this device only: no system-scope
  // fence (a cache writeback and invalidation per record) is needed
  for (hipEvent_t *e : {&E->ev_fork, &E->ev_join, &E->ev_q1, &E->ev_g1})
    HIP_TRY(hipEventCreateWithFlags(e, hipEventDisableTiming | hipEventDisableSystemFence));
  const msm::Plan &p = E->plan;
  for (int l = 1; l <= p.L; ++l) {
    E->gd[l] = {p.M[l][0], p.M[l][1], p.M[l][2]};
    E->sd[l] = {p.S[l][0], p.S[l][1], p.S[l][2]};
  }
  E->sd[p.L] = E->gd[p.L];
  // the pair-search cell grid sizes the binning buffers below, so it must be
  // known before anything is allocated (upload_box recomputes the same values)
  {
    std::vector<int4> sten;
    fill_params(p, p.N, E->P);
    pair_geometry(p, E->P, sten);
  }
  if (E->P.ncell <= 0) return fail(MSM_ERR_ARG, "FF_build: empty pair-search cell grid");
  long N = p.N;
  int nb = msm::bin_blocks((int)N), np = msm::pair_blocks((int)N);
  TRY(dalloc(&E->d_pos, 3 * N));
  TRY(dalloc(&E->d_q, N));
  TRY(dalloc(&E->d_F, 3 * N));
  TRY(dalloc(&E->d_phi, N));
  TRY(dalloc(&E->d_out, 8));
  TRY(dalloc(&E->d_qpart, 2 * nb));
  TRY(dalloc(&E->d_scal, 4));
  TRY(dalloc(&E->d_epair, std::max(np, E->P.ncell)));
  TRY(dalloc(&E->d_egrid, std::max<long>(msm::grid_blocks(E->gd[1].size()),
                                          msm::prolong_blocks(E->gd[1]))));
  TRY(dalloc(&E->d_cell_of, N));
  // bucket rows of several times the mean occupancy (a cell past its row
  // still works, through the overflow list, only more slowly)
  E->bcap = 16 * (int)std::ceil((4.0 * (double)N / E->P.ncell + 8.0) / 16.0);
  TRY(dalloc(&E->d_bucket, (size_t)E->P.ncell * E->bcap));
  TRY(dalloc(&E->d_ovf, N));
  TRY(dalloc(&E->d_ovfn, 2));
  TRY(dalloc(&E->d_Fp, 3 * N));
  TRY(dalloc(&E->d_phip, N));
  // the scan reads and writes whole int4 (padding of cnt zero)
  const size_t ncell4 = ((size_t)E->P.ncell + 3) / 4 * 4;
  TRY(dalloc(&E->d_cnt, ncell4));
  TRY(dalloc(&E->d_start, ncell4 + 4));
  TRY(dalloc(&E->d_sidx, N));
  TRY(dalloc(&E->d_flags, 4));
  // counters that each evaluation leaves cleared for the next one
  HIP_TRY(hipMemset(E->d_cnt, 0, ncell4 * sizeof(int)));
  HIP_TRY(hipMemset(E->d_flags, 0, 4 * sizeof(int)));
  HIP_TRY(hipMemset(E->d_ovfn, 0, 2 * sizeof(int)));
  TRY(dalloc(&E->d_spos, N));
  TRY(dalloc(&E->d_count, E->P.ncell));
  if (msm::spread_by_placements(E->P, E->gd[1])) {
    TRY(dalloc(&E->d_sbase, N));
    TRY(dalloc(&E->d_st, N));
  }
  for (int l = 1; l <= p.L; ++l) {
    TRY(dalloc(&E->d_ql[l], E->gd[l].size()));
    TRY(dalloc(&E->d_el[l], E->gd[l].size()));
    if (l < p.L) TRY(dalloc(&E->d_qb[l], E->gd[l].size()));  // BF16 stencil operand
  }
  TRY(dalloc(&E->d_topwork, 4 * E->gd[p.L].size()));
  TRY(plan_levels(*E));
  TRY(upload_box(*E));
  return MSM_OK;
}

// ---- the evaluation chain ------------------------------------------------
// Binning + counting sort of atoms into cells (replaces the stored
// neighbour list of forcefield0.c:21-164).
// fork (optional): bound to the completion of the last binning kernel, for
// the pair stream to wait on.
void stage_prepare(Engine &E, const double *pos, const double *q, hipStream_t s,
                   hipEvent_t fork = nullptr) {
  // d_cnt and d_ovfn[0] are zero here: set at allocation, cleared again by
  // every scan
  msm::launch_bin(E.P, pos, q, E.d_cell_of, E.d_bucket, E.bcap, E.d_ovfn, E.d_ovf, E.d_cnt,
                  E.d_qpart, E.d_flags, s);
  msm::launch_scan(E.P, E.d_cnt, E.d_start, E.d_qpart, msm::bin_blocks(E.P.N), E.d_scal,
                   E.d_ovfn, s);
  // the placement pass also writes the atoms' spreading placements
  const bool place = msm::spread_by_placements(E.P, E.gd[1]);
  msm::launch_place(E.P, pos, q, E.d_cell_of, E.d_start, E.d_bucket, E.bcap, E.d_ovfn,
                    E.d_ovf, E.d_spos, E.d_sidx, s, fork, E.gd[1],
                    place ? E.d_sbase : nullptr, place ? E.d_st : nullptr);
}

void stage_pairs(Engine &E, int precision, double *F, double *phi, hipStream_t s,
                 hipEvent_t done = nullptr) {
  if (precision == MSM_PREC_MIXED)
    msm::launch_pair_fp32(E.P, E.d_spos, E.d_sidx, E.d_start, E.d_sten, E.nsten, F, phi,
                          E.d_epair, s, done);
  else
    msm::launch_pair_fp64(E.P, E.d_spos, E.d_sidx, E.d_start, E.d_sten, E.nsten, F, phi,
                          E.d_epair, s, done);
}

// LDS of one stencil launch: its own levels' slabs only (a workgroup that
// asks for more than the pair kernel leaves free on a CU waits until pair
// workgroups retire there).
size_t g2g_launch_lds(int pr, const msm::G2GArgs &a) {
  size_t m = 0;
  for (int k = 0; k < a.nlev; ++k) m = std::max(m, msm::g2g_lds_bytes(pr, a.lv[k]));
  return m;
}

// The levels [first, last) of the level-stencil launch description, as a
// launch of their own (task numbering rebased to the first of them).
msm::G2GArgs g2g_levels(const msm::G2GArgs &A, int first, int last) {
  msm::G2GArgs R{};
  R.nlev = 0;
  R.ntasks = 0;
  for (int k = first; k < last && k < A.nlev; ++k) {
    R.lv[R.nlev] = A.lv[k];
    R.lv[R.nlev].task0 = R.ntasks;
    R.ntasks += msm::g2g_level_tasks(A.lv[k]);
    ++R.nlev;
  }
  return R;
}

// The grid chain: spread to level 1, restrict to the top, top-level DFT
// convolution, then prolongate and add each level's stencil pass down to
// e^1.  The level-1 stencil pass, the largest, needs q^1 only: it runs on the
// third stream beside the restrictions, the coarser levels and their
// prolongations, and is joined before the level-1 prolongation.
int stage_grids(Engine &E, int precision, int layout, Timing &T, hipStream_t s) {
  const msm::Plan &p = E.plan;
  const int L = p.L;
  const int pr = precision == MSM_PREC_MIXED ? msm::MSM_G2G_BF16X3 : msm::MSM_G2G_FP64;
  int t = T.open(s);
  // the BF16 stencils read q^l as bf16 hi/lo pairs, written with q^l
  const bool split = pr == msm::MSM_G2G_BF16X3;
  // ev_q1 (q^1 ready) rides on the spreading kernel's dispatch
  msm::launch_spread(E.P, E.gd[1], E.d_spos, E.d_start, E.d_ql[1], s,
                     split && L >= 2 ? E.d_qb[1] : nullptr, E.d_sbase, E.d_st,
                     L >= 2 ? E.ev_q1 : nullptr, /*placed=*/true);
  T.close(t, "anterpolation", s);
  if (L >= 2) {
    HIP_TRY(hipStreamWaitEvent(E.stream3, E.ev_q1, 0));
    t = T.open(E.stream3);
    // beside four or more levels the coarse chain outlasts this pass, which
    // then leaves issue priority to the pair kernel
    msm::G2GArgs a1 = g2g_levels(E.g2g[pr], 0, 1);
    a1.urgent = L <= 3;
    msm::launch_g2g_tiles(pr, a1, g2g_launch_lds(pr, a1), E.stream3, E.ev_g1);
    T.close(t, "grid2grid", E.stream3);
  }
  for (int l = 2; l <= L; ++l) {
    t = T.open(s);
    msm::launch_restrict(E.P, E.gd[l - 1], E.gd[l], E.d_ql[l - 1], E.d_ql[l], s,
                         split && l < L ? E.d_qb[l] : nullptr);
    T.close(t, "restriction", s);
  }
  if (L >= 3) {  // stencil passes of the levels 2 .. L-1 in one launch
    t = T.open(s);
    msm::G2GArgs ac = g2g_levels(E.g2g[pr], 1, L - 1);
    ac.urgent = 1;
    msm::launch_g2g_tiles(pr, ac, g2g_launch_lds(pr, ac), s);
    T.close(t, "grid2grid", s);
  }
  const int top_map = layout != MSM_LAYOUT_ASIS ? 0
                      : L >= 2                  ? (E.plan.fft ? msm::kTopAsisStride : 0)
                                                : (E.plan.fft ? 0 : msm::kTopAsisFlat);
  // a small top level rides on the prolongation onto level L-1 (>= 2)
  const bool fused_top = L >= 3 && msm::prolong_top_fits(E.gd[L - 1], E.gd[L]);
  if (fused_top) {
    t = T.open(s);
    msm::launch_prolong_top(E.P, E.gd[L - 1], E.gd[L], E.g2g[pr].lv[L - 2].G, E.d_part[pr][L - 1],
                            E.d_ql[L], E.d_khat[L], E.d_tw, E.d_el[L], E.d_el[L - 1], top_map, s);
    T.close(t, "grid2grid_top", s);
  } else {
    t = T.open(s);
    msm::launch_top(E.P, E.gd[L], E.d_ql[L], E.d_khat[L], E.d_tw, E.d_topwork, E.d_el[L],
                    top_map, s);
    T.close(t, "grid2grid_top", s);
  }
  for (int l = fused_top ? L - 2 : L - 1; l >= 1; --l) {
    if (l == 1) HIP_TRY(hipStreamWaitEvent(s, E.ev_g1, 0));
    t = T.open(s);
    // the finest level also writes the grid-energy partials (sum q1 e1)
    msm::launch_prolong_sum(E.P, E.gd[l], E.gd[l + 1], E.g2g[pr].lv[l - 1].G, E.d_part[pr][l],
                            E.d_el[l + 1], E.d_el[l], l == 1 ? E.d_ql[1] : nullptr,
                            l == 1 ? E.d_egrid : nullptr, s);
    T.close(t, "prolongation", s);
  }
  return MSM_OK;
}

// Grid energy: with L >= 2 the level-1 prolongation has written one partial
// of sum q1 e1 per workgroup; with L = 1 a separate pass does (it also
// carries the reference HEAD's L = 1 index map for the as-is layout).
int grid_energy_parts(const Engine &E) {
  return E.plan.L >= 2 ? msm::prolong_blocks(E.gd[1]) : msm::grid_blocks(E.gd[1].size());
}

msm::FinalizeArgs finalize_args(Engine &E, double *user) {
  return msm::FinalizeArgs{E.d_epair, E.P.ncell, E.d_egrid, grid_energy_parts(E), E.d_scal,
                           E.d_flags, E.d_out, user == E.d_out ? nullptr : user};
}

// Grid-energy partials for L = 1 (the level-1 prolongation writes them
// otherwise); part of the chain, before the join with the pair stream.
void stage_grid_energy(Engine &E, int layout, hipStream_t s) {
  if (E.plan.L == 1)
    msm::launch_grid_energy(E.P, E.gd[1], E.d_ql[1], E.d_el[1], E.d_egrid,
                            layout == MSM_LAYOUT_ASIS, s);
}

// After the join: with F or phi wanted, add the short-range parts (written
// by the pair kernel into d_Fp / d_phip) and reduce the energy in the same
// launch; otherwise only the reduction.
// `out` (caller's buffer of 4 doubles, may be E.d_out) receives the terms
// from the reduction kernel itself.
void stage_energy(Engine &E, double *F, double *phi, double *out, hipStream_t s) {
  if (F || phi)
    msm::launch_combine(E.P, E.d_Fp, E.d_phip, F, phi, finalize_args(E, out), s);
  else
    msm::launch_finalize(E.P, finalize_args(E, out), s);
}

int check_built(FF *ff) {
  if (!ff) return fail(MSM_ERR_ARG, "FF is NULL");
  FFX *x = X(ff);
  if (!x->built || !x->eng) return fail(MSM_ERR_STATE, "FF_build has not completed successfully");
  int dev = 0;
  HIP_TRY(hipGetDevice(&dev));
  if (dev != x->eng->device) HIP_TRY(hipSetDevice(x->eng->device));
  return MSM_OK;
}

// Pairs on the pair stream, concurrently with the grid chain; the caller
// joins before the reduction that adds the pair results (k_combine or
// k_finalize).  The binning bound ev_fork to its last kernel (stage_prepare
// with E.ev_fork); ev_join rides on the pair kernel.
int fork_pairs(Engine &E, int precision, double *F, double *phi, Timing &T) {
  HIP_TRY(hipStreamWaitEvent(E.stream2, E.ev_fork, 0));
  int t = T.open(E.stream2);
  stage_pairs(E, precision, F, phi, E.stream2, E.ev_join);
  T.close(t, "partcl2partcl", E.stream2);
  return MSM_OK;
}

// Full evaluation on device pointers; out receives {E, const, pair, grid}.
int eval_chain(FFX *x, const double *pos, const double *q, double *out, double *F, double *phi,
               hipStream_t s, Timing &T) {
  Engine &E = *x->eng;
  int t = T.open(s);
  stage_prepare(E, pos, q, s, E.ev_fork);
  T.close(t, "nlist", s);
  // the pair kernel runs beside the whole grid chain including the
  // interpolation (the chain's kernels are sized -- registers, LDS -- to fit
  // next to its resident wavefronts); it writes the short-range parts of F
  // and phi to its own buffers, added after the join
  TRY(fork_pairs(E, x->precision, F ? E.d_Fp : nullptr, phi ? E.d_phip : nullptr, T));
  TRY(stage_grids(E, x->precision, x->layout, T, s));
  stage_grid_energy(E, x->layout, s);
  if (F || phi) {
    t = T.open(s);
    msm::launch_gather(E.P, E.gd[1], E.d_spos, E.d_sidx, E.d_start, E.d_el[1], E.d_scal, F, phi, false, s);
    T.close(t, "interpolation", s);
  }
  HIP_TRY(hipStreamWaitEvent(s, E.ev_join, 0));
  t = T.open(s);
  stage_energy(E, F, phi, out, s);
  T.close(t, "energy", s);
  HIP_TRY(hipGetLastError());
  if (T.on) ++E.timed_evals;
  E.kept = E.kept_pairs = false;
  return MSM_OK;
}

void publish_timing(FFX *x, Timing &T) {
  if (!T.on || T.iv.empty()) return;
  FF *ff = &x->ff;
  ff->time_partcl2partcl = ff->time_anterpolation = ff->time_interpolation = 0;
  ff->time_nlist = 0;
  for (int l = 0; l < 10; ++l)
    ff->time_grid2grid[l] = ff->time_restriction[l] = ff->time_prolongation[l] =
        ff->time_padding[l] = 0;
  int L = x->eng->plan.L, lr = 2, lg = L - 1, lc = 1;
  for (auto &v : T.iv) {
    float ms = 0;
    hipEventElapsedTime(&ms, v.a, v.b);
    double sec = ms * 1e-3;
    const std::string &t = v.tag;
    if (t == "nlist") ff->time_nlist = sec;
    else if (t == "partcl2partcl") ff->time_partcl2partcl = sec;
    else if (t == "anterpolation") ff->time_anterpolation = sec;
    else if (t == "restriction") { if (lr <= 9) ff->time_restriction[lr] = sec; ++lr; }
    else if (t == "grid2grid_top") { if (L <= 9) ff->time_grid2grid[L] = sec; }
    else if (t == "grid2grid") {  // level 1, then the levels 2 .. L-1 in one launch
      if (lc <= 9) ff->time_grid2grid[lc] = sec;
      ++lc;
    }
    else if (t == "prolongation") { if (lg >= 1 && lg <= 9) ff->time_prolongation[lg] = sec; --lg; }
    else if (t == "interpolation") ff->time_interpolation = sec;
  }
}

// Stage names reported by msm_stage_times, in order.
const char *kStageNames[] = {"nlist",         "partcl2partcl", "anterpolation",
                             "restriction",   "grid2grid",     "grid2grid_top",
                             "prolongation",  "energy",        "interpolation"};
constexpr int kNumStages = 9;

void publish_plan(FFX *x) {
  FF *ff = &x->ff;
  const msm::Plan &p = x->eng->plan;
  std::memcpy(ff->A, p.A, sizeof ff->A);
  std::memcpy(ff->Ai, p.Ai, sizeof ff->Ai);
  ff->detA = p.detA;
  ff->relCutoff = p.relCutoff;
  ff->nLim = p.nLim;
  ff->ntc = 7;
  ff->tmax = 4;
  ff->estErr = p.estErr;
  ff->orderAcc = p.nu;
  ff->maxLevel = p.L;
  for (int a = 0; a < 3; ++a) ff->topGridDim[a] = p.M[p.L][a];
  ff->cutoff = p.aCut[0];
  ff->coeff1 = p.coeff1;
  ff->coeff2 = p.coeff2;
  ff->nlist = nullptr;
  x->h_tau.assign(p.tau, p.tau + p.nu);
  x->h_Q.assign(p.Q, p.Q + p.nu / 2 * p.nu);
  x->h_J.assign(p.J, p.J + p.nu + 1);
  x->h_aCut.assign(p.aCut, p.aCut + p.L + 1);
  ff->tau = x->h_tau.data();
  ff->Q = x->h_Q.data();
  ff->J = x->h_J.data();
  ff->aCut = x->h_aCut.data();
  ff->dsigma = nullptr;
  ff->sigmad = nullptr;
  for (int a = 0; a < 3; ++a) {
    x->h_cL[a] = p.cL[a];
    ff->cL[a] = x->h_cL[a].data();
  }
  x->h_khat.assign(p.L + 1, std::vector<double>());
  x->h_khat_ptr.assign(p.L + 1, nullptr);
  for (int l = 1; l <= p.L; ++l) {
    x->h_khat[l] = p.khat[l];
    x->h_khat_ptr[l] = x->h_khat[l].data();
  }
  ff->khat = x->h_khat_ptr.data();
  free(x->h_omegap);
  x->h_omegap = (double *(*)[3])calloc(p.L + 1, sizeof(double *[3]));
  x->h_omegap_store.clear();
  x->h_omegap_store.reserve(3 * (p.L + 1));
  for (int l = 1; l <= p.L; ++l)
    for (int a = 0; a < 3; ++a) {
      x->h_omegap_store.push_back(p.omegap[l][a]);
      x->h_omegap[l][a] = x->h_omegap_store.back().data();
    }
  ff->omegap = x->h_omegap;
}

// The reference's neighbour-list statistics for pairs out to a = a0 + margin
// (forcefield0.c:21-36,109-163): the capacity it allocates for the list and
// the number of pairs i < j whose nearest image lies within a.  The capacity
// is c N + 2.1 (4/3 pi a^3) N (N - 1) / (2 det A), truncated, with c = 200 in
// the source as it stands (forcefield0.c:32) and c = 3 in the build that
// produced the published runs (every nlist_len in results/TableII.json,
// Figure4.json and Figure6.json): the as-is layout reports the first, the
// fixed layout the second.
void publish_nlist_len(FFX *x) {
  FF *ff = &x->ff;
  const double a = ff->aCut[0] + x->margin, N = ff->N;
  const double c = x->layout == MSM_LAYOUT_ASIS ? 200.0 : 3.0;
  const double v = c * N + 2.1 * ((4.0 / 3.0) * M_PI * a * a * a * 0.5 * N * (N - 1.0) / ff->detA);
  ff->nlist_len = v < (double)INT_MAX ? (int)v : INT_MAX;
}

int count_pairs_dev(FFX *x, const double *d_pos, const double *d_q, double margin,
                    long long *count) {
  Engine &E = *x->eng;
  hipStream_t s = E.stream;
  const double R = E.plan.aCut[0] + margin;
  *count = 0;
  if (!(R > 0.0)) return MSM_OK;
  const int4 *cols = E.d_sten;
  int ncol = E.nsten;
  int4 *d_cols = nullptr;  // columns for a radius other than the cutoff
  if (margin != 0.0) {
    std::vector<int4> c;
    cell_columns(E.plan, E.P, R, c);
    TRY(dalloc(&d_cols, c.size()));
    HIP_TRY(hipMemcpy(d_cols, c.data(), c.size() * sizeof(int4), hipMemcpyHostToDevice));
    cols = d_cols, ncol = (int)c.size();
  }
  stage_prepare(E, d_pos, d_q, s);
  unsigned *cc = E.d_count;
  msm::launch_pair_count(E.P, E.d_spos, E.d_start, cols, ncol, R * R, cc, s);
  msm::launch_fill_int(E.d_flags, 1, 0, s);  // no energy reduction follows to clear it
  std::vector<unsigned> h(E.P.ncell);
  HIP_TRY(hipMemcpyAsync(h.data(), cc, h.size() * sizeof(unsigned), hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  unsigned long long c = 0;
  for (unsigned v : h) c += v;
  if (d_cols) hipFree(d_cols);
  *count = (long long)c;
  return MSM_OK;
}

int stage_host_inputs(FFX *x, double (*position)[3]) {
  Engine &E = *x->eng;
  size_t N = E.plan.N;
  HIP_TRY(hipMemcpyAsync(E.d_pos, position, 3 * N * sizeof(double), hipMemcpyHostToDevice, E.stream));
  HIP_TRY(hipMemcpyAsync(E.d_q, x->ff.q, N * sizeof(double), hipMemcpyHostToDevice, E.stream));
  return MSM_OK;
}

int build_impl(FF *ff, int N, double (*position)[3], double edges[3][3], double margin = 0.0) {
  FFX *x = X(ff);
  x->built = false;
  g_err.clear();
  if (x->eng) {
    delete x->eng;
    x->eng = nullptr;
  }
  if (N != ff->N) return fail(MSM_ERR_ARG, "FF_build: N differs from the N given to FF_new");
  if (!ff->q) return fail(MSM_ERR_ARG, "FF_build: charges are NULL");
  if (!position) return fail(MSM_ERR_ARG, "FF_build: positions are NULL");
  msm::BuildRequest r;
  r.N = N;
  r.q = ff->q;
  std::memcpy(r.edges, edges, sizeof r.edges);
  r.nu = ff->orderAcc;
  r.L = ff->maxLevel;
  for (int a = 0; a < 3; ++a) r.Mtop[a] = ff->topGridDim[0] > 0 ? ff->topGridDim[a] : -1;
  r.a0 = ff->cutoff;
  r.estErr = ff->estErr;
  r.method = x->method == MSM_METHOD_PME ? msm::kPme : msm::kMsm;
  r.tolDir = x->tolDir;
  r.complete_images = x->layout != MSM_LAYOUT_ASIS;
  r.fft = ff->FFT || x->method == MSM_METHOD_PME;
  msm::Plan plan;
  std::string err;
  if (x->method == MSM_METHOD_PME && x->layout == MSM_LAYOUT_ASIS)
    return fail(MSM_ERR_ARG, "FF_build: the as-is index layout reproduces the MSM reference's "
                             "quirks and does not apply to the PME method");
  auto t0 = std::chrono::steady_clock::now();
  if (!msm::build_plan(r, plan, err)) return fail(MSM_ERR_ARG, "FF_build: " + err);
  if (!r.fft && r.Mtop[0] <= 0 && plan.L != (r.L > 0 ? r.L : 1))
    std::fprintf(stderr, "L=%d is replaced by L=%d\n", r.L > 0 ? r.L : 1, plan.L);
  double tstencil =
      std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  // the reference's FF_build writes the parameters its rules chose into the
  // public fields (forcefield0.c:287-348); they are the host plan's, so they
  // are visible before (and whether or not) the device part succeeds
  std::memcpy(ff->A, plan.A, sizeof ff->A);
  std::memcpy(ff->Ai, plan.Ai, sizeof ff->Ai);
  ff->detA = plan.detA;
  ff->estErr = plan.estErr;
  ff->orderAcc = plan.nu;
  ff->maxLevel = plan.L;
  for (int a = 0; a < 3; ++a) ff->topGridDim[a] = plan.M[plan.L][a];
  ff->cutoff = plan.aCut[0];
  TRY(create_engine(x, plan));
  publish_plan(x);
  for (int l = 1; l <= plan.L && l < 10; ++l) ff->time_stencil[l] = tstencil / plan.L;
  // neighbour statistics for the reference's JSON fields; the FF counts as
  // built only once this first GPU pass has succeeded
  x->built = true;
  long long cnt = 0;
  auto t1 = std::chrono::steady_clock::now();
  int rc = stage_host_inputs(x, position);
  if (rc == MSM_OK) rc = count_pairs_dev(x, x->eng->d_pos, x->eng->d_q, margin, &cnt);
  if (rc != MSM_OK) {
    x->built = false;
    return rc;
  }
  ff->time_nlist = std::chrono::duration<double>(std::chrono::steady_clock::now() - t1).count();
  ff->nlist_interaction_count = (int)cnt;
  x->margin = margin;
  publish_nlist_len(x);
  return MSM_OK;
}

int energy_host(FF *ff, double (*force)[3], double (*position)[3], double *phi_out, double *E_out) {
  TRY(check_built(ff));
  FFX *x = X(ff);
  Engine &E = *x->eng;
  if (!force || !position) return fail(MSM_ERR_ARG, "FF_energy: force and position must be non-NULL");
  Timing &T = E.T;
  T.clear();
  E.timed_evals = 0;
  T.on = x->timing != 0;
  TRY(stage_host_inputs(x, position));
  int rc = eval_chain(x, E.d_pos, E.d_q, E.d_out, E.d_F, phi_out ? E.d_phi : nullptr, E.stream, T);
  if (rc != MSM_OK) {
    T.clear();
    return rc;
  }
  size_t N = E.plan.N;
  double out[5];  // {E, const, pair, grid, non-finite input count}
  HIP_TRY(hipMemcpyAsync(force, E.d_F, 3 * N * sizeof(double), hipMemcpyDeviceToHost, E.stream));
  if (phi_out) HIP_TRY(hipMemcpyAsync(phi_out, E.d_phi, N * sizeof(double), hipMemcpyDeviceToHost, E.stream));
  HIP_TRY(hipMemcpyAsync(out, E.d_out, sizeof out, hipMemcpyDeviceToHost, E.stream));
  HIP_TRY(hipStreamSynchronize(E.stream));
  const int flags = out[4] != 0.0;
  publish_timing(x, T);
  T.clear();
  E.timed_evals = 0;
  if (flags) return fail(MSM_ERR_ARG, "FF_energy: non-finite position or charge");
  std::memcpy(E.last_terms, out + 1, 3 * sizeof(double));
  *E_out = out[0];
  return MSM_OK;
}

double my_errest(FF *ff, int nu) {
  // _FF_get_errEst (forcefield0.c:553-582) on the public fields
  double ax = std::sqrt(ff->A[0][0] * ff->A[0][0] + ff->A[0][1] * ff->A[0][1] + ff->A[0][2] * ff->A[0][2]);
  double ay = std::sqrt(ff->A[1][0] * ff->A[1][0] + ff->A[1][1] * ff->A[1][1] + ff->A[1][2] * ff->A[1][2]);
  double az = std::sqrt(ff->A[2][0] * ff->A[2][0] + ff->A[2][1] * ff->A[2][1] + ff->A[2][2] * ff->A[2][2]);
  double M1[3];
  for (int a = 0; a < 3; ++a) M1[a] = ff->topGridDim[a] * std::pow(2.0, ff->maxLevel - 1);
  double Q2 = 0;
  for (int i = 0; i < ff->N; ++i) Q2 += ff->q[i] * ff->q[i];
  double Fref = (Q2 / (double)ff->N) / std::pow(ff->detA / (double)ff->N, 5. / 6.);
  double h1 = std::fmax(std::fmax(ax / M1[0], ay / M1[1]), az / M1[2]);
  double dF = Q2 * (1 - std::pow(4.0, -ff->maxLevel)) * std::pow(h1, nu - 1) *
              std::pow(ff->detA, -1. / 3.) * std::pow(ff->N, -2.0 / 3.0);
  return dF / Fref;
}

// ---- grid2grid without allocation -----------------------------------------
// Scratch of msm_grid2grid_device_noalloc: the unpadded input grid and the
// stencil result, kNoAllocGrid doubles each.  Every GPU the library runs on
// holds one copy (8 MiB) until the process ends.  In host code `scratch`
// only names the symbol; kernels get the device address that
// hipGetSymbolAddress returns for the current device.
constexpr long kNoAllocGrid = 1L << 19;  // grid points per call (80^3 = 512,000 fit)
__device__ __align__(256) double scratch[2 * kNoAllocGrid];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double *g_scratch[kMaxDevices] = {};

// Device address of `scratch` on device `dev` (looked up once per device,
// with that device current).
int scratch_on(int dev, double **out) {
  if (dev < 0 || dev >= kMaxDevices) return fail(MSM_ERR_ARG, "device index out of range");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch[dev]) {
    void *ptr = nullptr;
    HIP_TRY(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    g_scratch[dev] = static_cast<double *>(ptr);
  }
  *out = g_scratch[dev];
  return MSM_OK;
}

// q[m] = ql[m + of] of the reference's padded level array (dims g + pad).
__global__ __launch_bounds__(256) void k_unpad(msm::GridDims g, int3 pad, int3 of,
                                               const double *__restrict__ ql,
                                               double *__restrict__ q) {
  const long e = (long)blockIdx.x * 256 + threadIdx.x;
  if (e >= g.size()) return;
  const int k = (int)(e % g.z), j = (int)((e / g.z) % g.y), i = (int)(e / ((long)g.y * g.z));
  q[e] = ql[((long)(i + of.x) * (g.y + pad.y) + (j + of.y)) * (g.z + pad.z) + k + of.z];
}

__global__ __launch_bounds__(256) void k_accumulate(long n, const double *__restrict__ e,
                                                    double *__restrict__ el) {
  const long i = (long)blockIdx.x * 256 + threadIdx.x;
  if (i < n) el[i] += e[i];
}

}  // namespace

// ============================================================ forcefield.h
extern "C" {

FF *FF_new(int N, double *q) {
  FFX *x = new FFX();
  std::memset(&x->ff, 0, sizeof(FF));
  FF *ff = &x->ff;
  ff->FFT = true;
  ff->N = N;
  ff->q = q;
  ff->maxLevel = -1;
  ff->orderAcc = -1;
  ff->topGridDim[0] = ff->topGridDim[1] = ff->topGridDim[2] = -1;
  ff->cutoff = -1;
  ff->estErr = -1;
  return ff;
}

void FF_set_orderAcc(FF *ff, int orderAcc) { ff->orderAcc = orderAcc; }
void FF_set_maxLevel(FF *ff, int maxLevel) { ff->maxLevel = maxLevel; }
void FF_set_topGridDim(FF *ff, int topGridDim[3]) {
  for (int a = 0; a < 3; ++a) ff->topGridDim[a] = topGridDim[a];
}
void FF_set_estErr(FF *ff, double estErr) { ff->estErr = estErr; }
double FF_get_estErr(FF *ff) { return ff->estErr; }
void FF_set_cutoff(FF *ff, double cutoff) { ff->cutoff = cutoff; }
double FF_get_cutoff(FF *ff) { return ff->cutoff; }
// false selects the reference's DFT build (forcefield0.c:209-215): its own
// default grid and level rules, no rounding to 7-smooth sizes; the top level
// is the same circular convolution either way.
void FF_set_FFT(FF *ff, bool FFT) { ff->FFT = FFT; }
int FF_get_orderAcc(FF *ff) { return ff->orderAcc; }
int FF_get_maxLevel(FF *ff) { return ff->maxLevel; }
void FF_get_topGridDim(FF *ff, int topGridDim[3]) {
  for (int a = 0; a < 3; ++a) topGridDim[a] = ff->topGridDim[a];
}
bool FF_get_FFT(FF *ff) { return ff->FFT; }

// margin widens only the neighbour-list statistics, as in the reference (pairs
// are re-binned every evaluation, so no list is kept).
void FF_build(FF *ff, int N, double (*position)[3], double edges[3][3], double margin) {
  if (!ff) return;
  if (build_impl(ff, N, position, edges, margin) != MSM_OK)
    std::fprintf(stderr, "msmpb_hip: %s\n", g_err.c_str());
}

void FF_rebuild(FF *ff, double edges[3][3], double (*position)[3], double margin) {
  if (check_built(ff) != MSM_OK) {
    std::fprintf(stderr, "msmpb_hip: %s\n", g_err.c_str());
    return;
  }
  FFX *x = X(ff);
  std::string err;
  Engine &E = *x->eng;
  if (!msm::rebuild_plan(E.plan, edges, err, /*stencils=*/false) || rebuild_box(E) != MSM_OK) {
    if (!err.empty()) g_err = "FF_rebuild: " + err;
    x->built = false;
    std::fprintf(stderr, "msmpb_hip: %s\n", g_err.c_str());
    return;
  }
  publish_plan(x);
  if (position) {
    long long cnt = 0;
    if (stage_host_inputs(x, position) == MSM_OK &&
        count_pairs_dev(x, E.d_pos, E.d_q, margin, &cnt) == MSM_OK) {
      ff->nlist_interaction_count = (int)cnt;
      x->margin = margin;
      publish_nlist_len(x);
    }
  }
}

double FF_energy(FF *ff, double (*force)[3], double (*position)[3], double *weight) {
  if (weight && ff && ff->maxLevel > 0) {
    for (int l = 0; l <= ff->maxLevel; ++l)
      if (weight[l] != 1.0) {
        fail(MSM_ERR_ARG, "FF_energy: level weights other than 1 are not supported");
        std::fprintf(stderr, "msmpb_hip: %s\n", g_err.c_str());
        return NAN;
      }
  }
  double E = NAN;
  if (energy_host(ff, force, position, nullptr, &E) != MSM_OK) {
    std::fprintf(stderr, "msmpb_hip: %s\n", g_err.c_str());
    return NAN;
  }
  ff->errEst = FF_get_errEst(ff);
  return E;
}

void FF_delete(FF *ff) {
  if (!ff) return;
  FFX *x = X(ff);
  delete x->eng;
  free(x->h_omegap);
  delete x;
}

double FF_get_deltaF(FF *ff) {
  // forcefield0.c:521-551: the same a-priori model with fitted constants C
  static const double C[4] = {0.825169, 5.16701, 73.6508, 1977.23};
  int nu = ff->orderAcc;
  double Q2 = 0;
  for (int i = 0; i < ff->N; ++i) Q2 += ff->q[i] * ff->q[i];
  double Fref = (Q2 / (double)ff->N) / std::pow(ff->detA / (double)ff->N, 5. / 6.);
  return C[(nu - 4) / 2] * my_errest(ff, nu) * Fref / std::pow(ff->cutoff, nu);
}

double _FF_get_errEst(FF *ff, int nu) { return my_errest(ff, nu); }

double FF_get_errEst(FF *ff) { return my_errest(ff, ff->orderAcc) / std::pow(ff->cutoff, ff->orderAcc); }

// tic/toc stack of CPU-clock timers (forcefield1.c:35-56)
double msm4g_tictocmanager(int push) {
  static clock_t stack[100];
  static int top = 0;
  if (push) {
    if (top < 100) stack[top++] = clock();
    return 0.0;
  }
  if (top == 0) return 0.0;
  clock_t now = clock();
  return (double)(now - stack[--top]) / CLOCKS_PER_SEC;
}
void msm4g_tic(void) { msm4g_tictocmanager(1); }
double msm4g_toc(void) { return msm4g_tictocmanager(0); }

double invert(Matrix *A) {
  double a[3][3], ai[3][3];
  std::memcpy(a, A, sizeof a);
  double d = msm::invert3(a, ai);
  std::memcpy(A, ai, sizeof ai);
  return d;
}

void omegap(FF *ff) {
  // recompute the anti-blurring weights of a built FF (forcefield0.c:726-767)
  if (check_built(ff) != MSM_OK) return;
  FFX *x = X(ff);
  std::string err;
  msm::Plan &p = x->eng->plan;
  double edges[3][3];
  std::memcpy(edges, p.A, sizeof edges);
  if (msm::rebuild_plan(p, edges, err)) publish_plan(x);
}

Vector prod(Matrix m, Vector v) {
  Vector r = {m.xx * v.x + m.xy * v.y + m.xz * v.z, m.yx * v.x + m.yy * v.y + m.yz * v.z,
              m.zx * v.x + m.zy * v.y + m.zz * v.z};
  return r;
}

Vector prodT(Matrix m, Vector v) {
  Vector r = {m.xx * v.x + m.yx * v.y + m.zx * v.z, m.xy * v.x + m.yy * v.y + m.zy * v.z,
              m.xz * v.x + m.yz * v.y + m.zz * v.z};
  return r;
}

void grid2grid(FF *ff, const int l, const Triple gd, double *el, const double *ql, const Triple sd,
               const double *kh) {
  // reference layout (forcefield0.c:898-931): ql padded by
  // max(min(dmax, M), nu) per axis with the interior at max(pad/2, nu/2)
  if (check_built(ff) != MSM_OK) return;
  FFX *x = X(ff);
  Engine &E = *x->eng;
  int nu = E.plan.nu, dmax = 2 * E.plan.nLim + 1;
  int g[3] = {gd.x, gd.y, gd.z}, pad[3], of[3];
  for (int a = 0; a < 3; ++a) {
    pad[a] = std::max(std::min(dmax, g[a]), nu);
    of[a] = std::max(pad[a] / 2, nu / 2);
  }
  long n = (long)g[0] * g[1] * g[2], ns = (long)sd.x * sd.y * sd.z;
  std::vector<double> q(n), e(n);
  for (int i = 0; i < g[0]; ++i)
    for (int j = 0; j < g[1]; ++j)
      for (int k = 0; k < g[2]; ++k)
        q[((long)i * g[1] + j) * g[2] + k] =
            ql[((long)(i + of[0]) * (g[1] + pad[1]) + (j + of[1])) * (g[2] + pad[2]) + k + of[2]];
  double *dq = nullptr, *de = nullptr, *dk = nullptr;
  bool ok = hipMalloc(&dq, n * 8) == hipSuccess && hipMalloc(&de, n * 8) == hipSuccess &&
            hipMalloc(&dk, ns * 8) == hipSuccess;
  if (ok) {
    hipMemcpy(dq, q.data(), n * 8, hipMemcpyHostToDevice);
    hipMemcpy(dk, kh, ns * 8, hipMemcpyHostToDevice);
    msm::launch_level_fp64(E.P, {g[0], g[1], g[2]}, {1, 1, 1}, {sd.x, sd.y, sd.z}, dk, dq, nullptr,
                           de, E.stream);
    hipMemcpyAsync(e.data(), de, n * 8, hipMemcpyDeviceToHost, E.stream);
    ok = hipStreamSynchronize(E.stream) == hipSuccess;
  }
  if (dq) hipFree(dq);
  if (de) hipFree(de);
  if (dk) hipFree(dk);
  if (!ok) {
    std::fprintf(stderr, "msmpb_hip: grid2grid failed\n");
    return;
  }
  for (long m = 0; m < n; ++m) el[m] += e[m];
  (void)l;
}

int msm_grid2grid_device_noalloc(FF *ff, int l, Triple gd, double *d_el, const double *d_ql,
                                 Triple sd, const double *d_kh, void *stream) {
  // grid2grid (forcefield0.c:898-931) on device pointers, with its two
  // scratch grids in the library's static device array
  TRY(check_built(ff));
  (void)l;
  if (!d_el || !d_ql || !d_kh) return fail(MSM_ERR_ARG, "grid2grid: NULL device pointer");
  const int g[3] = {gd.x, gd.y, gd.z}, st[3] = {sd.x, sd.y, sd.z};
  for (int a = 0; a < 3; ++a)
    if (g[a] < 1 || st[a] < 1) return fail(MSM_ERR_ARG, "grid2grid: grid and stencil dims must be positive");
  const long n = (long)g[0] * g[1] * g[2];
  if (n > kNoAllocGrid)
    return fail(MSM_ERR_ARG, "grid2grid: " + std::to_string(n) + " grid points exceed the " +
                                 std::to_string(kNoAllocGrid) +
                                 " of the no-allocation entry point's scratch");
  // the stencil kernel's LDS slab (launch_level_fp64): rows x pitch doubles
  int LZ = std::min((g[2] + 7) / 8, 16), LY = std::min(64 / LZ, g[1]);
  const long lds = 8L * (LY + st[1] - 1) * (LZ * 8 + st[2] - 1 + 8);
  if (lds > 65536) return fail(MSM_ERR_ARG, "grid2grid: stencil too large for one workgroup's LDS");
  Engine &E = *X(ff)->eng;
  double *scr = nullptr;
  TRY(scratch_on(E.device, &scr));
  const int nu = E.plan.nu, dmax = 2 * E.plan.nLim + 1;
  int pad[3], of[3];
  for (int a = 0; a < 3; ++a) {
    pad[a] = std::max(std::min(dmax, g[a]), nu);
    of[a] = std::max(pad[a] / 2, nu / 2);
  }
  const msm::GridDims G = {g[0], g[1], g[2]};
  const hipStream_t s = static_cast<hipStream_t>(stream);
  const unsigned nb = (unsigned)((n + 255) / 256);
  double *q = scr, *e = scr + kNoAllocGrid;
  k_unpad<<<nb, 256, 0, s>>>(G, make_int3(pad[0], pad[1], pad[2]), make_int3(of[0], of[1], of[2]),
                             d_ql, q);
  msm::launch_level_fp64(E.P, G, {1, 1, 1}, {st[0], st[1], st[2]}, d_kh, q, nullptr, e, s);
  k_accumulate<<<nb, 256, 0, s>>>(n, e, d_el);
  HIP_TRY(hipGetLastError());
  return MSM_OK;
}

// ============================================================ msmpb_hip.h
const char *FF_last_error(void) { return g_err.c_str(); }

int FF_set_precision(FF *ff, int precision) {
  if (!ff) return fail(MSM_ERR_ARG, "FF is NULL");
  if (precision != MSM_PREC_FP64 && precision != MSM_PREC_MIXED)
    return fail(MSM_ERR_ARG, "precision must be MSM_PREC_FP64 or MSM_PREC_MIXED");
  X(ff)->precision = precision;
  return MSM_OK;
}
int FF_get_precision(FF *ff) { return ff ? X(ff)->precision : -1; }

int FF_set_method(FF *ff, int method, double tolDir) {
  if (!ff) return fail(MSM_ERR_ARG, "FF is NULL");
  if (method != MSM_METHOD_MSM && method != MSM_METHOD_PME)
    return fail(MSM_ERR_ARG, "method must be MSM_METHOD_MSM or MSM_METHOD_PME");
  X(ff)->method = method;
  X(ff)->tolDir = tolDir;
  return MSM_OK;
}

double FF_get_beta(FF *ff) {
  if (!ff || !X(ff)->eng) return 0.0;
  return X(ff)->eng->plan.beta;
}

int FF_set_layout_compat(FF *ff, int layout) {
  if (!ff) return fail(MSM_ERR_ARG, "FF is NULL");
  if (layout != MSM_LAYOUT_FIXED && layout != MSM_LAYOUT_ASIS)
    return fail(MSM_ERR_ARG, "layout must be MSM_LAYOUT_FIXED or MSM_LAYOUT_ASIS");
  if (layout == MSM_LAYOUT_ASIS && X(ff)->method == MSM_METHOD_PME)
    return fail(MSM_ERR_ARG, "the as-is index layout does not apply to the PME method");
  FFX *x = X(ff);
  if (layout == MSM_LAYOUT_ASIS && x->built && x->eng && !msm::reference_images_hold(x->eng->plan))
    return fail(MSM_ERR_ARG,
                "FF_set_layout_compat: the as-is layout keeps the reference's image rules, which "
                "need the cutoff below half the smallest face-to-face width of a triclinic box");
  x->layout = layout;
  if (x->built && x->eng && x->eng->plan.method == msm::kMsm) {
    // the layout also selects the level kernels' image range: regenerate
    // the stencils of a built FF when that changes them
    msm::Plan &p = x->eng->plan;
    const double old[3] = {p.kas[0], p.kas[1], p.kas[2]};
    p.complete_images = layout != MSM_LAYOUT_ASIS;
    msm::kappa_image_bound(p);
    if (p.L >= 2 && (old[0] != p.kas[0] || old[1] != p.kas[1] || old[2] != p.kas[2])) {
      std::string err;
      double edges[3][3];
      std::memcpy(edges, p.A, sizeof edges);
      if (!msm::rebuild_plan(p, edges, err, /*stencils=*/false) || rebuild_box(*x->eng) != MSM_OK) {
        x->built = false;
        if (!err.empty()) g_err = "FF_set_layout_compat: " + err;
        return MSM_ERR_ARG;
      }
      publish_plan(x);
    }
    publish_nlist_len(x);
  }
  return MSM_OK;
}

int FF_set_stage_timing(FF *ff, int enable) {
  if (!ff) return fail(MSM_ERR_ARG, "FF is NULL");
  X(ff)->timing = enable;
  return MSM_OK;
}

int FF_get_energy_terms(FF *ff, double terms[3]) {
  TRY(check_built(ff));
  std::memcpy(terms, X(ff)->eng->last_terms, 3 * sizeof(double));
  return MSM_OK;
}

double FF_energy_potential(FF *ff, double (*force)[3], double (*position)[3], double *potential) {
  double E = NAN;
  if (!potential) {
    fail(MSM_ERR_ARG, "FF_energy_potential: potential is NULL");
    return NAN;
  }
  if (energy_host(ff, force, position, potential, &E) != MSM_OK) return NAN;
  return E;
}

int msm_eval_device(FF *ff, const double *d_pos, const double *d_q, double *d_energy,
                    double *d_force, double *d_phi, void *stream) {
  TRY(check_built(ff));
  if (!d_pos || !d_q || !d_energy) return fail(MSM_ERR_ARG, "msm_eval_device: NULL input");
  FFX *x = X(ff);
  hipStream_t s = (hipStream_t)stream;  // NULL: the default stream, as in HIP
  Timing &T = x->eng->T;
  T.on = x->timing != 0;
  TRY(eval_chain(x, d_pos, d_q, d_energy, d_force, d_phi, s, T));
  return MSM_OK;
}

int msm_stage_times(FF *ff, double *out, int n) {
  TRY(check_built(ff));
  Engine &E = *X(ff)->eng;
  Timing &T = E.T;
  for (int k = 0; k < n; ++k) out[k] = 0.0;
  HIP_TRY(hipStreamSynchronize(E.stream2));
  for (auto &v : T.iv) {
    HIP_TRY(hipEventSynchronize(v.b));
    float ms = 0;
    HIP_TRY(hipEventElapsedTime(&ms, v.a, v.b));
    for (int j = 0; j < kNumStages && j < n; ++j)
      if (v.tag == kStageNames[j]) out[j] += ms * 1e-3;
  }
  int evals = E.timed_evals;
  T.clear();
  E.timed_evals = 0;
  return evals;
}

const char *msm_stage_name(int k) { return (k >= 0 && k < kNumStages) ? kStageNames[k] : ""; }

int msm_forward_device(FF *ff, const double *d_pos, const double *d_q, double *d_energy, int keep,
                       void *stream) {
  TRY(check_built(ff));
  if (!d_pos || !d_q || !d_energy) return fail(MSM_ERR_ARG, "msm_forward_device: NULL input");
  FFX *x = X(ff);
  Engine &E = *x->eng;
  hipStream_t s = (hipStream_t)stream;  // NULL: the default stream, as in HIP
  if (keep != MSM_KEEP_NONE && keep != MSM_KEEP_ALL && keep != MSM_KEEP_PAIRS)
    return fail(MSM_ERR_ARG, "msm_forward_device: keep must be MSM_KEEP_NONE, _ALL or _PAIRS");
  Timing T;
  // with state kept for the adjoint the pair kernel also writes its
  // per-atom sums (the short-range parts of -dE/dr and dE/dq)
  const bool kp = keep != MSM_KEEP_NONE;
  stage_prepare(E, d_pos, d_q, s, E.ev_fork);
  TRY(fork_pairs(E, x->precision, kp ? E.d_Fp : nullptr, kp ? E.d_phip : nullptr, T));
  TRY(stage_grids(E, x->precision, x->layout, T, s));
  stage_grid_energy(E, x->layout, s);
  HIP_TRY(hipStreamWaitEvent(s, E.ev_join, 0));
  stage_energy(E, nullptr, nullptr, d_energy, s);
  HIP_TRY(hipGetLastError());
  E.kept = keep == MSM_KEEP_ALL;
  E.kept_pairs = kp;
  E.kept_pos = d_pos;
  E.kept_q = d_q;
  return MSM_OK;
}

int msm_adjoint_device(FF *ff, const double *d_pos, const double *d_q, double *d_force,
                       double *d_phi, int recompute, void *stream) {
  TRY(check_built(ff));
  if (!d_pos || !d_q) return fail(MSM_ERR_ARG, "msm_adjoint_device: NULL input");
  FFX *x = X(ff);
  Engine &E = *x->eng;
  hipStream_t s = (hipStream_t)stream;  // NULL: the default stream, as in HIP
  Timing T;
  const bool same = E.kept_pos == d_pos && E.kept_q == d_q;
  if (recompute || !E.kept || !same) {
    // grid levels are not stored: rebuild the ordering and e^1 from scratch
    stage_prepare(E, d_pos, d_q, s);
    msm::launch_fill_int(E.d_flags, 1, 0, s);  // no energy reduction follows to clear it
    TRY(stage_grids(E, x->precision, x->layout, T, s));
    E.kept = false;
  }
  if (E.kept_pairs && same) {
    // the forward's pair kernel left the short-range parts (indexed by atom,
    // independent of the ordering): interpolate the long-range parts, then
    // add them (the reduction rides along and rewrites E.d_out only)
    if (d_force || d_phi) {
      msm::launch_gather(E.P, E.gd[1], E.d_spos, E.d_sidx, E.d_start, E.d_el[1], E.d_scal,
                         d_force, d_phi, false, s);
      msm::launch_combine(E.P, d_force ? E.d_Fp : nullptr, d_phi ? E.d_phip : nullptr, d_force,
                          d_phi, finalize_args(E, E.d_out), s);
    }
  } else {
    stage_pairs(E, x->precision, d_force, d_phi, s);  // the pair kernel writes, gather adds
    if (d_force || d_phi)
      msm::launch_gather(E.P, E.gd[1], E.d_spos, E.d_sidx, E.d_start, E.d_el[1], E.d_scal,
                         d_force, d_phi, true, s);
  }
  HIP_TRY(hipGetLastError());
  return MSM_OK;
}

int msm_grid_dims(FF *ff, int level, int dims[3]) {
  TRY(check_built(ff));
  Engine &E = *X(ff)->eng;
  if (level < 1 || level > E.plan.L) return fail(MSM_ERR_ARG, "level out of range");
  dims[0] = E.gd[level].x, dims[1] = E.gd[level].y, dims[2] = E.gd[level].z;
  return MSM_OK;
}

int msm_stencil_dims(FF *ff, int level, int dims[3]) {
  TRY(check_built(ff));
  Engine &E = *X(ff)->eng;
  if (level < 1 || level > E.plan.L) return fail(MSM_ERR_ARG, "level out of range");
  dims[0] = E.sd[level].x, dims[1] = E.sd[level].y, dims[2] = E.sd[level].z;
  return MSM_OK;
}

int msm_copy_grid(FF *ff, int level, int which, double *host_out) {
  TRY(check_built(ff));
  Engine &E = *X(ff)->eng;
  if (level < 1 || level > E.plan.L) return fail(MSM_ERR_ARG, "level out of range");
  const double *src = which ? E.d_el[level] : E.d_ql[level];
  HIP_TRY(hipMemcpyAsync(host_out, src, E.gd[level].size() * sizeof(double),
                         hipMemcpyDeviceToHost, E.stream));
  HIP_TRY(hipStreamSynchronize(E.stream));
  return MSM_OK;
}

int msm_copy_stencil(FF *ff, int level, double *host_out) {
  TRY(check_built(ff));
  Engine &E = *X(ff)->eng;
  if (level < 1 || level > E.plan.L) return fail(MSM_ERR_ARG, "level out of range");
  const std::vector<double> &k = E.plan.khat[level];
  std::memcpy(host_out, k.data(), k.size() * sizeof(double));
  return MSM_OK;
}

int msm_count_pairs(FF *ff, double (*position)[3], long long *count) {
  TRY(check_built(ff));
  if (!position || !count) return fail(MSM_ERR_ARG, "msm_count_pairs: NULL argument");
  FFX *x = X(ff);
  TRY(stage_host_inputs(x, position));
  return count_pairs_dev(x, x->eng->d_pos, x->eng->d_q, 0.0, count);
}


// ============================================================ msm_bridge.h
void *msmb_create(int N, double *q, int method, int orderAcc, int maxLevel,
                  const int *topGridDim, double cutoff, double tolDir) {
  FF *ff = FF_new(N, q);
  if (orderAcc > 0) ff->orderAcc = orderAcc;
  if (maxLevel > 0) ff->maxLevel = maxLevel;
  if (topGridDim && topGridDim[0] > 0)
    for (int a = 0; a < 3; ++a) ff->topGridDim[a] = topGridDim[a];
  if (cutoff > 0) ff->cutoff = cutoff;
  if (FF_set_method(ff, method, tolDir) != MSM_OK) {
    FF_delete(ff);
    return nullptr;
  }
  return ff;
}

int msmb_build(void *h, double (*position)[3], double edges[3][3]) {
  if (!h) return fail(MSM_ERR_ARG, "handle is NULL");
  FF *ff = (FF *)h;
  return build_impl(ff, ff->N, position, edges);
}

int msmb_rebuild(void *h, double edges[3][3], double (*position)[3]) {
  if (!h) return fail(MSM_ERR_ARG, "handle is NULL");
  FF *ff = (FF *)h;
  g_err.clear();
  FF_rebuild(ff, edges, position, 0.0);
  return X(ff)->built ? MSM_OK : MSM_ERR_ARG;
}

double msmb_energy(void *h, double (*force)[3], double (*position)[3], double *potential) {
  double E = NAN;
  if (energy_host((FF *)h, force, position, potential, &E) != MSM_OK) return NAN;
  return E;
}

int msmb_configure(void *h, int precision, int timing) {
  if (!h) return fail(MSM_ERR_ARG, "handle is NULL");
  TRY(FF_set_precision((FF *)h, precision));
  return FF_set_stage_timing((FF *)h, timing);
}

int msmb_info(void *h, int ints[5], double dbl[5]) {
  TRY(check_built((FF *)h));
  const FF *ff = (const FF *)h;
  const msm::Plan &p = X((FF *)h)->eng->plan;
  ints[0] = p.nu;
  for (int a = 0; a < 3; ++a) ints[1 + a] = p.M[p.L][a];
  ints[4] = ff->nlist_interaction_count;
  dbl[0] = p.aCut[0];
  dbl[1] = p.beta;
  dbl[2] = p.coeff1;
  dbl[3] = p.coeff2;
  dbl[4] = p.detA;
  return MSM_OK;
}

int msmb_tables(void *h, const double **Q, const double *cL[3], const double **khat) {
  TRY(check_built((FF *)h));
  const FF *ff = (const FF *)h;
  if (Q) *Q = ff->Q;
  if (cL)
    for (int a = 0; a < 3; ++a) cL[a] = ff->cL[a];
  if (khat) *khat = ff->khat[ff->maxLevel];
  return MSM_OK;
}

int msmb_times(void *h, double t[6]) {
  if (!h) return fail(MSM_ERR_ARG, "handle is NULL");
  const FF *ff = (const FF *)h;
  const int L = ff->maxLevel > 0 && ff->maxLevel <= 9 ? ff->maxLevel : 1;
  double st = 0;
  for (int l = 1; l <= L; ++l) st += ff->time_stencil[l];
  t[0] = ff->time_nlist;
  t[1] = st;
  t[2] = ff->time_partcl2partcl;
  t[3] = ff->time_anterpolation;
  t[4] = ff->time_grid2grid[L];
  t[5] = ff->time_interpolation;
  return MSM_OK;
}

void msmb_delete(void *h) { FF_delete((FF *)h); }

}  // extern "C"
