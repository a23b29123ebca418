// Host code under AddressSanitizer and UndefinedBehaviorSanitizer (CPU).
//
// Builds and rebuilds evaluation plans with the product's plan code
// (msmpb_amd/csrc/msm_plan.cpp: parameter rules, tables, stencils, the PME
// method) for every order and level count, including inputs that must be
// rejected, and runs both CPU oracles on small systems.  Built and run by
// tests/test_sanitize.py; prints "ok" on success.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "../../msmpb_amd/csrc/msm_plan.h"
extern "C" {
#include "../../oracle/msm_oracle.h"
#include "../../oracle/pme_oracle.h"
}

int main() {
  const int N = 240;
  const double box[3] = {12.0, 13.0, 14.5};
  std::mt19937_64 rng(11);
  std::uniform_real_distribution<double> u(0.0, 1.0);
  std::vector<double> pos(3 * N), q(N);
  double qs = 0;
  for (int i = 0; i < N; ++i) {
    for (int a = 0; a < 3; ++a) pos[3 * i + a] = u(rng) * box[a];
    q[i] = u(rng) - 0.5;
    qs += q[i];
  }
  for (double &v : q) v -= qs / N;
  double edges[3][3] = {{box[0], 0, 0}, {0, box[1], 0}, {0, 0, box[2]}};
  double scaled[3][3] = {{box[0] * 1.01, 0, 0}, {0, box[1] * 1.01, 0}, {0, 0, box[2] * 1.01}};
  int built = 0, refused = 0;
  for (int method : {msm::kMsm, msm::kPme})
    for (int nu : {3, 4, 6, 8, 10, 12})
      for (int L : {1, 2, 3})
        for (int M : {-1, 4, 8, 11, 12})
          for (bool fft : {true, false}) {
            msm::BuildRequest r;
            r.N = N;
            r.q = q.data();
            r.fft = fft;
            std::memcpy(r.edges, edges, sizeof edges);
            r.nu = nu;
            r.L = L;
            for (int a = 0; a < 3; ++a) r.Mtop[a] = M;
            r.a0 = method == msm::kPme ? 6.0 : 5.0;
            r.method = method;
            msm::Plan p;
            std::string err;
            if (!msm::build_plan(r, p, err)) {
              if (err.empty()) return 2;
              ++refused;
              continue;
            }
            ++built;
            if (!msm::rebuild_plan(p, scaled, err, true) || !msm::rebuild_plan(p, edges, err, false))
              return 3;
          }
  char err[512];
  int dft_built = 0;
  for (int nu : {4, 6, 8, 10}) {
    int Mt[3] = {8, 8, 8};
    orc_t *o = orc_build(N, q.data(), pos.data(), &edges[0][0], nu, 2, Mt, 5.0, -1, 0, 1, err);
    if (!o) return 4;
    std::vector<double> F(3 * N), phi(N);
    double terms[3];
    double E = orc_energy(o, pos.data(), q.data(), F.data(), phi.data(), terms);
    orc_free(o);
    // the DFT build: automatic levels and grid (refused, with a message, where
    // the grid is too coarse for the order), and its as-is top-level read
    for (int compat : {0, 1})
      for (int L : {-1, 1}) {
        int M11[3] = {11, 11, 11};
        err[0] = 0;
        orc_t *d = orc_build(N, q.data(), pos.data(), &edges[0][0], nu, L,
                             L < 0 ? nullptr : M11, 5.0, -1, compat, 0, err);
        if (!d) {
          if (!err[0]) return 8;
          continue;
        }
        ++dft_built;
        if (!std::isfinite(orc_energy(d, pos.data(), q.data(), F.data(), phi.data(), terms)))
          return 9;
        orc_free(d);
      }
    int Mp[3] = {16, 18, 20};
    pmo_t *p = pmo_build(N, q.data(), pos.data(), &edges[0][0], nu, Mp, 6.0, 1e-6, err);
    if (!p) return 5;
    double Ep = pmo_energy(p, pos.data(), q.data(), F.data(), phi.data(), terms);
    pmo_free(p);
    if (!std::isfinite(E) || !std::isfinite(Ep)) return 6;
  }
  std::printf("ok built=%d refused=%d dft=%d\n", built, refused, dft_built);
  return built > 0 && refused > 0 && dft_built > 0 ? 0 : 7;
}
