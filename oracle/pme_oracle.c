/*
 * pme_oracle.c -- CPU restatement of msmpb's smooth particle-mesh Ewald
 * comparator (code/pme0.c, code/pme1.c of hkayabilisim/msmpb).
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker for the PME method of
 * the HIP library in msmpb_amd/: only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load it.  The product library never links
 * it.
 *
 * The arithmetic follows the reference stage by stage (citations per
 * function), written independently:
 *   - B-spline weights by the Cox-de Boor recursion of the cardinal B-spline
 *     instead of the reference's stored polynomial pieces (same values to
 *     rounding);
 *   - the real-space pairs by a cell list instead of the stored neighbour
 *     list (the same pairs: i < j with the nearest image inside a0, every
 *     image inside a0 summed);
 *   - the reciprocal convolution by a separable naive DFT instead of FFTW.
 */
#include "pme_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

struct pmo {
  int N, nu, M[3];
  double A[3][3], Ai[3][3], detA;
  double a0, tolDir, beta;
  double coeff1, coeff2;
  double *cL[3];   /* 1 / Phi^(2 pi k / M)^2, k = 0 .. M-1 */
  double *khat;    /* M^3 */
  double *qg, *eg; /* last evaluation's grids */
  int pair_count;
};

static double real_space(const pmo_t *o, const double *pos, const double *q, double *F, double *phi,
                         int *count);

static double pi_(void) { return 4.0 * atan(1.0); }
static int pmod(int a, int m) { int r = a % m; return r < 0 ? r + m : r; }

static double inv3(const double a[3][3], double ai[3][3]) {
  double c[3][3];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++)
      c[i][j] = a[(j + 1) % 3][(i + 1) % 3] * a[(j + 2) % 3][(i + 2) % 3] -
                a[(j + 2) % 3][(i + 1) % 3] * a[(j + 1) % 3][(i + 2) % 3];
  double det = a[0][0] * c[0][0] + a[0][1] * c[1][0] + a[0][2] * c[2][0];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) ai[i][j] = c[i][j] / det;
  return fabs(det);
}

/* M_nu(i + t) for i = 0 .. nu-1 and t in [0, 1): the recursion
 * M_k(x) = (x M_{k-1}(x) + (k - x) M_{k-1}(x - 1)) / (k - 1), M_1 = 1 on
 * [0, 1); derivative M_nu'(x) = M_{nu-1}(x) - M_{nu-1}(x - 1).  The
 * reference evaluates the same function from polynomial pieces
 * (pme1.c:274-298 with the table of pme0.c:265-296). */
void pmo_bspline(int nu, double t, double *w, double *dw) {
  double b[12] = {0}, prev[12] = {0};
  b[0] = 1.0;
  for (int k = 2; k <= nu; k++) {
    memcpy(prev, b, sizeof b);
    if (k == nu && dw)
      for (int i = 0; i < nu; i++) dw[i] = prev[i] - (i > 0 ? prev[i - 1] : 0.0);
    for (int i = 0; i < k; i++) {
      double x = i + t;
      b[i] = (x * prev[i] + (k - x) * (i > 0 ? prev[i - 1] : 0.0)) / (k - 1);
    }
  }
  for (int i = 0; i < nu; i++) w[i] = b[i];
}

/* erfc(b) = c by Newton's method from b = 0, stopped when the residual no
 * longer shrinks (pme0.c:209-224); returns b. */
static double newton_erfc(double c) {
  double b = 0.0, res = erfc(b) - c, old = 2.0 * res;
  while (fabs(res) < fabs(old)) {
    double d = -2.0 / sqrt(pi_()) * exp(-b * b);
    b -= res / d;
    old = res;
    res = erfc(b) - c;
  }
  return b;
}

/* Default grid from the reference spacing table (pme0.c:226-263). */
static int default_grid(pmo_t *o, char *err) {
  static const double href[2][9] = {
      {0.70, 0.77, 0.83, 0.91, 1.00, 1.11, 1.21, 1.33, 1.48},
      {1.05, 1.18, 1.29, 1.43, 1.54, 1.67, 1.82, 2.00, 2.22}};
  int i = (o->nu - 4) / 2, j = (int)round((o->a0 - 6.0) / 0.5);
  if (i > 1 || j < 0 || j > 8) {
    if (err) snprintf(err, 256, "no default PME grid for order %d and cutoff %g", o->nu, o->a0);
    return -1;
  }
  double beta_ref = newton_erfc(1e-5 * o->a0) / o->a0;
  double hstar = pow(o->detA / o->N, 1.0 / 3.0);
  double h = pow(o->tolDir / (1e-5 * hstar), 1.0 / o->nu) * (beta_ref / o->beta) * href[i][j];
  for (int a = 0; a < 3; a++) {
    double ea = sqrt(o->A[a][0] * o->A[a][0] + o->A[a][1] * o->A[a][1] + o->A[a][2] * o->A[a][2]);
    o->M[a] = (int)ceil(ea / h);
  }
  return 0;
}

/* cL (pme0.c:444-471): Phi_n = M_nu(nu/2 + n), Phi^(th) = Phi_0 + 2 sum_n
 * Phi_n cos(n th), c(k) = 1 / Phi^(2 pi k / M)^2. */
static void build_cL(pmo_t *o) {
  double w[12];
  pmo_bspline(o->nu, 0.0, w, NULL); /* w[i] = M_nu(i) */
  for (int a = 0; a < 3; a++) {
    int Ma = o->M[a];
    free(o->cL[a]);
    o->cL[a] = (double *)malloc(Ma * sizeof(double));
    for (int k = 0; k < Ma; k++) {
      double th = 2.0 * pi_() * k / Ma, v = w[o->nu / 2];
      for (int n = 1; n < o->nu / 2; n++) v += 2.0 * w[o->nu / 2 + n] * cos(n * th);
      o->cL[a][k] = 1.0 / (v * v);
    }
  }
}

/* Box-dependent parts (FF_rebuild, pme0.c:365-412 and dALp1 :473-503):
 * constants and the reciprocal kernel
 * khat(k) = exp(-pi^2 |m|^2 / beta^2) / (pi V |m|^2) cL_x cL_y cL_z M^3,
 * m = Ai^T k over the centred frequencies, 0 at k = 0. */
static void box_dependent(pmo_t *o) {
  double pi = pi_(), a0 = o->a0, b = o->beta;
  int lim[3];
  for (int a = 0; a < 3; a++) {
    double as = sqrt(o->Ai[0][a] * o->Ai[0][a] + o->Ai[1][a] * o->Ai[1][a] + o->Ai[2][a] * o->Ai[2][a]);
    lim[a] = (int)ceil(a0 * as) + 1;
  }
  o->coeff1 = -2.0 * b / sqrt(pi);
  for (int x = -lim[0]; x <= lim[0]; x++)
    for (int y = -lim[1]; y <= lim[1]; y++)
      for (int z = -lim[2]; z <= lim[2]; z++) {
        if (!x && !y && !z) continue;
        double r2 = 0;
        for (int a = 0; a < 3; a++) {
          double v = o->A[a][0] * x + o->A[a][1] * y + o->A[a][2] * z;
          r2 += v * v;
        }
        if (r2 >= a0 * a0) continue;
        double r = sqrt(r2);
        o->coeff1 += erfc(b * r) / r;
      }
  o->coeff2 = pi / (b * b * o->detA);
  build_cL(o);
  int *g = o->M;
  size_t n = (size_t)g[0] * g[1] * g[2];
  free(o->khat);
  o->khat = (double *)calloc(n, sizeof(double));
  for (int i = 0; i < g[0]; i++)
    for (int j = 0; j < g[1]; j++)
      for (int k = 0; k < g[2]; k++) {
        int kv[3] = {i <= (g[0] - 1) / 2 ? i : i - g[0], j <= (g[1] - 1) / 2 ? j : j - g[1],
                     k <= (g[2] - 1) / 2 ? k : k - g[2]};
        double m2 = 0;
        for (int a = 0; a < 3; a++) {
          double v = o->Ai[0][a] * kv[0] + o->Ai[1][a] * kv[1] + o->Ai[2][a] * kv[2];
          m2 += v * v;
        }
        if (m2 == 0.0) continue;
        double c = o->cL[0][i] * o->cL[1][j] * o->cL[2][k];
        o->khat[((size_t)i * g[1] + j) * g[2] + k] =
            exp(-pi * pi * m2 / (b * b)) / (pi * o->detA * m2) * c * (double)n;
      }
}

pmo_t *pmo_build(int N, const double *q, const double *pos, const double *edges, int nu,
                 const int *M, double a0, double tolDir, char *err) {
  (void)q;
  pmo_t *o = (pmo_t *)calloc(1, sizeof(pmo_t));
  o->N = N;
  memcpy(o->A, edges, sizeof o->A);
  o->detA = inv3(o->A, o->Ai);
  o->nu = nu > 0 ? nu : 6;
  o->a0 = a0 > 0 ? a0 : 8.0;
  o->tolDir = tolDir > 0 ? tolDir : 1e-5;
  if (o->nu < 4 || o->nu > 10 || o->nu % 2 || !(o->detA > 0)) {
    if (err) snprintf(err, 256, "invalid order %d or singular box", o->nu);
    pmo_free(o);
    return NULL;
  }
  o->beta = newton_erfc(o->tolDir * o->a0) / o->a0;
  if (M && M[0] > 0) {
    for (int a = 0; a < 3; a++) o->M[a] = M[a];
  } else if (default_grid(o, err)) {
    pmo_free(o);
    return NULL;
  }
  for (int a = 0; a < 3; a++)
    if (o->M[a] < o->nu) {
      if (err) snprintf(err, 256, "grid dimension %d below the order %d", o->M[a], o->nu);
      pmo_free(o);
      return NULL;
    }
  box_dependent(o);
  /* pairs i < j with the nearest image inside a0: the neighbour-list
   * statistic of neighborlist (pme0.c:21-162) */
  real_space(o, pos, NULL, NULL, NULL, &o->pair_count);
  return o;
}

/* Real-space part (partcl2partcl, pme1.c:111-180): every pair i < j with
 * the nearest image inside a0 and every image of it inside a0:
 * E += q_i q_j erfc(beta r)/r,
 * F_i += q_i q_j (-(erfc(beta r)/r + 2 beta/sqrt(pi) exp(-beta^2 r^2))/r^2) r,
 * F_j -= the same, r = r_j - r_i.  Visited by a cell list of cells at
 * least a0 wide (all pairs where the box holds fewer than 3 cells). */
static double real_space(const pmo_t *o, const double *pos, const double *q, double *F, double *phi,
                         int *count) {
  int N = o->N;
  double a0 = o->a0, a02 = a0 * a0, b = o->beta, c2 = 2.0 * b / sqrt(pi_());
  double E = 0;
  int lim[3], nc[3];
  *count = 0;
  for (int a = 0; a < 3; a++) {
    /* images p of the nearest one with |p_a| <= ceil(a0 |as_a| - 1/2)
     * (pme1.c:126-128) */
    double as = sqrt(o->Ai[0][a] * o->Ai[0][a] + o->Ai[1][a] * o->Ai[1][a] + o->Ai[2][a] * o->Ai[2][a]);
    lim[a] = (int)ceil(a0 * as - 0.5);
    /* cells at least a0 wide across: face-to-face width 1 / |row a of Ai| */
    double rs = sqrt(o->Ai[a][0] * o->Ai[a][0] + o->Ai[a][1] * o->Ai[a][1] + o->Ai[a][2] * o->Ai[a][2]);
    nc[a] = (int)floor(1.0 / (a0 * rs));
    if (nc[a] < 3) nc[a] = 1;
  }
  int ncell = nc[0] * nc[1] * nc[2];
  int *head = (int *)malloc(ncell * sizeof(int)), *next = (int *)malloc(N * sizeof(int));
  int *cell = (int *)malloc(N * 3 * sizeof(int));
  for (int c = 0; c < ncell; c++) head[c] = -1;
  for (int i = N - 1; i >= 0; i--) {
    int ci = 0;
    for (int a = 0; a < 3; a++) {
      double s = o->Ai[a][0] * pos[3 * i] + o->Ai[a][1] * pos[3 * i + 1] + o->Ai[a][2] * pos[3 * i + 2];
      s -= floor(s);
      int m = (int)(s * nc[a]);
      if (m >= nc[a]) m = nc[a] - 1;
      cell[3 * i + a] = m;
      ci = ci * nc[a] + m;
    }
    next[i] = head[ci];
    head[ci] = i;
  }
  int r1[3];
  for (int a = 0; a < 3; a++) r1[a] = nc[a] >= 3 ? 1 : 0;
  for (int i = 0; i < N; i++) {
    for (int dx = -r1[0]; dx <= r1[0]; dx++)
      for (int dy = -r1[1]; dy <= r1[1]; dy++)
        for (int dz = -r1[2]; dz <= r1[2]; dz++) {
          int cj = (pmod(cell[3 * i] + dx, nc[0]) * nc[1] + pmod(cell[3 * i + 1] + dy, nc[1])) * nc[2] +
                   pmod(cell[3 * i + 2] + dz, nc[2]);
          for (int j = head[cj]; j >= 0; j = next[j]) {
            if (j <= i) continue;
            double d[3], s[3], r0[3];
            for (int a = 0; a < 3; a++) d[a] = pos[3 * j + a] - pos[3 * i + a];
            for (int a = 0; a < 3; a++) s[a] = o->Ai[a][0] * d[0] + o->Ai[a][1] * d[1] + o->Ai[a][2] * d[2];
            for (int a = 0; a < 3; a++) s[a] = floor(s[a] + 0.5);
            for (int a = 0; a < 3; a++) r0[a] = d[a] - (o->A[a][0] * s[0] + o->A[a][1] * s[1] + o->A[a][2] * s[2]);
            if (r0[0] * r0[0] + r0[1] * r0[1] + r0[2] * r0[2] >= a02) continue;
            ++*count;
            if (!q) continue;
            double sum = 0, f[3] = {0, 0, 0};
            for (int px = -lim[0]; px <= lim[0]; px++)
              for (int py = -lim[1]; py <= lim[1]; py++)
                for (int pz = -lim[2]; pz <= lim[2]; pz++) {
                  double r[3], r2 = 0;
                  for (int a = 0; a < 3; a++) {
                    r[a] = r0[a] + o->A[a][0] * px + o->A[a][1] * py + o->A[a][2] * pz;
                    r2 += r[a] * r[a];
                  }
                  if (r2 >= a02) continue;
                  double rr = sqrt(r2), v = erfc(b * rr) / rr;
                  sum += v;
                  double cf = -(v + c2 * exp(-b * b * r2)) / r2;
                  for (int a = 0; a < 3; a++) f[a] += cf * r[a];
                }
            double qq = q[i] * q[j];
            E += qq * sum;
            for (int a = 0; a < 3; a++) {
              F[3 * i + a] += qq * f[a];
              F[3 * j + a] -= qq * f[a];
            }
            if (phi) {
              phi[i] += q[j] * sum;
              phi[j] += q[i] * sum;
            }
          }
        }
  }
  free(head);
  free(next);
  free(cell);
  return E;
}

/* Weights of atom i on the grid: fractional position t = M s (s wrapped to
 * [0, 1)), base m = floor(t), weight M_nu(n + nu/2 + t - m) at grid point
 * (m - n) mod M for n = -nu/2 .. nu/2-1 (pme1.c:183-215). */
static void atom_weights(const pmo_t *o, const double *r, int m[3], double w[3][12], double dw[3][12]) {
  for (int a = 0; a < 3; a++) {
    double s = o->Ai[a][0] * r[0] + o->Ai[a][1] * r[1] + o->Ai[a][2] * r[2];
    s -= floor(s);
    double t = o->M[a] * s, fl = floor(t);
    m[a] = (int)fl;
    pmo_bspline(o->nu, t - fl, w[a], dw[a]);
  }
}

static void dft_axis(double *re, double *im, const int *g, int axis, int sign) {
  int M = g[axis];
  long stride = axis == 0 ? (long)g[1] * g[2] : axis == 1 ? g[2] : 1;
  long n = (long)g[0] * g[1] * g[2];
  double *tr = (double *)malloc(M * sizeof(double)), *ti = (double *)malloc(M * sizeof(double));
  double *cs = (double *)malloc(M * sizeof(double)), *sn = (double *)malloc(M * sizeof(double));
  for (int k = 0; k < M; k++) {
    cs[k] = cos(2.0 * pi_() * k / M);
    sn[k] = sign * sin(2.0 * pi_() * k / M);
  }
  for (long base = 0; base < n; base++) {
    int c = (int)((base / stride) % M);
    if (c) continue; /* first element of a line */
    for (int k = 0; k < M; k++) {
      double sr = 0, si = 0;
      for (int j = 0, kj = 0; j < M; j++, kj = kj + k >= M ? kj + k - M : kj + k) {
        double xr = re[base + j * stride], xi = im[base + j * stride];
        sr += xr * cs[kj] - xi * sn[kj];
        si += xr * sn[kj] + xi * cs[kj];
      }
      tr[k] = sr, ti[k] = si;
    }
    for (int k = 0; k < M; k++) re[base + k * stride] = tr[k], im[base + k * stride] = ti[k];
  }
  free(tr);
  free(ti);
  free(cs);
  free(sn);
}

double pmo_energy(pmo_t *o, const double *pos, const double *q, double *F, double *phi, double *terms) {
  int N = o->N, nu = o->nu, *g = o->M;
  long n = (long)g[0] * g[1] * g[2];
  /* constant part (pme1.c:62-67) */
  double qs = 0, q2 = 0;
  for (int i = 0; i < N; i++) qs += q[i], q2 += q[i] * q[i];
  double Ec = 0.5 * q2 * o->coeff1 - 0.5 * qs * qs * o->coeff2;
  memset(F, 0, 3 * (size_t)N * sizeof(double));
  if (phi)
    for (int i = 0; i < N; i++) phi[i] = q[i] * o->coeff1 - qs * o->coeff2;
  int cnt;
  double Ep = real_space(o, pos, q, F, phi, &cnt);
  /* charge assignment (anterpolate, pme1.c:183-215) */
  free(o->qg);
  free(o->eg);
  o->qg = (double *)calloc(n, sizeof(double));
  o->eg = (double *)calloc(n, sizeof(double));
  for (int i = 0; i < N; i++) {
    int m[3];
    double w[3][12], dw[3][12];
    atom_weights(o, pos + 3 * i, m, w, dw);
    for (int a = 0; a < nu; a++)
      for (int b = 0; b < nu; b++)
        for (int c = 0; c < nu; c++) {
          long x = pmod(m[0] + nu / 2 - a, g[0]), y = pmod(m[1] + nu / 2 - b, g[1]), z = pmod(m[2] + nu / 2 - c, g[2]);
          o->qg[(x * g[1] + y) * g[2] + z] += q[i] * w[0][a] * w[1][b] * w[2][c];
        }
  }
  /* reciprocal convolution (FFTg2g, pme1.c:217-229) */
  double *re = (double *)malloc(n * sizeof(double)), *im = (double *)calloc(n, sizeof(double));
  memcpy(re, o->qg, n * sizeof(double));
  for (int a = 2; a >= 0; a--) dft_axis(re, im, g, a, -1);
  for (long e = 0; e < n; e++) re[e] *= o->khat[e], im[e] *= o->khat[e];
  for (int a = 0; a < 3; a++) dft_axis(re, im, g, a, 1);
  double Eg = 0;
  for (long e = 0; e < n; e++) {
    o->eg[e] = re[e] / (double)n;
    Eg += o->qg[e] * o->eg[e];
  }
  Eg *= 0.5;
  free(re);
  free(im);
  /* field at the atoms (interpolate, pme1.c:232-272): E = -Ai^T grad */
  for (int i = 0; i < N; i++) {
    int m[3];
    double w[3][12], dw[3][12], ef[3] = {0, 0, 0}, pv = 0;
    atom_weights(o, pos + 3 * i, m, w, dw);
    for (int a = 0; a < nu; a++)
      for (int b = 0; b < nu; b++)
        for (int c = 0; c < nu; c++) {
          long x = pmod(m[0] + nu / 2 - a, g[0]), y = pmod(m[1] + nu / 2 - b, g[1]), z = pmod(m[2] + nu / 2 - c, g[2]);
          double e = o->eg[(x * g[1] + y) * g[2] + z];
          ef[0] -= dw[0][a] * w[1][b] * w[2][c] * e;
          ef[1] -= w[0][a] * dw[1][b] * w[2][c] * e;
          ef[2] -= w[0][a] * w[1][b] * dw[2][c] * e;
          pv += w[0][a] * w[1][b] * w[2][c] * e;
        }
    for (int a = 0; a < 3; a++) ef[a] *= g[a];
    for (int a = 0; a < 3; a++)
      F[3 * i + a] += q[i] * (o->Ai[0][a] * ef[0] + o->Ai[1][a] * ef[1] + o->Ai[2][a] * ef[2]);
    if (phi) phi[i] += pv;
  }
  if (terms) terms[0] = Ec, terms[1] = Ep, terms[2] = Eg;
  return Ec + Ep + Eg;
}

void pmo_info(const pmo_t *o, int *info, double *dinfo) {
  info[0] = o->nu;
  for (int a = 0; a < 3; a++) info[1 + a] = o->M[a];
  info[4] = o->pair_count;
  dinfo[0] = o->a0;
  dinfo[1] = o->beta;
  dinfo[2] = o->coeff1;
  dinfo[3] = o->coeff2;
  dinfo[4] = o->detA;
  dinfo[5] = o->tolDir;
}

void pmo_copy_khat(const pmo_t *o, double *out) {
  memcpy(out, o->khat, (size_t)o->M[0] * o->M[1] * o->M[2] * sizeof(double));
}

void pmo_copy_grid(const pmo_t *o, int which, double *out) {
  const double *src = which ? o->eg : o->qg;
  size_t n = (size_t)o->M[0] * o->M[1] * o->M[2];
  if (src) memcpy(out, src, n * sizeof(double));
  else memset(out, 0, n * sizeof(double));
}

void pmo_free(pmo_t *o) {
  if (!o) return;
  for (int a = 0; a < 3; a++) free(o->cL[a]);
  free(o->khat);
  free(o->qg);
  free(o->eg);
  free(o);
}
