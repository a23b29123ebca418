/*
 * msmpb_cli.c -- command-line driver with the argv and JSON output of the
 * reference's `msmpb` tool (hkayabilisim/msmpb code/msmpb.c:11-380), running
 * the evaluation on the GPU through libmsmpb_hip.so.
 *
 *   msmpb <data> [--a0 a] [--nu n] [-M m] [-L l] [--perturb]
 *         [--repl x y z] [--edge a11 .. a33] [--bincoor file]
 *         [--potfile f] [--accfile f]
 *         [--precision fp64|mixed] [--layout fixed|asis]   (extensions)
 *
 * Reads <data>.ini (and <data>.acc / <data>.pot when present), prints one
 * JSON object on stdout with the reference's keys (always valid JSON), and
 * writes msmpb.acc (-F/q, three blocks) and msmpb.pot in the working
 * directory.
 */
#define _POSIX_C_SOURCE 200809L
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "msmpb_hip.h"

struct O o = {"csCl"};

static void usage(void) {
  fprintf(stderr,
          "Usage: msmpb dataFile [--a0 cutoff] [--nu accuracyOrder] [-M grid-spacing]\n"
          "  [-L numberOfLevels] [--perturb] [--repl replx reply replz]\n"
          "  [--edge a11 a12 a13 a21 a22 a23 a31 a32 a33] [--bincoor file]\n"
          "  [--potfile .pot file] [--accfile .acc file]\n"
          "  [--precision fp64|mixed] [--layout fixed|asis]\n");
  exit(1);
}

static double wall(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

/* NAMD binary coordinates: int32 N then 3N doubles (msmpb.c:387-415). */
static int read_bincoor(const char *path, int n, double *xyz) {
  FILE *f = fopen(path, "rb");
  if (!f) {
    fprintf(stderr, "Unable to open file \"%s\" for reading\n", path);
    return -1;
  }
  int m = 0;
  int ok = fread(&m, sizeof m, 1, f) == 1 && m == n &&
           fread(xyz, sizeof(double), 3 * (size_t)n, f) == 3 * (size_t)n;
  if (!ok) fprintf(stderr, "Bad or mismatched binary coordinate file \"%s\"\n", path);
  fclose(f);
  return ok ? 0 : -1;
}

static void need(int i, int k, int argc) {
  if (i + k >= argc) usage();
}

int main(int argc, char **argv) {
  if (argc < 2) usage();
  const char *name = argv[1];
  char ini[1024], accf[1024], potf[1024], binf[1024] = "";
  snprintf(ini, sizeof ini, "%s.ini", name);
  snprintf(accf, sizeof accf, "%s.acc", name);
  snprintf(potf, sizeof potf, "%s.pot", name);
  double a0 = -1, edge_in[9];
  int nu = -1, L = -1, Mtop = -1, perturb = 0, edge_given = 0, rep[3] = {1, 1, 1};
  int precision = MSM_PREC_FP64, layout = MSM_LAYOUT_FIXED;
  for (int i = 2; i < argc; ++i) {
    if (!strcmp(argv[i], "--a0")) need(i, 1, argc), a0 = atof(argv[++i]);
    else if (!strcmp(argv[i], "--nu")) need(i, 1, argc), nu = atoi(argv[++i]);
    else if (!strcmp(argv[i], "-M")) need(i, 1, argc), Mtop = (int)atof(argv[++i]);
    else if (!strcmp(argv[i], "-L")) need(i, 1, argc), L = atoi(argv[++i]);
    else if (!strcmp(argv[i], "--perturb")) perturb = 1;
    else if (!strcmp(argv[i], "--repl")) {
      need(i, 3, argc);
      for (int k = 0; k < 3; ++k) rep[k] = atoi(argv[++i]);
    } else if (!strcmp(argv[i], "--edge")) {
      need(i, 9, argc);
      for (int k = 0; k < 9; ++k) edge_in[k] = atof(argv[++i]);
      edge_given = 1;
    } else if (!strcmp(argv[i], "--potfile")) need(i, 1, argc), snprintf(potf, sizeof potf, "%s", argv[++i]);
    else if (!strcmp(argv[i], "--accfile")) need(i, 1, argc), snprintf(accf, sizeof accf, "%s", argv[++i]);
    else if (!strcmp(argv[i], "--bincoor")) need(i, 1, argc), snprintf(binf, sizeof binf, "%s", argv[++i]);
    else if (!strcmp(argv[i], "--precision")) {
      need(i, 1, argc);
      ++i;
      precision = !strcmp(argv[i], "mixed") ? MSM_PREC_MIXED : MSM_PREC_FP64;
    } else if (!strcmp(argv[i], "--layout")) {
      need(i, 1, argc);
      ++i;
      layout = !strcmp(argv[i], "asis") ? MSM_LAYOUT_ASIS : MSM_LAYOUT_FIXED;
    }
  }

  /* .ini: box row, N, then q x y z mass per line (msmpb.c:93-118) */
  FILE *f = fopen(ini, "r");
  if (!f) {
    fprintf(stderr, "cannot open %s\n", ini);
    return 1;
  }
  double edge[3][3];
  char line[512];
  if (!fgets(line, sizeof line, f) ||
      sscanf(line, "%lf%lf%lf%lf%lf%lf%lf%lf%lf", &edge[0][0], &edge[0][1], &edge[0][2],
             &edge[1][0], &edge[1][1], &edge[1][2], &edge[2][0], &edge[2][1], &edge[2][2]) != 9) {
    fprintf(stderr, "bad box line in %s\n", ini);
    return 1;
  }
  if (edge_given) memcpy(edge, edge_in, sizeof edge);
  int N = 0;
  if (!fgets(line, sizeof line, f) || sscanf(line, "%d", &N) != 1 || N <= 0) {
    fprintf(stderr, "bad atom count in %s\n", ini);
    return 1;
  }
  int nrep = N * rep[0] * rep[1] * rep[2];
  double *q = calloc(nrep, sizeof *q), *mass = calloc(nrep, sizeof *mass);
  double (*r)[3] = calloc(nrep, sizeof *r), (*F)[3] = calloc(nrep, sizeof *F);
  double (*acc)[3] = calloc(N, sizeof *acc);
  for (int i = 0; i < N; ++i) {
    if (!fgets(line, sizeof line, f) ||
        sscanf(line, "%lf%lf%lf%lf%lf", &q[i], &r[i][0], &r[i][1], &r[i][2], &mass[i]) < 4) {
      fprintf(stderr, "bad atom line %d in %s\n", i + 1, ini);
      return 1;
    }
  }
  fclose(f);
  if (binf[0] && read_bincoor(binf, N, &r[0][0]) != 0) return 1;

  /* replicate along the box rows (msmpb.c:130-154) */
  int m = N;
  for (int a = 0; a < rep[0]; ++a)
    for (int b = 0; b < rep[1]; ++b)
      for (int c = 0; c < rep[2]; ++c) {
        if (!a && !b && !c) continue;
        for (int n = 0; n < N; ++n, ++m) {
          for (int k = 0; k < 3; ++k)
            r[m][k] = r[n][k] + a * edge[0][k] + b * edge[1][k] + c * edge[2][k];
          q[m] = q[n];
          mass[m] = mass[n];
        }
      }
  for (int k = 0; k < 3; ++k) edge[0][k] *= rep[0], edge[1][k] *= rep[1], edge[2][k] *= rep[2];
  double pert[3] = {0, 0, 0};
  if (perturb) { /* random rigid shift of all atoms (msmpb.c:156-167) */
    srand((unsigned)time(NULL));
    rand();
    for (int k = 0; k < 3; ++k) pert[k] = (2.0 * rand() / (double)RAND_MAX - 1.0) * edge[k][k];
    for (int i = 0; i < nrep; ++i)
      for (int k = 0; k < 3; ++k) r[i][k] += pert[k];
  }

  FF *ff = FF_new(nrep, q);
  if (L > 0) FF_set_maxLevel(ff, L);
  if (Mtop > 0) {
    int M3[3] = {Mtop, Mtop, Mtop};
    FF_set_topGridDim(ff, M3);
  }
  if (nu > 0) FF_set_orderAcc(ff, nu);
  if (a0 > 0) FF_set_cutoff(ff, a0);
  FF_set_precision(ff, precision);
  FF_set_layout_compat(ff, layout);
  FF_set_stage_timing(ff, 1);
  double t0 = wall();
  FF_build(ff, nrep, r, edge, 0.0);
  double time_build = wall() - t0;
  if (FF_get_maxLevel(ff) <= 0 || FF_last_error()[0]) {
    fprintf(stderr, "msmpb: build failed: %s\n", FF_last_error());
    return 2;
  }
  t0 = wall();
  double energy = FF_energy(ff, F, r, NULL);
  double time_energy = wall() - t0;
  if (!isfinite(energy)) {
    fprintf(stderr, "msmpb: evaluation failed: %s\n", FF_last_error());
    return 3;
  }
  int Mo[3];
  FF_get_topGridDim(ff, Mo);
  int Lv = FF_get_maxLevel(ff);

  /* JSON with the reference's keys (msmpb.c:186-360) */
  double sum = ff->time_partcl2partcl + ff->time_nlist;
  printf("{\n");
  printf("\"time_partcl2partcl\" : %10.8f,\n", ff->time_partcl2partcl);
  printf("\"time_nlist\" : %10.8f,\n", ff->time_nlist);
  for (int l = 1; l <= Lv && l < 10; ++l) {
    printf("\"time_grid2grid_atlevel%d\" : %10.8f,\n", l, ff->time_grid2grid[l]);
    printf("\"time_padding_atlevel%d\" : %10.8f,\n", l, ff->time_padding[l]);
    printf("\"time_stencil_atlevel%d\" : %10.8f,\n", l, ff->time_stencil[l]);
    sum += ff->time_grid2grid[l] + ff->time_padding[l] + ff->time_stencil[l];
  }
  for (int l = 2; l <= Lv && l < 10; ++l) {
    printf("\"time_restrictionfrom%dto%d\" : %10.8f,\n", l - 1, l, ff->time_restriction[l]);
    sum += ff->time_restriction[l];
  }
  for (int l = Lv - 1; l >= 1; --l) {
    printf("\"time_prolongationfrom%dto%d\" : %10.8f,\n", l + 1, l, ff->time_prolongation[l]);
    sum += ff->time_prolongation[l];
  }
  printf("\"time_anterpolation\" : %10.8f,\n", ff->time_anterpolation);
  printf("\"time_interpolation\" : %10.8f,\n", ff->time_interpolation);
  sum += ff->time_anterpolation + ff->time_interpolation;
  printf("\"time_build\" : %10.8f,\n", time_build);
  printf("\"time_energy\" : %10.8f,\n", time_energy);
  printf("\"time_longrange\" : %10.8f,\n", time_energy - ff->time_partcl2partcl);
  printf("\"time_other\" : %10.8f,\n", time_build + time_energy - sum);
  printf("\"time_total\" : %10.8f,\n", time_build + time_energy);
  printf("\"data\" : \"%s\",\n", name);
  printf("\"NumberOfLevels\" : %d,\n", Lv);
  printf("\"Nrep\" : %d,\n", nrep);
  printf("\"N\" : %d,\n", N);
  printf("\"repl\" : [%d, %d, %d],\n", rep[0], rep[1], rep[2]);
  printf("\"Perturbationx\" : %.5f,\n", pert[0]);
  printf("\"Perturbationy\" : %.5f,\n", pert[1]);
  printf("\"Perturbationz\" : %.5f,\n", pert[2]);
  printf("\"nu\" : %d,\n", FF_get_orderAcc(ff));
  printf("\"ntc\" : %d,\n", ff->ntc);
  printf("\"nlist_len\" : %d,\n", ff->nlist_len);
  printf("\"nlist_interaction_count\" : %d,\n", ff->nlist_interaction_count);
  printf("\"tmax\" : %d,\n", ff->tmax);
  printf("\"cutoff\" : %f,\n", FF_get_cutoff(ff));
  printf("\"margin\" : %f,\n", 0.0);
  for (int k = 0; k < 3; ++k)
    printf("\"Edge row%d\" : [%5.2f, %5.2f, %5.2f],\n", k + 1, edge[k][0], edge[k][1], edge[k][2]);
  printf("\"TopLevelMx\" : %d,\n", Mo[0]);
  printf("\"TopLevelMy\" : %d,\n", Mo[1]);
  printf("\"TopLevelMz\" : %d,\n", Mo[2]);
  printf("\"precision\" : \"%s\",\n", precision == MSM_PREC_MIXED ? "mixed" : "fp64");
  printf("\"layout\" : \"%s\",\n", layout == MSM_LAYOUT_ASIS ? "asis" : "fixed");
  printf("\"utotal\" : %.16e,\n", energy);

  f = fopen(accf, "r");
  if (f) { /* accelerations -> forces, errors against them (msmpb.c:254-312) */
    int ok = fgets(line, sizeof line, f) != NULL;
    for (int k = 0; k < 3 && ok; ++k)
      for (int i = 0; i < N && ok; ++i) ok = fgets(line, sizeof line, f) && sscanf(line, "%lf", &acc[i][k]) == 1;
    fclose(f);
    if (ok) {
      for (int i = 0; i < N; ++i)
        for (int k = 0; k < 3; ++k) acc[i][k] *= -q[i];
      double Q2 = 0;
      for (int i = 0; i < N; ++i) Q2 += q[i] * q[i];
      double Fref = (Q2 / N) / pow(ff->detA / N, 2. / 3.), Uref = (Q2 / N) / pow(ff->detA / N, 1. / 3.);
      double err2 = 0;
      for (int i = 0; i < nrep; ++i) {
        int n = i % N;
        for (int k = 0; k < 3; ++k) err2 += (acc[n][k] - F[i][k]) * (acc[n][k] - F[i][k]);
      }
      double dF = sqrt(err2 / N);
      printf("\"deltaF\" : %25.16e,\n", dF);
      printf("\"deltaF/Fref\" : %25.16e,\n", dF / Fref);
      printf("\"deltaFest\" : %25.16e,\n", FF_get_deltaF(ff));
      printf("\"deltaFest/Fref\" : %25.16e,\n", FF_get_errEst(ff));
      printf("\"fref\" : %25.16e,\n", Fref);
      printf("\"uref\" : %25.16e,\n", Uref);
      printf("\"Q2\" : %25.16e,\n", Q2);
    }
  }
  /* intramolecular Coulomb of consecutive 3-site molecules (msmpb.c:316-339) */
  double uintra = 0;
  for (int i = 0; i + 2 < N; i += 3) {
    double d01 = 0, d02 = 0, d12 = 0;
    for (int k = 0; k < 3; ++k) {
      d01 += (r[i][k] - r[i + 1][k]) * (r[i][k] - r[i + 1][k]);
      d02 += (r[i][k] - r[i + 2][k]) * (r[i][k] - r[i + 2][k]);
      d12 += (r[i + 1][k] - r[i + 2][k]) * (r[i + 1][k] - r[i + 2][k]);
    }
    uintra += q[i] * q[i + 1] / sqrt(d01) + q[i] * q[i + 2] / sqrt(d02) + q[i + 1] * q[i + 2] / sqrt(d12);
  }
  printf("\"uintra\" : %25.16e,\n", uintra);
  f = fopen(potf, "r");
  double pref = 0;
  int havepot = f && fgets(line, sizeof line, f) && sscanf(line, "%lf", &pref) == 1;
  if (f) fclose(f);
  if (havepot) {
    pref *= rep[0] * rep[1] * rep[2];
    printf("\"utotal_kcal_mol\" : %25.16e,\n", 332.0636 * (energy - uintra));
    printf("\"poterror\" : %25.16e,\n", fabs(pref - energy) / fabs(pref));
    printf("\"potabserror\" : %25.16f\n", fabs(energy - pref));
  } else {
    printf("\"utotal_kcal_mol\" : %25.16e\n", 332.0636 * (energy - uintra));
  }
  printf("}");

  FILE *fa = fopen("msmpb.acc", "w");
  if (fa) {
    fprintf(fa, "%d\n", N);
    for (int k = 0; k < 3; ++k)
      for (int i = 0; i < nrep; ++i) fprintf(fa, "%-25.16f\n", -F[i][k] / q[i]);
    fclose(fa);
  }
  FILE *fp = fopen("msmpb.pot", "w");
  if (fp) {
    fprintf(fp, "%25.16e\n", energy);
    fclose(fp);
  }
  FF_delete(ff);
  free(q), free(mass), free(r), free(F), free(acc);
  return 0;
}
