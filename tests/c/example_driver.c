/*
 * A library caller in the manner of the reference's code/example.c:20-39:
 * it uses forcefield.h only, sets the cutoff and the number of levels and
 * leaves the order of accuracy and the top-level grid to FF_build's rules,
 * then evaluates once.  Written for the test-suite (tests/test_pins.py
 * compiles and links it, tests/test_z_large_configs.py runs it on the GPU).
 *
 *   example_driver <input.ini> <cutoff> <maxLevel> <forces.out>
 *
 * input.ini: the reference's .ini layout (line 1: nine edge numbers, row
 * major; line 2: N; then N lines "q x y z ..." -- the same columns as
 * example.c's h2o.dat).  Prints the parameters FF_build chose and the
 * energy; writes the N x 3 forces as text.
 */
#include <stdio.h>
#include <stdlib.h>

#include "forcefield.h"

struct O o = {"csCl"}; /* callers define the optional-output hook, as example.c does */

int main(int argc, char **argv) {
  if (argc != 5) {
    fprintf(stderr, "usage: %s input.ini cutoff maxLevel forces.out\n", argv[0]);
    return 2;
  }
  FILE *in = fopen(argv[1], "r");
  if (!in) {
    perror(argv[1]);
    return 2;
  }
  double edges[3][3];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++)
      if (fscanf(in, "%lf", &edges[i][j]) != 1) return 3;
  int N = 0;
  if (fscanf(in, "%d", &N) != 1 || N <= 0) return 3;
  double *q = malloc(N * sizeof(double));
  double(*r)[3] = malloc(N * sizeof(double[3]));
  char line[256];
  if (!fgets(line, sizeof line, in)) return 3; /* rest of the N line */
  for (int i = 0; i < N; i++) {
    if (!fgets(line, sizeof line, in) ||
        sscanf(line, "%lf %lf %lf %lf", q + i, r[i], r[i] + 1, r[i] + 2) != 4)
      return 3;
  }
  fclose(in);

  FF *ff = FF_new(N, q);
  FF_set_cutoff(ff, atof(argv[2]));
  FF_set_maxLevel(ff, atoi(argv[3]));
  FF_build(ff, N, r, edges, 0.0);

  int dim[3];
  FF_get_topGridDim(ff, dim);
  printf("error estimate = %.17g\n", FF_get_estErr(ff));
  printf("number of levels = %d\n", FF_get_maxLevel(ff));
  printf("short-range cutoff = %.17g\n", FF_get_cutoff(ff));
  printf("order of accuracy = %d\n", FF_get_orderAcc(ff));
  printf("using FFT = %s\n", FF_get_FFT(ff) ? "true" : "false");
  printf("top level grid is %d x %d x %d\n", dim[0], dim[1], dim[2]);

  double(*F)[3] = malloc(N * sizeof(double[3]));
  double E = FF_energy(ff, F, r, NULL);
  printf("energy = %.17g\n", E);
  FILE *out = fopen(argv[4], "w");
  if (!out) {
    perror(argv[4]);
    return 2;
  }
  for (int i = 0; i < N; i++) fprintf(out, "%.17g %.17g %.17g\n", F[i][0], F[i][1], F[i][2]);
  fclose(out);
  FF_delete(ff);
  free(F);
  free(q);
  free(r);
  return E == E ? 0 : 1; /* NaN: the build or the evaluation failed */
}
