/*
 * Memory and undefined-behaviour self-test of the oracle (test infrastructure).
 *
 * Built with -fsanitize=address,undefined by `make selftest` and run by
 * tests/test_oracle_sanitized.py.  It drives every entry point of oracle.h on
 * small grids: the four 3-D transforms, the three solvers (3m also with a
 * translational part), diffusion, propagation in both directions, two SCFT
 * iterations, and the grids the reference rejects.  Any out-of-bounds access,
 * leak or undefined operation makes the sanitizers abort; the program itself
 * also checks that the results are finite.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

static unsigned long long rng_state = 0x9E3779B97F4A7C15ull;

static double uniform(void) {
  rng_state ^= rng_state << 13;
  rng_state ^= rng_state >> 7;
  rng_state ^= rng_state << 17;
  return (double)(rng_state >> 11) / 9007199254740992.0 - 0.5;
}

static int all_finite(const double* p, size_t n) {
  for (size_t i = 0; i < n; ++i)
    if (!isfinite(p[i])) return 0;
  return 1;
}

static int fail(const char* what) {
  fprintf(stderr, "selftest: %s\n", what);
  return 1;
}

static int transforms(void) {
  const size_t nx = 4, ny = 6, nz = 8;
  const size_t n = nx * ny * nz, nc = nx * ny * (nz / 2 + 1) * 2;
  double* a = malloc(n * sizeof(double));
  double* b = malloc(n * sizeof(double));
  double* c = malloc(nc * sizeof(double));
  if (!a || !b || !c) return fail("allocation");
  for (size_t i = 0; i < n; ++i) a[i] = uniform();
  orc_dct2_3d(a, b, nx, ny, nz);
  orc_dct3_3d(b, a, nx, ny, nz);
  orc_rfft_3d(a, c, nx, ny, nz);
  orc_irfft_3d(c, b, nx, ny, nz);
  const int ok = all_finite(a, n) && all_finite(b, n) && all_finite(c, nc);
  free(a);
  free(b);
  free(c);
  return ok ? 0 : fail("non-finite transform output");
}

static int solver_case(int kind, const size_t grid[3], const double* g) {
  char err[256];
  orc_solver* s = orc_solver_create(kind, grid, g, err, sizeof(err));
  if (!s) return fail(err);
  const double cell[6] = {4.0, 4.1, 4.3, 0, 0, 0};
  orc_set_cell_para(s, cell);
  orc_set_contour_step(s, 0.05);
  const size_t n = orc_get_space_size_phy(s);
  double* io = orc_get_iomat(s);
  for (size_t i = 0; i < n; ++i) io[i] = 1.0 + 0.1 * uniform();
  orc_diffusion(s);
  int r = all_finite(io, n) ? 0 : fail("non-finite diffusion output");

  const size_t Ns = 5;
  double* q = malloc((Ns + 1) * n * sizeof(double));
  double* E = malloc(n * sizeof(double));
  double* init = malloc(n * sizeof(double));
  double *wA = malloc(n * sizeof(double)), *wB = malloc(n * sizeof(double));
  double *pA = malloc(n * sizeof(double)), *pB = malloc(n * sizeof(double));
  double trace[2 * ORC_TRACE_COLS];
  if (!q || !E || !init || !wA || !wB || !pA || !pB) return fail("allocation");
  for (size_t i = 0; i < n; ++i) {
    E[i] = exp(-0.05 * uniform());
    init[i] = 1.0;
    wA[i] = 2.0 * uniform();
    wB[i] = -wA[i];
  }
  orc_propagate(s, q, init, E, Ns, 1);
  orc_propagate(s, q, init, E, Ns, -1);
  if (!r && !all_finite(q, (Ns + 1) * n)) r = fail("non-finite propagator");
  if (!r && orc_scft_simple_iterate(s, 0.3, 0.05, 15.0, 4.0, 4.1, 4.3, 0.1, 2, wA, wB, pA, pB,
                                    trace) != 0)
    r = fail("scft returned an error");
  if (!r && !(all_finite(trace, 2 * ORC_TRACE_COLS) && all_finite(pA, n) && all_finite(wB, n)))
    r = fail("non-finite scft output");
  free(q);
  free(E);
  free(init);
  free(wA);
  free(wB);
  free(pA);
  free(pB);
  orc_solver_destroy(s);
  return r;
}

int main(void) {
  orc_set_num_threads(2);
  int r = transforms();
  const size_t p1[3] = {8, 12, 16}, red[3] = {8, 12, 16};
  const double g[9] = {0.5, 0.5, 0, 0, 0, 0, 0, 0, 0};
  r = r ? r : solver_case(ORC_P1, p1, NULL);
  r = r ? r : solver_case(ORC_PMMM, red, NULL);
  r = r ? r : solver_case(ORC_M3, red, NULL);
  r = r ? r : solver_case(ORC_M3, red, g);
  /* grids the reference rejects: NULL, no crash */
  const size_t bad_p1[3] = {8, 8, 12}, bad_red[3] = {7, 8, 16};
  char err[128];
  if (!r && orc_solver_create(ORC_P1, bad_p1, NULL, err, sizeof(err)) != NULL)
    r = fail("P1 accepted Nz % 8 != 0");
  if (!r && orc_solver_create(ORC_PMMM, bad_red, NULL, err, sizeof(err)) != NULL)
    r = fail("Pmmm accepted an odd grid");
  if (!r) printf("selftest ok\n");
  return r;
}
