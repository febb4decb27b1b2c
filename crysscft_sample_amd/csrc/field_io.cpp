// Field files and the full <-> reduced grid maps (host only).
//
// output_fields / input_fields of crysSCFT_sample.cpp:271-292: one text line
// "phiA phiB wA wB" per grid point, row-major with z fastest.  The
// extraction of the reduced fields each solver works on follows the init
// lambdas of main() (crysSCFT_sample.cpp:202-258); the expansion is its
// inverse for fields with the mirrors f(N-1-j) = f(j) the 3m and Pmmm solvers
// assume.
#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "crys_solver.h"

namespace crys {
int set_error(int code, const char* msg);
}

namespace {

int check_kind_grid(int kind, const size_t g[3]) {
  if (!g) return crys::set_error(CRYS_ERR_INVALID_ARGUMENT, "null grid");
  if (kind != CRYS_P1 && kind != CRYS_3M && kind != CRYS_PMMM_DCT)
    return crys::set_error(CRYS_ERR_INVALID_ARGUMENT, "unknown solver kind");
  for (int a = 0; a < 3; ++a) {
    if (g[a] == 0) return crys::set_error(CRYS_ERR_BAD_GRID, "zero grid length");
    if (kind != CRYS_P1 && g[a] % 2)
      return crys::set_error(CRYS_ERR_BAD_GRID, "3m and Pmmm need even grid lengths");
  }
  return CRYS_OK;
}

// Source index on the full grid of reduced coordinate j on an axis of full
// length N, for the solver kind.
inline size_t src_index(int kind, size_t j) { return kind == CRYS_3M ? 2 * j : j; }

}  // namespace

extern "C" {

int crys_output_fields(const double* phiA, const double* phiB, const double* wA, const double* wB,
                       size_t size, const char* filename, int digits) {
  if (!phiA || !phiB || !wA || !wB || !filename)
    return crys::set_error(CRYS_ERR_INVALID_ARGUMENT, "null argument");
  FILE* f = fopen(filename, "w");
  if (!f) return crys::set_error(CRYS_ERR_IO, strerror(errno));
  // digits <= 0: the default std::ostream precision (6) of the reference
  const int p = digits > 0 ? digits : 6;
  for (size_t i = 0; i < size; ++i)
    if (fprintf(f, "%.*g %.*g %.*g %.*g\n", p, phiA[i], p, phiB[i], p, wA[i], p, wB[i]) < 0) {
      fclose(f);
      return crys::set_error(CRYS_ERR_IO, "write failed");
    }
  if (fclose(f) != 0) return crys::set_error(CRYS_ERR_IO, "close failed");
  return CRYS_OK;
}

int crys_input_fields(double* phiA, double* phiB, double* wA, double* wB, size_t size,
                      const char* filename) {
  if (!phiA || !phiB || !wA || !wB || !filename)
    return crys::set_error(CRYS_ERR_INVALID_ARGUMENT, "null argument");
  FILE* f = fopen(filename, "r");
  if (!f) return crys::set_error(CRYS_ERR_IO, strerror(errno));
  for (size_t i = 0; i < size; ++i) {
    if (fscanf(f, "%lf %lf %lf %lf", &phiA[i], &phiB[i], &wA[i], &wB[i]) != 4) {
      fclose(f);
      char msg[96];
      snprintf(msg, sizeof(msg), "file ends or is malformed at line %zu of %zu", i + 1, size);
      return crys::set_error(CRYS_ERR_IO, msg);
    }
  }
  fclose(f);
  return CRYS_OK;
}

int crys_extract_fields(int kind, const size_t g[3], const double* full, double* reduced) {
  int r = check_kind_grid(kind, g);
  if (r != CRYS_OK) return r;
  if (!full || !reduced) return crys::set_error(CRYS_ERR_INVALID_ARGUMENT, "null argument");
  if (kind == CRYS_P1) {
    memcpy(reduced, full, g[0] * g[1] * g[2] * sizeof(double));
    return CRYS_OK;
  }
  const size_t X = g[0] / 2, Y = g[1] / 2, Z = g[2] / 2;
  for (size_t x = 0; x < X; ++x)
    for (size_t y = 0; y < Y; ++y)
      for (size_t z = 0; z < Z; ++z)
        reduced[(x * Y + y) * Z + z] =
            full[(src_index(kind, x) * g[1] + src_index(kind, y)) * g[2] + src_index(kind, z)];
  return CRYS_OK;
}

int crys_expand_fields(int kind, const size_t g[3], const double* reduced, double* full) {
  int r = check_kind_grid(kind, g);
  if (r != CRYS_OK) return r;
  if (!full || !reduced) return crys::set_error(CRYS_ERR_INVALID_ARGUMENT, "null argument");
  if (kind == CRYS_P1) {
    memcpy(full, reduced, g[0] * g[1] * g[2] * sizeof(double));
    return CRYS_OK;
  }
  const size_t X = g[0] / 2, Y = g[1] / 2, Z = g[2] / 2;
  const size_t M[3] = {X, Y, Z};
  for (size_t i = 0; i < g[0]; ++i)
    for (size_t j = 0; j < g[1]; ++j)
      for (size_t k = 0; k < g[2]; ++k) {
        const size_t c[3] = {i, j, k};
        size_t red[3];
        for (int a = 0; a < 3; ++a) {
          const size_t Na = 2 * M[a];
          if (kind == CRYS_PMMM_DCT) {
            // first octant, mirror f(N-1-j) = f(j)
            red[a] = c[a] < M[a] ? c[a] : Na - 1 - c[a];
          } else {
            // even sites f[2j] = e[j]; odd sites f[2j+1] = e[M-1-j]
            red[a] = c[a] % 2 == 0 ? c[a] / 2 : M[a] - 1 - (c[a] - 1) / 2;
          }
        }
        full[(i * g[1] + j) * g[2] + k] = reduced[(red[0] * Y + red[1]) * Z + red[2]];
      }
  return CRYS_OK;
}

}  // extern "C"
