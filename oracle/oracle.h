/*
 * CPU FP64 oracle for the crysSCFT propagator path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference's
 * algorithm (crysSCFT_sample.cpp, solver_p1.h, solver_3m.h, solver_Pmmm.h,
 * solver_utilities.h) with self-written transforms that follow FFTW's published
 * definitions of REDFT10 / REDFT01 / r2c / c2r.  It is the checker that the
 * HIP library is compared with, and the timed CPU baseline of bench.py.  The
 * product library never links, loads or calls it.
 *
 * Pinning (what makes the oracle trustworthy without a buildable reference,
 * which needs FFTW3 that is absent here):
 *   - orc_dct2_3d / orc_dct3_3d / orc_rfft_3d / orc_irfft_3d: scipy/numpy
 *     golden fixtures (tests/golden) and brute-force definitions on tiny grids.
 *   - Pmmm / P1 / 3m diffusion: cross-solver identity on mirror-symmetric
 *     fields (the reference's own acceptance criterion, Readme.md:28-76),
 *     plane-wave eigenvector tests, scipy fixtures.
 *   - Driver: uniform-field closed form Q = exp(-ds(a NsA + b NsB)) with
 *     phiA == fA, phiB == fB; mean(phiA) = fA and mean(phiB) = fB for any
 *     field; d lnQ / dwA(r) = -(ds NsA / fA) phiA(r) / n by finite differences.
 *   - print_energy terms: closed forms on uniform fields (enthalpic xN fA fB,
 *     -ln Q, entropic -(a fA + b fB), ksi term 0, MaxIncompressibility 0),
 *     truncated NsA included; for any fields the identity
 *     F - H = <xN t^2/4> - <(v + xN d)^2/(4 xN)> with the density-free
 *     H = <(wA-wB)^2/(4 xN) - (wA+wB)/2 + xN/4> - ln Q, t = 1-phiA-phiB,
 *     d = phiA-phiB, v = wA-wB; F = H at the fixed point of the mixing.
 *   - calc_new_fields: closed form on uniform fields (pre-update ksi); for any
 *     fields wA+wB moves by -acc xN t and wA-wB by -acc (v + xN d).
 *     A copy of this file with a sign flipped in any print_energy term, the
 *     densities swapped in a term, or the wrong density in the mixing fails
 *     these checks.
 *   - 3m with non-zero translational parts: the phase-folded 3m equals P1
 *     applied to the field expanded by the 8 symmetry operations, then
 *     reduced to the even sub-lattice (6 translation sets, two of them with
 *     shifted mirror planes that fix the direction of the translations).
 *   - Memory safety: oracle/selftest.c under AddressSanitizer and
 *     UndefinedBehaviorSanitizer.
 */
#ifndef CRYS_ORACLE_H
#define CRYS_ORACLE_H
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_P1 = 0, ORC_M3 = 1, ORC_PMMM = 2 };

/* ---- transforms (FFTW conventions, unnormalised) ---- */
/* REDFT10 along every axis: Y_k = 2 sum_j X_j cos(pi (j+1/2) k / n). */
void orc_dct2_3d(const double* in, double* out, size_t nx, size_t ny, size_t nz);
/* REDFT01 along every axis: Y_k = X_0 + 2 sum_{j>=1} X_j cos(pi j (k+1/2) / n). */
void orc_dct3_3d(const double* in, double* out, size_t nx, size_t ny, size_t nz);
/* r2c, forward sign e^{-i}; out is complex interleaved nx*ny*(nz/2+1). */
void orc_rfft_3d(const double* in, double* out, size_t nx, size_t ny, size_t nz);
/* c2r, backward sign e^{+i}; in is complex interleaved nx*ny*(nz/2+1). */
void orc_irfft_3d(const double* in, double* out, size_t nx, size_t ny, size_t nz);

/* ---- solvers (CrysPSSolverBase restated, solver_base.h:8-51) ---- */
typedef struct orc_solver orc_solver;
/* Returns NULL and fills err (if non-NULL) on a grid the reference rejects. */
orc_solver* orc_solver_create(int kind, const size_t grid_log[3], const double g[9],
                              char* err, size_t errlen);
void orc_solver_destroy(orc_solver* s);
void orc_set_cell_para(orc_solver* s, const double cell[6]);
void orc_set_contour_step(orc_solver* s, double ds);
void orc_diffusion(orc_solver* s);
double* orc_get_iomat(orc_solver* s);
size_t orc_get_space_size_phy(const orc_solver* s);
void orc_get_grid_size_phy(const orc_solver* s, size_t out[3]);
void orc_get_cell_para(const orc_solver* s, double out[6]);
double orc_get_ds(const orc_solver* s);

/* solve_diffusion_equation (crysSCFT_sample.cpp:373-396). q has (Ns+1)*n. */
void orc_propagate(orc_solver* s, double* q, const double* init, const double* E,
                   size_t Ns, int direction);

/* diblock_scft_simple_iterate (crysSCFT_sample.cpp:426-531).
 * wA, wB: in = initial fields, out = fields after Nstep mixing updates.
 * phiA, phiB: out, densities of the last iteration.
 * trace: Nstep rows of ORC_TRACE_COLS doubles:
 *   F, enthalpic, -lnQ, <-wA phiA - wB phiB>, ksi term, MaxIncompressibility, Q
 * Returns 0, or -1 on bad arguments. */
#define ORC_TRACE_COLS 7
int orc_scft_simple_iterate(orc_solver* s, double fA, double ds, double xN,
                            double Lx, double Ly, double Lz, double acceptance,
                            size_t Nstep, double* wA, double* wB,
                            double* phiA, double* phiB, double* trace);

/* Number of OpenMP threads the transforms use (for the CPU baseline). */
void orc_set_num_threads(int n);
int orc_get_num_threads(void);

#ifdef __cplusplus
}
#endif
#endif
