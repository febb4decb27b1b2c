/*
 * crys_solver.h -- C ABI of the MI355X (gfx950) propagator library
 * libcrys_solver.so.
 *
 * The library is a drop-in for the reference's solver plugin API
 * (CrysPSSolverBase, solver_base.h:8-51, and its three implementations
 * CrysPSSolverP1 / CrysPSSolver3m / CrysPSSolverPmmm) and for the SCFT
 * iteration that calls it (diblock_scft_simple_iterate,
 * crysSCFT_sample.cpp:426-531).  All buffers the handle owns live in device
 * memory (HBM); host staging is explicit.  Every call returns a status code
 * and never exits the process; crys_last_error() describes the last failure
 * on the calling thread.
 *
 * Threading: calls on one handle are serialised by the caller; different
 * handles may be used concurrently from different threads (each handle owns
 * its own HIP stream).
 */
#ifndef CRYS_SOLVER_H
#define CRYS_SOLVER_H

#include <stddef.h>

#if defined(__GNUC__)
#define CRYS_API __attribute__((visibility("default")))
#else
#define CRYS_API
#endif

#ifdef __cplusplus
extern "C" {
#endif

typedef struct crys_solver crys_solver_t;
typedef struct crys_scft crys_scft_t;

/* Solver kinds: CrysPSSolverP1 (solver_p1.h:14), CrysPSSolver3m
 * (solver_3m.h:14), CrysPSSolverPmmm "SolverPmmmDCT" (solver_Pmmm.h:13). */
enum crys_kind { CRYS_P1 = 0, CRYS_3M = 1, CRYS_PMMM_DCT = 2 };

/* Arithmetic of the propagators.  CRYS_FP64 is the plain method. */
enum crys_precision { CRYS_FP64 = 0, CRYS_FP32 = 1 };

/* Creation flags. */
enum crys_flags {
  CRYS_FLAG_NONE = 0,
  /* PmmmDCT only: honour Lx, Ly, Lz.  Without it the solver reproduces the
   * reference, which keeps only cell_para[0] (solver_Pmmm.h:101). */
  CRYS_FLAG_PMMM_ORTHORHOMBIC = 1
};

enum crys_status {
  CRYS_OK = 0,
  CRYS_ERR_INVALID_ARGUMENT = 1, /* NULL pointer, bad direction, ds <= 0, ... */
  CRYS_ERR_BAD_GRID = 2,         /* the grid the reference's set_grid() rejects */
  CRYS_ERR_NO_CONTOUR_STEP = 3,  /* diffusion before set_contour_step */
  CRYS_ERR_UNSUPPORTED = 4,      /* a combination the library does not run */
  CRYS_ERR_HIP = 5,              /* a HIP runtime call failed */
  CRYS_ERR_OUT_OF_MEMORY = 6,
  CRYS_ERR_DIVERGED = 7,         /* non-finite free energy or Q */
  CRYS_ERR_IO = 8                /* a field file could not be read or written */
};

/* Description of the last error on this thread ("" if none). */
CRYS_API const char* crys_last_error(void);
/* Library version string. */
CRYS_API const char* crys_version(void);
/* Number of visible HIP devices (0 when none); never fails. */
CRYS_API int crys_device_count(void);
/* hipDeviceSynchronize on the current device. */
CRYS_API int crys_device_synchronize(void);

/* ---- solver plugin API ------------------------------------------------ */

/* Constructors CrysPSSolverP1(grid) (solver_p1.h:16),
 * CrysPSSolver3m(grid, translational_part) (solver_3m.h:17) and
 * CrysPSSolverPmmm(grid) (solver_Pmmm.h:16), with the grid checks of their
 * set_grid() (solver_p1.h:95-125, solver_3m.h:73-113, solver_Pmmm.h:66-96)
 * returning CRYS_ERR_BAD_GRID instead of exit(1).
 * grid_log: logical grid {Nx, Ny, Nz}.  translational_part: 9 doubles for
 * CRYS_3M (the translations of the -z, -y, -x mirrors in cell fractions, as
 * generate_m3_operation, solver_utilities.h:55-66), may be NULL (all zero);
 * ignored for other kinds.  With non-zero parts the solver works on the
 * 2x2y2z sub-lattice (1/8 of the grid) like the reference's folded
 * coefficients (solver_3m.h:172-255): its diffusion is a short sum of
 * separable terms, each the P1 operator of every axis restricted to the
 * sub-lattice and the images of one symmetry operation; a translation that is
 * not a whole number of grid points, or operations that do not map the
 * sub-lattice one to one onto the grid, give CRYS_ERR_UNSUPPORTED.
 * precision: enum crys_precision.
 * flags: enum crys_flags. */
CRYS_API int crys_solver_create(int kind, const size_t grid_log[3], const double translational_part[9],
                       int precision, int flags, crys_solver_t** out);
/* ~CrysPSSolver* (solver_p1.h:55-62, solver_3m.h:55-71, solver_Pmmm.h:54-64). */
CRYS_API int crys_solver_destroy(crys_solver_t* h);

/* set_cell_para (solver_base.h:16): {Lx, Ly, Lz, alpha, beta, gamma}; the
 * angles are reserved, as in the reference.  Early-return and cubic-only
 * rules of solver_3m.h:115-121 / solver_Pmmm.h:98-104 are kept; P1 always
 * resets (solver_p1.h:64-69).  A change of cell drops the cached operators
 * and resets ds to 0. */
CRYS_API int crys_set_cell_para(crys_solver_t* h, const double cell_para[6]);
/* set_contour_step (solver_base.h:18): operators are cached per ds value,
 * like the reference's std::map (solver_p1.h:71-81, solver_Pmmm.h:106-123). */
CRYS_API int crys_set_contour_step(crys_solver_t* h, double ds);
/* diffusion() (solver_base.h:20): iomat <- T^-1 K T iomat, normalisation
 * included, on the device-resident iomat; asynchronous on the handle's
 * stream.  E = exp(-w ds/2) is applied by the caller, as in the reference. */
CRYS_API int crys_diffusion(crys_solver_t* h);
/* Times `reps` back-to-back diffusion() calls on the handle's stream with HIP
 * events and writes the mean milliseconds per call: the device counterpart of
 * the reference's triplet clock (std::clock around diffusion(),
 * crysSCFT_sample.cpp:389-391, summed at :512).  Changes the iomat. */
CRYS_API int crys_diffusion_time(crys_solver_t* h, int reps, double* ms_per_call);

/* Getters (solver_base.h:22-28). */
CRYS_API int crys_get_cell_para(const crys_solver_t* h, double out[6]);
CRYS_API int crys_get_grid_size_log(const crys_solver_t* h, size_t out[3]);
CRYS_API int crys_get_grid_size_phy(const crys_solver_t* h, size_t out[3]);
CRYS_API int crys_get_space_size_log(const crys_solver_t* h, size_t* out);
CRYS_API int crys_get_space_size_phy(const crys_solver_t* h, size_t* out);
CRYS_API int crys_get_ds(const crys_solver_t* h, double* out);
/* get_iomat_for_diffusion (solver_base.h:28): a DEVICE pointer to
 * space_size_phy doubles, owned by the handle. */
CRYS_API int crys_get_iomat_for_diffusion(const crys_solver_t* h, double** out_device_ptr);
CRYS_API int crys_get_kind(const crys_solver_t* h, int* out);
CRYS_API int crys_get_precision(const crys_solver_t* h, int* out);

/* Explicit host staging of iomat (space_size_phy doubles); synchronous. */
CRYS_API int crys_iomat_upload(crys_solver_t* h, const double* host);
CRYS_API int crys_iomat_download(crys_solver_t* h, double* host);
/* Wait for the handle's stream. */
CRYS_API int crys_solver_synchronize(crys_solver_t* h);
/* The handle's HIP stream (hipStream_t), for callers that order work. */
CRYS_API int crys_solver_get_stream(const crys_solver_t* h, void** out_stream);

/* solve_diffusion_equation (crysSCFT_sample.cpp:373-396) on device buffers:
 * q_hist[s0] = init; then Ns times: iomat *= E; diffusion(); iomat *= E;
 * q_hist[s0 + direction] = iomat.  s0 = 0 for direction +1 and Ns for -1.
 * E_dev, init_dev: space_size_phy doubles; q_hist_dev: (Ns+1) of them.
 * On return iomat holds the last state.  Asynchronous on the handle's stream. */
CRYS_API int crys_propagate(crys_solver_t* h, const double* E_dev, const double* init_dev, size_t Ns,
                   int direction, double* q_hist_dev);

/* crys_propagate without any allocation.  No counterpart in the reference: it
 * is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The same
 * kernels and results as crys_propagate, but the half-finished contour step is
 * staged in the library's static device field of 2^21 doubles (16 MiB per
 * GPU, held until the process ends) instead of the solver's own scratch field;
 * solvers with more physical points (none that crys_solver_create accepts
 * today) get CRYS_ERR_UNSUPPORTED.  The call makes no hipMalloc or hipFree.
 * All calls on one GPU share that field: calls that overlap on one GPU (two
 * streams, two threads) race and must be serialised by the caller. */
CRYS_API int crys_propagate_noalloc(crys_solver_t* h, const double* E_dev, const double* init_dev,
                                    size_t Ns, int direction, double* q_hist_dev);

/* ---- SCFT iteration (diblock_scft_simple_iterate) ---------------------- */

/* Columns of one trace row, one row per iteration (print_energy,
 * crysSCFT_sample.cpp:398-424, recorded every iteration at full precision). */
enum crys_trace_col {
  CRYS_TRACE_F = 0,          /* FreeEnergy */
  CRYS_TRACE_ENTHALPIC = 1,  /* <xN phiA phiB> */
  CRYS_TRACE_MINUS_LNQ = 2,  /* -ln Q */
  CRYS_TRACE_ENTROPIC = 3,   /* <-wA phiA - wB phiB> */
  CRYS_TRACE_KSI = 4,        /* <-ksi (1 - phiA - phiB)> */
  CRYS_TRACE_MAXINCOMP = 5,  /* max |1 - phiA - phiB| */
  CRYS_TRACE_Q = 6,          /* single-chain partition function */
  CRYS_TRACE_COLS = 7
};

/* Sets up an AB-diblock run on `solver` (not owned): NsA = (size_t)(fA/ds),
 * NsB = (size_t)((1-fA)/ds) (crysSCFT_sample.cpp:443-444), calls
 * set_cell_para({Lx,Ly,Lz}) and set_contour_step(ds) (:466-467) and the
 * warm-up diffusion (:471), and allocates the device fields.  The contour
 * storage (the (NsA+NsB+1)-slot forward and backward propagator histories
 * of :447-460, or checkpoints, see crys_scft_set_checkpointing) is allocated
 * on first use. */
CRYS_API int crys_scft_create(crys_solver_t* solver, double fA, double ds, double xN, double Lx, double Ly,
                     double Lz, double acceptance, crys_scft_t** out);
CRYS_API int crys_scft_destroy(crys_scft_t* h);
/* Contour checkpointing.  No counterpart in the reference, whose driver keeps
 * every q(s) of the four chains (crysSCFT_sample.cpp:447-460).  every > 0
 * keeps the forward chain only at the slots 0, every, 2 every, ... and
 * recomputes the segments between them, so that an iteration (and
 * crys_scft_evaluate, stress included) holds ceil(Ns/every) + 2 (every + 1)
 * fields instead of 2 (Ns + 1), for one more propagation of the forward
 * chain.  every == 0 picks the interval: full histories when they fit in
 * budget_bytes (0: 90% of the free device memory when the storage is
 * allocated), else every = ceil(sqrt(Ns)).  Fails with
 * CRYS_ERR_OUT_OF_MEMORY when a nonzero budget cannot hold the storage.
 * Results equal the full-history ones to rounding. */
CRYS_API int crys_scft_set_checkpointing(crys_scft_t* h, size_t every, size_t budget_bytes);
/* The interval in use (0: full histories) and the contour storage in bytes. */
CRYS_API int crys_scft_get_checkpointing(crys_scft_t* h, size_t* every, size_t* storage_bytes);
/* Initial fields in the solver's physical layout (init_func, :463). */
CRYS_API int crys_scft_set_fields(crys_scft_t* h, const double* wA_host, const double* wB_host);
/* Runs nstep iterations of :476-508 on the device.  trace_host (may be NULL)
 * receives nstep rows of CRYS_TRACE_COLS doubles.  Synchronous. */
CRYS_API int crys_scft_run(crys_scft_t* h, size_t nstep, double* trace_host);
/* Fields after the last iteration (any pointer may be NULL).  phi are the
 * last iteration's densities, w the fields after its mixing update. */
CRYS_API int crys_scft_get_fields(crys_scft_t* h, double* phiA, double* phiB, double* wA, double* wB);
CRYS_API int crys_scft_get_contour(const crys_scft_t* h, size_t* NsA, size_t* NsB);

/* One-shot mirror of diblock_scft_simple_iterate (crysSCFT_sample.cpp:426-531)
 * with the fields passed explicitly instead of through init_func.  Output
 * pointers may be NULL. */
CRYS_API int crys_diblock_scft_simple_iterate(crys_solver_t* solver, double fA, double ds, double xN,
                                     double Lx, double Ly, double Lz, double acceptance,
                                     size_t nstep, const double* wA_init, const double* wB_init,
                                     double* trace, double* phiA, double* phiB, double* wA_out,
                                     double* wB_out);

/* ---- many independent SCFT instances ----------------------------------- */

/* One diblock_scft_simple_iterate run (crysSCFT_sample.cpp:426-531): solver
 * kind (CRYS_3M with zero translational parts), flags (enum crys_flags),
 * logical grid with every physical axis of at most 64 points, cell lengths,
 * block fraction fA, xN, contour step ds (NsA = (size_t)(fA/ds),
 * NsB = (size_t)((1-fA)/ds), :443-444) and mixing acceptance. */
typedef struct crys_instance {
  int kind;
  int flags;
  size_t grid_log[3];
  double cell[3];
  double fA, xN, ds, acceptance;
} crys_instance_t;

typedef struct crys_batch crys_batch_t;

/* Allocates all instances in HBM and builds their operators.  Every launch
 * then advances one axis pass of one contour step for all instances whose
 * contour is not finished, both chains of each; one SCFT iteration of the
 * whole batch is one hipGraph.  No counterpart in the reference, which runs
 * one instance per call of diblock_scft_simple_iterate. */
CRYS_API int crys_batch_create(const crys_instance_t* instances, size_t count, int precision,
                               crys_batch_t** out);
CRYS_API int crys_batch_destroy(crys_batch_t* b);
/* Initial fields of instance i (its physical layout; init_func, :463). */
CRYS_API int crys_batch_set_fields(crys_batch_t* b, size_t i, const double* wA, const double* wB);
/* nstep iterations of every instance; trace (may be NULL) receives
 * count * nstep rows of CRYS_TRACE_COLS doubles, instance-major. */
CRYS_API int crys_batch_run(crys_batch_t* b, size_t nstep, double* trace);
CRYS_API int crys_batch_get_fields(crys_batch_t* b, size_t i, double* phiA, double* phiB,
                                   double* wA, double* wB);

/* How a batch advances its contour steps.  No counterpart in the reference,
 * which runs one instance at a time.  PHASES: one launch per step for every
 * instance (possible when each instance has its three axes in one length
 * class); PASSES: three per-axis launches per step; AUTO: the phases while the
 * batch holds at most 1792 (instance, chain, plane) items, else the passes. */
enum crys_batch_path {
  CRYS_BATCH_PATH_AUTO = 0,
  CRYS_BATCH_PATH_PHASES = 1,
  CRYS_BATCH_PATH_PASSES = 2
};
/* Selects the path (CRYS_ERR_UNSUPPORTED for PHASES on a batch that cannot
 * run them).  Results agree to rounding on both paths. */
CRYS_API int crys_batch_set_path(crys_batch_t* b, int path);
/* The path in use: CRYS_BATCH_PATH_PHASES or CRYS_BATCH_PATH_PASSES. */
CRYS_API int crys_batch_get_path(const crys_batch_t* b, int* path);
/* Times the contour steps of one iteration on each possible path (reps
 * repetitions after a warm-up, HIP events on the batch's stream), selects the
 * faster and returns it in *path; ms (may be NULL) receives the mean
 * milliseconds of the phases (-1 when not possible) and of the passes.  The
 * timed steps write only the histories and scratch fields, so the batch's
 * fields and trace are unchanged. */
CRYS_API int crys_batch_autotune(crys_batch_t* b, int reps, int* path, double ms[2]);
/* The launch shape of the phases: 0 runs one workgroup per (instance, chain,
 * plane) item; G > 0 runs a persistent grid of min(G, items) workgroups, each
 * taking every G-th item and copying the next item's plane into LDS while it
 * computes the current one (the default, with G = the device's compute
 * units).  crys_batch_autotune times 0 and 1, 2 and 4 workgroups per CU and
 * keeps the fastest; ms[0] is then the fastest phases' time.  No counterpart
 * in the reference.  Results are identical for every G. */
CRYS_API int crys_batch_set_phase_grid(crys_batch_t* b, int workgroups);
CRYS_API int crys_batch_get_phase_grid(const crys_batch_t* b, int* workgroups);
CRYS_API int crys_batch_get_info(const crys_batch_t* b, size_t i, size_t* n_phy, size_t* NsA,
                                 size_t* NsB);

/* ---- field files and grid maps (host only) ------------------------------ */

/* output_fields (crysSCFT_sample.cpp:271-282): one line "phiA phiB wA wB" per
 * point.  digits <= 0 writes the reference's default 6 significant digits;
 * 17 makes the file an exact restart point. */
CRYS_API int crys_output_fields(const double* phiA, const double* phiB, const double* wA,
                                const double* wB, size_t size, const char* filename, int digits);
/* input_fields (crysSCFT_sample.cpp:284-292).  A missing or short file is
 * CRYS_ERR_IO (the reference leaves the arrays unchanged silently). */
CRYS_API int crys_input_fields(double* phiA, double* phiB, double* wA, double* wB, size_t size,
                               const char* filename);
/* The reduced field a solver works on, from the full logical grid
 * (crysSCFT_sample.cpp:202-258): P1 copies, 3m takes f[2x][2y][2z], PmmmDCT
 * takes the first octant. */
CRYS_API int crys_extract_fields(int kind, const size_t grid_log[3], const double* full,
                                 double* reduced);
/* Inverse of crys_extract_fields for fields with the mirrors f(N-1-j) = f(j):
 * Pmmm reflects the octant, 3m fills odd sites from f[2j+1] = e[M-1-j]. */
CRYS_API int crys_expand_fields(int kind, const size_t grid_log[3], const double* reduced,
                                double* full);

/* ---- free energy, gradient and cell stress ------------------------------ */

/* Evaluates, for the fields (wA, wB) (host arrays of space_size_phy doubles;
 * both NULL = the handle's current fields) on the solver's current cell:
 *   H = (1/n) sum_r [ (wA-wB)^2 / (4 xN) - (wA+wB)/2 + xN/4 ] - ln Q[wA, wB],
 * the field-theoretic free energy of the incompressible melt (it equals the
 * FreeEnergy of print_energy, crysSCFT_sample.cpp:398-424, at a saddle point,
 * where ksi = (wA+wB-xN)/2 as calc_ksi :341-346 sets it), and
 *   dH_dwA[r] = ((wA-wB)/(2 xN) - 1/2 + (ds NsA / fA) phiA[r]) / n, B alike,
 * from the forward and backward (adjoint) propagators of those fields, and
 *   dH_dL[a] = -(1/Q) dQ/dL_a = -(1/(n Q)) sum_s < E q^+_{s+1}, dD/dL_a E q_s >,
 * the cell stress used for box relaxation, from the adjoint inner products
 * with the L derivative of the per-axis operator.  For a PmmmDCT solver
 * without CRYS_FLAG_PMMM_ORTHORHOMBIC every axis uses cell_para[0]
 * (solver_Pmmm.h:101) and the derivative with respect to that single length
 * is dH_dL[0] + dH_dL[1] + dH_dL[2].  trace_row (7 doubles, enum
 * crys_trace_col) receives the print_energy terms of the same fields.  The
 * fields are not mixed.  Any output may be NULL; the stress is computed only
 * when dH_dL is not NULL.  Synchronous. */
CRYS_API int crys_scft_evaluate(crys_scft_t* h, const double* wA, const double* wB, double* H,
                                double* dH_dwA, double* dH_dwB, double dH_dL[3],
                                double* trace_row);
/* crys_scft_evaluate with the fields wA, wB and the gradients dH_dwA, dH_dwB
 * in DEVICE memory (space_size_phy doubles each, the solver's physical
 * layout), for callers whose tensors already live on the GPU (the torch
 * binding); H, dH_dL and trace_row stay host values.  The inputs must be
 * ready when the call is made (work that writes them on another stream is
 * finished); the call returns after the outputs are written.  Same results
 * as crys_scft_evaluate. */
CRYS_API int crys_scft_evaluate_device(crys_scft_t* h, const double* wA_dev,
                                       const double* wB_dev, double* H, double* dH_dwA_dev,
                                       double* dH_dwB_dev, double dH_dL[3], double* trace_row);

/* ---- introspection (host only, no device needed) ------------------------ */

/* The per-axis operator the device applies for one axis of a solver: D of
 * length m (m = n_logical for CRYS_P1, n_logical/2 otherwise), written as
 * out[i*m + j] = D[j][i] (m*m doubles).  The diffusion() of the solver is
 * Dx (x) Dy (x) Dz; this exposes the factor that replaces the reference's
 * transform/K/inverse-transform triplet (solver_p1.h:184-202,
 * solver_3m.h:281-324, solver_Pmmm.h:194-215) for checks on the host. */
CRYS_API int crys_axis_operator(int kind, size_t n_logical, double L, double ds, double* out);

/* ---- measurement -------------------------------------------------------- */

/* Times one contour step of both chains of an SCFT handle on its stream with
 * HIP events: `reps` whole steps back to back, then `reps` launches of each
 * kernel of the step alone.  Writes the mean milliseconds per step, the mean
 * milliseconds per launch of each kernel in issue order (ms_per_launch[k],
 * k < *nkernels; at most 3 entries) and the number of kernels per step (2 when
 * the z and y passes are fused on x-slabs, else 3).  Any output may be null.
 * No counterpart in the reference, which times diffusion() with std::clock
 * (crysSCFT_sample.cpp:389-391). */
/* Times the propagation of both chains as one SCFT iteration runs it: the
 * Ns-step (or, with checkpointing, one segment's) launch sequence on the
 * handle's own buffers, captured in a hipGraph and launched `reps` times
 * between HIP events after a warm-up.  Writes the mean milliseconds per
 * contour step (graph time / (reps * steps)) and the number of launches per
 * step (rounded).  Any output may be null.  No counterpart in the reference. */
CRYS_API int crys_scft_time_propagation(crys_scft_t* h, int reps, double* ms_per_step,
                                        int* launches_per_step);
CRYS_API int crys_scft_time_step(crys_scft_t* h, int reps, double* ms_per_step,
                                 double* ms_per_launch, int* nkernels);

/* Synchronisation between contour steps, measured.  No counterpart in the
 * reference.  Every phase of the one-launch contour step ends with all
 * workgroups' planes written and starts by reading planes written on the
 * other XCDs; crys_time_grid_sync times that exchange (each of `workgroups`
 * workgroups of 1024 threads, one per CU, writes 64 KB, then reads the 64 KB
 * that the workgroup 37 ids away wrote) for `rounds` rounds, synchronised by
 * one of: */
enum crys_sync_variant {
  CRYS_SYNC_LAUNCH = 0,    /* a kernel launch per round (what the iteration uses) */
  CRYS_SYNC_EMPTY = 1,     /* empty launches of the same shape, no exchange */
  CRYS_SYNC_FENCED = 2,    /* persistent kernel, one counter, agent-scope fences */
  CRYS_SYNC_TWO_LEVEL = 3, /* persistent, per-XCD counters then a root, agent-scope fences */
  CRYS_SYNC_BYPASS = 4,    /* persistent, one counter, exchange with agent-scope accesses */
  CRYS_SYNC_BARRIER = 5,   /* persistent, FENCED without the exchange */
  CRYS_SYNC_XCD = 6        /* persistent, groups of the workgroups on one XCD (read from the
                              hardware): one counter per XCD, no L2 fence; the partner is
                              the workgroup 5 places on in the same XCD, read through the
                              shared L2 with loads that miss the CU's L1 */
};
/* Writes the mean microseconds per round, the number of exchanged values a
 * reader found not yet written (must be 0) and the number of bounded spins
 * that timed out (must be 0; a time-out ends every workgroup; with
 * CRYS_SYNC_XCD a workgroup beyond workgroups / 8 on one XCD counts as a
 * time-out too).  8..256 workgroups, a multiple of 8.  Allocates its own
 * buffers. */
CRYS_API int crys_time_grid_sync(int variant, int workgroups, int rounds, double* us_per_round,
                                 unsigned* not_visible, unsigned* timeouts);

#ifdef __cplusplus
}
#endif
#endif /* CRYS_SOLVER_H */
