/*
 * hpc_ode_amd C ABI: batched stiff-chemistry BDF integration on one MI355X (gfx950).
 *
 * This is the drop-in boundary for the reference's per-cell CVODE path. The reference
 * exposes that path through the C++ Integrator interface
 * (include/ODEIntegrator/Integrators/Integrator.hpp:22-40: init / integrate(t0, t) / clean)
 * with CVodeIntegrator (src/Integrators/CVodeIntegrator.cpp:34-155) as its serial
 * implementation and CVodeIntegratorGPU (src/Integrators/CVodeIntegratorGPU.cpp:21-190)
 * as its GPU one. The functions below are what such an Integrator calls: the C++
 * class BDFIntegratorHIP (hpc_ode_amd/host/BDFIntegratorHIP.hpp) is the in-tree caller.
 *
 * Conventions: plain pointers and sizes; every function returns 0 on success or a
 * negative HPO_E* code, with a message available from hpo_last_error() (thread-local).
 * State rows are y = [T, Y_0 .. Y_{K-2}] (K species; the last species is recovered as
 * 1 - sum Y, as at src/Integrators/CVodeIntegrator.cpp:90-96). Calls on one handle must
 * not overlap in time (the caller serialises them); different handles are independent.
 */
#ifndef HPC_ODE_AMD_H
#define HPC_ODE_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HPO_OK 0
#define HPO_E_INVALID (-1)   /* bad argument */
#define HPO_E_PARSE (-2)     /* mechanism file could not be read or is unsupported */
#define HPO_E_DEVICE (-3)    /* HIP runtime error */
#define HPO_E_NOMEM (-4)     /* device memory exhausted */
#define HPO_E_UNSUPPORTED (-5)

/* per-cell status codes: CVODE's return codes (cvode.h) */
#define HPO_CELL_SUCCESS 0
#define HPO_CELL_TOO_MUCH_WORK (-1)
#define HPO_CELL_TOO_MUCH_ACC (-2)
#define HPO_CELL_ERR_FAILURE (-3)
#define HPO_CELL_CONV_FAILURE (-4)
#define HPO_CELL_ILL_INPUT (-22)
#define HPO_CELL_TOO_CLOSE (-27)

#define HPO_NSTATS 8 /* per cell: nst, nfe, nje, nsetups, nni, ncfn, netf, q_last */

typedef struct hpo_mech hpo_mech;
typedef struct hpo_batch hpo_batch;

typedef struct {
  double rtol;             /* scalar relative tolerance (CVodeSStolerances) */
  double atol;             /* scalar absolute tolerance */
  int32_t maxsteps;        /* CVodeSetMaxNumSteps; reference uses 10000 */
  int32_t lu_precision;    /* 0: FP64 Newton-matrix LU; 1: FP32 Newton matrix (Gauss-Jordan inverse
                              with the LU's partial pivoting) + FP64 residual refinement */
  int32_t refine_sweeps;   /* refinement sweeps per linear solve when lu_precision == 1 */
  int32_t waves_per_block; /* cells resident per workgroup (1..4); 0 = default */
  int32_t blocks_per_cu;   /* persistent workgroups per CU; 0 = as many as fit */
  int32_t sort_cells;      /* work-queue order: 0 cell index; 1 descending initial temperature;
                              2 (hpo_batch_integrate only) the order of the last hpo_batch_autotune */
  int32_t lu_layout;       /* with lu_precision == 1: 0 explicit inverse (Gauss-Jordan, solves as
                              matrix-vector products); 1 in-place LU (triangular solves) */
  int32_t cell_layout;     /* 0 automatic (one wave per cell when a wave kernel exists for the state
                              size, else one workgroup per cell); 1 one wave per cell; 2 one
                              256-thread workgroup per cell (Gauss-Jordan Newton matrix across the
                              workgroup; states of more than 64 variables, e.g. nDodecane_Reitz) */
  int32_t reactor;         /* 0 constant pressure (the reference's CONP, header.h:6): each cell's P is
                              its pressure in Pa; 1 constant volume (CONV, dydt.c:135-263): each
                              cell's P is its density in kg/m^3, held through the interval, and
                              the temperature equation uses internal energies and cv */
  int32_t fp64_fallback;   /* with lu_precision == 1: 1 (default) re-integrates in FP64, from its
                              input state, every cell whose FP32 run ends in a solver failure
                              (convergence or error-test failure, too much work or accuracy),
                              in a second launch on the same stream; 0 reports the FP32 failure.
                              An in-place call (y_in and y_out overlap) first copies the input
                              rows to device memory the library keeps for it */
} hpo_opts;

typedef struct {
  double kernel_ms;        /* integration kernel time, HIP events on the call's stream */
  double total_ms;         /* whole call including queue reset and any host copies */
  int64_t ncells;
  int32_t grid, block;     /* persistent launch shape */
  int32_t lds_bytes;       /* dynamic LDS per workgroup */
  int32_t fallback_cells;  /* cells re-integrated in FP64 after an FP32 failure (fp64_fallback) */
} hpo_timing;

/* Universal gas constant of the library's thermochemistry, J/(kmol K); e.g. for the density a
 * constant-volume cell starting at (T, Y, P) keeps: rho = P / (HPO_GAS_CONSTANT T sum_k Y_k/W_k). */
#define HPO_GAS_CONSTANT 8314.4621

/* Defaults: rtol 1e-6, atol 1e-10, maxsteps 10000 (src/Integrators/CVodeIntegrator.cpp:107,122). */
void hpo_opts_default(hpo_opts *o);

/* ---- mechanism: replaces the compiled-in pyJac code (src/Mechanism/CPU, src/Mechanism/GPU) ---- */
int hpo_mech_load_yaml(const char *path, hpo_mech **out);
/* Cantera YAML, Cantera CTML/XML for a path ending in .xml or .ctml, or Cantera CTI for a path
 * ending in .cti (the reference ships all three forms in resources/mechanisms; e.g.
 * h2_li_2004.xml, the mechanism of ref_data/res_h2.csv, ref_data/README.md:15). */
int hpo_mech_load(const char *path, hpo_mech **out);
void hpo_mech_destroy(hpo_mech *m);
/* n_state = n_species (T plus n_species - 1 mass fractions) */
int hpo_mech_info(const hpo_mech *m, int32_t *n_species, int32_t *n_state, int32_t *n_reactions);
int hpo_mech_species_name(const hpo_mech *m, int32_t k, char *buf, int32_t buflen);
int hpo_mech_molecular_weights(const hpo_mech *m, double *W /* [n_species] */);
/* Mass-specific enthalpy h(T, Y) = sum_k Y_k H_k(T) / W_k in J/kg from the mechanism's NASA7
 * polynomials, computed on the host: T[n], Y[n][ldY] with all n_species mass fractions per row.
 * The quantity the reference's Cantera integrator writes to the mesh's ENTHA
 * (thermo->enthalpy_mass(), src/Integrators/CanteraIntegrator.cpp:93). */
int hpo_mech_enthalpy_mass(const hpo_mech *m, int64_t n, const double *T, const double *Y, int64_t ldY,
                           double *h);
/* Flat host tables (for cross-checking readers): see hpc_ode_amd/csrc/mechanism.hpp. */
int hpo_mech_tables(const hpo_mech *m, double *nasa /*[K][14]*/, double *tmid /*[K]*/,
                    double *arrhenius /*[R][8]: lnA,b,theta,A,lnA0,b0,theta0,A0*/,
                    int32_t *kind /*[R]*/, int32_t *reversible /*[R]*/);

/* ---- resident cell batch (device memory owned by the handle) ---- */
int hpo_batch_create(const hpo_mech *m, int64_t capacity, hpo_batch **out);
void hpo_batch_destroy(hpo_batch *b);
/* Copies cells [first, first+count) from host. y: [count][ldy]; P and dt per cell, or NULL to
 * use the scalar for every cell (the reference's single pressure / dt: Integrator.hpp:13-19). */
int hpo_batch_put(hpo_batch *b, int64_t first, int64_t count, const double *y, int64_t ldy,
                  const double *P, double P_scalar, const double *dt, double dt_scalar);
/* Integrates cells [0, count) from t = 0 to their dt, in place on the device. */
int hpo_batch_integrate(hpo_batch *b, int64_t count, const hpo_opts *o, void *hip_stream,
                        hpo_timing *timing);
int hpo_batch_get(const hpo_batch *b, int64_t first, int64_t count, double *y, int64_t ldy,
                  int32_t *status /* [count] or NULL */, int32_t *stats /* [count][HPO_NSTATS] or NULL */);

/* ---- device-pointer entry: all arrays already in device memory (e.g. torch tensors) ---- */
/* d_order: work-queue order of the cells (a permutation of 0..ncells-1) or NULL.
 * d_prof: [ncells][HPO_NPROF] shader-clock cycles per phase (RHS, Jacobian, LU factorization,
 * LU solves, whole cell) or NULL. With timing != NULL the call waits for the kernel;
 * otherwise it returns after the launch. With timing == NULL the call enqueues its work on
 * hip_stream: a work-queue counter reset, the integration and, with the FP64 fallback armed
 * (lu_precision 1, fp64_fallback 1), a copy of the input rows first when the call runs in
 * place, then a failed-cell count and the FP64 launch; kernels and one device copy only. It can
 * be captured into a HIP graph once an uncaptured call of at least as many cells, in place or
 * not alike, has run on the same mechanism. The graph uses the mechanism's counter, scratch and
 * row copy: its replays must not overlap other calls on that mechanism, and a later call with
 * more cells (which grows them) invalidates it.
 * Uncaptured calls on one mechanism may come from several host threads and streams: they share
 * the mechanism's counter and scratch, so each launch waits on the device for the previous one
 * (an event, when the stream differs) and they never overlap. Launches that must overlap need
 * separate mechanism handles (or the server, whose slots own their counters and scratch). */
#define HPO_NPROF 5
int hpo_integrate_device(const hpo_mech *m, int64_t ncells, const double *d_y_in, double *d_y_out,
                         int64_t ld, const double *d_P, const double *d_dt, int32_t *d_status,
                         int32_t *d_stats, const int32_t *d_order, int64_t *d_prof, const hpo_opts *o,
                         void *hip_stream, hpo_timing *timing);
/* hpo_integrate_device for callers that cannot allocate (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc). No counterpart in the reference. It calls no
 * hipMalloc, hipFree or allocating routine: the work-queue counter and the per-wave scratch come
 * from a fixed static device array of 192 MiB per GPU, held until the process ends, enough for a
 * whole-device launch of the 53-variable GRI-3.0 kernels. A launch that would need more is
 * rejected (HPO_E_INVALID, the message gives the bytes). The mechanism's tables must already be
 * on the current device (any allocating integration call uploads them; hpo_integrate_device with
 * ncells = 0 does nothing else). Every call on one GPU shares that GPU's array and clears its
 * counter with a kernel on hip_stream: calls that overlap on one GPU race, so order them on one
 * stream. It returns after the launch (no timing, no profile). In place (d_y_in and d_y_out
 * overlap) it runs without the FP64 fallback, which would need a copy of the input rows. */
int hpo_integrate_device_noalloc(const hpo_mech *m, int64_t ncells, const double *d_y_in, double *d_y_out,
                                 int64_t ld, const double *d_P, const double *d_dt, int32_t *d_status,
                                 int32_t *d_stats, const int32_t *d_order, const hpo_opts *o,
                                 void *hip_stream);

/* ---- autotuner for uneven workloads (persistent kernel + work queue) ----
 * No counterpart in the reference (its GPU path integrates one batch with shared steps,
 * src/Integrators/CVodeIntegratorGPU.cpp:60-108). Times launch shapes and work-queue orders on
 * an evenly strided sample of the workload (integrated into scratch buffers: the inputs are not
 * modified) and returns the fastest.
 * Orders: 0 = cell index, 1 = initial temperature descending, 2 = predicted cost descending
 * (cost model: mean BDF step count of the sample cells per (T0, log dt) bin).
 * o: in = tolerances and LU precision to tune for; out = waves_per_block, blocks_per_cu set,
 * sort_cells cleared. d_order_out: [ncells] device array, receives the chosen order (a
 * permutation of 0..ncells-1) to pass to hpo_integrate_device. */
#define HPO_TUNE_MAX 16
typedef struct {
  int32_t n_candidates, best;
  int32_t waves_per_block[HPO_TUNE_MAX], blocks_per_cu[HPO_TUNE_MAX], order_mode[HPO_TUNE_MAX];
  int32_t lu_layout[HPO_TUNE_MAX];
  int32_t cell_layout[HPO_TUNE_MAX];   /* 1 one wave per cell, 2 one workgroup per cell */
  double sample_ms[HPO_TUNE_MAX];
  int64_t sample_cells;
  double tune_ms;          /* wall time of the whole tuning call */
} hpo_tune_report;
int hpo_autotune(const hpo_mech *m, int64_t ncells, const double *d_y, int64_t ld, const double *d_P,
                 const double *d_dt, int64_t sample_cells, hpo_opts *o, int32_t *d_order_out,
                 hpo_tune_report *report, void *hip_stream);

/* The same tuning on a resident batch (cells [0, count)). The order is kept in the batch and
 * used by hpo_batch_integrate when o->sort_cells == 2; o comes back with sort_cells = 2. */
int hpo_batch_autotune(hpo_batch *b, int64_t count, int64_t sample_cells, hpo_opts *o,
                       hpo_tune_report *report, void *hip_stream);

/* ---- serving loop: continuous batching of substep requests over resident cell streams ----
 * No counterpart in the reference, which integrates one mesh per application run
 * (app/main.cpp:53-57). For callers that issue chemistry substeps as a CFD solver does:
 * a stream is a set of cells of one mechanism that stays resident on the device between
 * substeps (a CFD partition); opening one is admitted against the server's HBM budget.
 * hpo_server_submit enqueues one substep of a stream. A scheduler thread packs all eligible
 * pending substeps of one mechanism into one persistent-kernel launch whenever one of the
 * server's launch slots is free, so requests that arrive while others run join the next
 * launch, and launches on different slots run on the device at the same time. A stream has at
 * most one substep in flight; its next substep starts from the state the previous one left.
 * Each cell is integrated by one wave with the same instruction sequence, so a request's
 * results do not depend on what it was batched with. */
typedef struct hpo_server hpo_server;
typedef struct {
  int32_t n_slots;           /* launches that may be in flight at once (one HIP stream each);
                                1..16, 0 = 8 */
  int32_t cu_share_pct;      /* largest share of the device's workgroup slots one launch may
                                take, in percent (1..100); 0 = 100. Below 100, a launch that
                                arrives while a large one runs finds free slots at once. */
  int64_t max_batch_cells;   /* cells per launch; 0 = 262144 */
  int64_t hbm_budget_bytes;  /* device memory for resident streams; 0 = 90% of free HBM at creation */
  hpo_opts opts;             /* integration options for every launch. Defaults here differ from
                                hpo_opts_default in two fields: sort_cells = 1 orders each
                                launch's cells by descending initial temperature on the device
                                (0 takes them in packing order), and lu_precision = 1 (FP32
                                Newton matrix, FP64 refinement, FP64 fallback) */
  int64_t launch_cells;      /* a batch of more cells runs as consecutive launches of at most
                                this many on its slot, so that other slots' launches reach the
                                device between them; 0 = one launch per batch */
} hpo_server_opts;
typedef struct {
  double latency_ms;         /* submit -> results available to the caller */
  double queue_ms;           /* submit -> launch of the batch that carried it */
  double run_ms;             /* launch -> completion of that batch (copies included) */
  int64_t ncells;
  int64_t batch_cells;       /* cells in that launch */
  int32_t batch_requests;    /* requests in that launch */
  int32_t failed_cells;      /* cells whose solver status is not HPO_CELL_SUCCESS */
} hpo_request_info;
#define HPO_E_BUSY (-6)      /* HBM budget exhausted: close a stream and retry */
#define HPO_E_TIMEOUT (-7)
void hpo_server_opts_default(hpo_server_opts *o);
int hpo_server_create(hpo_mech *const *mechs, int32_t n_mechs, const hpo_server_opts *o, hpo_server **out);
void hpo_server_destroy(hpo_server *s); /* finishes the launches in flight first */
/* y0: [ncells][ld] host rows [T, Y_0..Y_{K-2}]; P: [ncells] Pa or NULL for P_scalar. */
int hpo_server_open_stream(hpo_server *s, int32_t mech, int64_t ncells, const double *y0, int64_t ld,
                           const double *P, double P_scalar, int64_t *stream_id);
int hpo_server_close_stream(hpo_server *s, int64_t stream_id); /* HPO_E_BUSY while a substep is pending */
/* One substep of dt seconds. y_out ([ncells][ld_out]) and status_out ([ncells]) receive the
 * result when the request completes; either may be NULL (state stays on the device only).
 * The buffers must stay valid until hpo_server_wait returns for this request. */
int hpo_server_submit(hpo_server *s, int64_t stream_id, double dt, double *y_out, int64_t ld_out,
                      int32_t *status_out, int64_t *request_id);
/* Waits for a request (timeout_s < 0: forever); HPO_E_TIMEOUT if it has not completed. A request
 * can be waited for once. */
int hpo_server_wait(hpo_server *s, int64_t request_id, double timeout_s, hpo_request_info *info);
/* Copies a stream's resident state (after its last completed substep) to the host. */
int hpo_server_stream_get(hpo_server *s, int64_t stream_id, double *y, int64_t ld);
int hpo_server_hbm_used(const hpo_server *s, int64_t *used_bytes, int64_t *budget_bytes);

/* ---- component entry points (host arrays), used by parity tests ---- */
/* f = dydt(y) for n states: y, f are [n][n_state]; the pyJac dydt of dydt.c:7-133. Runs the
 * wave-per-cell code when the mechanism has it (at most 63 species), else the block-per-cell code. */
int hpo_eval_rhs(const hpo_mech *m, int64_t n, const double *y, const double *P, double *f);
/* J = d f / d y, [n][n_state][n_state] row-major; what eval_jacob (jacob.c:3-371) computes. */
int hpo_eval_jac(const hpo_mech *m, int64_t n, const double *y, const double *P, double *J);
/* hpo_eval_rhs (what = 0) or hpo_eval_jac (what = 1) with the cell layout and the reactor named
 * as in hpo_opts::cell_layout (0 auto, 1 wave, 2 block) and hpo_opts::reactor (0 constant
 * pressure, P = pressure; 1 constant volume, P = density); HPO_E_UNSUPPORTED when the mechanism
 * has no tables or kernel of that layout. */
int hpo_eval_layout(const hpo_mech *m, int32_t cell_layout, int32_t reactor, int32_t what, int64_t n,
                    const double *y, const double *P, double *out);
/* Batched dense solve with the kernel's register-resident Newton-matrix code: M [n][ns][ns]
 * row-major, b, x [n][ns]. fp32 = 0: FP64 partial-pivot LU and triangular solves; fp32 = 1: the
 * FP32 Gauss-Jordan inverse (same pivots) applied as a matrix-vector product, no refinement. */
int hpo_lu_solve(int32_t ns, int32_t fp32, int64_t n, const double *M, const double *b, double *x,
                 int32_t *ok);
/* hpo_lu_solve with the batched layout chosen, in place of the reference GPU path's MAGMA
 * batched LU (magma_dgetrf_batched / magma_dgetrs_batched via SUNLinSol_MagmaDense,
 * src/Integrators/CVodeIntegratorGPU.cpp:86-91): layout 0 = one wave per matrix (the
 * register code above, fp32 0 or 1); layout 1 = one 256-thread workgroup per matrix, the
 * augmented matrix [M | b] in LDS, FP64 only (fp32 = 1 returns HPO_E_UNSUPPORTED); layout 2 =
 * the block-per-cell kernel's Newton-matrix code (ns 53 or 100): Gauss-Jordan inverse with one
 * matrix row per thread of a 256-thread workgroup, FP64 or FP32, applied to b without refinement.
 * kernel_ms (may be NULL): the launch's device time from HIP events. */
int hpo_lu_solve_layout(int32_t ns, int32_t fp32, int32_t layout, int64_t n, const double *M,
                        const double *b, double *x, int32_t *ok, float *kernel_ms);

const char *hpo_last_error(void);
/* Device properties used for sizing: CU count, free/total HBM bytes. */
int hpo_device_info(int32_t *n_cu, int64_t *free_bytes, int64_t *total_bytes);

#ifdef __cplusplus
}
#endif
#endif
