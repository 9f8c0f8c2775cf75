/*
 * warp_reactor.h -- C ABI of the MI355X (gfx950) primordial-chemistry ROS2S
 * grid advance.
 *
 * The library replaces the reference's device runner: the host loop
 * run_cells_cuda (reproducer.cpp:7791-7975) with its two kernels
 * prepare_grid_timestep_kernel (7696-7734) and
 * advance_collapse_gridwide_kernel (7736-7785), keeping their semantics:
 * per-cell perturbation and free-fall dt, the grid-wide minimum dt, the
 * gridwide-independent ROS2S advance, first-failure-wins error codes and the
 * 1000-step cap.  All arithmetic is IEEE FP64.
 *
 * Conventions
 *  - Functions return an int status: the reference's IntegratorResult values
 *    unchanged (WR_SUCCESS = 1, negative integrator codes), or a distinct
 *    WR_ERR_* code (<= -100) for API/HIP errors.  wr_last_error() returns the
 *    message of the most recent error on the calling thread.
 *  - Plain pointers and sizes only.  Host pointers unless stated.
 *  - A context owns all device memory and one HIP stream; calls on a context
 *    are synchronous and must come from one host thread at a time.
 */
#ifndef WARP_REACTOR_H
#define WARP_REACTOR_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define WR_ABI_VERSION 2

#define WR_NUM_SPEC 14
#define WR_NEQS 15

/* IntegratorResult (reproducer.cpp:30-37) and API errors. */
enum wr_status {
    WR_SUCCESS = 1,
    WR_BAD_INPUTS = -1,
    WR_DT_UNDERFLOW = -2,
    WR_TOO_MANY_STEPS = -4,
    WR_TOO_MUCH_ACCURACY_REQUESTED = -5,
    WR_LU_DECOMPOSITION_ERROR = -7,
    WR_ERR_HIP = -100,      /* a HIP runtime call failed */
    WR_ERR_ARGUMENT = -101, /* invalid argument (null pointer, size) */
    WR_ERR_CAPACITY = -102, /* more cells than the context was created for */
    WR_ERR_IO = -103        /* file I/O error */
};

/* burn_t (reproducer.cpp:259-264): 136 bytes. */
typedef struct wr_burn_state {
    double rho;
    double T;
    double e;
    double xn[WR_NUM_SPEC];
} wr_burn_state;

/* IntegratorStats (reproducer.cpp:7433-7441): 56 bytes. */
typedef struct wr_integrator_stats {
    uint64_t internal_steps;
    uint64_t rhs_calls;
    uint64_t jacobian_calls;
    uint64_t decompositions;
    uint64_t linear_solves;
    uint64_t accepted_steps;
    uint64_t rejected_steps;
} wr_integrator_stats;

/* CollapseState (reproducer.cpp:7455-7461): 216 bytes, same layout. */
typedef struct wr_collapse_state {
    wr_burn_state current;
    double time;
    double density_driver;
    int32_t completed_steps;
    int32_t reserved_;
    wr_integrator_stats stats;
} wr_collapse_state;

/* PackedFinalState (reproducer.cpp:7488-7505): 172 bytes, little endian,
 * no padding.  i = cell % N, j = (cell / N) % N, k = cell / N^2. */
#pragma pack(push, 1)
typedef struct wr_packed_final_state {
    int32_t cell;
    int32_t i;
    int32_t j;
    int32_t k;
    int32_t completed_steps;
    double time;
    double density_driver;
    double rho;
    double T;
    double e;
    double xn[WR_NUM_SPEC];
} wr_packed_final_state;
#pragma pack(pop)

/* Why a run stopped. */
enum wr_stop_reason {
    WR_STOP_NONE = 0,           /* still running / not run */
    WR_STOP_STEP_LIMIT = 1,     /* max_steps outer steps completed (reproducer.cpp:7859) */
    WR_STOP_NO_CANDIDATE = 2,   /* no cell offered a dt (7894-7896) */
    WR_STOP_NONE_INTEGRATED = 3,/* advance integrated no cell (7954-7956) */
    WR_STOP_FAILURE = 4         /* a cell failed; run terminal (7881-7886, 7948-7953) */
};

/* Outcome of a run, filled by wr_run_info(). */
typedef struct wr_run_info {
    int32_t status;             /* wr_status */
    int32_t stop_reason;        /* wr_stop_reason */
    int32_t completed_global_steps;
    int32_t failed_step;        /* outer step of the failure, or -1 */
    int32_t failed_phase;       /* 1 = prepare, 2 = advance, 0 = none */
    int32_t failed_cell;        /* cell whose failure ended the run, or -1 */
    double grid_time;           /* physical time of the grid after the run */
    uint64_t cell_advances;     /* sum over steps of integrated cells */
    int32_t limiting_cell;      /* (dt, lowest cell) holder of the last planned step, or -1 */
    int32_t reserved_;
} wr_run_info;

typedef struct wr_context wr_context;

/* ---- version / errors ------------------------------------------------------ */
int wr_abi_version(void);
const char *wr_last_error(void);

/* ---- drop-in runner ----------------------------------------------------------
 * Replaces run_cells_cuda (reproducer.cpp:7791): uploads the cells, runs the
 * gridwide-independent outer loop (<= 1000 steps) and downloads them.  On any
 * HIP error prints "<action> failed: <message>" to stderr and returns
 * WR_BAD_INPUTS, as the reference does (7688-7694). */
int wr_run_cells(wr_collapse_state *cells, int num_cells, int perturb,
                 int *completed_global_steps);

/* ---- context ---------------------------------------------------------------- */
int wr_context_create(int device, int capacity_cells, wr_context **out);
void wr_context_destroy(wr_context *ctx);
/* cells: host memory, or device memory of the context's device (the copy
 * direction follows from unified addressing; device data must be complete
 * before the call, and the call returns after its copy has finished). */
int wr_upload_cells(wr_context *ctx, const wr_collapse_state *cells, int num_cells);
int wr_download_cells(wr_context *ctx, wr_collapse_state *cells, int num_cells);

/* Run the outer loop on the resident cells from global step
 * `completed_global_steps` and grid time `grid_time` (normally 0, 0) for at
 * most `max_steps` steps (<= 0: 1000).  Resume semantics: the outer-loop
 * index that the perturbation schedule reads (every 20th step,
 * reproducer.cpp:7564) starts at the completed count, so a run split into
 * pieces perturbs at the same steps as one run.  (run_cells_cuda itself starts
 * its loop index at 0 whatever count the caller passes, 7859; wr_run_cells
 * and wr_run_from(..., 0) do the same.)  Host-side control is device
 * resident: the host only polls a completion flag every few steps. */
int wr_run(wr_context *ctx, int perturb, int max_steps, int completed_global_steps,
           double grid_time);
/* wr_run with an explicit loop index for the first step: first_step = 0 is
 * run_cells_cuda's loop (reproducer.cpp:7859-7959) for any completed count;
 * first_step = completed_global_steps is wr_run. */
int wr_run_from(wr_context *ctx, int perturb, int max_steps, int completed_global_steps,
                double grid_time, int first_step);
int wr_run_info_get(const wr_context *ctx, wr_run_info *out);
/* dt_grid of each outer step of the last wr_run (the host loop's dt_grid,
 * reproducer.cpp:7888-7897), in step order: copies min(count, max_entries)
 * values, *count = number of steps recorded (at most 4096; a last step that
 * planned a dt but integrated no cell is recorded too).  With the initial
 * states this determines every cell's trajectory independently of the others,
 * so single cells can be replayed and checked. */
int wr_dt_history(const wr_context *ctx, double *out, int max_entries, int *count);

/* One kernel of the outer step each, for callers that drive the loop:
 * prepare (reproducer.cpp:7696) + device min reduction (7888-7896);
 * *dt_grid is DBL_MAX when no cell offers a candidate.  `*limiting_cell` is
 * the lowest cell index holding the minimum (or -1). */
int wr_prepare(wr_context *ctx, int completed_global_steps, double grid_time, int step,
               int perturb, double *dt_grid, int *limiting_cell);
/* No counterpart in the reference: wr_prepare for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  The prepare phase of one outer step on
 * the caller's DEVICE array of CollapseState records (the layout of the
 * reference's prepare_grid_timestep_kernel, reproducer.cpp:7696-7734) and the
 * grid-wide (dt, lowest cell) minimum (7888-7893), enqueued on `stream` (a
 * hipStream_t; NULL: the default stream) and not waited for.  Results go to
 * caller-given DEVICE memory: *dt_grid (DBL_MAX when no cell offers a
 * candidate or a cell failed), *limiting_cell (-1 then) and *failure (1 or
 * the first failure code); they equal wr_prepare's on the same cells.  The
 * call allocates nothing: its per-block partials and failure word live in a
 * fixed device array of the library (768 KiB per GPU), so at most 2^24 cells
 * (larger counts return WR_ERR_ARGUMENT), and calls that overlap on one GPU
 * race on it. */
int wr_prepare_cells_noalloc(wr_collapse_state *cells, int num_cells, int completed_global_steps,
                             double grid_time, int step, int perturb, double *dt_grid,
                             int *limiting_cell, int *failure, void *stream);
/* advance (reproducer.cpp:7736); *integrated = number of cells integrated. */
int wr_advance(wr_context *ctx, int completed_global_steps, double next_grid_time,
               double dt_grid, int *integrated);

/* Driver policy of subsequent wr_run calls.  WR_POLICY_GRIDWIDE (default):
 * the gridwide-independent loop of run_cells_cuda (reproducer.cpp:7791-7975),
 * one grid-wide dt per outer step.  WR_POLICY_LEGACY_LOCAL: the reference's
 * CPU driver (reproducer.cpp:8628-8648): every outer step each cell that has
 * not stopped applies its perturbation and takes one step of its own
 * free-fall dt (advance_collapse_step, 7646-7685); the run stops when every
 * cell stopped or at the first failure, and, as the reference's sequential
 * cell loop stops at the lowest failing cell, the cells after it keep their
 * state of that step's start.  grid_time is not used (each cell keeps its own
 * time); wr_dt_history records no steps.  The legacy-local policy runs on the
 * queue-fed kernel in FP64 (FP32 factorisation: WR_ERR_ARGUMENT).
 * WR_POLICY_GRIDWIDE_CTA32: the gridwide-cta32 policy of the reference's Mojo
 * front end (reproducer.mojo:2829-3093, driver 3604-3621; --driver-policy
 * gridwide-cta32): the gridwide outer step with the density updates evaluated
 * as (dt * driver) / tff, and the cells advanced in tiles of 32 consecutive
 * cells that share one ROS2S step controller (tile error = the largest cell
 * error, a singular factorisation in any cell halves h for the tile, an
 * accepted final trial ends exactly at dt_grid; the Jacobian counter counts
 * every trial).  A failing tile ends the run with the code of the lowest
 * failing tile (failed_cell = its first cell); the Mojo driver then raises,
 * so the cells' states after a failed step are not defined.  FP64 reference
 * integrator only (FP32 factorisation, Jacobian reuse: WR_ERR_ARGUMENT). */
enum wr_policy { WR_POLICY_GRIDWIDE = 0, WR_POLICY_LEGACY_LOCAL = 1, WR_POLICY_GRIDWIDE_CTA32 = 2 };
int wr_set_policy(wr_context *ctx, int policy);

/* Advance kernel of subsequent steps: 1 (default) = queue-fed persistent
 * kernel (lanes take the next cell when theirs finishes; register-budgeted
 * network code); 0 = one thread per cell.  Results are identical. */
int wr_set_advance_kernel(wr_context *ctx, int kernel);

/* Launch shape of the queue-fed kernel (no counterpart in the reference,
 * whose advance is one thread per cell, reproducer.cpp:7736): threads per
 * block (64..256, multiple of 64), resident blocks per CU (0: as many as the
 * kernel's occupancy allows), and the queue positions reserved for block 0's
 * first wave, which then runs alone on its CU (0 or 64; the most expensive
 * cells of the last step).  Results do not depend on it; the autotuner
 * (warp-reactor_amd/autotune.py) picks it by timing.  Defaults 256, 0, 64. */
int wr_set_queue_launch(wr_context *ctx, int threads_per_block, int blocks_per_cu,
                        int solo_lanes);
/* Wave-cooperative advance of the most expensive cells (no counterpart in the
 * reference, whose advance is one thread per cell, reproducer.cpp:7736): in
 * every outer step the first `cells` queue positions -- the cells that took
 * the most ROS2S attempts in the previous step -- are advanced by whole
 * 64-lane waves, one cell per wave, taken by wave 0 of the first `waves`
 * blocks of the queue-fed kernel; the network runs as cooperative programs
 * over the wave's LDS, the LU and solves at wave level.  The single-lane
 * path takes the other cells, and the cooperative waves join it when their
 * cells are done.  Results are bit-identical to the single-lane path.
 * Applies to the queue-fed FP64 kernel with the reference's Jacobian rule
 * and the gridwide policy; cells = 0 or waves = 0 turns it off.  cells = -1
 * (the default) takes one cooperative cell per 2,048 cells of the context,
 * at most 128, with as many waves (none under 2,048 cells). */
int wr_set_coop(wr_context *ctx, int cells, int waves);
/* The cells and waves the setting gives for the context's grid, and the LDS
 * bytes a cooperative wave uses. */
int wr_get_coop(const wr_context *ctx, int *cells, int *waves, int *lds_bytes);

/* The current shape and the grid size last used (0 before the first launch). */
int wr_get_queue_launch(const wr_context *ctx, int *threads_per_block, int *blocks_per_cu,
                        int *solo_lanes, int *blocks);

/* Factorisation of I - gamma h J in the queue-fed kernel of subsequent steps
 * (no counterpart in the reference, which factorises in FP64 only:
 * reproducer.cpp:7159-7171).  mode 0 (default): FP64, results identical to the
 * reference arithmetic.  mode 1: E = fac I - J rounded to FP32 and factorised
 * in FP32; every stage solve takes `sweeps` (1..4) refinement sweeps with the
 * residual b - E x in FP64; the RHS, energy equation, error norm and step
 * controller stay FP64.  The unknowns are equilibrated by exact powers of two,
 * 2^ilogb(atol_i + rtol_i |y_i|), before rounding to FP32.  FP64 factors are
 * used for the rest of an attempt when the FP32 factors are singular or a
 * refinement does not converge (last correction > 1e-6 (atol_i + rtol_i |y_i|)
 * in some component).  Mode 1 needs advance kernel 1 (runs otherwise fail
 * with WR_ERR_ARGUMENT). */
int wr_set_lu_precision(wr_context *ctx, int mode, int sweeps);
/* Jacobian reuse in the queue-fed kernel of subsequent gridwide steps (a study
 * with no counterpart in the reference, which evaluates J at the head of every
 * outer iteration of integrate and reuses it only on rejection,
 * reproducer.cpp:7236-7260).  age_max = 1 (default): the reference.
 * age_max = k (2..64): J is evaluated at a burn's first attempt and again once
 * k steps have been accepted with it (the attempts in between evaluate f
 * alone, so ROS2S runs as a W-method with a lagged matrix); an attempt
 * rejected (error or singular matrix) with a J from an earlier state evaluates
 * J at the current state before the retry.  The results differ from the
 * reference's by what the controller makes of the lagged J.  Needs advance
 * kernel 1 with FP64 factors and the gridwide policy (runs otherwise fail with
 * WR_ERR_ARGUMENT). */
int wr_set_jacobian_reuse(wr_context *ctx, int age_max);
/* Fallback counts of the last wr_run in mode 1: attempts whose FP32 factors
 * were singular, stage solves whose refinement did not converge. */
int wr_mixed_fallbacks(const wr_context *ctx, unsigned long long *lu_fallbacks,
                       unsigned long long *refine_fallbacks);

/* Integrator step limit of the burns of subsequent runs (<= 0: 1e7, the
 * configure_ros2s value, reproducer.cpp:7589).  Lower values force
 * WR_TOO_MANY_STEPS failures, for testing the failure path. */
int wr_set_integrator_max_steps(wr_context *ctx, int max_steps);

/* wr_run enqueues its outer steps 8 at a time as one captured HIP graph
 * (default 1), or as individual launches (0).  Runs with timing enabled use
 * individual launches.  Results are identical. */
int wr_set_graph(wr_context *ctx, int enable);

/* Time the advance kernel launches of subsequent wr_run calls with HIP events
 * recorded on the context's stream (off by default). */
int wr_set_timing(wr_context *ctx, int enable);
/* Total device time (ms) of the advance launches of the last timed wr_run
 * that integrated cells, and their count. */
int wr_last_advance_timing(const wr_context *ctx, double *total_ms, int *launches);

/* ---- discrete adjoint (no counterpart in the reference) ----------------------
 * Sensitivities of the collapse advance with its controller decisions frozen
 * at a recorded run: dt_grid of every outer step and the step sizes of every
 * cell's accepted ROS2S attempts (the forward path is reproducer.cpp:7736-7785,
 * 7205-7347).  Parameters: each cell's state at its first advance and the rate
 * multipliers M_r = exp(theta_r) of the network's reactions (recovered by the
 * code generator from the reference's rate code; wr_adjoint_ntheta() of them):
 * M_r scales reaction r in the species rates, in the energy rate where the
 * network books its heat, and in the Jacobian.
 * States are [18][n] arrays, field-major: xn[0..13], T, e, rho,
 * density_driver; states, seeds and gradients may be host memory or device
 * memory of the context's device (as for wr_upload_cells), multipliers are
 * host memory.  e of the first state does not enter the map (eos_rt
 * recomputes it from T).  With wr_adjoint_set_perturb(ctx, 1) the recording
 * applies the perturbations of a run whose loop index starts at 0 (steps 20,
 * 40, ...; reproducer.cpp:7564-7580), and they are part of the map. */
int wr_adjoint_ntheta(void);
/* Run `nsteps` outer steps of the uploaded cells with multipliers M (no
 * perturbation, every cell must integrate in every step) and record the tape:
 * checkpoints of every step and up to `hcap` accepted step sizes per cell.
 * The cells' final state stays in the context (wr_download_cells).  Returns
 * WR_SUCCESS, a failure code, WR_ERR_CAPACITY (tape full) or WR_ERR_ARGUMENT
 * (a cell stopped integrating). */
int wr_adjoint_record(wr_context *ctx, int nsteps, const double *multipliers, int hcap,
                      int *completed_global_steps);
/* The frozen map from s0 (NULL: the recorded first state) with multipliers
 * (NULL: the recorded ones) to sN. */
int wr_adjoint_replay(wr_context *ctx, const double *s0, const double *multipliers, double *sN);
/* gN = dL/d(final state) -> g0 = dL/d(first state), mbar [R][n] = dL/dM_r per
 * cell, at the recorded state and multipliers. */
int wr_adjoint_backward(wr_context *ctx, const double *gN, double *g0, double *mbar);
/* Recordings count per context: *generation = the number of successful
 * wr_adjoint_record calls so far (0: none).  wr_adjoint_backward_gen is
 * wr_adjoint_backward for the recording with that number, and returns
 * WR_ERR_ARGUMENT when the context's tape is from another one (a later
 * forward on the same context overwrote it).  No counterpart in the
 * reference. */
int wr_adjoint_generation(const wr_context *ctx, unsigned long long *generation);
int wr_adjoint_backward_gen(wr_context *ctx, unsigned long long generation, const double *gN,
                            double *g0, double *mbar);
/* The recorded tape: hoff [nsteps+1][n] (first tape index of each step),
 * htape [hcap][n], dts [nsteps], ckpt [nsteps][18][n]; any may be NULL. */
int wr_adjoint_tape(const wr_context *ctx, int32_t *hoff, double *htape, double *dts,
                    double *ckpt);
/* The backward pass takes the recorded steps in windows from the last one
 * backwards, each window's accepted steps stored at once in one of two
 * buffers (each sized to an eighth of the free device memory at the call,
 * 1-32 GiB, and windows to at most a quarter of the steps); max_steps > 0
 * caps the steps per window (0: the budget alone).  Results do not depend on
 * it. */
int wr_adjoint_set_window(wr_context *ctx, int max_steps);
/* (new) Whether subsequent wr_adjoint_record calls apply the density
 * perturbations (0: no, the default; 1: yes, as wr_run with perturb = 1 from
 * loop index 0).  The recording then also keeps each cell's state before every
 * perturbation step (18 doubles per cell per 20 steps), and replay and
 * backward differentiate through the perturbations. */
int wr_adjoint_set_perturb(wr_context *ctx, int perturb);
/* (new) Cells per 64-lane wave of the backward pass's reverse sweep, 1..64;
 * 0 (the default) means 64.  Results do not depend on it. */
int wr_adjoint_set_lanes(wr_context *ctx, int cells_per_wave);
/* Device time (HIP events on the context's stream) of the last record,
 * replay or backward call, ms. */
int wr_adjoint_last_ms(const wr_context *ctx, double *ms);

/* ---- device building blocks (batched; for verification) --------------------
 * y: n x 15 row-major host arrays; outputs likewise.  Each runs the same
 * device functions the advance kernel uses. */
int wr_eval_rhs(wr_context *ctx, const double *y, int n, double *dydt);
/* Device libm on x[n]: op 0 exp, 1 log, 2 cbrt, 3 sqrt as the advance kernel
 * evaluates them; 4 exp, 5 log, 6 cbrt of the ROCm device library. */
int wr_eval_math(wr_context *ctx, int op, const double *x, int n, double *out);
int wr_eval_jac(wr_context *ctx, const double *y, int n, double *jac /* n x 225 */);
int wr_eval_lu(wr_context *ctx, double *a /* n x 225, in/out */, int n, int32_t *ipvt /* n x 15 */,
               int32_t *info /* n */);
int wr_eval_lu_solve(wr_context *ctx, const double *lu, const int32_t *ipvt, int n,
                     double *x /* n x 15, in/out */);
/* One stage solve of the FP32-factorisation form (wr_set_lu_precision mode 1)
 * per row: E = fac I - J equilibrated by 2^ilogb(atol_i + rtol_i |y_i|) and
 * factorised in FP32 (factors to lu32, n x 225 floats), then x (in: b) solved
 * with `sweeps` (0..4) FP64-residual refinement sweeps.  conv[i]: -1 singular
 * FP32 factors, 1 converged, 0 not converged. */
int wr_eval_solve_mixed(wr_context *ctx, const double *jac, const double *fac, const double *y,
                        int n, int sweeps, double *x, int32_t *conv, float *lu32);
/* RODAS<PrimordialChem>::integrate from t = 0 to tout[i] with first step
 * dt0[i] and the collapse tolerances; counters n x 7 = n_step, n_rhs, n_jac,
 * n_decomp, n_solve, n_accept, n_reject; t_end/dt_end as the integrator
 * leaves them; result n codes.  max_steps <= 0: 1e7. */
int wr_eval_integrate(wr_context *ctx, double *y, const double *tout, const double *dt0, int n,
                      int max_steps, int32_t *counters, double *t_end, double *dt_end,
                      int32_t *result);

/* ---- initial states, packed final states, comparison (host) ------------------ */
/* make_collapse_state (reproducer.cpp:7533-7546) for n cells. */
int wr_make_collapse_states(wr_collapse_state *cells, int num_cells);
/* Benchmark-recipe initial states (no counterpart in the reference, which
 * only has make_collapse_state): log-normal hydrogen density
 * n_H = exp(sigma_ln_nh * z) cm^-3 with z standard normal (Box-Muller on
 * SplitMix64 of seed ^ cell), temperature t_min, or log-uniform in
 * [t_min, t_max] when t_max > t_min; ionisation 1e-4, H2 1e-6, He mass
 * fraction 0.24, D/H 3e-5, electrons by charge balance, then normalised and
 * e from the EOS as make_initial_state does (7533-7541).
 *   16^3 recipe: seed 0, sigma 0.1, T = 300 K.
 *   128^3 recipe: seed 1, sigma 1.0, T log-uniform in [100, 8000] K. */
int wr_make_recipe_states(wr_collapse_state *cells, int num_cells, uint64_t seed,
                          double sigma_ln_nh, double t_min, double t_max);
/* make_packed_final_state (8289-8305). */
int wr_pack_final_states(const wr_collapse_state *cells, int num_cells, int grid_dim,
                         wr_packed_final_state *out);
/* write_final_states (8357-8378), rotating an existing file to
 * PATH.old.NNNNNN first (8307-8355). */
int wr_write_final_states(const wr_collapse_state *cells, int num_cells, int grid_dim,
                          const char *path);
/* read_final_states (8380-8414): exact size check. */
int wr_read_final_states(const char *path, int expected_records, wr_packed_final_state *out);

typedef struct wr_comparison_summary {
    int32_t pass;
    int32_t failures;           /* number of failing fields */
    double max_deuterium_species_rel_error;
    double max_non_deuterium_species_rel_error;
    double max_thermodynamic_rel_error;
    double max_rho_rel_error;
    char first_failure[512];
} wr_comparison_summary;

/* compare_final_states (8416-8549) against a packed file; when csv_path is
 * non-null and the comparison fails, writes the failures CSV (8100-8142). */
int wr_compare_final_states(const wr_collapse_state *cells, int num_cells, int grid_dim,
                            const char *path, const char *csv_path,
                            wr_comparison_summary *out);

#ifdef __cplusplus
}
#endif

#endif /* WARP_REACTOR_H */
