/*
 * seisfem_gpu.h -- C ABI of the MI355X time loop for SeisFEM's explicit
 * elastic M-PML masslump solver.
 *
 * The library takes the arrays the reference's elastic_wave() builds during
 * setup (time_evolution/elastic_wave.c:136-313) -- the seven node-indexed CSR
 * matrices M, S1..S6, the lumped mass, the Dirichlet flags, the four elastic
 * constants and the six M-PML profiles -- and replaces the step loop
 * (elastic_wave.c:446-818) with hand-written HIP kernels on one GPU.
 *
 * Layouts are the reference's: 0-based CSR (int row pointers of length
 * node_num+1, ascending int columns, double values), node vectors double[N],
 * Dirichlet flags int[N] (1 = clamp), station nodes int[] (0-based).
 * All pointers passed to the sfg_* calls are HOST pointers; device memory is
 * owned by the library.  Functions return SFG_OK (0) or a negative SFG_E_*
 * code; sfg_last_error() gives a message.  No function calls exit().
 */
#ifndef SEISFEM_GPU_H
#define SEISFEM_GPU_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    SFG_OK = 0,
    SFG_E_ARG = -1,      /* invalid argument / inconsistent sizes */
    SFG_E_HIP = -2,      /* HIP runtime error (message in sfg_last_error) */
    SFG_E_NOMEM = -4,
    SFG_E_ENERGY = -6,   /* energy guard fired: sum(U^2) or sum(W^2) > 1e7, or NaN
                            (elastic_wave.c:779-785 exits with status 1 instead) */
    SFG_E_ZERO_MASS = -7 /* a zero lumped-mass entry was met (elastic_wave.c:726-727) */
};

/* Arithmetic of a run.  FP64: bit-identical to the reference's FP64 loop.
 * FP32: single precision state and arithmetic.  FP32_KAHAN: FP32 with a
 * compensated displacement update.  FP32_KAHAN16: the same update with its
 * compensation terms kept as bfloat16 (half the bytes of the terms). */
enum { SFG_FP64 = 0, SFG_FP32 = 1, SFG_FP32_KAHAN = 2, SFG_FP32_KAHAN16 = 3 };

/* Last error message of the calling thread ("" if none). */
const char *sfg_last_error(void);
/* Selects the HIP device used by later calls of this thread (default 0). */
int sfg_set_device(int device);

/* ---- drop-in replacements of the reference's hot-path entry points ------- */

/* y = A x on the GPU, row sums sequential in column order: bit-identical to
 * the reference.  Replaces sparse_matrix/csr_matvec.c:1.  Host pointers;
 * errors are reported through sfg_last_error() (the reference returns void). */
void csr_matvec(int Bp_size, int *Bp, int *Bj, double *Bx, double *x, double *y);

/* M-PML damping profiles, computed in FP64 on the GPU with the reference's
 * region tests and operand order.  Replaces pml/abc_mpml.c:1-4 (same
 * arguments).  Every node gets the value of its own coordinates, which is
 * what the reference's per-element visits give; nodes that no element
 * references are set the same way (the reference leaves them unset). */
void abc_mpml(int node_num, int element_num, int element_order, int *element_node,
              double **node_xy, int mpml_nx, int mpml_ny, double h, double xmin, double xmax,
              double ymin, double ymax, double vp_max, int use_mpml_xmin, int use_mpml_xmax,
              int use_mpml_ymin, int use_mpml_ymax, double *mpml_dx, double *mpml_dy,
              double *mpml_dxx, double *mpml_dyy, double *mpml_dxx_pyx, double *mpml_dyy_pxy);

/* Ricker wavelet magnitude*(1-2a)exp(-a), a = (pi f0 (t-t0))^2.
 * Replaces source_receiver/seismic_source.c:1 (host function, same result). */
double seismic_source(double f0, double t0, double magnitide, double t);

/* The whole elastic_wave() with the reference's 26 arguments
 * (time_evolution/elastic_wave.c:1-3): setup (:136-313) with the host setup of
 * seisfem_host.h, then every shot at once on the GPU, writing
 * ./outputfile/seismogram_{u,w}_shot_N.txt ("%f   " rows) and
 * wavefield_{u,w}_shot_N.txt ("%f\t" rows every 200 steps) as the reference
 * does.  Only solver "masslump" runs; other solvers and errors set
 * sfh_driver_status() != 0 instead of exiting. */
void elastic_wave(char *type, int node_num, int element_num, int element_order, int *element_node,
                  double **node_xy, int nnz, int csr_p_size, int step, double dt, double f0, double t0,
                  double edge_size, double xmin, double xmax, double ymin, double ymax, int pml_nx,
                  int pml_ny, int src_num, int *src_node, int rec_num, int *rec_node, char *solver,
                  int use_exterior_mesh, int free_surface_code);
/* main() of seisfem/seisfem.c:28-299 for the run directory workdir (par.txt,
 * mesh/, outputfile/); chdir()s into it.  Returns 0 or a negative code. */
int seisfem_run(const char *workdir);
/* Precision (SFG_*) and an extra full-precision binary seismogram file
 * (outputfile/seismogram_uw_shot_N.f64) for later elastic_wave() calls. */
void sfh_driver_options(int precision, int binary_output);
/* Status of the last elastic_wave() call (0 = success). */
int sfh_driver_status(void);

/* ---- operators, media and runs -------------------------------------------- */

typedef struct sfg_operator sfg_operator; /* M, S1..S6 of one mesh on the device */
typedef struct sfg_medium sfg_medium;     /* per-node coefficients of one model  */
typedef struct sfg_run sfg_run;           /* a batch of shots being stepped      */

/* Uploads the seven CSR matrices (index 0 = M, 1..6 = S1..S6, the order of
 * elastic_wave.c:236-313).  Each keeps its own zero-eliminated pattern on
 * input; the device stores them on the union pattern (explicit zeros added),
 * which leaves every row sum bit-identical. */
int sfg_operator_create(sfg_operator **op, int node_num, const int *const csr_p[7],
                        const int *const csr_j[7], const double *const csr_x[7]);
int sfg_operator_nnz(const sfg_operator *op); /* union-pattern nnz */
void sfg_operator_destroy(sfg_operator *op);

/* Per-node model: c[4] = c11, c13, c33, c44 (elastic_wave.c:206, :585-657),
 * mass_lump (HRZ, :236-238), dirichlet flags (:136-155), mpml[6] = dx, dy,
 * dxx, dyy, dxx_pyx, dyy_pxy (:227-228).  The FP32 copy folds the Dirichlet
 * flag into 1/mass_lump (computed in FP64 on the GPU). */
int sfg_medium_create(sfg_medium **md, const sfg_operator *op, const double *const c[4],
                      const double *mass_lump, const int *dirichlet, const double *const mpml[6]);
void sfg_medium_destroy(sfg_medium *md);
/* Per-row factor on the M-products of this medium (default 1).  The consistent
 * mass samples rho at the row node (assemble/mass_sparse_q4.c:252), so a model
 * whose density differs from the one the operator's M was assembled with uses
 * row_scale[i] = rho_model[i] / rho_operator[i].  Applied with per-shot media
 * (sfg_run_create_media); equal to the reference up to rounding (bit-identical
 * only when the factor is 1). */
int sfg_medium_set_row_scale(sfg_medium *md, const double *row_scale);

typedef struct {
    int precision;          /* SFG_FP64 / SFG_FP32 / SFG_FP32_KAHAN / SFG_FP32_KAHAN16 */
    int nshots;             /* shots sharing the operator (right-hand sides) */
    const int *src_node;    /* [nshots] source node per shot (elastic_wave.c:448) */
    const int *nsteps;      /* [nshots] executed steps per shot, or NULL = max_steps for all */
    int max_steps;          /* executed steps (the reference executes step-2) */
    int rec_num;
    const int *rec_node;    /* [rec_num] */
    double dt, f0, t0;
    double amplitude;       /* 1e10 in the reference (elastic_wave.c:526) */
    double angle_deg;       /* 90 in the reference (elastic_wave.c:66) */
    int block;              /* threads per block, 0 = tuned default */
    int shots_per_lane_group; /* SP in {1,4,16,32,64}: consecutive shots sharing one row
                                 in a wavefront (64/SP rows per wave); 0 = chosen from nshots */
    int guard_every;        /* energy-guard interval in steps, 0 = 50 */
    int use_graph;          /* 1 = replay captured hipGraphs of guard_every steps */
    int xcd_remap;          /* 0 = default (on): each XCD steps one contiguous row range; -1 = off
                               (round-robin); Q > 0 = stripes: all XCDs sweep the rows together, XCD x
                               taking row blocks x*Q .. x*Q+Q-1 of every 8*Q */
    /* Optional distributed sources (no counterpart in the reference, whose only
     * source is the point force at src_node): shot s adds wavelet * weight to
     * the u / w rows of src_count[s] nodes.  NULL src_count = point forces. */
    const int *src_count;      /* [nshots] */
    const int *src_nodes;      /* [sum of src_count] */
    const double *src_weight;  /* [sum of src_count][2]: u and w weights */
    int kernel_variant;        /* 0 = default (1), 1 = packed values + 2-way gather, 2 = separate value
                                  arrays, 3 = 1 + streaming hints on data used once per step,
                                  4 = 3 + own-row fields kept from the diagonal gather and own-row
                                  accelerations (p in the FP32 forms) / Kahan terms loaded first,
                                  5 and 6 = 4 with 3 and 4 nonzeros' gathers in flight per trip
                                  (6: 2 per trip on rows with an M-PML profile in the FP32 forms),
                                  7 = LDS-staged gathers: each
                                  block copies the fields of its rows' neighbour nodes into LDS once
                                  (16, 32 or 64 shots per lane group only), 8 = sliced layout for one
                                  shot per lane group: the matrix is stored so that nonzero k of the
                                  64 rows of a wave is one coalesced load per matrix (rows sorted by
                                  length within windows of sell_sigma rows of the row order).  Same results (bit for
                                  bit in FP64; in the FP32 forms variant 7 can differ in the sign
                                  of a zero); a multi-run launch runs 4 to 6 as 5 in the FP32 forms
                                  and as 4 in FP64, 3 as 3 for 32- and 64-shot groups (else as
                                  1), 7 as 1, 8 as 8 */
    const int *row_order;      /* [node_num] permutation: the order in which the step kernel's
                                  waves take rows (e.g. sfh_locality_order), NULL = node order.
                                  Changes only which rows run together, not any result */
    int sell_sigma;            /* kernel_variants 8 and 9: rows per sorting window, a multiple of
                                  64 (dividing the block for 9); 0 = 64, each slice of 64 rows
                                  sorted on its own */
} sfg_shot_desc;

/* Creates a run with all state zeroed (elastic_wave.c:466-502). */
int sfg_run_create(sfg_run **run, const sfg_operator *op, const sfg_medium *md,
                   const sfg_shot_desc *desc);
/* As sfg_run_create, with one medium per shot (media[nshots], all built on
 * op): many models of one mesh share the seven operator matrices while each
 * shot reads its own coefficients (stored [coef][node][shot] on the device).
 * The reference runs one model per process; this batches them. */
int sfg_run_create_media(sfg_run **run, const sfg_operator *op, const sfg_medium *const *media,
                         const sfg_shot_desc *desc);
/* Executes up to nsteps further steps (fewer at the end).  Returns the number
 * of steps executed so far (>= 0) or an error code. */
int sfg_run_advance(sfg_run *run, int nsteps);
/* Returns the run to its state at creation without reallocating: every field,
 * seismogram and energy record zeroed and the step counter at 0, so the same
 * shots can be stepped again (queued on the run's stream, no host wait).
 * Not for a run stepped by a multi-run. */
int sfg_run_reset(sfg_run *run);
/* Waits for queued work; returns the executed step count. */
int sfg_run_sync(sfg_run *run);
/* Receiver displacement U_now/W_now[rec_node] per executed step
 * (elastic_wave.c:787-795), host arrays [nshots][max_steps][rec_num]. */
int sfg_run_seismograms(sfg_run *run, double *u, double *w);
/* Current U = U1+U2+U3 and W = W1+W2+W3 of one shot (snapshot rows,
 * elastic_wave.c:797-806), host arrays [node_num]. */
int sfg_run_snapshot(sfg_run *run, int shot, double *U, double *W);
/* sfg_run_snapshot without any device allocation, for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  No counterpart in the reference.  U and W are staged in
 * a fixed device array of 2 * SFG_NOALLOC_MAX_NODES doubles (64 MiB, one copy
 * per GPU, held until the process ends); runs on meshes with more nodes get
 * SFG_E_ARG.  Every call on one GPU shares that array: calls from one process
 * are serialised by the library, so calls that overlap on a GPU only wait for
 * each other.  The run must live on the calling thread's current device.
 * Same values as sfg_run_snapshot, bit for bit. */
#define SFG_NOALLOC_MAX_NODES (1 << 22)
int sfg_run_snapshot_noalloc(sfg_run *run, int shot, double *U, double *W);
/* Energy guard values sum(U^2), sum(W^2) per check: [nchecks][nshots][2]. */
int sfg_run_energy(sfg_run *run, double *energy, int max_checks);
/* Average device time per step kernel (ms) measured with HIP events on the
 * run's stream over the last sfg_run_advance call (0 if unavailable). */
double sfg_run_last_ms_per_step(const sfg_run *run);
/* Algorithmic HBM bytes of one step kernel launch (operator + coefficients +
 * state traffic of the active shots). */
double sfg_run_step_bytes(const sfg_run *run, int active_shots);
void sfg_run_destroy(sfg_run *run);

/* ---- several runs in one launch per step ---------------------------------- */

/* Groups runs on different meshes / media / shot counts (e.g. many models of
 * uneven size and step count) so that they step together: each step, every
 * batched run large enough to fill the GPU alone takes its own step launch,
 * the smaller batched runs (32 or 64 shots per lane group) share one launch,
 * and the single-shot runs share one launch per layout (kernel variant 9 in a
 * launch of its own, with the largest halo's LDS); all of them read one device
 * step counter and the launches of guard_every steps are captured in one
 * hipGraph.  The runs must share precision, max_steps (use per-shot nsteps for
 * uneven lengths), block size and guard interval and not have stepped yet.
 * While grouped, a run is advanced only through the group; its seismograms
 * and snapshots are read as usual.  No counterpart in the reference, which
 * runs one model per process. */
typedef struct sfg_multi sfg_multi;
int sfg_multi_create(sfg_multi **multi, sfg_run *const *runs, int nruns);
/* Executes up to nsteps further steps of every run; returns the step count so far. */
int sfg_multi_advance(sfg_multi *multi, int nsteps);
double sfg_multi_last_ms_per_step(const sfg_multi *multi);
/* Persistent schedule: blocks_per_cu > 0 replaces the per-step launches by one
 * cooperative launch per guard interval, in which blocks_per_cu x (CU count)
 * resident blocks take each step's tiles from a queue and meet at a grid-wide
 * barrier between steps; 0 returns to the graph-replayed per-step launches.
 * Results are identical bit for bit.  Needs every run in one shared launch
 * (single-shot runs in the per-row or sliced layouts, or only batched runs
 * too small to fill the GPU alone); fails if that many blocks cannot be
 * resident at once.  No counterpart in the reference. */
int sfg_multi_set_persistent(sfg_multi *multi, int blocks_per_cu);
/* Algorithmic bytes of the launch of step `step` (shots still active then). */
double sfg_multi_step_bytes(const sfg_multi *multi, int step);
void sfg_multi_destroy(sfg_multi *multi);

/* ---- time-step estimate --------------------------------------------------- */

/* Largest eigenvalue of M_L^{-1} K, K the elastic operator of the loop with the
 * M-PML terms removed (c11 S1 + c44 S2 | c13 S3 + c44 S4 on u, likewise on w)
 * and clamped rows dropped, by FP64 power iteration on the GPU (stops when the
 * estimate changes by <= tol relatively for 5 iterations, or at max_iter).
 * dt_max = 2/sqrt(lambda_max) is the stability limit of the reference's
 * Newmark update (alpha = 1, delta = 1.5 is stable for omega*dt <= 2).
 * No counterpart in the reference: dt is a user input there (seisfem/seisfem.c:144).
 * Returns the iteration count or an error code.  Power iteration approaches
 * lambda_max from below, so apply a safety factor to dt_max. */
int sfg_estimate_dt(const sfg_operator *op, const sfg_medium *md, int max_iter, double tol,
                    double *lambda_max, double *dt_max);

/* Copy bandwidth of the device: `reps` timed launches of a plain copy kernel
 * (stream_copy_kernel) over two buffers of `bytes` each (a positive multiple
 * of 16).  *gbs = 2 * bytes / time per copy.  The yardstick for the HBM
 * roofline and for scaling a profiler's byte counters; no counterpart in the
 * reference. */
int sfg_stream_copy(size_t bytes, int reps, double *gbs, double *ms_per_copy);

#ifdef __cplusplus
}
#endif
#endif
