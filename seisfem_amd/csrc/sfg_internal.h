// Internal declarations shared by the HIP kernels and the C-ABI layer.
#pragma once
#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>

namespace sfg {

// Field order of the 26-array state of one split-field M-PML shot
// (reference time_evolution/elastic_wave.c:78-87).
enum Field : int {
    U1 = 0, U2, U3, W1, W2, W3,         // split displacements
    U1T, U2T, U3T, W1T, W2T, W3T,       // split velocities
    U1TT, U2TT, U3TT, W1TT, W2TT, W3TT, // split accelerations (step n)
    LX1, LX2, LX3, LX4, LY1, LY2, LY3, LY4, // M-PML memory fields
    NFIELD
};

// Per-node coefficient order (c11,c13,c33,c44 from model_elastic_parameter.c:7,
// M-PML profiles from abc_mpml.c:4, then the lumped mass term).
// MROW scales the M-products of a row: the consistent mass samples the density
// at the row node (assemble/mass_sparse_q4.c:252), so a medium whose density
// differs from the one the operator was assembled with is the operator's M with
// each row multiplied by rho_medium / rho_operator.  Used with per-shot media.
enum Coef : int {
    C11 = 0, C13, C33, C44, DX, DY, DXX, DYY, DXXP, DYYP, MASS, MROW, NCOEF
};

// FP32 forms keep p = v + (0.5 - alpha) dt a in the acceleration slots: with
// the reference's alpha = 1, delta = 1.5 (so 1 - delta = 0.5 - alpha) the
// Newmark update becomes d' = d + p dt + alpha a' dt^2, p' = p + a' dt and
// v' = p + delta a' dt, so the acceleration is not kept and the velocity only
// where an M-PML row gathers it.
constexpr double NEWMARK_ALPHA = 1.0, NEWMARK_DELTA = 1.5;  // elastic_wave.c:97
static_assert(1.0 - NEWMARK_DELTA == 0.5 - NEWMARK_ALPHA, "p-form needs delta = alpha + 0.5");

// Device-side control block: the step counter lets one captured graph replay
// any chunk of steps.  A step kernel launched with slot p reads its step t from
// step[p] and one thread of it stores t + 1 into step[p ^ 1], which the next
// launch (slot p ^ 1) reads; launches on one stream run one after the other,
// so no block of a launch can see its own store.  Executed steps =
// max(step[0], step[1]).
struct Ctrl {
    int step[2];
    int err;            // bit 0: zero lumped mass met
    int pad;
};

__device__ __forceinline__ int executed_steps(const Ctrl *c)
{
    const int a = __hip_atomic_load(&c->step[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const int b = __hip_atomic_load(&c->step[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return a > b ? a : b;
}

// Where value (node i, field f, shot s) of a state array lives.  Field-major,
// [field][node][shot], for lane groups of 1 or 4 shots (a wave reads one field
// of 64 or 16 consecutive rows); node-major, [node][shot group][field][shot in
// group], for groups of 16, 32 or 64 shots, so that a wave's gather of one
// neighbour reads all that node's fields from one contiguous block instead of
// one short piece from each of 20 arrays tens of MB apart.
struct StateLayout {
    int gs;         // shots per group when node-major, 0 = field-major
    int ng;         // shot groups (node-major)
    int nf;         // fields per node in this array (26 state, 6 Kahan terms)
    int S;          // padded shot count
    size_t F;       // n * S: field stride of the field-major form
};

__host__ __device__ inline size_t state_index(const StateLayout &L, int i, int f, int s)
{
    if (L.gs) return (((size_t)i * L.ng + s / L.gs) * L.nf + f) * L.gs + s % L.gs;
    return (size_t)f * L.F + (size_t)i * L.S + s;
}

// Lane groups of SP shots use the node-major layout from 16 shots up.
constexpr bool node_major(int sp) { return sp >= 16; }

template <class T>
struct StepArgs {
    int n;                  // nodes (rows)
    int S;                  // shot stride of the state (padded shot count)
    int rows_per_block;
    int nsteps_max;
    int xcd_remap;          // 1: contiguous block ranges per XCD
    int xcd_stripe;         // Q > 0: XCD x takes tiles x*Q .. x*Q+Q-1 of every 8*Q (overrides xcd_remap)
    int sp, cps, nbx;       // lane grouping, per-shot coefficients, row blocks (multi-run dispatch)
    const int *rp, *ci;     // union CSR pattern
    const int *order;       // row taken by logical row i (a permutation), or null = i
    // LDS-staged gathers (node-major layouts): block b of rows stages the 20
    // gathered fields of its halo nodes hnode[hptr[b] .. hptr[b+1]) into LDS,
    // slot h at h * lds_stride; kslot[k] = slot of column ci[k] in the halo of
    // the block holding its row, rslot[i] = slot of row i itself.
    const int *hptr, *hnode, *kslot, *rslot;
    int lds_stride;
    // Sliced layout for one row per lane (kernel variant 7, SP = 1): wave w of
    // block b steps slice s = b * waves_per_block + w, whose 64 lanes take rows
    // sell_row[64 s + lane] (-1 = none).  The slice's nonzero slots are
    // sell_off[s] .. sell_off[s+1]-1; slot k of a lane holds its row's k-th
    // nonzero (column order), column sell_ci[64 k + lane] and the value of
    // matrix m (M, S1..S6) at sell_val[(7 k + m) * 64 + lane]; the slots past a
    // row's length hold column = the row itself and value 0 (a sum plus an
    // exact zero is unchanged).  Rows are sorted by length within windows of
    // the row order so that a slice's rows are of about the same length.
    const int *sell_row, *sell_off, *sell_ci;
    const T *sell_val;
    // LDS-staged sliced layout (kernel variant 8): block b stages the 20
    // gathered fields of its halo nodes hnode[hptr[b] .. hptr[b+1]) into LDS as
    // [field][slot] with field stride lds_stride; sell_ci then holds LDS slots
    // and sell_own[64 s + lane] the slot of the lane's own row.
    const int *sell_own;
    const T *A[7];          // M, S1..S6 on the union pattern (one array each)
    const T *Ap;            // the same values packed per nonzero: [nnz][8] = M, S1..S6, 0
    int variant;            // 0: separate value arrays; 1: packed values, gather loop unrolled x2
    const T *coef;          // NCOEF x n (x S when coefficients are per shot)
    const int *flag;        // Dirichlet flags (exact mode), may be null otherwise
    T *state;               // [2][NFIELD][n][S]
    T *comp;                // [2][6][n][S] Kahan compensation of U1..W3 (or null)
    // FP32 forms: U = u1 + u2 + u3 and W = w1 + w2 + w3 of each state buffer,
    // stored by the step that produces it, [2][2][n][S] in the state's layout
    // with 2 fields (null in exact mode).  pml[i] = 2 when one of the six M-PML
    // profiles of row i is nonzero for some shot, 1 when row i is a column of
    // such a row, else 0.  Rows below 2 gather only U, W and the 8 memory
    // fields of their neighbours; rows at 0 keep no velocity (null in exact mode).
    T *uw;
    const int *pml;
    // 1: rows of kind 0 (their split fields are gathered by no M-PML row) keep
    // only U, W (in uw), p_u = p1+p2+p3 and p_w = p4+p5+p6 (in the slots of p1
    // and p4) and the memory fields: their split updates are linear, so their
    // sums follow U' = U + p_u dt + alpha (a1+a2+a3) dt^2, p_u' = p_u + (a1+a2+a3) dt
    int agg;
    size_t fstride;         // n * S
    const int *active;      // active[t] = shots that execute step t (shots sorted)
    // sources grouped by node: entries src_ptr[i]..src_ptr[i+1]-1 of node i hold
    // (shot, u weight, w weight); a point force has weights sin/cos(angle)
    const int *src_ptr, *src_shot;
    const double *src_wt;   // [entries][2]
    const double *wavelet;  // Ricker value per step (time (t+3)*dt)
    double src_u, src_w;    // force direction factors sin/cos(angle)
    T dt, ca, alpha, cd, delta; // Newmark: ca = 0.5 - alpha, cd = 1 - delta
    const int *rec_ptr, *rec_idx; // node -> receiver slots (CSR over nodes)
    int rec_num;
    double *seis_u, *seis_w; // [S][nsteps_max][rec_num]
    Ctrl *ctrl;
};

// Launchers (defined in sfg_kernels.hip).  precision: 0 exact FP64, 1 FP32, 2 FP32+Kahan,
// 3 FP32+Kahan with the compensation terms stored as bfloat16.
// slot: parity of the step being launched (see Ctrl).
hipError_t launch_step(int precision, int sp, bool coef_per_shot, const void *args, int nblk_x,
                       int nblk_y, int block, int slot, size_t lds_bytes, hipStream_t st);
hipError_t allow_step_lds(int precision, int sp, bool coef_per_shot, size_t lds_bytes);
// single_shot_only: every run has one shot per lane group (a kernel with only those forms)
hipError_t launch_multi_step(int precision, const void *args_dev, const int *tile_start_dev, int nruns, int ntiles,
                             int block, int xcd_remap, int slot, bool single_shot_only, hipStream_t st);
// Persistent schedule of a multi-run group (cooperative launch of `blocks`
// resident blocks running nsteps steps; queue: device int[16], zero)
hipError_t launch_multi_persistent(int precision, const void *args_dev, const int *tile_start_dev, int nruns,
                                   int ntiles, int block, int blocks, int slot, int nsteps, int *queue,
                                   bool single_shot_only, hipStream_t st);
hipError_t multi_persistent_blocks_per_cu(int precision, bool single_shot_only, int block, int *per_cu);
// multi-run launch of single-shot runs in the LDS-staged sliced layout (kernel variant 8)
hipError_t launch_multi_step_lds(int precision, const void *args_dev, const int *tile_start_dev, int nruns,
                                 int ntiles, int block, int xcd_remap, int slot, size_t lds_bytes, hipStream_t st);
hipError_t allow_multi_step_lds(int precision, size_t lds_bytes);
// uw (FP32 forms, may be null): U and W are read from it instead of being summed
// from the split fields.  shot_steps[s] (s < L.S): steps shot s executes; a shot
// that has finished is read from the buffer of its own last step
hipError_t launch_guard(int precision, const void *state, const void *uw, int n, const StateLayout &L,
                        const Ctrl *ctrl, const int *shot_steps, double *energy, int nslots, int every,
                        hipStream_t st);
hipError_t launch_gather_uw(int precision, const void *state, const void *uw, int n, const StateLayout &L,
                            int shot, int parity, double *U, double *W, hipStream_t st);
hipError_t launch_spmv64(int n, const int *rp, const int *cj, const double *a, const double *x,
                         double *y, hipStream_t st);
// M-PML geometry: damping-layer edges, thicknesses, d0 and cross factors.
struct MpmlGeo {
    double x0, x1, y0, y1, tx, ty, d0x, d0y, pxy, pyx;
    int use_xmin, use_xmax, use_ymin, use_ymax;
};
hipError_t launch_mpml(int n, const double *x, const double *y, const MpmlGeo &g, double *out6,
                       hipStream_t st);
hipError_t launch_to_f32(const double *src, float *dst, size_t n, hipStream_t st);
// FP32 medium from the FP64 one: coefficients rounded once, MASS -> 1/mass_lump
// (formed in FP64) or 0 on clamped / zero-mass nodes
hipError_t launch_medium_f32(int n, const double *coef64, const int *flag, float *coef32, hipStream_t st);
hipError_t launch_fill_zero(void *p, size_t bytes, hipStream_t st);
// Values of the sliced layout: out[(7 k + m) * 64 + lane] = vals[m * nnz + src[64 k + lane]] for
// src >= 0, else 0, over nslots = 64 * (slots); elem = 4 (float) or 8 (double)
hipError_t launch_sell_values(const void *vals, size_t nnz, const int *src, size_t nslots, int elem, void *out,
                              hipStream_t st);
// pml[i] = 2 when coefficient DX..DYYP of row i is nonzero for one of the ns
// shots (coef is [NCOEF][n][ns]; ns = 1: one medium for every shot), 1 when
// row i is a column of such a row in the pattern rp/ci, else 0
hipError_t launch_pml_rows(int precision, const void *coef, int n, int ns, const int *rp, const int *ci, int *pml,
                           hipStream_t st);
// dst[c][i][s] = srcs[min(s, nsrc-1)][c*n + i] for c < ncoef, i < n, s < S; elem = 4 or 8 bytes
hipError_t launch_pack_per_shot(const void *const *srcs_dev, int nsrc, int ncoef, int n, int S, int elem,
                                void *dst, hipStream_t st);
hipError_t launch_elastic_apply(int n, const int *rp, const int *ci, const double *A, size_t nnz,
                                const double *coef, const int *flag, const double *x, double *y,
                                double *sumsq, hipStream_t st);
hipError_t launch_scale(size_t n, double *x, const double *sumsq, hipStream_t st);
hipError_t launch_stream_copy(const void *src, void *dst, size_t n16, hipStream_t st);

} // namespace sfg
