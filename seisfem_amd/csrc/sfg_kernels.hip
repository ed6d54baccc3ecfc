// HIP kernels for the explicit elastic M-PML masslump time loop (gfx950).
//
// One launch = one time step of every active shot.  Each lane owns one
// (row, shot) pair: it walks the row of the union CSR pattern once, forms the
// 32 distinct products of the reference's 44 csr_matvec calls (20 with the
// consistent mass M, two with each of S1..S6), then does the row-local RHS,
// lumped-mass divide + Dirichlet clamp, Newmark update and receiver store,
// writing the other half of a double-buffered state.
//
// Lane mapping: SP lanes of a wavefront share one row and hold consecutive
// shots (state layout [field][node][shot], shots innermost), so a wave covers
// 64/SP rows; shot groups of SP are the grid's y dimension.  SP = 1 is the
// single-shot path (one row per lane); SP = 64 gives one row per wave with
// the matrix row read once as scalars and broadcast to 64 shots; SP = 16 or
// 32 trade that broadcast for a smaller per-node footprint in L2.
//
// Precision modes:
//   0  FP64, bit-identical to the reference: every sum is sequential in
//      column order, no FMA (the file is built with -ffp-contract=off), the
//      RHS / divide / update keep the reference's operand order
//      (time_evolution/elastic_wave.c:582-592, 648-658, 681-729, 754-778);
//   1  FP32 storage and arithmetic (FMA in the dot products);
//   2  FP32 with a Kahan-compensated displacement update.
#include <hip/hip_cooperative_groups.h>

#include "sfg_internal.h"

namespace sfg {

template <bool EXACT, class T>
__device__ __forceinline__ T mac(T acc, T a, T x)
{
    if constexpr (EXACT) return acc + a * x;
    else return __builtin_fmaf(a, x, acc);
}

// Loads and stores of data no other lane of this step touches again (a row's
// own accelerations, the Kahan terms, every output) may carry the streaming
// (nontemporal) hint, so the lines they occupy do not push the neighbours'
// fields, which nine rows gather, out of L2.
template <bool NT, class T>
__device__ __forceinline__ T ld_once(const T *p)
{
    if constexpr (NT) return __builtin_nontemporal_load(p);
    else return *p;
}

template <bool NT, class T>
__device__ __forceinline__ void st_once(T *p, T v)
{
    if constexpr (NT) __builtin_nontemporal_store(v, p);
    else *p = v;
}

// bfloat16 halves of a 32-bit word (precision 3 keeps the Kahan terms of two
// displacement fields per word): the low half is field 2k, the high half 2k+1.
__device__ __forceinline__ float bf16_lo(uint32_t w) { return __uint_as_float(w << 16); }
__device__ __forceinline__ float bf16_hi(uint32_t w) { return __uint_as_float(w & 0xffff0000u); }
// round to nearest even (the terms are finite: a Kahan term is the rounding
// error of an FP32 sum)
__device__ __forceinline__ uint32_t bf16_bits(float f)
{
    const uint32_t u = __float_as_uint(f);
    return (u + 0x7fffu + ((u >> 16) & 1u)) >> 16;
}

// Read-only arrays of a step (pattern, values, coefficients, schedules) are
// read through the constant address space: they do not change during a
// launch, which lets wave-uniform reads (one row per wave) go through the
// scalar unit instead of the vector-memory path the gathers saturate.
template <class T>
__device__ __forceinline__ const __attribute__((address_space(4))) T *ro(const T *p)
{
    return (const __attribute__((address_space(4))) T *)p;
}

// 16-byte vector of T (staging copies into LDS)
template <class T> struct vec16;
template <> struct vec16<float> { using type = float4; };
template <> struct vec16<double> { using type = double2; };

// Position of state field f among the 20 gathered fields (U1..W3, U1t..W3t,
// Lx1..Ly4; the accelerations are not gathered).
__device__ __forceinline__ constexpr int gathered_slot(int f) { return f < U1TT ? f : f - 6; }

// Workgroups are dispatched round-robin over the 8 XCDs; this map makes each
// XCD walk one contiguous range of logical blocks (shot chunk, row block), so
// neighbouring rows share that XCD's L2.  A bijection on [0, nb) for any nb.
__device__ __forceinline__ unsigned int xcd_contiguous(unsigned int bid, unsigned int nb)
{
    const unsigned int q = nb / 8, r = nb % 8;
    const unsigned int x = bid % 8, l = bid / 8;
    return x * q + (x < r ? x : r) + l;
}

// Stripes: tiles are taken in runs of 8*Q; within a run XCD x (workgroup ids
// x, x+8, ...) steps tiles x*Q .. x*Q+Q-1.  All XCDs then sweep the rows
// together, each keeping one stripe of consecutive row blocks, so a row's
// neighbours on the adjacent grid lines were fetched into the same XCD's L2.
// A bijection on [0, nb): the ids past the last whole run map to themselves.
__device__ __forceinline__ unsigned int xcd_striped(unsigned int bid, unsigned int nb, unsigned int q)
{
    const unsigned int span = 8u * q;
    if (bid >= (nb / span) * span) return bid;
    const unsigned int x = bid % 8, k = bid / 8;
    return (k / q) * span + x * q + k % q;
}

// Step number of a launch with the given slot; the first thread of the grid
// hands t + 1 to the next launch (the other slot).  One store per launch: a
// per-block completion count on one address serialises tens of thousands of
// atomics per step at one L2 channel.
__device__ __forceinline__ int begin_step(Ctrl *ctrl, int slot)
{
    const int t = __hip_atomic_load(&ctrl->step[slot], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0)
        __hip_atomic_store(&ctrl->step[slot ^ 1], t + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return t;
}

// One tile (logical block bx of rows, shot chunk by) of step t.
// Variant V: 0 separate value arrays; 1 values packed per nonzero, gather loop
// unrolled x2; 2 as 1 with streaming hints on data used once; 3 as 2, and the
// row's own 20 gathered fields are kept from the diagonal nonzero instead of
// being loaded again, and its accelerations and Kahan terms are loaded before
// the gathers; 4 and 5 as 3 with 3 and 4 nonzeros' gathers in flight per trip;
// 6 (node-major layouts only) stages the 20 gathered fields of the block's halo
// nodes into LDS once, and every row of the block gathers from there: each
// node is read from L2 once per block instead of once per row that needs it.
// FP32 forms: the products of U1..W3 and U1t..W3t with M, and of U and W with
// S5 and S6, enter a row's equations only multiplied by its M-PML profiles.
// A row whose six profiles are all zero (outside the damping layers) leaves
// them at 0 and gathers only U, W (stored by the producing step) and Lx1..Ly4:
// 10 values per neighbour instead of 20.  Each skipped term is then an exact
// zero in the same expressions, so the results equal those of the full walk.
// The FP32 forms also update displacements with the p-form of the Newmark
// step (NEWMARK_ALPHA): a row far from the damping layers reads and writes
// displacements, p and memory fields only; and with a.agg a row that no M-PML
// row gathers from keeps its split fields as sums (StepArgs::agg).
template <class T, int PREC, int SP, bool CPS, int V>
__device__ __forceinline__ void step_tile(const StepArgs<T> &a, int t, int bx, int by)
{
    constexpr bool EXACT = (PREC == 0);
    constexpr bool KAHAN = (PREC >= 2);
    constexpr bool C16 = (PREC == 3);           // Kahan terms stored as bfloat16 pairs
    constexpr bool LDSG = V == 6;
    // 7: the sliced layout of one row per lane (StepArgs::sell_*): nonzero k of
    // the 64 rows of a wave is one coalesced load per matrix; 8: the same with
    // the 20 gathered fields of the block's halo nodes staged in LDS once
    // (field-major state), the slices' columns holding LDS slots
    constexpr bool LDS1 = V == 8;
    constexpr bool SELL = V == 7 || V == 8;
    constexpr bool LDSX = LDSG || LDS1;          // neighbours' fields come from LDS
    static_assert(!SELL || SP == 1, "the sliced layout holds one row per lane");
    constexpr bool PACKED = V >= 1 && !SELL;
    // sliced walks: slots whose loads are in flight ahead of their use
    constexpr int SELL_AHEAD = EXACT ? 3 : 4;
    constexpr bool NT = V >= 2;
    constexpr bool PRE = V >= 3;                 // own-row accelerations / Kahan terms first
    constexpr bool CAP = V >= 3 && !LDSX;        // own-row fields from the diagonal gather
    constexpr int UN = V == 5 ? 4 : V == 4 ? 3 : 2;  // nonzeros per trip of the gather loop
    // the full 20-field walk of M-PML rows in variant 5 keeps 2 per trip: its
    // registers, not the short walk's, set the kernel's occupancy
    constexpr int UNP = (V == 5 && !EXACT) ? 2 : UN;
    constexpr bool WUW = !EXACT;                 // store U and W of the new state
    constexpr bool SKIP = !EXACT && !LDSX;       // interior rows gather 10 values per neighbour
    constexpr int RPW = (SP >= 64) ? 1 : 64 / SP;
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int wpb = blockDim.x >> 6;
    int row = (bx * wpb + wave) * RPW + (SP >= 64 ? 0 : lane / SP);
    if constexpr (SP >= 64) row = __builtin_amdgcn_readfirstlane(row);
    const int slice = bx * wpb + wave;           // sliced layout: the wave's slice of 64 rows
    if constexpr (SELL) row = ro(a.sell_row)[(size_t)slice * 64 + lane];
    constexpr int GS = SP >= 64 ? 64 : SP;      // shots of one lane group
    const int ls = SP >= 64 ? lane : lane % SP;  // this lane's shot within the group
    const int s = by * GS + ls;
    const int nact = (t < a.nsteps_max) ? ro(a.active)[t] : 0;
    extern __shared__ __align__(16) unsigned char sfg_lds_raw[];
    T *lds = reinterpret_cast<T *>(sfg_lds_raw);
    if constexpr (LDSG) {
        static_assert(node_major(SP), "LDS staging needs the node-major layout");
        // a finished shot group skips the whole block (uniform, before the barrier)
        if (by * GS >= nact) return;
        using V16 = typename vec16<T>::type;
        constexpr int VE = 16 / sizeof(T);
        constexpr int GV = GS / VE;              // 16-byte vectors per field of one group
        const T *inb = a.state + (size_t)(t & 1) * NFIELD * a.fstride;
        const size_t NGr = (size_t)(a.S / GS);
        const int h0 = a.hptr[bx], nh = a.hptr[bx + 1] - h0;
        const int total = nh * 20 * GV;
        // BATCH independent 16-byte loads per thread are in flight before their
        // LDS stores, so the copy costs a few memory round trips, not one per element
        constexpr int BATCH = 8;
        const int nt = blockDim.x;
        for (int e0 = threadIdx.x; e0 < total; e0 += BATCH * nt) {
            V16 buf[BATCH];
            int dst[BATCH];
            // no branch per element (it would send the batch to scratch): an index
            // past the end is clamped to the last element, whose copy is then
            // stored twice with the same value
#pragma unroll
            for (int b = 0; b < BATCH; ++b) {
                const int e = min(e0 + b * nt, total - 1);
                const int v = e % GV, f20 = (e / GV) % 20, h = e / (GV * 20);
                const int f = f20 < 12 ? f20 : f20 + 6;
                const size_t g = ((size_t)a.hnode[h0 + h] * NGr + by) * (NFIELD * GS) + (size_t)f * GS + v * VE;
                buf[b] = *reinterpret_cast<const V16 *>(inb + g);
                dst[b] = h * a.lds_stride + f20 * GS + v * VE;
            }
#pragma unroll
            for (int b = 0; b < BATCH; ++b) *reinterpret_cast<V16 *>(lds + dst[b]) = buf[b];
        }
        __syncthreads();
    }
    if constexpr (LDS1) {
        // a finished shot skips the whole block (uniform, before the barrier)
        if (by >= nact) return;
        // each thread copies the 20 gathered fields of halo nodes h, h + blockDim.x, ...
        // (20 independent loads in flight; consecutive threads read consecutive halo
        // nodes, which are sorted) to LDS [field][slot]
        const T *inb = a.state + (size_t)(t & 1) * NFIELD * a.fstride;
        const int h0 = ro(a.hptr)[bx], nh = ro(a.hptr)[bx + 1] - h0;
        for (int h = threadIdx.x; h < nh; h += blockDim.x) {
            const size_t node = (size_t)a.hnode[h0 + h] * a.S + by;
            T buf[20];
#pragma unroll
            for (int f20 = 0; f20 < 20; ++f20) buf[f20] = inb[(size_t)(f20 < 12 ? f20 : f20 + 6) * a.fstride + node];
#pragma unroll
            for (int f20 = 0; f20 < 20; ++f20) lds[f20 * a.lds_stride + h] = buf[f20];
        }
        __syncthreads();
    }
    if constexpr (SELL) {
        if (row < 0) return;                     // a lane past the slices' last row
    } else if (row >= a.n) {
        return;
    }
    if (!SELL && a.order) {
        // rows in locality order: the waves in flight hold a compact patch of the mesh
        row = ro(a.order)[row];
        if constexpr (SP >= 64) row = __builtin_amdgcn_readfirstlane(row);
    }

    if (s < nact) {
        const size_t F = a.fstride;
        const int S = a.S;
        const T *in = a.state + (size_t)(t & 1) * NFIELD * F;
        T *out = a.state + (size_t)((t + 1) & 1) * NFIELD * F;
        // element (node, field f) of this lane's shot is at node_base(node) + f * FS
        constexpr bool NM = node_major(SP);
        const size_t FS = NM ? (size_t)GS : F;
        auto node_base = [&](int node) __attribute__((always_inline)) -> size_t {
            if constexpr (NM) return ((size_t)node * (S / GS) + by) * (NFIELD * GS) + ls;
            else return (size_t)node * S + s;
        };
        // U and W: 2 fields in the same layout as the state
        auto uw_base = [&](int node) __attribute__((always_inline)) -> size_t {
            if constexpr (NM) return ((size_t)node * (S / GS) + by) * (2 * GS) + ls;
            else return (size_t)node * S + s;
        };
        const T *uwin = WUW ? a.uw + (size_t)(t & 1) * 2 * F : nullptr;
        T *uwout = WUW ? a.uw + (size_t)((t + 1) & 1) * 2 * F : nullptr;
        int pk = 2;                              // row kind: see StepArgs::pml
        if constexpr (!EXACT) pk = ro(a.pml)[row];
        const bool interior = SKIP && pk < 2;    // short gather
        const bool keepv = EXACT || pk != 0;     // velocity stored
        // row of kind 0 that keeps only the sums U, W, p_u, p_w (see StepArgs::agg)
        const bool agg = SKIP && a.agg && pk == 0;
        // Kahan terms: 6 fields in the same layout as the state, or (C16) 3
        // words of two bfloat16 terms each
        constexpr int NC = C16 ? 3 : 6;
        const size_t C6 = (size_t)NC * F;
        const T *cin = KAHAN ? a.comp + (size_t)(t & 1) * C6 : nullptr;
        T *cout = KAHAN ? a.comp + (size_t)((t + 1) & 1) * C6 : nullptr;
        const uint32_t *win = reinterpret_cast<const uint32_t *>(cin);
        uint32_t *wout = reinterpret_cast<uint32_t *>(cout);
        size_t co;
        if constexpr (NM) co = ((size_t)row * (S / GS) + by) * (NC * GS) + ls;
        else co = node_base(row);
        // own-row data nobody else gathers: with CAP it is requested before the
        // gathers so that its latency overlaps theirs
        // (summed rows: p_u, p_w in the slots of p1 and p4, their Kahan terms in
        // word 0 as a bfloat16 pair, or in terms 0 and 1; kept here at 0 and 3)
        T pre_acc[6], pre_c[6];
        auto load_terms = [&]() __attribute__((always_inline)) {
            if (agg) {
                if constexpr (C16) {
                    const uint32_t w = ld_once<NT>(win + co);
                    pre_c[0] = bf16_lo(w);
                    pre_c[3] = bf16_hi(w);
                } else if constexpr (KAHAN) {
                    pre_c[0] = ld_once<NT>(cin + co);
                    pre_c[3] = ld_once<NT>(cin + FS + co);
                }
            } else {
                if constexpr (C16) {
#pragma unroll
                    for (int k = 0; k < 3; ++k) {
                        const uint32_t w = ld_once<NT>(win + k * FS + co);
                        pre_c[2 * k] = bf16_lo(w);
                        pre_c[2 * k + 1] = bf16_hi(w);
                    }
                } else if constexpr (KAHAN) {
#pragma unroll
                    for (int q = 0; q < 6; ++q) pre_c[q] = ld_once<NT>(cin + q * FS + co);
                }
            }
        };
        if constexpr (PRE) {
            const size_t o0 = node_base(row);
            if (agg) {
                pre_acc[0] = ld_once<NT>(in + U1TT * FS + o0);
                pre_acc[3] = ld_once<NT>(in + W1TT * FS + o0);
            } else {
#pragma unroll
                for (int q = 0; q < 6; ++q) pre_acc[q] = ld_once<NT>(in + (U1TT + q) * FS + o0);
            }
            load_terms();
        }

        // ---- the 32 distinct products over one walk of the row ------------
        T s1u = 0, s1w = 0, s2u = 0, s2w = 0, s3u = 0, s3w = 0;
        T s4u = 0, s4w = 0, s5u = 0, s5w = 0, s6u = 0, s6w = 0;
        T mu1t = 0, mu1 = 0, mlx1 = 0, mlx2 = 0, mu2t = 0, mu2 = 0, mu3t = 0, mu3 = 0, mlx3 = 0, mlx4 = 0;
        T mw1t = 0, mw1 = 0, mly1 = 0, mly2 = 0, mw2t = 0, mw2 = 0, mw3t = 0, mw3 = 0, mly3 = 0, mly4 = 0;
        // nonzeros k0 .. k1-1 of the row, or slots k0 .. k1-1 of its slice
        const int k0 = SELL ? ro(a.sell_off)[slice] : ro(a.rp)[row];
        const int k1 = SELL ? ro(a.sell_off)[slice + 1] : ro(a.rp)[row + 1];
        // sliced layout: slots of a lane past its own row's length are padding
        // (column = the row itself, or its LDS slot; values 0); they are handed
        // over as such without loading them, so a cache line of the slot arrays
        // that only padding lanes would touch is not fetched (a slice's rows are
        // sorted by length, so those lanes are contiguous)
        const int rlen = SELL ? ro(a.rp)[row + 1] - ro(a.rp)[row] : 0;
        const int pad_j = LDS1 ? ro(a.sell_own)[(size_t)slice * 64 + lane] : row;
        const size_t o = node_base(row);
        T cur[NFIELD];
        T own_u = 0, own_w = 0;                  // summed rows: the row's own U, W
        bool got = false;
        // one slot of the sliced layout: column (or LDS slot) and the 7 values
        struct Slot {
            int j;
            T v[7];
        };
        // nonzero k; the sliced walk hands over its slot already loaded
        auto nonzero = [&](int k, auto... pre) __attribute__((always_inline)) {
            int j;
            T am, a1, a2, a3, a4, a5, a6;
            if constexpr (sizeof...(pre) > 0) {
                const Slot &q = (pre, ...);
                j = q.j;
                am = q.v[0]; a1 = q.v[1]; a2 = q.v[2]; a3 = q.v[3]; a4 = q.v[4]; a5 = q.v[5]; a6 = q.v[6];
            } else if constexpr (PACKED) {
                j = LDSG ? 0 : ro(a.ci)[k];
                const auto v = ro(a.Ap) + (size_t)k * 8;
                am = v[0]; a1 = v[1]; a2 = v[2]; a3 = v[3]; a4 = v[4]; a5 = v[5]; a6 = v[6];
            } else {
                j = LDSG ? 0 : ro(a.ci)[k];
                am = ro(a.A[0])[k];
                a1 = ro(a.A[1])[k]; a2 = ro(a.A[2])[k]; a3 = ro(a.A[3])[k];
                a4 = ro(a.A[4])[k]; a5 = ro(a.A[5])[k]; a6 = ro(a.A[6])[k];
            }
            // field f of neighbour j: global state, or the block's LDS copy
            const T *x;
            size_t xs;
            if constexpr (LDSG) {
                x = lds + ro(a.kslot)[k] * a.lds_stride + ls;
                xs = GS;
            } else if constexpr (LDS1) {
                x = lds + j;                     // j is the neighbour's LDS slot
                xs = a.lds_stride;
            } else {
                x = in + node_base(j);
                xs = FS;
            }
            auto fld = [&](int f) __attribute__((always_inline)) {
                return LDSX ? x[gathered_slot(f) * xs] : x[f * xs];
            };
            const T u1 = fld(U1), u2 = fld(U2), u3 = fld(U3);
            const T w1 = fld(W1), w2 = fld(W2), w3 = fld(W3);
            const T u1t = fld(U1T), u2t = fld(U2T), u3t = fld(U3T);
            const T w1t = fld(W1T), w2t = fld(W2T), w3t = fld(W3T);
            const T lx1 = fld(LX1), lx2 = fld(LX2), lx3 = fld(LX3), lx4 = fld(LX4);
            const T ly1 = fld(LY1), ly2 = fld(LY2), ly3 = fld(LY3), ly4 = fld(LY4);
            if constexpr (CAP) {
                if (j == row) {
                    cur[U1] = u1; cur[U2] = u2; cur[U3] = u3; cur[W1] = w1; cur[W2] = w2; cur[W3] = w3;
                    cur[U1T] = u1t; cur[U2T] = u2t; cur[U3T] = u3t;
                    cur[W1T] = w1t; cur[W2T] = w2t; cur[W3T] = w3t;
                    cur[LX1] = lx1; cur[LX2] = lx2; cur[LX3] = lx3; cur[LX4] = lx4;
                    cur[LY1] = ly1; cur[LY2] = ly2; cur[LY3] = ly3; cur[LY4] = ly4;
                    got = true;
                }
            }
            // U_now / W_now as recombined at the end of the previous step (:775-776)
            const T U = u1 + u2 + u3;
            const T W = w1 + w2 + w3;
            s1u = mac<EXACT>(s1u, a1, U); s1w = mac<EXACT>(s1w, a1, W);
            s2u = mac<EXACT>(s2u, a2, U); s2w = mac<EXACT>(s2w, a2, W);
            s3u = mac<EXACT>(s3u, a3, U); s3w = mac<EXACT>(s3w, a3, W);
            s4u = mac<EXACT>(s4u, a4, U); s4w = mac<EXACT>(s4w, a4, W);
            s5u = mac<EXACT>(s5u, a5, U); s5w = mac<EXACT>(s5w, a5, W);
            s6u = mac<EXACT>(s6u, a6, U); s6w = mac<EXACT>(s6w, a6, W);
            mu1t = mac<EXACT>(mu1t, am, u1t); mu1 = mac<EXACT>(mu1, am, u1);
            mlx1 = mac<EXACT>(mlx1, am, lx1); mlx2 = mac<EXACT>(mlx2, am, lx2);
            mu2t = mac<EXACT>(mu2t, am, u2t); mu2 = mac<EXACT>(mu2, am, u2);
            mu3t = mac<EXACT>(mu3t, am, u3t); mu3 = mac<EXACT>(mu3, am, u3);
            mlx3 = mac<EXACT>(mlx3, am, lx3); mlx4 = mac<EXACT>(mlx4, am, lx4);
            mw1t = mac<EXACT>(mw1t, am, w1t); mw1 = mac<EXACT>(mw1, am, w1);
            mly1 = mac<EXACT>(mly1, am, ly1); mly2 = mac<EXACT>(mly2, am, ly2);
            mw2t = mac<EXACT>(mw2t, am, w2t); mw2 = mac<EXACT>(mw2, am, w2);
            mw3t = mac<EXACT>(mw3t, am, w3t); mw3 = mac<EXACT>(mw3, am, w3);
            mly3 = mac<EXACT>(mly3, am, ly3); mly4 = mac<EXACT>(mly4, am, ly4);
        };
        // interior row: U, W and the memory fields of neighbour j
        auto nonzero_in = [&](int k, auto... pre) __attribute__((always_inline)) {
            int j;
            T am, a1, a2, a3, a4;
            if constexpr (sizeof...(pre) > 0) {
                const Slot &q = (pre, ...);
                j = q.j;
                am = q.v[0]; a1 = q.v[1]; a2 = q.v[2]; a3 = q.v[3]; a4 = q.v[4];
            } else if constexpr (PACKED) {
                j = ro(a.ci)[k];
                const auto v = ro(a.Ap) + (size_t)k * 8;
                am = v[0]; a1 = v[1]; a2 = v[2]; a3 = v[3]; a4 = v[4];
            } else {
                j = ro(a.ci)[k];
                am = ro(a.A[0])[k];
                a1 = ro(a.A[1])[k]; a2 = ro(a.A[2])[k]; a3 = ro(a.A[3])[k]; a4 = ro(a.A[4])[k];
            }
            const T *x = in + node_base(j);
            const T *xu = uwin + uw_base(j);
            const T U = xu[0], W = xu[FS];
            const T lx1 = x[LX1 * FS], lx2 = x[LX2 * FS], lx3 = x[LX3 * FS], lx4 = x[LX4 * FS];
            const T ly1 = x[LY1 * FS], ly2 = x[LY2 * FS], ly3 = x[LY3 * FS], ly4 = x[LY4 * FS];
            if constexpr (CAP) {
                if (j == row) {
                    cur[LX1] = lx1; cur[LX2] = lx2; cur[LX3] = lx3; cur[LX4] = lx4;
                    cur[LY1] = ly1; cur[LY2] = ly2; cur[LY3] = ly3; cur[LY4] = ly4;
                    own_u = U;
                    own_w = W;
                    got = true;
                }
            }
            s1u = mac<EXACT>(s1u, a1, U); s1w = mac<EXACT>(s1w, a1, W);
            s2u = mac<EXACT>(s2u, a2, U); s2w = mac<EXACT>(s2w, a2, W);
            s3u = mac<EXACT>(s3u, a3, U); s3w = mac<EXACT>(s3w, a3, W);
            s4u = mac<EXACT>(s4u, a4, U); s4w = mac<EXACT>(s4w, a4, W);
            mlx1 = mac<EXACT>(mlx1, am, lx1); mlx2 = mac<EXACT>(mlx2, am, lx2);
            mlx3 = mac<EXACT>(mlx3, am, lx3); mlx4 = mac<EXACT>(mlx4, am, lx4);
            mly1 = mac<EXACT>(mly1, am, ly1); mly2 = mac<EXACT>(mly2, am, ly2);
            mly3 = mac<EXACT>(mly3, am, ly3); mly4 = mac<EXACT>(mly4, am, ly4);
        };
        // sliced layout: slot k's column and its first NV values (the rest 0)
        auto load_slot = [&](int k, auto nv) __attribute__((always_inline)) {
            constexpr int NV = decltype(nv)::value;
            Slot q;
            if (k - k0 < rlen) {
                q.j = ld_once<true>(a.sell_ci + (size_t)k * 64 + lane);
                const T *v = a.sell_val + (size_t)k * (7 * 64) + lane;
#pragma unroll
                for (int m = 0; m < 7; ++m) q.v[m] = m < NV ? ld_once<true>(v + 64 * m) : T(0);
            } else {
                q.j = pad_j;
#pragma unroll
                for (int m = 0; m < 7; ++m) q.v[m] = T(0);
            }
            return q;
        };
        auto walk = [&](auto &&nz, auto per_trip, auto nv) __attribute__((always_inline)) {
            constexpr int U = decltype(per_trip)::value;
            if constexpr (SELL) {
                // the slots' loads run SELL_AHEAD slots ahead of their use (a ring of
                // registers), so a wave keeps several slots' loads in flight.  Every
                // lane walks to the slice's longest row (its padded slots add exact
                // zeros, loaded or not): stopping each lane at its own row measured
                // slower, the divergent ring costing more than the lines it skips.
                constexpr int PA = SELL_AHEAD;
                Slot ring[PA];
#pragma unroll
                for (int p = 0; p < PA; ++p)
                    if (k0 + p < k1) ring[p] = load_slot(k0 + p, nv);
                for (int kb = k0; kb < k1; kb += PA) {
#pragma unroll
                    for (int p = 0; p < PA; ++p) {
                        const int k = kb + p;
                        if (k < k1) {
                            const Slot q = ring[p];
                            if (k + PA < k1) ring[p] = load_slot(k + PA, nv);
                            nz(k, q);
                        }
                    }
                }
            } else if constexpr (PACKED) {
                // U nonzeros per trip: their gathers are in flight together; each
                // sum still runs in column order
                int k = k0;
                for (; k + U - 1 < k1; k += U) {
#pragma unroll
                    for (int u = 0; u < U; ++u) nz(k + u);
                }
                for (; k < k1; ++k) nz(k);
            } else {
                for (int k = k0; k < k1; ++k) nz(k);
            }
        };
        if (interior) {
            // the row's own displacements are not gathered here
            if constexpr (PRE) {
                if (!agg) {
#pragma unroll
                    for (int f = U1; f <= W3; ++f) cur[f] = ld_once<NT>(in + f * FS + o);
                }
            }
            walk(nonzero_in, std::integral_constant<int, UN>{}, std::integral_constant<int, 5>{});
        } else {
            walk(nonzero, std::integral_constant<int, UNP>{}, std::integral_constant<int, 7>{});
        }

        // ---- per-node coefficients --------------------------------------
        const size_t cs = CPS ? (size_t)a.n * S : (size_t)a.n;
        const size_t ci = CPS ? (size_t)row * S + s : (size_t)row;
        const auto cf = ro(a.coef);
        const T c11 = cf[C11 * cs + ci], c13 = cf[C13 * cs + ci];
        const T c33 = cf[C33 * cs + ci], c44 = cf[C44 * cs + ci];
        const T dx = cf[DX * cs + ci], dy = cf[DY * cs + ci];
        const T dxx = cf[DXX * cs + ci], dyy = cf[DYY * cs + ci];
        const T dxxp = cf[DXXP * cs + ci], dyyp = cf[DYYP * cs + ci];
        const T mass = cf[MASS * cs + ci];
        const T dt = a.dt;
        if constexpr (CPS) {
            // per-shot density: scale this row of the operator's M
            const T mr = cf[MROW * cs + ci];
            mu1t *= mr; mu1 *= mr; mlx1 *= mr; mlx2 *= mr; mu2t *= mr; mu2 *= mr;
            mu3t *= mr; mu3 *= mr; mlx3 *= mr; mlx4 *= mr;
            mw1t *= mr; mw1 *= mr; mly1 *= mr; mly2 *= mr; mw2t *= mr; mw2 *= mr;
            mw3t *= mr; mw3 *= mr; mly3 *= mr; mly4 *= mr;
        }

        // source term: the reference adds (i == source_node) * point_source * sin|cos(angle)
        // to every row, so a row without a source still adds (0 * ps) * factor
        const double ps = ro(a.wavelet)[t];
        double su = 0.0 * ps * a.src_u, sw = 0.0 * ps * a.src_w;
        for (int q = ro(a.src_ptr)[row]; q < ro(a.src_ptr)[row + 1]; ++q)
            if (ro(a.src_shot)[q] == s) {
                su = 1.0 * ps * ro(a.src_wt)[2 * q];
                sw = 1.0 * ps * ro(a.src_wt)[2 * q + 1];
            }
        const T fu = (T)su;
        const T fw = (T)sw;

        // ---- RHS of the 14 split equations (operand order of :585-591, :651-657)
        const T ru1 = -c11 * s1u - T(2.0) * dx * mu1t - dx * dx * mu1 + mlx1 + mlx2 + fu;
        const T ru2 = -c13 * s3w - c44 * s4w - dx * mu2t - dy * mu2t - dx * dy * mu2;
        const T ru3 = -c44 * s2u - T(2.0) * dy * mu3t - dy * dy * mu3 + mlx3 + mlx4;
        const T ru4 = -dt * c11 * dxx * s5u - dt * dx * mlx1 + mlx1;
        const T ru5 = -dt * c44 * dyyp * s5w - dt * dy * mlx2 + mlx2;
        const T ru6 = -dt * c13 * dxxp * s6w - dt * dx * mlx3 + mlx3;
        const T ru7 = -dt * c44 * dyy * s6u - dt * dy * mlx4 + mlx4;
        const T rw1 = -c44 * s1w - T(2.0) * dx * mw1t - dx * dx * mw1 + mly1 + mly2 + fw;
        const T rw2 = -c44 * s3u - c13 * s4u - dx * mw2t - dy * mw2t - dx * dy * mw2;
        const T rw3 = -c33 * s2w - T(2.0) * dy * mw3t - dy * dy * mw3 + mly3 + mly4;
        const T rw4 = -dt * c44 * dxx * s5w - dt * dx * mly1 + mly1;
        const T rw5 = -dt * c13 * dyyp * s5u - dt * dy * mly2 + mly2;
        const T rw6 = -dt * c44 * dxxp * s6u - dt * dx * mly3 + mly3;
        const T rw7 = -dt * c33 * dyy * s6w - dt * dy * mly4 + mly4;

        // ---- own-row state ----------------------------------------------
        if constexpr (LDSG) {
            const T *x = lds + ro(a.rslot)[row] * a.lds_stride + ls;
#pragma unroll
            for (int f = 0; f < NFIELD; ++f)
                if (f < U1TT || f > W3TT) cur[f] = x[gathered_slot(f) * GS];
        } else if constexpr (LDS1) {
            const T *x = lds + a.sell_own[(size_t)slice * 64 + lane];
#pragma unroll
            for (int f = 0; f < NFIELD; ++f)
                if (f < U1TT || f > W3TT) cur[f] = x[gathered_slot(f) * a.lds_stride];
        } else if (interior) {
            if constexpr (!PRE) {
                if (!agg) {
#pragma unroll
                    for (int f = U1; f <= W3; ++f) cur[f] = in[f * FS + o];
                }
            }
            if (!CAP || !got) {
#pragma unroll
                for (int f = LX1; f < NFIELD; ++f) cur[f] = in[f * FS + o];
                if (agg) {
                    own_u = uwin[uw_base(row)];
                    own_w = uwin[uw_base(row) + FS];
                }
            }
        } else if (!CAP || !got) {
            // (every row of an assembled operator holds its diagonal, so with
            // CAP this runs only for a pattern without one)
#pragma unroll
            for (int f = 0; f < NFIELD; ++f)
                if (f < U1TT || f > W3TT) cur[f] = in[f * FS + o];
        }
        if (agg) {
            cur[U1TT] = PRE ? pre_acc[0] : ld_once<NT>(in + U1TT * FS + o);
            cur[W1TT] = PRE ? pre_acc[3] : ld_once<NT>(in + W1TT * FS + o);
        } else {
#pragma unroll
            for (int q = 0; q < 6; ++q) cur[U1TT + q] = PRE ? pre_acc[q] : ld_once<NT>(in + (U1TT + q) * FS + o);
        }

        // ---- lumped-mass solve with Dirichlet clamp (:681-729) ------------
        T an[6], ln[8];
        if constexpr (EXACT) {
            if (mass != T(0)) {
                if (ro(a.flag)[ci] == 1) {
#pragma unroll
                    for (int q = 0; q < 6; ++q) an[q] = T(0);
#pragma unroll
                    for (int q = 0; q < 8; ++q) ln[q] = T(0);
                } else {
                    const T m1 = T(1.0) * mass;
                    an[0] = ru1 / m1; an[1] = ru2 / m1; an[2] = ru3 / m1;
                    ln[0] = ru4 / m1; ln[1] = ru5 / m1; ln[2] = ru6 / m1; ln[3] = ru7 / m1;
                    an[3] = rw1 / m1; an[4] = rw2 / m1; an[5] = rw3 / m1;
                    ln[4] = rw4 / m1; ln[5] = rw5 / m1; ln[6] = rw6 / m1; ln[7] = rw7 / m1;
                }
            } else {
                // zero lumped mass: the reference leaves the outputs untouched
#pragma unroll
                for (int q = 0; q < 6; ++q) an[q] = cur[U1TT + q];
#pragma unroll
                for (int q = 0; q < 8; ++q) ln[q] = cur[LX1 + q];
                atomicOr(&a.ctrl->err, 1);
            }
        } else {
            // mass holds 1/mass_lump with the Dirichlet flag folded in as 0
            an[0] = ru1 * mass; an[1] = ru2 * mass; an[2] = ru3 * mass;
            ln[0] = ru4 * mass; ln[1] = ru5 * mass; ln[2] = ru6 * mass; ln[3] = ru7 * mass;
            an[3] = rw1 * mass; an[4] = rw2 * mass; an[5] = rw3 * mass;
            ln[4] = rw4 * mass; ln[5] = rw5 * mass; ln[6] = rw6 * mass; ln[7] = rw7 * mass;
        }

        // ---- Newmark update (:757-776) ---------------------------------------
        const T ca = a.ca, al = a.alpha, cd = a.cd, de = a.delta;
        T dn[6], cn[6];
        T sum_u, sum_w;                          // U and W of the new state
        if constexpr (C16 && !PRE) load_terms();
        if (agg) {
            // the three split updates of a summed row, added: U' = U + p_u dt +
            // alpha (a1 + a2 + a3) dt^2, p_u' = p_u + (a1 + a2 + a3) dt, and for W
            const T an_s[2] = {an[0] + an[1] + an[2], an[3] + an[4] + an[5]};
            T res[2];
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const T d = h ? own_w : own_u;
                const T p = cur[U1TT + 3 * h];
                if constexpr (KAHAN) {
                    const T inc = p * dt + al * an_s[h] * dt * dt;
                    const T y = inc - ((PRE || C16) ? pre_c[3 * h] : ld_once<NT>(cin + h * FS + co));
                    const T sm = d + y;
                    cn[3 * h] = (sm - d) - y;
                    if constexpr (!C16) st_once<NT>(cout + h * FS + co, cn[3 * h]);
                    res[h] = sm;
                } else {
                    res[h] = d + p * dt + al * an_s[h] * dt * dt;
                }
                st_once<NT>(out + (U1TT + 3 * h) * FS + o, p + (de + ca) * an_s[h] * dt);
            }
            sum_u = res[0];
            sum_w = res[1];
            if constexpr (C16) st_once<NT>(wout + co, bf16_bits(cn[0]) | (bf16_bits(cn[3]) << 16));
        } else {
#pragma unroll
            for (int q = 0; q < 6; ++q) {
                const T d = cur[U1 + q];
                if constexpr (EXACT) {
                    const T v = cur[U1T + q], acc = cur[U1TT + q];
                    dn[q] = d + v * dt + (ca * acc + al * an[q]) * dt * dt;
                    st_once<NT>(out + (U1T + q) * FS + o, v + (cd * acc + de * an[q]) * dt);
                    st_once<NT>(out + (U1TT + q) * FS + o, an[q]);
                } else {
                    // p-form: the acceleration slot holds p = v + ca dt a
                    const T p = cur[U1TT + q];
                    if constexpr (KAHAN) {
                        const T inc = p * dt + al * an[q] * dt * dt;
                        const T y = inc - ((PRE || C16) ? pre_c[q] : ld_once<NT>(cin + q * FS + co));
                        const T sum = d + y;
                        cn[q] = (sum - d) - y;
                        if constexpr (!C16) st_once<NT>(cout + q * FS + co, cn[q]);
                        dn[q] = sum;
                    } else {
                        dn[q] = d + p * dt + al * an[q] * dt * dt;
                    }
                    if (keepv) st_once<NT>(out + (U1T + q) * FS + o, p + de * an[q] * dt);
                    st_once<NT>(out + (U1TT + q) * FS + o, p + (de + ca) * an[q] * dt);
                }
                st_once<NT>(out + (U1 + q) * FS + o, dn[q]);
            }
            if constexpr (C16) {
#pragma unroll
                for (int k = 0; k < 3; ++k)
                    st_once<NT>(wout + k * FS + co, bf16_bits(cn[2 * k]) | (bf16_bits(cn[2 * k + 1]) << 16));
            }
            // the sums every neighbour forms from the split fields, in the same order
            sum_u = dn[0] + dn[1] + dn[2];
            sum_w = dn[3] + dn[4] + dn[5];
        }
#pragma unroll
        for (int q = 0; q < 8; ++q) st_once<NT>(out + (LX1 + q) * FS + o, ln[q]);
        if constexpr (WUW) {
            const size_t ob = uw_base(row);
            st_once<NT>(uwout + ob, sum_u);
            st_once<NT>(uwout + ob + FS, sum_w);
        }

        // ---- receivers sitting on this node --------------------------------
        const int r0 = ro(a.rec_ptr)[row], r1 = ro(a.rec_ptr)[row + 1];
        if (r0 < r1) {
            const double Un = (double)sum_u;
            const double Wn = (double)sum_w;
            for (int q = r0; q < r1; ++q) {
                const size_t w = ((size_t)s * a.nsteps_max + t) * a.rec_num + ro(a.rec_idx)[q];
                a.seis_u[w] = Un;
                a.seis_w[w] = Wn;
            }
        }
    }

}

// The grid holds every (row block, shot group) tile; a launch hands out only
// the tiles of the shot groups still active at its step (shots are sorted by
// decreasing step count, so they are the first groups) and the surplus blocks
// exit at once.  With the XCD remap this keeps the remaining work spread over
// all 8 XCDs when shots finish, instead of leaving it on the XCDs that own the
// first groups.
template <class T, int PREC, int SP, bool CPS, int V>
__global__ __launch_bounds__(256) void step_kernel(const StepArgs<T> a, int slot)
{
    const int t = begin_step(a.ctrl, slot);
    constexpr unsigned int GS = SP >= 64 ? 64 : SP;
    const unsigned int nact = (t < a.nsteps_max) ? (unsigned int)ro(a.active)[t] : 0u;
    const unsigned int groups = min(gridDim.y, (nact + GS - 1) / GS);
    const unsigned int nb = gridDim.x * groups;
    unsigned int bid = blockIdx.y * gridDim.x + blockIdx.x;
    if (bid >= nb) return;
    if (a.xcd_stripe > 0) bid = xcd_striped(bid, nb, (unsigned int)a.xcd_stripe);
    else if (a.xcd_remap) bid = xcd_contiguous(bid, nb);
    step_tile<T, PREC, SP, CPS, V>(a, t, (int)(bid % gridDim.x), (int)(bid / gridDim.x));
}

// Several runs (own mesh, medium, shots and step counts) in one launch per
// step.  tile_start[r] is the first logical block of run r (runs in order,
// tile_start[nruns] = total); a run's tiles are its (row block, shot chunk)
// grid, row blocks fastest.  Runs with one row per lane (SP = 1) and runs with
// 32- or 64-shot lane groups can be mixed; every run shares the step counter
// of args[0].ctrl, and a run past its last step has no active shot.
// SP1: only runs of one shot per lane group (their forms alone set the registers).
template <class T, int PREC, bool SP1>
__device__ __forceinline__ void multi_tile(const StepArgs<T> *__restrict__ args, const int *__restrict__ tile_start,
                                           int nruns, int t, int bid)
{
    int r = 0;
    while (r + 1 < nruns && bid >= tile_start[r + 1]) ++r;
    const StepArgs<T> &a = args[r];
    const int local = bid - tile_start[r];
    const int bx = local % a.nbx, by = local / a.nbx;
    // Kernel variants of one launch: separate value arrays (0), packed values
    // (1; also runs 6), and for 32- and 64-shot groups streaming hints (2) and
    // the form with own-row data first (runs 3, 4 and 5): 3 nonzeros per trip
    // in the FP32 forms, 2 in FP64.  The launch's register count is that of
    // its largest form; deeper unrolls would cost every run a wave per SIMD
    // (FP64 with 3 per trip: 256 VGPRs, one wave).
    constexpr int D = PREC == 0 ? 3 : 4;
    const int v = a.variant;
    const int w = v == 0 ? 0 : (v >= 3 && v <= 5) ? 3 : v == 2 ? 2 : 1;   // 7 (sliced) only with sp == 1
#define SFG_T(SPV, C, VV) step_tile<T, PREC, SPV, C, VV>(a, t, bx, by)
#define SFG_G(SPV, C)                                  \
    do {                                               \
        if (w == 3) SFG_T(SPV, C, D);                  \
        else if (w == 2) SFG_T(SPV, C, 2);             \
        else if (w) SFG_T(SPV, C, 1);                  \
        else SFG_T(SPV, C, 0);                         \
    } while (0)
    if (SP1 || a.sp == 1) {
        // single-shot runs in the sliced layout (7) or the per-row layouts
        if (a.cps) { if (v == 7) SFG_T(1, true, 7); else if (w) SFG_T(1, true, 1); else SFG_T(1, true, 0); }
        else { if (v == 7) SFG_T(1, false, 7); else if (w) SFG_T(1, false, 1); else SFG_T(1, false, 0); }
    } else if constexpr (!SP1) {
        if (a.sp == 32) {
            if (a.cps) SFG_G(32, true);
            else SFG_G(32, false);
        } else {
            if (a.cps) SFG_G(64, true);
            else SFG_G(64, false);
        }
    }
#undef SFG_G
#undef SFG_T
}

template <class T, int PREC, bool SP1>
__global__ __launch_bounds__(256) void multi_step_kernel(const StepArgs<T> *__restrict__ args,
                                                         const int *__restrict__ tile_start, int nruns,
                                                         int xcd_remap, int slot)
{
    Ctrl *ctrl = args[0].ctrl;
    const int t = begin_step(ctrl, slot);
    const unsigned int nb = gridDim.x;
    unsigned int bid = blockIdx.x;
    // xcd_remap: 0 round-robin, 1 contiguous ranges, Q + 1 stripes of Q tiles
    if (xcd_remap > 1) bid = xcd_striped(bid, nb, (unsigned int)(xcd_remap - 1));
    else if (xcd_remap) bid = xcd_contiguous(bid, nb);
    multi_tile<T, PREC, SP1>(args, tile_start, nruns, t, (int)bid);
}

// Persistent form of multi_step_kernel: a cooperative grid of resident blocks
// runs nsteps steps in one launch, meeting at a grid-wide barrier between
// steps (a step's gathers read what other blocks wrote in the previous one).
// A step's tiles are split into 8 contiguous ranges, one per XCD, each with a
// queue counter; a block takes PQ_CHUNK tiles at a time from the range of the
// XCD it runs on (hardware register XCC_ID), then from the other ranges in
// turn until all are exhausted.  Every tile is taken exactly once whatever XCD
// a block reports, since each range's counter only grows.  queue[2][8]: the
// counters by step parity; those of the next step are cleared during the
// current one (nothing uses them after the previous barrier) and all are left
// at zero.  The step counter in args[0].ctrl ends at t0 + nsteps in the slot
// the next launch reads.
constexpr int PQ_CHUNK = 4;

template <class T, int PREC, bool SP1>
__global__ __launch_bounds__(256, 2) void multi_persistent_kernel(const StepArgs<T> *__restrict__ args,
                                                                  const int *__restrict__ tile_start, int nruns,
                                                                  int ntiles, int slot, int nsteps, int *queue)
{
    cooperative_groups::grid_group grid = cooperative_groups::this_grid();
    Ctrl *ctrl = args[0].ctrl;
    const int t0 = __hip_atomic_load(&ctrl->step[slot], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
    // HW_REG_XCC_ID (id 20), bits 3:0
    const int xcc = (int)(__builtin_amdgcn_s_getreg(20 | (3 << 11)) & 7u);
    __shared__ int tile_s;
    for (int s = 0; s < nsteps; ++s) {
        int *q = queue + 8 * (s & 1);
        if (lead)
            for (int k = 0; k < 8; ++k)
                __hip_atomic_store(queue + 8 * ((s + 1) & 1) + k, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        for (int k = 0; k < 8; ++k) {
            const int x = (xcc + k) & 7;
            const int lo = (int)((long long)ntiles * x / 8), hi = (int)((long long)ntiles * (x + 1) / 8);
            for (;;) {
                if (threadIdx.x == 0)
                    tile_s = lo + PQ_CHUNK * __hip_atomic_fetch_add(q + x, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                __syncthreads();
                const int first = tile_s;
                __syncthreads();
                if (first >= hi) break;
                for (int tile = first; tile < min(hi, first + PQ_CHUNK); ++tile)
                    multi_tile<T, PREC, SP1>(args, tile_start, nruns, t0 + s, tile);
            }
        }
        grid.sync();
    }
    if (lead) {
        __hip_atomic_store(&ctrl->step[slot ^ (nsteps & 1)], t0 + nsteps, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        for (int k = 0; k < 16; ++k) __hip_atomic_store(queue + k, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

template <bool SP1>
static const void *persistent_fn(int precision)
{
    switch (precision) {
    case 0: return (const void *)multi_persistent_kernel<double, 0, SP1>;
    case 1: return (const void *)multi_persistent_kernel<float, 1, SP1>;
    case 2: return (const void *)multi_persistent_kernel<float, 2, SP1>;
    case 3: return (const void *)multi_persistent_kernel<float, 3, SP1>;
    default: return nullptr;
    }
}

hipError_t multi_persistent_blocks_per_cu(int precision, bool single_shot_only, int block, int *per_cu)
{
    const void *f = single_shot_only ? persistent_fn<true>(precision) : persistent_fn<false>(precision);
    if (!f) return hipErrorInvalidValue;
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(per_cu, f, block, 0);
}

hipError_t launch_multi_persistent(int precision, const void *args_dev, const int *tile_start_dev, int nruns,
                                   int ntiles, int block, int blocks, int slot, int nsteps, int *queue,
                                   bool single_shot_only, hipStream_t st)
{
    const void *f = single_shot_only ? persistent_fn<true>(precision) : persistent_fn<false>(precision);
    if (!f) return hipErrorInvalidValue;
    void *kargs[] = {(void *)&args_dev, (void *)&tile_start_dev, (void *)&nruns, (void *)&ntiles,
                     (void *)&slot, (void *)&nsteps, (void *)&queue};
    return hipLaunchCooperativeKernel(f, dim3(blocks), dim3(block), kargs, 0, st);
}

// The single-shot runs of a multi-run in the LDS-staged sliced layout (kernel
// variant 8) get a launch of their own per step: their tiles run at this
// form's register count and take its dynamic LDS (the largest halo of these
// runs), not the batched forms' register count.  Same tile table scheme as
// multi_step_kernel; every run shares args[0].ctrl.
template <class T, int PREC>
__global__ __launch_bounds__(256) void multi_step_lds_kernel(const StepArgs<T> *__restrict__ args,
                                                             const int *__restrict__ tile_start, int nruns,
                                                             int xcd_remap, int slot)
{
    Ctrl *ctrl = args[0].ctrl;
    const int t = begin_step(ctrl, slot);
    const unsigned int nb = gridDim.x;
    unsigned int bid = blockIdx.x;
    if (xcd_remap > 1) bid = xcd_striped(bid, nb, (unsigned int)(xcd_remap - 1));
    else if (xcd_remap) bid = xcd_contiguous(bid, nb);
    int r = 0;
    while (r + 1 < nruns && (int)bid >= tile_start[r + 1]) ++r;
    const StepArgs<T> &a = args[r];
    const int local = (int)bid - tile_start[r];
    const int bx = local % a.nbx, by = local / a.nbx;
    if (a.cps) step_tile<T, PREC, 1, true, 8>(a, t, bx, by);
    else step_tile<T, PREC, 1, false, 8>(a, t, bx, by);
}

hipError_t launch_multi_step_lds(int precision, const void *args_dev, const int *tile_start_dev, int nruns,
                                 int ntiles, int block, int xcd_remap, int slot, size_t lds_bytes, hipStream_t st)
{
#define SFG_ML(T, P)                                                                                        \
    hipLaunchKernelGGL((multi_step_lds_kernel<T, P>), dim3(ntiles), dim3(block), lds_bytes, st,            \
                       (const StepArgs<T> *)args_dev, tile_start_dev, nruns, xcd_remap, slot)
    switch (precision) {
    case 0: SFG_ML(double, 0); break;
    case 1: SFG_ML(float, 1); break;
    case 2: SFG_ML(float, 2); break;
    case 3: SFG_ML(float, 3); break;
    default: return hipErrorInvalidValue;
    }
#undef SFG_ML
    return hipGetLastError();
}

hipError_t allow_multi_step_lds(int precision, size_t lds_bytes)
{
    const void *f = nullptr;
    switch (precision) {
    case 0: f = (const void *)multi_step_lds_kernel<double, 0>; break;
    case 1: f = (const void *)multi_step_lds_kernel<float, 1>; break;
    case 2: f = (const void *)multi_step_lds_kernel<float, 2>; break;
    case 3: f = (const void *)multi_step_lds_kernel<float, 3>; break;
    default: return hipErrorInvalidValue;
    }
    return hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes);
}

template <bool SP1>
static hipError_t launch_multi_step_t(int precision, const void *args_dev, const int *tile_start_dev, int nruns,
                                      int ntiles, int block, int xcd_remap, int slot, hipStream_t st)
{
#define SFG_MM(T, P)                                                                                          \
    hipLaunchKernelGGL((multi_step_kernel<T, P, SP1>), dim3(ntiles), dim3(block), 0, st,                      \
                       (const StepArgs<T> *)args_dev, tile_start_dev, nruns, xcd_remap, slot)
    switch (precision) {
    case 0: SFG_MM(double, 0); break;
    case 1: SFG_MM(float, 1); break;
    case 2: SFG_MM(float, 2); break;
    case 3: SFG_MM(float, 3); break;
    default: return hipErrorInvalidValue;
    }
#undef SFG_MM
    return hipGetLastError();
}

hipError_t launch_multi_step(int precision, const void *args_dev, const int *tile_start_dev, int nruns, int ntiles,
                             int block, int xcd_remap, int slot, bool single_shot_only, hipStream_t st)
{
    return single_shot_only
               ? launch_multi_step_t<true>(precision, args_dev, tile_start_dev, nruns, ntiles, block, xcd_remap, slot, st)
               : launch_multi_step_t<false>(precision, args_dev, tile_start_dev, nruns, ntiles, block, xcd_remap, slot,
                                            st);
}

template <class T, int PREC, int V>
static hipError_t launch_step_v(int sp, bool cps, const StepArgs<T> &a, dim3 g, int block, int slot, size_t lds,
                                hipStream_t st)
{
#define SFG_L(SPV, C) hipLaunchKernelGGL((step_kernel<T, PREC, SPV, C, V>), g, dim3(block), lds, st, a, slot)
    switch (sp) {
    case 1: if (cps) SFG_L(1, true); else SFG_L(1, false); break;
    case 4: if (cps) SFG_L(4, true); else SFG_L(4, false); break;
    case 16: if (cps) SFG_L(16, true); else SFG_L(16, false); break;
    case 32: if (cps) SFG_L(32, true); else SFG_L(32, false); break;
    case 64: if (cps) SFG_L(64, true); else SFG_L(64, false); break;
    default: return hipErrorInvalidValue;
    }
#undef SFG_L
    return hipGetLastError();
}

// LDS-staged variant: node-major lane groups only
template <class T, int PREC>
static hipError_t launch_step_lds(int sp, bool cps, const StepArgs<T> &a, dim3 g, int block, int slot, size_t lds,
                                  hipStream_t st)
{
#define SFG_L(SPV, C) hipLaunchKernelGGL((step_kernel<T, PREC, SPV, C, 6>), g, dim3(block), lds, st, a, slot)
    switch (sp) {
    case 16: if (cps) SFG_L(16, true); else SFG_L(16, false); break;
    case 32: if (cps) SFG_L(32, true); else SFG_L(32, false); break;
    case 64: if (cps) SFG_L(64, true); else SFG_L(64, false); break;
    default: return hipErrorInvalidValue;
    }
#undef SFG_L
    return hipGetLastError();
}

template <class T, int PREC>
static hipError_t launch_step_t(int sp, bool cps, const void *args, dim3 g, int block, int slot, size_t lds,
                                hipStream_t st)
{
    const StepArgs<T> &a = *static_cast<const StepArgs<T> *>(args);
    switch (a.variant) {
    case 0: return launch_step_v<T, PREC, 0>(sp, cps, a, g, block, slot, 0, st);
    case 1: return launch_step_v<T, PREC, 1>(sp, cps, a, g, block, slot, 0, st);
    case 2: return launch_step_v<T, PREC, 2>(sp, cps, a, g, block, slot, 0, st);
    case 3: return launch_step_v<T, PREC, 3>(sp, cps, a, g, block, slot, 0, st);
    case 4: return launch_step_v<T, PREC, 4>(sp, cps, a, g, block, slot, 0, st);
    case 5: return launch_step_v<T, PREC, 5>(sp, cps, a, g, block, slot, 0, st);
    case 6: return launch_step_lds<T, PREC>(sp, cps, a, g, block, slot, lds, st);
    case 7:
        // sliced layout: one row per lane only
        if (sp != 1) return hipErrorInvalidValue;
        if (cps) hipLaunchKernelGGL((step_kernel<T, PREC, 1, true, 7>), g, dim3(block), 0, st, a, slot);
        else hipLaunchKernelGGL((step_kernel<T, PREC, 1, false, 7>), g, dim3(block), 0, st, a, slot);
        return hipGetLastError();
    case 8:
        // sliced layout with the halo staged in LDS: one row per lane only
        if (sp != 1) return hipErrorInvalidValue;
        if (cps) hipLaunchKernelGGL((step_kernel<T, PREC, 1, true, 8>), g, dim3(block), lds, st, a, slot);
        else hipLaunchKernelGGL((step_kernel<T, PREC, 1, false, 8>), g, dim3(block), lds, st, a, slot);
        return hipGetLastError();
    default: return hipErrorInvalidValue;
    }
}

// Allows the LDS-staged step kernel of this (precision, SP, per-shot
// coefficients) to take lds_bytes of dynamic LDS.
template <class T, int PREC>
static hipError_t allow_lds_t(int sp, bool cps, int bytes)
{
    const void *f = nullptr;
    switch (sp) {
    case 1: f = cps ? (const void *)step_kernel<T, PREC, 1, true, 8> : (const void *)step_kernel<T, PREC, 1, false, 8>; break;
    case 16: f = cps ? (const void *)step_kernel<T, PREC, 16, true, 6> : (const void *)step_kernel<T, PREC, 16, false, 6>; break;
    case 32: f = cps ? (const void *)step_kernel<T, PREC, 32, true, 6> : (const void *)step_kernel<T, PREC, 32, false, 6>; break;
    case 64: f = cps ? (const void *)step_kernel<T, PREC, 64, true, 6> : (const void *)step_kernel<T, PREC, 64, false, 6>; break;
    default: return hipErrorInvalidValue;
    }
    return hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
}

hipError_t allow_step_lds(int precision, int sp, bool cps, size_t lds_bytes)
{
    switch (precision) {
    case 0: return allow_lds_t<double, 0>(sp, cps, (int)lds_bytes);
    case 1: return allow_lds_t<float, 1>(sp, cps, (int)lds_bytes);
    case 2: return allow_lds_t<float, 2>(sp, cps, (int)lds_bytes);
    case 3: return allow_lds_t<float, 3>(sp, cps, (int)lds_bytes);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_step(int precision, int sp, bool cps, const void *args, int nbx, int nby,
                       int block, int slot, size_t lds_bytes, hipStream_t st)
{
    const dim3 g(nbx, nby);
    switch (precision) {
    case 0: return launch_step_t<double, 0>(sp, cps, args, g, block, slot, lds_bytes, st);
    case 1: return launch_step_t<float, 1>(sp, cps, args, g, block, slot, lds_bytes, st);
    case 2: return launch_step_t<float, 2>(sp, cps, args, g, block, slot, lds_bytes, st);
    case 3: return launch_step_t<float, 3>(sp, cps, args, g, block, slot, lds_bytes, st);
    default: return hipErrorInvalidValue;
    }
}

// ---- energy guard: sum(U^2), sum(W^2) per shot (reference :777-785) ----------
// Threads are laid out [rows][shots] (shots innermost, as the state), each
// accumulates a strided set of nodes, then a block reduction per shot.
template <class T>
__global__ __launch_bounds__(256) void guard_kernel(const T *state, const T *uw, int n, const StateLayout L,
                                                    const Ctrl *ctrl, const int *shot_steps, double *energy,
                                                    int nslots, int every, int nodes_per_block)
{
    __shared__ double red_u[256], red_w[256];
    const int t = executed_steps(ctrl);
    const int S = L.S;
    const int sg = S < 64 ? S : 64;                 // shots covered by one block
    const int rstep = blockDim.x / sg;
    const int sl = threadIdx.x % sg, rl = threadIdx.x / sg;
    const int s = blockIdx.y * sg + sl;
    // a shot that stopped after k < t steps keeps its last state in buffer k & 1
    const int done = s < S ? min(t, shot_steps[s]) : t;
    const T *in = state + (size_t)(done & 1) * NFIELD * L.F;
    // U and W kept beside the state (FP32 forms), in its layout with 2 fields
    const T *sums = uw ? uw + (size_t)(done & 1) * 2 * L.F : nullptr;
    StateLayout L2 = L;
    L2.nf = 2;
    double eu = 0.0, ew = 0.0;
    const int n0 = blockIdx.x * nodes_per_block;
    const int n1 = min(n, n0 + nodes_per_block);
    if (rl < rstep && s < S) {
        for (int i = n0 + rl; i < n1; i += rstep) {
            double U, W;
            if (sums) {
                U = (double)sums[state_index(L2, i, 0, s)];
                W = (double)sums[state_index(L2, i, 1, s)];
            } else {
                U = (double)(in[state_index(L, i, U1, s)] + in[state_index(L, i, U2, s)] +
                             in[state_index(L, i, U3, s)]);
                W = (double)(in[state_index(L, i, W1, s)] + in[state_index(L, i, W2, s)] +
                             in[state_index(L, i, W3, s)]);
            }
            eu += U * U;
            ew += W * W;
        }
    }
    red_u[threadIdx.x] = eu;
    red_w[threadIdx.x] = ew;
    __syncthreads();
    if (rl == 0 && s < S) {
        for (int q = 1; q < rstep; ++q) {
            eu += red_u[q * sg + sl];
            ew += red_w[q * sg + sl];
        }
        // check c covers the steps ((c-1)*every, c*every]
        const int slot = min((t + every - 1) / every, nslots - 1);
        atomicAdd(&energy[((size_t)slot * S + s) * 2 + 0], eu);
        atomicAdd(&energy[((size_t)slot * S + s) * 2 + 1], ew);
    }
}

hipError_t launch_guard(int precision, const void *state, const void *uw, int n, const StateLayout &L,
                        const Ctrl *ctrl, const int *shot_steps, double *energy, int nslots, int every,
                        hipStream_t st)
{
    const int npb = 2048;
    const int sg = L.S < 64 ? L.S : 64;
    const dim3 g((n + npb - 1) / npb, (L.S + sg - 1) / sg);
    if (precision == 0)
        hipLaunchKernelGGL(guard_kernel<double>, g, dim3(256), 0, st, (const double *)state, (const double *)uw,
                           n, L, ctrl, shot_steps, energy, nslots, every, npb);
    else
        hipLaunchKernelGGL(guard_kernel<float>, g, dim3(256), 0, st, (const float *)state, (const float *)uw,
                           n, L, ctrl, shot_steps, energy, nslots, every, npb);
    return hipGetLastError();
}

// ---- U = U1+U2+U3, W = W1+W2+W3 of one shot (snapshots, :775-776, :797-806) --
template <class T>
__global__ void gather_uw_kernel(const T *state, const T *uw, int n, const StateLayout L, int shot, int parity,
                                 double *U, double *W)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (uw) {
        // U and W kept beside the state (FP32 forms)
        const T *sums = uw + (size_t)parity * 2 * L.F;
        StateLayout L2 = L;
        L2.nf = 2;
        U[i] = (double)sums[state_index(L2, i, 0, shot)];
        W[i] = (double)sums[state_index(L2, i, 1, shot)];
        return;
    }
    const T *in = state + (size_t)parity * NFIELD * L.F;
    U[i] = (double)(in[state_index(L, i, U1, shot)] + in[state_index(L, i, U2, shot)] +
                    in[state_index(L, i, U3, shot)]);
    W[i] = (double)(in[state_index(L, i, W1, shot)] + in[state_index(L, i, W2, shot)] +
                    in[state_index(L, i, W3, shot)]);
}

hipError_t launch_gather_uw(int precision, const void *state, const void *uw, int n, const StateLayout &L,
                            int shot, int parity, double *U, double *W, hipStream_t st)
{
    const dim3 g((n + 255) / 256);
    if (precision == 0)
        hipLaunchKernelGGL(gather_uw_kernel<double>, g, dim3(256), 0, st, (const double *)state,
                           (const double *)uw, n, L, shot, parity, U, W);
    else
        hipLaunchKernelGGL(gather_uw_kernel<float>, g, dim3(256), 0, st, (const float *)state,
                           (const float *)uw, n, L, shot, parity, U, W);
    return hipGetLastError();
}

// ---- y = A x, one row per lane, sequential row sums (csr_matvec.c:8-17) -------
__global__ void spmv64_kernel(int n, const int *rp, const int *cj, const double *a, const double *x,
                              double *y)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double t = 0.0;
    for (int k = rp[i]; k < rp[i + 1]; ++k) t = t + a[k] * x[cj[k]];
    y[i] = t;
}

hipError_t launch_spmv64(int n, const int *rp, const int *cj, const double *a, const double *x,
                         double *y, hipStream_t st)
{
    hipLaunchKernelGGL(spmv64_kernel, dim3((n + 255) / 256), dim3(256), 0, st, n, rp, cj, a, x, y);
    return hipGetLastError();
}

// ---- M-PML damping profiles in FP64 (pml/abc_mpml.c:50-199) ---------------------
// out6 = 6 arrays of n: dx, dy, dxx, dyy, dxx_pyx, dyy_pxy.  The region tests
// and expressions keep the reference's order; every node gets its value from
// its own coordinates, which is what the reference's per-element visits give.
__global__ void mpml_kernel(int n, const double *X, const double *Y, const MpmlGeo g, double *out)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const double x0 = g.x0, x1 = g.x1, y0 = g.y0, y1 = g.y1;
    const double tx = g.tx, ty = g.ty, d0x = g.d0x, d0y = g.d0y;
    const double pxy = g.pxy, pyx = g.pyx;
    const int uxn = g.use_xmin, uxx = g.use_xmax, uyn = g.use_ymin, uyx = g.use_ymax;
    const double x = X[p], y = Y[p];
    double dx = 0.0, dy = 0.0, dxx = 0.0, dyy = 0.0;
    if ((x <= x0) && (y >= y0) && (y <= y1) && (uxn == 1)) {          // left strip
        dx = d0x * ((x0 - x) / tx) * ((x0 - x) / tx);
        dy = pyx * dx;
        dxx = -d0x * 2 * (x0 - x) / (tx * tx);
        dyy = 0.0;
    } else if ((x >= x1) && (y >= y0) && (y <= y1) && (uxx == 1)) {   // right strip
        dx = d0x * ((x - x1) / tx) * ((x - x1) / tx);
        dy = pyx * dx;
        dxx = d0x * 2 * (x - x1) / (tx * tx);
        dyy = 0.0;
    } else if ((y <= y0) && (x >= x0) && (x <= x1) && (uyn == 1)) {   // low-y strip
        dy = d0y * ((y0 - y) / ty) * ((y0 - y) / ty);
        dx = pxy * dy;
        dxx = 0.0;
        dyy = -d0y * 2 * (y0 - y) / (ty * ty);
    } else if ((y >= y1) && (x >= x0) && (x <= x1) && (uyx == 1)) {   // high-y strip
        dy = d0y * ((y - y1) / ty) * ((y - y1) / ty);
        dx = pxy * dy;
        dxx = 0.0;
        dyy = d0y * 2 * (y - y1) / (ty * ty);
    } else if ((x <= x0) && (y <= y0) && ((uxn == 1) || (uyn == 1))) { // corner 12
        dx = d0x * ((x0 - x) / tx) * ((x0 - x) / tx) + pxy * d0y * ((y0 - y) / ty) * ((y0 - y) / ty);
        dy = d0y * ((y0 - y) / ty) * ((y0 - y) / ty) + pyx * d0x * ((x0 - x) / tx) * ((x0 - x) / tx);
        dxx = -d0x * 2 * (x0 - x) / (tx * tx);
        dyy = -d0y * 2 * (y0 - y) / (ty * ty);
    } else if ((x <= x0) && (y >= y1) && ((uxn == 1) || (uyx == 1))) { // corner 14
        dx = d0x * ((x0 - x) / tx) * ((x0 - x) / tx) + pxy * d0y * ((y - y1) / ty) * ((y - y1) / ty);
        dy = d0y * ((y - y1) / ty) * ((y - y1) / ty) + pyx * d0x * ((x0 - x) / tx) * ((x0 - x) / tx);
        dyy = d0y * 2 * (y - y1) / (ty * ty);
        dxx = -d0x * 2 * (x0 - x) / (tx * tx);
    } else if ((x >= x1) && (y <= y0) && ((uxx == 1) || (uyn == 1))) { // corner 23
        if (uyx == 1) { // the reference tests the ymax switch here (abc_mpml.c:147)
            dx = d0x * ((x - x1) / tx) * ((x - x1) / tx) + pxy * d0y * ((y0 - y) / ty) * ((y0 - y) / ty);
            dxx = d0x * 2 * (x - x1) / (tx * tx);
            dy = d0y * ((y0 - y) / ty) * ((y0 - y) / ty) + pyx * d0x * ((x - x1) / tx) * ((x - x1) / tx);
        } else {
            dx = pxy * d0y * ((y0 - y) / ty) * ((y0 - y) / ty);
            dxx = 0.0;
            dy = d0y * ((y0 - y) / ty) * ((y0 - y) / ty);
        }
        dyy = -d0y * 2 * (y0 - y) / (ty * ty);
    } else if ((x >= x1) && (y >= y1) && ((uxx == 1) || (uyx == 1))) { // corner 34
        if (uyx == 1) {
            dx = d0x * ((x - x1) / tx) * ((x - x1) / tx) + pxy * d0y * ((y - y1) / ty) * ((y - y1) / ty);
            dy = d0y * ((y - y1) / ty) * ((y - y1) / ty) + pyx * d0x * ((x - x1) / tx) * ((x - x1) / tx);
            dxx = d0x * 2 * (x - x1) / (tx * tx);
        } else {
            dx = pxy * d0y * ((y - y1) / ty) * ((y - y1) / ty);
            dy = d0y * ((y - y1) / ty) * ((y - y1) / ty);
            dxx = 0.0;
        }
        dyy = d0y * 2 * (y - y1) / (ty * ty);
    }
    out[0 * (size_t)n + p] = dx;
    out[1 * (size_t)n + p] = dy;
    out[2 * (size_t)n + p] = dxx;
    out[3 * (size_t)n + p] = dyy;
    out[4 * (size_t)n + p] = pyx * dxx;
    out[5 * (size_t)n + p] = pxy * dyy;
}

hipError_t launch_mpml(int n, const double *x, const double *y, const MpmlGeo &g, double *out6,
                       hipStream_t st)
{
    hipLaunchKernelGGL(mpml_kernel, dim3((n + 255) / 256), dim3(256), 0, st, n, x, y, g, out6);
    return hipGetLastError();
}

__global__ void to_f32_kernel(const double *src, float *dst, size_t n)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        dst[i] = (float)src[i];
}

hipError_t launch_to_f32(const double *src, float *dst, size_t n, hipStream_t st)
{
    hipLaunchKernelGGL(to_f32_kernel, dim3(1024), dim3(256), 0, st, src, dst, n);
    return hipGetLastError();
}

__global__ void fill_zero_kernel(uint32_t *p, size_t nwords)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < nwords;
         i += (size_t)gridDim.x * blockDim.x)
        p[i] = 0u;
}

template <class T>
__global__ void pml_rows_kernel(const T *coef, int n, int ns, int *pml)
{
    const size_t cs = (size_t)n * ns;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        int any = 0;
        for (int c = DX; c <= DYYP; ++c)
            for (int s = 0; s < ns; ++s) any |= coef[c * cs + (size_t)i * ns + s] != T(0);
        pml[i] = any ? 2 : 0;
    }
}

// Marks the columns of M-PML rows that are not M-PML rows themselves.  Threads
// may store 1 to the same entry at once; a 2 is never overwritten.
__global__ void pml_ring_kernel(int n, const int *rp, const int *ci, int *pml)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        if (pml[i] != 2) continue;
        for (int k = rp[i]; k < rp[i + 1]; ++k) {
            const int j = ci[k];
            if (pml[j] != 2) pml[j] = 1;
        }
    }
}

hipError_t launch_pml_rows(int precision, const void *coef, int n, int ns, const int *rp, const int *ci, int *pml,
                           hipStream_t st)
{
    const int blocks = (n + 255) / 256 < 4096 ? (n + 255) / 256 : 4096;
    if (n <= 0) return hipSuccess;
    if (precision == 0)
        hipLaunchKernelGGL(pml_rows_kernel<double>, dim3(blocks), dim3(256), 0, st, (const double *)coef, n, ns, pml);
    else
        hipLaunchKernelGGL(pml_rows_kernel<float>, dim3(blocks), dim3(256), 0, st, (const float *)coef, n, ns, pml);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(pml_ring_kernel, dim3(blocks), dim3(256), 0, st, n, rp, ci, pml);
    return hipGetLastError();
}

__global__ void medium_f32_kernel(int n, const double *c64, const int *flag, float *c32)
{
    const size_t N = (size_t)n;
    for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < (size_t)NCOEF * N;
         q += (size_t)gridDim.x * blockDim.x) {
        const int k = (int)(q / N);
        const size_t i = q % N;
        if (k == MASS) {
            const double m = c64[q];
            c32[q] = (flag[i] == 1 || m == 0.0) ? 0.0f : (float)(1.0 / m);
        } else {
            c32[q] = (float)c64[q];
        }
    }
}

hipError_t launch_medium_f32(int n, const double *coef64, const int *flag, float *coef32, hipStream_t st)
{
    hipLaunchKernelGGL(medium_f32_kernel, dim3(1024), dim3(256), 0, st, n, coef64, flag, coef32);
    return hipGetLastError();
}

// Per-shot coefficients: shot-innermost copy of each medium's [ncoef][n] array.
// Padded shots (s >= nsrc) repeat the last medium so every lane reads finite data.
template <class E>
__global__ void pack_per_shot_kernel(const E *const *srcs, int nsrc, int ncoef, int n, int S, E *dst)
{
    const size_t total = (size_t)ncoef * n * S;
    for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < total; q += (size_t)gridDim.x * blockDim.x) {
        const int s = (int)(q % S);
        const size_t ci = q / S; // c * n + i
        dst[q] = srcs[s < nsrc ? s : nsrc - 1][ci];
    }
}

hipError_t launch_pack_per_shot(const void *const *srcs_dev, int nsrc, int ncoef, int n, int S, int elem,
                                void *dst, hipStream_t st)
{
    if (elem == 8)
        hipLaunchKernelGGL(pack_per_shot_kernel<uint64_t>, dim3(2048), dim3(256), 0, st,
                           (const uint64_t *const *)srcs_dev, nsrc, ncoef, n, S, (uint64_t *)dst);
    else if (elem == 4)
        hipLaunchKernelGGL(pack_per_shot_kernel<uint32_t>, dim3(2048), dim3(256), 0, st,
                           (const uint32_t *const *)srcs_dev, nsrc, ncoef, n, S, (uint32_t *)dst);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

template <class E>
__global__ void sell_values_kernel(const E *vals, size_t nnz, const int *src, size_t nslots, E *out)
{
    for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < nslots; q += (size_t)gridDim.x * blockDim.x) {
        const int k = src[q];
        const size_t base = (q / 64) * (7 * 64) + q % 64;
#pragma unroll
        for (int m = 0; m < 7; ++m) out[base + (size_t)m * 64] = k >= 0 ? vals[(size_t)m * nnz + k] : E(0);
    }
}

hipError_t launch_sell_values(const void *vals, size_t nnz, const int *src, size_t nslots, int elem, void *out,
                              hipStream_t st)
{
    if (nslots % 64) return hipErrorInvalidValue;
    if (elem == 8)
        hipLaunchKernelGGL(sell_values_kernel<double>, dim3(2048), dim3(256), 0, st, (const double *)vals, nnz, src,
                           nslots, (double *)out);
    else if (elem == 4)
        hipLaunchKernelGGL(sell_values_kernel<float>, dim3(2048), dim3(256), 0, st, (const float *)vals, nnz, src,
                           nslots, (float *)out);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

hipError_t launch_fill_zero(void *p, size_t bytes, hipStream_t st)
{
    // every buffer here is a multiple of 4 bytes
    hipLaunchKernelGGL(fill_zero_kernel, dim3(2048), dim3(256), 0, st, (uint32_t *)p, bytes / 4);
    return hipGetLastError();
}


// ---- dt estimate: y = M_L^{-1} K x for the elastic operator without M-PML ----------
// K [u; w] from the d = 0 reduction of the RHS (elastic_wave.c:585-587, 651-653):
//   (K x)_u = c11 S1 u + c44 S2 u + c13 S3 w + c44 S4 w
//   (K x)_w = c44 S1 w + c33 S2 w + c44 S3 u + c13 S4 u
// Clamped (Dirichlet) rows get 0.  x, y hold [u (n) ; w (n)].  sumsq[0] += |y|^2.
__global__ void elastic_apply_kernel(int n, const int *rp, const int *ci, const double *A,
                                     size_t nnz, const double *coef, const int *flag,
                                     const double *x, double *y, double *sumsq)
{
    __shared__ double red[256];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double acc = 0.0;
    if (i < n) {
        double s1u = 0, s1w = 0, s2u = 0, s2w = 0, s3u = 0, s3w = 0, s4u = 0, s4w = 0;
        for (int k = rp[i]; k < rp[i + 1]; ++k) {
            const int j = ci[k];
            const double u = x[j], w = x[(size_t)n + j];
            const double a1 = A[1 * nnz + k], a2 = A[2 * nnz + k], a3 = A[3 * nnz + k], a4 = A[4 * nnz + k];
            s1u += a1 * u; s1w += a1 * w; s2u += a2 * u; s2w += a2 * w;
            s3u += a3 * u; s3w += a3 * w; s4u += a4 * u; s4w += a4 * w;
        }
        const size_t N = (size_t)n;
        const double c11 = coef[C11 * N + i], c13 = coef[C13 * N + i];
        const double c33 = coef[C33 * N + i], c44 = coef[C44 * N + i];
        const double m = coef[MASS * N + i];
        double yu = 0.0, yw = 0.0;
        if (flag[i] != 1 && m != 0.0) {
            yu = (c11 * s1u + c44 * s2u + c13 * s3w + c44 * s4w) / m;
            yw = (c44 * s1w + c33 * s2w + c44 * s3u + c13 * s4u) / m;
        }
        y[i] = yu;
        y[N + i] = yw;
        acc = yu * yu + yw * yw;
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int o = blockDim.x / 2; o > 0; o >>= 1) {
        if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicAdd(sumsq, red[0]);
}

__global__ void scale_kernel(size_t n, double *x, const double *sumsq)
{
    const double s = sumsq[0] > 0.0 ? 1.0 / sqrt(sumsq[0]) : 0.0;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        x[i] *= s;
}

hipError_t launch_elastic_apply(int n, const int *rp, const int *ci, const double *A, size_t nnz,
                                const double *coef, const int *flag, const double *x, double *y,
                                double *sumsq, hipStream_t st)
{
    hipLaunchKernelGGL(elastic_apply_kernel, dim3((n + 255) / 256), dim3(256), 0, st, n, rp, ci, A, nnz,
                       coef, flag, x, y, sumsq);
    return hipGetLastError();
}

hipError_t launch_scale(size_t n, double *x, const double *sumsq, hipStream_t st)
{
    hipLaunchKernelGGL(scale_kernel, dim3(1024), dim3(256), 0, st, n, x, sumsq);
    return hipGetLastError();
}

// Plain 16-byte-per-lane copy: the yardstick for the HBM roofline and for
// scaling the byte counters of a profiler (it moves exactly 2 x n x 16 bytes).
// One 16-byte element per thread and a grid over the whole array: on MI355X
// this form reached 6.2 TB/s for 1 GiB, against 4.5-5.1 TB/s for grid-stride
// loops of 2k-32k blocks and 5.1 TB/s for hipMemcpy device to device.
__global__ __launch_bounds__(256) void stream_copy_kernel(const float4 *__restrict__ src,
                                                          float4 *__restrict__ dst, size_t n)
{
    const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

hipError_t launch_stream_copy(const void *src, void *dst, size_t n16, hipStream_t st)
{
    const size_t blocks = (n16 + 255) / 256;
    if (blocks > 0x7fffffffu) return hipErrorInvalidValue;
    hipLaunchKernelGGL(stream_copy_kernel, dim3((unsigned int)blocks), dim3(256), 0, st, (const float4 *)src,
                       (float4 *)dst, n16);
    return hipGetLastError();
}

} // namespace sfg
