// Surrogate MLP on the GPU: bf16 MFMA GEMM (FP32 accumulate), GELU, bias, MSE,
// Adam with FP32 master weights, and bf16 inference.  No counterpart in the
// reference; it learns layered-model parameters -> receiver traces from shots
// the FEM loop generates.
//
// Network: x[nin] -> (Linear + GELU) x nhidden_layers -> Linear -> y[nout].
// Weights W_l [out][in] (row-major, "in" contiguous), bias b_l [out].
// All GEMMs are C[M][N] = A[M][K] . B[N][K]^T with A and B bf16:
//   forward   z = h . W^T          (A = h [batch][in],  B = W [out][in])
//   dX        dh = dz . W          (A = dz [batch][out], B = W [out][in] read k-major)
//   dW        dW = dz^T . h        (A = dz^T [out][batch], B = h^T [in][batch])
// dX reads W as stored (a k-major operand, transposed on its way into LDS), so
// no W^T copy is kept; dW uses transposed bf16 copies of dz and h.  Widths are
// padded to multiples of 32 (zeros); M / N / K edges are bounds-checked.
#include <hip/hip_runtime.h>
#include <hip/hip_bf16.h>
#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "seisfem_gpu.h"
#include "seisfem_surrogate.h"

namespace {

typedef float f32x16 __attribute__((ext_vector_type(16)));

__device__ __forceinline__ uint16_t f2bf(float f)
{
    // round to nearest even (finite inputs; NaN kept quiet)
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

// GELU, tanh form
__device__ __forceinline__ float gelu(float x)
{
    const float k = 0.7978845608028654f; // sqrt(2/pi)
    return 0.5f * x * (1.0f + tanhf(k * (x + 0.044715f * x * x * x)));
}
__device__ __forceinline__ float gelu_grad(float x)
{
    const float k = 0.7978845608028654f;
    const float u = k * (x + 0.044715f * x * x * x);
    const float th = tanhf(u);
    const float du = k * (1.0f + 3.0f * 0.044715f * x * x);
    return 0.5f * (1.0f + th) + 0.5f * x * (1.0f - th * th) * du;
}

// Adam hyper-parameters of one update, bias corrections included
struct AdamArgs {
    float lr, b1, b2, eps, bc1, bc2;
};

// one Adam update of weight w with gradient g; m and v are updated in place
__device__ __forceinline__ float adam_update(float w, float g, float &m, float &v, const AdamArgs &a)
{
    const float mm = a.b1 * m + (1.0f - a.b1) * g;
    const float vv = a.b2 * v + (1.0f - a.b2) * g * g;
    m = mm;
    v = vv;
    return w - a.lr * (mm / a.bc1) / (sqrtf(vv / a.bc2) + a.eps);
}

// What the GEMM does with a finished element v = (A . B^T)[row][col] (C row stride ldc):
//   EPI_STORE      C = v
//   EPI_BIAS       C = v + bias[col]                 (H = bf16 of it, if given)
//   EPI_BIAS_GELU  C = v + bias[col], H = bf16(gelu(C))
//   EPI_MSE        d = v + bias[col] - T[tidx[row]][col]; C = scale d and H = bf16(scale d) on
//                  columns < nvalid (0 beyond); *loss += sum d^2 (the mean-squared-error
//                  gradient of the output layer, targets read in place by sample index)
//   EPI_ADAM       v is the gradient of weight W[row][col] (row stride ldc, moments Mw, Vw
//                  alike): one Adam step on rows < arows, columns < acols, and the bf16
//                  copy W16 refreshed; C is not written (row-piece epilogue only)
enum Epi { EPI_STORE = 0, EPI_BIAS = 1, EPI_BIAS_GELU = 2, EPI_MSE = 3, EPI_ADAM = 4 };
struct EpiArgs {
    int kind = EPI_STORE;
    const float *bias = nullptr;
    uint16_t *H = nullptr;
    int ldh = 0;
    const float *T = nullptr;
    const int *tidx = nullptr;
    int ldt = 0;
    int nvalid = 0;
    float scale = 0.0f;
    double *loss = nullptr;
    float *W = nullptr, *Mw = nullptr, *Vw = nullptr;
    uint16_t *W16 = nullptr;
    int arows = 0, acols = 0;
    AdamArgs ad{};
    // a product split over K (gridDim.z > 1): 0 = the parts are added into C with
    // atomics; > 0 = part z is stored (EPI_STORE) at C + z * slice
    size_t slice = 0;
    // EPI_MSE with db: the block holds every row (M <= BM), so it also sums the
    // gradient over the rows into db[col] (col < nvalid) instead of storing C, and
    // with bw updates the bias bw by Adam (moments bm, bv; arguments ad) from it
    float *db = nullptr, *bw = nullptr, *bm = nullptr, *bv = nullptr;
};

// ---- C = A . B^T, bf16 in, fp32 out -----------------------------------------
// Block tile BM x BN, K in steps of BK = 64, NW waves in (NW / WN) x WN, each
// wave's part as 32 x 32 tiles of v_mfma_f32_32x32x16_bf16.  The LDS
// tiles are double-buffered: while a wave multiplies step k out of one buffer,
// its share of step k+1 is loaded from global memory into registers, then
// stored into the other buffer; one barrier per step.  blockIdx.z splits K
// into chunks of kchunk (a multiple of BK): with more than one chunk, chunk z's
// partial product is stored at C + z * ep.slice (ep.slice > 0), or added to C
// (zeroed beforehand) with atomics; no epilogue other than the store applies.
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
constexpr int BK = 64, LDK = BK + 8;

// TB: B is stored [K][N] (ldb >= N) instead of [N][K]; it is read along N in
// 16-byte pieces and transposed while it is written into the LDS tile.
template <int BM, int BN, int NW, int WN, bool TB>
__global__ __launch_bounds__(NW * 64) void gemm_nt_kernel(int M, int N, int K, int kchunk, const uint16_t *__restrict__ A,
                                                      int lda, const uint16_t *__restrict__ B, int ldb,
                                                      float *__restrict__ C, int ldc, const EpiArgs ep)
{
    constexpr int WM = NW / WN;                         // waves along M and N
    constexpr int TM = BM / WM / 32, TN = BN / WN / 32; // 32 x 32 tiles per wave
    constexpr int NT = NW * 64;
    constexpr int LA = BM * BK / 8 / NT, LB = BN * BK / 8 / NT;  // 16-byte loads per thread and step
    static_assert(TM >= 1 && TN >= 1 && LA >= 1 && LB >= 1, "tile shape");
    __shared__ __align__(16) uint16_t As[2][BM * LDK];
    __shared__ __align__(16) uint16_t Bs[2][BN * LDK];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave / WN, wn = wave % WN;
    const int m0 = blockIdx.y * BM, n0 = blockIdx.x * BN;
    const int kb = blockIdx.z * kchunk, ke = min(K, kb + kchunk);
    f32x16 acc[TM][TN];
#pragma unroll
    for (int a = 0; a < TM; ++a)
#pragma unroll
        for (int b = 0; b < TN; ++b)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[a][b][r] = 0.0f;
    // Loads run two K steps ahead of the products: step t + 2 is fetched into
    // register set t % 2 while step t is multiplied out of LDS buffer t % 2, and
    // set (t + 1) % 2 (fetched a step earlier) is stored into the other buffer.
    uint4 ra[2][LA], rb[2][LB];
    // A (and B unless TB): 16-byte piece q of an [rows][BK] tile is row q / 8, k
    // 8 (q % 8) ..  With TB the B tile [BK][BN] is read along N: thread t takes
    // columns 8 (t % (BN / 8)) .. and the LB consecutive k rows LB (t / (BN / 8))
    // .., and writes each column's LB values as one piece of the tile row.  In
    // that tile the 16-byte chunk j of row n sits at chunk j ^ (n / 8 % 8), so
    // that the 8 columns of a thread and their neighbours' spread over the banks.
    // Every load is issued, from a clamped in-bounds address when its piece lies
    // past an edge, and such a piece is replaced by zeros when it is stored to
    // LDS (bit i of ok[set] marks piece i valid): a load skipped by a branch, or
    // a select right after it, would make the wait before a register set's store
    // wait for every load in flight, the next steps' included.
    uint32_t ok[2] = {0u, 0u};
    auto load_rows = [&](const uint16_t *X, int ld, int lim, int r0, int q, int k0, bool &valid)
        __attribute__((always_inline)) {
        const int row = q >> 3, kk = k0 + (q & 7) * 8, g = r0 + row;
        valid = g < lim && kk < ke;
        return *reinterpret_cast<const uint4 *>(X + (size_t)min(g, lim - 1) * ld + min(kk, ke - 8));
    };
    const int tcg = tid % (BN / 8), tkg = tid / (BN / 8);
    auto fetch = [&](auto S, int k0) __attribute__((always_inline)) {
        constexpr int st = decltype(S)::value;
        uint32_t m = 0;
        bool valid;
#pragma unroll
        for (int i = 0; i < LA; ++i) {
            ra[st][i] = load_rows(A, lda, M, m0, tid + NT * i, k0, valid);
            m |= (uint32_t)valid << i;
        }
#pragma unroll
        for (int i = 0; i < LB; ++i) {
            if constexpr (TB) {
                const int kk = k0 + tkg * LB + i, g = n0 + tcg * 8;
                rb[st][i] = *reinterpret_cast<const uint4 *>(B + (size_t)min(kk, ke - 1) * ldb + min(g, N - 8));
                valid = g < N && kk < ke;
            } else {
                rb[st][i] = load_rows(B, ldb, N, n0, tid + NT * i, k0, valid);
            }
            m |= (uint32_t)valid << (LA + i);
        }
        ok[st] = m;
    };
    auto kept = [&](const uint4 &v, uint32_t m, int bit) __attribute__((always_inline)) {
        const uint32_t z = 0u - ((m >> bit) & 1u);   // all ones for a valid piece (no branch)
        return make_uint4(v.x & z, v.y & z, v.z & z, v.w & z);
    };
    auto half = [](const uint4 &v, int e) __attribute__((always_inline)) {
        const uint32_t w = e < 2 ? v.x : e < 4 ? v.y : e < 6 ? v.z : v.w;
        return (uint32_t)((w >> (16 * (e & 1))) & 0xffffu);
    };
    auto stash = [&](auto S, int buf) __attribute__((always_inline)) {
        constexpr int st = decltype(S)::value;
#pragma unroll
        for (int i = 0; i < LA; ++i) {
            const int q = tid + NT * i;
            *reinterpret_cast<uint4 *>(&As[buf][(q >> 3) * LDK + (q & 7) * 8]) = kept(ra[st][i], ok[st], i);
        }
        if constexpr (TB) {
            static_assert(LB == 1 || LB == 2 || LB == 4, "k-major piece");
            uint4 rv[LB];
#pragma unroll
            for (int i = 0; i < LB; ++i) rv[i] = kept(rb[st][i], ok[st], LA + i);
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                const int kq = tkg * LB;
                uint16_t *dst = &Bs[buf][(tcg * 8 + e) * LDK + (((kq >> 3) ^ (tcg & 7)) << 3) + (kq & 7)];
                if constexpr (LB == 1) {
                    *dst = (uint16_t)half(rv[0], e);
                } else if constexpr (LB == 2) {
                    *reinterpret_cast<uint32_t *>(dst) = half(rv[0], e) | (half(rv[1], e) << 16);
                } else {
                    *reinterpret_cast<uint2 *>(dst) = make_uint2(half(rv[0], e) | (half(rv[1], e) << 16),
                                                                 half(rv[2], e) | (half(rv[3], e) << 16));
                }
            }
        } else {
#pragma unroll
            for (int i = 0; i < LB; ++i) {
                const int q = tid + NT * i;
                *reinterpret_cast<uint4 *>(&Bs[buf][(q >> 3) * LDK + (q & 7) * 8]) = kept(rb[st][i], ok[st], LA + i);
            }
        }
    };
    using I0 = std::integral_constant<int, 0>;
    using I1 = std::integral_constant<int, 1>;
    // The loop has no branch but its own: the step count is rounded up to even
    // (an extra step multiplies tiles of zeros, which adds exact zeros), and the
    // fetches past the last step load clamped addresses whose pieces are zeroed.
    // The compiler can then wait before a store for the older register set's
    // loads only, with the newer set's still in flight.
    const int nk = ((ke - kb + BK - 1) / BK + 1) & ~1;
    fetch(I0{}, kb);
    fetch(I1{}, kb + BK);
    stash(I0{}, 0);
    __syncthreads();
    auto step = [&](int t, auto S) __attribute__((always_inline)) {
        constexpr int buf = decltype(S)::value;
        fetch(S, kb + (t + 2) * BK);
        // keep the loads at the top of the step (the scheduler would sink them
        // below the products and the other set's store in the first step of a pair)
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int kk = 0; kk < BK; kk += 16) {
            // operands of v_mfma_f32_32x32x16_bf16: lane l holds row (col) l % 32 and
            // k = kk + 8 (l / 32) .. + 7
            const int kof = kk + 8 * (lane >> 5);
            bf16x8 af[TM], bfr[TN];
#pragma unroll
            for (int a = 0; a < TM; ++a)
                af[a] = *reinterpret_cast<const bf16x8 *>(&As[buf][(wm * (BM / WM) + a * 32 + (lane & 31)) * LDK + kof]);
#pragma unroll
            for (int b = 0; b < TN; ++b) {
                const int n = wn * (BN / WN) + b * 32 + (lane & 31);
                const int kb_ = TB ? (((kof >> 3) ^ ((n >> 3) & 7)) << 3) : kof;
                bfr[b] = *reinterpret_cast<const bf16x8 *>(&Bs[buf][n * LDK + kb_]);
            }
#pragma unroll
            for (int a = 0; a < TM; ++a)
#pragma unroll
                for (int b = 0; b < TN; ++b)
                    acc[a][b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(af[a], bfr[b], acc[a][b], 0, 0, 0);
        }
        stash(std::integral_constant<int, 1 - buf>{}, 1 - buf);
        __syncthreads();
    };
    for (int t = 0; t < nk; t += 2) {
        step(t, I0{});
        step(t + 1, I1{});
    }
    // epilogue: lane l, register r -> row (r/4)*8 + (l/32)*4 + r%4, column l%32.
    // Per 32 x 32 tile, every load an epilogue needs is issued before any of its
    // stores (the stores could alias the loads, so loads interleaved with them
    // would each wait out a full memory latency).  One unrolled walk over the
    // tiles per epilogue kind keeps acc in registers.
    auto tiles = [&](auto f) __attribute__((always_inline)) {
#pragma unroll
        for (int a = 0; a < TM; ++a)
#pragma unroll
            for (int b = 0; b < TN; ++b)
                f(acc[a][b], n0 + wn * (BN / WN) + b * 32 + (lane & 31), m0 + wm * (BM / WM) + a * 32 + (lane >> 5) * 4);
    };
    // row of register r in a tile whose lane rows start at rb
#define SFG_ROW(r) (rb + ((r) >> 2) * 8 + ((r) & 3))
    if (gridDim.z > 1) C += (size_t)blockIdx.z * ep.slice;
    if (gridDim.z > 1 && ep.slice == 0) {
        tiles([&](const f32x16 &t, int col, int rb) __attribute__((always_inline)) {
#pragma unroll
            for (int r = 0; r < 16; ++r)
                if (col < N && SFG_ROW(r) < M) atomicAdd(&C[(size_t)SFG_ROW(r) * ldc + col], t[r]);
        });
    } else if (N % 4 == 0 && ldc % 4 == 0 && (reinterpret_cast<uintptr_t>(C) & 15) == 0 &&
               (ep.kind == EPI_STORE || (reinterpret_cast<uintptr_t>(ep.bias) & 15) == 0) &&
               (!ep.H || (ep.ldh % 4 == 0 && (reinterpret_cast<uintptr_t>(ep.H) & 7) == 0)) &&
               (ep.kind != EPI_MSE ||
                (ep.ldt % 4 == 0 && ep.nvalid % 4 == 0 && (reinterpret_cast<uintptr_t>(ep.T) & 15) == 0))) {
        // Each wave turns its tile rows round in LDS (the A buffers, idle after the
        // last step's barrier), 32 rows at a time, so that the epilogue reads and
        // writes whole rows in 4-element pieces (16 bytes of C, bias, targets,
        // weights and moments, 8 bytes of bf16) rather than single elements: a
        // wave's 32 x WC slab is written to LDS element by element in the register
        // layout, then read back along the rows.  The arithmetic per element is that of the paths below.
        constexpr int WC = BN / WN, SLD = WC + (WC >= 64 ? 8 : 4); // floats per staged row
        constexpr int PR = WC / 4, PL = 32 * PR / 64, RS = 64 / PR;   // pieces per row and per lane, rows apart
        static_assert(NW * 32 * SLD * 4 <= (int)sizeof(As), "staging slab fits in the A buffers");
        static_assert(64 % PR == 0, "a lane keeps its columns");
        float *S = reinterpret_cast<float *>(&As[0][0]) + wave * 32 * SLD;
        // lane: columns col .. col + 3 of rows r0 + a * 32 + i * RS
        const int pc = lane % PR, pr = lane / PR;
        const int col = n0 + wn * WC + pc * 4, r0 = m0 + wm * (BM / WM) + pr;
        const bool mse = ep.kind == EPI_MSE, act = ep.kind == EPI_BIAS_GELU;
        const int nv = mse ? ep.nvalid : N; // a multiple of 4: a piece is all valid or all not
        const bool adam = ep.kind == EPI_ADAM, biased = ep.kind != EPI_STORE && !adam;
        float4 bc = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        if (biased) bc = *reinterpret_cast<const float4 *>(&ep.bias[min(col, N - 4)]);
        double lsum = 0.0;
        float cs[4] = {0.0f, 0.0f, 0.0f, 0.0f}; // the lane's column sums of the loss gradient (ep.db)
#pragma unroll
        for (int a = 0; a < TM; ++a) {
#pragma unroll
            for (int b = 0; b < TN; ++b)
#pragma unroll
                for (int r = 0; r < 16; ++r)
                    S[((r >> 2) * 8 + (lane >> 5) * 4 + (r & 3)) * SLD + b * 32 + (lane & 31)] = acc[a][b][r];
            // targets: the pass's sample indices, then its 16-byte target pieces, all
            // from in-bounds addresses (clamped; what lies past an edge is not used),
            // all in flight together
            float4 tv[PL];
            if (mse) {
                int ti[PL];
#pragma unroll
                for (int i = 0; i < PL; ++i) ti[i] = ep.tidx[min(r0 + a * 32 + i * RS, M - 1)];
#pragma unroll
                for (int i = 0; i < PL; ++i)
                    tv[i] = *reinterpret_cast<const float4 *>(ep.T + (size_t)ti[i] * ep.ldt + min(col, nv - 4));
            }
            // Adam: the pass's weights and moments, likewise all issued before any store
            float4 aw[PL], am[PL], av[PL];
            if (adam) {
#pragma unroll
                for (int i = 0; i < PL; ++i) {
                    const size_t o = (size_t)min(r0 + a * 32 + i * RS, ep.arows - 1) * ldc + min(col, N - 4);
                    aw[i] = *reinterpret_cast<const float4 *>(ep.W + o);
                    am[i] = *reinterpret_cast<const float4 *>(ep.Mw + o);
                    av[i] = *reinterpret_cast<const float4 *>(ep.Vw + o);
                }
            }
            __syncthreads();
#pragma unroll
            for (int i = 0; i < PL; ++i) {
                const int row = r0 + a * 32 + i * RS;
                const float4 v = *reinterpret_cast<const float4 *>(&S[(pr + i * RS) * SLD + pc * 4]);
                if (row >= M || col >= N) continue;
                if (adam) {
                    if (row >= ep.arows) continue;
                    float4 w4 = aw[i], m4 = am[i], v4 = av[i];
                    float *pw = &w4.x, *pm = &m4.x, *pv = &v4.x;
                    const float *pg = &v.x;
#pragma unroll
                    for (int e = 0; e < 4; ++e)
                        if (col + e < ep.acols) pw[e] = adam_update(pw[e], pg[e], pm[e], pv[e], ep.ad);
                    const size_t o = (size_t)row * ldc + col;
                    *reinterpret_cast<float4 *>(ep.W + o) = w4;
                    *reinterpret_cast<float4 *>(ep.Mw + o) = m4;
                    *reinterpret_cast<float4 *>(ep.Vw + o) = v4;
                    *reinterpret_cast<uint2 *>(ep.W16 + o) = make_uint2(f2bf(w4.x) | ((uint32_t)f2bf(w4.y) << 16),
                                                                        f2bf(w4.z) | ((uint32_t)f2bf(w4.w) << 16));
                    continue;
                }
                float4 o = v;
                if (biased) o = make_float4(v.x + bc.x, v.y + bc.y, v.z + bc.z, v.w + bc.w);
                if (mse) {
                    const float d[4] = {o.x - tv[i].x, o.y - tv[i].y, o.z - tv[i].z, o.w - tv[i].w};
                    float g[4] = {0.0f, 0.0f, 0.0f, 0.0f};
                    if (col < nv) {
#pragma unroll
                        for (int e = 0; e < 4; ++e) {
                            lsum += (double)d[e] * d[e];
                            g[e] = ep.scale * d[e];
                            cs[e] += g[e];
                        }
                    }
                    o = make_float4(g[0], g[1], g[2], g[3]);
                }
                if (!ep.db) *reinterpret_cast<float4 *>(&C[(size_t)row * ldc + col]) = o;
                if (ep.H) {
                    const float h[4] = {act ? gelu(o.x) : o.x, act ? gelu(o.y) : o.y, act ? gelu(o.z) : o.z,
                                        act ? gelu(o.w) : o.w};
                    *reinterpret_cast<uint2 *>(&ep.H[(size_t)row * ep.ldh + col]) =
                        make_uint2(f2bf(h[0]) | ((uint32_t)f2bf(h[1]) << 16), f2bf(h[2]) | ((uint32_t)f2bf(h[3]) << 16));
                }
            }
            __syncthreads();
        }
        if (mse) {
            // one atomic per block: the waves' sums meet in LDS (free after the last
            // pass's barrier); so do the column sums of the gradient (ep.db): first
            // over the lanes that hold the same columns, then over the WM waves
            static_assert(NW * 8 + WM * BN * 4 + 128 <= (int)sizeof(Bs), "loss and column sums fit");
            for (int o = 32; o > 0; o >>= 1) lsum += __shfl_xor(lsum, o, 64);
            double *ws = reinterpret_cast<double *>(&Bs[0][0]);
            float *red = reinterpret_cast<float *>(&Bs[0][0]) + 32;
            if (lane == 0) ws[wave] = lsum;
            if (ep.db) {
#pragma unroll
                for (int o = PR; o < 64; o <<= 1)
#pragma unroll
                    for (int e = 0; e < 4; ++e) cs[e] += __shfl_xor(cs[e], o, 64);
                if (pr == 0)
#pragma unroll
                    for (int e = 0; e < 4; ++e) red[wm * BN + wn * WC + pc * 4 + e] = cs[e];
            }
            __syncthreads();
            if (tid == 0) {
                double t = 0.0;
                for (int w = 0; w < NW; ++w) t += ws[w];
                atomicAdd(ep.loss, t);
            }
            if (ep.db) {
                for (int c = tid; c < BN; c += NT) {
                    const int gc = n0 + c;
                    if (gc >= ep.nvalid) continue;
                    float g = red[c];
                    for (int w = 1; w < WM; ++w) g += red[w * BN + c];
                    ep.db[gc] = g;
                    if (ep.bw) ep.bw[gc] = adam_update(ep.bw[gc], g, ep.bm[gc], ep.bv[gc], ep.ad);
                }
            }
        }
    } else if (ep.kind == EPI_STORE) {
        tiles([&](const f32x16 &t, int col, int rb) __attribute__((always_inline)) {
#pragma unroll
            for (int r = 0; r < 16; ++r)
                if (col < N && SFG_ROW(r) < M) C[(size_t)SFG_ROW(r) * ldc + col] = t[r];
        });
    } else if (ep.kind == EPI_MSE) {
        double lsum = 0.0;
        tiles([&](const f32x16 &t, int col, int rb) __attribute__((always_inline)) {
            const bool cv = col < ep.nvalid;
            const float bc = col < N ? ep.bias[col] : 0.0f;
            int ti[16];
            float tv[16];
#pragma unroll
            for (int r = 0; r < 16; ++r) ti[r] = (cv && SFG_ROW(r) < M) ? ep.tidx[SFG_ROW(r)] : 0;
#pragma unroll
            for (int r = 0; r < 16; ++r) tv[r] = (cv && SFG_ROW(r) < M) ? ep.T[(size_t)ti[r] * ep.ldt + col] : 0.0f;
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                if (col < N && SFG_ROW(r) < M) {
                    float g = 0.0f;
                    if (cv) {
                        const float d = (t[r] + bc) - tv[r];
                        lsum += (double)d * d;
                        g = ep.scale * d;
                    }
                    C[(size_t)SFG_ROW(r) * ldc + col] = g;
                    ep.H[(size_t)SFG_ROW(r) * ep.ldh + col] = f2bf(g);
                }
            }
        });
        for (int o = 32; o > 0; o >>= 1) lsum += __shfl_xor(lsum, o, 64);
        if (lane == 0) atomicAdd(ep.loss, lsum);
    } else if (ep.kind != EPI_ADAM) { // (the host sends EPI_ADAM only where the row pieces apply)
        const bool act = ep.kind == EPI_BIAS_GELU;
        tiles([&](const f32x16 &t, int col, int rb) __attribute__((always_inline)) {
            const float bc = col < N ? ep.bias[col] : 0.0f;
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                if (col < N && SFG_ROW(r) < M) {
                    const float v = t[r] + bc;
                    C[(size_t)SFG_ROW(r) * ldc + col] = v;
                    if (ep.H) ep.H[(size_t)SFG_ROW(r) * ep.ldh + col] = f2bf(act ? gelu(v) : v);
                }
            }
        });
    }
#undef SFG_ROW
}

// C[r][c] = 0 for r < rows, c < cols (row stride ld)
__global__ void fill_zero_f32_kernel(int rows, int cols, int ld, float *p)
{
    if (ld == cols && (reinterpret_cast<uintptr_t>(p) & 15) == 0) {
        const size_t step = (size_t)gridDim.x * gridDim.y * blockDim.x;
        const size_t t0 = ((size_t)blockIdx.y * gridDim.x + blockIdx.x) * blockDim.x + threadIdx.x;
        // contiguous rows: float4 stores, then the last n % 4 elements
        const size_t n = (size_t)rows * cols, n4 = n / 4;
        for (size_t i = t0; i < n4; i += step) reinterpret_cast<float4 *>(p)[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        if (t0 < n - 4 * n4) p[4 * n4 + t0] = 0.0f;
        return;
    }
    for (int r = blockIdx.y; r < rows; r += gridDim.y)
        for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < cols; c += gridDim.x * blockDim.x) p[(size_t)r * ld + c] = 0.0f;
}

// Zeroes up to 8 contiguous, 16-byte aligned float buffers in one launch
// (blockIdx.y picks the buffer).
struct ZeroList {
    float *p[8];
    size_t n[8];
    int count;
};
__global__ void zero_list_kernel(ZeroList z)
{
    if ((int)blockIdx.y >= z.count) return;
    float *p = z.p[blockIdx.y];
    const size_t n = z.n[blockIdx.y], n4 = n / 4, step = (size_t)gridDim.x * blockDim.x;
    const size_t t0 = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    for (size_t i = t0; i < n4; i += step) reinterpret_cast<float4 *>(p)[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (t0 < n - 4 * n4) p[4 * n4 + t0] = 0.0f;
}

// dst[c][r] = src[r][c] for r < rows, c < cols (bf16), for two such matrices in
// one launch (both operands of a weight gradient): blocks 0 .. nb0 - 1 take the
// 32 x 32 tiles of the first, row of tiles by row, the rest those of the second
struct Transpose {
    int rows, cols;
    const uint16_t *src;
    int lds;
    uint16_t *dst;
    int ldd;
};
__global__ void transpose_bf16_kernel(Transpose t0, Transpose t1, int nb0)
{
    const bool second = (int)blockIdx.x >= nb0;
    const Transpose &t = second ? t1 : t0;
    const int rows = t.rows, cols = t.cols, lds = t.lds, ldd = t.ldd;
    const uint16_t *src = t.src;
    uint16_t *dst = t.dst;
    const int b = (int)blockIdx.x - (second ? nb0 : 0), tx = (cols + 31) / 32;
    const int c0 = (b % tx) * 32, r0 = (b / tx) * 32;
    __shared__ uint16_t tile[32][33];
    for (int j = threadIdx.y; j < 32; j += blockDim.y) {
        const int r = r0 + j, c = c0 + threadIdx.x;
        tile[j][threadIdx.x] = (r < rows && c < cols) ? src[(size_t)r * lds + c] : (uint16_t)0;
    }
    __syncthreads();
    for (int j = threadIdx.y; j < 32; j += blockDim.y) {
        const int c = c0 + j, r = r0 + threadIdx.x;
        if (c < cols && r < rows) dst[(size_t)c * ldd + r] = tile[threadIdx.x][j];
    }
}

__global__ void f32_to_bf16_kernel(size_t n, const float *src, uint16_t *dst)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        dst[i] = f2bf(src[i]);
}

// rows of X (nin wide) for the given sample indices -> bf16 [batch][ldx], zero padded
// (and *zero = 0, the batch's loss accumulator, by the first thread)
__global__ void gather_rows_kernel(int batch, int width, int ld, const float *src, const int *idx, uint16_t *dst,
                                   float *dstf, double *zero)
{
    const int b = blockIdx.y;
    if (zero && b == 0 && blockIdx.x == 0 && threadIdx.x == 0) *zero = 0.0;
    if (b >= batch) return;
    const size_t s = (size_t)idx[b];
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < ld; c += gridDim.x * blockDim.x) {
        const float v = c < width ? src[s * width + c] : 0.0f;
        if (dst) dst[(size_t)b * ld + c] = f2bf(v);
        if (dstf) dstf[(size_t)b * ld + c] = v;
    }
}

// dz = dh * gelu'(z) (fp32 and bf16 copies), dh the sum of the nparts parts of
// a product split over K (part p at parts + p * slice), added in part order
__global__ void gelu_back_kernel(size_t n, const float *parts, int nparts, size_t slice, const float *z, float *dz,
                                 uint16_t *dz16)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        float dh = parts[i];
        for (int p = 1; p < nparts; ++p) dh += parts[p * slice + i];
        const float g = dh * gelu_grad(z[i]);
        dz[i] = g;
        dz16[i] = f2bf(g);
    }
}

// A hidden layer's forward epilogue after a product split over K: z = (sum of the
// nparts parts, in part order) + bias[c], h = bf16(gelu(z)); rows x cols, row
// stride cols for all of them
__global__ void bias_gelu_parts_kernel(int rows, int cols, const float *parts, int nparts, size_t slice,
                                       const float *bias, float *z, uint16_t *h)
{
    const int n = rows * cols;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        float v = parts[i];
        for (int p = 1; p < nparts; ++p) v += parts[p * slice + i];
        const float zz = v + bias[i % cols];
        z[i] = zz;
        h[i] = f2bf(gelu(zz));
    }
}

// db[c] = sum_b dz[b][c]: a block of 64 x 4 threads takes 64 columns, each of
// its 4 rows of threads a quarter of the batch (4 independent partial sums in
// flight), then the 4 partial sums meet in LDS.  With bw given, the bias bw is
// then updated by Adam (moments bm, bv) from that gradient.
__global__ __launch_bounds__(256) void bias_grad_kernel(int batch, int cols, int ld, const float *dz, float *db,
                                                        float *bw, float *bm, float *bv, AdamArgs ad)
{
    __shared__ float part[4][64];
    const int cx = threadIdx.x & 63, ry = threadIdx.x >> 6;
    const int c = blockIdx.x * 64 + cx;
    float s = 0.0f;
    if (c < cols) {
        float p[4] = {0.0f, 0.0f, 0.0f, 0.0f};
        int b = ry;
        for (; b + 12 < batch; b += 16)
#pragma unroll
            for (int u = 0; u < 4; ++u) p[u] += dz[(size_t)(b + 4 * u) * ld + c];
        for (; b < batch; b += 4) p[0] += dz[(size_t)b * ld + c];
        s = (p[0] + p[1]) + (p[2] + p[3]);
    }
    part[ry][cx] = s;
    __syncthreads();
    if (ry == 0 && c < cols) {
        const float g = (part[0][cx] + part[1][cx]) + (part[2][cx] + part[3][cx]);
        db[c] = g;
        if (bw) bw[c] = adam_update(bw[c], g, bm[c], bv[c], ad);
    }
}

int pad32(int x) { return (x + 31) / 32 * 32; }

thread_local std::string g_mlp_err;
int mfail(int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_mlp_err = buf;
    return code;
}
#define MCHK(expr)                                                                                 \
    do {                                                                                           \
        hipError_t e_ = (expr);                                                                    \
        if (e_ != hipSuccess) return mfail(SFG_E_HIP, "%s: %s", #expr, hipGetErrorString(e_));     \
    } while (0)

// C = A . B^T, A [M][K] and B [N][K] (or, tb, [K][N]), epilogue ep; a plain
// store with too few blocks splits K (the partial products meet in C by atomics).
// Tiles: 256 x 128 (eight waves of 64 x 64, one block per CU) for a product whose M
// is one or two tiles of the batch (256 x 32000 x 1024: 28.7 us against 31.0 us for
// 128 x 128); 128 x 128 (four waves of 64 x 64) for other large products (dW of the
// output layer, M = 32000: faster than 256 x 128); 64 x 64 below 256 tiles, the
// K = 32000 product (dX of the output layer) split 8 ways (faster than 256 x 128 split
// 63, and than 128 x 128 split 32: 0.520 against 0.511 ms per training step; and
// than 16 parts: 0.474 against 0.469 ms).
// Loading three K steps ahead instead of two measured no faster.
struct GemmShape {
    bool big, wide;
    int bm, bn, splits, kchunk;
};
GemmShape gemm_shape(int M, int N, int K, int kind)
{
    GemmShape g;
    g.big = (long)((M + 127) / 128) * ((N + 127) / 128) >= 256;
    g.wide = g.big && M % 256 == 0 && M <= 512;
    g.bm = g.wide ? 256 : g.big ? 128 : 64;
    g.bn = g.big ? 128 : 64;
    const long tiles = (long)((M + g.bm - 1) / g.bm) * ((N + g.bn - 1) / g.bn);
    const int ksteps = (K + BK - 1) / BK;
    int splits = 1;
    if (kind == EPI_STORE && tiles < 256 && !g.wide) splits = (int)std::min<long>(ksteps, (512 + tiles - 1) / tiles);
    g.kchunk = ((ksteps + splits - 1) / splits) * BK;
    g.splits = (K + g.kchunk - 1) / g.kchunk;
    return g;
}

// C = A . B^T with the epilogue ep.  A product split over K adds its partial
// products into C: zeroed here first, unless the caller has (zeroed).
hipError_t gemm(hipStream_t st, int M, int N, int K, const uint16_t *A, int lda, const uint16_t *B, int ldb, float *C,
                int ldc, const EpiArgs &ep, bool tb = false, bool zeroed = false)
{
    if (K % 8 || lda % 8 || ldb % 8 || M <= 0 || N <= 0 || K <= 0) return hipErrorInvalidValue;
    if (tb && N % 8) return hipErrorInvalidValue;
    if (ep.kind == EPI_MSE && !ep.H) return hipErrorInvalidValue;
    if (ep.kind == EPI_ADAM) {
        // the row-piece epilogue must apply: 16-byte rows of C (unused), W and its moments
        auto a16 = [](const void *p) { return (reinterpret_cast<uintptr_t>(p) & 15) == 0; };
        if (N % 4 || ldc % 4 || !a16(C) || !ep.W || !a16(ep.W) || !ep.Mw || !a16(ep.Mw) || !ep.Vw || !a16(ep.Vw) ||
            !ep.W16 || (reinterpret_cast<uintptr_t>(ep.W16) & 7) || ep.H || ep.arows < 1 || ep.arows > M ||
            ep.acols < 1 || ep.acols > N)
            return hipErrorInvalidValue;
    }
    const GemmShape gs = gemm_shape(M, N, K, ep.kind);
    if (ep.db) {
        // the row-piece loss epilogue must apply, in blocks that hold every row
        auto al = [](const void *p, uintptr_t a) { return (reinterpret_cast<uintptr_t>(p) & (a - 1)) == 0; };
        if (ep.kind != EPI_MSE || M > gs.bm || N % 4 || ldc % 4 || !al(C, 16) || !ep.bias || !al(ep.bias, 16) ||
            ep.ldt % 4 || ep.nvalid % 4 || ep.nvalid > N || !ep.T || !al(ep.T, 16) || ep.ldh % 4 || !al(ep.H, 8) ||
            (ep.bw && (!ep.bm || !ep.bv)))
            return hipErrorInvalidValue;
    }
    const bool wide = gs.wide, big = gs.big;
    const int bm = gs.bm, bn = gs.bn, splits = gs.splits, kchunk = gs.kchunk;
    if (ep.slice && (ep.kind != EPI_STORE || ep.slice < (size_t)M * ldc || ep.slice % 4)) return hipErrorInvalidValue;
    if (splits > 1 && !zeroed && !ep.slice)
        hipLaunchKernelGGL(fill_zero_f32_kernel, dim3(1024), dim3(256), 0, st, M, N, ldc, C);
    const dim3 g((N + bn - 1) / bn, (M + bm - 1) / bm, splits);
#define SFG_GEMM(BMV, TBV)                                                                                        \
    hipLaunchKernelGGL((gemm_nt_kernel<BMV, BMV, 4, 2, TBV>), g, dim3(256), 0, st, M, N, K, kchunk, A, lda, B, ldb, C, \
                       ldc, ep)
    if (wide) {
        if (tb)
            hipLaunchKernelGGL((gemm_nt_kernel<256, 128, 8, 2, true>), g, dim3(512), 0, st, M, N, K, kchunk, A, lda, B,
                               ldb, C, ldc, ep);
        else
            hipLaunchKernelGGL((gemm_nt_kernel<256, 128, 8, 2, false>), g, dim3(512), 0, st, M, N, K, kchunk, A, lda, B,
                               ldb, C, ldc, ep);
    } else if (big) {
        if (tb) SFG_GEMM(128, true); else SFG_GEMM(128, false);
    } else {
        if (tb) SFG_GEMM(64, true); else SFG_GEMM(64, false);
    }
#undef SFG_GEMM
    return hipGetLastError();
}

hipError_t transpose2(hipStream_t st, const Transpose &a, const Transpose &b)
{
    const int nb0 = ((a.cols + 31) / 32) * ((a.rows + 31) / 32), nb1 = ((b.cols + 31) / 32) * ((b.rows + 31) / 32);
    hipLaunchKernelGGL(transpose_bf16_kernel, dim3(nb0 + nb1), dim3(32, 8), 0, st, a, b, nb0);
    return hipGetLastError();
}

} // namespace

// ---------------------------------------------------------------------------
// the network object

struct sfg_mlp {
    int nin = 0, hid = 0, nh = 0, nout = 0, batch = 0, nl = 0;
    std::vector<int> din, dout; // true widths per layer
    std::vector<int> pin, pout; // padded widths
    // per layer: master weights, grads, Adam moments (fp32 [pout][pin]), bias (+grad, moments)
    std::vector<float *> W, G, Mw, Vw, b, gb, Mb, Vb;
    std::vector<uint16_t *> W16;
    // activations: h16[l] input of layer l (bf16 [batch][pin]); z[l] pre-activation fp32 [batch][pout]
    std::vector<uint16_t *> h16, hT16;
    std::vector<float *> z, dz, dh;
    std::vector<uint16_t *> dz16, dzT16;
    float *part = nullptr; // parts of the products split over K (dX; hidden layers' forward)
    size_t part_n = 0;
    bool out_db = false;   // the last forward formed the output layer's bias gradient
    float *y = nullptr;
    double *loss = nullptr;
    int *idx = nullptr;   // sample indices of one batch
    int *order = nullptr; // an epoch's sample order (ndata)
    float *Xd = nullptr, *Yd = nullptr;
    int ndata = 0;
    long step = 0;
    hipStream_t st = nullptr;
};

static void mlp_free(sfg_mlp *m)
{
    auto F = [](auto &vec) {
        for (auto *p : vec)
            if (p) (void)hipFree(p);
        vec.clear();
    };
    F(m->W); F(m->G); F(m->Mw); F(m->Vw); F(m->b); F(m->gb); F(m->Mb); F(m->Vb);
    F(m->W16); F(m->h16); F(m->hT16); F(m->z); F(m->dz); F(m->dh); F(m->dz16); F(m->dzT16);
    for (void *p : {(void *)m->y, (void *)m->part, (void *)m->loss, (void *)m->idx, (void *)m->order, (void *)m->Xd,
                    (void *)m->Yd})
        if (p) (void)hipFree(p);
    if (m->st) (void)hipStreamDestroy(m->st);
    delete m;
}

template <class T>
static hipError_t zalloc(T **p, size_t count)
{
    hipError_t e = hipMalloc((void **)p, std::max<size_t>(count, 1) * sizeof(T));
    if (e == hipSuccess) e = hipMemset(*p, 0, std::max<size_t>(count, 1) * sizeof(T));
    return e;
}

extern "C" const char *sfg_mlp_last_error(void) { return g_mlp_err.c_str(); }

extern "C" int sfg_mlp_create(sfg_mlp **out, int nin, int hidden, int nhidden_layers, int nout, int batch,
                              unsigned long long seed)
{
    g_mlp_err.clear();
    if (!out || nin <= 0 || hidden <= 0 || nhidden_layers <= 0 || nout <= 0 || batch <= 0)
        return mfail(SFG_E_ARG, "mlp: bad sizes");
    *out = nullptr;
    sfg_mlp *m = new sfg_mlp;
    m->nin = nin; m->hid = hidden; m->nh = nhidden_layers; m->nout = nout; m->batch = batch;
    m->nl = nhidden_layers + 1;
    for (int l = 0; l < m->nl; ++l) {
        m->din.push_back(l == 0 ? nin : hidden);
        m->dout.push_back(l == m->nl - 1 ? nout : hidden);
        m->pin.push_back(pad32(m->din.back()));
        m->pout.push_back(pad32(m->dout.back()));
    }
    hipError_t e = hipStreamCreateWithFlags(&m->st, hipStreamNonBlocking);
    // He-style uniform init from a fixed-seed generator, on the host
    uint64_t s = seed * 0x9E3779B97F4A7C15ull + 1;
    auto rnd = [&]() {
        s ^= s >> 12; s ^= s << 25; s ^= s >> 27;
        return (double)((s * 2685821657736338717ull) >> 11) * (1.0 / 9007199254740992.0);
    };
    for (int l = 0; l < m->nl && e == hipSuccess; ++l) {
        const size_t nw = (size_t)m->pout[l] * m->pin[l];
        float *W, *G, *Mw, *Vw, *b, *gb, *Mb, *Vb;
        uint16_t *W16;
        e = zalloc(&W, nw);
        if (e == hipSuccess) e = zalloc(&G, nw);
        if (e == hipSuccess) e = zalloc(&Mw, nw);
        if (e == hipSuccess) e = zalloc(&Vw, nw);
        if (e == hipSuccess) e = zalloc(&b, m->pout[l]);
        if (e == hipSuccess) e = zalloc(&gb, m->pout[l]);
        if (e == hipSuccess) e = zalloc(&Mb, m->pout[l]);
        if (e == hipSuccess) e = zalloc(&Vb, m->pout[l]);
        if (e == hipSuccess) e = zalloc(&W16, nw);
        m->W.push_back(W); m->G.push_back(G); m->Mw.push_back(Mw); m->Vw.push_back(Vw);
        m->b.push_back(b); m->gb.push_back(gb); m->Mb.push_back(Mb); m->Vb.push_back(Vb);
        m->W16.push_back(W16);
        if (e != hipSuccess) break;
        std::vector<float> hw(nw, 0.0f);
        const double a = std::sqrt(6.0 / m->din[l]) * (l == m->nl - 1 ? 0.5 : 1.0);
        for (int r = 0; r < m->dout[l]; ++r)
            for (int c = 0; c < m->din[l]; ++c) hw[(size_t)r * m->pin[l] + c] = (float)((2.0 * rnd() - 1.0) * a);
        e = hipMemcpy(W, hw.data(), sizeof(float) * nw, hipMemcpyHostToDevice);
    }
    for (int l = 0; l < m->nl && e == hipSuccess; ++l) {
        uint16_t *h, *hT, *d16, *dT;
        float *zz, *dd, *dhh;
        e = zalloc(&h, (size_t)batch * m->pin[l]);
        if (e == hipSuccess) e = zalloc(&hT, (size_t)batch * m->pin[l]);
        if (e == hipSuccess) e = zalloc(&zz, (size_t)batch * m->pout[l]);
        if (e == hipSuccess) e = zalloc(&dd, (size_t)batch * m->pout[l]);
        if (e == hipSuccess) e = zalloc(&dhh, (size_t)batch * m->pin[l]);
        if (e == hipSuccess) e = zalloc(&d16, (size_t)batch * m->pout[l]);
        if (e == hipSuccess) e = zalloc(&dT, (size_t)batch * m->pout[l]);
        m->h16.push_back(h); m->hT16.push_back(hT); m->z.push_back(zz); m->dz.push_back(dd);
        m->dh.push_back(dhh); m->dz16.push_back(d16); m->dzT16.push_back(dT);
    }
    if (e == hipSuccess) e = zalloc(&m->y, (size_t)batch * m->pout.back());
    // room for the parts of the largest product split over K
    for (int l = 0; l < m->nl; ++l) {
        if (l > 0)
            m->part_n = std::max(m->part_n, (size_t)gemm_shape(batch, m->pin[l], m->pout[l], EPI_STORE).splits *
                                                batch * m->pin[l]);
        if (l < m->nl - 1)
            m->part_n = std::max(m->part_n, (size_t)gemm_shape(batch, m->pout[l], m->pin[l], EPI_STORE).splits *
                                                batch * m->pout[l]);
    }
    if (e == hipSuccess) e = zalloc(&m->part, m->part_n);
    if (e == hipSuccess) e = zalloc(&m->loss, 1);
    if (e == hipSuccess) e = zalloc(&m->idx, batch);
    // bf16 weights from the master copy
    for (int l = 0; l < m->nl && e == hipSuccess; ++l) {
        const size_t nw = (size_t)m->pout[l] * m->pin[l];
        hipLaunchKernelGGL(f32_to_bf16_kernel, dim3(1024), dim3(256), 0, m->st, nw, m->W[l], m->W16[l]);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(m->st);
    if (e != hipSuccess) {
        mlp_free(m);
        return mfail(SFG_E_HIP, "mlp create: %s", hipGetErrorString(e));
    }
    *out = m;
    return SFG_OK;
}

extern "C" void sfg_mlp_destroy(sfg_mlp *m)
{
    if (m) mlp_free(m);
}

extern "C" long long sfg_mlp_param_count(const sfg_mlp *m)
{
    long long c = 0;
    for (int l = 0; l < m->nl; ++l) c += (long long)m->dout[l] * m->din[l] + m->dout[l];
    return c;
}

// forward of the current batch (h16[0] holds its input, loaded by load_batch,
// which also zeroes *loss).  Without targets the output goes to y; with targets
// Y (rows by the batch's sample indices idx, on the device) the output layer's
// epilogue turns it into the loss gradient dz16 (and dz) of the last layer, adds
// the squared errors to *loss, and where it can forms the output bias's gradient
// (and with ad takes its Adam step).
static int forward(sfg_mlp *m, const float *Y = nullptr, const int *idx = nullptr, const AdamArgs *ad = nullptr)
{
    m->out_db = false;
    for (int l = 0; l < m->nl; ++l) {
        const bool last = l == m->nl - 1;
        if (!last) {
            const GemmShape gs = gemm_shape(m->batch, m->pout[l], m->pin[l], EPI_STORE);
            if (gs.splits > 1) {
                // few tiles and a long K: the product in parts over K, then bias + GELU
                EpiArgs ep;
                ep.slice = (size_t)m->batch * m->pout[l];
                MCHK(gemm(m->st, m->batch, m->pout[l], m->pin[l], m->h16[l], m->pin[l], m->W16[l], m->pin[l], m->part,
                          m->pout[l], ep));
                hipLaunchKernelGGL(bias_gelu_parts_kernel, dim3(1024), dim3(256), 0, m->st, m->batch, m->pout[l],
                                   m->part, gs.splits, ep.slice, m->b[l], m->z[l], m->h16[l + 1]);
                MCHK(hipGetLastError());
                continue;
            }
        }
        EpiArgs ep;
        ep.bias = m->b[l];
        float *C = m->z[l];
        if (!last) {
            ep.kind = EPI_BIAS_GELU;
            ep.H = m->h16[l + 1];
            ep.ldh = m->pin[l + 1];
        } else if (!Y) {
            ep.kind = EPI_BIAS;
            C = m->y;
        } else {
            ep.kind = EPI_MSE;
            C = m->dz[l];
            ep.H = m->dz16[l];
            ep.ldh = m->pout[l];
            ep.T = Y;
            ep.tidx = idx;
            ep.ldt = m->nout;
            ep.nvalid = m->nout;
            ep.scale = 2.0f / ((float)m->batch * (float)m->nout);
            ep.loss = m->loss;
            // one block holds the whole batch: the bias gradient (and with ad the
            // bias's Adam step) in the same epilogue, the FP32 gradient not stored
            if (m->batch <= gemm_shape(m->batch, m->pout[l], m->pin[l], EPI_MSE).bm && m->nout % 4 == 0) {
                ep.db = m->gb[l];
                if (ad) {
                    ep.bw = m->b[l];
                    ep.bm = m->Mb[l];
                    ep.bv = m->Vb[l];
                    ep.ad = *ad;
                }
                m->out_db = true;
            }
        }
        MCHK(gemm(m->st, m->batch, m->pout[l], m->pin[l], m->h16[l], m->pin[l], m->W16[l], m->pin[l], C, m->pout[l],
                  ep));
    }
    return SFG_OK;
}

// backward from dz of the last layer (in dz[nl-1] / dz16[nl-1]): weight gradients
// to G, bias gradients to gb.  With ad each layer's bias is also updated by Adam
// as its gradient is formed, and its weights right after their gradient (a
// layer's dX product reads its bf16 weights before they change).
static int backward(sfg_mlp *m, const AdamArgs *ad = nullptr)
{
    const int B = m->batch;
    // without an Adam step, the dW products split over K (narrow layers) add into
    // their outputs: all of those are zeroed by one launch up front (the dX
    // products split over K store their parts, which gelu_back adds up)
    ZeroList zl{};
    std::vector<char> zw(m->nl, 0);
    for (int l = m->nl - 1; l >= 0 && zl.count < 8; --l) {
        if (!ad && gemm_shape(m->pout[l], m->pin[l], B, EPI_STORE).splits > 1) {
            zl.p[zl.count] = m->G[l];
            zl.n[zl.count++] = (size_t)m->pout[l] * m->pin[l];
            zw[l] = 1;
        }
    }
    if (zl.count) {
        hipLaunchKernelGGL(zero_list_kernel, dim3(256, zl.count), dim3(256), 0, m->st, zl);
        MCHK(hipGetLastError());
    }
    for (int l = m->nl - 1; l >= 0; --l) {
        if (l > 0) {
            // dh = dz . W (W [pout][pin] read k-major as stored) ; dz_{l-1} = dh * gelu'(z_{l-1});
            // split over K, the product's parts go to m->part and gelu_back adds them up
            const GemmShape gs = gemm_shape(B, m->pin[l], m->pout[l], EPI_STORE);
            const size_t n = (size_t)B * m->pin[l];
            EpiArgs ep;
            float *dh = m->dh[l];
            if (gs.splits > 1) {
                ep.slice = n;
                dh = m->part;
            }
            MCHK(gemm(m->st, B, m->pin[l], m->pout[l], m->dz16[l], m->pout[l], m->W16[l], m->pin[l], dh, m->pin[l], ep,
                      true));
            hipLaunchKernelGGL(gelu_back_kernel, dim3(1024), dim3(256), 0, m->st, n, dh, gs.splits, ep.slice,
                               m->z[l - 1], m->dz[l - 1], m->dz16[l - 1]);
            MCHK(hipGetLastError());
        }
        if (!(l == m->nl - 1 && m->out_db)) { // (else formed in the forward pass's loss epilogue)
            hipLaunchKernelGGL(bias_grad_kernel, dim3((m->pout[l] + 63) / 64), dim3(256), 0, m->st, B, m->dout[l],
                               m->pout[l], m->dz[l], m->gb[l], ad ? m->b[l] : nullptr, m->Mb[l], m->Vb[l],
                               ad ? *ad : AdamArgs{});
            MCHK(hipGetLastError());
        }
        // dW = dz^T . h from transposed copies (A = dz^T [pout][B], B = h^T [pin][B]); K = batch.
        // Reading both operands k-major instead measured 2.4x slower (transposing into LDS
        // element by element costs as much as the products at this small K)
        MCHK(transpose2(m->st, Transpose{B, m->pout[l], m->dz16[l], m->pout[l], m->dzT16[l], B},
                        Transpose{B, m->pin[l], m->h16[l], m->pin[l], m->hT16[l], B}));
        if (ad) {
            // the weights' Adam step in the dW product's epilogue: the gradient is never stored
            EpiArgs ep;
            ep.kind = EPI_ADAM;
            ep.W = m->W[l];
            ep.Mw = m->Mw[l];
            ep.Vw = m->Vw[l];
            ep.W16 = m->W16[l];
            ep.arows = m->dout[l];
            ep.acols = m->din[l];
            ep.ad = *ad;
            MCHK(gemm(m->st, m->pout[l], m->pin[l], B, m->dzT16[l], B, m->hT16[l], B, m->G[l], m->pin[l], ep));
        } else {
            MCHK(gemm(m->st, m->pout[l], m->pin[l], B, m->dzT16[l], B, m->hT16[l], B, m->G[l], m->pin[l], EpiArgs{},
                      false, zw[l]));
        }
    }
    return SFG_OK;
}

// sample indices (host, range-checked) -> m->idx
static int upload_idx(sfg_mlp *m, const int *idx_host, int nsrc)
{
    if (nsrc <= 0) return mfail(SFG_E_ARG, "mlp: no data");
    for (int b = 0; b < m->batch; ++b)
        if (idx_host[b] < 0 || idx_host[b] >= nsrc) return mfail(SFG_E_ARG, "mlp: sample index out of range");
    MCHK(hipMemcpyAsync(m->idx, idx_host, sizeof(int) * m->batch, hipMemcpyHostToDevice, m->st));
    return SFG_OK;
}

// rows idx[batch] (device) of X -> bf16 input h16[0]
static int load_batch(sfg_mlp *m, const int *idx, const float *X)
{
    hipLaunchKernelGGL(gather_rows_kernel, dim3(1, m->batch), dim3(64), 0, m->st, m->batch, m->nin, m->pin[0], X,
                       idx, m->h16[0], (float *)nullptr, m->loss);
    MCHK(hipGetLastError());
    return SFG_OK;
}

// Adam's step count and bias corrections for the next update
static AdamArgs next_adam(sfg_mlp *m, float lr)
{
    const float b1 = 0.9f, b2 = 0.999f;
    m->step += 1;
    return AdamArgs{lr, b1, b2, 1e-8f, 1.0f - powf(b1, (float)m->step), 1.0f - powf(b2, (float)m->step)};
}

static int read_loss(sfg_mlp *m, double *loss)
{
    double h = 0.0;
    MCHK(hipMemcpyAsync(&h, m->loss, sizeof(double), hipMemcpyDeviceToHost, m->st));
    MCHK(hipStreamSynchronize(m->st));
    *loss = h / ((double)m->batch * m->nout);
    return SFG_OK;
}

extern "C" int sfg_mlp_set_data(sfg_mlp *m, const float *X, const float *Y, int n)
{
    g_mlp_err.clear();
    if (!m || !X || !Y || n <= 0) return mfail(SFG_E_ARG, "mlp data: bad arguments");
    MCHK(hipStreamSynchronize(m->st));
    if (m->Xd) (void)hipFree(m->Xd);
    if (m->Yd) (void)hipFree(m->Yd);
    if (m->order) (void)hipFree(m->order);
    m->Xd = m->Yd = nullptr;
    m->order = nullptr;
    MCHK(hipMalloc((void **)&m->order, sizeof(int) * (size_t)n));
    MCHK(hipMalloc((void **)&m->Xd, sizeof(float) * (size_t)n * m->nin));
    MCHK(hipMalloc((void **)&m->Yd, sizeof(float) * (size_t)n * m->nout));
    MCHK(hipMemcpy(m->Xd, X, sizeof(float) * (size_t)n * m->nin, hipMemcpyHostToDevice));
    MCHK(hipMemcpy(m->Yd, Y, sizeof(float) * (size_t)n * m->nout, hipMemcpyHostToDevice));
    m->ndata = n;
    return SFG_OK;
}

// One Adam step on the samples idx[batch] of the resident data; returns the batch loss.
extern "C" int sfg_mlp_train_step(sfg_mlp *m, const int *idx, float lr, double *loss)
{
    g_mlp_err.clear();
    if (!m || !idx || !m->Xd) return mfail(SFG_E_ARG, "mlp train: no data");
    int rc = upload_idx(m, idx, m->ndata);
    if (!rc) rc = load_batch(m, m->idx, m->Xd);
    AdamArgs ad{};
    if (!rc) ad = next_adam(m, lr);
    if (!rc) rc = forward(m, m->Yd, m->idx, &ad);
    if (!rc) rc = backward(m, &ad);
    double h = 0.0;
    if (!rc) rc = read_loss(m, &h); // also ends the host index buffer's use
    if (!rc && loss) *loss = h;
    return rc;
}

// Runs `epochs` epochs over the resident data with a host-shuffled order,
// cosine-decayed learning rate from lr0 to 0 over total_steps (0 = this call).
extern "C" int sfg_mlp_train(sfg_mlp *m, int epochs, float lr0, long long total_steps, unsigned long long seed,
                             double *last_loss, double *ms_per_step)
{
    g_mlp_err.clear();
    if (!m || !m->Xd || epochs <= 0) return mfail(SFG_E_ARG, "mlp train: bad arguments");
    const int n = m->ndata, B = m->batch;
    const int per_epoch = n / B;
    if (per_epoch <= 0) return mfail(SFG_E_ARG, "mlp train: fewer samples than one batch");
    const long long T = total_steps > 0 ? total_steps : (long long)per_epoch * epochs;
    std::vector<int> order(n);
    uint64_t s = seed * 0x9E3779B97F4A7C15ull + 7;
    hipEvent_t e0, e1;
    MCHK(hipEventCreate(&e0));
    MCHK(hipEventCreate(&e1));
    MCHK(hipEventRecord(e0, m->st));
    long long done = 0;
    for (int ep = 0; ep < epochs; ++ep) {
        for (int i = 0; i < n; ++i) order[i] = i;
        for (int i = n - 1; i > 0; --i) {
            s ^= s >> 12; s ^= s << 25; s ^= s >> 27;
            const int j = (int)((s * 2685821657736338717ull) % (uint64_t)(i + 1));
            std::swap(order[i], order[j]);
        }
        // the epoch's order goes up once; the host copy is reused only after the
        // stream has drained at the end of the epoch
        MCHK(hipMemcpyAsync(m->order, order.data(), sizeof(int) * (size_t)n, hipMemcpyHostToDevice, m->st));
        for (int k = 0; k < per_epoch; ++k) {
            const double prog = std::min(1.0, (double)(m->step) / (double)T);
            const float lr = (float)(lr0 * 0.5 * (1.0 + std::cos(M_PI * prog)));
            const int *idx = m->order + (size_t)k * B;
            int rc = load_batch(m, idx, m->Xd);
            AdamArgs ad{};
            if (!rc) ad = next_adam(m, lr);
            if (!rc) rc = forward(m, m->Yd, idx, &ad);
            if (!rc) rc = backward(m, &ad);
            if (rc) return rc;
            ++done;
        }
        MCHK(hipStreamSynchronize(m->st));
    }
    MCHK(hipEventRecord(e1, m->st));
    MCHK(hipEventSynchronize(e1));
    float ms = 0.f;
    MCHK(hipEventElapsedTime(&ms, e0, e1));
    if (ms_per_step) *ms_per_step = done ? ms / done : 0.0;
    if (last_loss) {
        double h = 0.0;
        MCHK(hipMemcpy(&h, m->loss, sizeof(double), hipMemcpyDeviceToHost));
        *last_loss = h / ((double)B * m->nout);
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return SFG_OK;
}

// bf16 inference: Y[n][nout] = net(X[n][nin]) (host arrays), batch by batch
extern "C" int sfg_mlp_predict(sfg_mlp *m, const float *X, int n, float *Y, double *ms_total)
{
    g_mlp_err.clear();
    if (!m || !X || !Y || n <= 0) return mfail(SFG_E_ARG, "mlp predict: bad arguments");
    const int B = m->batch, L = m->nl - 1;
    float *Xd = nullptr;
    MCHK(hipMalloc((void **)&Xd, sizeof(float) * (size_t)n * m->nin));
    MCHK(hipMemcpy(Xd, X, sizeof(float) * (size_t)n * m->nin, hipMemcpyHostToDevice));
    std::vector<int> idx(B);
    std::vector<float> yb((size_t)B * m->pout[L]);
    hipEvent_t e0, e1;
    MCHK(hipEventCreate(&e0));
    MCHK(hipEventCreate(&e1));
    float total = 0.f;
    for (int s0 = 0; s0 < n; s0 += B) {
        const int cnt = std::min(B, n - s0);
        for (int b = 0; b < B; ++b) idx[b] = s0 + std::min(b, cnt - 1);
        int rc = upload_idx(m, idx.data(), n);
        if (!rc) rc = load_batch(m, m->idx, Xd);
        MCHK(hipEventRecord(e0, m->st));
        if (!rc) rc = forward(m);
        MCHK(hipEventRecord(e1, m->st));
        if (rc) { (void)hipFree(Xd); return rc; }
        MCHK(hipMemcpyAsync(yb.data(), m->y, sizeof(float) * yb.size(), hipMemcpyDeviceToHost, m->st));
        MCHK(hipStreamSynchronize(m->st));
        float ms = 0.f;
        MCHK(hipEventElapsedTime(&ms, e0, e1));
        total += ms;
        for (int b = 0; b < cnt; ++b)
            std::memcpy(Y + (size_t)(s0 + b) * m->nout, yb.data() + (size_t)b * m->pout[L], sizeof(float) * m->nout);
    }
    (void)hipFree(Xd);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (ms_total) *ms_total = total;
    return SFG_OK;
}

// Loss and gradients for one given batch without updating (gradient checks).
// Xb [batch][nin], Yb [batch][nout] host; grads of layer l: dW [dout][din], db [dout]
extern "C" int sfg_mlp_gradients(sfg_mlp *m, const float *Xb, const float *Yb, double *loss, int layer, float *dW,
                                 float *db)
{
    g_mlp_err.clear();
    if (!m || !Xb || !Yb || layer < 0 || layer >= m->nl) return mfail(SFG_E_ARG, "mlp grads: bad arguments");
    float *Xd = nullptr, *Yd = nullptr;
    MCHK(hipMalloc((void **)&Xd, sizeof(float) * (size_t)m->batch * m->nin));
    MCHK(hipMalloc((void **)&Yd, sizeof(float) * (size_t)m->batch * m->nout));
    MCHK(hipMemcpy(Xd, Xb, sizeof(float) * (size_t)m->batch * m->nin, hipMemcpyHostToDevice));
    MCHK(hipMemcpy(Yd, Yb, sizeof(float) * (size_t)m->batch * m->nout, hipMemcpyHostToDevice));
    std::vector<int> idx(m->batch);
    for (int b = 0; b < m->batch; ++b) idx[b] = b;
    int rc = upload_idx(m, idx.data(), m->batch);
    if (!rc) rc = load_batch(m, m->idx, Xd);
    if (!rc) rc = forward(m, Yd, m->idx);
    if (!rc) rc = backward(m);
    double h = 0.0;
    if (!rc) rc = read_loss(m, &h);
    if (!rc && loss) *loss = h;
    if (!rc) {
        std::vector<float> g((size_t)m->pout[layer] * m->pin[layer]), gb(m->pout[layer]);
        hipError_t e = hipStreamSynchronize(m->st);
        if (e == hipSuccess) e = hipMemcpy(g.data(), m->G[layer], sizeof(float) * g.size(), hipMemcpyDeviceToHost);
        if (e == hipSuccess) e = hipMemcpy(gb.data(), m->gb[layer], sizeof(float) * gb.size(), hipMemcpyDeviceToHost);
        if (e != hipSuccess) rc = mfail(SFG_E_HIP, "mlp grads: %s", hipGetErrorString(e));
        for (int r = 0; !rc && r < m->dout[layer]; ++r) {
            if (dW) std::memcpy(dW + (size_t)r * m->din[layer], g.data() + (size_t)r * m->pin[layer],
                                sizeof(float) * m->din[layer]);
            if (db) db[r] = gb[r];
        }
    }
    (void)hipFree(Xd);
    (void)hipFree(Yd);
    return rc;
}

// Master weights of layer l -> W [dout][din], b [dout] (host)
extern "C" int sfg_mlp_get_layer(sfg_mlp *m, int layer, float *W, float *b)
{
    g_mlp_err.clear();
    if (!m || layer < 0 || layer >= m->nl) return mfail(SFG_E_ARG, "mlp get: bad layer");
    std::vector<float> g((size_t)m->pout[layer] * m->pin[layer]), gb(m->pout[layer]);
    MCHK(hipStreamSynchronize(m->st));
    MCHK(hipMemcpy(g.data(), m->W[layer], sizeof(float) * g.size(), hipMemcpyDeviceToHost));
    MCHK(hipMemcpy(gb.data(), m->b[layer], sizeof(float) * gb.size(), hipMemcpyDeviceToHost));
    for (int r = 0; r < m->dout[layer]; ++r) {
        if (W) std::memcpy(W + (size_t)r * m->din[layer], g.data() + (size_t)r * m->pin[layer],
                           sizeof(float) * m->din[layer]);
        if (b) b[r] = gb[r];
    }
    return SFG_OK;
}

// W [dout][din], b [dout] (host) -> master weights of layer l and their bf16 copy
// (the Adam moments are left as they are)
extern "C" int sfg_mlp_set_layer(sfg_mlp *m, int layer, const float *W, const float *b)
{
    g_mlp_err.clear();
    if (!m || layer < 0 || layer >= m->nl || !W || !b) return mfail(SFG_E_ARG, "mlp set: bad arguments");
    std::vector<float> g((size_t)m->pout[layer] * m->pin[layer], 0.0f), gb(m->pout[layer], 0.0f);
    for (int r = 0; r < m->dout[layer]; ++r) {
        std::memcpy(g.data() + (size_t)r * m->pin[layer], W + (size_t)r * m->din[layer],
                    sizeof(float) * m->din[layer]);
        gb[r] = b[r];
    }
    MCHK(hipStreamSynchronize(m->st));
    MCHK(hipMemcpy(m->W[layer], g.data(), sizeof(float) * g.size(), hipMemcpyHostToDevice));
    MCHK(hipMemcpy(m->b[layer], gb.data(), sizeof(float) * gb.size(), hipMemcpyHostToDevice));
    hipLaunchKernelGGL(f32_to_bf16_kernel, dim3(1024), dim3(256), 0, m->st, g.size(), m->W[layer], m->W16[layer]);
    MCHK(hipGetLastError());
    MCHK(hipStreamSynchronize(m->st));
    return SFG_OK;
}

// Plain bf16 GEMM C = A . B^T on host arrays (A [M][K], B [N][K] as fp32, rounded
// to bf16 here; K padded to 32 inside).  For kernel tests.
extern "C" int sfg_gemm_bf16_nt(int M, int N, int K, const float *A, const float *B, float *C)
{
    g_mlp_err.clear();
    if (M <= 0 || N <= 0 || K <= 0 || !A || !B || !C) return mfail(SFG_E_ARG, "gemm: bad arguments");
    const int Kp = pad32(K);
    std::vector<float> ap((size_t)M * Kp, 0.f), bp((size_t)N * Kp, 0.f);
    for (int i = 0; i < M; ++i) std::memcpy(&ap[(size_t)i * Kp], A + (size_t)i * K, sizeof(float) * K);
    for (int i = 0; i < N; ++i) std::memcpy(&bp[(size_t)i * Kp], B + (size_t)i * K, sizeof(float) * K);
    float *da = nullptr, *db = nullptr, *dc = nullptr;
    uint16_t *ha = nullptr, *hb = nullptr;
    MCHK(hipMalloc((void **)&da, sizeof(float) * ap.size()));
    MCHK(hipMalloc((void **)&db, sizeof(float) * bp.size()));
    MCHK(hipMalloc((void **)&dc, sizeof(float) * (size_t)M * N));
    MCHK(hipMalloc((void **)&ha, sizeof(uint16_t) * ap.size()));
    MCHK(hipMalloc((void **)&hb, sizeof(uint16_t) * bp.size()));
    MCHK(hipMemcpy(da, ap.data(), sizeof(float) * ap.size(), hipMemcpyHostToDevice));
    MCHK(hipMemcpy(db, bp.data(), sizeof(float) * bp.size(), hipMemcpyHostToDevice));
    hipLaunchKernelGGL(f32_to_bf16_kernel, dim3(256), dim3(256), 0, nullptr, ap.size(), da, ha);
    hipLaunchKernelGGL(f32_to_bf16_kernel, dim3(256), dim3(256), 0, nullptr, bp.size(), db, hb);
    MCHK(gemm(nullptr, M, N, Kp, ha, Kp, hb, Kp, dc, N, EpiArgs{}));
    MCHK(hipMemcpy(C, dc, sizeof(float) * (size_t)M * N, hipMemcpyDeviceToHost));
    for (void *p : {(void *)da, (void *)db, (void *)dc, (void *)ha, (void *)hb}) (void)hipFree(p);
    return SFG_OK;
}

// pseudo-random bf16 in [-1, 1) (hash of the index)
__global__ void fill_bf16_kernel(size_t n, uint32_t seed, uint16_t *p)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        uint32_t h = (uint32_t)i * 2654435761u ^ seed;
        h ^= h >> 15; h *= 2246822519u; h ^= h >> 13;
        p[i] = f2bf((float)(h >> 8) * (2.0f / 16777216.0f) - 1.0f);
    }
}

extern "C" int sfg_gemm_bf16_time(int M, int N, int K, int reps, double *ms)
{
    g_mlp_err.clear();
    if (M <= 0 || N <= 0 || K <= 0 || K % 32 || reps <= 0 || !ms) return mfail(SFG_E_ARG, "gemm time: bad arguments");
    uint16_t *a = nullptr, *b = nullptr;
    float *c = nullptr;
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t e = hipMalloc((void **)&a, sizeof(uint16_t) * (size_t)M * K);
    if (e == hipSuccess) e = hipMalloc((void **)&b, sizeof(uint16_t) * (size_t)N * K);
    if (e == hipSuccess) e = hipMalloc((void **)&c, sizeof(float) * (size_t)M * N);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(fill_bf16_kernel, dim3(1024), dim3(256), 0, st, (size_t)M * K, 1u, a);
        hipLaunchKernelGGL(fill_bf16_kernel, dim3(1024), dim3(256), 0, st, (size_t)N * K, 2u, b);
        e = gemm(st, M, N, K, a, K, b, K, c, N, EpiArgs{}); // warm-up
    }
    if (e == hipSuccess) e = hipEventRecord(e0, st);
    for (int r = 0; r < reps && e == hipSuccess; ++r) e = gemm(st, M, N, K, a, K, b, K, c, N, EpiArgs{});
    if (e == hipSuccess) e = hipEventRecord(e1, st);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float t = 0.0f;
    if (e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
    if (e == hipSuccess) *ms = (double)t / reps;
    if (e1) (void)hipEventDestroy(e1);
    if (e0) (void)hipEventDestroy(e0);
    if (st) (void)hipStreamDestroy(st);
    for (void *p : {(void *)a, (void *)b, (void *)c})
        if (p) (void)hipFree(p);
    if (e != hipSuccess) return mfail(SFG_E_HIP, "gemm time: %s", hipGetErrorString(e));
    return SFG_OK;
}
