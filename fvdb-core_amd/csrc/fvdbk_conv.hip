// Sparse-convolution contraction for gfx950: implicit GEMM through a kernel map on
// MFMA (bf16/f16 with v_mfma_f32_16x16x32_*, fp32 with v_mfma_f32_16x16x4f32), a
// scalar fp64 form, and the split-K weight gradient with a compensated reduction.
#include <hip/hip_runtime.h>

#include <mutex>

#include "fvdbk_common.h"
#include "fvdbk_conv_frag.h"
#include "fvdbk_conv_tile.h"

namespace fvdbk {

// ---------------------------------------------------------------------------------
// forward / dgrad implicit GEMM on MFMA
//
// D[m][r] = sum_k sum_c a[k][m][c] * in[map[r][k]][c]  is computed as W^T X^T per
// 16x16 MFMA tile: the A operand is a 16(m) x CH(c) slice of a[k], the B operand the
// gathered CH(c) x 16(rows) slice of the input, each lane loading 16 contiguous bytes
// of one row (weights: row m, inputs: gathered voxel row).  A workgroup is 4
// wavefronts; wave w owns NI*16 output rows and all MI*16 output channels of the
// workgroup's m-tile.  The kernel-map entries are read per offset; an offset whose
// rows are all -1 in a wavefront is skipped by the whole wavefront.
// ---------------------------------------------------------------------------------

constexpr int CONV_WAVES = 4;

template <typename T, int MI, int NI, bool VEC>
__global__ __launch_bounds__(256) void k_conv_mfma(const T* __restrict__ X, int cin, const T* __restrict__ A,
                                                   const int32_t* __restrict__ map, int64_t n_rows, int K, int M,
                                                   T* __restrict__ Y, const T* __restrict__ bias, int relu) {
    using F = Frag<T>;
    using V = typename F::vec;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int g = lane >> 4, r16 = lane & 15;
    const int64_t row_w = ((int64_t)blockIdx.x * CONV_WAVES + wave) * (NI * 16);
    const int m0 = blockIdx.y * (MI * 16);

    floatx4 acc[MI][NI];
#pragma unroll
    for (int i = 0; i < MI; ++i)
#pragma unroll
        for (int j = 0; j < NI; ++j) acc[i][j] = floatx4{0.f, 0.f, 0.f, 0.f};

    int64_t row[NI];
#pragma unroll
    for (int j = 0; j < NI; ++j) row[j] = row_w + j * 16 + r16;

    for (int k = 0; k < K; ++k) {
        int32_t src[NI];
        bool any = false;
#pragma unroll
        for (int j = 0; j < NI; ++j) {
            src[j] = row[j] < n_rows ? map[row[j] * K + k] : -1;
            any |= src[j] >= 0;
        }
        if (!__any(any)) continue;
        const T* __restrict__ arow[MI];
        bool mval[MI];
#pragma unroll
        for (int i = 0; i < MI; ++i) {
            const int m = m0 + i * 16 + r16;
            mval[i] = m < M;
            arow[i] = A + ((int64_t)k * M + (mval[i] ? m : 0)) * cin;
        }
        const T* __restrict__ xrow[NI];
#pragma unroll
        for (int j = 0; j < NI; ++j) xrow[j] = X + (int64_t)(src[j] >= 0 ? src[j] : 0) * cin;

        for (int c0 = 0; c0 < cin; c0 += F::CH) {
            const int c = c0 + g * F::EPL;
            V b[NI], a[MI];
#pragma unroll
            for (int j = 0; j < NI; ++j) b[j] = src[j] >= 0 ? load_frag<T, VEC>(xrow[j], c, cin) : zero_frag<T>();
#pragma unroll
            for (int i = 0; i < MI; ++i) a[i] = mval[i] ? load_frag<T, VEC>(arow[i], c, cin) : zero_frag<T>();
#pragma unroll
            for (int i = 0; i < MI; ++i)
#pragma unroll
                for (int j = 0; j < NI; ++j) mfma(a[i], b[j], acc[i][j]);
        }
    }

    // lane (g, r16) of tile (i, j) holds D[m = 4g + e][r = r16], e = 0..3
#pragma unroll
    for (int j = 0; j < NI; ++j) {
        if (row[j] >= n_rows) continue;
        T* __restrict__ y = Y + row[j] * M;
#pragma unroll
        for (int i = 0; i < MI; ++i) {
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int m = m0 + i * 16 + 4 * g + e;
                if (m < M) y[m] = epilogue<T>(acc[i][j][e], bias, m, relu);
            }
        }
    }
}

// Scalar form for any dtype (fp64 accumulates in fp64): one thread per output element.
template <typename T, typename Acc>
__global__ void k_conv_generic(const T* __restrict__ X, int cin, const T* __restrict__ A,
                               const int32_t* __restrict__ map, int64_t n_rows, int K, int M, T* __restrict__ Y,
                               const T* __restrict__ bias, int relu) {
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= n_rows * M) return;
    const int64_t r = idx / M;
    const int m = (int)(idx - r * M);
    Acc acc = 0;
    for (int k = 0; k < K; ++k) {
        const int32_t s = map[r * K + k];
        if (s < 0) continue;
        const T* __restrict__ x = X + (int64_t)s * cin;
        const T* __restrict__ a = A + ((int64_t)k * M + m) * cin;
        for (int c = 0; c < cin; ++c) acc += (Acc)x[c] * (Acc)a[c];
    }
    if (bias) acc += (Acc)bias[m];
    if (relu && !(acc > (Acc)0) && acc == acc) acc = 0;
    Y[idx] = (T)acc;
}

template <typename T, int MI, bool VEC>
static void launch_conv_mfma(const T* X, int cin, const T* A, const int32_t* map, int64_t n_rows, int K, int M, T* Y,
                             const T* bias, int relu, hipStream_t s) {
    constexpr int NI = 2;
    const dim3 grid((unsigned)((n_rows + CONV_WAVES * NI * 16 - 1) / (CONV_WAVES * NI * 16)),
                    (unsigned)((M + MI * 16 - 1) / (MI * 16)));
    hipLaunchKernelGGL((k_conv_mfma<T, MI, NI, VEC>), grid, dim3(256), 0, s, X, cin, A, map, n_rows, K, M, Y, bias,
                       relu);
}

template <typename T>
static int conv_mfma_dispatch(const T* X, int64_t n_in, int cin, const T* A, const int32_t* map, const int32_t* perm,
                              int64_t n_rows, int K, int M, T* Y, const T* bias, int relu, hipStream_t s) {
    if (conv_tile_dispatch(X, n_in, cin, A, map, perm, n_rows, K, M, Y, bias, relu, s)) {
        FVDBK_LAUNCH_CHECK("k_conv_tile");
        return FVDBK_OK;
    }
    const bool vec = (cin % Frag<T>::EPL) == 0 && ((uintptr_t)X % 16) == 0 && ((uintptr_t)A % 16) == 0;
    const int mi = M <= 16 ? 1 : (M <= 32 ? 2 : 4);
    if (vec) {
        if (mi == 1) launch_conv_mfma<T, 1, true>(X, cin, A, map, n_rows, K, M, Y, bias, relu, s);
        else if (mi == 2) launch_conv_mfma<T, 2, true>(X, cin, A, map, n_rows, K, M, Y, bias, relu, s);
        else launch_conv_mfma<T, 4, true>(X, cin, A, map, n_rows, K, M, Y, bias, relu, s);
    } else {
        if (mi == 1) launch_conv_mfma<T, 1, false>(X, cin, A, map, n_rows, K, M, Y, bias, relu, s);
        else if (mi == 2) launch_conv_mfma<T, 2, false>(X, cin, A, map, n_rows, K, M, Y, bias, relu, s);
        else launch_conv_mfma<T, 4, false>(X, cin, A, map, n_rows, K, M, Y, bias, relu, s);
    }
    FVDBK_LAUNCH_CHECK("k_conv_mfma");
    return FVDBK_OK;
}

// ---------------------------------------------------------------------------------
// weight gradient: split-K over the neighbour-map pairs of each offset
//
// gw[k][c][o] = sum_p in[pi][c] * gout[po][o].  Offset k's pairs (ordered by output row)
// are cut into rounds of WG_ROWS pairs and its n_k = chunk_start[k+1] - chunk_start[k]
// partials take them interleaved: partial s sums rounds s, s + n_k, s + 2 n_k, ...  With
// n_k proportional to the offset's pair count, the workgroups of every offset then sweep
// the output rows together, so at any time they share a window of a few thousand rows
// (the rows' next readers are the other offsets' workgroups, a few microseconds later)
// instead of each reading its own stretch of the grid once: the 16-bit 128 x 128 layer of
// 1.1M rows fetched 3.4 GB from HBM per launch with contiguous chunks (2% L2 hits).
// A workgroup takes one partial of one offset and a (c, o) tile; rows of both operands are gathered into LDS
// (16-byte loads), then read back as MFMA fragments whose reduction index is the
// pair: for 16-bit types lane group g, element e takes pair row  step*CH + 8g + e  of
// the stage through LDS transpose reads (two 8-byte reads per fragment); for fp32
// element e takes row  step*CH + 4e + g.  A and B use the same assignment, so every
// pair meets itself.  Partials are fp32.
// ---------------------------------------------------------------------------------

constexpr int WG_ROWS = 64;  // pairs staged per LDS round

// Segment swizzle of the 16-bit wgrad tiles (rows of NSEG 16-byte segments): the lanes
// of one transpose read cover rows st*32 + 8g + (0..3) of a stage, 8 consecutive banks
// per row; the swizzle puts the rows read by each half-wavefront in distinct 8-bank
// groups.  It is even (a lane's two segments stay adjacent) and depends on row bits 0,
// 1 and 3 only, so rows r and r + 4 share it.
template <int NSEG>
__device__ __host__ constexpr int wg_swizzle(int r) {
    return NSEG >= 16 ? 2 * ((r & 3) | (((r >> 3) & 1) << 2))
           : NSEG == 8 ? 2 * (((r >> 1) & 1) | (((r >> 3) & 1) << 1))
           : NSEG == 4 ? 2 * ((r >> 3) & 1)
                       : 0;
}

// Software pipeline: each thread owns fixed (pair row, channel group) slots of the
// staged tiles.  While round r is contracted from LDS, the rows of round r+1 are in flight
// into registers (their pair indices were loaded one round earlier) and the pair indices
// of round r+2 are loaded; a round's rows are written to LDS at its start.  The 16-bit
// gathers come mostly from the last-level cache, several microseconds away under load,
// so the rate is set by the bytes each CU keeps in flight: two workgroups of 32 KB rounds
// per CU (a third register set for a second round ahead costs half the occupancy and
// measured slower).
template <typename T, int MI, int NI, bool VEC>
__global__ __launch_bounds__(256) void k_wgrad_pipe(const T* __restrict__ X, int cin, const T* __restrict__ G,
                                                    int cout, const int32_t* __restrict__ pairs,
                                                    const int64_t* __restrict__ k_offsets,
                                                    const int32_t* __restrict__ chunk_start, int K,
                                                    int transposed, float* __restrict__ ws) {
    using F = Frag<T>;
    using V = typename F::vec;
    constexpr int TM = 2 * MI * 16;
    constexpr int TN = 2 * NI * 16;
    // 16-bit tiles: unpadded rows, 16-byte segment s of row r stored at s ^ wg_swizzle(r)
    // (bank-conflict-free transpose reads); fp32 tiles: rows padded by 32 bytes
    constexpr bool TR = sizeof(T) == 2;
    constexpr int PADB = TR ? 0 : 32;
    constexpr int XP = TM + PADB / (int)sizeof(T);
    constexpr int GP = TN + PADB / (int)sizeof(T);
    constexpr int XSEG = TM / F::EPL, GSEG = TN / F::EPL;
    constexpr int XPT = (WG_ROWS * XSEG + 255) / 256, GPT = (WG_ROWS * GSEG + 255) / 256;
    __shared__ __attribute__((aligned(16))) T sX[WG_ROWS * XP];
    __shared__ __attribute__((aligned(16))) T sG[WG_ROWS * GP];

    const int id = blockIdx.x;
    int lo = 0, hi = K;  // chunk_start[lo] <= id < chunk_start[hi]
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (chunk_start[mid] <= id) lo = mid;
        else hi = mid;
    }
    const int k = lo;
    // partial s of offset k takes rounds s, s + n_k, ...: round j starts at pair p0 + j * stride
    const int64_t stride = (int64_t)(chunk_start[k + 1] - chunk_start[k]) * WG_ROWS;
    const int64_t p0 = k_offsets[k] + (int64_t)(id - chunk_start[k]) * WG_ROWS;
    const int64_t p1 = k_offsets[k + 1];

    const int n_ct = (cin + TM - 1) / TM;
    const int c0 = (blockIdx.y % n_ct) * TM;
    const int o0 = (blockIdx.y / n_ct) * TN;
    const int ci = transposed ? 1 : 0, co = 1 - ci;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int g = lane >> 4, r16 = lane & 15;
    const int wm = (wave & 1) * MI * 16, wn = (wave >> 1) * NI * 16;

    floatx4 acc[MI][NI];
#pragma unroll
    for (int i = 0; i < MI; ++i)
#pragma unroll
        for (int j = 0; j < NI; ++j) acc[i][j] = floatx4{0.f, 0.f, 0.f, 0.f};

    // VEC (channel counts multiples of the 16-byte fragment, 16-byte aligned rows): every
    // round issues the same loads, an entry past the partial's last pair (or past the
    // channels) loading row 0 and being zeroed when stored, so there are no branches around
    // loads and the wait before a round's stores covers exactly its rows.
    int32_t xi[XPT], gi[GPT];  // pair indices of the round after next
    V xv[XPT], gv[GPT];        // rows of the next round
    uint32_t vmask;            // bit q (X) / XPT + q (G): the entry is a real pair row
    auto load_idx = [&](int64_t p, int32_t (&xx)[XPT], int32_t (&gg)[GPT]) __attribute__((always_inline)) {
#pragma unroll
        for (int q = 0; q < XPT; ++q) {
            const int rr = (tid + 256 * q) / XSEG;
            const int64_t pp = p + rr < p1 ? p + rr : p1 - 1;
            xx[q] = pairs[2 * pp + ci];
        }
#pragma unroll
        for (int q = 0; q < GPT; ++q) {
            const int rr = (tid + 256 * q) / GSEG;
            const int64_t pp = p + rr < p1 ? p + rr : p1 - 1;
            gg[q] = pairs[2 * pp + co];
        }
    };
    auto load_rows = [&](int64_t p, const int32_t (&xx)[XPT], const int32_t (&gg)[GPT], V (&xr)[XPT],
                         V (&gr)[GPT], uint32_t& m) __attribute__((always_inline)) {
        m = 0u;
#pragma unroll
        for (int q = 0; q < XPT; ++q) {
            const int e = tid + 256 * q, rr = e / XSEG;
            const int c = c0 + (e % XSEG) * F::EPL;
            const bool ok = e < WG_ROWS * XSEG && p + rr < p1 && c < cin;
            m |= (uint32_t)ok << q;
            if constexpr (VEC) xr[q] = *reinterpret_cast<const V*>(X + (ok ? (int64_t)xx[q] * cin + c : 0));
            else xr[q] = ok ? load_frag<T, false>(X + (int64_t)xx[q] * cin, c, cin) : zero_frag<T>();
        }
#pragma unroll
        for (int q = 0; q < GPT; ++q) {
            const int e = tid + 256 * q, rr = e / GSEG;
            const int o = o0 + (e % GSEG) * F::EPL;
            const bool ok = e < WG_ROWS * GSEG && p + rr < p1 && o < cout;
            m |= (uint32_t)ok << (XPT + q);
            if constexpr (VEC) gr[q] = *reinterpret_cast<const V*>(G + (ok ? (int64_t)gg[q] * cout + o : 0));
            else gr[q] = ok ? load_frag<T, false>(G + (int64_t)gg[q] * cout, o, cout) : zero_frag<T>();
        }
    };
    auto store_rows = [&](const V (&xr)[XPT], const V (&gr)[GPT], uint32_t m) __attribute__((always_inline)) {
        const V z = zero_frag<T>();
#pragma unroll
        for (int q = 0; q < XPT; ++q) {
            const int e = tid + 256 * q, r = e / XSEG;
            const int seg = TR ? (e % XSEG) ^ wg_swizzle<XSEG>(r) : e % XSEG;
            if (e < WG_ROWS * XSEG) *reinterpret_cast<V*>(&sX[r * XP + seg * F::EPL]) = ((m >> q) & 1u) ? xr[q] : z;
        }
#pragma unroll
        for (int q = 0; q < GPT; ++q) {
            const int e = tid + 256 * q, r = e / GSEG;
            const int seg = TR ? (e % GSEG) ^ wg_swizzle<GSEG>(r) : e % GSEG;
            if (e < WG_ROWS * GSEG)
                *reinterpret_cast<V*>(&sG[r * GP + seg * F::EPL]) = ((m >> (XPT + q)) & 1u) ? gr[q] : z;
        }
    };
    auto contract = [&]() __attribute__((always_inline)) {
#pragma unroll
        for (int st = 0; st < WG_ROWS / F::CH; ++st) {
            V a[MI], b[NI];
            if constexpr (TR) {
                // element e of lane group g is pair row st*CH + 8g + e: two transpose reads,
                // rows tr and tr + 4 (the same swizzle: it depends on bits 0-1 and 3 of the row)
                const int tr = st * F::CH + 8 * g + (r16 >> 2);
                const int half = (r16 & 1) * 4;  // element offset inside the 16-byte segment
                const int sx = wg_swizzle<XSEG>(tr), sg = wg_swizzle<GSEG>(tr);
#pragma unroll
                for (int i = 0; i < MI; ++i) {
                    const int seg = ((wm + i * 16) / 8 + ((r16 & 3) >> 1)) ^ sx;
                    a[i] = lds_frag_tr16<V>(&sX[tr * XP + seg * 8 + half], XP);
                }
#pragma unroll
                for (int j = 0; j < NI; ++j) {
                    const int seg = ((wn + j * 16) / 8 + ((r16 & 3) >> 1)) ^ sg;
                    b[j] = lds_frag_tr16<V>(&sG[tr * GP + seg * 8 + half], GP);
                }
            } else {
                // element e of lane group g is pair row st*CH + 4e + g
#pragma unroll
                for (int i = 0; i < MI; ++i)
#pragma unroll
                    for (int e = 0; e < F::EPL; ++e) a[i][e] = sX[(st * F::CH + 4 * e + g) * XP + wm + i * 16 + r16];
#pragma unroll
                for (int j = 0; j < NI; ++j)
#pragma unroll
                    for (int e = 0; e < F::EPL; ++e) b[j][e] = sG[(st * F::CH + 4 * e + g) * GP + wn + j * 16 + r16];
            }
#pragma unroll
            for (int i = 0; i < MI; ++i)
#pragma unroll
                for (int j = 0; j < NI; ++j) mfma(a[i], b[j], acc[i][j]);
        }
    };
    const int64_t n_rounds = p0 < p1 ? (p1 - p0 + stride - 1) / stride : 0;
    if (n_rounds > 0) {
        load_idx(p0, xi, gi);
        load_rows(p0, xi, gi, xv, gv, vmask);
        load_idx(p0 + stride, xi, gi);
    }
    for (int64_t j = 0; j < n_rounds; ++j) {
        store_rows(xv, gv, vmask);
        __syncthreads();
        load_rows(p0 + (j + 1) * stride, xi, gi, xv, gv, vmask);  // rows of round j + 1
        load_idx(p0 + (j + 2) * stride, xi, gi);                  // indices of round j + 2
        contract();
        __syncthreads();
    }

    float* __restrict__ out = ws + (int64_t)id * cin * cout;
#pragma unroll
    for (int i = 0; i < MI; ++i)
#pragma unroll
        for (int j = 0; j < NI; ++j)
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int c = c0 + wm + i * 16 + 4 * g + e;
                const int o = o0 + wn + j * 16 + r16;
                if (c < cin && o < cout) out[(int64_t)c * cout + o] = acc[i][j][e];
            }
}

// Ring stages of k_wgrad_dma: four rounds of WGD_ROWS pairs x 128 channels for each
// operand, 8 KB per stage and operand, as separate LDS objects so that the compiler tells
// the stage being read from the stages being filled and does not wait for their copies.
constexpr int WGD_ROWS = 32;
static __shared__ __attribute__((aligned(16))) uint4 wgd_x0[512], wgd_x1[512], wgd_x2[512], wgd_x3[512];
static __shared__ __attribute__((aligned(16))) uint4 wgd_g0[512], wgd_g1[512], wgd_g2[512], wgd_g3[512];
// Source of the copies of pairs past a partial's end: one row of 128 zero channels (device
// memory, zero-initialised at load, never written).
static __device__ __attribute__((aligned(16))) uint4 wgd_zero_row[16];

// Direct-to-LDS form of k_wgrad_pipe for 16-bit layers whose channel counts are
// multiples of 128 (128 x 128 tiles): the rows of each 32-pair round are copied by the
// load unit (global_load_lds_dwordx4) into a ring of four LDS stages, three rounds ahead
// of the round being contracted, instead of passing through registers one round ahead.
// Two workgroups per CU (64 KB of LDS each) then keep six 16 KB rounds in flight per CU
// instead of two 32 KB ones; the rows come mostly from the last-level cache, whose
// latency sets the rate (planes + spheres, 1.1M rows, 6.7M pairs, 128 -> 128 bf16 on
// MI355X, A/B in one process: 0.477-0.480 ms against 0.492-0.510 ms for k_wgrad_pipe;
// four 64-pair stages at one workgroup per CU ran 0.527 ms).
// Offset k's pairs are cut into rounds of WGD_ROWS pairs taken interleaved by its
// partials, as in k_wgrad_pipe with rounds of 32.  Every round issues the same copies: a
// pair past the partial's end copies a row of zeros.  Pair indices are loaded two rounds
// before their rows, so the compiler's wait for them (and the wait for a stage before it
// is read) counts exactly the loads issued after.  Swizzled tiles and transpose reads as
// k_wgrad_pipe.
template <typename T>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2)))
void k_wgrad_dma(const T* __restrict__ X, int cin, const T* __restrict__ G, int cout, const int32_t* __restrict__ pairs,
                 const int64_t* __restrict__ k_offsets, const int32_t* __restrict__ chunk_start, int K, int transposed,
                 float* __restrict__ ws) {
    static_assert(sizeof(T) == 2, "16-bit operands");
    using F = Frag<T>;
    using V = typename F::vec;
    constexpr int MI = 4, NI = 4, TM = 128, TN = 128, SEG = 16, RP = WGD_ROWS;
    constexpr int PT = RP * SEG / 256;  // copies per thread and operand

    const int id = blockIdx.x;
    int lo = 0, hi = K;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (chunk_start[mid] <= id) lo = mid;
        else hi = mid;
    }
    const int k = lo;
    const int64_t stride = (int64_t)(chunk_start[k + 1] - chunk_start[k]) * RP;
    const int64_t p0 = k_offsets[k] + (int64_t)(id - chunk_start[k]) * RP;
    const int64_t p1 = k_offsets[k + 1];

    const int n_ct = cin / TM;
    const int c0 = (blockIdx.y % n_ct) * TM;
    const int o0 = (blockIdx.y / n_ct) * TN;
    const int ci = transposed ? 1 : 0, co = 1 - ci;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int g = lane >> 4, r16 = lane & 15;
    const int wm = (wave & 1) * MI * 16, wn = (wave >> 1) * NI * 16;

    // copy q of this thread: stage element e = tid + 256 q is row e / 16 at 16-byte
    // position e % 16, which holds the row's segment (e % 16) ^ swizzle(row)
    int rr[PT], so[PT];
#pragma unroll
    for (int q = 0; q < PT; ++q) {
        const int e = tid + 256 * q, r = e / SEG;
        rr[q] = r;
        so[q] = ((e % SEG) ^ wg_swizzle<SEG>(r)) * 8;  // element offset of the segment this lane copies
    }
    const T* const zrow = reinterpret_cast<const T*>(wgd_zero_row);
    floatx4 acc[MI][NI];
#pragma unroll
    for (int i = 0; i < MI; ++i)
#pragma unroll
        for (int j = 0; j < NI; ++j) acc[i][j] = floatx4{0.f, 0.f, 0.f, 0.f};

    auto sx = [&](auto sc) -> T* {
        constexpr int s = decltype(sc)::value & 3;
        return reinterpret_cast<T*>(s == 0 ? wgd_x0 : s == 1 ? wgd_x1 : s == 2 ? wgd_x2 : wgd_x3);
    };
    auto sg = [&](auto sc) -> T* {
        constexpr int s = decltype(sc)::value & 3;
        return reinterpret_cast<T*>(s == 0 ? wgd_g0 : s == 1 ? wgd_g1 : s == 2 ? wgd_g2 : wgd_g3);
    };
    int32_t xi[2][PT], gi[2][PT];  // pair indices, two rounds in flight
    auto load_idx = [&](int64_t j, int32_t (&xx)[PT], int32_t (&gg)[PT]) __attribute__((always_inline)) {
        const int64_t p = p0 + j * stride;
#pragma unroll
        for (int q = 0; q < PT; ++q) {
            const int64_t pp = p + rr[q] < p1 ? p + rr[q] : p1 - 1;
            xx[q] = pairs[2 * pp + ci];
            gg[q] = pairs[2 * pp + co];
        }
    };
    // rows of round j into stage sc (the lanes of copy q of wavefront w fill 1 KB from
    // element 256 q + 64 w on)
    auto load_rows = [&](int64_t j, const int32_t (&xx)[PT], const int32_t (&gg)[PT], auto sc)
                         __attribute__((always_inline)) {
        const int64_t p = p0 + j * stride;
        T* const dx = sx(sc);
        T* const dg = sg(sc);
#pragma unroll
        for (int q = 0; q < PT; ++q) {
            const bool ok = p + rr[q] < p1;
            const T* const srcx = ok ? X + (int64_t)xx[q] * cin + c0 + so[q] : zrow + so[q];
            const T* const srcg = ok ? G + (int64_t)gg[q] * cout + o0 + so[q] : zrow + so[q];
            const int e0 = 256 * q + 64 * wave;
            lds_dma16(srcx, dx + e0 * 8);
            lds_dma16(srcg, dg + e0 * 8);
        }
    };
    auto contract = [&](auto sc) __attribute__((always_inline)) {
        const T* const sX = sx(sc);
        const T* const sG = sg(sc);
#pragma unroll
        for (int st = 0; st < RP / F::CH; ++st) {
            V a[MI], b[NI];
            const int tr = st * F::CH + 8 * g + (r16 >> 2);
            const int half = (r16 & 1) * 4;
            const int sw = wg_swizzle<SEG>(tr);
#pragma unroll
            for (int i = 0; i < MI; ++i) {
                const int seg = ((wm + i * 16) / 8 + ((r16 & 3) >> 1)) ^ sw;
                a[i] = lds_frag_tr16<V>(&sX[tr * TM + seg * 8 + half], TM);
            }
#pragma unroll
            for (int j = 0; j < NI; ++j) {
                const int seg = ((wn + j * 16) / 8 + ((r16 & 3) >> 1)) ^ sw;
                b[j] = lds_frag_tr16<V>(&sG[tr * TN + seg * 8 + half], TN);
            }
#pragma unroll
            for (int i = 0; i < MI; ++i)
#pragma unroll
                for (int j = 0; j < NI; ++j) mfma(a[i], b[j], acc[i][j]);
        }
    };
    using I0 = std::integral_constant<int, 0>;
    using I1 = std::integral_constant<int, 1>;
    using I2 = std::integral_constant<int, 2>;
    const int64_t n_rounds = p0 < p1 ? (p1 - p0 + stride - 1) / stride : 0;
    if (n_rounds > 0) {
        // issue order: idx 0, idx 1, rows 0, idx 2, rows 1, idx 3, rows 2; then round j issues
        // idx j+4 and rows j+3, so that when round j waits for rows j the loads after them are
        // always idx j+2, rows j+1, idx j+3, rows j+2 (2 PT loads each)
        load_idx(0, xi[0], gi[0]);
        load_idx(1, xi[1], gi[1]);
        load_rows(0, xi[0], gi[0], I0{});
        load_idx(2, xi[0], gi[0]);
        load_rows(1, xi[1], gi[1], I1{});
        load_idx(3, xi[1], gi[1]);
        load_rows(2, xi[0], gi[0], I2{});
        // round j (u = j mod 4 fixes the stage and the index sets)
        auto round = [&](int64_t j, auto uc) __attribute__((always_inline)) {
            constexpr int u = decltype(uc)::value;
            __builtin_amdgcn_sched_barrier(0);
            wait_vm_lds<4 * 2 * PT>();  // rows j landed (idx j+2, rows j+1, idx j+3, rows j+2 may be in flight)
            __syncthreads();            // ... in every wavefront, and stage j+3 = j-1 is no longer read
            load_idx(j + 4, xi[u & 1], gi[u & 1]);
            load_rows(j + 3, xi[(u + 1) & 1], gi[(u + 1) & 1], std::integral_constant<int, u + 3>{});
            contract(uc);
        };
        int64_t j = 0;
#pragma unroll 1
        for (; j + 4 <= n_rounds; j += 4) {
            round(j, I0{});
            round(j + 1, I1{});
            round(j + 2, I2{});
            round(j + 3, std::integral_constant<int, 3>{});
        }
        if (j < n_rounds) round(j, I0{});
        if (j + 1 < n_rounds) round(j + 1, I1{});
        if (j + 2 < n_rounds) round(j + 2, I2{});
        wait_vm_lds<0>();  // the copies issued past the last round land before the workgroup ends
    }

    float* __restrict__ out = ws + (int64_t)id * cin * cout;
#pragma unroll
    for (int i = 0; i < MI; ++i)
#pragma unroll
        for (int j = 0; j < NI; ++j)
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int c = c0 + wm + i * 16 + 4 * g + e;
                const int o = o0 + wn + j * 16 + r16;
                out[(int64_t)c * cout + o] = acc[i][j][e];
            }
}

// fp64 (and any dtype) scalar weight gradient: one thread per (chunk, c, o).
template <typename T, typename Acc>
__global__ void k_wgrad_generic(const T* __restrict__ X, int cin, const T* __restrict__ G, int cout,
                                const int32_t* __restrict__ pairs, const int64_t* __restrict__ k_offsets,
                                const int32_t* __restrict__ chunk_start, int K, int64_t total_chunks,
                                int transposed, Acc* __restrict__ ws) {
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t per = (int64_t)cin * cout;
    if (idx >= total_chunks * per) return;
    const int id = (int)(idx / per);
    const int c = (int)((idx % per) / cout), o = (int)(idx % cout);
    int lo = 0, hi = K;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (chunk_start[mid] <= id) lo = mid;
        else hi = mid;
    }
    // the same interleaved rounds as k_wgrad_pipe
    const int64_t stride = (int64_t)(chunk_start[lo + 1] - chunk_start[lo]) * WG_ROWS;
    const int64_t p1 = k_offsets[lo + 1];
    const int ci = transposed ? 1 : 0, co = 1 - ci;
    Acc acc = 0;
    for (int64_t r = k_offsets[lo] + (int64_t)(id - chunk_start[lo]) * WG_ROWS; r < p1; r += stride)
        for (int64_t p = r; p < r + WG_ROWS && p < p1; ++p)
            acc += (Acc)X[(int64_t)pairs[2 * p + ci] * cin + c] * (Acc)G[(int64_t)pairs[2 * p + co] * cout + o];
    ws[idx] = acc;
}

// Sum of the chunk partials of each offset, in chunk order.  fp32 uses Neumaier's
// compensated summation; fp64 a plain sum.
template <typename Acc>
__global__ void k_wgrad_reduce(const Acc* __restrict__ ws, const int32_t* __restrict__ chunk_start, int K, int cin,
                               int cout, Acc* __restrict__ gw) {
    const int64_t per = (int64_t)cin * cout;
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (int64_t)K * per) return;
    const int k = (int)(idx / per);
    const int64_t e = idx - (int64_t)k * per;
    Acc sum = 0, comp = 0;
    for (int c = chunk_start[k]; c < chunk_start[k + 1]; ++c) {
        const Acc x = ws[(int64_t)c * per + e];
        const Acc t = sum + x;
        if (sizeof(Acc) == 4) {
            if (fabsf((float)sum) >= fabsf((float)x)) comp += (sum - t) + x;
            else comp += (x - t) + sum;
        }
        sum = t;
    }
    gw[idx] = sum + comp;
}

__global__ void k_fill_zero_u32(uint32_t* __restrict__ p, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

template <typename T, int MI, int NI>
static void launch_wgrad_mfma(const T* X, int cin, const T* G, int cout, const int32_t* pairs, const int64_t* koff,
                              const int32_t* cstart, int K, int64_t total_chunks, int transposed, float* ws,
                              hipStream_t s) {
    constexpr int TM = 2 * MI * 16, TN = 2 * NI * 16;
    const dim3 grid((unsigned)total_chunks, (unsigned)(((cin + TM - 1) / TM) * ((cout + TN - 1) / TN)));
    const bool vec = cin % Frag<T>::EPL == 0 && cout % Frag<T>::EPL == 0 && (uintptr_t)X % 16 == 0 &&
                     (uintptr_t)G % 16 == 0;
    if constexpr (sizeof(T) == 2 && MI == 4 && NI == 4) {
        if (vec && cin % 128 == 0 && cout % 128 == 0) {
            hipLaunchKernelGGL((k_wgrad_dma<T>), grid, dim3(256), 0, s, X, cin, G, cout, pairs, koff, cstart, K,
                               transposed, ws);
            return;
        }
    }
    if (vec)
        hipLaunchKernelGGL((k_wgrad_pipe<T, MI, NI, true>), grid, dim3(256), 0, s, X, cin, G, cout, pairs, koff, cstart,
                           K, transposed, ws);
    else
        hipLaunchKernelGGL((k_wgrad_pipe<T, MI, NI, false>), grid, dim3(256), 0, s, X, cin, G, cout, pairs, koff,
                           cstart, K, transposed, ws);
}

template <typename T>
static int wgrad_mfma_dispatch(const T* X, int cin, const T* G, int cout, const int32_t* pairs, const int64_t* koff,
                               const int32_t* cstart, int K, int chunk, int64_t total_chunks, int transposed,
                               float* ws, float* gw, hipStream_t s) {
    const int mi = cin <= 32 ? 1 : (cin <= 64 ? 2 : 4);
    const int ni = cout <= 32 ? 1 : (cout <= 64 ? 2 : 4);
#define FVDBK_WG(MI_, NI_) \
    launch_wgrad_mfma<T, MI_, NI_>(X, cin, G, cout, pairs, koff, cstart, K, total_chunks, transposed, ws, s)
    if (mi == 1 && ni == 1) FVDBK_WG(1, 1);
    else if (mi == 1 && ni == 2) FVDBK_WG(1, 2);
    else if (mi == 1 && ni == 4) FVDBK_WG(1, 4);
    else if (mi == 2 && ni == 1) FVDBK_WG(2, 1);
    else if (mi == 2 && ni == 2) FVDBK_WG(2, 2);
    else if (mi == 2 && ni == 4) FVDBK_WG(2, 4);
    else if (mi == 4 && ni == 1) FVDBK_WG(4, 1);
    else if (mi == 4 && ni == 2) FVDBK_WG(4, 2);
    else FVDBK_WG(4, 4);
#undef FVDBK_WG
    FVDBK_LAUNCH_CHECK("k_wgrad_mfma");
    const int64_t n = (int64_t)K * cin * cout;
    hipLaunchKernelGGL(k_wgrad_reduce<float>, dim3(grid_1d(n, 256)), dim3(256), 0, s, ws, cstart, K, cin, cout, gw);
    FVDBK_LAUNCH_CHECK("k_wgrad_reduce");
    return FVDBK_OK;
}

static int conv_entry(fvdbk_dtype dtype, const void* in, int64_t n_in, int32_t c_in, const void* a,
                      const int32_t* map, const int32_t* row_order, int64_t n_rows, int32_t K, int32_t M,
                      const void* bias, int relu, void* out, hipStream_t stream) {
    FVDBK_REQUIRE(c_in > 0 && M > 0 && K > 0, FVDBK_EINVAL, "channel counts and kernel volume must be positive");
    FVDBK_REQUIRE(n_in >= 0 && n_rows >= 0, FVDBK_EINVAL, "negative row count");
    FVDBK_REQUIRE(n_rows < ((int64_t)1 << 31), FVDBK_EINVAL, "too many output rows");
    if (n_rows == 0) return FVDBK_OK;
    FVDBK_REQUIRE(in && a && map && out, FVDBK_EINVAL, "null pointer");
    // the row order only changes which rows share a workgroup, never a row's value, so
    // paths without tiles ignore it
    switch (dtype) {
    case FVDBK_BF16:
        return conv_mfma_dispatch<__bf16>((const __bf16*)in, n_in, c_in, (const __bf16*)a, map, row_order, n_rows, K, M,
                                          (__bf16*)out, (const __bf16*)bias, relu, stream);
    case FVDBK_F16:
        return conv_mfma_dispatch<_Float16>((const _Float16*)in, n_in, c_in, (const _Float16*)a, map, row_order, n_rows, K,
                                            M, (_Float16*)out, (const _Float16*)bias, relu, stream);
    case FVDBK_F32:
        return conv_mfma_dispatch<float>((const float*)in, n_in, c_in, (const float*)a, map, row_order, n_rows, K, M,
                                         (float*)out, (const float*)bias, relu, stream);
    case FVDBK_F64: {
        const int64_t n = n_rows * M;
        hipLaunchKernelGGL((k_conv_generic<double, double>), dim3(grid_1d(n, 256)), dim3(256), 0, stream,
                           (const double*)in, c_in, (const double*)a, map, n_rows, K, M, (double*)out,
                           (const double*)bias, relu);
        FVDBK_LAUNCH_CHECK("k_conv_generic");
        return FVDBK_OK;
    }
    default:
        return set_error(FVDBK_ETYPE, "unsupported dtype %d", (int)dtype);
    }
}

__global__ void k_row_keys(const int32_t* __restrict__ map, int64_t n_rows, int K, int32_t* __restrict__ keys) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n_rows) return;
    int32_t key = 0;
    const int kk = K < 31 ? K : 31;
    for (int k = 0; k < kk; ++k) key |= (map[r * K + k] >= 0 ? 1 : 0) << k;
    keys[r] = key;
}

}  // namespace fvdbk

// ---------------------------------------------------------------------------------
// library-owned scratch of the no-allocation weight gradient
//
// File-scope device array: every GPU that loads the library holds its own copy for the
// life of the process.  Host code never passes the array's name to a kernel (on the host
// it is only a stand-in for the symbol); it uses the device address that
// hipGetSymbolAddress returns, looked up once per device with that device current.
// ---------------------------------------------------------------------------------
static __device__ __align__(256) unsigned char scratch[FVDBK_WGRAD_SCRATCH_BYTES];

namespace fvdbk {
namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
void* g_scratch_ptr[kMaxDevices] = {};

// Device address of the current device's scratch array (looked up on its first use).
int scratch_address(void** out) {
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices)
        return set_error(FVDBK_EHIP, "no current HIP device for the wgrad scratch");
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[dev]) {
        void* ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            return set_error(FVDBK_EHIP, "hipGetSymbolAddress(scratch) on device %d: %s", dev, hipGetErrorString(e));
        g_scratch_ptr[dev] = ptr;
    }
    *out = g_scratch_ptr[dev];
    return FVDBK_OK;
}
}  // namespace
}  // namespace fvdbk

using namespace fvdbk;

extern "C" {

int fvdbk_conv_implicit_gemm(fvdbk_dtype dtype, const void* in, int64_t n_in, int32_t c_in, const void* a,
                             const int32_t* map, int64_t n_rows, int32_t K, int32_t M, void* out, uint32_t flags,
                             hipStream_t stream) {
    (void)flags;
    return conv_entry(dtype, in, n_in, c_in, a, map, nullptr, n_rows, K, M, nullptr, 0, out, stream);
}

int fvdbk_conv_implicit_gemm_sorted(fvdbk_dtype dtype, const void* in, int64_t n_in, int32_t c_in, const void* a,
                                    const int32_t* map, const int32_t* row_order, int64_t n_rows, int32_t K,
                                    int32_t M, void* out, uint32_t flags, hipStream_t stream) {
    (void)flags;
    FVDBK_REQUIRE(row_order, FVDBK_EINVAL, "null row order");
    return conv_entry(dtype, in, n_in, c_in, a, map, row_order, n_rows, K, M, nullptr, 0, out, stream);
}

int fvdbk_conv_implicit_gemm_fused(fvdbk_dtype dtype, const void* in, int64_t n_in, int32_t c_in, const void* a,
                                   const int32_t* map, const int32_t* row_order, int64_t n_rows, int32_t K, int32_t M,
                                   const void* bias, uint32_t flags, void* out, hipStream_t stream) {
    FVDBK_REQUIRE((flags & ~(uint32_t)FVDBK_CONV_RELU) == 0, FVDBK_EINVAL, "unknown epilogue flags 0x%x", flags);
    return conv_entry(dtype, in, n_in, c_in, a, map, row_order, n_rows, K, M, bias,
                      (flags & FVDBK_CONV_RELU) ? 1 : 0, out, stream);
}

int fvdbk_conv_implicit_gemm_cat2(fvdbk_dtype dtype, const void* in0, const void* in1, int64_t n_in, int32_t c_half,
                                  const void* a, const int32_t* map, const int32_t* row_order, int64_t n_rows,
                                  int32_t K, int32_t M, const void* bias, uint32_t flags, void* out,
                                  hipStream_t stream) {
    FVDBK_REQUIRE((flags & ~(uint32_t)FVDBK_CONV_RELU) == 0, FVDBK_EINVAL, "unknown epilogue flags 0x%x", flags);
    FVDBK_REQUIRE(c_half > 0 && M > 0 && K > 0, FVDBK_EINVAL, "channel counts and kernel volume must be positive");
    FVDBK_REQUIRE(n_in >= 0 && n_rows >= 0, FVDBK_EINVAL, "negative row count");
    FVDBK_REQUIRE(n_rows < ((int64_t)1 << 31), FVDBK_EINVAL, "too many output rows");
    if (n_rows == 0) return FVDBK_OK;
    FVDBK_REQUIRE(in0 && in1 && a && map && out, FVDBK_EINVAL, "null pointer");
    const int relu = (flags & FVDBK_CONV_RELU) ? 1 : 0;
    bool launched = false;
    switch (dtype) {
    case FVDBK_BF16:
        launched = conv_tile_dispatch2((const __bf16*)in0, (const __bf16*)in1, n_in, 2 * c_half, (const __bf16*)a, map,
                                       row_order, n_rows, K, M, (__bf16*)out, (const __bf16*)bias, relu, stream);
        break;
    case FVDBK_F16:
        launched = conv_tile_dispatch2((const _Float16*)in0, (const _Float16*)in1, n_in, 2 * c_half,
                                       (const _Float16*)a, map, row_order, n_rows, K, M, (_Float16*)out,
                                       (const _Float16*)bias, relu, stream);
        break;
    case FVDBK_F32:
    case FVDBK_F64:
        break;
    default:
        return set_error(FVDBK_ETYPE, "unsupported dtype %d", (int)dtype);
    }
    if (!launched)
        return set_error(FVDBK_EUNSUPPORTED, "no two-source kernel for dtype %d, %d + %d -> %d channels, K = %d%s",
                         (int)dtype, c_half, c_half, M, K, row_order ? " (row order)" : "");
    FVDBK_LAUNCH_CHECK("k_conv_tile (two sources)");
    return FVDBK_OK;
}

int fvdbk_kmap_row_keys(const int32_t* map, int64_t n_rows, int32_t K, int32_t* keys, hipStream_t stream) {
    FVDBK_REQUIRE(K > 0, FVDBK_EINVAL, "kernel volume must be positive");
    if (n_rows == 0) return FVDBK_OK;
    hipLaunchKernelGGL(k_row_keys, dim3(grid_1d(n_rows, 256)), dim3(256), 0, stream, map, n_rows, (int)K, keys);
    FVDBK_LAUNCH_CHECK("k_row_keys");
    return FVDBK_OK;
}

size_t fvdbk_conv_wgrad_workspace_size(fvdbk_dtype dtype, int64_t total_chunks, int32_t c_in, int32_t c_out) {
    const size_t el = dtype == FVDBK_F64 ? 8 : 4;
    return (size_t)(total_chunks > 0 ? total_chunks : 0) * (size_t)c_in * (size_t)c_out * el;
}

int fvdbk_conv_wgrad(fvdbk_dtype dtype, const void* in, int32_t c_in, const void* gout, int32_t c_out,
                     const int32_t* pairs, const int64_t* k_offsets, const int32_t* chunk_start, int32_t K,
                     int32_t chunk, int64_t total_chunks, int32_t transposed, void* workspace, size_t workspace_bytes,
                     void* gw, hipStream_t stream) {
    FVDBK_REQUIRE(c_in > 0 && c_out > 0 && K > 0, FVDBK_EINVAL, "channel counts and kernel volume must be positive");
    FVDBK_REQUIRE(chunk > 0 && chunk % WG_ROWS == 0, FVDBK_EINVAL, "chunk must be a positive multiple of %d", WG_ROWS);
    FVDBK_REQUIRE(workspace_bytes >= fvdbk_conv_wgrad_workspace_size(dtype, total_chunks, c_in, c_out), FVDBK_ENOSPC,
                  "wgrad workspace too small");
    FVDBK_REQUIRE(total_chunks < ((int64_t)1 << 31), FVDBK_EINVAL, "too many chunks");
    const int64_t n = (int64_t)K * c_in * c_out;
    if (total_chunks == 0) {
        // no pairs at all: the gradient is zero (a fill kernel, not a memset node)
        const int64_t words = dtype == FVDBK_F64 ? 2 * n : n;
        hipLaunchKernelGGL(k_fill_zero_u32, dim3(grid_1d(words, 256)), dim3(256), 0, stream, (uint32_t*)gw, words);
        FVDBK_LAUNCH_CHECK("k_fill_zero_u32");
        return FVDBK_OK;
    }
    switch (dtype) {
    case FVDBK_BF16:
        return wgrad_mfma_dispatch<__bf16>((const __bf16*)in, c_in, (const __bf16*)gout, c_out, pairs, k_offsets,
                                           chunk_start, K, chunk, total_chunks, transposed, (float*)workspace,
                                           (float*)gw, stream);
    case FVDBK_F16:
        return wgrad_mfma_dispatch<_Float16>((const _Float16*)in, c_in, (const _Float16*)gout, c_out, pairs,
                                             k_offsets, chunk_start, K, chunk, total_chunks, transposed,
                                             (float*)workspace, (float*)gw, stream);
    case FVDBK_F32:
        return wgrad_mfma_dispatch<float>((const float*)in, c_in, (const float*)gout, c_out, pairs, k_offsets,
                                          chunk_start, K, chunk, total_chunks, transposed, (float*)workspace,
                                          (float*)gw, stream);
    case FVDBK_F64: {
        const int64_t m = total_chunks * c_in * c_out;
        hipLaunchKernelGGL((k_wgrad_generic<double, double>), dim3(grid_1d(m, 256)), dim3(256), 0, stream,
                           (const double*)in, c_in, (const double*)gout, c_out, pairs, k_offsets, chunk_start, K,
                           total_chunks, transposed, (double*)workspace);
        FVDBK_LAUNCH_CHECK("k_wgrad_generic");
        hipLaunchKernelGGL(k_wgrad_reduce<double>, dim3(grid_1d(n, 256)), dim3(256), 0, stream,
                           (const double*)workspace, chunk_start, K, c_in, c_out, (double*)gw);
        FVDBK_LAUNCH_CHECK("k_wgrad_reduce");
        return FVDBK_OK;
    }
    default:
        return set_error(FVDBK_ETYPE, "unsupported dtype %d", (int)dtype);
    }
}

size_t fvdbk_conv_wgrad_scratch_bytes(void) { return (size_t)FVDBK_WGRAD_SCRATCH_BYTES; }

int fvdbk_conv_wgrad_noalloc(fvdbk_dtype dtype, const void* in, int32_t c_in, const void* gout, int32_t c_out,
                             const int32_t* pairs, const int64_t* k_offsets, const int32_t* chunk_start, int32_t K,
                             int32_t chunk, int64_t total_chunks, int32_t transposed, void* gw, hipStream_t stream) {
    const size_t need = fvdbk_conv_wgrad_workspace_size(dtype, total_chunks, c_in, c_out);
    FVDBK_REQUIRE(need <= (size_t)FVDBK_WGRAD_SCRATCH_BYTES, FVDBK_ENOSPC,
                  "wgrad partials need %zu bytes (%lld chunks x %d x %d), over the %zu-byte scratch of "
                  "fvdbk_conv_wgrad_noalloc: use larger chunks or fvdbk_conv_wgrad",
                  need, (long long)total_chunks, (int)c_in, (int)c_out, (size_t)FVDBK_WGRAD_SCRATCH_BYTES);
    void* ws = nullptr;
    const int rc = scratch_address(&ws);
    if (rc != FVDBK_OK) return rc;
    // every chunk writes its whole partial before the reduction reads it: nothing to clear
    return fvdbk_conv_wgrad(dtype, in, c_in, gout, c_out, pairs, k_offsets, chunk_start, K, chunk, total_chunks,
                            transposed, ws, (size_t)FVDBK_WGRAD_SCRATCH_BYTES, gw, stream);
}

}  // extern "C"
