// Sparse-convolution contraction for gfx950: implicit GEMM through a kernel map on
// MFMA (bf16/f16 with v_mfma_f32_16x16x32_*, fp32 with v_mfma_f32_16x16x4f32), a
// scalar fp64 form, and the split-K weight gradient with a compensated reduction.
#include <hip/hip_runtime.h>

#include "fvdbk_common.h"

namespace fvdbk {

typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));

// Per element type: a lane's MFMA operand fragment is 16 bytes (EPL elements) and
// one "chunk" of channels (CH = 4 lane groups x EPL) is consumed per fragment.
template <typename T> struct Frag;
template <> struct Frag<__bf16> {
    using vec = bf16x8;
    static constexpr int EPL = 8;
    static constexpr int CH = 32;
};
template <> struct Frag<_Float16> {
    using vec = f16x8;
    static constexpr int EPL = 8;
    static constexpr int CH = 32;
};
template <> struct Frag<float> {
    using vec = floatx4;
    static constexpr int EPL = 4;
    static constexpr int CH = 16;
};

__device__ inline void mfma(const bf16x8& a, const bf16x8& b, floatx4& c) {
    c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}
__device__ inline void mfma(const f16x8& a, const f16x8& b, floatx4& c) {
    c = __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
}
// fp32: the lane's 4 elements feed 4 MFMAs; MFMA j reduces over element j of every
// lane group, so the 4 instructions together cover the 16 channels of the chunk.
__device__ inline void mfma(const floatx4& a, const floatx4& b, floatx4& c) {
    c = __builtin_amdgcn_mfma_f32_16x16x4f32(a[0], b[0], c, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f32_16x16x4f32(a[1], b[1], c, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f32_16x16x4f32(a[2], b[2], c, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f32_16x16x4f32(a[3], b[3], c, 0, 0, 0);
}

template <typename T> __device__ inline T from_f32(float v) { return (T)v; }

template <typename T, bool VEC>
__device__ inline typename Frag<T>::vec load_frag(const T* __restrict__ p, int c, int cin) {
    using V = typename Frag<T>::vec;
    V v;
    if (VEC) {
        if (c < cin) return *reinterpret_cast<const V*>(p + c);
#pragma unroll
        for (int e = 0; e < Frag<T>::EPL; ++e) v[e] = (T)0;
        return v;
    }
#pragma unroll
    for (int e = 0; e < Frag<T>::EPL; ++e) v[e] = (c + e < cin) ? p[c + e] : (T)0;
    return v;
}

template <typename T>
__device__ inline typename Frag<T>::vec zero_frag() {
    typename Frag<T>::vec v;
#pragma unroll
    for (int e = 0; e < Frag<T>::EPL; ++e) v[e] = (T)0;
    return v;
}

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
                                                   T* __restrict__ Y) {
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
                if (m < M) y[m] = from_f32<T>(acc[i][j][e]);
            }
        }
    }
}

// ---------------------------------------------------------------------------------
// LDS-tiled implicit GEMM (the main forward/dgrad kernel)
//
// A workgroup owns RT = 4*NI*16 consecutive output rows (or rows perm[i] of a row
// permutation) and MI*16 output channels.  At start it stages the rows' kernel-map
// entries in LDS (coalesced) and ORs their validity into a per-tile tap mask; only
// taps with at least one valid row are visited.  For each visited tap the a[k] tile
// (MI*16 x CIN) is double-buffered in LDS: the next tap's tile is fetched into
// registers while the current tap is contracted, then written to the other buffer
// before the single barrier of the iteration.  Input rows are gathered straight from
// global memory into MFMA operand registers (16 bytes per lane), all channel chunks of
// a tap at once (CIN is a compile-time constant).
// ---------------------------------------------------------------------------------

constexpr int TILE_MAX_K = 32;  // taps whose map tile is staged in LDS (3^3, 2^3, ...)

template <typename T, int MI, int NI, int CIN, int LA>
__global__ __launch_bounds__(256) void k_conv_tile(const T* __restrict__ X, const T* __restrict__ A,
                                                   const int32_t* __restrict__ map,
                                                   const int32_t* __restrict__ perm, int64_t n_rows, int K, int M,
                                                   int n_mt, int64_t n_tiles, T* __restrict__ Y) {
    using F = Frag<T>;
    using V = typename F::vec;
    constexpr int RT = 4 * NI * 16;                     // rows per workgroup
    constexpr int WP = CIN + 16 / (int)sizeof(T);       // LDS row pitch of a tap tile (16-byte pad)
    constexpr int WT = MI * 16 * WP;                    // elements of one tap tile
    constexpr int NCH = CIN / F::CH;                    // channel chunks per tap
    constexpr int SEGS = MI * 16 * (CIN / F::EPL);      // 16-byte segments of a tap tile
    constexpr int SEG_PER_T = (SEGS + 255) / 256;
    extern __shared__ __attribute__((aligned(16))) unsigned char tile_smem[];
    T* sW = reinterpret_cast<T*>(tile_smem);
    int32_t* sMap = reinterpret_cast<int32_t*>(tile_smem + 2 * WT * sizeof(T));
    __shared__ unsigned long long s_kmask;

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int g = lane >> 4, r16 = lane & 15;
    // XCD-aware tile order: workgroups go to the 8 XCDs round-robin, so hand each XCD a
    // contiguous run of (row tile, channel tile) pairs; the channel tiles of one row
    // tile are adjacent and share the gathered rows in that XCD's L2.
    const int64_t per = (n_tiles + 7) / 8;
    const int64_t t = (int64_t)(blockIdx.x % 8) * per + blockIdx.x / 8;
    if (t >= n_tiles) return;
    const int64_t row0 = (t / n_mt) * RT;
    const int m0 = (int)(t % n_mt) * (MI * 16);

    if (tid == 0) s_kmask = 0ull;
    __syncthreads();
    // kernel-map tile -> LDS, and the tile's tap mask
    unsigned long long mk = 0ull;
    if (perm == nullptr) {
        const int64_t base = row0 * K;
        const int64_t lim = n_rows * K;
        for (int e = tid; e < RT * K; e += 256) {
            const int32_t v = (base + e < lim) ? map[base + e] : -1;
            sMap[e] = v;
            if (v >= 0) mk |= 1ull << (e % K);
        }
    } else {
        for (int e = tid; e < RT * K; e += 256) {
            const int i = e / K, k = e - i * K;
            const int64_t r = row0 + i;
            const int32_t v = r < n_rows ? map[(int64_t)perm[r] * K + k] : -1;
            sMap[e] = v;
            if (v >= 0) mk |= 1ull << k;
        }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) mk |= __shfl_xor(mk, off);
    if (lane == 0 && mk) atomicOr(&s_kmask, mk);

    floatx4 acc[MI][NI];
#pragma unroll
    for (int i = 0; i < MI; ++i)
#pragma unroll
        for (int j = 0; j < NI; ++j) acc[i][j] = floatx4{0.f, 0.f, 0.f, 0.f};

    // tap tile k -> registers (zero rows beyond M)
    auto fetch = [&](int k, V (&w)[SEG_PER_T]) {
#pragma unroll
        for (int q = 0; q < SEG_PER_T; ++q) {
            const int sgi = tid + q * 256;
            const int m = sgi / (CIN / F::EPL), c = (sgi - m * (CIN / F::EPL)) * F::EPL;
            if (sgi < SEGS && m0 + m < M) w[q] = *reinterpret_cast<const V*>(A + ((int64_t)k * M + m0 + m) * CIN + c);
            else w[q] = zero_frag<T>();
        }
    };
    auto store = [&](int buf, const V (&w)[SEG_PER_T]) {
#pragma unroll
        for (int q = 0; q < SEG_PER_T; ++q) {
            const int sgi = tid + q * 256;
            if (sgi < SEGS) {
                const int m = sgi / (CIN / F::EPL), c = (sgi - m * (CIN / F::EPL)) * F::EPL;
                *reinterpret_cast<V*>(sW + buf * WT + m * WP + c) = w[q];
            }
        }
    };

    // gathered input fragments of tap k (16 bytes per lane per row tile and chunk)
    auto gather = [&](int k, V (&b)[NI][NCH]) -> bool {
        bool any = false;
#pragma unroll
        for (int j = 0; j < NI; ++j) {
            const int32_t src = sMap[(wave * NI * 16 + j * 16 + r16) * K + k];
            any |= src >= 0;
            const T* xr = X + (int64_t)(src >= 0 ? src : 0) * CIN + g * F::EPL;
#pragma unroll
            for (int ch = 0; ch < NCH; ++ch)
                b[j][ch] = src >= 0 ? *reinterpret_cast<const V*>(xr + ch * F::CH) : zero_frag<T>();
        }
        return __any(any);
    };
    auto pop = [](unsigned long long& m) -> int {
        if (!m) return -1;
        const int k = __builtin_ctzll(m);
        m &= m - 1;
        return k;
    };

    auto contract = [&](int buf, const V (&b)[NI][NCH]) {
        const T* wb = sW + buf * WT + r16 * WP + g * F::EPL;
#pragma unroll
        for (int ch = 0; ch < NCH; ++ch) {
#pragma unroll
            for (int i = 0; i < MI; ++i) {
                const V a = *reinterpret_cast<const V*>(wb + i * 16 * WP + ch * F::CH);
#pragma unroll
                for (int j = 0; j < NI; ++j) mfma(a, b[j][ch], acc[i][j]);
            }
        }
    };

    __syncthreads();
    unsigned long long todo = s_kmask;
    if (todo != 0ull) {
        // Software pipeline over the active taps, two phases with swapped register
        // sets (no copies): while tap t0 is contracted from LDS buffer `buf` and
        // registers bA/bB, the weights of the tap after next are fetched into
        // registers and the next tap's input rows are gathered; the next tap's weights
        // (fetched one phase earlier) go to the other LDS buffer before the barrier.
        int t1 = pop(todo), t2;
        const int t0 = t1;
        t1 = pop(todo);
        V wA[SEG_PER_T], wB[SEG_PER_T];
        V bA[NI][NCH], bB[NI][NCH];
        fetch(t0, wA);
        store(0, wA);
        if (t1 >= 0) fetch(t1, wA);
        gather(t0, bA);
        __syncthreads();
        int buf = 0;
        for (;;) {
            // phase A: contract bA; prefetch into bB / wB
            t2 = pop(todo);
            if (t2 >= 0) fetch(t2, wB);
            if (t1 >= 0) gather(t1, bB);
            contract(buf, bA);
            if (t1 < 0) break;
            store(buf ^ 1, wA);
            __syncthreads();
            buf ^= 1;
            t1 = t2;
            // phase B: contract bB; prefetch into bA / wA
            t2 = pop(todo);
            if (t2 >= 0) fetch(t2, wA);
            if (t1 >= 0) gather(t1, bA);
            contract(buf, bB);
            if (t1 < 0) break;
            store(buf ^ 1, wB);
            __syncthreads();
            buf ^= 1;
            t1 = t2;
        }
    }

    // lane (g, r16) of tile (i, j) holds D[m = 4g + e][row = r16]
#pragma unroll
    for (int j = 0; j < NI; ++j) {
        const int64_t ri = row0 + wave * NI * 16 + j * 16 + r16;
        if (ri >= n_rows) continue;
        const int64_t r = perm ? (int64_t)perm[ri] : ri;
        T* __restrict__ y = Y + r * M;
#pragma unroll
        for (int i = 0; i < MI; ++i) {
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int m = m0 + i * 16 + 4 * g + e;
                if (m < M) y[m] = from_f32<T>(acc[i][j][e]);
            }
        }
    }
}

// Fully unrolled variant for a compile-time kernel volume KT (27 = 3^3, 8 = 2^3):
// every tap is visited (no per-tile skipping), and the pipeline registers (weights two
// taps ahead, gathered rows one tap ahead) are indexed by compile-time tap numbers, so
// the compiler keeps them in place.  Gathers are branch-free: a missing neighbour
// loads row 0 and is zeroed by a select.
template <typename T, int MI, int NI, int CIN, int KT, int LX>
__global__ __launch_bounds__(256) void k_conv_tile_k(const T* __restrict__ X, const T* __restrict__ A,
                                                     const int32_t* __restrict__ map,
                                                     const int32_t* __restrict__ perm, int64_t n_rows, int M, int n_mt,
                                                     int64_t n_tiles, T* __restrict__ Y) {
    using F = Frag<T>;
    using V = typename F::vec;
    constexpr int RT = 4 * NI * 16;
    constexpr int WP = CIN + 16 / (int)sizeof(T);
    constexpr int WT = MI * 16 * WP;
    constexpr int NCH = CIN / F::CH;
    constexpr int SEGS = MI * 16 * (CIN / F::EPL);
    constexpr int SEG_PER_T = (SEGS + 255) / 256;
    extern __shared__ __attribute__((aligned(16))) unsigned char tile_smem[];
    T* sW = reinterpret_cast<T*>(tile_smem);
    int32_t* sMap = reinterpret_cast<int32_t*>(tile_smem + 2 * WT * sizeof(T));

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int g = lane >> 4, r16 = lane & 15;
    const int64_t per = (n_tiles + 7) / 8;
    const int64_t t = (int64_t)(blockIdx.x % 8) * per + blockIdx.x / 8;
    if (t >= n_tiles) return;
    const int64_t row0 = (t / n_mt) * RT;
    const int m0 = (int)(t % n_mt) * (MI * 16);

    if (perm == nullptr) {
        const int64_t base = row0 * KT;
        const int64_t lim = n_rows * KT;
        for (int e = tid; e < RT * KT; e += 256) sMap[e] = (base + e < lim) ? map[base + e] : -1;
    } else {
        for (int e = tid; e < RT * KT; e += 256) {
            const int i = e / KT, k = e - i * KT;
            const int64_t r = row0 + i;
            sMap[e] = r < n_rows ? map[(int64_t)perm[r] * KT + k] : -1;
        }
    }

    floatx4 acc[MI][NI];
#pragma unroll
    for (int i = 0; i < MI; ++i)
#pragma unroll
        for (int j = 0; j < NI; ++j) acc[i][j] = floatx4{0.f, 0.f, 0.f, 0.f};

    V w[3][SEG_PER_T];
    V b[LX + 1][NI][NCH];
    auto fetch = [&](int k, V (&wr)[SEG_PER_T]) {
#pragma unroll
        for (int q = 0; q < SEG_PER_T; ++q) {
            const int sgi = tid + q * 256;
            const int m = sgi / (CIN / F::EPL), c = (sgi - m * (CIN / F::EPL)) * F::EPL;
            const bool ok = sgi < SEGS && m0 + m < M;
            const V v = *reinterpret_cast<const V*>(A + ((int64_t)k * M + (ok ? m0 + m : 0)) * CIN + (ok ? c : 0));
            wr[q] = ok ? v : zero_frag<T>();
        }
    };
    auto store = [&](int buf, const V (&wr)[SEG_PER_T]) {
#pragma unroll
        for (int q = 0; q < SEG_PER_T; ++q) {
            const int sgi = tid + q * 256;
            if (SEGS % 256 == 0 || sgi < SEGS) {
                const int m = sgi / (CIN / F::EPL), c = (sgi - m * (CIN / F::EPL)) * F::EPL;
                *reinterpret_cast<V*>(sW + buf * WT + m * WP + c) = wr[q];
            }
        }
    };
    auto gather = [&](int k, V (&bb)[NI][NCH]) {
#pragma unroll
        for (int j = 0; j < NI; ++j) {
            const int32_t src = sMap[(wave * NI * 16 + j * 16 + r16) * KT + k];
            const T* xr = X + (int64_t)(src >= 0 ? src : 0) * CIN + g * F::EPL;
#pragma unroll
            for (int ch = 0; ch < NCH; ++ch) {
                const V v = *reinterpret_cast<const V*>(xr + ch * F::CH);
                bb[j][ch] = src >= 0 ? v : zero_frag<T>();
            }
        }
    };

    __syncthreads();  // map tile visible
    fetch(0, w[0]);
    store(0, w[0]);
    if (KT > 1) fetch(1, w[1]);
#pragma unroll
    for (int k = 0; k < LX && k < KT; ++k) gather(k, b[k]);
    __syncthreads();
#pragma unroll
    for (int k = 0; k < KT; ++k) {
        // weights two taps ahead, input rows LX taps ahead
        if (k + 2 < KT) fetch(k + 2, w[(k + 2) % 3]);
        if (k + LX < KT) gather(k + LX, b[(k + LX) % (LX + 1)]);
        const T* wb = sW + (k % 2) * WT + r16 * WP + g * F::EPL;
#pragma unroll
        for (int ch = 0; ch < NCH; ++ch) {
#pragma unroll
            for (int i = 0; i < MI; ++i) {
                const V a = *reinterpret_cast<const V*>(wb + i * 16 * WP + ch * F::CH);
#pragma unroll
                for (int j = 0; j < NI; ++j) mfma(a, b[k % (LX + 1)][j][ch], acc[i][j]);
            }
        }
        if (k + 1 < KT) {
            store((k + 1) % 2, w[(k + 1) % 3]);
            __syncthreads();
        }
    }

#pragma unroll
    for (int j = 0; j < NI; ++j) {
        const int64_t ri = row0 + wave * NI * 16 + j * 16 + r16;
        if (ri >= n_rows) continue;
        const int64_t r = perm ? (int64_t)perm[ri] : ri;
        T* __restrict__ y = Y + r * M;
#pragma unroll
        for (int i = 0; i < MI; ++i) {
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int m = m0 + i * 16 + 4 * g + e;
                if (m < M) y[m] = from_f32<T>(acc[i][j][e]);
            }
        }
    }
}

template <typename T, int MI, int NI, int CIN>
static size_t tile_lds_bytes(int K) {
    constexpr int WP = CIN + 16 / (int)sizeof(T);
    return (size_t)2 * MI * 16 * WP * sizeof(T) + (size_t)4 * NI * 16 * K * 4;
}

template <typename T, int MI, int NI, int CIN>
static void launch_tile(const T* X, const T* A, const int32_t* map, const int32_t* perm, int64_t n_rows, int K, int M,
                        T* Y, hipStream_t s) {
    const int64_t n_rt = (n_rows + 64 * NI - 1) / (64 * NI);
    const int n_mt = (M + MI * 16 - 1) / (MI * 16);
    const int64_t n_tiles = n_rt * n_mt;
    const dim3 grid((unsigned)((n_tiles + 7) / 8 * 8));
    const size_t lds = tile_lds_bytes<T, MI, NI, CIN>(K);
    // gather lookahead of the unrolled kernel: 2 taps for 16-bit types, 1 for fp32
    constexpr int LX = sizeof(T) == 2 ? 2 : 1;
    if (K == 27) {
        hipLaunchKernelGGL((k_conv_tile_k<T, MI, NI, CIN, 27, LX>), grid, dim3(256), lds, s, X, A, map, perm, n_rows,
                           M, n_mt, n_tiles, Y);
    } else if (K == 8) {
        hipLaunchKernelGGL((k_conv_tile_k<T, MI, NI, CIN, 8, LX>), grid, dim3(256), lds, s, X, A, map, perm, n_rows,
                           M, n_mt, n_tiles, Y);
    } else {
        hipLaunchKernelGGL((k_conv_tile<T, MI, NI, CIN, 1>), grid, dim3(256), lds, s, X, A, map, perm, n_rows, K, M,
                           n_mt, n_tiles, Y);
    }
}

constexpr size_t TILE_LDS_LIMIT = 65536;

// Channel tile for a compile-time CIN: 64 output channels x 128 rows when the LDS
// fits (the best shape measured on MI355X), else 32 x 256 or 16 x 256 rows.
template <typename T, int CIN>
static bool try_tile(const T* X, const T* A, const int32_t* map, const int32_t* perm, int64_t n_rows, int K, int M,
                     T* Y, hipStream_t s) {
    if (M > 32 && tile_lds_bytes<T, 4, 2, CIN>(K) <= TILE_LDS_LIMIT) {
        launch_tile<T, 4, 2, CIN>(X, A, map, perm, n_rows, K, M, Y, s);
    } else if (M > 16 && tile_lds_bytes<T, 2, 4, CIN>(K) <= TILE_LDS_LIMIT) {
        launch_tile<T, 2, 4, CIN>(X, A, map, perm, n_rows, K, M, Y, s);
    } else if (tile_lds_bytes<T, 1, 4, CIN>(K) <= TILE_LDS_LIMIT) {
        launch_tile<T, 1, 4, CIN>(X, A, map, perm, n_rows, K, M, Y, s);
    } else {
        return false;
    }
    return true;
}

// LDS-tiled path for the channel counts it is compiled for; false if not applicable.
template <typename T>
static bool conv_tile_dispatch(const T* X, int cin, const T* A, const int32_t* map, const int32_t* perm,
                               int64_t n_rows, int K, int M, T* Y, hipStream_t s) {
    if (K > TILE_MAX_K || ((uintptr_t)X % 16) != 0 || ((uintptr_t)A % 16) != 0) return false;
    switch (cin) {
    case 32: return try_tile<T, 32>(X, A, map, perm, n_rows, K, M, Y, s);
    case 64: return try_tile<T, 64>(X, A, map, perm, n_rows, K, M, Y, s);
    case 128: return try_tile<T, 128>(X, A, map, perm, n_rows, K, M, Y, s);
    case 256: return try_tile<T, 256>(X, A, map, perm, n_rows, K, M, Y, s);
    default: return false;
    }
}

// Scalar form for any dtype (fp64 accumulates in fp64): one thread per output element.
template <typename T, typename Acc>
__global__ void k_conv_generic(const T* __restrict__ X, int cin, const T* __restrict__ A,
                               const int32_t* __restrict__ map, int64_t n_rows, int K, int M, T* __restrict__ Y) {
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
    Y[idx] = (T)acc;
}

template <typename T, int MI, bool VEC>
static void launch_conv_mfma(const T* X, int cin, const T* A, const int32_t* map, int64_t n_rows, int K, int M, T* Y,
                             hipStream_t s) {
    constexpr int NI = 2;
    const dim3 grid((unsigned)((n_rows + CONV_WAVES * NI * 16 - 1) / (CONV_WAVES * NI * 16)),
                    (unsigned)((M + MI * 16 - 1) / (MI * 16)));
    hipLaunchKernelGGL((k_conv_mfma<T, MI, NI, VEC>), grid, dim3(256), 0, s, X, cin, A, map, n_rows, K, M, Y);
}

template <typename T>
static int conv_mfma_dispatch(const T* X, int cin, const T* A, const int32_t* map, int64_t n_rows, int K, int M, T* Y,
                              hipStream_t s) {
    if (conv_tile_dispatch<T>(X, cin, A, map, nullptr, n_rows, K, M, Y, s)) {
        FVDBK_LAUNCH_CHECK("k_conv_tile");
        return FVDBK_OK;
    }
    const bool vec = (cin % Frag<T>::EPL) == 0 && ((uintptr_t)X % 16) == 0 && ((uintptr_t)A % 16) == 0;
    const int mi = M <= 16 ? 1 : (M <= 32 ? 2 : 4);
    if (vec) {
        if (mi == 1) launch_conv_mfma<T, 1, true>(X, cin, A, map, n_rows, K, M, Y, s);
        else if (mi == 2) launch_conv_mfma<T, 2, true>(X, cin, A, map, n_rows, K, M, Y, s);
        else launch_conv_mfma<T, 4, true>(X, cin, A, map, n_rows, K, M, Y, s);
    } else {
        if (mi == 1) launch_conv_mfma<T, 1, false>(X, cin, A, map, n_rows, K, M, Y, s);
        else if (mi == 2) launch_conv_mfma<T, 2, false>(X, cin, A, map, n_rows, K, M, Y, s);
        else launch_conv_mfma<T, 4, false>(X, cin, A, map, n_rows, K, M, Y, s);
    }
    FVDBK_LAUNCH_CHECK("k_conv_mfma");
    return FVDBK_OK;
}

// ---------------------------------------------------------------------------------
// weight gradient: split-K over the neighbour-map pairs of each offset
//
// gw[k][c][o] = sum_p in[pi][c] * gout[po][o].  A workgroup takes one chunk of the
// pairs of one offset and a (c, o) tile; rows of both operands are gathered into LDS
// (16-byte loads), then read back as MFMA fragments whose reduction index is the
// pair: lane group g, element e takes pair row  step*CH + 4e + g  of the stage (the
// same assignment for A and B, so every pair meets itself).  Partials are fp32.
// ---------------------------------------------------------------------------------

constexpr int WG_ROWS = 64;  // pairs staged per LDS round

template <typename T, int MI, int NI>
__global__ __launch_bounds__(256) void k_wgrad_mfma(const T* __restrict__ X, int cin, const T* __restrict__ G,
                                                    int cout, const int32_t* __restrict__ pairs,
                                                    const int64_t* __restrict__ k_offsets,
                                                    const int32_t* __restrict__ chunk_start, int K, int chunk,
                                                    int transposed, float* __restrict__ ws) {
    using F = Frag<T>;
    constexpr int TM = 2 * MI * 16;  // c tile of the workgroup
    constexpr int TN = 2 * NI * 16;  // o tile
    constexpr int PADB = 32;         // bytes of row padding (spreads the 4 lane groups over banks)
    constexpr int XP = TM + PADB / (int)sizeof(T);
    constexpr int GP = TN + PADB / (int)sizeof(T);
    __shared__ __attribute__((aligned(16))) T sX[WG_ROWS * XP];
    __shared__ __attribute__((aligned(16))) T sG[WG_ROWS * GP];

    const int id = blockIdx.x;
    // offset of this chunk: last k with chunk_start[k] <= id
    int lo = 0, hi = K;  // invariant chunk_start[lo] <= id < chunk_start[hi]
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (chunk_start[mid] <= id) lo = mid;
        else hi = mid;
    }
    const int k = lo;
    const int64_t p0 = k_offsets[k] + (int64_t)(id - chunk_start[k]) * chunk;
    const int64_t p1 = min(p0 + chunk, k_offsets[k + 1]);

    const int n_ct = (cin + TM - 1) / TM;
    const int c0 = (blockIdx.y % n_ct) * TM;
    const int o0 = (blockIdx.y / n_ct) * TN;
    const int ci = transposed ? 1 : 0, co = 1 - ci;

    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int g = lane >> 4, r16 = lane & 15;
    const int wm = (wave & 1) * MI * 16, wn = (wave >> 1) * NI * 16;

    floatx4 acc[MI][NI];
#pragma unroll
    for (int i = 0; i < MI; ++i)
#pragma unroll
        for (int j = 0; j < NI; ++j) acc[i][j] = floatx4{0.f, 0.f, 0.f, 0.f};

    const bool vec_x = (cin % F::EPL) == 0;
    const bool vec_g = (cout % F::EPL) == 0;

    for (int64_t p = p0; p < p1; p += WG_ROWS) {
        // stage rows [p, p + WG_ROWS) of both operands
        constexpr int XSEG = TM / F::EPL, GSEG = TN / F::EPL;
        for (int e = threadIdx.x; e < WG_ROWS * XSEG; e += 256) {
            const int rr = e / XSEG, sg = e - rr * XSEG;
            const int c = c0 + sg * F::EPL;
            typename F::vec v = zero_frag<T>();
            if (p + rr < p1) {
                const int32_t s = pairs[2 * (p + rr) + ci];
                const T* src = X + (int64_t)s * cin;
                if (vec_x) v = load_frag<T, true>(src, c, cin);
                else v = load_frag<T, false>(src, c, cin);
            }
            *reinterpret_cast<typename F::vec*>(&sX[rr * XP + sg * F::EPL]) = v;
        }
        for (int e = threadIdx.x; e < WG_ROWS * GSEG; e += 256) {
            const int rr = e / GSEG, sg = e - rr * GSEG;
            const int o = o0 + sg * F::EPL;
            typename F::vec v = zero_frag<T>();
            if (p + rr < p1) {
                const int32_t t = pairs[2 * (p + rr) + co];
                const T* src = G + (int64_t)t * cout;
                if (vec_g) v = load_frag<T, true>(src, o, cout);
                else v = load_frag<T, false>(src, o, cout);
            }
            *reinterpret_cast<typename F::vec*>(&sG[rr * GP + sg * F::EPL]) = v;
        }
        __syncthreads();
#pragma unroll
        for (int st = 0; st < WG_ROWS / F::CH; ++st) {
            typename F::vec a[MI], b[NI];
#pragma unroll
            for (int i = 0; i < MI; ++i)
#pragma unroll
                for (int e = 0; e < F::EPL; ++e) a[i][e] = sX[(st * F::CH + 4 * e + g) * XP + wm + i * 16 + r16];
#pragma unroll
            for (int j = 0; j < NI; ++j)
#pragma unroll
                for (int e = 0; e < F::EPL; ++e) b[j][e] = sG[(st * F::CH + 4 * e + g) * GP + wn + j * 16 + r16];
#pragma unroll
            for (int i = 0; i < MI; ++i)
#pragma unroll
                for (int j = 0; j < NI; ++j) mfma(a[i], b[j], acc[i][j]);
        }
        __syncthreads();
    }

    // D[m = c][n = o]: lane holds m = 4g + e, n = r16
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

// fp64 (and any dtype) scalar weight gradient: one thread per (chunk, c, o).
template <typename T, typename Acc>
__global__ void k_wgrad_generic(const T* __restrict__ X, int cin, const T* __restrict__ G, int cout,
                                const int32_t* __restrict__ pairs, const int64_t* __restrict__ k_offsets,
                                const int32_t* __restrict__ chunk_start, int K, int chunk, int64_t total_chunks,
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
    const int64_t p0 = k_offsets[lo] + (int64_t)(id - chunk_start[lo]) * chunk;
    const int64_t p1 = min(p0 + chunk, k_offsets[lo + 1]);
    const int ci = transposed ? 1 : 0, co = 1 - ci;
    Acc acc = 0;
    for (int64_t p = p0; p < p1; ++p)
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
            if (fabs((double)sum) >= fabs((double)x)) comp += (sum - t) + x;
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
                              const int32_t* cstart, int K, int chunk, int64_t total_chunks, int transposed,
                              float* ws, hipStream_t s) {
    constexpr int TM = 2 * MI * 16, TN = 2 * NI * 16;
    const unsigned tiles = (unsigned)(((cin + TM - 1) / TM) * ((cout + TN - 1) / TN));
    hipLaunchKernelGGL((k_wgrad_mfma<T, MI, NI>), dim3((unsigned)total_chunks, tiles), dim3(256), 0, s, X, cin, G,
                       cout, pairs, koff, cstart, K, chunk, transposed, ws);
}

template <typename T>
static int wgrad_mfma_dispatch(const T* X, int cin, const T* G, int cout, const int32_t* pairs, const int64_t* koff,
                               const int32_t* cstart, int K, int chunk, int64_t total_chunks, int transposed,
                               float* ws, float* gw, hipStream_t s) {
    const int mi = cin <= 32 ? 1 : (cin <= 64 ? 2 : 4);
    const int ni = cout <= 32 ? 1 : (cout <= 64 ? 2 : 4);
#define FVDBK_WG(MI_, NI_) \
    launch_wgrad_mfma<T, MI_, NI_>(X, cin, G, cout, pairs, koff, cstart, K, chunk, total_chunks, transposed, ws, s)
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

}  // namespace fvdbk

using namespace fvdbk;

extern "C" {

int fvdbk_conv_implicit_gemm(fvdbk_dtype dtype, const void* in, int64_t n_in, int32_t c_in, const void* a,
                             const int32_t* map, int64_t n_rows, int32_t K, int32_t M, void* out, uint32_t flags,
                             hipStream_t stream) {
    (void)flags;
    FVDBK_REQUIRE(c_in > 0 && M > 0 && K > 0, FVDBK_EINVAL, "channel counts and kernel volume must be positive");
    FVDBK_REQUIRE(n_in >= 0 && n_rows >= 0, FVDBK_EINVAL, "negative row count");
    FVDBK_REQUIRE(n_rows < ((int64_t)1 << 31), FVDBK_EINVAL, "too many output rows");
    if (n_rows == 0) return FVDBK_OK;
    FVDBK_REQUIRE(in && a && map && out, FVDBK_EINVAL, "null pointer");
    switch (dtype) {
    case FVDBK_BF16:
        return conv_mfma_dispatch<__bf16>((const __bf16*)in, c_in, (const __bf16*)a, map, n_rows, K, M, (__bf16*)out,
                                          stream);
    case FVDBK_F16:
        return conv_mfma_dispatch<_Float16>((const _Float16*)in, c_in, (const _Float16*)a, map, n_rows, K, M,
                                            (_Float16*)out, stream);
    case FVDBK_F32:
        return conv_mfma_dispatch<float>((const float*)in, c_in, (const float*)a, map, n_rows, K, M, (float*)out,
                                         stream);
    case FVDBK_F64: {
        const int64_t n = n_rows * M;
        hipLaunchKernelGGL((k_conv_generic<double, double>), dim3(grid_1d(n, 256)), dim3(256), 0, stream,
                           (const double*)in, c_in, (const double*)a, map, n_rows, K, M, (double*)out);
        FVDBK_LAUNCH_CHECK("k_conv_generic");
        return FVDBK_OK;
    }
    default:
        return set_error(FVDBK_ETYPE, "unsupported dtype %d", (int)dtype);
    }
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
                           chunk, total_chunks, transposed, (double*)workspace);
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

}  // extern "C"
