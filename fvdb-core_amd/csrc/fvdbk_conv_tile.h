// LDS-tiled implicit-GEMM kernels (the main forward/dgrad contraction) and their
// launch logic.  Included by one source file per element type, which instantiates
// conv_tile_dispatch_t for that type only, so the instances compile in parallel.
#pragma once
#include <type_traits>

#include "fvdbk_common.h"
#include "fvdbk_conv_frag.h"

namespace fvdbk {

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

// f(integral_constant<int, I>) for I = B .. E-1, in order (compile-time loop indices)
template <int B, int E, typename Fn>
__device__ inline void static_for(Fn&& f) {
    if constexpr (B < E) {
        f(std::integral_constant<int, B>{});
        static_for<B + 1, E>(f);
    }
}

template <typename T, int E>
struct StoreVec {
    typedef T type __attribute__((ext_vector_type(E)));
};

// Output tile of one wavefront (NI*16 rows x MI*16 channels) in LDS, pitch padded by
// 16 bytes.  The kernels reuse their weight/map LDS for it after the main loop.
template <typename T, int MI, int NI>
constexpr size_t out_tile_lds_bytes() {
    return (size_t)4 * NI * 16 * (MI * 16 + 16 / sizeof(T)) * sizeof(T);
}

// Epilogue of the tiled kernels: the wavefront's accumulators (lane (g, r16) of block
// (i, j) holds channel m0 + i*16 + 4g + e of row j*16 + r16) get the bias/ReLU, are
// rounded and written to its LDS slice, then read back row-major so that each store
// instruction writes whole rows (16 bytes per lane, consecutive lanes along a row).
// Rows at or beyond n_rows are skipped; perm maps tile rows to output rows.  The
// caller has made the LDS free (barrier after the last read of the weight tiles).
// The wavefront's slice s holds NI*16 rows of MI*16 + 16/sizeof(T) elements.
template <typename T, int MI, int NI>
__device__ inline void store_tile_slice(T* __restrict__ Y, const floatx4 (&acc)[MI][NI], T* s, int lane,
                                        int64_t row_base, const int32_t* __restrict__ perm, int64_t n_rows, int M,
                                        int m0, const T* __restrict__ bias, int relu) {
    constexpr int TW = MI * 16;
    constexpr int OP = TW + 16 / (int)sizeof(T);
    constexpr int RW = NI * 16;
    constexpr int E = 16 / (int)sizeof(T);  // elements per 16-byte lane segment
    constexpr int LPR = TW / E;             // lanes per row
    constexpr int RPP = 64 / LPR;           // rows per pass
    static_assert(LPR <= 64 && RW % RPP == 0, "output tile shape");
    using V4 = typename StoreVec<T, 4>::type;
    using VE = typename StoreVec<T, E>::type;
    const int g = lane >> 4, r16 = lane & 15;
    // this lane's bias values (channels m0 + i*16 + 4g + e), loaded together up front
    float bv[MI][4];
#pragma unroll
    for (int i = 0; i < MI; ++i) {
        const int mb = m0 + i * 16 + 4 * g;
        if (bias != nullptr && mb + 4 <= M && (reinterpret_cast<uintptr_t>(bias + mb) & (4 * sizeof(T) - 1)) == 0) {
            const V4 b4 = *reinterpret_cast<const V4*>(bias + mb);
#pragma unroll
            for (int e = 0; e < 4; ++e) bv[i][e] = (float)b4[e];
        } else {
#pragma unroll
            for (int e = 0; e < 4; ++e) bv[i][e] = (bias != nullptr && mb + e < M) ? (float)bias[mb + e] : 0.f;
        }
    }
#pragma unroll
    for (int j = 0; j < NI; ++j)
#pragma unroll
        for (int i = 0; i < MI; ++i) {
            V4 v;
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int m = m0 + i * 16 + 4 * g + e;
                // the epilogue of epilogue<T>(): bias, ReLU (NaN passes), one rounding
                float x = acc[i][j][e];
                if (bias != nullptr) x += bv[i][e];
                if (relu && !(x > 0.f) && x == x) x = 0.f;
                v[e] = m < M ? from_f32<T>(x) : (T)0.f;
            }
            *reinterpret_cast<V4*>(s + (j * 16 + r16) * OP + i * 16 + 4 * g) = v;
        }
    const bool vec = ((reinterpret_cast<uintptr_t>(Y) | (uintptr_t)((size_t)M * sizeof(T))) & 15u) == 0;
    const int rr = lane / LPR, cc = (lane % LPR) * E;
    const int c = m0 + cc;
#pragma unroll
    for (int p = 0; p < RW / RPP; ++p) {
        const int lr = p * RPP + rr;
        const int64_t ri = row_base + lr;
        if (ri >= n_rows) continue;
        const int64_t r = perm ? (int64_t)perm[ri] : ri;
        const VE v = *reinterpret_cast<const VE*>(s + lr * OP + cc);
        T* __restrict__ y = Y + r * M + c;
        if (vec && c + E <= M) {
            *reinterpret_cast<VE*>(y) = v;
        } else {
#pragma unroll
            for (int q = 0; q < E; ++q)
                if (c + q < M) y[q] = v[q];
        }
    }
}

template <typename T, int MI, int NI>
__device__ inline void store_tile(T* __restrict__ Y, const floatx4 (&acc)[MI][NI], unsigned char* smem, int wave,
                                  int lane, int64_t row_base, const int32_t* __restrict__ perm, int64_t n_rows, int M,
                                  int m0, const T* __restrict__ bias, int relu) {
    constexpr size_t slice = (size_t)NI * 16 * (MI * 16 + 16 / sizeof(T));
    store_tile_slice<T, MI, NI>(Y, acc, reinterpret_cast<T*>(smem) + wave * slice, lane, row_base, perm, n_rows, M, m0,
                                bias, relu);
}

template <typename T, int MI, int NI, int CIN, int LA>
__global__ __launch_bounds__(256) void k_conv_tile(const T* __restrict__ X, const T* __restrict__ A,
                                                   const int32_t* __restrict__ map,
                                                   const int32_t* __restrict__ perm, int64_t n_rows, int K, int M,
                                                   int n_mt, int64_t n_tiles, T* __restrict__ Y, const T* __restrict__ bias, int relu) {
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
            const int mc = m0 + m;
            if (sgi < SEGS && mc < M) w[q] = *reinterpret_cast<const V*>(A + ((int64_t)k * M + mc) * CIN + c);
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

    __syncthreads();  // every wavefront is done with the weight tiles
    store_tile<T, MI, NI>(Y, acc, tile_smem, wave, lane, row0 + wave * NI * 16, perm, n_rows, M, m0, bias, relu);
}

// Fully unrolled variant for a compile-time kernel volume KT (27 = 3^3, 8 = 2^3): the
// pipeline registers (weights two taps ahead, gathered rows LX taps ahead) are indexed
// by compile-time tap numbers, so the compiler keeps them in place.
// Rows are consecutive (tree order) and every tap is visited: a tile of neighbouring
// rows almost always uses all of them.  Rows through a tap-mask ordering run the
// compacted-step form of the wide-channel kernel below instead.
// WPE > 0 asks the register allocator for WPE waves per SIMD (128 registers per
// lane at 4): latency hiding, not registers, limits this kernel where it fits.
template <typename T, int MI, int NI, int CIN, int KT, int LX, int WPE>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(WPE > 0 ? WPE : 1, WPE > 0 ? WPE : 8)))
void k_conv_tile_k(const T* __restrict__ X, const T* __restrict__ A,
                                                     const int32_t* __restrict__ map,
                                                     int64_t n_rows, int M, int n_mt,
                                                     int64_t n_tiles, T* __restrict__ Y, const T* __restrict__ bias, int relu) {
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

    // kernel-map tile -> LDS
    {
        const int64_t base = row0 * KT;
        const int64_t lim = n_rows * KT;
        for (int e = tid; e < RT * KT; e += 256) {
            const int32_t v = (base + e < lim) ? map[base + e] : -1;
            sMap[e] = v;
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
            const int mc = m0 + m;
            const bool ok = sgi < SEGS && mc < M;
            const V v = *reinterpret_cast<const V*>(A + ((int64_t)k * M + (ok ? mc : 0)) * CIN + (ok ? c : 0));
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
    // Lanes of missing neighbours are masked off the load (they keep zeros), so the
    // memory pipeline moves only the rows that exist.
    auto gather = [&](int k, V (&bb)[NI][NCH]) {
#pragma unroll
        for (int j = 0; j < NI; ++j) {
            const int32_t src = sMap[(wave * NI * 16 + j * 16 + r16) * KT + k];
#pragma unroll
            for (int ch = 0; ch < NCH; ++ch) bb[j][ch] = zero_frag<T>();
            if (src >= 0) {
                const T* xr = X + (int64_t)src * CIN + g * F::EPL;
#pragma unroll
                for (int ch = 0; ch < NCH; ++ch) bb[j][ch] = *reinterpret_cast<const V*>(xr + ch * F::CH);
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
        {
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
        }
        if (k + 1 < KT) {
            store((k + 1) % 2, w[(k + 1) % 3]);
            __syncthreads();
        }
    }

    __syncthreads();  // every wavefront is done with the weight tiles
    store_tile<T, MI, NI>(Y, acc, tile_smem, wave, lane, row0 + wave * NI * 16, nullptr, n_rows, M, m0, bias, relu);
}

// Workgroup barrier without a memory fence.  The caller has waited for the counters it
// needs (s_waitcnt); the scheduling barriers keep every instruction on its side.
__device__ inline void wg_barrier() {
    __builtin_amdgcn_sched_barrier(0);
    asm volatile("s_barrier" ::: "memory");
    __builtin_amdgcn_sched_barrier(0);
}

// Copies 16 bytes per lane from global memory to LDS: lane i's bytes land at
// lds + 16*i (global_load_lds_dwordx4).  A non-template wrapper: called with pointers
// whose type depends on a kernel's template parameter, the builtin made the host pass
// drop the kernel's launch stub.
__device__ inline void lds_dma16(const void* src, void* lds) {
    __builtin_amdgcn_global_load_lds(src, lds, 16, 0, 0);
}

// s_waitcnt vmcnt(N) lgkmcnt(0), N a compile-time count below 64 (gfx9 encoding: vmcnt
// bits [3:0] and [15:14], expcnt [6:4] left at its maximum, lgkmcnt [11:8]).
template <int N>
__device__ inline void wait_vm_lds() {
    static_assert(N >= 0 && N < 64, "vmcnt range");
    __builtin_amdgcn_s_waitcnt((N & 15) | ((N >> 4) << 14) | 0x70);
}
// The same for a count that is constant once the caller's loop is unrolled.
__device__ inline void wait_vm_lds(int n) {
    switch (n) {
#define FVDBK_WAIT_CASE(x) \
    case x: wait_vm_lds<x>(); break;
        FVDBK_WAIT_CASE(1) FVDBK_WAIT_CASE(2) FVDBK_WAIT_CASE(3) FVDBK_WAIT_CASE(4) FVDBK_WAIT_CASE(5)
        FVDBK_WAIT_CASE(6) FVDBK_WAIT_CASE(7) FVDBK_WAIT_CASE(8) FVDBK_WAIT_CASE(9) FVDBK_WAIT_CASE(10)
        FVDBK_WAIT_CASE(11) FVDBK_WAIT_CASE(12) FVDBK_WAIT_CASE(13) FVDBK_WAIT_CASE(14) FVDBK_WAIT_CASE(15)
        FVDBK_WAIT_CASE(16) FVDBK_WAIT_CASE(17) FVDBK_WAIT_CASE(18) FVDBK_WAIT_CASE(19) FVDBK_WAIT_CASE(20)
#undef FVDBK_WAIT_CASE
    default: wait_vm_lds<0>(); break;
    }
}

// WSKIP: a wavefront none of whose 32 rows has a step's tap skips that step's weight reads and
// MFMAs (the gathers and the weight staging still run, so the load counts per step stay
// fixed); a skipped step would have added exact zeros, so results are bit-identical for
// finite weights.
// Direct-to-LDS form of k_conv_tile_k for the most common narrow layers: 16-bit features,
// 64 input channels, KT = 27 or 8 taps (3x3x3, 2x2x2), rows in tree order, outputs in
// whole 64-channel tiles.
// Wavefronts of the same shape (32 rows x 64 output channels) and MFMA order, so the
// results are bit-identical to k_conv_tile_k, in workgroups of NW = 4 or 8 wavefronts
// (128 or 256 rows); the memory pipeline differs:
//  * weights: each tap's 64 x 64 tile is copied global -> LDS by the load unit
//    (global_load_lds_dwordx4, 8/NW per lane), two taps ahead into a ring of three
//    buffers, with no registers and no LDS store instructions.  Rows are 128 bytes
//    unpadded; the 16-byte segment s of row r sits at position s ^ (r & 7), so the
//    MFMA operand reads of 8 consecutive rows hit distinct banks.  The buffers are
//    separate LDS objects so that the compiler can tell the staging target from the
//    buffer being read and does not wait for every outstanding load.
//  * gathers: raw buffer loads over the feature rows; a missing neighbour takes an
//    offset beyond the buffer, which loads zeros without a branch, so every tap issues
//    exactly NI*2 gathers + 8/NW staging loads.  With that fixed count the end of each tap
//    waits (vmcnt) only for the loads the next tap needs, and the gathers run LX = 3
//    taps ahead within 120 registers (4 waves per SIMD).
//  * epilogue: the output tile goes through LDS (wavefront slices in the weight
//    buffers and the map tile) and is stored as whole rows.
// With 8 wavefronts each weight tile is staged once for 256 rows, halving the staging
// loads per row (they were a third of the kernel's vector memory instructions).
// MI = 8: 128 output channels per workgroup (wavefronts of 32 rows x 128 channels), so
// each gathered row feeds twice the MFMAs and layers of 128k output channels gather
// every row half as often; 16 KB weight tiles, gathers LX = 2 taps ahead, and the
// output leaves in two 64-channel halves through slices of the ring.
// NCC = 2 (128 input channels): each pipeline step is one (tap, 64-channel chunk) pair,
// chunk fastest -- the accumulation order of k_conv_tile_c, so results are bit-identical
// to it (54 steps for 3x3x3).
// Features must span fewer than 2^31 bytes (32-bit buffer offsets); callers check.
// Weight rings of k_conv_tile_dma: three separate LDS objects of 8 KB (MI = 4) or 16 KB
// (MI = 8) (file scope: a static __shared__ array inside a kernel template does not link;
// a kernel is allocated only the objects it uses).
static __shared__ __attribute__((aligned(16))) uint4 dma_ring0[512], dma_ring1[512], dma_ring2[512];
static __shared__ __attribute__((aligned(16))) uint4 dma8_ring0[1024], dma8_ring1[1024], dma8_ring2[1024];

template <typename T, int LX, int NW, int KT, int MI = 4, int NCC = 1, bool WSKIP = false>
__global__ __launch_bounds__(64 * NW) __attribute__((amdgpu_waves_per_eu(4, 4)))
void k_conv_tile_dma(const T* __restrict__ X, int64_t n_in, const T* __restrict__ A, const int32_t* __restrict__ map,
                     int64_t n_rows, int M, int n_mt, int64_t n_tiles, T* __restrict__ Y, const T* __restrict__ bias,
                     int relu) {
    static_assert(sizeof(T) == 2, "16-bit features");
    using F = Frag<T>;
    using V = typename F::vec;
    // CIN = NCC chunks of CC = 64 channels; a pipeline step v is (tap v / NCC, chunk v % NCC)
    constexpr int NI = 2, CC = 64, CIN = CC * NCC, KV = KT * NCC;
    constexpr int RT = NW * NI * 16;            // rows per workgroup
    constexpr int WT = MI * 16 * CC;           // elements of one step's weight tile (8 or 16 KB)
    constexpr int NCH = CC / F::CH;
    constexpr int GPT = NI * NCH;              // gather loads per tap and lane
    constexpr int OUT_SLICE = NI * 16 * (4 * 16 + 8);  // one 64-channel output slice
    static_assert(NW == 4 || NW == 8, "4 or 8 wavefronts");
    static_assert(MI == 4 || MI == 8, "64 or 128 output channels");
    static_assert(KV > LX, "more steps than the gather lookahead");
    static_assert(NCC == 1 || NCC == 2, "64 or 128 input channels");
    static_assert(OUT_SLICE <= WT, "output slices fit the weight buffers");
    static_assert(MI == 4 || 3 * OUT_SLICE <= WT, "three slices per 16 KB buffer");
    static_assert(WT * sizeof(T) == (MI == 4 ? sizeof(dma_ring0) : sizeof(dma8_ring0)), "ring buffer size");
    extern __shared__ __attribute__((aligned(16))) unsigned char tile_smem[];
    T* const sW0 = reinterpret_cast<T*>(MI == 4 ? dma_ring0 : dma8_ring0);
    T* const sW1 = reinterpret_cast<T*>(MI == 4 ? dma_ring1 : dma8_ring1);
    T* const sW2 = reinterpret_cast<T*>(MI == 4 ? dma_ring2 : dma8_ring2);
    int32_t* sMap = reinterpret_cast<int32_t*>(tile_smem);
    auto wbuf = [&](int v) -> T* {
        const int r = v % 3;
        return r == 0 ? sW0 : (r == 1 ? sW1 : sW2);
    };

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int g = lane >> 4, r16 = lane & 15;
    const int64_t per = (n_tiles + 7) / 8;
    const int64_t t = (int64_t)(blockIdx.x % 8) * per + blockIdx.x / 8;
    if (t >= n_tiles) return;
    const int64_t row0 = (t / n_mt) * RT;
    const int m0 = (int)(t % n_mt) * (MI * 16);
    {
        const int64_t base = row0 * KT;
        const int64_t lim = n_rows * KT;
        for (int e = tid; e < RT * KT; e += 64 * NW) sMap[e] = (base + e < lim) ? map[base + e] : -1;
    }
    floatx4 acc[MI][NI];
#pragma unroll
    for (int i = 0; i < MI; ++i)
#pragma unroll
        for (int j = 0; j < NI; ++j) acc[i][j] = floatx4{0.f, 0.f, 0.f, 0.f};

    // staging: the MI*16 weight rows are 2*MI blocks of 8; wavefront w copies blocks
    // w*SQ .. +SQ-1, lane -> (row, position)
    constexpr int SQ = 2 * MI / NW;            // staging loads per lane per tap
    const T* wsrc[SQ];
#pragma unroll
    for (int q = 0; q < SQ; ++q) {
        const int row = (wave * SQ + q) * 8 + (lane >> 3);
        const int seg = (lane & 7) ^ (lane >> 3);
        wsrc[q] = A + (int64_t)(m0 + row) * CIN + seg * 8;
    }
    auto stage = [&](int v) {
        const int k = v / NCC, c = v % NCC;
        T* dst = wbuf(v) + wave * SQ * 512;
#pragma unroll
        for (int q = 0; q < SQ; ++q)
            lds_dma16(wsrc[q] + (int64_t)k * M * CIN + c * CC, dst + q * 512);
    };
    const int bytes = (int)(n_in * CIN * (int64_t)sizeof(T));
    const __amdgpu_buffer_rsrc_t xr = __builtin_amdgcn_make_buffer_rsrc((void*)X, (short)0, bytes, 0x00020000);
    // A missing neighbour's offset: at or beyond the buffer's size (features span at most
    // 0x7FFFFF00 bytes, checked by the launchers), so the load returns zeros, and with
    // room below 2^31 for the chunk offsets added to it (at most 64 bytes).
    constexpr int MISSING = 0x7FFFFF00;
    V b[LX + 1][NI][NCH];
    // WSKIP: whether any of the 16 rows of group j has step v's tap (wavefront-uniform); a
    // wavefront none of whose rows has the tap skips the step's MFMAs, which would add exact zeros
    bool has[LX + 1][NI];
    auto gather = [&](int v, V (&bb)[NI][NCH], bool (&hh)[NI]) {
        const int k = v / NCC, c = v % NCC;
#pragma unroll
        for (int j = 0; j < NI; ++j) {
            const int32_t src = sMap[(wave * NI * 16 + j * 16 + r16) * KT + k];
            const int off = src >= 0 ? src * (CIN * (int)sizeof(T)) + c * (CC * (int)sizeof(T)) + g * 16 : MISSING;
            if constexpr (WSKIP) hh[j] = __any(src >= 0);
#pragma unroll
            for (int ch = 0; ch < NCH; ++ch)
                bb[j][ch] = __builtin_bit_cast(V, __builtin_amdgcn_raw_buffer_load_b128(xr, off + ch * 64, 0, 0));
        }
    };
    // The wait at the end of step v: issue order is stage 0, stage 1, gathers 0..LX-1 (all
    // complete before the loop), then per step t: stage t+2, gathers t+LX.  It lets through
    // only the loads issued after both the staging of step v+1 (its weights must have landed
    // before the barrier) and the gathers of step v+1 (issued in step v+1-LX), so that step
    // v+1's MFMAs find their operands complete: with direct-to-LDS copies in flight the
    // compiler's own wait for gathered registers would be vmcnt(0), which would also wait for
    // the loads issued for later steps.
    auto step_wait = [&](int v) -> int {
        auto G = [&](int tt) { return tt + LX < KV ? GPT : 0; };
        auto D = [&](int tt) { return tt + 2 < KV ? SQ : 0; };
        const int st = v == 0 ? D(0) + G(0) : G(v - 1) + D(v) + G(v);
        if (v + 1 < LX) return st;
        int ga = 0;
        for (int tt = v + 2 - LX; tt <= v; ++tt) ga += D(tt) + G(tt);
        return ga < st ? ga : st;
    };

    __syncthreads();  // map tile visible
    stage(0);
    stage(1);
#pragma unroll
    for (int v = 0; v < LX; ++v) gather(v, b[v], has[v]);
    wait_vm_lds<0>();
    wg_barrier();
#pragma unroll
    for (int v = 0; v < KV; ++v) {
        if (v + 2 < KV) stage(v + 2);
        if (v + LX < KV) gather(v + LX, b[(v + LX) % (LX + 1)], has[(v + LX) % (LX + 1)]);
        const T* wb = wbuf(v);
        const int sl = v % (LX + 1);
        bool any = true;
        if constexpr (WSKIP) {
            any = false;
#pragma unroll
            for (int j = 0; j < NI; ++j) any |= has[sl][j];
        }
        if (any) {
#pragma unroll
            for (int ch = 0; ch < NCH; ++ch) {
#pragma unroll
                for (int i = 0; i < MI; ++i) {
                    const int pos = (ch * 4 + g) ^ (r16 & 7);
                    const V a = *reinterpret_cast<const V*>(wb + (i * 16 + r16) * CC + pos * 8);
#pragma unroll
                    for (int j = 0; j < NI; ++j) mfma(a, b[sl][j][ch], acc[i][j]);
                }
            }
        }
        if (v + 1 < KV) {
            // step v+1's weights have landed in every wavefront; nobody reads step v's buffer
            // again before stage(v+3) overwrites it
            __builtin_amdgcn_sched_barrier(0);
            wait_vm_lds(step_wait(v));
            wg_barrier();
        }
    }
    __syncthreads();  // every wavefront is done with the weight buffers and the map tile
    if constexpr (MI == 4) {
        T* slice = wave == 0 ? sW0
                   : wave == 1 ? sW1
                   : wave == 2 ? sW2
                               : reinterpret_cast<T*>(tile_smem) + (wave - 3) * OUT_SLICE;
        store_tile_slice<T, MI, NI>(Y, acc, slice, lane, row0 + wave * NI * 16, nullptr, n_rows, M, m0, bias, relu);
    } else {
        // three slices per ring buffer; the wavefront stores its two 64-channel halves in turn
        T* const ring = wave < 3 ? sW0 : (wave < 6 ? sW1 : sW2);
        T* slice = ring + (wave % 3) * OUT_SLICE;
        using Half = floatx4[4][NI];
        store_tile_slice<T, 4, NI>(Y, *reinterpret_cast<const Half*>(&acc[0][0]), slice, lane, row0 + wave * NI * 16,
                                   nullptr, n_rows, M, m0, bias, relu);
        store_tile_slice<T, 4, NI>(Y, *reinterpret_cast<const Half*>(&acc[4][0]), slice, lane, row0 + wave * NI * 16,
                                   nullptr, n_rows, M, m0 + 64, bias, relu);
    }
}

// k_conv_tile_dma over rows taken through a tap-mask ordering `perm` (row i of the tile is
// output row perm[row0 + i]), with compacted steps: the prologue ORs the tile's map
// entries into a tap mask, and the pipeline runs only the (tap, chunk) steps of the taps
// that at least one of the tile's rows uses, so a skipped tap costs neither its weight
// staging, its gathers, its barrier nor its MFMAs; inside a step a wavefront none of whose
// 32 rows has the tap skips its MFMAs as well.  Same wavefront shape, LDS layout and
// accumulation order (taps ascending, chunks ascending) as k_conv_tile_dma, and a skipped
// step would have added exact zeros, so the output is bit-identical to the tree-order
// kernel for finite weights, whatever the order.
// The step count is known only at run time: the loop body is U = lcm(3, LX+1) steps (a
// multiple of the 3-entry weight ring and the LX+1-entry gather ring, so ring slots stay
// compile-time), unguarded while the group and the loads it issues lie inside the step
// sequence, and guarded in the last groups, where the per-step load count, and with it the
// wait before each barrier, depends on how many steps remain.  WPE waves per SIMD: 4
// (128 registers per lane) or 2 (256, room for deeper gather lookahead).
// TWO (NCC = 2): the input is the channel concatenation [X | X1] of two arrays of 64
// channels each, read from both without forming it (chunk 0 from X, chunk 1 from X1).
template <typename T, int LX, int NW, int KT, int MI, int NCC, int WPE = 4, bool TWO = false>
__global__ __launch_bounds__(64 * NW) __attribute__((amdgpu_waves_per_eu(WPE, WPE)))
void k_conv_tile_dma_perm(const T* __restrict__ X, const T* __restrict__ X1, int64_t n_in, const T* __restrict__ A,
                          const int32_t* __restrict__ map, const int32_t* __restrict__ perm, int64_t n_rows, int M,
                          int n_mt, int64_t n_tiles, T* __restrict__ Y, const T* __restrict__ bias, int relu) {
    static_assert(sizeof(T) == 2, "16-bit features");
    static_assert(KT <= 32, "tap masks are 32 bits");
    using F = Frag<T>;
    using V = typename F::vec;
    constexpr int NI = 2, CC = 64, CIN = CC * NCC, U = LX == 3 ? 12 : 6;
    constexpr int RT = NW * NI * 16;
    constexpr int WT = MI * 16 * CC;
    constexpr int NCH = CC / F::CH;
    constexpr int GPT = NI * NCH;
    constexpr int SQ = 2 * MI / NW;
    constexpr int OUT_SLICE = NI * 16 * (4 * 16 + 8);
    static_assert(NW == 4 || NW == 8, "4 or 8 wavefronts");
    static_assert(MI == 4 || MI == 8, "64 or 128 output channels");
    static_assert(NCC == 1 || NCC == 2, "64 or 128 input channels");
    static_assert(LX >= 1 && LX <= 3, "gathers one to three steps ahead");
    static_assert(U % (LX + 1) == 0 && U % 3 == 0, "rings must divide the unrolled step count");
    static_assert(OUT_SLICE <= WT && (MI == 4 || 3 * OUT_SLICE <= WT), "output slices fit the weight buffers");
    static_assert(WT * sizeof(T) == (MI == 4 ? sizeof(dma_ring0) : sizeof(dma8_ring0)), "ring buffer size");
    extern __shared__ __attribute__((aligned(16))) unsigned char tile_smem[];
    __shared__ unsigned int s_tmask;
    T* const sW0 = reinterpret_cast<T*>(MI == 4 ? dma_ring0 : dma8_ring0);
    T* const sW1 = reinterpret_cast<T*>(MI == 4 ? dma_ring1 : dma8_ring1);
    T* const sW2 = reinterpret_cast<T*>(MI == 4 ? dma_ring2 : dma8_ring2);
    int32_t* sMap = reinterpret_cast<int32_t*>(tile_smem);
    auto wring = [&](auto rc) -> T* {
        constexpr int r = decltype(rc)::value % 3;
        return r == 0 ? sW0 : (r == 1 ? sW1 : sW2);
    };

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int g = lane >> 4, r16 = lane & 15;
    const int64_t per = (n_tiles + 7) / 8;
    const int64_t t = (int64_t)(blockIdx.x % 8) * per + blockIdx.x / 8;
    if (t >= n_tiles) return;
    const int64_t row0 = (t / n_mt) * RT;
    const int m0 = (int)(t % n_mt) * (MI * 16);
    if (tid == 0) s_tmask = 0u;
    __syncthreads();
    {
        unsigned int tm = 0u;
        for (int e = tid; e < RT * KT; e += 64 * NW) {
            const int i = e / KT, k = e - i * KT;
            const int64_t r = row0 + i;
            const int32_t v = r < n_rows ? map[(int64_t)perm[r] * KT + k] : -1;
            sMap[e] = v;
            if (v >= 0) tm |= 1u << k;
        }
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) tm |= __shfl_xor(tm, off);
        if (lane == 0 && tm) atomicOr(&s_tmask, tm);
    }
    __syncthreads();  // map tile and tap mask visible
    // the weight and gather streams walk the tile's taps with uniform bit iterators: the
    // lowest set bit is the tap of the stream's next step, cleared after its last chunk
    unsigned int fit = (unsigned int)__builtin_amdgcn_readfirstlane((int)s_tmask), git = fit;
    const int kv = __popc(fit) * NCC;
    auto next_tap = [](unsigned int& it, int v) -> int {
        const int tap = __builtin_ctz(it);
        if ((v & (NCC - 1)) == NCC - 1) it &= it - 1u;
        return tap;
    };

    floatx4 acc[MI][NI];
#pragma unroll
    for (int i = 0; i < MI; ++i)
#pragma unroll
        for (int j = 0; j < NI; ++j) acc[i][j] = floatx4{0.f, 0.f, 0.f, 0.f};

    // staging: the MI*16 weight rows are 2*MI blocks of 8; wavefront w copies blocks
    // w*SQ .. +SQ-1, lane -> (row, swizzled 16-byte position); 32-bit element offsets
    // over a wavefront-uniform base
    int woff[SQ];
#pragma unroll
    for (int q = 0; q < SQ; ++q) {
        const int row = (wave * SQ + q) * 8 + (lane >> 3);
        const int seg = (lane & 7) ^ (lane >> 3);
        woff[q] = (m0 + row) * CIN + seg * 8;
    }
    // weights of step v into ring buffer `dst` (called once per step, in step order)
    auto stage = [&](int v, T* dst) {
        const int k = next_tap(fit, v), c = v & (NCC - 1);
        const T* src = A + (int64_t)k * M * CIN + c * CC;
        dst += wave * SQ * 512;
#pragma unroll
        for (int q = 0; q < SQ; ++q) lds_dma16(src + woff[q], dst + q * 512);
    };
    static_assert(!TWO || NCC == 2, "two sources of one chunk each");
    constexpr int RB = (TWO ? CC : CIN) * (int)sizeof(T);  // bytes per row of a source
    const int bytes = (int)(n_in * RB);
    const __amdgpu_buffer_rsrc_t xr = __builtin_amdgcn_make_buffer_rsrc((void*)X, (short)0, bytes, 0x00020000);
    const __amdgpu_buffer_rsrc_t xr1 =
        __builtin_amdgcn_make_buffer_rsrc((void*)(TWO ? X1 : X), (short)0, bytes, 0x00020000);
    constexpr int MISSING = 0x7FFFFF00;  // beyond the features (checked by the launcher): loads zeros
    V b[LX + 1][NI][NCH];
    bool has[LX + 1][NI];
    // input rows of step v (called once per step, in step order)
    auto gather = [&](int v, V (&bb)[NI][NCH], bool (&hh)[NI]) {
        const int k = next_tap(git, v), c = v & (NCC - 1);
#pragma unroll
        for (int j = 0; j < NI; ++j) {
            const int32_t src = sMap[(wave * NI * 16 + j * 16 + r16) * KT + k];
            const int off = src >= 0 ? src * RB + (TWO ? 0 : c * (CC * (int)sizeof(T))) + g * 16 : MISSING;
            hh[j] = __any(src >= 0);
#pragma unroll
            for (int ch = 0; ch < NCH; ++ch)
                bb[j][ch] = __builtin_bit_cast(
                    V, __builtin_amdgcn_raw_buffer_load_b128(TWO && c != 0 ? xr1 : xr, off + ch * 64, 0, 0));
        }
    };
    // one step; u = v mod U fixes the ring slots.  GUARD: the step may be among the last
    // three, so the loads it issues and the wait before its barrier depend on kv.
    auto step = [&](int v, auto uc, auto gc) {
        constexpr int u = decltype(uc)::value;
        constexpr bool GUARD = decltype(gc)::value;
        if (!GUARD || v + 2 < kv) stage(v + 2, wring(std::integral_constant<int, u + 2>{}));
        if (!GUARD || v + LX < kv) gather(v + LX, b[(u + LX) % (LX + 1)], has[(u + LX) % (LX + 1)]);
        const T* wb = wring(uc);
        constexpr int sl = u % (LX + 1);
        bool any = false;
#pragma unroll
        for (int j = 0; j < NI; ++j) any |= has[sl][j];
        if (any) {
#pragma unroll
            for (int ch = 0; ch < NCH; ++ch) {
#pragma unroll
                for (int i = 0; i < MI; ++i) {
                    const int pos = (ch * 4 + g) ^ (r16 & 7);
                    const V a = *reinterpret_cast<const V*>(wb + (i * 16 + r16) * CC + pos * 8);
#pragma unroll
                    for (int j = 0; j < NI; ++j) mfma(a, b[sl][j][ch], acc[i][j]);
                }
            }
        }
        if (!GUARD || v + 1 < kv) {
            // step v+1's weights have landed in every wavefront and its gathered rows are
            // complete (the same wait as k_conv_tile_dma's: only loads issued after both)
            __builtin_amdgcn_sched_barrier(0);
            if constexpr (!GUARD) {
                constexpr int GA = (LX - 1) * (GPT + SQ), ST = 2 * GPT + SQ;
                wait_vm_lds<GA < ST ? GA : ST>();
            } else {
                int n = (v - 1 + LX < kv ? GPT : 0) + (v + 2 < kv ? SQ : 0) + (v + LX < kv ? GPT : 0);
                if (v + 1 >= LX) {
                    int ga = 0;
#pragma unroll
                    for (int d = 2 - LX; d <= 0; ++d) ga += (v + d + 2 < kv ? SQ : 0) + (v + d + LX < kv ? GPT : 0);
                    n = ga < n ? ga : n;
                }
                wait_vm_lds(n);
            }
            wg_barrier();
        }
    };
    // U steps from v0 (a multiple of U); guarded groups stop at the last step
    auto group = [&](int v0, auto gc) {
        constexpr bool GUARD = decltype(gc)::value;
        static_for<0, U>([&](auto uc) {
            if (!GUARD || v0 + decltype(uc)::value < kv) step(v0 + decltype(uc)::value, uc, gc);
        });
    };

    if (kv > 0) stage(0, sW0);
    if (kv > 1) stage(1, sW1);
#pragma unroll
    for (int v = 0; v < LX; ++v)
        if (v < kv) gather(v, b[v], has[v]);
    wait_vm_lds<0>();
    wg_barrier();
    int v0 = 0;
    // unguarded groups: every step v0..v0+U-1 issues its staging (v + 2 < kv) and gathers
    // (v + LX < kv) and ends with a barrier
    constexpr int AHEAD = LX > 2 ? LX : 2;
#pragma unroll 1
    for (; v0 + U - 1 + AHEAD < kv; v0 += U) group(v0, std::false_type{});
#pragma unroll 1
    for (; v0 < kv; v0 += U) group(v0, std::true_type{});
    __syncthreads();  // every wavefront is done with the weight buffers and the map tile
    if constexpr (MI == 4) {
        T* slice = wave == 0 ? sW0
                   : wave == 1 ? sW1
                   : wave == 2 ? sW2
                               : reinterpret_cast<T*>(tile_smem) + (wave - 3) * OUT_SLICE;
        store_tile_slice<T, MI, NI>(Y, acc, slice, lane, row0 + wave * NI * 16, perm, n_rows, M, m0, bias, relu);
    } else {
        T* const ring = wave < 3 ? sW0 : (wave < 6 ? sW1 : sW2);
        T* slice = ring + (wave % 3) * OUT_SLICE;
        using Half = floatx4[4][NI];
        store_tile_slice<T, 4, NI>(Y, *reinterpret_cast<const Half*>(&acc[0][0]), slice, lane, row0 + wave * NI * 16,
                                   perm, n_rows, M, m0, bias, relu);
        store_tile_slice<T, 4, NI>(Y, *reinterpret_cast<const Half*>(&acc[4][0]), slice, lane, row0 + wave * NI * 16,
                                   perm, n_rows, M, m0 + 64, bias, relu);
    }
}

// Launches k_conv_tile_dma when the layer fits it (K = 27 or 8); false otherwise.  Large
// layers take 8-wavefront workgroups of 256 rows, which stage each weight tile once for
// twice the rows (conv A of the 2.3M-voxel sphere batch: -7..-12% on MI355X); smaller
// ones keep 128-row workgroups so that the launch still fills the chip.  Large 2x2x2
// layers with 128k output channels take 128-channel workgroups (conv B of the sphere
// batch, 64->128: 0.155 -> 0.129 ms on MI355X).  Dynamic LDS holds the map tile, and at
// the end the output slices of wavefronts 3.. (wavefronts 0-2 use the weight ring; with
// 128-channel workgroups every slice is in the ring).
template <typename T>
static bool launch_tile_dma(const T* X, int64_t n_in, const T* A, const int32_t* map, int64_t n_rows, int K, int M,
                            T* Y, const T* bias, int relu, hipStream_t s) {
    if ((K != 27 && K != 8) || M % 64 != 0 || n_in * 64 * (int64_t)sizeof(T) > 0x7FFFFF00 ||
        ((uintptr_t)Y % 16) != 0)
        return false;
    const int n_mt = M / 64;
    auto go = [&](auto nw, auto kt) {
        constexpr int NW = decltype(nw)::value, KT = decltype(kt)::value;
        const int64_t n_tiles = (n_rows + 32 * NW - 1) / (32 * NW) * n_mt;
        const dim3 grid((unsigned)((n_tiles + 7) / 8 * 8));
        const size_t map_bytes = (size_t)32 * NW * KT * 4;
        const size_t out_bytes = (size_t)(NW - 3) * 2 * 16 * (64 + 8) * sizeof(T);
        const size_t lds = map_bytes > out_bytes ? map_bytes : out_bytes;
        // 3x3x3 maps skip, per wavefront, the steps none of its 32 rows has (conv A of the sphere
        // batch: 0.69 -> 0.65 ms on MI355X); 2x2x2 maps run every step (their skip form spills)
        hipLaunchKernelGGL((k_conv_tile_dma<T, 3, NW, KT, 4, 1, KT == 27>), grid, dim3(64 * NW), lds, s, X, n_in, A,
                           map, n_rows, M, n_mt, n_tiles, Y, bias, relu);
    };
    using W4 = std::integral_constant<int, 4>;
    using W8 = std::integral_constant<int, 8>;
    const bool wide = (n_rows + 255) / 256 * n_mt >= 1024;
    if (K == 8 && M % 128 == 0 && (n_rows + 255) / 256 * (M / 128) >= 1024) {
        constexpr int NW = 8, KT = 8;
        const int n_mt8 = M / 128;
        const int64_t n_tiles = (n_rows + 32 * NW - 1) / (32 * NW) * n_mt8;
        const dim3 grid((unsigned)((n_tiles + 7) / 8 * 8));
        const size_t lds = (size_t)32 * NW * KT * 4;  // map tile (output slices live in the ring)
        hipLaunchKernelGGL((k_conv_tile_dma<T, 2, NW, KT, 8>), grid, dim3(64 * NW), lds, s, X, n_in, A, map, n_rows,
                           M, n_mt8, n_tiles, Y, bias, relu);
        return true;
    }
    if (K == 27) {
        if (wide) go(W8{}, std::integral_constant<int, 27>{});
        else go(W4{}, std::integral_constant<int, 27>{});
    } else {
        if (wide) go(W8{}, std::integral_constant<int, 8>{});
        else go(W4{}, std::integral_constant<int, 8>{});
    }
    return true;
}

// Direct-to-LDS contraction for 16-bit 3x3x3 layers with 128 input channels in tree order
// (NCC = 2, workgroups of 8 wavefronts); large layers only (>= 1024 such tiles), false
// otherwise.  Outputs in multiples of 128 channels take 128-channel workgroups one step
// ahead, other multiples of 64 take 64-channel workgroups three steps ahead.  Planes +
// spheres (1.1M rows), MI355X: 128 -> 128 0.71 ms against 1.04 ms for the tap-skipping
// kernel over rows in tap-mask order and 1.74 ms for the register-staged kernel in tree
// order; 128 -> 64 0.51 ms against 0.61-0.67 ms.
template <typename T>
static bool launch_tile_dma_wide(const T* X, int64_t n_in, const T* A, const int32_t* map, int64_t n_rows, int K,
                                 int M, T* Y, const T* bias, int relu, hipStream_t s) {
    if (K != 27 || M % 64 != 0 || n_in * 128 * (int64_t)sizeof(T) > 0x7FFFFF00 || ((uintptr_t)Y % 16) != 0)
        return false;
    constexpr int NW = 8, KT = 27;
    const size_t map_bytes = (size_t)32 * NW * KT * 4;
    if (M % 128 == 0) {
        if ((n_rows + 255) / 256 * (M / 128) < 1024) return false;
        const int n_mt = M / 128;
        const int64_t n_tiles = (n_rows + 32 * NW - 1) / (32 * NW) * n_mt;
        const dim3 grid((unsigned)((n_tiles + 7) / 8 * 8));
        // dynamic LDS: the map tile (output slices live in the ring)
        hipLaunchKernelGGL((k_conv_tile_dma<T, 1, NW, KT, 8, 2, true>), grid, dim3(64 * NW), map_bytes, s, X, n_in,
                           A, map, n_rows, M, n_mt, n_tiles, Y, bias, relu);
        return true;
    }
    if ((n_rows + 255) / 256 * (M / 64) < 1024) return false;
    const int n_mt = M / 64;
    const int64_t n_tiles = (n_rows + 32 * NW - 1) / (32 * NW) * n_mt;
    const dim3 grid((unsigned)((n_tiles + 7) / 8 * 8));
    const size_t out_bytes = (size_t)(NW - 3) * 2 * 16 * (64 + 8) * sizeof(T);
    hipLaunchKernelGGL((k_conv_tile_dma<T, 3, NW, KT, 4, 2, true>), grid, dim3(64 * NW),
                       map_bytes > out_bytes ? map_bytes : out_bytes, s, X, n_in, A, map, n_rows, M, n_mt, n_tiles, Y,
                       bias, relu);
    return true;
}

// Narrow layers (16-bit, 32 output channels, CIN = 32 or 64 input channels, KT taps) with
// the whole weight tensor resident in LDS.  At 32 channels a feature row is 64 bytes and a
// map row 4*KT bytes, so these layers are bound by moving rows, not by the matrix cores;
// with the weights resident nothing is staged per tap and the wavefronts never meet at a
// barrier after the prologue.
//  * Persistent workgroups of 16 wavefronts, one per CU (4 waves per SIMD).  The prologue
//    copies the KT x 32 x CIN weights global -> LDS with the load unit (54 KB at 32 input
//    channels, 108 KB at 64) and waits once.  Rows of 64 bytes are swizzled: 16-byte segment
//    s of row r sits at position s ^ (((r >> 1) ^ (r >> 3)) & 3), so the MFMA operand reads
//    of 8 (or 16) consecutive rows hit distinct banks.
//  * Each wavefront owns tiles of 32 rows (2 groups of 16), a contiguous range per
//    workgroup, and runs them independently.  A tile's map rows are read straight into
//    registers (lane (g, r) holds taps [g*TPL, (g+1)*TPL) of its row), the next tile's while
//    the current one runs, and each tap's neighbour index reaches the row's four lanes
//    through a lane shuffle.
//  * Gathers are raw buffer loads LX taps ahead; the lanes of a missing neighbour are
//    switched off the load and keep zeros (the compiler's waits count the loads, as no
//    direct-to-LDS copy is in flight; the texture data path then moves only rows that exist:
//    level 0 of the U-Net, 32 -> 32: 0.36 -> 0.28 ms against loading zeros from beyond the
//    buffer).  The MFMA order per output element is taps ascending, then 32-channel chunks,
//    as in the other tile kernels, so the results are bit-identical to theirs.
//  * The output tile goes through a wavefront-private LDS slice and leaves as whole rows.
// Features must span fewer than 2^31 bytes (32-bit buffer offsets); callers check.
// TWO (NCC = 2): the input is the channel concatenation [X | X1] of two arrays of CC
// channels each, read from both without forming it (chunk 0 from X, chunk 1 from X1).
template <typename T, int KT, int NCC, int LX, bool TWO = false>
__global__ __launch_bounds__(1024) __attribute__((amdgpu_waves_per_eu(4, 4)))
void k_conv_resident(const T* __restrict__ X, const T* __restrict__ X1, int64_t n_in, const T* __restrict__ A,
                     const int32_t* __restrict__ map, int64_t n_rows, int64_t n_wt, T* __restrict__ Y,
                     const T* __restrict__ bias, int relu) {
    static_assert(sizeof(T) == 2, "16-bit features");
    using F = Frag<T>;
    using V = typename F::vec;
    constexpr int NW = 16, MI = 2, NI = 2, M = 32, CC = 32, CIN = CC * NCC;
    constexpr int TPL = (KT + 3) / 4;                 // map entries per lane and row group
    constexpr int SEGS = KT * NCC * 128;              // 16-byte segments of the weights
    constexpr int W_BYTES = SEGS * 16;
    constexpr int SLICE = NI * 16 * (M + 8);           // output slice (elements) per wavefront
    static_assert(W_BYTES + NW * SLICE * (int)sizeof(T) <= 160 * 1024, "LDS budget");
    extern __shared__ __attribute__((aligned(16))) unsigned char res_smem[];
    T* const sW = reinterpret_cast<T*>(res_smem);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int g = lane >> 4, r16 = lane & 15;
    auto swz = [](int r) { return ((r >> 1) ^ (r >> 3)) & 3; };

    // weights -> LDS: segment p of block b = (tap, chunk) lands at b*128 + p; it holds
    // row r = p/4, position p%4 = s ^ swz(r) of segment s (8 channels) of that row
    for (int p0 = wave * 64; p0 < SEGS; p0 += NW * 64) {
        const int p = p0 + lane;
        if (p < SEGS) {
            const int b = p >> 7, q = p & 127, r = q >> 2, s = (q & 3) ^ swz(r);
            const int k = b / NCC, c = b % NCC;
            lds_dma16(A + ((int64_t)k * M + r) * CIN + c * CC + s * 8, res_smem + (size_t)p0 * 16);
        }
    }
    wait_vm_lds<0>();
    __syncthreads();

    const int64_t per = (n_wt + gridDim.x - 1) / gridDim.x;
    const int64_t t_begin = (int64_t)blockIdx.x * per;
    const int64_t t_end = t_begin + per < n_wt ? t_begin + per : n_wt;
    static_assert(!TWO || NCC == 2, "two sources of one chunk each");
    constexpr int RB = (TWO ? CC : CIN) * (int)sizeof(T);  // bytes per row of a source
    const int bytes = (int)(n_in * RB);
    const __amdgpu_buffer_rsrc_t xr = __builtin_amdgcn_make_buffer_rsrc((void*)X, (short)0, bytes, 0x00020000);
    const __amdgpu_buffer_rsrc_t xr1 =
        __builtin_amdgcn_make_buffer_rsrc((void*)(TWO ? X1 : X), (short)0, bytes, 0x00020000);
    T* const slice = reinterpret_cast<T*>(res_smem + W_BYTES) + wave * SLICE;

    // map entries of tile t: lane (g, r16) of row group j holds row t*32 + j*16 + r16, taps
    // g*TPL + e (-1 beyond the rows or the taps)
    auto load_map = [&](int64_t t, int32_t (&mr)[NI][TPL]) {
#pragma unroll
        for (int j = 0; j < NI; ++j) {
            const int64_t row = t * 32 + j * 16 + r16;
            const bool ok = row < n_rows;
            const int32_t* src = map + (ok ? row : 0) * KT;
#pragma unroll
            for (int e = 0; e < TPL; ++e) {
                const int k = g * TPL + e;
                const int32_t v = src[k < KT ? k : KT - 1];
                mr[j][e] = (ok && k < KT) ? v : -1;
            }
        }
    };
    int32_t cur[NI][TPL], nxt[NI][TPL];
    int64_t t = t_begin + wave;
    if (t < t_end) load_map(t, nxt);
    for (; t < t_end; t += NW) {
#pragma unroll
        for (int j = 0; j < NI; ++j)
#pragma unroll
            for (int e = 0; e < TPL; ++e) cur[j][e] = nxt[j][e];

        floatx4 acc[MI][NI];
#pragma unroll
        for (int i = 0; i < MI; ++i)
#pragma unroll
            for (int j = 0; j < NI; ++j) acc[i][j] = floatx4{0.f, 0.f, 0.f, 0.f};
        V b[LX + 1][NI][NCC];
        auto gather = [&](auto kc, V (&bb)[NI][NCC]) {
            constexpr int k = decltype(kc)::value;
#pragma unroll
            for (int j = 0; j < NI; ++j) {
                const int32_t src = __shfl(cur[j][k % TPL], (k / TPL) * 16 + r16);
                const int off = src * RB + g * 16;
#pragma unroll
                for (int c = 0; c < NCC; ++c) bb[j][c] = zero_frag<T>();
                if (src >= 0) {
#pragma unroll
                    for (int c = 0; c < NCC; ++c)
                        bb[j][c] = __builtin_bit_cast(
                            V, TWO ? __builtin_amdgcn_raw_buffer_load_b128(c == 0 ? xr : xr1, off, 0, 0)
                                   : __builtin_amdgcn_raw_buffer_load_b128(xr, off + c * 64, 0, 0));
                }
            }
        };
        static_for<0, (LX < KT ? LX : KT)>([&](auto kc) { gather(kc, b[decltype(kc)::value]); });
        static_for<0, KT>([&](auto kc) {
            constexpr int k = decltype(kc)::value;
            if constexpr (k + LX < KT) gather(std::integral_constant<int, k + LX>{}, b[(k + LX) % (LX + 1)]);
            // the next tile's map, halfway through the taps: loads complete in issue order, so
            // issued at the start it would hold up the first gathers
            if constexpr (k == KT / 2) {
                if (t + NW < t_end) load_map(t + NW, nxt);
            }
#pragma unroll
            for (int c = 0; c < NCC; ++c) {
#pragma unroll
                for (int i = 0; i < MI; ++i) {
                    const int r = i * 16 + r16;
                    const V a = *reinterpret_cast<const V*>(sW + ((k * NCC + c) * 128 + r * 4 + (g ^ swz(r))) * 8);
#pragma unroll
                    for (int j = 0; j < NI; ++j) mfma(a, b[k % (LX + 1)][j][c], acc[i][j]);
                }
            }
        });
        store_tile_slice<T, MI, NI>(Y, acc, slice, lane, t * 32, nullptr, n_rows, M, 0, bias, relu);
    }
}

// Launches k_conv_resident for 16-bit layers of 32 output channels, 32 or 64 input
// channels and 27 or 8 taps in tree order; false if the layer does not fit it.  U-Net
// level 0 (3.2M rows), MI355X, against the register-staged tile kernel: 3x3x3 32 -> 32
// 0.45 -> 0.28 ms, 64 -> 32 0.92 -> 0.68 ms, 2x2x2 transposed 64 -> 32 0.29 -> 0.14 ms.
template <typename T>
static bool launch_resident(const T* X, int64_t n_in, int cin, const T* A, const int32_t* map, int64_t n_rows, int K,
                            int M, T* Y, const T* bias, int relu, hipStream_t s, const T* X1 = nullptr) {
    if (M != 32 || (cin != 32 && cin != 64) || (K != 27 && K != 8) ||
        n_in * cin * (int64_t)sizeof(T) > 0x7FFFFF00 || ((uintptr_t)Y % 16) != 0 || n_rows <= 0)
        return false;
    if (X1 != nullptr && (cin != 64 || K != 27)) return false;  // two sources: the 3x3x3 64-channel form
    const int64_t n_wt = (n_rows + 31) / 32;
    const int64_t want = (n_wt + 15) / 16;
    const int n_cu = device_cu_count();
    const unsigned grid = (unsigned)(want < n_cu ? want : n_cu);
    auto go = [&](auto kt, auto ncc, auto lx, auto two) {
        constexpr int KT = decltype(kt)::value, NCC = decltype(ncc)::value, LX = decltype(lx)::value;
        constexpr bool TWO = decltype(two)::value;
        const size_t lds = (size_t)KT * NCC * 128 * 16 + (size_t)16 * 2 * 16 * (32 + 8) * sizeof(T);
        // more than 64 KB of dynamic LDS: raise the kernel's limit once
        static const hipError_t attr = hipFuncSetAttribute(
            reinterpret_cast<const void*>(&k_conv_resident<T, KT, NCC, LX, TWO>),
            hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        (void)attr;
        hipLaunchKernelGGL((k_conv_resident<T, KT, NCC, LX, TWO>), dim3(grid), dim3(1024), lds, s, X, X1, n_in, A,
                           map, n_rows, n_wt, Y, bias, relu);
    };
    using F0 = std::false_type;
    using I1 = std::integral_constant<int, 1>;
    using I2 = std::integral_constant<int, 2>;
    using I4 = std::integral_constant<int, 4>;
    using I8 = std::integral_constant<int, 8>;
    using I27 = std::integral_constant<int, 27>;
    if (X1 != nullptr) {
        go(I27{}, I2{}, I1{}, std::true_type{});
        return true;
    }
    if (K == 27) {
        if (cin == 32) go(I27{}, I1{}, std::integral_constant<int, 3>{}, F0{});
        else go(I27{}, I2{}, I1{}, F0{});
    } else {
        if (cin == 32) go(I8{}, I1{}, I4{}, F0{});
        else go(I8{}, I2{}, I2{}, F0{});
    }
    return true;
}

// Direct-to-LDS contraction of 16-bit 3x3x3 layers over rows in a tap-mask order
// (k_conv_tile_dma_perm, workgroups of 8 wavefronts = 256 rows): 64 or 128 input
// channels, outputs in multiples of 64 channels (128-channel workgroups one step ahead
// where the output is a multiple of 128, else 64-channel workgroups two steps ahead).
// False if the layer does not fit it.
template <typename T>
static bool launch_tile_dma_perm(const T* X, int64_t n_in, int cin, const T* A, const int32_t* map,
                                 const int32_t* perm, int64_t n_rows, int K, int M, T* Y, const T* bias, int relu,
                                 hipStream_t s, const T* X1 = nullptr) {
    if (K != 27 || (cin != 64 && cin != 128) || M % 64 != 0 || n_in * cin * (int64_t)sizeof(T) > 0x7FFFFF00 ||
        ((uintptr_t)Y % 16) != 0)
        return false;
    if (X1 != nullptr && cin != 128) return false;  // two sources: the 128-channel (2 x 64) form
    constexpr int NW = 8, KT = 27;
    const size_t map_bytes = (size_t)32 * NW * KT * 4;
    auto go = [&](auto li, auto mi, auto ncc, auto wpe, auto two) {
        constexpr int LX = decltype(li)::value, MI = decltype(mi)::value, NCC = decltype(ncc)::value;
        constexpr int WPE = decltype(wpe)::value;
        constexpr bool TWO = decltype(two)::value;
        const int n_mt = M / (MI * 16);
        const int64_t n_tiles = (n_rows + 32 * NW - 1) / (32 * NW) * n_mt;
        const dim3 grid((unsigned)((n_tiles + 7) / 8 * 8));
        const size_t out_bytes = MI == 4 ? (size_t)(NW - 3) * 2 * 16 * (64 + 8) * sizeof(T) : 0;
        hipLaunchKernelGGL((k_conv_tile_dma_perm<T, LX, NW, KT, MI, NCC, WPE, TWO>), grid, dim3(64 * NW),
                           map_bytes > out_bytes ? map_bytes : out_bytes, s, X, X1, n_in, A, map, perm, n_rows, M,
                           n_mt, n_tiles, Y, bias, relu);
    };
    using I1 = std::integral_constant<int, 1>;
    using I2 = std::integral_constant<int, 2>;
    using I4 = std::integral_constant<int, 4>;
    using I8 = std::integral_constant<int, 8>;
    using F0 = std::false_type;
    if (cin == 128) {
        if (X1 != nullptr) {
            if (M % 128 == 0) go(I1{}, I8{}, I2{}, I4{}, std::true_type{});
            else go(I2{}, I4{}, I2{}, I4{}, std::true_type{});
        } else {
            if (M % 128 == 0) go(I1{}, I8{}, I2{}, I4{}, F0{});
            else go(I2{}, I4{}, I2{}, I4{}, F0{});
        }
    } else {
        if (M % 128 == 0) go(I1{}, I8{}, I1{}, I4{}, F0{});
        else go(I2{}, I4{}, I1{}, I4{}, F0{});
    }
    return true;
}

// Wide-channel form (CIN = NCC x CC): each pipeline step is one (tap, channel chunk)
// pair, chunk fastest, so a step has exactly the registers and LDS of the CC-channel
// kernel above (weights two steps ahead, gathered rows LX steps ahead, one barrier per
// step) however wide the input is.  Up to 64 steps are fully unrolled; longer
// sequences run a loop whose body is 6 steps unrolled (a multiple of the 3-entry weight
// ring, the 2-entry LDS ring and the LX+1-entry gather ring), which bounds the code
// size.  The accumulation order per row is taps ascending, then chunks ascending,
// then MFMA chunks: fixed, so results do not depend on the launch.
// SKIP (rows taken through a tap-mask ordering `perm`): the steps run over the taps
// that at least one row of the tile uses, listed in LDS at the start, so a skipped tap
// costs neither its gathers, its weight fetch, its barrier nor its MFMAs.  A skipped tap
// would have added exact zeros, so the result is bit-identical to the full sequence
// for finite weights.
// TWO: the input is the channel concatenation [X | X1] of two arrays of CIN / 2 channels,
// read from both without forming it (chunks 0 .. NCC/2-1 from X, the rest from X1).
template <typename T, int MI, int NI, int CC, int NCC, int KT, int LX, int WPE, bool SKIP = false, bool TWO = false>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(WPE > 0 ? WPE : 1, WPE > 0 ? WPE : 8)))
void k_conv_tile_c(const T* __restrict__ X, const T* __restrict__ A, const int32_t* __restrict__ map,
                   const int32_t* __restrict__ perm, int64_t n_rows, int M, int n_mt, int64_t n_tiles,
                   T* __restrict__ Y, const T* __restrict__ bias, int relu, const T* __restrict__ X1 = nullptr) {
    using F = Frag<T>;
    using V = typename F::vec;
    constexpr int CIN = CC * NCC;
    constexpr int KV = KT * NCC;
    constexpr int U = 6;
    constexpr int RT = 4 * NI * 16;
    constexpr int WP = CC + 16 / (int)sizeof(T);
    constexpr int WT = MI * 16 * WP;
    constexpr int NCH = CC / F::CH;
    constexpr int SEGS = MI * 16 * (CC / F::EPL);
    constexpr int SEG_PER_T = (SEGS + 255) / 256;
    static_assert(LX >= 1 && U % (LX + 1) == 0, "gather ring must divide the unrolled step count");
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

    static_assert(!SKIP || KT <= 32, "tap masks are 32 bits");
    static_assert((NCC & (NCC - 1)) == 0, "chunk count must be a power of two");
    __shared__ unsigned int s_tmask;
    if (SKIP) {
        if (tid == 0) s_tmask = 0u;
        __syncthreads();
    }
    if (perm == nullptr) {
        const int64_t base = row0 * KT;
        const int64_t lim = n_rows * KT;
        for (int e = tid; e < RT * KT; e += 256) sMap[e] = (base + e < lim) ? map[base + e] : -1;
    } else {
        unsigned int tm = 0u;
        for (int e = tid; e < RT * KT; e += 256) {
            const int i = e / KT, k = e - i * KT;
            const int64_t r = row0 + i;
            const int32_t v = r < n_rows ? map[(int64_t)perm[r] * KT + k] : -1;
            sMap[e] = v;
            if (SKIP && v >= 0) tm |= 1u << k;
        }
        if (SKIP) {
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) tm |= __shfl_xor(tm, off);
            if (lane == 0 && tm) atomicOr(&s_tmask, tm);
        }
    }
    __syncthreads();  // map tile and tap mask visible
    // Steps: all (tap, chunk) pairs, or those of the taps the tile uses.  The fetch and
    // gather streams each walk the tile's taps with a uniform (scalar) bit iterator: the
    // lowest set bit is the tap of the stream's next step, cleared after its last chunk.
    unsigned int fit = 0u, git = 0u;
    int kv = KV;
    if (SKIP) {
        fit = git = (unsigned int)__builtin_amdgcn_readfirstlane((int)s_tmask);
        kv = __popc(fit) * NCC;
    }
    auto next_tap = [&](unsigned int& it, int v) -> int {
        if constexpr (SKIP) {
            const int tap = __builtin_ctz(it);
            if ((v & (NCC - 1)) == NCC - 1) it &= it - 1u;
            return tap;
        } else {
            return v / NCC;
        }
    };

    floatx4 acc[MI][NI];
#pragma unroll
    for (int i = 0; i < MI; ++i)
#pragma unroll
        for (int j = 0; j < NI; ++j) acc[i][j] = floatx4{0.f, 0.f, 0.f, 0.f};

    V w[3][SEG_PER_T];
    V b[LX + 1][NI][NCH];
    // weights of step v (called once per step, in step order)
    auto fetch = [&](int v, V (&wr)[SEG_PER_T]) {
        const int tap = next_tap(fit, v), c0 = (v & (NCC - 1)) * CC;
#pragma unroll
        for (int q = 0; q < SEG_PER_T; ++q) {
            const int sgi = tid + q * 256;
            const int m = sgi / (CC / F::EPL), c = (sgi - m * (CC / F::EPL)) * F::EPL;
            const int mc = m0 + m;
            const bool ok = sgi < SEGS && mc < M;
            const V vv =
                *reinterpret_cast<const V*>(A + ((int64_t)tap * M + (ok ? mc : 0)) * CIN + c0 + (ok ? c : 0));
            wr[q] = ok ? vv : zero_frag<T>();
        }
    };
    auto store = [&](int buf, const V (&wr)[SEG_PER_T]) {
#pragma unroll
        for (int q = 0; q < SEG_PER_T; ++q) {
            const int sgi = tid + q * 256;
            if (SEGS % 256 == 0 || sgi < SEGS) {
                const int m = sgi / (CC / F::EPL), c = (sgi - m * (CC / F::EPL)) * F::EPL;
                *reinterpret_cast<V*>(sW + buf * WT + m * WP + c) = wr[q];
            }
        }
    };
    // gathered input rows of step v (called once per step, in step order)
    static_assert(!TWO || NCC % 2 == 0, "two sources of NCC / 2 chunks each");
    constexpr int XS = TWO ? CIN / 2 : CIN;  // row stride of a source
    auto gather = [&](int v, V (&bb)[NI][NCH]) {
        const int tap = next_tap(git, v), cc = v & (NCC - 1);
        const T* const xs = TWO && cc >= NCC / 2 ? X1 : X;
        const int c0 = (TWO ? cc & (NCC / 2 - 1) : cc) * CC;
#pragma unroll
        for (int j = 0; j < NI; ++j) {
            const int32_t src = sMap[(wave * NI * 16 + j * 16 + r16) * KT + tap];
#pragma unroll
            for (int ch = 0; ch < NCH; ++ch) bb[j][ch] = zero_frag<T>();
            if (src >= 0) {
                const T* xr = xs + (int64_t)src * XS + c0 + g * F::EPL;
#pragma unroll
                for (int ch = 0; ch < NCH; ++ch) bb[j][ch] = *reinterpret_cast<const V*>(xr + ch * F::CH);
            }
        }
    };
    // one step; u = v mod U selects the ring entries at compile time
    auto step = [&](int v, auto uc) {
        constexpr int u = decltype(uc)::value;
        if (v + 2 < kv) fetch(v + 2, w[(u + 2) % 3]);
        if (v + LX < kv) gather(v + LX, b[(u + LX) % (LX + 1)]);
        const T* wb = sW + (u % 2) * WT + r16 * WP + g * F::EPL;
        if constexpr (SKIP && (WPE > 0 || MI >= 8)) {
            // Under the 4-waves-per-SIMD register budget, and for 128-channel tiles, the
            // loop's scheduler would read one weight fragment at a time, each waiting for
            // the last: read all fragments of a channel chunk before its MFMAs so the LDS
            // reads overlap (measured on MI355X; the other shapes schedule better alone).
#pragma unroll
            for (int ch = 0; ch < NCH; ++ch) {
                V a[MI];
#pragma unroll
                for (int i = 0; i < MI; ++i) a[i] = *reinterpret_cast<const V*>(wb + i * 16 * WP + ch * F::CH);
                __builtin_amdgcn_sched_barrier(0);  // keep the reads ahead of the MFMAs
#pragma unroll
                for (int i = 0; i < MI; ++i)
#pragma unroll
                    for (int j = 0; j < NI; ++j) mfma(a[i], b[u % (LX + 1)][j][ch], acc[i][j]);
            }
        } else {
#pragma unroll
            for (int ch = 0; ch < NCH; ++ch) {
#pragma unroll
                for (int i = 0; i < MI; ++i) {
                    const V a = *reinterpret_cast<const V*>(wb + i * 16 * WP + ch * F::CH);
#pragma unroll
                    for (int j = 0; j < NI; ++j) mfma(a, b[u % (LX + 1)][j][ch], acc[i][j]);
                }
            }
        }
        if (v + 1 < kv) {
            store((u + 1) % 2, w[(u + 1) % 3]);
            __syncthreads();
        }
    };

    if (kv > 0) {
        fetch(0, w[0]);
        store(0, w[0]);
    }
    if (kv > 1) fetch(1, w[1]);
#pragma unroll
    for (int v = 0; v < LX && v < KV; ++v)
        if (v < kv) gather(v, b[v]);
    __syncthreads();
    // a step count known at compile time lets whole six-step groups drop their checks
    constexpr bool FULL = !SKIP && KV % U == 0;
    auto six = [&](int v0) {
        step(v0 + 0, std::integral_constant<int, 0>{});
        if (FULL || v0 + 1 < kv) step(v0 + 1, std::integral_constant<int, 1>{});
        if (FULL || v0 + 2 < kv) step(v0 + 2, std::integral_constant<int, 2>{});
        if (FULL || v0 + 3 < kv) step(v0 + 3, std::integral_constant<int, 3>{});
        if (FULL || v0 + 4 < kv) step(v0 + 4, std::integral_constant<int, 4>{});
        if (FULL || v0 + 5 < kv) step(v0 + 5, std::integral_constant<int, 5>{});
    };
    if constexpr (!SKIP && KV <= 64) {
        // fully unrolled: step numbers and taps are compile-time
#pragma unroll
        for (int v0 = 0; v0 < KV; v0 += U) six(v0);
    } else {
#pragma unroll 1
        for (int v0 = 0; v0 < kv; v0 += U) six(v0);
    }

    __syncthreads();  // every wavefront is done with the weight tiles
    store_tile<T, MI, NI>(Y, acc, tile_smem, wave, lane, row0 + wave * NI * 16, perm, n_rows, M, m0, bias, relu);
}

template <typename T, int MI, int NI, int CIN>
static size_t tile_lds_bytes(int K) {
    constexpr int WP = CIN + 16 / (int)sizeof(T);
    const size_t main_loop = (size_t)2 * MI * 16 * WP * sizeof(T) + (size_t)4 * NI * 16 * K * 4;
    return main_loop > out_tile_lds_bytes<T, MI, NI>() ? main_loop : out_tile_lds_bytes<T, MI, NI>();
}

template <typename T, int MI, int NI, int CIN>
static void launch_tile(const T* X, const T* A, const int32_t* map, const int32_t* perm, int64_t n_rows, int K, int M,
                        T* Y, const T* bias, int relu, hipStream_t s) {
    const int64_t n_rt = (n_rows + 64 * NI - 1) / (64 * NI);
    const int n_mt = (M + MI * 16 - 1) / (MI * 16);
    const int64_t n_tiles = n_rt * n_mt;
    const dim3 grid((unsigned)((n_tiles + 7) / 8 * 8));
    const size_t lds = tile_lds_bytes<T, MI, NI, CIN>(K);
    // Gather lookahead and occupancy of the unrolled kernel (measured on MI355X):
    // 16-bit types up to 64 channels in 64x128 tiles run best one tap ahead at 4
    // waves per SIMD; wider tiles keep the default allocation, 2 taps ahead for
    // 16-bit types and 1 for fp32.
    constexpr bool W4 = sizeof(T) == 2 && MI <= 4 && (CIN <= 32 || (CIN == 64 && NI == 2) || (CIN == 64 && MI == 1));
    constexpr int LX = W4 ? 1 : (sizeof(T) == 2 ? 2 : 1);
    constexpr int WPE = W4 ? 4 : 0;
    // rows through a tap-mask ordering: the compacted-step kernel (one chunk per tap),
    // which skips the weights, barriers, gathers and MFMAs of the taps a tile does not use
    if (K == 27 && perm) {
        hipLaunchKernelGGL((k_conv_tile_c<T, MI, NI, CIN, 1, 27, LX, WPE, true>), grid, dim3(256), lds, s, X, A, map,
                           perm, n_rows, M, n_mt, n_tiles, Y, bias, relu);
    } else if (K == 27) {
        hipLaunchKernelGGL((k_conv_tile_k<T, MI, NI, CIN, 27, LX, WPE>), grid, dim3(256), lds, s, X, A, map, n_rows,
                           M, n_mt, n_tiles, Y, bias, relu);
    } else if (K == 8 && perm) {
        hipLaunchKernelGGL((k_conv_tile_c<T, MI, NI, CIN, 1, 8, LX, WPE, true>), grid, dim3(256), lds, s, X, A, map,
                           perm, n_rows, M, n_mt, n_tiles, Y, bias, relu);
    } else if (K == 8) {
        hipLaunchKernelGGL((k_conv_tile_k<T, MI, NI, CIN, 8, LX, WPE>), grid, dim3(256), lds, s, X, A, map, n_rows,
                           M, n_mt, n_tiles, Y, bias, relu);
    } else {
        hipLaunchKernelGGL((k_conv_tile<T, MI, NI, CIN, 1>), grid, dim3(256), lds, s, X, A, map, perm, n_rows, K, M,
                           n_mt, n_tiles, Y, bias, relu);
    }
}

constexpr size_t TILE_LDS_LIMIT = 65536;

// Channel tile for a compile-time CIN: 64 output channels x 128 rows when the LDS
// fits (the best shape measured on MI355X; 64 rows for fp32), else 32 x 64 rows for
// fp32 and 32 x 256 or 16 x 256 rows otherwise.
template <typename T, int CIN>
static bool try_tile(const T* X, int64_t n_in, const T* A, const int32_t* map, const int32_t* perm, int64_t n_rows,
                     int K, int M, T* Y, const T* bias, int relu, hipStream_t s) {
    if constexpr (CIN == 64 && sizeof(T) == 2) {
        // 128-channel tiles for outputs in multiples of 128 channels and 27 taps
        if (M % 128 == 0 && K == 27) {
            launch_tile<T, 8, 2, CIN>(X, A, map, perm, n_rows, K, M, Y, bias, relu, s);
            return true;
        }
        // tree-order rows, whole 64-channel output tiles: direct-to-LDS weights, deep gathers
        // (2x2x2 stride-2 64->128 of the sphere batch: 0.19 -> 0.16 ms on MI355X)
        if (perm == nullptr && launch_tile_dma<T>(X, n_in, A, map, n_rows, K, M, Y, bias, relu, s)) return true;
    }
    if (sizeof(T) == 4 && M > 32 && tile_lds_bytes<T, 4, 1, CIN>(K) <= TILE_LDS_LIMIT) {
        // fp32 with more than 32 outputs in 64-row tiles (2.3M-voxel sphere batch, 3x3x3
        // 64->64: 4.10 -> 3.72 ms through the windowed row order, 4.32 -> 4.23 ms in tree
        // order; the fp32 U-Net 19.6 -> 19.2 ms)
        launch_tile<T, 4, 1, CIN>(X, A, map, perm, n_rows, K, M, Y, bias, relu, s);
    } else if (M > 32 && tile_lds_bytes<T, 4, 2, CIN>(K) <= TILE_LDS_LIMIT) {
        launch_tile<T, 4, 2, CIN>(X, A, map, perm, n_rows, K, M, Y, bias, relu, s);
    } else if (sizeof(T) == 4 && M > 16 && tile_lds_bytes<T, 2, 1, CIN>(K) <= TILE_LDS_LIMIT) {
        // fp32 with 17-32 outputs in 64-row tiles (one 338k-voxel grid, 3x3x3 32->32: 0.255 ms
        // in 256-row tiles, 0.225 in 128-row, 0.203 in 64-row; the fp32 U-Net 18.1 -> 17.5 ms
        // against 128-row tiles)
        launch_tile<T, 2, 1, CIN>(X, A, map, perm, n_rows, K, M, Y, bias, relu, s);
    } else if (M > 16 && tile_lds_bytes<T, 2, 4, CIN>(K) <= TILE_LDS_LIMIT) {
        launch_tile<T, 2, 4, CIN>(X, A, map, perm, n_rows, K, M, Y, bias, relu, s);
    } else if (tile_lds_bytes<T, 1, 4, CIN>(K) <= TILE_LDS_LIMIT) {
        launch_tile<T, 1, 4, CIN>(X, A, map, perm, n_rows, K, M, Y, bias, relu, s);
    } else {
        return false;
    }
    return true;
}

// Wide inputs (CIN = NCC x 64) with K = 27 or 8: the chunked kernel, tile shape and
// occupancy chosen as for 64 input channels.
template <typename T, int MI, int NI, int NCC>
static void launch_tile_c(const T* X, const T* A, const int32_t* map, const int32_t* perm, int64_t n_rows, int K,
                          int M, T* Y, const T* bias, int relu, hipStream_t s) {
    constexpr int CC = 64;
    const int64_t n_rt = (n_rows + 64 * NI - 1) / (64 * NI);
    const int n_mt = (M + MI * 16 - 1) / (MI * 16);
    const int64_t n_tiles = n_rt * n_mt;
    const dim3 grid((unsigned)((n_tiles + 7) / 8 * 8));
    const size_t lds = tile_lds_bytes<T, MI, NI, CC>(K);
    // Occupancy and lookahead measured per shape on MI355X: 16-bit tiles of 64 output
    // channels run at 4 waves per SIMD one step ahead, except 27 taps x 2 chunks, which
    // is faster with the default allocation two steps ahead, as are 128-channel tiles;
    // fp32 runs one step ahead.
    constexpr bool W4_27 = sizeof(T) == 2 && MI <= 4 && (NI == 2 || MI == 1) && NCC != 2;
    constexpr bool W4_8 = sizeof(T) == 2 && MI <= 4 && (NI == 2 || MI == 1);
    constexpr int LX27 = W4_27 ? 1 : (sizeof(T) == 2 ? 2 : 1);
    constexpr int LX8 = W4_8 ? 1 : (sizeof(T) == 2 ? 2 : 1);
    // rows through a tap-mask ordering skip the taps a tile does not use
    if (K == 27 && perm)
        hipLaunchKernelGGL((k_conv_tile_c<T, MI, NI, CC, NCC, 27, LX27, W4_27 ? 4 : 0, true>), grid, dim3(256), lds, s,
                           X, A, map, perm, n_rows, M, n_mt, n_tiles, Y, bias, relu);
    else if (K == 27)
        hipLaunchKernelGGL((k_conv_tile_c<T, MI, NI, CC, NCC, 27, LX27, W4_27 ? 4 : 0>), grid, dim3(256), lds, s, X, A,
                           map, perm, n_rows, M, n_mt, n_tiles, Y, bias, relu);
    else if (perm)
        hipLaunchKernelGGL((k_conv_tile_c<T, MI, NI, CC, NCC, 8, LX8, W4_8 ? 4 : 0, true>), grid, dim3(256), lds, s, X,
                           A, map, perm, n_rows, M, n_mt, n_tiles, Y, bias, relu);
    else
        hipLaunchKernelGGL((k_conv_tile_c<T, MI, NI, CC, NCC, 8, LX8, W4_8 ? 4 : 0>), grid, dim3(256), lds, s, X, A,
                           map, perm, n_rows, M, n_mt, n_tiles, Y, bias, relu);
}

// Two-source form of launch_tile_c for 16-bit 3x3x3 layers of 256 input channels (two
// sources of 128) and more than 32 (not 128k) output channels: the 64-channel-tile
// instance the single-source dispatch picks for that shape, reading [X | X1].
template <typename T>
static bool try_tile_c_two(const T* X, const T* X1, const T* A, const int32_t* map, const int32_t* perm,
                           int64_t n_rows, int K, int M, T* Y, const T* bias, int relu, hipStream_t s) {
    constexpr int MI = 4, NI = 2, NCC = 4, CC = 64;
    if (sizeof(T) != 2 || K != 27 || M <= 32) return false;
    const int64_t n_rt = (n_rows + 64 * NI - 1) / (64 * NI);
    const int n_mt = (M + MI * 16 - 1) / (MI * 16);
    const int64_t n_tiles = n_rt * n_mt;
    const dim3 grid((unsigned)((n_tiles + 7) / 8 * 8));
    const size_t lds = tile_lds_bytes<T, MI, NI, CC>(K);
    constexpr int LX27 = 1, WPE = 4;  // the single-source choice for this shape (W4_27)
    if (perm)
        hipLaunchKernelGGL((k_conv_tile_c<T, MI, NI, CC, NCC, 27, LX27, WPE, true, true>), grid, dim3(256), lds, s, X,
                           A, map, perm, n_rows, M, n_mt, n_tiles, Y, bias, relu, X1);
    else
        hipLaunchKernelGGL((k_conv_tile_c<T, MI, NI, CC, NCC, 27, LX27, WPE, false, true>), grid, dim3(256), lds, s,
                           X, A, map, perm, n_rows, M, n_mt, n_tiles, Y, bias, relu, X1);
    return true;
}

// 16-bit outputs in multiples of 128 channels take 128-channel tiles (each gathered
// row feeds twice the MFMAs) where that measured faster: two chunks per tap, or 8 taps.
template <typename T>
static bool wide_m_tile(int M, int K, int ncc) {
    return sizeof(T) == 2 && M % 128 == 0 && (ncc == 2 || K == 8);
}

template <typename T, int NCC>
static bool try_tile_c(const T* X, const T* A, const int32_t* map, const int32_t* perm, int64_t n_rows, int K, int M,
                       T* Y, const T* bias, int relu, hipStream_t s) {
    if constexpr (sizeof(T) == 2) {
        if (wide_m_tile<T>(M, K, NCC)) {
            launch_tile_c<T, 8, 2, NCC>(X, A, map, perm, n_rows, K, M, Y, bias, relu, s);
            return true;
        }
    }
    // fp32 with more than 32 outputs in 64-row tiles (1.1M-voxel 3x3x3 128->128 layer through
    // the windowed row order: forward and dgrad 5.88 -> 5.31 ms)
    if (sizeof(T) == 4 && M > 32) launch_tile_c<T, 4, 1, NCC>(X, A, map, perm, n_rows, K, M, Y, bias, relu, s);
    else if (M > 32) launch_tile_c<T, 4, 2, NCC>(X, A, map, perm, n_rows, K, M, Y, bias, relu, s);
    else if (M > 16) launch_tile_c<T, 2, 4, NCC>(X, A, map, perm, n_rows, K, M, Y, bias, relu, s);
    else launch_tile_c<T, 1, 4, NCC>(X, A, map, perm, n_rows, K, M, Y, bias, relu, s);
    return true;
}

// Looped kernel only (runtime K), for wide inputs with other kernel volumes.
template <typename T, int CIN>
static bool try_tile_looped(const T* X, const T* A, const int32_t* map, const int32_t* perm, int64_t n_rows, int K,
                            int M, T* Y, const T* bias, int relu, hipStream_t s) {
    auto go = [&](auto mi, auto ni) {
        constexpr int MI = decltype(mi)::value, NI = decltype(ni)::value;
        const int64_t n_rt = (n_rows + 64 * NI - 1) / (64 * NI);
        const int n_mt = (M + MI * 16 - 1) / (MI * 16);
        const int64_t n_tiles = n_rt * n_mt;
        const dim3 grid((unsigned)((n_tiles + 7) / 8 * 8));
        const size_t lds = tile_lds_bytes<T, MI, NI, CIN>(K);
        hipLaunchKernelGGL((k_conv_tile<T, MI, NI, CIN, 1>), grid, dim3(256), lds, s, X, A, map, perm, n_rows, K, M,
                           n_mt, n_tiles, Y, bias, relu);
    };
    if (M > 32 && tile_lds_bytes<T, 4, 2, CIN>(K) <= TILE_LDS_LIMIT) go(std::integral_constant<int, 4>{}, std::integral_constant<int, 2>{});
    else if (M > 16 && tile_lds_bytes<T, 2, 4, CIN>(K) <= TILE_LDS_LIMIT) go(std::integral_constant<int, 2>{}, std::integral_constant<int, 4>{});
    else if (tile_lds_bytes<T, 1, 4, CIN>(K) <= TILE_LDS_LIMIT) go(std::integral_constant<int, 1>{}, std::integral_constant<int, 4>{});
    else return false;
    return true;
}

// LDS-tiled path for the channel counts it is compiled for; false if not applicable.
template <typename T>
inline bool conv_tile_dispatch_t(const T* X, int64_t n_in, int cin, const T* A, const int32_t* map,
                                 const int32_t* perm, int64_t n_rows, int K, int M, T* Y, const T* bias, int relu,
                                 hipStream_t s) {
    if (K > TILE_MAX_K || ((uintptr_t)X % 16) != 0 || ((uintptr_t)A % 16) != 0) return false;
    if constexpr (sizeof(T) == 2) {
        // rows through a tap-mask ordering: direct-to-LDS kernel with compacted steps
        if (perm != nullptr && launch_tile_dma_perm<T>(X, n_in, cin, A, map, perm, n_rows, K, M, Y, bias, relu, s))
            return true;
        // narrow layers in tree order: weights resident in LDS
        if (perm == nullptr && launch_resident<T>(X, n_in, cin, A, map, n_rows, K, M, Y, bias, relu, s)) return true;
    }
    switch (cin) {
    case 32: return try_tile<T, 32>(X, n_in, A, map, perm, n_rows, K, M, Y, bias, relu, s);
    case 64: return try_tile<T, 64>(X, n_in, A, map, perm, n_rows, K, M, Y, bias, relu, s);
    case 128:
        if constexpr (sizeof(T) == 2) {
            if (perm == nullptr && launch_tile_dma_wide<T>(X, n_in, A, map, n_rows, K, M, Y, bias, relu, s)) return true;
        }
        if (K == 27 || K == 8) return try_tile_c<T, 2>(X, A, map, perm, n_rows, K, M, Y, bias, relu, s);
        return try_tile_looped<T, 128>(X, A, map, perm, n_rows, K, M, Y, bias, relu, s);
    case 256:
        if (K == 27 || K == 8) return try_tile_c<T, 4>(X, A, map, perm, n_rows, K, M, Y, bias, relu, s);
        return try_tile_looped<T, 256>(X, A, map, perm, n_rows, K, M, Y, bias, relu, s);
    default: return false;
    }
}

// Contraction over the channel concatenation [X | X1] of two arrays of cin / 2 channels
// each, without forming it, where the kernel the single-source dispatch picks for the
// shape has a two-source form (16-bit 3x3x3 layers: 64 inputs and 32 outputs in tree
// order, 128 inputs through a row order, 256 inputs and more than 32 outputs); false
// otherwise.  Same instance and accumulation order as for the concatenated input, so the
// result is bit-identical to it.
template <typename T>
inline bool conv_tile_dispatch2_t(const T* X, const T* X1, int64_t n_in, int cin, const T* A, const int32_t* map,
                                  const int32_t* perm, int64_t n_rows, int K, int M, T* Y, const T* bias, int relu,
                                  hipStream_t s) {
    if constexpr (sizeof(T) != 2) {
        return false;
    } else {
        if (K != 27 || ((uintptr_t)X % 16) != 0 || ((uintptr_t)X1 % 16) != 0 || ((uintptr_t)A % 16) != 0)
            return false;
        if (perm == nullptr && cin == 64 && M == 32)
            return launch_resident<T>(X, n_in, cin, A, map, n_rows, K, M, Y, bias, relu, s, X1);
        if (perm != nullptr && cin == 128)
            return launch_tile_dma_perm<T>(X, n_in, cin, A, map, perm, n_rows, K, M, Y, bias, relu, s, X1);
        if (cin == 256) return try_tile_c_two<T>(X, X1, A, map, perm, n_rows, K, M, Y, bias, relu, s);
        return false;
    }
}

// Per element type (one translation unit each): launches the LDS-tiled kernel and
// returns true, or returns false when the shape has no tiled instance.
bool conv_tile_dispatch(const __bf16* X, int64_t n_in, int cin, const __bf16* A, const int32_t* map, const int32_t* perm,
                        int64_t n_rows, int K, int M, __bf16* Y, const __bf16* bias, int relu, hipStream_t s);
bool conv_tile_dispatch(const _Float16* X, int64_t n_in, int cin, const _Float16* A, const int32_t* map, const int32_t* perm,
                        int64_t n_rows, int K, int M, _Float16* Y, const _Float16* bias, int relu, hipStream_t s);
bool conv_tile_dispatch(const float* X, int64_t n_in, int cin, const float* A, const int32_t* map, const int32_t* perm,
                        int64_t n_rows, int K, int M, float* Y, const float* bias, int relu, hipStream_t s);
bool conv_tile_dispatch2(const __bf16* X, const __bf16* X1, int64_t n_in, int cin, const __bf16* A,
                         const int32_t* map, const int32_t* perm, int64_t n_rows, int K, int M, __bf16* Y,
                         const __bf16* bias, int relu, hipStream_t s);
bool conv_tile_dispatch2(const _Float16* X, const _Float16* X1, int64_t n_in, int cin, const _Float16* A,
                         const int32_t* map, const int32_t* perm, int64_t n_rows, int K, int M, _Float16* Y,
                         const _Float16* bias, int relu, hipStream_t s);

}  // namespace fvdbk
