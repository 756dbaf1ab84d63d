// Grid index build, point voxelisation, conv target-grid candidates, kernel map and
// gather-scatter neighbour map for gfx950.  All integer work; results are bit-exact
// and independent of launch shape.
#include <hip/hip_runtime.h>

#include <climits>
#include <stdarg.h>
#include <stdio.h>

#include <string>

#include <hipcub/block/block_radix_sort.hpp>
#include <hipcub/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>
#include <rocprim/iterator/counting_iterator.hpp>
#include <rocprim/iterator/transform_iterator.hpp>

#include "fvdbk_common.h"

namespace fvdbk {

static thread_local std::string g_last_error;

int set_error(int code, const char* fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return code;
}

int check_launch(const char* what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return set_error(FVDBK_EHIP, "%s: %s", what, hipGetErrorString(e));
    return FVDBK_OK;
}

// ---------------------------------------------------------------------------------
// grid build
// ---------------------------------------------------------------------------------

__global__ void k_leaf_keys(const int32_t* __restrict__ ijk, const int32_t* __restrict__ bidx, int64_t n,
                            uint64_t* __restrict__ keys, int32_t* __restrict__ bad) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int32_t b = bidx ? bidx[i] : 0;
    if (b < 0) {
        keys[i] = kSkipKey;
        return;
    }
    const int32_t x = ijk[3 * i], y = ijk[3 * i + 1], z = ijk[3 * i + 2];
    if (b >= kMaxGrids || x < kCoordMin || x > kCoordMax || y < kCoordMin || y > kCoordMax || z < kCoordMin ||
        z > kCoordMax) {
        atomicAdd(bad, 1);
        keys[i] = kSkipKey;
        return;
    }
    keys[i] = leaf_key(b, x, y, z);
}

// Unique leaf keys of the input voxels, unsorted.  A voxel's key is dropped when the
// lane before it holds the same key (inputs derived from a grid arrive in tree order,
// so runs are long); the rest claim a slot of an open-addressing set by a 64-bit CAS
// (kEmptyKey = free).  Each new key is appended to keys_out: a workgroup reserves
// room for its new keys with one atomic on counters[0] and lanes take consecutive
// places by a ballot prefix count.  Workgroups are 1024 lanes: every workgroup's
// append hits the same counter, and with 256-lane workgroups those atomics took two
// thirds of the kernel (83 -> 47 us for 2.3M voxels on MI355X).  counters[1] counts voxels out of range (grid index
// >= n_grids or coordinate outside [-2^19, 2^19)); grid index < 0 skips a voxel.
constexpr uint64_t kEmptyKey = ~0ull;

constexpr int DEDUPE_THREADS = 1024;

__global__ __launch_bounds__(DEDUPE_THREADS) void k_leaf_dedupe(const int32_t* __restrict__ ijk,
                                                     const int32_t* __restrict__ bidx, int64_t n, int32_t n_grids,
                                                     unsigned long long* __restrict__ set, int64_t cap,
                                                     uint64_t* __restrict__ keys_out,
                                                     unsigned int* __restrict__ counters) {
    __shared__ unsigned int s_cnt[DEDUPE_THREADS / 64];
    __shared__ unsigned int s_base;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint64_t key = kSkipKey;
    if (i < n) {
        const int32_t b = bidx ? bidx[i] : 0;
        if (b >= 0) {
            const int32_t x = ijk[3 * i], y = ijk[3 * i + 1], z = ijk[3 * i + 2];
            if (b >= n_grids || b >= kMaxGrids || x < kCoordMin || x > kCoordMax || y < kCoordMin ||
                y > kCoordMax || z < kCoordMin || z > kCoordMax)
                atomicAdd(&counters[1], 1u);
            else
                key = leaf_key(b, x, y, z);
        }
    }
    const uint64_t prev = __shfl_up(key, 1);
    bool fresh = false;
    if (key != kSkipKey && (lane == 0 || prev != key)) {
        uint64_t slot = mix64(key) & (uint64_t)(cap - 1);
        for (;;) {
            const unsigned long long old = atomicCAS(&set[slot], kEmptyKey, (unsigned long long)key);
            if (old == kEmptyKey) {
                fresh = true;
                break;
            }
            if (old == key) break;
            slot = (slot + 1) & (uint64_t)(cap - 1);
        }
    }
    const uint64_t ball = __ballot(fresh);
    if (lane == 0) s_cnt[wave] = (unsigned int)__popcll(ball);
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned int tot = 0u;
        for (int w = 0; w < DEDUPE_THREADS / 64; ++w) tot += s_cnt[w];
        s_base = tot ? atomicAdd(&counters[0], tot) : 0u;
    }
    __syncthreads();
    if (fresh) {
        unsigned int off = s_base;
        for (int w = 0; w < wave; ++w) off += s_cnt[w];
        off += (unsigned int)__popcll(ball & ((1ull << lane) - 1ull));
        keys_out[off] = key;
    }
}

__global__ void k_fill_u64(unsigned long long* __restrict__ p, int64_t n, unsigned long long v) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void k_hash_clear(fvdbk_hslot* __restrict__ h, int64_t cap) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= cap) return;
    h[i].key = 0;
    h[i].leaf = -1;
    h[i].pad = 0;
}

// Unique keys: a slot is claimed by a CAS on its leaf field; the key is stored after
// the claim (no reader runs during the build).
__global__ void k_hash_insert(const uint64_t* __restrict__ keys, int64_t n, fvdbk_hslot* __restrict__ h,
                              int64_t cap) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint64_t key = keys[i];
    uint64_t s = mix64(key) & (uint64_t)(cap - 1);
    for (;;) {
        if (atomicCAS(&h[s].leaf, -1, (int32_t)i) == -1) {
            h[s].key = key;
            return;
        }
        s = (s + 1) & (uint64_t)(cap - 1);
    }
}

__global__ void k_zero_u64(uint64_t* __restrict__ p, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

// Mask bits of every input voxel.  Inputs usually arrive in tree order (candidates
// derived from an existing grid), so neighbouring lanes often hit the same mask word:
// lanes OR their bits with the lanes before them that share the word (segmented
// shuffle scan) and only the last lane of each run issues the atomic.
__global__ __launch_bounds__(256) void k_leaf_masks(const int32_t* __restrict__ ijk, const int32_t* __restrict__ bidx,
                                                    int64_t n, const fvdbk_hslot* __restrict__ h, int64_t cap,
                                                    uint64_t* __restrict__ masks) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int lane = threadIdx.x & 63;
    int64_t word = -1;
    uint64_t bits = 0;
    if (i < n) {
        const int32_t b = bidx ? bidx[i] : 0;
        const int32_t x = ijk[3 * i], y = ijk[3 * i + 1], z = ijk[3 * i + 2];
        if (b >= 0 && b < kMaxGrids && x >= kCoordMin && x <= kCoordMax && y >= kCoordMin && y <= kCoordMax &&
            z >= kCoordMin && z <= kCoordMax) {
            const int32_t leaf = hash_find(h, cap, leaf_key(b, x, y, z));
            if (leaf >= 0) {
                const int bit = ((x & 7) << 6) | ((y & 7) << 3) | (z & 7);
                word = (int64_t)leaf * 8 + (bit >> 6);
                bits = 1ull << (bit & 63);
            }
        }
    }
    // inclusive segmented OR over runs of equal `word` (runs are contiguous lanes)
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint64_t ob = __shfl_up(bits, d);
        const int64_t ow = __shfl_up(word, d);
        if (lane >= d && ow == word) bits |= ob;
    }
    const int64_t next = __shfl_down(word, 1);
    const bool tail = lane == 63 || next != word;
    if (word >= 0 && tail) atomicOr((unsigned long long*)&masks[word], bits);
}

// Coarsening by power-of-two factors (1 << sh[a] per axis a = x, y, z) straight from a
// fine grid's leaves instead of its voxels.  Every voxel of a fine leaf falls into one
// coarse leaf: the fine leaf spans 8 aligned coordinates per axis, and 8 >> sh (or a
// single coordinate) of them stay within an aligned block of 8 coarse coordinates.
// Per fine leaf, one coarse voxel of that coarse leaf (the leaf origin shifted down),
// for the leaf de-duplication.
__global__ void k_coarse_reps(const uint64_t* __restrict__ fine_keys, int64_t n, int sx, int sy, int sz,
                              int32_t* __restrict__ rep_ijk, int32_t* __restrict__ rep_bidx) {
    const int64_t l = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= n) return;
    int32_t b, x, y, z;
    decode_leaf_key(fine_keys[l], b, x, y, z);
    rep_ijk[3 * l + 0] = x >> sx;  // arithmetic shift: floor division
    rep_ijk[3 * l + 1] = y >> sy;
    rep_ijk[3 * l + 2] = z >> sz;
    rep_bidx[l] = b;
}

// Coarse leaf masks from the fine masks: one thread per (fine leaf, x slab w).  The
// slab's 64 bits (y, z) map to coarse bits (((y0 + y) >> sy) & 7, ((z0 + z) >> sz) & 7)
// of coarse slab ((x0 + w) >> sx) & 7; the thread ORs them over its set bits and
// issues one atomic OR (slabs that share a coarse slab merge there).
__global__ void k_coarse_masks(const uint64_t* __restrict__ fine_keys, const uint64_t* __restrict__ fine_masks,
                               int64_t n, int sx, int sy, int sz, const fvdbk_hslot* __restrict__ h, int64_t cap,
                               uint64_t* __restrict__ coarse_masks) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n * 8) return;
    const int64_t l = t >> 3;
    const int w = (int)(t & 7);
    uint64_t m = fine_masks[l * 8 + w];
    if (m == 0ull) return;
    int32_t b, x0, y0, z0;
    decode_leaf_key(fine_keys[l], b, x0, y0, z0);
    const int32_t cx = (x0 + w) >> sx, cy0 = y0 >> sy, cz0 = z0 >> sz;
    const int32_t leaf = hash_find(h, cap, leaf_key(b, cx, cy0, cz0));
    if (leaf < 0) return;  // not reached: every coarse leaf was inserted
    uint64_t acc = 0ull;
    while (m) {
        const int p = __ffsll((unsigned long long)m) - 1;
        m &= m - 1ull;
        const int y = p >> 3, z = p & 7;
        acc |= 1ull << (((((y0 + y) >> sy) & 7) << 3) | (((z0 + z) >> sz) & 7));
    }
    atomicOr((unsigned long long*)&coarse_masks[(int64_t)leaf * 8 + (cx & 7)], (unsigned long long)acc);
}

__global__ void k_leaf_counts(const uint64_t* __restrict__ masks, int64_t n_leaves, int32_t* __restrict__ counts) {
    const int64_t l = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= n_leaves) return;
    int c = 0;
#pragma unroll
    for (int w = 0; w < 8; ++w) c += __popcll(masks[l * 8 + w]);
    counts[l] = c;
}

// Active voxels of leaf l (the popcount of its 8 mask words), as a scan input.
struct LeafPopcount {
    const uint64_t* masks;
    __host__ __device__ int64_t operator()(int64_t l) const {
        int64_t c = 0;
#pragma unroll
        for (int w = 0; w < 8; ++w) c += __popcll(masks[l * 8 + w]);
        return c;
    }
};

// leaf_start[l] = leaf_incl[l] - (voxels of leaf l): the exclusive scan from the inclusive one.
__global__ void k_leaf_start(const uint64_t* __restrict__ masks, const int64_t* __restrict__ incl, int64_t n_leaves,
                             int64_t* __restrict__ start) {
    const int64_t l = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (l < n_leaves) start[l] = incl[l] - LeafPopcount{masks}(l);
}

__global__ void k_zero_i32(int32_t* __restrict__ p, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

// One wavefront per leaf: lane l handles bit l of each of the 8 mask words.
__global__ __launch_bounds__(256) void k_finalize(const uint64_t* __restrict__ keys,
                                                  const uint64_t* __restrict__ masks,
                                                  const int64_t* __restrict__ leaf_start, int64_t n_leaves,
                                                  int32_t* __restrict__ word_base, int32_t* __restrict__ ijk_out,
                                                  int32_t* __restrict__ jidx_out) {
    const int lane = threadIdx.x & 63;
    const int64_t l = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if (l >= n_leaves) return;
    int32_t b, ox, oy, oz;
    decode_leaf_key(keys[l], b, ox, oy, oz);
    int64_t base = leaf_start[l];
    for (int w = 0; w < 8; ++w) {
        const uint64_t m = masks[l * 8 + w];
        if (lane == 0) word_base[l * 8 + w] = (int32_t)base;
        if ((m >> lane) & 1ull) {
            const int64_t v = base + __popcll(m & ((1ull << lane) - 1ull));
            const int bit = w * 64 + lane;
            ijk_out[3 * v + 0] = ox + (bit >> 6);
            ijk_out[3 * v + 1] = oy + ((bit >> 3) & 7);
            ijk_out[3 * v + 2] = oz + (bit & 7);
            if (jidx_out) jidx_out[v] = b;
        }
        base += __popcll(m);
    }
}

// Voxels per grid: leaves are sorted by grid (the grid index is the key's top field),
// so grid b owns the leaves between the first keys >= b<<51 and >= (b+1)<<51.
__global__ void k_grid_counts(const uint64_t* __restrict__ keys, const int64_t* __restrict__ leaf_start,
                              int64_t n_leaves, int64_t n_voxels, int32_t n_grids, int32_t* __restrict__ counts) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= n_grids) return;
    auto lower = [&](uint64_t key) {
        int64_t lo = 0, hi = n_leaves;
        while (lo < hi) {
            const int64_t mid = (lo + hi) >> 1;
            if (keys[mid] < key) lo = mid + 1;
            else hi = mid;
        }
        return lo;
    };
    const int64_t l0 = lower((uint64_t)b << 51), l1 = lower((uint64_t)(b + 1) << 51);
    const int64_t v0 = l0 < n_leaves ? leaf_start[l0] : n_voxels;
    const int64_t v1 = l1 < n_leaves ? leaf_start[l1] : n_voxels;
    counts[b] = (int32_t)(v1 - v0);
}

// Voxel offsets of the grids from the inclusive scan of the leaf counts:
// offsets[b] = voxels in leaves of grids < b, b = 0..n_grids (offsets[n_grids] = all).
__global__ void k_grid_offsets(const uint64_t* __restrict__ keys, const int64_t* __restrict__ leaf_incl,
                               int64_t n_leaves, int32_t n_grids, int64_t* __restrict__ offsets) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b > n_grids) return;
    const uint64_t key = (uint64_t)b << 51;
    int64_t lo = 0, hi = n_leaves;
    while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if (keys[mid] < key) lo = mid + 1;
        else hi = mid;
    }
    offsets[b] = lo > 0 ? leaf_incl[lo - 1] : 0;
}

// ---------------------------------------------------------------------------------
// queries
// ---------------------------------------------------------------------------------

__device__ inline int32_t voxel_index(const fvdbk_grid& g, int32_t b, int32_t x, int32_t y, int32_t z) {
    if (x < kCoordMin || x > kCoordMax || y < kCoordMin || y > kCoordMax || z < kCoordMin || z > kCoordMax)
        return -1;
    const int32_t leaf = hash_find(g.hash, g.hash_cap, leaf_key(b, x, y, z));
    if (leaf < 0) return -1;
    const int bit = ((x & 7) << 6) | ((y & 7) << 3) | (z & 7);
    const uint64_t m = g.leaf_masks[(int64_t)leaf * 8 + (bit >> 6)];
    if (!((m >> (bit & 63)) & 1ull)) return -1;
    return g.leaf_word_base[(int64_t)leaf * 8 + (bit >> 6)] + __popcll(m & ((1ull << (bit & 63)) - 1ull));
}

__global__ void k_ijk_to_index(fvdbk_grid g, const int32_t* __restrict__ ijk, const int32_t* __restrict__ bidx,
                               int64_t n, int32_t* __restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int32_t b = bidx ? bidx[i] : 0;
    if (b < 0 || b >= g.n_grids || g.n_leaves == 0) {
        out[i] = -1;
        return;
    }
    out[i] = voxel_index(g, b, ijk[3 * i], ijk[3 * i + 1], ijk[3 * i + 2]);
}

// 64-bit form: the index within the query's own grid (minus offsets[b]) when offsets is given.
__global__ void k_ijk_to_index64(fvdbk_grid g, const int32_t* __restrict__ ijk, const int32_t* __restrict__ bidx,
                                 int64_t n, const int64_t* __restrict__ offsets, int64_t* __restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int32_t b = bidx ? bidx[i] : 0;
    if (b < 0 || b >= g.n_grids || g.n_leaves == 0) {
        out[i] = -1;
        return;
    }
    const int32_t v = voxel_index(g, b, ijk[3 * i], ijk[3 * i + 1], ijk[3 * i + 2]);
    out[i] = v < 0 ? (int64_t)-1 : (int64_t)v - (offsets ? offsets[b] : 0);
}

template <typename T>
__global__ void k_points_to_ijk(const T* __restrict__ pts, const int32_t* __restrict__ bidx, int64_t n,
                                const double* __restrict__ xf, int32_t* __restrict__ ijk) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int32_t b = bidx ? bidx[i] : 0;
    const double* t = xf + 6 * b;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const T v = pts[3 * i + d] * (T)t[d] + (T)t[3 + d];
        ijk[3 * i + d] = (int32_t)floor(v + (T)0.5);
    }
}

// ---------------------------------------------------------------------------------
// conv target-grid candidates
// ---------------------------------------------------------------------------------

struct CandGeom {
    int coarse;
    int lo[3], hi[3];  // offset range per kernel dim (d,h,w) as in the reference loops
    int ks[3];
    int st[3];         // (d,h,w)
    int kvol;
};

__global__ void k_conv_candidates(const int32_t* __restrict__ ijk, const int32_t* __restrict__ jidx, int64_t n,
                                  CandGeom g, int32_t* __restrict__ cijk, int32_t* __restrict__ cjidx) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int32_t b = jidx ? jidx[i] : 0;
    const int32_t x = ijk[3 * i], y = ijk[3 * i + 1], z = ijk[3 * i + 2];
    if (g.coarse) {
        // floor(ijk / branching factor), factor indexed by coordinate axis
        cijk[3 * i + 0] = floordiv(x, g.st[0]);
        cijk[3 * i + 1] = floordiv(y, g.st[1]);
        cijk[3 * i + 2] = floordiv(z, g.st[2]);
        cjidx[i] = b;
        return;
    }
    int64_t c = i * g.kvol;
    for (int di = g.lo[0]; di <= g.hi[0]; ++di)
        for (int dj = g.lo[1]; dj <= g.hi[1]; ++dj)
            for (int dk = g.lo[2]; dk <= g.hi[2]; ++dk, ++c) {
                const int32_t X = x + dk, Y = y + dj, Z = z + di;
                if (X % g.st[2] != 0 || Y % g.st[1] != 0 || Z % g.st[0] != 0) {
                    cjidx[c] = -1;
                    cijk[3 * c] = cijk[3 * c + 1] = cijk[3 * c + 2] = 0;
                    continue;
                }
                cijk[3 * c + 0] = X / g.st[2];
                cijk[3 * c + 1] = Y / g.st[1];
                cijk[3 * c + 2] = Z / g.st[0];
                cjidx[c] = b;
            }
}

// ---------------------------------------------------------------------------------
// kernel map
// ---------------------------------------------------------------------------------

// Per coordinate axis a (x, y, z): the probe coordinate of row coordinate c for
// kernel tap t in [0, ext[a]) is q = (c*mul[a] + off0[a] + sgn*t) / div[a], valid
// when the division is exact.
struct ProbeGeom {
    int K;
    int ext[3];   // taps per coordinate axis (x: ks[2], y: ks[1], z: ks[0])
    int mul[3];
    int div[3];
    int shift[3]; // log2(div) when div is a power of two, else -1
    int off0[3];
    int sgn;
    int n_blocks; // workgroups holding a row leaf (the launch may be padded to a multiple of 8)
    int tail_len; // map entries past the last row (padding rows) that the launch's last
                  // workgroup sets to -1 (it is always a padding workgroup), or 0
    int64_t tail_off;
};

// Workgroup -> work-item order that gives each of the 8 XCDs (which take workgroups
// round-robin) one contiguous range of the launch, so spatially adjacent tiles share an
// L2.  The grid is a multiple of 8 workgroups; the map is a bijection of [0, grid), and
// items >= n (returned as -1) are padding -- the last workgroup always is one when the
// grid holds more than n workgroups.
__device__ inline int64_t xcd_swizzle(int64_t bid, int64_t n) {
    const int64_t per = (int64_t)gridDim.x / 8;
    const int64_t t = (bid % 8) * per + bid / 8;
    return t < n ? t : -1;
}

// Row leaves per workgroup (one wavefront each).  One: leaves differ widely in size
// (1..512 voxels), and a workgroup holds its slots until its largest leaf is done.
constexpr int KMAP_WAVES = 1;
constexpr int KMAP_STAGE_BYTES = 16384; // LDS per wavefront for the staged neighbourhood
constexpr int KMAP_STAGE_PER_LEAF = 128; // staged bytes per probe leaf: 16 (mask word, word base) pairs
constexpr int KMAP_ROW_BYTES = 2048;     // the row leaf's voxels: 512 packed local coordinates

// One wavefront per row-grid leaf.  The probe-grid leaves
// that the leaf's taps can reach form a small box (nb_cap leaves at most, fixed per
// launch); each lane stages one of them (hash lookup, then its 8 mask words and word
// bases) into the wavefront's LDS slice.  The leaf's active voxels are listed with a
// ballot compaction of its mask (bit position per voxel rank, no ijk reads), and the
// leaf's block of the map (cnt*K consecutive int32) is written by consecutive lanes.
// With nb_cap == 0 (box too large to stage) every pair does a hash probe instead.
// UNIT: every axis has mul = div = 1 (forward maps of stride-1 convolutions), so a probe
// coordinate is the row coordinate plus the tap offset and every division is exact.
template <bool UNIT>
__global__ __launch_bounds__(64 * KMAP_WAVES) void k_kmap(fvdbk_grid rows, fvdbk_grid probe, ProbeGeom g, int nb_cap,
                                              int32_t* __restrict__ map) {
    extern __shared__ __attribute__((aligned(16))) unsigned char kmap_smem[];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int64_t blk = xcd_swizzle(blockIdx.x, g.n_blocks);
    if (blk < 0 && g.tail_len > 0 && blockIdx.x == gridDim.x - 1) {
        for (int i = threadIdx.x; i < g.tail_len; i += 64 * KMAP_WAVES) map[g.tail_off + i] = -1;
        return;
    }
    // the row leaf is wavefront-uniform: its key, masks and box live in scalar registers
    const int64_t L = blk < 0 ? rows.n_leaves : blk * KMAP_WAVES + __builtin_amdgcn_readfirstlane(wave);
    // LDS slice of this wavefront: per staged leaf, 16 pairs (32-bit mask word, index of
    // the word's first active voxel) -- 32-bit words keep the lookup's shifts and popcounts
    // 32-bit and one 8-byte read fetches both -- then the row leaf's voxels in rank order
    // as packed local coordinates x | y << 8 | z << 16
    unsigned char* base_ptr = kmap_smem + (size_t)wave * (nb_cap * KMAP_STAGE_PER_LEAF + KMAP_ROW_BYTES);
    uint2* s_mb = reinterpret_cast<uint2*>(base_ptr);
    uint32_t* s_vox = reinterpret_cast<uint32_t*>(base_ptr + nb_cap * KMAP_STAGE_PER_LEAF);
    // per-tap offsets (x, y, z), shared by the workgroup
    int32_t* s_off = reinterpret_cast<int32_t*>(kmap_smem +
                                                (size_t)KMAP_WAVES * (nb_cap * KMAP_STAGE_PER_LEAF + KMAP_ROW_BYTES));
    // tap k -> (tx, ty, tz), z fastest; the quotients by the small extents come from float
    // reciprocals: k < 4096, so (k + 0.5) / e is at least 0.5 / e from an integer, far more
    // than the rounding error of the product, and the truncation is exact
    const float r2 = 1.0f / (float)g.ext[2], r12 = 1.0f / (float)(g.ext[1] * g.ext[2]);
    for (int k = threadIdx.x; k < g.K; k += 64 * KMAP_WAVES) {
        const int tx = (int)(((float)k + 0.5f) * r12);
        const int rem = k - tx * g.ext[1] * g.ext[2];
        const int ty = (int)(((float)rem + 0.5f) * r2);
        const int tz = rem - ty * g.ext[2];
        s_off[3 * k + 0] = g.off0[0] + g.sgn * tx;
        s_off[3 * k + 1] = g.off0[1] + g.sgn * ty;
        s_off[3 * k + 2] = g.off0[2] + g.sgn * tz;
    }

    const bool live = L < rows.n_leaves;
    int32_t b = 0, o[3] = {0, 0, 0};
    int lo[3] = {0, 0, 0}, ext[3] = {1, 1, 1};
    int64_t vbase = 0;
    int cnt = 0;
    bool staged = false;
    if (live) {
        decode_leaf_key(rows.leaf_keys[L], b, o[0], o[1], o[2]);
        vbase = rows.leaf_word_base[L * 8];
        // ballot compaction of the leaf's active bits into rank order
        int rank = 0;
        for (int w = 0; w < 8; ++w) {
            const uint64_t m = rows.leaf_masks[L * 8 + w];
            if ((m >> lane) & 1ull)
                s_vox[rank + __popcll(m & ((1ull << lane) - 1ull))] =
                    (uint32_t)w | ((uint32_t)(lane >> 3) << 8) | ((uint32_t)(lane & 7) << 16);
            rank += __popcll(m);
        }
        cnt = rank;
        for (int a = 0; a < 3; ++a) {
            const int span = g.sgn * (g.ext[a] - 1);
            const int off_lo = g.off0[a] + (span < 0 ? span : 0);
            const int off_hi = g.off0[a] + (span > 0 ? span : 0);
            const int q_lo = UNIT ? o[a] + off_lo : floordiv(o[a] * g.mul[a] + off_lo, g.div[a]);
            const int q_hi = UNIT ? o[a] + 7 + off_hi : floordiv((o[a] + 7) * g.mul[a] + off_hi, g.div[a]);
            lo[a] = q_lo >> 3;
            ext[a] = (q_hi >> 3) - lo[a] + 1;
        }
        const int nn = ext[0] * ext[1] * ext[2];
        staged = nn <= nb_cap;
        if (staged) {
            // box leaf n -> (dx, dy, dz), z fastest (n < nb_cap: exact float quotients as above)
            const float e2r = 1.0f / (float)ext[2], e12r = 1.0f / (float)(ext[1] * ext[2]);
            for (int n = lane; n < nn; n += 64) {
                const int dx = (int)(((float)n + 0.5f) * e12r);
                const int nrem = n - dx * ext[1] * ext[2];
                const int dy = (int)(((float)nrem + 0.5f) * e2r);
                const int dz = nrem - dy * ext[2];
                const int32_t lx = (lo[0] + dx) * 8, ly = (lo[1] + dy) * 8, lz = (lo[2] + dz) * 8;
                int32_t leaf = -1;
                if (probe.n_leaves > 0 && lx >= kCoordMin && lx <= kCoordMax && ly >= kCoordMin && ly <= kCoordMax &&
                    lz >= kCoordMin && lz <= kCoordMax)
                    leaf = hash_find(probe.hash, probe.hash_cap, leaf_key(b, lx, ly, lz));
                uint64_t mw[8];
                int32_t bw[8];
#pragma unroll
                for (int w = 0; w < 8; ++w) {
                    mw[w] = leaf >= 0 ? probe.leaf_masks[(int64_t)leaf * 8 + w] : 0ull;
                    bw[w] = leaf >= 0 ? probe.leaf_word_base[(int64_t)leaf * 8 + w] : 0;
                }
#pragma unroll
                for (int w = 0; w < 8; ++w) {
                    const uint32_t lo32 = (uint32_t)mw[w], hi32 = (uint32_t)(mw[w] >> 32);
                    s_mb[n * 16 + 2 * w] = make_uint2(lo32, (uint32_t)bw[w]);
                    s_mb[n * 16 + 2 * w + 1] = make_uint2(hi32, (uint32_t)(bw[w] + __popc(lo32)));
                }
            }
        }
    }
    __syncthreads();
    if (!live) return;

    const int K = g.K;
    const int total = cnt * K;
    // the leaf's first voxel is wavefront-uniform: keep the block's base address scalar
    const int64_t vb = (int64_t)(((uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)(vbase >> 32)) << 32) |
                                 (uint32_t)__builtin_amdgcn_readfirstlane((int)vbase));
    int32_t* __restrict__ out = map + vb * K;
    // probe coordinate of the row voxel with packed local coordinates `v` for tap offsets
    // `of`; false when the division is inexact
    auto probe_q = [&](uint32_t v, const int (&of)[3], int32_t (&q)[3]) -> bool {
        const int c[3] = {o[0] + (int)(v & 255u), o[1] + (int)((v >> 8) & 255u), o[2] + (int)(v >> 16)};
        if constexpr (UNIT) {
#pragma unroll
            for (int a = 0; a < 3; ++a) q[a] = c[a] + of[a];
            return true;
        }
        bool ok = true;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const int num = c[a] * g.mul[a] + of[a];
            if (g.shift[a] >= 0) {
                if (num & ((1 << g.shift[a]) - 1)) ok = false;
                q[a] = num >> g.shift[a];
            } else {
                if (num % g.div[a] != 0) ok = false;
                q[a] = num / g.div[a];
            }
        }
        return ok;
    };
    // staged lookup: 32-bit mask word of the probe voxel, then its index or -1
    auto staged_index = [&](bool ok, const int32_t (&q)[3], uint32_t m, int32_t wb, int sh) -> int32_t {
        (void)q;
        return (ok && ((m >> sh) & 1u)) ? wb + __popc(m & ((1u << sh) - 1u)) : -1;
    };
    auto staged_word = [&](const int32_t (&q)[3], int& sh) -> int {
        const int n = (((q[0] >> 3) - lo[0]) * ext[1] + ((q[1] >> 3) - lo[1])) * ext[2] + ((q[2] >> 3) - lo[2]);
        const int qb = ((q[0] & 7) << 6) | ((q[1] & 7) << 3) | (q[2] & 7);
        sh = qb & 31;
        return n * 16 + (qb >> 5);
    };
    if (staged && K <= 64) {
        // each lane keeps one tap; a pass covers U groups of 64/K consecutive rows with
        // every LDS read unconditional (row index clamped), so the reads of the U groups
        // overlap
        constexpr int U = 4;
        const int rpi = 64 / K, lk = lane % K, lr = lane / K;
        const bool act = lane < rpi * K;
        const int of[3] = {s_off[3 * lk + 0], s_off[3 * lk + 1], s_off[3 * lk + 2]};
        const int last_r = cnt - 1;
        if constexpr (UNIT) {
            // Box coordinates P = local + (o - 8 lo + offset) per axis, packed one byte each
            // (P < 8 * ext <= 255): one add per row, then the box leaf n and the bit inside
            // it each come out of one 4 x u8 dot product.
            const uint32_t add = (uint32_t)(o[0] - 8 * lo[0] + of[0]) | ((uint32_t)(o[1] - 8 * lo[1] + of[1]) << 8) |
                                 ((uint32_t)(o[2] - 8 * lo[2] + of[2]) << 16);
            const uint32_t leaf_w = (uint32_t)(ext[1] * ext[2]) | ((uint32_t)ext[2] << 8) | (1u << 16);
            // stores through a buffer resource over the leaf's block: rows past the leaf and
            // idle lanes take an offset beyond it, which the hardware drops, so no branch
            const __amdgpu_buffer_rsrc_t ob =
                __builtin_amdgcn_make_buffer_rsrc((void*)out, (short)0, cnt * K * 4, 0x00020000);
            const uint32_t lane_off = act ? (uint32_t)lk * 4u : 0x80000000u;
            for (int r0 = lr; r0 < cnt; r0 += rpi * U) {
                uint32_t P[U];
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    const int rr = r0 + rpi * u;
                    P[u] = s_vox[rr < last_r ? rr : last_r] + add;
                }
                uint2 mb[U];
                int sh[U];
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    const uint32_t n = __builtin_amdgcn_udot4((P[u] >> 3) & 0x1F1F1Fu, leaf_w, 0u, false);
                    const uint32_t qb = __builtin_amdgcn_udot4(P[u] & 0x070707u, 0x00010840u, 0u, false);
                    sh[u] = (int)(qb & 31u);
                    mb[u] = s_mb[(n << 4) | (qb >> 5)];
                }
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    const int rr = r0 + rpi * u;
                    const uint32_t m = mb[u].x;
                    // index if the probe bit is set, else all ones (-1), without a branch
                    const uint32_t idx = (mb[u].y + __popc(m & ((1u << sh[u]) - 1u))) | (((m >> sh[u]) & 1u) - 1u);
                    __builtin_amdgcn_raw_buffer_store_b32(idx, ob, lane_off + (uint32_t)(rr * K * 4), 0, 0);
                }
            }
            return;
        }
        // branch-free stores, as in the stride-1 loop above
        const __amdgpu_buffer_rsrc_t ob = __builtin_amdgcn_make_buffer_rsrc((void*)out, (short)0, cnt * K * 4, 0x00020000);
        const uint32_t lane_off = act ? (uint32_t)lk * 4u : 0x80000000u;
        for (int r0 = lr; r0 < cnt; r0 += rpi * U) {
            uint32_t vox[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int rr = r0 + rpi * u;
                vox[u] = s_vox[rr < last_r ? rr : last_r];
            }
            int32_t q[U][3];
            int word[U], sh[U];
            bool ok[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                ok[u] = probe_q(vox[u], of, q[u]);
                word[u] = ok[u] ? staged_word(q[u], sh[u]) : 0;
                if (!ok[u]) sh[u] = 0;
            }
            uint2 mb[U];
#pragma unroll
            for (int u = 0; u < U; ++u) mb[u] = s_mb[word[u]];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int rr = r0 + rpi * u;
                const int32_t idx = staged_index(ok[u], q[u], mb[u].x, (int32_t)mb[u].y, sh[u]);
                __builtin_amdgcn_raw_buffer_store_b32((uint32_t)idx, ob, lane_off + (uint32_t)(rr * K * 4), 0, 0);
            }
        }
        return;
    }
    const int qstep = 64 / K, rstep = 64 % K;
    int r = lane / K, k = lane - (lane / K) * K;
    for (int idx = lane; idx < total; idx += 64) {
        const uint32_t vox = s_vox[r];
        const int of[3] = {s_off[3 * k + 0], s_off[3 * k + 1], s_off[3 * k + 2]};
        int32_t q[3];
        const bool ok = probe_q(vox, of, q);
        int32_t res = -1;
        if (staged) {
            int sh = 0;
            const int word = ok ? staged_word(q, sh) : 0;
            const uint2 mb = s_mb[word];
            res = staged_index(ok, q, mb.x, (int32_t)mb.y, sh);
        } else if (ok && probe.n_leaves > 0) {
            res = voxel_index(probe, b, q[0], q[1], q[2]);
        }
        out[idx] = res;
        r += qstep;
        k += rstep;
        if (k >= K) {
            k -= K;
            ++r;
        }
    }
}

// Upper bound on the probe leaves one row leaf can reach (per launch).
static int kmap_box_bound(const ProbeGeom& g) {
    int nb = 1;
    for (int a = 0; a < 3; ++a) {
        // probe coordinates span at most (7*mul + ext - 1)/div + 1 values
        const int span_q = (7 * g.mul[a] + g.ext[a] - 1) / g.div[a] + 1;
        nb *= (span_q + 7) / 8 + 1;
    }
    return nb;
}

// Maps of 2x2x2 kernels with stride 2 (K = 8, tap k = 4 dx + 2 dy + dz, dz fastest), one
// wavefront per row leaf as in k_kmap, with the geometry fixed so each row is a few LDS reads:
//   DOWN (rows: coarse voxels c, probe: fine grid): tap (dx, dy, dz) -> fine voxel 2c + d.  The
//        row leaf's children fill the 2x2x2 fine leaves at 2 * origin + 8 * (0|1), staged in LDS;
//        a row's 8 children lie in one of them, the four (dy, dz) of each dx in one 32-bit half
//        of a mask word.
//   UP (rows: fine voxels v, probe: coarse grid; the transposed map): the one valid tap of v is
//        its position d = v & 1 inside its parent v >> 1, and every parent of the row leaf lies
//        in the single coarse leaf at (origin >> 1) rounded down to a multiple of 8.
// Entries are bit-identical to k_kmap's.  Each row's 8 entries are written as two 16-byte stores.
// A workgroup takes KMAP_S2_WAVES row leaves, one per wavefront (U-Net topology of 3.2M voxels:
// 123 us per forward for all six 2x2x2 maps with one leaf per workgroup, 129 us with four;
// k_kmap took 253 us for them).
constexpr int KMAP_S2_WAVES = 1;  // row leaves per workgroup of k_kmap_s2, one per wavefront

template <bool UP>
__global__ __launch_bounds__(64 * KMAP_S2_WAVES) void k_kmap_s2(fvdbk_grid rows, fvdbk_grid probe, int n_groups,
                                                                int tail_len, int32_t* __restrict__ map) {
    __shared__ uint32_t s_vox_w[KMAP_S2_WAVES][512];
    __shared__ uint2 s_mb_w[KMAP_S2_WAVES][8 * 16];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int64_t grp = xcd_swizzle(blockIdx.x, n_groups);
    if (grp < 0) {
        // the launch's last workgroup (always a padding one when tail_len > 0) sets the map
        // entries of the padding rows past the last voxel to -1
        if (tail_len > 0 && blockIdx.x == gridDim.x - 1)
            for (int i = threadIdx.x; i < tail_len; i += 64 * KMAP_S2_WAVES) map[rows.n_voxels * 8 + i] = -1;
        return;
    }
    const int64_t L = grp * KMAP_S2_WAVES + __builtin_amdgcn_readfirstlane(wave);
    const bool live = L < rows.n_leaves;
    uint32_t* s_vox = s_vox_w[wave];
    uint2* s_mb = s_mb_w[wave];
    int32_t b = 0, o[3] = {0, 0, 0};
    int64_t vbase = 0;
    int cnt = 0;
    if (live) {
        decode_leaf_key(rows.leaf_keys[L], b, o[0], o[1], o[2]);
        vbase = rows.leaf_word_base[L * 8];
        int rank = 0;
        for (int w = 0; w < 8; ++w) {
            const uint64_t m = rows.leaf_masks[L * 8 + w];
            if ((m >> lane) & 1ull)
                s_vox[rank + __popcll(m & ((1ull << lane) - 1ull))] =
                    (uint32_t)w | ((uint32_t)(lane >> 3) << 8) | ((uint32_t)(lane & 7) << 16);
            rank += __popcll(m);
        }
        cnt = rank;
        // stage the probe leaves: 8 fine leaves (DOWN) or the one coarse leaf (UP)
        constexpr int NS = UP ? 1 : 8;
        if (lane < NS) {
            int32_t q[3];
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                const int d = UP ? 0 : (lane >> (2 - a)) & 1;
                q[a] = UP ? ((o[a] >> 1) & ~7) : 2 * o[a] + 8 * d;
            }
            int32_t leaf = -1;
            if (probe.n_leaves > 0 && q[0] >= kCoordMin && q[0] <= kCoordMax && q[1] >= kCoordMin &&
                q[1] <= kCoordMax && q[2] >= kCoordMin && q[2] <= kCoordMax)
                leaf = hash_find(probe.hash, probe.hash_cap, leaf_key(b, q[0], q[1], q[2]));
#pragma unroll
            for (int w = 0; w < 8; ++w) {
                const uint64_t m = leaf >= 0 ? probe.leaf_masks[(int64_t)leaf * 8 + w] : 0ull;
                const uint32_t bw = leaf >= 0 ? (uint32_t)probe.leaf_word_base[(int64_t)leaf * 8 + w] : 0u;
                const uint32_t lo32 = (uint32_t)m, hi32 = (uint32_t)(m >> 32);
                s_mb[lane * 16 + 2 * w] = make_uint2(lo32, bw);
                s_mb[lane * 16 + 2 * w + 1] = make_uint2(hi32, bw + __popc(lo32));
            }
        }
    }
    __syncthreads();
    if (!live) return;
    int4* out = reinterpret_cast<int4*>(map + vbase * 8);
    // index of bit `sh` of a staged 32-bit half-word, or -1
    auto idx = [](uint2 hw, int sh) -> int32_t {
        return ((hw.x >> sh) & 1u) ? (int32_t)(hw.y + __popc(hw.x & ((1u << sh) - 1u))) : -1;
    };
    for (int r = lane; r < cnt; r += 64) {
        const uint32_t v = s_vox[r];
        const int lx = (int)(v & 255u), ly = (int)((v >> 8) & 255u), lz = (int)(v >> 16);
        int32_t e[8];
        if constexpr (!UP) {
            // children 2 (o + l) + d: fine leaf (l >> 2) of the box, local 2 (l & 3) + d
            const int n = ((lx >> 2) << 2) | ((ly >> 2) << 1) | (lz >> 2);
            const int fx = 2 * (lx & 3), fy = 2 * (ly & 3), fz = 2 * (lz & 3);
#pragma unroll
            for (int dx = 0; dx < 2; ++dx) {
                // bits (fy + dy) * 8 + fz + dz of word fx + dx, all in its half fy >> 2
                const uint2 hw = s_mb[n * 16 + 2 * (fx + dx) + (fy >> 2)];
                const int sh0 = ((fy & 3) << 3) + fz;
#pragma unroll
                for (int dy = 0; dy < 2; ++dy)
#pragma unroll
                    for (int dz = 0; dz < 2; ++dz) e[dx * 4 + dy * 2 + dz] = idx(hw, sh0 + dy * 8 + dz);
            }
        } else {
            // parent local coordinates in the staged coarse leaf
            const int px = ((o[0] >> 1) & 7) + (lx >> 1), py = ((o[1] >> 1) & 7) + (ly >> 1),
                      pz = ((o[2] >> 1) & 7) + (lz >> 1);
            const int bit = px * 64 + py * 8 + pz;
            const int32_t p = idx(s_mb[bit >> 5], bit & 31);
            const int k = (lx & 1) * 4 + (ly & 1) * 2 + (lz & 1);
#pragma unroll
            for (int t = 0; t < 8; ++t) e[t] = t == k ? p : -1;
        }
        out[2 * r] = make_int4(e[0], e[1], e[2], e[3]);
        out[2 * r + 1] = make_int4(e[4], e[5], e[6], e[7]);
    }
}

__global__ void k_fill_i32(int32_t* __restrict__ p, int64_t n, int32_t v) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// ---------------------------------------------------------------------------------
// gather-scatter neighbour map: ordered compaction with wavefront ballots
// ---------------------------------------------------------------------------------

constexpr int NB_THREADS = 256;

__global__ __launch_bounds__(NB_THREADS) void k_nbmap_counts(const int32_t* __restrict__ map, int64_t n_rows,
                                                             int K, int32_t* __restrict__ counts) {
    const int64_t r0 = (int64_t)blockIdx.x * FVDBK_NBMAP_ROWS;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    __shared__ int32_t s_cnt[4];
    for (int k = 0; k < K; ++k) {
        int c = 0;
        for (int it = 0; it < FVDBK_NBMAP_ROWS / NB_THREADS; ++it) {
            const int64_t r = r0 + it * NB_THREADS + threadIdx.x;
            const bool v = r < n_rows && map[r * K + k] >= 0;
            c += __popcll(__ballot(v));
        }
        if (lane == 0) s_cnt[wave] = c;
        __syncthreads();
        if (threadIdx.x == 0) counts[(int64_t)blockIdx.x * K + k] = s_cnt[0] + s_cnt[1] + s_cnt[2] + s_cnt[3];
        __syncthreads();
    }
}

__global__ __launch_bounds__(NB_THREADS) void k_nbmap_emit(const int32_t* __restrict__ map, int64_t n_rows, int K,
                                                           const int64_t* __restrict__ offsets,
                                                           int32_t* __restrict__ pairs) {
    const int64_t r0 = (int64_t)blockIdx.x * FVDBK_NBMAP_ROWS;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    __shared__ int32_t s_cnt[4];
    for (int k = 0; k < K; ++k) {
        int64_t pos = offsets[(int64_t)blockIdx.x * K + k];
        for (int it = 0; it < FVDBK_NBMAP_ROWS / NB_THREADS; ++it) {
            const int64_t r = r0 + it * NB_THREADS + threadIdx.x;
            const int32_t s = r < n_rows ? map[r * K + k] : -1;
            const uint64_t bal = __ballot(s >= 0);
            if (lane == 0) s_cnt[wave] = __popcll(bal);
            __syncthreads();
            int before = 0;
            for (int w = 0; w < wave; ++w) before += s_cnt[w];
            if (s >= 0) {
                const int64_t p = pos + before + __popcll(bal & ((1ull << lane) - 1ull));
                pairs[2 * p] = s;
                pairs[2 * p + 1] = (int32_t)r;
            }
            pos += s_cnt[0] + s_cnt[1] + s_cnt[2] + s_cnt[3];
            __syncthreads();
        }
    }
}


// ---------------------------------------------------------------------------
// Row order by tap mask within windows of consecutive rows.  One workgroup owns
// ORDER_THREADS * ITEMS consecutive rows (8192, or 2048 for windows of up to 2048 rows, so
// that small grids still spread over many workgroups): it builds each row's key (window index within the
// block above the inverted tap mask of the first min(K, 27) taps, so windows never mix
// and masks sort descending), runs a block-wide radix sort over just those key bits
// (stable: ties keep row order), and writes the rows in sorted order.  Rows past
// n_rows take the largest key and, being last in the block, stay behind every real row.
// ---------------------------------------------------------------------------
constexpr int ORDER_THREADS = 1024;
constexpr int ORDER_BLOCK = ORDER_THREADS * 8;  // largest block (and window)
constexpr int ORDER_SMALL_BLOCK = ORDER_THREADS * 2;
constexpr int ORDER_MASK_BITS = 27;

// KT > 0: compile-time kernel volume (the tap loads of a row are unrolled and issued
// together); KT = 0: runtime K.  ORDER_ITEMS rows per lane.
template <int KT, int ORDER_ITEMS>
__global__ __launch_bounds__(ORDER_THREADS) void k_row_order_window(const int32_t* __restrict__ map, int64_t n_rows,
                                                                    int K_rt, int log2_window,
                                                                    int32_t* __restrict__ order) {
    constexpr int ORDER_BLOCK = ORDER_THREADS * ORDER_ITEMS;
    constexpr int ORDER_LOG2_BLOCK = ORDER_ITEMS == 8 ? 13 : (ORDER_ITEMS == 2 ? 11 : -1);
    static_assert(ORDER_LOG2_BLOCK > 0 && ORDER_BLOCK == 1 << ORDER_LOG2_BLOCK, "block size");
    using Sort = hipcub::BlockRadixSort<uint32_t, ORDER_THREADS, ORDER_ITEMS, int32_t>;
    __shared__ union {
        typename Sort::TempStorage sort;
        uint32_t keys[ORDER_BLOCK];
    } sm;
    const int K = KT > 0 ? KT : K_rt;
    const int kb = K < ORDER_MASK_BITS ? K : ORDER_MASK_BITS;
    const int sub_bits = ORDER_LOG2_BLOCK - log2_window;
    const uint32_t full = (1u << kb) - 1u;
    const int64_t base = (int64_t)blockIdx.x * ORDER_BLOCK;
    // keys, one row per lane (neighbouring lanes read neighbouring map rows)
#pragma unroll
    for (int it = 0; it < ORDER_ITEMS; ++it) {
        const int local = threadIdx.x + it * ORDER_THREADS;
        const int64_t r = base + local;
        uint32_t key = (sub_bits + kb == 32) ? 0xFFFFFFFFu : ((1u << (sub_bits + kb)) - 1u);
        if (r < n_rows) {
            const int32_t* m = map + r * K;
            uint32_t mask = 0u;
            if constexpr (KT > 0) {
#pragma unroll
                for (int k = 0; k < (KT < ORDER_MASK_BITS ? KT : ORDER_MASK_BITS); ++k)
                    mask |= (m[k] >= 0 ? 1u : 0u) << k;
            } else {
                for (int k = 0; k < kb; ++k) mask |= (m[k] >= 0 ? 1u : 0u) << k;
            }
            key = ((uint32_t)(local >> log2_window) << kb) | (full - mask);
        }
        sm.keys[local] = key;
    }
    __syncthreads();
    // blocked arrangement: lane t holds rows t*ITEMS .. t*ITEMS+ITEMS-1
    uint32_t key[ORDER_ITEMS];
    int32_t row[ORDER_ITEMS];
#pragma unroll
    for (int i = 0; i < ORDER_ITEMS; ++i) {
        const int local = threadIdx.x * ORDER_ITEMS + i;
        key[i] = sm.keys[local];
        row[i] = (int32_t)(base + local);
    }
    __syncthreads();  // the key array and the sort's storage share memory
    Sort(sm.sort).Sort(key, row, 0, sub_bits + kb);
#pragma unroll
    for (int i = 0; i < ORDER_ITEMS; ++i) {
        const int64_t pos = base + threadIdx.x * ORDER_ITEMS + i;
        if (pos < n_rows) order[pos] = row[i];
    }
}

}  // namespace fvdbk

using namespace fvdbk;

extern "C" {

const char* fvdbk_last_error(void) { return g_last_error.c_str(); }

const char* fvdbk_version(void) { return "fvdbk 0.1 gfx950"; }

int fvdbk_grid_leaf_keys(const int32_t* ijk, const int32_t* batch_idx, int64_t n, uint64_t* keys, int32_t* bad_count,
                         hipStream_t stream) {
    FVDBK_REQUIRE(n >= 0, FVDBK_EINVAL, "negative voxel count");
    if (n == 0) return FVDBK_OK;
    FVDBK_REQUIRE(ijk && keys && bad_count, FVDBK_EINVAL, "null pointer");
    hipLaunchKernelGGL(k_leaf_keys, dim3(grid_1d(n, 256)), dim3(256), 0, stream, ijk, batch_idx, n, keys, bad_count);
    FVDBK_LAUNCH_CHECK("k_leaf_keys");
    return FVDBK_OK;
}

int64_t fvdbk_grid_dedupe_capacity(int64_t n) {
    int64_t cap = 1024;
    while (cap < 2 * n + 2) cap <<= 1;
    return cap;
}

int fvdbk_grid_leaf_dedupe(const int32_t* ijk, const int32_t* batch_idx, int64_t n, int32_t n_grids,
                           uint64_t* set, int64_t set_cap, uint64_t* leaf_keys, uint32_t* counters,
                           hipStream_t stream) {
    FVDBK_REQUIRE(n >= 0 && n_grids >= 0, FVDBK_EINVAL, "grid_leaf_dedupe: negative size");
    FVDBK_REQUIRE(set_cap >= 2 * n + 2 && (set_cap & (set_cap - 1)) == 0, FVDBK_EINVAL,
                  "grid_leaf_dedupe: set capacity must be a power of two >= 2n+2");
    FVDBK_REQUIRE(set && counters && (n == 0 || (ijk && leaf_keys)), FVDBK_EINVAL, "grid_leaf_dedupe: null argument");
    hipLaunchKernelGGL(k_fill_u64, dim3(grid_1d(set_cap, 256)), dim3(256), 0, stream,
                       reinterpret_cast<unsigned long long*>(set), set_cap, (unsigned long long)kEmptyKey);
    FVDBK_LAUNCH_CHECK("k_fill_u64");
    if (n == 0) return FVDBK_OK;
    hipLaunchKernelGGL(k_leaf_dedupe, dim3(grid_1d(n, DEDUPE_THREADS)), dim3(DEDUPE_THREADS), 0, stream, ijk,
                       batch_idx, n, n_grids,
                       reinterpret_cast<unsigned long long*>(set), set_cap, leaf_keys,
                       reinterpret_cast<unsigned int*>(counters));
    FVDBK_LAUNCH_CHECK("k_leaf_dedupe");
    return FVDBK_OK;
}

int fvdbk_grid_offsets(const uint64_t* leaf_keys, const int64_t* leaf_incl, int64_t n_leaves, int32_t n_grids,
                       int64_t* offsets, hipStream_t stream) {
    FVDBK_REQUIRE(n_leaves >= 0 && n_grids >= 0 && n_grids <= kMaxGrids, FVDBK_EINVAL, "grid_offsets: bad size");
    FVDBK_REQUIRE(offsets && (n_leaves == 0 || (leaf_keys && leaf_incl)), FVDBK_EINVAL, "grid_offsets: null argument");
    hipLaunchKernelGGL(k_grid_offsets, dim3(grid_1d(n_grids + 1, 64)), dim3(64), 0, stream, leaf_keys, leaf_incl,
                       n_leaves, n_grids, offsets);
    FVDBK_LAUNCH_CHECK("k_grid_offsets");
    return FVDBK_OK;
}

// Leaf keys of a batch of n_grids grids use bits 0..50 (root tile, upper and lower
// node digits) and the grid index from bit 51.
static int leaf_key_end_bit(int32_t n_grids) {
    int gb = 0;
    while ((1 << gb) < n_grids) ++gb;
    return 51 + gb;
}

size_t fvdbk_grid_sort_leaf_keys_workspace(int64_t n, int32_t n_grids) {
    if (n <= 0 || n > INT32_MAX || n_grids < 0 || n_grids > kMaxGrids) return 0;
    size_t bytes = 0;
    if (hipcub::DeviceRadixSort::SortKeys(nullptr, bytes, static_cast<const uint64_t*>(nullptr),
                                          static_cast<uint64_t*>(nullptr), (int)n, 0, leaf_key_end_bit(n_grids),
                                          (hipStream_t)0) != hipSuccess)
        return 0;
    return bytes;
}

int fvdbk_grid_sort_leaf_keys(const uint64_t* keys_in, uint64_t* keys_out, int64_t n, int32_t n_grids,
                              void* workspace, size_t workspace_bytes, hipStream_t stream) {
    FVDBK_REQUIRE(n >= 0 && n <= INT32_MAX, FVDBK_EINVAL, "leaf count %lld out of range", (long long)n);
    FVDBK_REQUIRE(n_grids >= 0 && n_grids <= kMaxGrids, FVDBK_EINVAL, "a grid batch holds at most %d grids",
                  kMaxGrids);
    if (n == 0) return FVDBK_OK;
    FVDBK_REQUIRE(keys_in && keys_out && keys_in != keys_out, FVDBK_EINVAL, "sort needs distinct key buffers");
    size_t need = fvdbk_grid_sort_leaf_keys_workspace(n, n_grids);
    FVDBK_REQUIRE(need > 0 && workspace != nullptr && workspace_bytes >= need, FVDBK_EINVAL,
                  "sort workspace of %zu bytes is smaller than the %zu needed", workspace_bytes, need);
    const hipError_t e = hipcub::DeviceRadixSort::SortKeys(workspace, need, keys_in, keys_out, (int)n, 0,
                                                           leaf_key_end_bit(n_grids), stream);
    FVDBK_REQUIRE(e == hipSuccess, FVDBK_EHIP, "leaf key sort: %s", hipGetErrorString(e));
    return FVDBK_OK;
}

int fvdbk_grid_hash_build(const uint64_t* leaf_keys, int64_t n_leaves, fvdbk_hslot* hash, int64_t hash_cap,
                          hipStream_t stream) {
    FVDBK_REQUIRE(hash_cap > 0 && (hash_cap & (hash_cap - 1)) == 0, FVDBK_EINVAL, "hash capacity must be a power of two");
    FVDBK_REQUIRE(n_leaves >= 0 && n_leaves < hash_cap, FVDBK_EINVAL, "hash capacity %lld too small for %lld leaves",
                  (long long)hash_cap, (long long)n_leaves);
    hipLaunchKernelGGL(k_hash_clear, dim3(grid_1d(hash_cap, 256)), dim3(256), 0, stream, hash, hash_cap);
    FVDBK_LAUNCH_CHECK("k_hash_clear");
    if (n_leaves == 0) return FVDBK_OK;
    hipLaunchKernelGGL(k_hash_insert, dim3(grid_1d(n_leaves, 256)), dim3(256), 0, stream, leaf_keys, n_leaves, hash,
                       hash_cap);
    FVDBK_LAUNCH_CHECK("k_hash_insert");
    return FVDBK_OK;
}

int fvdbk_grid_leaf_masks(const int32_t* ijk, const int32_t* batch_idx, int64_t n, const fvdbk_hslot* hash,
                          int64_t hash_cap, uint64_t* leaf_masks, int64_t n_leaves, hipStream_t stream) {
    if (n_leaves == 0) return FVDBK_OK;
    hipLaunchKernelGGL(k_zero_u64, dim3(grid_1d(n_leaves * 8, 256)), dim3(256), 0, stream, leaf_masks, n_leaves * 8);
    FVDBK_LAUNCH_CHECK("k_zero_u64");
    if (n == 0) return FVDBK_OK;
    hipLaunchKernelGGL(k_leaf_masks, dim3(grid_1d(n, 256)), dim3(256), 0, stream, ijk, batch_idx, n, hash, hash_cap,
                       leaf_masks);
    FVDBK_LAUNCH_CHECK("k_leaf_masks");
    return FVDBK_OK;
}

static int coarsen_shifts(const int32_t factor[3], int (&sh)[3]) {
    for (int a = 0; a < 3; ++a) {
        const int32_t f = factor[a];
        FVDBK_REQUIRE(f > 0 && f <= (1 << 16) && (f & (f - 1)) == 0, FVDBK_EINVAL,
                      "coarsening factor %d is not a power of two in [1, 65536]", (int)f);
        sh[a] = __builtin_ctz((unsigned)f);
    }
    return FVDBK_OK;
}

int fvdbk_grid_coarsen_reps(const uint64_t* fine_keys, int64_t n_fine_leaves, const int32_t factor[3],
                            int32_t* rep_ijk, int32_t* rep_bidx, hipStream_t stream) {
    int sh[3];
    const int rc = coarsen_shifts(factor, sh);
    if (rc != FVDBK_OK) return rc;
    if (n_fine_leaves == 0) return FVDBK_OK;
    hipLaunchKernelGGL(k_coarse_reps, dim3(grid_1d(n_fine_leaves, 256)), dim3(256), 0, stream, fine_keys,
                       n_fine_leaves, sh[0], sh[1], sh[2], rep_ijk, rep_bidx);
    FVDBK_LAUNCH_CHECK("k_coarse_reps");
    return FVDBK_OK;
}

int fvdbk_grid_coarsen_masks(const uint64_t* fine_keys, const uint64_t* fine_masks, int64_t n_fine_leaves,
                             const int32_t factor[3], const fvdbk_hslot* hash, int64_t hash_cap, uint64_t* leaf_masks,
                             int64_t n_leaves, hipStream_t stream) {
    int sh[3];
    const int rc = coarsen_shifts(factor, sh);
    if (rc != FVDBK_OK) return rc;
    if (n_leaves == 0) return FVDBK_OK;
    hipLaunchKernelGGL(k_zero_u64, dim3(grid_1d(n_leaves * 8, 256)), dim3(256), 0, stream, leaf_masks, n_leaves * 8);
    FVDBK_LAUNCH_CHECK("k_zero_u64");
    if (n_fine_leaves == 0) return FVDBK_OK;
    hipLaunchKernelGGL(k_coarse_masks, dim3(grid_1d(n_fine_leaves * 8, 256)), dim3(256), 0, stream, fine_keys,
                       fine_masks, n_fine_leaves, sh[0], sh[1], sh[2], hash, hash_cap, leaf_masks);
    FVDBK_LAUNCH_CHECK("k_coarse_masks");
    return FVDBK_OK;
}

int fvdbk_grid_leaf_counts(const uint64_t* leaf_masks, int64_t n_leaves, int32_t* counts, hipStream_t stream) {
    if (n_leaves == 0) return FVDBK_OK;
    hipLaunchKernelGGL(k_leaf_counts, dim3(grid_1d(n_leaves, 256)), dim3(256), 0, stream, leaf_masks, n_leaves,
                       counts);
    FVDBK_LAUNCH_CHECK("k_leaf_counts");
    return FVDBK_OK;
}

using LeafCountIter = rocprim::transform_iterator<rocprim::counting_iterator<int64_t>, LeafPopcount, int64_t>;

static LeafCountIter leaf_count_iter(const uint64_t* masks) {
    return LeafCountIter(rocprim::counting_iterator<int64_t>(0), LeafPopcount{masks});
}

size_t fvdbk_grid_leaf_scan_workspace(int64_t n_leaves) {
    if (n_leaves <= 0 || n_leaves > INT32_MAX) return 0;
    size_t bytes = 0;
    if (rocprim::inclusive_scan(nullptr, bytes, leaf_count_iter(nullptr), static_cast<int64_t*>(nullptr),
                                (size_t)n_leaves, rocprim::plus<int64_t>(), (hipStream_t)0) != hipSuccess)
        return 0;
    return bytes;
}

int fvdbk_grid_leaf_scan(const uint64_t* leaf_masks, int64_t n_leaves, int64_t* leaf_start, int64_t* leaf_incl,
                         void* workspace, size_t workspace_bytes, hipStream_t stream) {
    FVDBK_REQUIRE(n_leaves >= 0 && n_leaves <= INT32_MAX, FVDBK_EINVAL, "leaf count %lld out of range",
                  (long long)n_leaves);
    if (n_leaves == 0) return FVDBK_OK;
    FVDBK_REQUIRE(leaf_masks && leaf_start && leaf_incl, FVDBK_EINVAL, "leaf scan: null argument");
    size_t need = fvdbk_grid_leaf_scan_workspace(n_leaves);
    FVDBK_REQUIRE(need > 0 && workspace != nullptr && workspace_bytes >= need, FVDBK_EINVAL,
                  "leaf scan workspace of %zu bytes is smaller than the %zu needed", workspace_bytes, need);
    const hipError_t e = rocprim::inclusive_scan(workspace, need, leaf_count_iter(leaf_masks), leaf_incl,
                                                 (size_t)n_leaves, rocprim::plus<int64_t>(), stream);
    FVDBK_REQUIRE(e == hipSuccess, FVDBK_EHIP, "leaf scan: %s", hipGetErrorString(e));
    hipLaunchKernelGGL(k_leaf_start, dim3(grid_1d(n_leaves, 256)), dim3(256), 0, stream, leaf_masks, leaf_incl,
                       n_leaves, leaf_start);
    FVDBK_LAUNCH_CHECK("k_leaf_start");
    return FVDBK_OK;
}

int fvdbk_grid_finalize(const uint64_t* leaf_keys, const uint64_t* leaf_masks, const int64_t* leaf_start,
                        int64_t n_leaves, int64_t n_voxels, int32_t* leaf_word_base, int32_t* ijk_out,
                        int32_t* jidx_out, int32_t* grid_counts, int32_t n_grids, hipStream_t stream) {
    if (n_leaves > 0) {
        hipLaunchKernelGGL(k_finalize, dim3(grid_1d(n_leaves, 4)), dim3(256), 0, stream, leaf_keys, leaf_masks,
                           leaf_start, n_leaves, leaf_word_base, ijk_out, jidx_out);
        FVDBK_LAUNCH_CHECK("k_finalize");
    }
    if (grid_counts && n_grids > 0) {
        hipLaunchKernelGGL(k_grid_counts, dim3(grid_1d(n_grids, 64)), dim3(64), 0, stream, leaf_keys, leaf_start,
                           n_leaves, n_voxels, n_grids, grid_counts);
        FVDBK_LAUNCH_CHECK("k_grid_counts");
    }
    return FVDBK_OK;
}

int fvdbk_grid_ijk_to_index(const fvdbk_grid* grid, const int32_t* ijk, const int32_t* batch_idx, int64_t n,
                            int32_t* out_index, hipStream_t stream) {
    FVDBK_REQUIRE(grid, FVDBK_EINVAL, "null grid");
    if (n == 0) return FVDBK_OK;
    hipLaunchKernelGGL(k_ijk_to_index, dim3(grid_1d(n, 256)), dim3(256), 0, stream, *grid, ijk, batch_idx, n,
                       out_index);
    FVDBK_LAUNCH_CHECK("k_ijk_to_index");
    return FVDBK_OK;
}

int fvdbk_grid_ijk_to_index64(const fvdbk_grid* grid, const int32_t* ijk, const int32_t* batch_idx, int64_t n,
                              const int64_t* offsets, int64_t* out_index, hipStream_t stream) {
    FVDBK_REQUIRE(grid && n >= 0, FVDBK_EINVAL, "ijk_to_index64: bad arguments");
    if (n == 0) return FVDBK_OK;
    FVDBK_REQUIRE(ijk && out_index, FVDBK_EINVAL, "ijk_to_index64: null argument");
    hipLaunchKernelGGL(k_ijk_to_index64, dim3(grid_1d(n, 256)), dim3(256), 0, stream, *grid, ijk, batch_idx, n,
                       offsets, out_index);
    FVDBK_LAUNCH_CHECK("k_ijk_to_index64");
    return FVDBK_OK;
}

int fvdbk_points_to_ijk(fvdbk_dtype dtype, const void* points, const int32_t* batch_idx, int64_t n,
                        const double* transforms, int32_t* ijk_out, hipStream_t stream) {
    if (n == 0) return FVDBK_OK;
    if (dtype == FVDBK_F32) {
        hipLaunchKernelGGL(k_points_to_ijk<float>, dim3(grid_1d(n, 256)), dim3(256), 0, stream,
                           (const float*)points, batch_idx, n, transforms, ijk_out);
    } else if (dtype == FVDBK_F64) {
        hipLaunchKernelGGL(k_points_to_ijk<double>, dim3(grid_1d(n, 256)), dim3(256), 0, stream,
                           (const double*)points, batch_idx, n, transforms, ijk_out);
    } else {
        return set_error(FVDBK_ETYPE, "points must be float32 or float64");
    }
    FVDBK_LAUNCH_CHECK("k_points_to_ijk");
    return FVDBK_OK;
}

static bool is_coarse(const int32_t ks[3], const int32_t st[3]) {
    const bool one = st[0] == 1 && st[1] == 1 && st[2] == 1;
    const bool same = st[0] == ks[0] && st[1] == ks[1] && st[2] == ks[2];
    return one || same;
}

int64_t fvdbk_conv_grid_candidates_per_voxel(const int32_t ks[3], const int32_t stride[3]) {
    return is_coarse(ks, stride) ? 1 : (int64_t)ks[0] * ks[1] * ks[2];
}

int fvdbk_conv_grid_candidates(const int32_t* ijk, const int32_t* jidx, int64_t n, const int32_t ks[3],
                               const int32_t stride[3], int32_t* cand_ijk, int32_t* cand_jidx, hipStream_t stream) {
    for (int d = 0; d < 3; ++d) {
        FVDBK_REQUIRE(ks[d] > 0, FVDBK_EINVAL, "kernel_size must be strictly positive.");
        FVDBK_REQUIRE(stride[d] > 0, FVDBK_EINVAL, "stride must be strictly positive.");
    }
    if (n == 0) return FVDBK_OK;
    CandGeom g{};
    g.coarse = is_coarse(ks, stride) ? 1 : 0;
    g.kvol = ks[0] * ks[1] * ks[2];
    for (int d = 0; d < 3; ++d) {
        g.ks[d] = ks[d];
        g.st[d] = stride[d];
        if (ks[d] % 2 == 0) {
            g.lo[d] = 0;
            g.hi[d] = ks[d] - 1;
        } else {
            g.lo[d] = -(ks[d] - 1) / 2;
            g.hi[d] = (ks[d] - 1) / 2;
        }
    }
    hipLaunchKernelGGL(k_conv_candidates, dim3(grid_1d(n, 256)), dim3(256), 0, stream, ijk, jidx, n, g, cand_ijk,
                       cand_jidx);
    FVDBK_LAUNCH_CHECK("k_conv_candidates");
    return FVDBK_OK;
}

// Shared launch of the kernel-map kernel for a probe geometry (x, y, z order).
static int kmap_launch(const fvdbk_grid* rows, const fvdbk_grid* probe, ProbeGeom g, int32_t* map,
                       int64_t n_rows_padded, hipStream_t stream) {
    FVDBK_REQUIRE(rows->n_grids == probe->n_grids, FVDBK_EINVAL,
                  "source and target grids must hold the same number of grids (%d vs %d)", rows->n_grids,
                  probe->n_grids);
    FVDBK_REQUIRE(n_rows_padded >= rows->n_voxels, FVDBK_EINVAL, "map has fewer rows than the row grid has voxels");
    const int64_t K = g.K;
    FVDBK_REQUIRE(K <= 4096, FVDBK_EINVAL, "kernel volume %lld too large", (long long)K);
    FVDBK_REQUIRE(rows->n_voxels * K < ((int64_t)1 << 40), FVDBK_EINVAL, "kernel map too large");
    for (int a = 0; a < 3; ++a) {
        g.shift[a] = -1;
        for (int sh = 0; sh < 31; ++sh)
            if ((1 << sh) == g.div[a]) g.shift[a] = sh;
    }
    // 2x2x2 kernels with stride 2, both directions: the fixed-geometry kernel
    bool s2_down = g.K == 8 && g.sgn == 1, s2_up = g.K == 8 && g.sgn == -1;
    for (int a = 0; a < 3; ++a) {
        s2_down = s2_down && g.ext[a] == 2 && g.off0[a] == 0 && g.mul[a] == 2 && g.div[a] == 1;
        s2_up = s2_up && g.ext[a] == 2 && g.off0[a] == 0 && g.mul[a] == 1 && g.div[a] == 2;
    }
    if ((s2_down || s2_up) && ((uintptr_t)map % 16) == 0) {
        const int64_t tail = (n_rows_padded - rows->n_voxels) * K;
        const bool fold_tail = rows->n_leaves > 0 && tail > 0 && tail <= 8192;
        if (rows->n_leaves > 0) {
            const int ng = (int)((rows->n_leaves + KMAP_S2_WAVES - 1) / KMAP_S2_WAVES);
            const unsigned blocks = (unsigned)((ng + (fold_tail ? 1 : 0) + 7) / 8 * 8);
            const int tl = fold_tail ? (int)tail : 0;
            if (s2_down)
                hipLaunchKernelGGL(k_kmap_s2<false>, dim3(blocks), dim3(64 * KMAP_S2_WAVES), 0, stream, *rows, *probe,
                                   ng, tl, map);
            else
                hipLaunchKernelGGL(k_kmap_s2<true>, dim3(blocks), dim3(64 * KMAP_S2_WAVES), 0, stream, *rows, *probe,
                                   ng, tl, map);
            FVDBK_LAUNCH_CHECK("k_kmap_s2");
        }
        if (tail > 0 && !fold_tail) {
            hipLaunchKernelGGL(k_fill_i32, dim3(grid_1d(tail, 256)), dim3(256), 0, stream, map + rows->n_voxels * K,
                               tail, -1);
            FVDBK_LAUNCH_CHECK("k_fill_i32");
        }
        return FVDBK_OK;
    }
    if (rows->n_leaves > 0) {
        int nb = kmap_box_bound(g);
        if (nb * KMAP_STAGE_PER_LEAF + KMAP_ROW_BYTES > KMAP_STAGE_BYTES) nb = 0;  // too large to stage: probe the hash per pair
        const size_t lds = (size_t)KMAP_WAVES * (nb * KMAP_STAGE_PER_LEAF + KMAP_ROW_BYTES) + (size_t)K * 12;
        FVDBK_REQUIRE(lds <= 65536, FVDBK_EINVAL, "kernel volume %lld too large for the kernel-map launch",
                      (long long)K);
        g.n_blocks = (int)((rows->n_leaves + KMAP_WAVES - 1) / KMAP_WAVES);
        // a short padding tail is written by an extra (last) workgroup of the launch
        const int64_t tail = (n_rows_padded - rows->n_voxels) * K;
        const bool fold_tail = tail > 0 && tail <= 8192;
        g.tail_len = fold_tail ? (int)tail : 0;
        g.tail_off = rows->n_voxels * K;
        const unsigned blocks = (unsigned)((g.n_blocks + (fold_tail ? 1 : 0) + 7) / 8 * 8);
        const bool unit = g.mul[0] == 1 && g.mul[1] == 1 && g.mul[2] == 1 && g.div[0] == 1 && g.div[1] == 1 &&
                          g.div[2] == 1;
        if (unit)
            hipLaunchKernelGGL(k_kmap<true>, dim3(blocks), dim3(64 * KMAP_WAVES), lds, stream, *rows, *probe, g, nb,
                               map);
        else
            hipLaunchKernelGGL(k_kmap<false>, dim3(blocks), dim3(64 * KMAP_WAVES), lds, stream, *rows, *probe, g, nb,
                               map);
        FVDBK_LAUNCH_CHECK("k_kmap");
        if (fold_tail) return FVDBK_OK;
    }
    const int64_t tail = (n_rows_padded - rows->n_voxels) * K;
    if (tail > 0) {
        hipLaunchKernelGGL(k_fill_i32, dim3(grid_1d(tail, 256)), dim3(256), 0, stream, map + rows->n_voxels * K, tail,
                           -1);
        FVDBK_LAUNCH_CHECK("k_fill_i32");
    }
    return FVDBK_OK;
}

int fvdbk_kmap_build(const fvdbk_grid* rows, const fvdbk_grid* probe, const int32_t ks[3], const int32_t stride[3],
                     int32_t inverse, int32_t* map, int64_t n_rows_padded, hipStream_t stream) {
    FVDBK_REQUIRE(rows && probe, FVDBK_EINVAL, "null grid");
    FVDBK_REQUIRE(map || n_rows_padded == 0, FVDBK_EINVAL, "null kernel map");
    for (int d = 0; d < 3; ++d) {
        FVDBK_REQUIRE(ks[d] > 0, FVDBK_EINVAL, "Expect kernel size to be larger than {0,0,0}.");
        FVDBK_REQUIRE(stride[d] > 0, FVDBK_EINVAL, "Expect stride to be larger than 0.");
    }
    ProbeGeom g{};
    g.K = ks[0] * ks[1] * ks[2];
    // coordinate axis a (x,y,z) uses kernel dim 2-a of the (d,h,w) convention
    for (int a = 0; a < 3; ++a) {
        const int kd = 2 - a;
        const int kstart = (int)floor(-ks[kd] / 2.0 + 1.0);
        g.ext[a] = ks[kd];
        if (!inverse) {
            g.mul[a] = stride[kd];
            g.div[a] = 1;
            g.off0[a] = kstart;
        } else {
            g.mul[a] = 1;
            g.div[a] = stride[kd];
            g.off0[a] = -kstart;
        }
    }
    g.sgn = inverse ? -1 : 1;
    return kmap_launch(rows, probe, g, map, n_rows_padded, stream);
}

int fvdbk_kmap_row_order_windowed(const int32_t* map, int64_t n_rows, int32_t K, int32_t window, int32_t* order,
                                  hipStream_t stream) {
    FVDBK_REQUIRE(K > 0, FVDBK_EINVAL, "kernel volume must be positive");
    FVDBK_REQUIRE(window >= 256 && window <= ORDER_BLOCK && (window & (window - 1)) == 0, FVDBK_EINVAL,
                  "window must be a power of two in [256, %d], got %d", ORDER_BLOCK, (int)window);
    FVDBK_REQUIRE(n_rows >= 0 && n_rows < ((int64_t)1 << 31), FVDBK_EINVAL, "row count out of range");
    if (n_rows == 0) return FVDBK_OK;
    const int log2w = __builtin_ctz((unsigned)window);
    // windows of up to 2048 rows: 2048-row workgroups (4x the workgroups of 8192-row ones)
    const bool small = window <= ORDER_SMALL_BLOCK;
    const int64_t rows_per_block = small ? ORDER_SMALL_BLOCK : ORDER_BLOCK;
    const int64_t blocks = (n_rows + rows_per_block - 1) / rows_per_block;
    const dim3 grid((unsigned)blocks), block(ORDER_THREADS);
    if (small) {
        if (K == 27)
            hipLaunchKernelGGL((k_row_order_window<27, 2>), grid, block, 0, stream, map, n_rows, (int)K, log2w, order);
        else if (K == 8)
            hipLaunchKernelGGL((k_row_order_window<8, 2>), grid, block, 0, stream, map, n_rows, (int)K, log2w, order);
        else
            hipLaunchKernelGGL((k_row_order_window<0, 2>), grid, block, 0, stream, map, n_rows, (int)K, log2w, order);
    } else {
        if (K == 27)
            hipLaunchKernelGGL((k_row_order_window<27, 8>), grid, block, 0, stream, map, n_rows, (int)K, log2w, order);
        else if (K == 8)
            hipLaunchKernelGGL((k_row_order_window<8, 8>), grid, block, 0, stream, map, n_rows, (int)K, log2w, order);
        else
            hipLaunchKernelGGL((k_row_order_window<0, 8>), grid, block, 0, stream, map, n_rows, (int)K, log2w, order);
    }
    FVDBK_LAUNCH_CHECK("k_row_order_window");
    return FVDBK_OK;
}

int fvdbk_kmap_window(const fvdbk_grid* rows, const fvdbk_grid* probe, const int32_t window[3],
                      const int32_t stride[3], int32_t* map, int64_t n_rows_padded, hipStream_t stream) {
    FVDBK_REQUIRE(rows && probe, FVDBK_EINVAL, "null grid");
    FVDBK_REQUIRE(map || n_rows_padded == 0, FVDBK_EINVAL, "null window map");
    for (int a = 0; a < 3; ++a) {
        FVDBK_REQUIRE(window[a] > 0, FVDBK_EINVAL, "pooling_factor must be greater than 0");
        FVDBK_REQUIRE(stride[a] > 0, FVDBK_EINVAL, "stride must be greater than 0");
    }
    ProbeGeom g{};
    g.K = window[0] * window[1] * window[2];
    for (int a = 0; a < 3; ++a) {
        g.ext[a] = window[a];
        g.mul[a] = stride[a];
        g.div[a] = 1;
        g.off0[a] = 0;
    }
    g.sgn = 1;
    return kmap_launch(rows, probe, g, map, n_rows_padded, stream);
}

int fvdbk_nbmap_block_counts(const int32_t* map, int64_t n_rows, int32_t K, int32_t* counts, hipStream_t stream) {
    FVDBK_REQUIRE(K > 0, FVDBK_EINVAL, "kernel volume must be positive");
    if (n_rows == 0) return FVDBK_OK;
    const unsigned nblk = grid_1d(n_rows, FVDBK_NBMAP_ROWS);
    hipLaunchKernelGGL(k_nbmap_counts, dim3(nblk), dim3(NB_THREADS), 0, stream, map, n_rows, (int)K, counts);
    FVDBK_LAUNCH_CHECK("k_nbmap_counts");
    return FVDBK_OK;
}

int fvdbk_nbmap_emit(const int32_t* map, int64_t n_rows, int32_t K, const int64_t* block_offsets, int32_t* pairs,
                     hipStream_t stream) {
    FVDBK_REQUIRE(K > 0, FVDBK_EINVAL, "kernel volume must be positive");
    if (n_rows == 0) return FVDBK_OK;
    const unsigned nblk = grid_1d(n_rows, FVDBK_NBMAP_ROWS);
    hipLaunchKernelGGL(k_nbmap_emit, dim3(nblk), dim3(NB_THREADS), 0, stream, map, n_rows, (int)K, block_offsets,
                       pairs);
    FVDBK_LAUNCH_CHECK("k_nbmap_emit");
    return FVDBK_OK;
}

}  // extern "C"
