// Grid index build, point voxelisation, conv target-grid candidates, kernel map and
// gather-scatter neighbour map for gfx950.  All integer work; results are bit-exact
// and independent of launch shape.
#include <hip/hip_runtime.h>

#include <stdarg.h>
#include <stdio.h>

#include <string>

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

__global__ void k_leaf_masks(const int32_t* __restrict__ ijk, const int32_t* __restrict__ bidx, int64_t n,
                             const fvdbk_hslot* __restrict__ h, int64_t cap, uint64_t* __restrict__ masks) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int32_t b = bidx ? bidx[i] : 0;
    if (b < 0 || b >= kMaxGrids) return;
    const int32_t x = ijk[3 * i], y = ijk[3 * i + 1], z = ijk[3 * i + 2];
    if (x < kCoordMin || x > kCoordMax || y < kCoordMin || y > kCoordMax || z < kCoordMin || z > kCoordMax) return;
    const int32_t leaf = hash_find(h, cap, leaf_key(b, x, y, z));
    if (leaf < 0) return;  // cannot happen for keys produced by k_leaf_keys
    const int bit = ((x & 7) << 6) | ((y & 7) << 3) | (z & 7);
    atomicOr((unsigned long long*)&masks[(int64_t)leaf * 8 + (bit >> 6)], 1ull << (bit & 63));
}

__global__ void k_leaf_counts(const uint64_t* __restrict__ masks, int64_t n_leaves, int32_t* __restrict__ counts) {
    const int64_t l = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= n_leaves) return;
    int c = 0;
#pragma unroll
    for (int w = 0; w < 8; ++w) c += __popcll(masks[l * 8 + w]);
    counts[l] = c;
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
                                                  int32_t* __restrict__ jidx_out, int32_t* __restrict__ grid_counts,
                                                  int32_t n_grids) {
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
    if (grid_counts && lane == 0 && b < n_grids) atomicAdd(&grid_counts[b], (int32_t)(base - leaf_start[l]));
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
    int off0[3];
    int sgn;
};

constexpr int KMAP_THREADS = 256;
constexpr int KMAP_MAX_NB = 216;  // staged neighbourhood: up to 6x6x6 probe leaves

// One workgroup per row-grid leaf.  The probe-grid leaves that the leaf's taps can
// reach form a small box; their masks and word bases are staged in LDS once, then
// every (row voxel, tap) pair of the leaf is resolved from LDS and the map block of
// the leaf (contiguous rows) is written with consecutive threads on consecutive
// int32 entries.  Boxes larger than KMAP_MAX_NB leaves fall back to a hash probe per
// pair.
__global__ __launch_bounds__(KMAP_THREADS) void k_kmap(fvdbk_grid rows, fvdbk_grid probe, ProbeGeom g,
                                                       int32_t* __restrict__ map) {
    __shared__ uint64_t s_mask[KMAP_MAX_NB * 8];
    __shared__ int32_t s_base[KMAP_MAX_NB * 8];
    __shared__ int32_t s_ijk[512 * 3];

    const int64_t L = blockIdx.x;
    const int tid = threadIdx.x;
    int32_t b, o[3];
    decode_leaf_key(rows.leaf_keys[L], b, o[0], o[1], o[2]);

    const int64_t vbase = rows.leaf_word_base[L * 8];
    const int cnt = (int)(rows.leaf_word_base[L * 8 + 7] - vbase) + __popcll(rows.leaf_masks[L * 8 + 7]);

    int lo[3], ext[3];
    for (int a = 0; a < 3; ++a) {
        const int span = g.sgn * (g.ext[a] - 1);
        const int off_lo = g.off0[a] + (span < 0 ? span : 0);
        const int off_hi = g.off0[a] + (span > 0 ? span : 0);
        const int q_lo = floordiv(o[a] * g.mul[a] + off_lo, g.div[a]);
        const int q_hi = floordiv((o[a] + 7) * g.mul[a] + off_hi, g.div[a]);
        lo[a] = q_lo >> 3;
        ext[a] = (q_hi >> 3) - lo[a] + 1;
    }
    const int nn = ext[0] * ext[1] * ext[2];
    const bool staged = nn <= KMAP_MAX_NB;

    if (staged) {
        for (int n = tid; n < nn; n += KMAP_THREADS) {
            const int dz = n % ext[2], dy = (n / ext[2]) % ext[1], dx = n / (ext[2] * ext[1]);
            const int32_t lx = (lo[0] + dx) * 8, ly = (lo[1] + dy) * 8, lz = (lo[2] + dz) * 8;
            int32_t leaf = -1;
            if (probe.n_leaves > 0 && lx >= kCoordMin && lx <= kCoordMax && ly >= kCoordMin && ly <= kCoordMax &&
                lz >= kCoordMin && lz <= kCoordMax)
                leaf = hash_find(probe.hash, probe.hash_cap, leaf_key(b, lx, ly, lz));
            for (int w = 0; w < 8; ++w) {
                s_mask[n * 8 + w] = leaf >= 0 ? probe.leaf_masks[(int64_t)leaf * 8 + w] : 0ull;
                s_base[n * 8 + w] = leaf >= 0 ? probe.leaf_word_base[(int64_t)leaf * 8 + w] : 0;
            }
        }
    }
    for (int e = tid; e < 3 * cnt; e += KMAP_THREADS) s_ijk[e] = rows.ijk[vbase * 3 + e];
    __syncthreads();

    const int K = g.K;
    const int total = cnt * K;
    int32_t* __restrict__ out = map + vbase * K;
    for (int idx = tid; idx < total; idx += KMAP_THREADS) {
        const int r = idx / K;
        const int k = idx - r * K;
        const int tz = k % g.ext[2];
        const int ty = (k / g.ext[2]) % g.ext[1];
        const int tx = k / (g.ext[2] * g.ext[1]);
        const int t[3] = {tx, ty, tz};
        int32_t q[3];
        bool ok = true;
        for (int a = 0; a < 3; ++a) {
            const int num = s_ijk[r * 3 + a] * g.mul[a] + g.off0[a] + g.sgn * t[a];
            if (num % g.div[a] != 0) ok = false;
            q[a] = num / g.div[a];
        }
        int32_t res = -1;
        if (ok) {
            if (staged) {
                const int n = (((q[0] >> 3) - lo[0]) * ext[1] + ((q[1] >> 3) - lo[1])) * ext[2] + ((q[2] >> 3) - lo[2]);
                const int bit = ((q[0] & 7) << 6) | ((q[1] & 7) << 3) | (q[2] & 7);
                const uint64_t m = s_mask[n * 8 + (bit >> 6)];
                if ((m >> (bit & 63)) & 1ull)
                    res = s_base[n * 8 + (bit >> 6)] + __popcll(m & ((1ull << (bit & 63)) - 1ull));
            } else if (probe.n_leaves > 0) {
                res = voxel_index(probe, b, q[0], q[1], q[2]);
            }
        }
        out[idx] = res;
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

int fvdbk_grid_leaf_counts(const uint64_t* leaf_masks, int64_t n_leaves, int32_t* counts, hipStream_t stream) {
    if (n_leaves == 0) return FVDBK_OK;
    hipLaunchKernelGGL(k_leaf_counts, dim3(grid_1d(n_leaves, 256)), dim3(256), 0, stream, leaf_masks, n_leaves,
                       counts);
    FVDBK_LAUNCH_CHECK("k_leaf_counts");
    return FVDBK_OK;
}

int fvdbk_grid_finalize(const uint64_t* leaf_keys, const uint64_t* leaf_masks, const int64_t* leaf_start,
                        int64_t n_leaves, int32_t* leaf_word_base, int32_t* ijk_out, int32_t* jidx_out,
                        int32_t* grid_counts, int32_t n_grids, hipStream_t stream) {
    if (grid_counts && n_grids > 0) {
        hipLaunchKernelGGL(k_zero_i32, dim3(grid_1d(n_grids, 256)), dim3(256), 0, stream, grid_counts,
                           (int64_t)n_grids);
        FVDBK_LAUNCH_CHECK("k_zero_i32");
    }
    if (n_leaves == 0) return FVDBK_OK;
    hipLaunchKernelGGL(k_finalize, dim3(grid_1d(n_leaves, 4)), dim3(256), 0, stream, leaf_keys, leaf_masks,
                       leaf_start, n_leaves, leaf_word_base, ijk_out, jidx_out, grid_counts, n_grids);
    FVDBK_LAUNCH_CHECK("k_finalize");
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

int fvdbk_kmap_build(const fvdbk_grid* rows, const fvdbk_grid* probe, const int32_t ks[3], const int32_t stride[3],
                     int32_t inverse, int32_t* map, int64_t n_rows_padded, hipStream_t stream) {
    FVDBK_REQUIRE(rows && probe, FVDBK_EINVAL, "null grid");
    FVDBK_REQUIRE(map || n_rows_padded == 0, FVDBK_EINVAL, "null kernel map");
    for (int d = 0; d < 3; ++d) {
        FVDBK_REQUIRE(ks[d] > 0, FVDBK_EINVAL, "Expect kernel size to be larger than {0,0,0}.");
        FVDBK_REQUIRE(stride[d] > 0, FVDBK_EINVAL, "Expect stride to be larger than 0.");
    }
    FVDBK_REQUIRE(rows->n_grids == probe->n_grids, FVDBK_EINVAL,
                  "source and target grids must hold the same number of grids (%d vs %d)", rows->n_grids,
                  probe->n_grids);
    FVDBK_REQUIRE(n_rows_padded >= rows->n_voxels, FVDBK_EINVAL, "map has fewer rows than the row grid has voxels");
    const int64_t K = (int64_t)ks[0] * ks[1] * ks[2];
    FVDBK_REQUIRE(K <= 4096, FVDBK_EINVAL, "kernel volume %lld too large", (long long)K);
    FVDBK_REQUIRE(rows->n_voxels * K < ((int64_t)1 << 40), FVDBK_EINVAL, "kernel map too large");
    ProbeGeom g{};
    g.K = (int)K;
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
    if (rows->n_leaves > 0) {
        hipLaunchKernelGGL(k_kmap, dim3((unsigned)rows->n_leaves), dim3(KMAP_THREADS), 0, stream, *rows, *probe, g,
                           map);
        FVDBK_LAUNCH_CHECK("k_kmap");
    }
    const int64_t tail = (n_rows_padded - rows->n_voxels) * K;
    if (tail > 0) {
        hipLaunchKernelGGL(k_fill_i32, dim3(grid_1d(tail, 256)), dim3(256), 0, stream, map + rows->n_voxels * K, tail,
                           -1);
        FVDBK_LAUNCH_CHECK("k_fill_i32");
    }
    return FVDBK_OK;
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
