/*
 * CPU oracle for the integer half of fVDB's sparse-convolution path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP kernels of
 * the fvdb-core_amd package: only tests, the smoke check and the bench's CPU
 * baseline may load it.  The product path never calls into it.
 *
 * It restates, single-threaded and without any shared code with the GPU path:
 *
 *   - the voxel linear order of a NanoVDB OnIndex grid (the order in which the
 *     reference numbers active voxels; GridBatchImpl.h:178-181 voxelOffset plus
 *     leaf.getValue(v)-1 at ConvolutionKernelMap.cu:40,50).  Order within a grid is
 *     lexicographic on (root tile, upper-node child, lower-node child, leaf bit)
 *     where the root tile is (uint32(c) >> 12) per axis (the 64-bit root key of the
 *     GPU builder used by BuildGridFromIjk.cu:27-84), the upper child is
 *     (c >> 7) & 31, the lower child (c >> 3) & 15 and the leaf bit (c & 7),
 *     each compared x, then y, then z.
 *   - the convolution target grid of BuildGridForConv.cu:190-255 (CPU path) and
 *     CoarseIjkForFineGrid.cu:18-40 (floor(ijk / stride) when stride == 1 or
 *     stride == kernel size).
 *   - the kernel map of ConvolutionKernelMap.cu:61-102 (convolutionKernelMapCPU).
 *   - the gather-scatter neighbour map of SparseConvPackInfo.cpp:61-88.
 *
 * Parity status of each function is stated in the tests that pin it.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>

/* ------------------------------------------------------------------------- */
/* NanoVDB tree order                                                         */
/* ------------------------------------------------------------------------- */

static int cmp_u32(uint32_t a, uint32_t b) { return (a > b) - (a < b); }

/* Compare two coordinates in NanoVDB OnIndex tree order.  Returns <0, 0, >0. */
int oracle_cmp_order(const int32_t* a, const int32_t* b) {
    int c;
    for (int d = 0; d < 3; ++d) { /* root tile, unsigned key as the GPU builder forms it */
        c = cmp_u32(((uint32_t)a[d]) >> 12, ((uint32_t)b[d]) >> 12);
        if (c) return c;
    }
    for (int d = 0; d < 3; ++d) { /* child of the 4096^3 upper node: 32^3 table */
        c = cmp_u32((uint32_t)(a[d] >> 7) & 31u, (uint32_t)(b[d] >> 7) & 31u);
        if (c) return c;
    }
    for (int d = 0; d < 3; ++d) { /* child of the 128^3 lower node: 16^3 table */
        c = cmp_u32((uint32_t)(a[d] >> 3) & 15u, (uint32_t)(b[d] >> 3) & 15u);
        if (c) return c;
    }
    for (int d = 0; d < 3; ++d) { /* bit inside the 8^3 leaf: (x&7)<<6 | (y&7)<<3 | (z&7) */
        c = cmp_u32((uint32_t)a[d] & 7u, (uint32_t)b[d] & 7u);
        if (c) return c;
    }
    return 0;
}

static int qsort_cmp(const void* a, const void* b) {
    return oracle_cmp_order((const int32_t*)a, (const int32_t*)b);
}

/* Sort ijk[n][3] into tree order and drop duplicates in place; returns the count. */
int64_t oracle_sort_unique(int32_t* ijk, int64_t n) {
    if (n <= 0) return 0;
    qsort(ijk, (size_t)n, 3 * sizeof(int32_t), qsort_cmp);
    int64_t m = 1;
    for (int64_t i = 1; i < n; ++i) {
        if (oracle_cmp_order(ijk + 3 * i, ijk + 3 * (m - 1)) != 0) {
            if (m != i) memcpy(ijk + 3 * m, ijk + 3 * i, 3 * sizeof(int32_t));
            ++m;
        }
    }
    return m;
}

/* ------------------------------------------------------------------------- */
/* Coordinate hash (open addressing) used only by the oracle                  */
/* ------------------------------------------------------------------------- */

typedef struct {
    int64_t cap;     /* power of two */
    int32_t* keys;   /* [cap][3] */
    int64_t* vals;   /* -1 = empty */
} chash;

static uint64_t mix3(int32_t x, int32_t y, int32_t z) {
    uint64_t h = (uint64_t)(uint32_t)x * 0x9E3779B97F4A7C15ull;
    h ^= (uint64_t)(uint32_t)y * 0xC2B2AE3D27D4EB4Full + (h << 6) + (h >> 2);
    h ^= (uint64_t)(uint32_t)z * 0x165667B19E3779F9ull + (h << 6) + (h >> 2);
    h ^= h >> 31;
    return h;
}

static int chash_init(chash* t, int64_t n) {
    int64_t cap = 16;
    while (cap < 2 * n + 16) cap <<= 1;
    t->cap = cap;
    t->keys = (int32_t*)malloc((size_t)cap * 3 * sizeof(int32_t));
    t->vals = (int64_t*)malloc((size_t)cap * sizeof(int64_t));
    if (!t->keys || !t->vals) return -1;
    for (int64_t i = 0; i < cap; ++i) t->vals[i] = -1;
    return 0;
}

static void chash_free(chash* t) { free(t->keys); free(t->vals); }

static void chash_put(chash* t, const int32_t* c, int64_t v) {
    uint64_t h = mix3(c[0], c[1], c[2]) & (uint64_t)(t->cap - 1);
    for (;;) {
        if (t->vals[h] < 0) {
            t->keys[3 * h] = c[0]; t->keys[3 * h + 1] = c[1]; t->keys[3 * h + 2] = c[2];
            t->vals[h] = v;
            return;
        }
        if (t->keys[3 * h] == c[0] && t->keys[3 * h + 1] == c[1] && t->keys[3 * h + 2] == c[2]) {
            t->vals[h] = v;
            return;
        }
        h = (h + 1) & (uint64_t)(t->cap - 1);
    }
}

static int64_t chash_get(const chash* t, int32_t x, int32_t y, int32_t z) {
    uint64_t h = mix3(x, y, z) & (uint64_t)(t->cap - 1);
    for (;;) {
        if (t->vals[h] < 0) return -1;
        if (t->keys[3 * h] == x && t->keys[3 * h + 1] == y && t->keys[3 * h + 2] == z) return t->vals[h];
        h = (h + 1) & (uint64_t)(t->cap - 1);
    }
}

/* ------------------------------------------------------------------------- */
/* Convolution target grid                                                    */
/* ------------------------------------------------------------------------- */

static int32_t floordiv(int32_t a, int32_t b) {
    int32_t q = a / b;
    if ((a % b != 0) && ((a < 0) != (b < 0))) --q;
    return q;
}

/*
 * Upper bound on the number of candidate target voxels before de-duplication.
 * ks and stride use the reference's (d, h, w) convention: ks[2] and stride[2]
 * act on the x (first) coordinate, ks[0] and stride[0] on z.
 */
int64_t oracle_conv_grid_capacity(int64_t n, const int32_t ks[3], const int32_t stride[3]) {
    int coarse = (stride[0] == 1 && stride[1] == 1 && stride[2] == 1) ||
                 (stride[0] == ks[0] && stride[1] == ks[1] && stride[2] == ks[2]);
    return coarse ? n : n * (int64_t)ks[0] * ks[1] * ks[2];
}

/*
 * Target grid of a strided convolution over the voxels src[n][3].
 * Follows BuildGridForConv.cu:190-255: with stride == 1 or stride == kernel size the
 * target is {floor(ijk / stride)} (buildCoarseGridFromFineGridCPU, :25-61, with a
 * per-axis branching factor equal to the stride; note the branching factor is indexed
 * x,y,z there while the stride is (d,h,w), which only matters for anisotropic strides
 * and is reproduced as is).  Otherwise every offset d in [lo, hi] per axis (lo = 0,
 * hi = ks-1 for an even kernel size, lo = -(ks-1)/2, hi = (ks-1)/2 for an odd one) is
 * added and kept when divisible by the stride (C remainder semantics, :230-243).
 * Writes the sorted, unique target coordinates to out and returns their count.
 */
int64_t oracle_conv_grid(const int32_t* src, int64_t n, const int32_t ks[3], const int32_t stride[3],
                         int32_t* out) {
    int64_t m = 0;
    int is_one = stride[0] == 1 && stride[1] == 1 && stride[2] == 1;
    int is_ks = stride[0] == ks[0] && stride[1] == ks[1] && stride[2] == ks[2];
    if (is_one || is_ks) {
        for (int64_t i = 0; i < n; ++i) {
            /* branching factor indexed by coordinate axis (x uses stride[0]) */
            out[3 * m + 0] = floordiv(src[3 * i + 0], stride[0]);
            out[3 * m + 1] = floordiv(src[3 * i + 1], stride[1]);
            out[3 * m + 2] = floordiv(src[3 * i + 2], stride[2]);
            ++m;
        }
        return oracle_sort_unique(out, m);
    }
    int lower[3], upper[3];
    for (int d = 0; d < 3; ++d) {
        if (ks[d] % 2 == 0) { lower[d] = 0; upper[d] = ks[d] - 1; }
        else { lower[d] = -(ks[d] - 1) / 2; upper[d] = (ks[d] - 1) / 2; }
    }
    for (int64_t v = 0; v < n; ++v) {
        const int32_t* c = src + 3 * v;
        for (int di = lower[0]; di <= upper[0]; ++di)
            for (int dj = lower[1]; dj <= upper[1]; ++dj)
                for (int dk = lower[2]; dk <= upper[2]; ++dk) {
                    int32_t x = c[0] + dk, y = c[1] + dj, z = c[2] + di;
                    if (x % stride[2] != 0 || y % stride[1] != 0 || z % stride[0] != 0) continue;
                    out[3 * m + 0] = x / stride[2];
                    out[3 * m + 1] = y / stride[1];
                    out[3 * m + 2] = z / stride[0];
                    ++m;
                }
    }
    return oracle_sort_unique(out, m);
}

/* ------------------------------------------------------------------------- */
/* Kernel map                                                                 */
/* ------------------------------------------------------------------------- */

/*
 * out_in_map[t][kIdx] for one grid of a batch, following convolutionKernelMapCPU
 * (ConvolutionKernelMap.cu:61-102).  src/tgt are the grid's voxels in tree order;
 * src_base is the cumulative voxel offset of the grid in the source batch.  The
 * target voxel t probes (t.x*stride[2], t.y*stride[1], t.z*stride[0]) + (kz, ky, kx)
 * with kz over kernel dim 2 (outer loop), kx over kernel dim 0 (inner loop) and
 * kernelStart[d] = floor(-ks[d]/2 + 1).  map rows are written in target order.
 */
int oracle_kernel_map(const int32_t* src, int64_t n_src, int64_t src_base,
                      const int32_t* tgt, int64_t n_tgt,
                      const int32_t ks[3], const int32_t stride[3], int32_t* map) {
    chash h;
    if (chash_init(&h, n_src) != 0) return -1;
    for (int64_t i = 0; i < n_src; ++i) chash_put(&h, src + 3 * i, i);
    int st[3];
    for (int d = 0; d < 3; ++d) st[d] = (int)floor(-ks[d] / 2.0 + 1.0);
    const int K = ks[0] * ks[1] * ks[2];
    for (int64_t t = 0; t < n_tgt; ++t) {
        const int32_t sx = tgt[3 * t + 0] * stride[2];
        const int32_t sy = tgt[3 * t + 1] * stride[1];
        const int32_t sz = tgt[3 * t + 2] * stride[0];
        int k = 0;
        for (int kz = st[2]; kz < st[2] + ks[2]; ++kz)
            for (int ky = st[1]; ky < st[1] + ks[1]; ++ky)
                for (int kx = st[0]; kx < st[0] + ks[0]; ++kx, ++k) {
                    int64_t v = chash_get(&h, sx + kz, sy + ky, sz + kx);
                    map[t * K + k] = v < 0 ? -1 : (int32_t)(v + src_base);
                }
    }
    chash_free(&h);
    return 0;
}

/*
 * Gather-scatter neighbour map from a kernel map [n_tgt][K]
 * (SparseConvPackInfo.cpp:61-88): pairs ordered by (k, target row), each pair
 * (source index, target row).  nbsizes[k] counts the pairs of offset k.
 * Returns the number of pairs written (nbmap must hold n_tgt*K pairs).
 */
int64_t oracle_neighbor_map(const int32_t* map, int64_t n_tgt, int32_t K, int32_t* nbmap, int32_t* nbsizes) {
    int64_t p = 0;
    for (int32_t k = 0; k < K; ++k) {
        int32_t c = 0;
        for (int64_t t = 0; t < n_tgt; ++t) {
            int32_t s = map[t * K + k];
            if (s >= 0) {
                nbmap[2 * p] = s;
                nbmap[2 * p + 1] = (int32_t)t;
                ++p;
                ++c;
            }
        }
        nbsizes[k] = c;
    }
    return p;
}

/*
 * Transposed map (IGEMMBitOperations.cu:92-105): T[s][k] = t for every valid
 * map[t][k] = s; rows of T start at -1.
 */
void oracle_transpose_map(const int32_t* map, int64_t n_tgt, int32_t K, int32_t* tmap, int64_t n_src) {
    for (int64_t i = 0; i < n_src * K; ++i) tmap[i] = -1;
    for (int64_t t = 0; t < n_tgt; ++t)
        for (int32_t k = 0; k < K; ++k) {
            int32_t s = map[t * K + k];
            if (s >= 0) tmap[(int64_t)s * K + k] = (int32_t)t;
        }
}
