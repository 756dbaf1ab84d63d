/*
 * fvdbk — C ABI of the MI355X (gfx950) sparse-convolution kernels.
 *
 * Plain C: no torch types, only device pointers, sizes and a hipStream_t.  The
 * caller owns every buffer (inputs, outputs, workspaces) and the stream; no entry
 * point allocates device memory, synchronises the stream or reads device memory
 * on the host.  Every entry point returns FVDBK_OK (0) or a negative status and
 * records a message retrievable with fvdbk_last_error() (thread-local).  Status
 * classes mirror the reference's error classes: FVDBK_EINVAL is what the reference
 * raises with TORCH_CHECK_VALUE (Python ValueError), FVDBK_ETYPE with
 * TORCH_CHECK_TYPE (TypeError).
 *
 * Each declaration names the reference interface it replaces.  Reference paths are
 * relative to the fvdb-core repository root.
 */
#ifndef FVDBK_H
#define FVDBK_H

#include <stddef.h>
#include <stdint.h>

#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
extern "C" {
#endif

enum fvdbk_status {
    FVDBK_OK = 0,
    FVDBK_EINVAL = -1,      /* bad argument value: ValueError */
    FVDBK_ETYPE = -2,       /* unsupported dtype: TypeError */
    FVDBK_EHIP = -3,        /* HIP launch/runtime error */
    FVDBK_ENOSPC = -4,      /* workspace or LDS too small */
    FVDBK_EUNSUPPORTED = -5 /* valid but not implemented */
};

typedef enum fvdbk_dtype {
    FVDBK_F32 = 0,
    FVDBK_F64 = 1,
    FVDBK_F16 = 2,
    FVDBK_BF16 = 3
} fvdbk_dtype;

/* One open-addressing slot of a grid's leaf hash: key is the leaf's sort key,
 * leaf is its index in the sorted leaf arrays, -1 when the slot is empty. */
typedef struct fvdbk_hslot {
    uint64_t key;
    int32_t leaf;
    int32_t pad;
} fvdbk_hslot;

/*
 * Device-resident index of a batch of sparse grids (replaces the NanoVDB OnIndex
 * grid batch of src/fvdb/detail/GridBatchImpl.h:31-231 for the convolution path).
 *
 * Voxels are numbered in NanoVDB tree order: grid, then root tile, upper-node
 * child, lower-node child, leaf bit.  A leaf is an 8^3 brick with a 512-bit
 * activity mask; its sort key packs (grid, tile, upper child, lower child) so
 * that sorting keys sorts leaves in tree order.  leaf_word_base[l][w] is the
 * global index of the first active voxel of 64-bit word w of leaf l, so the index
 * of bit b of leaf l is leaf_word_base[l][b>>6] + popcount(mask[l][b>>6] & ((1<<(b&63))-1)).
 */
typedef struct fvdbk_grid {
    int64_t n_voxels;
    int64_t n_leaves;
    int64_t hash_cap;                /* power of two */
    int32_t n_grids;
    int32_t pad;
    const int32_t* ijk;              /* [n_voxels][3] in tree order */
    const uint64_t* leaf_keys;       /* [n_leaves] ascending */
    const uint64_t* leaf_masks;      /* [n_leaves][8] */
    const int32_t* leaf_word_base;   /* [n_leaves][8] */
    const fvdbk_hslot* hash;         /* [hash_cap] */
} fvdbk_grid;

/* Thread-local message of the last failing call on this thread. */
const char* fvdbk_last_error(void);

/* Library version string. */
const char* fvdbk_version(void);

/* ---------------------------------------------------------------------------
 * Grid index build (replaces nanovdb::tools::cuda::voxelsToGrid as called by
 * src/fvdb/detail/ops/BuildGridFromIjk.cu:27-84 and the per-voxel ijk/jidx export
 * of src/fvdb/detail/GridBatchImpl.cu).  A build is a short sequence because the
 * number of leaves is only known after de-duplicating keys:
 *   1. fvdbk_grid_leaf_keys     per input voxel: leaf sort key (INT64_MAX = skipped voxel)
 *   2. caller sorts + de-duplicates the keys (any device sort) -> n_leaves keys
 *   3. fvdbk_grid_hash_build    leaf hash over the unique keys
 *   4. fvdbk_grid_leaf_masks    512-bit masks (OR of the input voxels)
 *   5. fvdbk_grid_leaf_counts   active voxels per leaf
 *   6. caller exclusive-scans the counts -> leaf_start[n_leaves]
 *      (or 5-6 in one call: fvdbk_grid_leaf_scan, inclusive and exclusive scans)
 *   7. fvdbk_grid_finalize      word bases, ijk and grid index of every voxel
 * ------------------------------------------------------------------------- */

/* batch_idx may be NULL (all voxels in grid 0); entries < 0 are skipped.
 * Coordinates must lie in [-2^19, 2^19) and grid indices below 1024; *bad_count
 * (device int32) is incremented
 * per out-of-range voxel (the caller zeroes it and checks it). */
int fvdbk_grid_leaf_keys(const int32_t* ijk, const int32_t* batch_idx, int64_t n, uint64_t* keys,
                         int32_t* bad_count, hipStream_t stream);

/* Steps 1-2 without a sort of every voxel key: the unique leaf keys of the input
 * voxels, unsorted (the caller sorts only these n_leaves keys).  set is scratch of
 * set_cap u64 (fvdbk_grid_dedupe_capacity(n)), leaf_keys has room for n keys,
 * counters is device uint32[2] zeroed by the caller: counters[0] = n_leaves,
 * counters[1] = voxels with a grid index >= n_grids or a coordinate outside
 * [-2^19, 2^19).  Grid indices < 0 skip a voxel. */
int64_t fvdbk_grid_dedupe_capacity(int64_t n);
int fvdbk_grid_leaf_dedupe(const int32_t* ijk, const int32_t* batch_idx, int64_t n, int32_t n_grids,
                           uint64_t* set, int64_t set_cap, uint64_t* leaf_keys, uint32_t* counters,
                           hipStream_t stream);

/* Step 2's sort of the unique leaf keys into tree order: keys_out = keys_in sorted
 * ascending (keys_out must not overlap keys_in), an LSD radix sort over the key bits a
 * batch of n_grids grids can use (bits 0..50 plus the grid index bits).  workspace is
 * device scratch of fvdbk_grid_sort_leaf_keys_workspace(n, n_grids) bytes.  Replaces
 * the sort inside nanovdb::tools::cuda::voxelsToGrid (src/fvdb/detail/ops/
 * BuildGridFromIjk.cu:66-81), which orders voxels by the same tree levels. */
size_t fvdbk_grid_sort_leaf_keys_workspace(int64_t n, int32_t n_grids);
int fvdbk_grid_sort_leaf_keys(const uint64_t* keys_in, uint64_t* keys_out, int64_t n, int32_t n_grids,
                              void* workspace, size_t workspace_bytes, hipStream_t stream);

int fvdbk_grid_hash_build(const uint64_t* leaf_keys, int64_t n_leaves, fvdbk_hslot* hash, int64_t hash_cap,
                          hipStream_t stream);

/* leaf_masks [n_leaves][8] is cleared by the call. */
int fvdbk_grid_leaf_masks(const int32_t* ijk, const int32_t* batch_idx, int64_t n, const fvdbk_hslot* hash,
                          int64_t hash_cap, uint64_t* leaf_masks, int64_t n_leaves, hipStream_t stream);

/* Coarsening by power-of-two factors (x, y, z), from the fine grid's leaves instead of
 * its voxels: the coarse grid holds floor(ijk / factor) of every fine voxel, as
 * dispatchCoarseIJKForFineGrid + dispatchCreateNanoGridFromIJK build it
 * (src/fvdb/detail/ops/BuildCoarseGridFromFine.cu:24-29, CoarseIjkForFineGrid.cu:34;
 * conv target grids with stride == kernel size, BuildGridForConv.cu:133-134).
 *   fvdbk_grid_coarsen_reps   one coarse voxel per fine leaf (rep_ijk [n][3], rep_bidx [n]):
 *                             the input of step 1 / the leaf de-duplication in place of
 *                             the voxels (every voxel of a fine leaf lands in that coarse leaf)
 *   fvdbk_grid_coarsen_masks  step 4 for the coarse leaves: their masks from the fine
 *                             masks (leaf_masks [n_leaves][8] is cleared by the call)
 * Factors must be powers of two in [1, 65536] (FVDBK_EINVAL otherwise). */
int fvdbk_grid_coarsen_reps(const uint64_t* fine_keys, int64_t n_fine_leaves, const int32_t factor[3],
                            int32_t* rep_ijk, int32_t* rep_bidx, hipStream_t stream);
int fvdbk_grid_coarsen_masks(const uint64_t* fine_keys, const uint64_t* fine_masks, int64_t n_fine_leaves,
                             const int32_t factor[3], const fvdbk_hslot* hash, int64_t hash_cap, uint64_t* leaf_masks,
                             int64_t n_leaves, hipStream_t stream);

int fvdbk_grid_leaf_counts(const uint64_t* leaf_masks, int64_t n_leaves, int32_t* counts, hipStream_t stream);

/* Steps 5-6 in one call: leaf_incl[l] = voxels in leaves 0..l and leaf_start[l] = voxels in
 * leaves 0..l-1 (device int64 [n_leaves] each), a device scan over the leaves' mask popcounts.
 * The workspace (device scratch) is fvdbk_grid_leaf_scan_workspace(n_leaves) bytes. */
size_t fvdbk_grid_leaf_scan_workspace(int64_t n_leaves);
int fvdbk_grid_leaf_scan(const uint64_t* leaf_masks, int64_t n_leaves, int64_t* leaf_start, int64_t* leaf_incl,
                         void* workspace, size_t workspace_bytes, hipStream_t stream);

/* Writes leaf_word_base [n_leaves][8], ijk_out [n_voxels][3], jidx_out [n_voxels]
 * (grid index per voxel; may be NULL) and grid_counts [n_grids] (voxels per grid; may
 * be NULL).  leaf_start is the exclusive scan of the leaf counts, n_voxels their sum. */
int fvdbk_grid_finalize(const uint64_t* leaf_keys, const uint64_t* leaf_masks, const int64_t* leaf_start,
                        int64_t n_leaves, int64_t n_voxels, int32_t* leaf_word_base, int32_t* ijk_out,
                        int32_t* jidx_out, int32_t* grid_counts, int32_t n_grids, hipStream_t stream);

/* Voxel offsets of the grids, offsets[b] for b = 0..n_grids (offsets[n_grids] = total
 * voxels), from the sorted leaf keys and the inclusive scan of the leaf counts; lets
 * the caller read the total and the per-grid counts in one device->host copy. */
int fvdbk_grid_offsets(const uint64_t* leaf_keys, const int64_t* leaf_incl, int64_t n_leaves, int32_t n_grids,
                       int64_t* offsets, hipStream_t stream);

/* Linear index of query coordinates (-1 when inactive).  Replaces
 * src/fvdb/detail/ops/IjkToIndex.cu (GridBatch.ijk_to_index, cumulative indices). */
int fvdbk_grid_ijk_to_index(const fvdbk_grid* grid, const int32_t* ijk, const int32_t* batch_idx, int64_t n,
                            int32_t* out_index, hipStream_t stream);
/* The same as int64 (the reference's index type), with the index taken within the query's own
 * grid, minus offsets[batch_idx[i]], when offsets (device int64 [n_grids+1]) is not NULL
 * (GridBatch.ijk_to_index with cumulative=False); -1 stays -1. */
int fvdbk_grid_ijk_to_index64(const fvdbk_grid* grid, const int32_t* ijk, const int32_t* batch_idx, int64_t n,
                              const int64_t* offsets, int64_t* out_index, hipStream_t stream);

/* World points -> voxel coordinates, ijk = floor(p * scale + translate + 0.5) in the
 * point precision (src/fvdb/detail/ops/BuildGridFromPoints.cu:28-44 with
 * VoxelCoordTransform.h:88-107).  transforms is [n_grids][6] (scale xyz, translate
 * xyz) as double; points are float32 (dtype FVDBK_F32) or float64. */
int fvdbk_points_to_ijk(fvdbk_dtype dtype, const void* points, const int32_t* batch_idx, int64_t n,
                        const double* transforms, int32_t* ijk_out, hipStream_t stream);

/* ---------------------------------------------------------------------------
 * Convolution target grid (src/fvdb/detail/ops/BuildGridForConv.cu:66-255 and
 * CoarseIjkForFineGrid.cu:18-111).  Emits n * K' candidate coordinates with their
 * grid index (-1 = not a target) that step 1 of the grid build consumes;
 * K' = 1 when stride is all ones or equals the kernel size (floor(ijk/stride)),
 * else K' = ks[0]*ks[1]*ks[2].  ks/stride are in the reference's (d,h,w) order.
 * ------------------------------------------------------------------------- */
int64_t fvdbk_conv_grid_candidates_per_voxel(const int32_t ks[3], const int32_t stride[3]);

int fvdbk_conv_grid_candidates(const int32_t* ijk, const int32_t* jidx, int64_t n, const int32_t ks[3],
                               const int32_t stride[3], int32_t* cand_ijk, int32_t* cand_jidx,
                               hipStream_t stream);

/* ---------------------------------------------------------------------------
 * Kernel map (replaces ops::dispatchConvolutionKernelMap<kCUDA>,
 * src/fvdb/detail/ops/convolution/pack_info/ConvolutionKernelMap.h:16-21 and
 * ConvolutionKernelMap.cu:104-138).
 *
 * inverse == 0: rows = target grid, probe = source grid; map[t][k] = index of
 *   source voxel t*stride + kernelStart + offset(k), or -1 (the reference's
 *   out_in_map / kernel map).
 * inverse == 1: rows = source grid, probe = target grid; map[s][k] = t such that
 *   the forward map has map[t][k] = s, or -1 (the reference's transposed map,
 *   IGEMMBitOperations.cu:92-105, computed by probing instead of scattering).
 * Rows [rows->n_voxels, n_rows_padded) are filled with -1.  ks/stride in (d,h,w).
 * ------------------------------------------------------------------------- */
int fvdbk_kmap_build(const fvdbk_grid* rows, const fvdbk_grid* probe, const int32_t ks[3],
                     const int32_t stride[3], int32_t inverse, int32_t* map, int64_t n_rows_padded,
                     hipStream_t stream);

/* Window map of the pooling operators (replaces the per-voxel window loops of
 * src/fvdb/detail/ops/DownsampleGridMaxPool.cu:17-63 and DownsampleGridAvgPool.cu:17-62):
 * map[c][k] = index of probe voxel  c*stride + (i, j, l)  or -1, where tap
 * k = (i*window[1] + j)*window[2] + l, i.e. x outermost, as the reference's loops.
 * window and stride are in (x, y, z) order. */
int fvdbk_kmap_window(const fvdbk_grid* rows, const fvdbk_grid* probe, const int32_t window[3],
                      const int32_t stride[3], int32_t* map, int64_t n_rows_padded, hipStream_t stream);

/* ---------------------------------------------------------------------------
 * Pooling over a window map (fvdbk_kmap_window; rows = coarse voxels).  Replaces
 * the forward voxel callbacks of src/fvdb/detail/ops/DownsampleGridMaxPool.cu:17-63
 * and DownsampleGridAvgPool.cu:17-62.  in [N_fine][channels], out [n_rows][channels].
 *   fvdbk_pool_max: out = max over the window's voxels (-inf when none); argmax (may be
 *     NULL) = index of the first voxel in tap order holding the maximum, -1 when none
 *     (the voxel that receives the gradient, DownsampleGridMaxPool.cu:65-114).
 *   fvdbk_pool_avg: out = T(sum) * T(scale), the sum in fp32 (fp64 for fp64) in tap
 *     order; scale = 1 / window volume in the reference.
 * ------------------------------------------------------------------------- */
int fvdbk_pool_max(fvdbk_dtype dtype, const void* in, int32_t channels, const int32_t* map, int64_t n_rows,
                   int32_t K, void* out, int32_t* argmax, hipStream_t stream);
int fvdbk_pool_avg(fvdbk_dtype dtype, const void* in, int32_t channels, const int32_t* map, int64_t n_rows,
                   int32_t K, double scale, void* out, hipStream_t stream);

/* Row gather with fill (replaces the per-voxel copies of src/fvdb/detail/ops/UpsampleGridNearest.cu:17-96
 * and ops/Inject.cu): out[i] = src[idx[i]] where 0 <= idx[i] < n_src, else fill[i], or zeros when
 * fill is NULL; rows of row_bytes bytes (any element type), n rows out.  out must not overlap src
 * or fill.  Copies in 16-byte units when rows and pointers allow, else 4-byte, else bytes. */
int fvdbk_gather_rows(const void* src, int64_t n_src, int64_t row_bytes, const int32_t* idx, int64_t n,
                      const void* fill, void* out, hipStream_t stream);

/* Sparse rows <-> dense frame (replaces the voxel callbacks of src/fvdb/detail/ops/ReadIntoDense.cu:17-91
 * and ops/ReadFromDense.cu:15-89 behind GridBatch.write_to_dense_* / read_from_dense_*): voxel i of
 * grid b = jidx[i] (0 when jidx is NULL) sits at p = ijk[i] - origins[b] (ijk device int32 [n][3],
 * origins device int32 [n_grids][3]).  When 0 <= p < size (the frame's extent on x, y, z) channel c
 * of its row pairs with the dense element strides[0] b + strides[1] p.x + strides[2] p.y +
 * strides[3] p.z + strides[4] c (element strides: a [B, X, Y, Z, C] and a [B, C, Z, Y, X] tensor
 * alike).  rows [n][channels] row-major; elements of elem_bytes bytes (1, 2, 4 or 8).
 * write != 0: the frame gets the rows of the voxels inside (other elements untouched: clear it
 * first); write == 0: the rows get the frame's values, zeros for voxels outside. */
int fvdbk_dense_rows(int write, const int32_t* ijk, const int32_t* jidx, int64_t n, int32_t channels,
                     const int32_t* origins, const int64_t size[3], const int64_t strides[5], int32_t elem_bytes,
                     void* dense, void* rows, hipStream_t stream);

/* ---------------------------------------------------------------------------
 * Gather-scatter neighbour map (replaces SparseConvPackInfo::buildGatherScatter,
 * src/fvdb/SparseConvPackInfo.cpp:61-88): pairs (source index, target row)
 * ordered by (k, target row), nbsizes[k] = pairs of offset k.
 *   fvdbk_nbmap_block_counts: counts[b][k] for row blocks of FVDBK_NBMAP_ROWS rows
 *   caller: exclusive scan of counts in (k, block) order -> block_offsets [nblk][K]
 *   fvdbk_nbmap_emit: writes pairs [P][2]
 * ------------------------------------------------------------------------- */
#define FVDBK_NBMAP_ROWS 1024
int fvdbk_nbmap_block_counts(const int32_t* map, int64_t n_rows, int32_t K, int32_t* counts, hipStream_t stream);
int fvdbk_nbmap_emit(const int32_t* map, int64_t n_rows, int32_t K, const int64_t* block_offsets, int32_t* pairs,
                     hipStream_t stream);

/* ---------------------------------------------------------------------------
 * Implicit-GEMM contraction (replaces dispatchSparseConvolutionImplicitGEMM /
 * ...Sorted, src/fvdb/detail/ops/convolution/backend/SparseConvolutionImplicitGEMM.h
 * and SparseConvolutionImplicitGEMMSorted.h, and the forward of
 * dispatchSparseConvolutionKernelMap, backend/SparseConvolutionKernelMap.h:13-21).
 *
 *   out[r][m] = sum_k sum_c  in[map[r][k]][c] * a[k][m][c]      (map entries < 0 skipped)
 *
 * in  [n_in][c_in], a [K][M][c_in], map [n_rows][K] (row stride K), out [n_rows][M],
 * all of dtype `dtype`; accumulation in fp32 (fp64 for FVDBK_F64).  Forward conv:
 * a = W^T per offset ([K][Cout][Cin]); dgrad: map = transposed map, in = grad_out,
 * a = W ([K][Cin][Cout]).  Each output row is produced by one wavefront in a fixed
 * (k ascending, channel chunk ascending) order, so a row's value does not depend on
 * which other rows share the launch.  flags: reserved, pass 0.
 * ------------------------------------------------------------------------- */
int fvdbk_conv_implicit_gemm(fvdbk_dtype dtype, const void* in, int64_t n_in, int32_t c_in, const void* a,
                             const int32_t* map, int64_t n_rows, int32_t K, int32_t M, void* out, uint32_t flags,
                             hipStream_t stream);

/* Same contraction with the output rows visited in `row_order` (a permutation of
 * [0, n_rows)), replacing dispatchSparseConvolutionImplicitGEMMSorted
 * (backend/SparseConvolutionImplicitGEMMSorted.h, reorder_loc of
 * SparseConvPackInfo.cpp:117-128).  Rows that share a workgroup share its skipped taps,
 * so ordering rows by their tap mask (fvdbk_kmap_row_keys, then a stable descending
 * sort) removes most work on missing neighbours.  Output values are identical to the
 * unsorted call. */
int fvdbk_conv_implicit_gemm_sorted(fvdbk_dtype dtype, const void* in, int64_t n_in, int32_t c_in, const void* a,
                                    const int32_t* map, const int32_t* row_order, int64_t n_rows, int32_t K,
                                    int32_t M, void* out, uint32_t flags, hipStream_t stream);

/* Contraction with a fused output epilogue: out = act(conv + bias), computed on the
 * fp32 accumulator and rounded once to the output type.  bias is [M] in the feature
 * dtype or NULL; flags: FVDBK_CONV_RELU applies max(., 0).  row_order may be NULL.
 * Replaces the separate bias add and activation that follow a convolution in
 * fvdb/nn/modules.py:400-403 (SparseConv3d bias) and the U-Net's ReLU modules. */
#define FVDBK_CONV_RELU 1u
int fvdbk_conv_implicit_gemm_fused(fvdbk_dtype dtype, const void* in, int64_t n_in, int32_t c_in, const void* a,
                                   const int32_t* map, const int32_t* row_order, int64_t n_rows, int32_t K, int32_t M,
                                   const void* bias, uint32_t flags, void* out, hipStream_t stream);

/* The contraction of fvdbk_conv_implicit_gemm_fused over the channel concatenation
 * [in0 | in1] of two [n_in][c_half] feature arrays (c_in = 2 c_half), read from both
 * without forming it.  Replaces fvdb.jcat([a, b], dim=1) followed by the convolution of
 * the concatenation (the decoder pattern of docs/tutorials/simple_unet.md:200-245), which
 * writes and re-reads the concatenated features.  Where the kernel that
 * fvdbk_conv_implicit_gemm_fused picks for the shape has a two-source form (16-bit 3x3x3
 * layers: c_half = 32 with M = 32 and no row order, c_half = 64 with a row order,
 * c_half = 128 with M > 32) the result is bit-identical to that call on the concatenated
 * input; otherwise the call launches nothing and returns FVDBK_EUNSUPPORTED, and the
 * caller concatenates. */
int fvdbk_conv_implicit_gemm_cat2(fvdbk_dtype dtype, const void* in0, const void* in1, int64_t n_in, int32_t c_half,
                                  const void* a, const int32_t* map, const int32_t* row_order, int64_t n_rows,
                                  int32_t K, int32_t M, const void* bias, uint32_t flags, void* out,
                                  hipStream_t stream);

/* keys[r] = bit k set when map[r][k] >= 0, for the first min(K, 31) taps (the
 * reference's bitmask of IGEMMBitOperations.cu:14-36 with one split). */
int fvdbk_kmap_row_keys(const int32_t* map, int64_t n_rows, int32_t K, int32_t* keys, hipStream_t stream);

/* Row order for fvdbk_conv_implicit_gemm_sorted that keeps locality: each window of
 * `window` consecutive rows (a power of two in [256, 8192]) is ordered by tap mask,
 * descending over the first min(K, 27) taps, ties in row order; windows stay in
 * place.  The windowed form of the reference's reorder_loc (SparseConvPackInfo.cpp:117-128,
 * one global sort): rows of a tile share most taps while their neighbours stay close. */
int fvdbk_kmap_row_order_windowed(const int32_t* map, int64_t n_rows, int32_t K, int32_t window, int32_t* order,
                                  hipStream_t stream);

/* ---------------------------------------------------------------------------
 * Weight gradient (replaces dispatchSparseConvolutionImplicitGEMMGrad{,Sorted},
 * backend/SparseConvolutionImplicitGEMMGrad.h, and the kernel-gradient half of
 * dispatchSparseConvolutionKernelMapGrad, backend/SparseConvolutionKernelMap.h:23-30).
 *
 *   gw[k][c][o] = sum over pairs p of offset k:  in[pairs[p][ci]][c] * gout[pairs[p][co]][o]
 *
 * with (ci, co) = (0, 1), or (1, 0) when transposed.  Pairs are the neighbour map
 * (ordered by k, then output row).  Split-K: offset k has n_k = chunk_start[k+1] -
 * chunk_start[k] partials (chunk_start[k] = first partial of offset k, chunk_start[K] =
 * total_chunks, a device int32 array; the callers take n_k = ceil(pairs of k / chunk)).
 * The offset's pairs are cut into rounds of 64 (32 for 16-bit dtypes with c_in and c_out
 * multiples of 128) and partial s sums rounds s, s + n_k, s + 2 n_k, ..., so the partials
 * of all offsets sweep the output rows together.  Each
 * partial's fp32 product goes to workspace [total_chunks][c_in][c_out] and a second pass
 * adds the partials of each offset with Neumaier-compensated fp32 summation (fp64 plain
 * sums for FVDBK_F64), in partial order, so the result is deterministic.  k_offsets [K+1] (device int64)
 * are the pair offsets of each k.  gw is fp32 ([K][c_in][c_out]) for the 16/32-bit
 * dtypes and fp64 for FVDBK_F64.
 * ------------------------------------------------------------------------- */
size_t fvdbk_conv_wgrad_workspace_size(fvdbk_dtype dtype, int64_t total_chunks, int32_t c_in, int32_t c_out);

int fvdbk_conv_wgrad(fvdbk_dtype dtype, const void* in, int32_t c_in, const void* gout, int32_t c_out,
                     const int32_t* pairs, const int64_t* k_offsets, const int32_t* chunk_start, int32_t K,
                     int32_t chunk, int64_t total_chunks, int32_t transposed, void* workspace,
                     size_t workspace_bytes, void* gw, hipStream_t stream);

/* ---------------------------------------------------------------------------
 * Group normalization over the voxels of each grid (replaces the per-grid loop of
 * fvdb/nn/modules.py:407-432, which runs torch.nn.GroupNorm on each grid's [1, C, n] view).
 *
 * x, y [n][C] in `dtype`, grids contiguous: rows offsets[b] .. offsets[b+1] (device int64
 * [n_grids+1]); jidx (device int32 [n], the grid of each row) may be NULL for one grid.
 * Channels form G groups of C/G.  mean / rstd are [n_grids][G] in fp32 (fp64 for FVDBK_F64):
 * the mean and 1/sqrt(biased variance + eps) over each (grid, group)'s rows x channels.
 *   y = (x - mean) * rstd * gamma[c] + beta[c]   (gamma / beta [C] in dtype, or NULL),
 * then max(., 0) with FVDBK_CONV_RELU.  n_blocks = fvdbk_group_norm_blocks(host counts):
 * the statistics are reduced over blocks of FVDBK_GN_ROWS rows of one grid, each merged in
 * a fixed order, so results are deterministic and a grid's rows do not depend on the other
 * grids of the batch.  The backward pass takes dy (and y when the forward applied the
 * ReLU, else NULL) and writes dx [n][C] and, when not NULL, dgamma / dbeta [C] in fp32
 * (fp64 for FVDBK_F64).  The workspace (fvdbk_group_norm_workspace_size) is device scratch.
 * ------------------------------------------------------------------------- */
#define FVDBK_GN_ROWS 1024
int64_t fvdbk_group_norm_blocks(const int64_t* counts, int32_t n_grids);
size_t fvdbk_group_norm_workspace_size(fvdbk_dtype dtype, int64_t n_blocks, int32_t n_grids, int32_t channels,
                                       int32_t groups);
int fvdbk_group_norm_forward(fvdbk_dtype dtype, const void* x, int64_t n, int32_t C, int32_t G, const int64_t* offsets,
                             const int32_t* jidx, int32_t n_grids, int64_t n_blocks, const void* gamma,
                             const void* beta, double eps, uint32_t flags, void* y, void* mean, void* rstd, void* ws,
                             size_t ws_bytes, hipStream_t stream);
int fvdbk_group_norm_backward(fvdbk_dtype dtype, const void* x, const void* dy, const void* y, int64_t n, int32_t C,
                              int32_t G, const int64_t* offsets, const int32_t* jidx, int32_t n_grids,
                              int64_t n_blocks, const void* mean, const void* rstd, const void* gamma, void* dx,
                              void* dgamma, void* dbeta, void* ws, size_t ws_bytes, hipStream_t stream);

/* ---------------------------------------------------------------------------
 * Weight gradient that allocates nothing.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  Same arguments and result as fvdbk_conv_wgrad
 * without the workspace: the split-K partials live in a device array owned by the library,
 * FVDBK_WGRAD_SCRATCH_BYTES per GPU, held from the library's load until the process ends.
 * Calls whose partials (fvdbk_conv_wgrad_workspace_size) exceed it fail with FVDBK_ENOSPC
 * before any launch.  The entry point calls no hipMalloc/hipFree.  Every call on a GPU shares
 * that GPU's array, so calls that overlap on one GPU (different streams or threads) race;
 * the array keeps the last call's contents, and no call depends on them.
 * ------------------------------------------------------------------------- */
#define FVDBK_WGRAD_SCRATCH_BYTES (64u << 20)
size_t fvdbk_conv_wgrad_scratch_bytes(void);
int fvdbk_conv_wgrad_noalloc(fvdbk_dtype dtype, const void* in, int32_t c_in, const void* gout, int32_t c_out,
                             const int32_t* pairs, const int64_t* k_offsets, const int32_t* chunk_start, int32_t K,
                             int32_t chunk, int64_t total_chunks, int32_t transposed, void* gw, hipStream_t stream);

#ifdef __cplusplus
}
#endif

#endif /* FVDBK_H */
