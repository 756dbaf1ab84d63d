// PyTorch-ROCm extension over the fvdbk C ABI.
//
// This is the tensor-level layer that the reference implements as the templated
// ops::dispatch*<DeviceTag> functions on torch::Tensor (e.g.
// src/fvdb/detail/ops/convolution/pack_info/ConvolutionKernelMap.h:16-21 and the
// backend/*.h headers): it validates shapes and dtypes, allocates outputs with the
// caching allocator, takes the current HIP stream, and calls the C ABI.  Status codes
// become the reference's error classes: FVDBK_EINVAL -> ValueError
// (TORCH_CHECK_VALUE), FVDBK_ETYPE -> TypeError (TORCH_CHECK_TYPE), others ->
// RuntimeError.  Host synchronisations happen only where a size must be known on the
// host (grid build, neighbour-map size), as in the reference.
#include <torch/extension.h>

#include <c10/hip/HIPGuard.h>
#include <c10/hip/HIPStream.h>

#include <string>
#include <vector>

#include "fvdbk.h"

namespace {

void check_rc(int rc, const char* what) {
    if (rc == FVDBK_OK) return;
    const std::string msg = std::string(what) + ": " + fvdbk_last_error();
    if (rc == FVDBK_EINVAL) TORCH_CHECK_VALUE(false, msg);
    if (rc == FVDBK_ETYPE) TORCH_CHECK_TYPE(false, msg);
    TORCH_CHECK(false, msg);
}

hipStream_t stream_of(const at::Tensor& t) { return c10::hip::getCurrentHIPStream(t.device().index()).stream(); }

void require_hip(const at::Tensor& t, const char* name) {
    TORCH_CHECK_VALUE(t.is_cuda(), name, " must be on a HIP device (fvdb-core_amd has no CPU path)");
}

fvdbk_dtype to_dtype(at::ScalarType s) {
    switch (s) {
    case at::kFloat: return FVDBK_F32;
    case at::kDouble: return FVDBK_F64;
    case at::kHalf: return FVDBK_F16;
    case at::kBFloat16: return FVDBK_BF16;
    default: TORCH_CHECK_TYPE(false, "unsupported floating point type ", s);
    }
    return FVDBK_F32;
}

// Grid tensors as held by the Python GridBatch: ijk [N,3] i32, leaf_keys [L] i64,
// leaf_masks [L,8] i64, word_base [L,8] i32, hash [cap,2] i64.
fvdbk_grid make_grid(const at::Tensor& ijk, const at::Tensor& keys, const at::Tensor& masks, const at::Tensor& wbase,
                     const at::Tensor& hash, int64_t n_grids) {
    fvdbk_grid g{};
    g.n_voxels = ijk.size(0);
    g.n_leaves = keys.size(0);
    g.hash_cap = hash.size(0);
    g.n_grids = (int32_t)n_grids;
    g.ijk = ijk.data_ptr<int32_t>();
    g.leaf_keys = reinterpret_cast<const uint64_t*>(keys.data_ptr<int64_t>());
    g.leaf_masks = reinterpret_cast<const uint64_t*>(masks.data_ptr<int64_t>());
    g.leaf_word_base = wbase.data_ptr<int32_t>();
    g.hash = reinterpret_cast<const fvdbk_hslot*>(hash.data_ptr<int64_t>());
    return g;
}

std::vector<int32_t> vec3(const std::vector<int64_t>& v, const char* name) {
    TORCH_CHECK_VALUE(v.size() == 3, name, " must have 3 entries");
    return {(int32_t)v[0], (int32_t)v[1], (int32_t)v[2]};
}

int64_t next_pow2(int64_t x) {
    int64_t p = 16;
    while (p < x) p <<= 1;
    return p;
}

}  // namespace

namespace {

// Steps after the leaf de-duplication, shared by the builds from voxels and from a fine
// grid's leaves: sort the unique keys (tree order), hash them, fill the masks
// (`fill_masks(hash, cap, masks, L)`), count, scan, grid offsets (one device->host
// read) and finalize.  Returns [ijk, jidx, grid offsets (host int64 [B+1]), grid
// offsets (device), leaf_keys, leaf_masks, word_base, hash].
template <typename FillMasks>
std::vector<at::Tensor> finish_grid(const at::Tensor& lkeys, int64_t L, int64_t n_grids, hipStream_t s,
                                    FillMasks fill_masks) {
    auto o32 = lkeys.options().dtype(at::kInt);
    auto o64 = lkeys.options().dtype(at::kLong);
    // tree order of the leaves: radix-sort only the unique keys
    at::Tensor ukeys = at::empty({L}, o64);
    const size_t ws_bytes = fvdbk_grid_sort_leaf_keys_workspace(L, (int32_t)n_grids);
    at::Tensor ws = at::empty({(int64_t)std::max<size_t>(ws_bytes, 1)}, lkeys.options().dtype(at::kByte));
    check_rc(fvdbk_grid_sort_leaf_keys(reinterpret_cast<const uint64_t*>(lkeys.data_ptr<int64_t>()),
                                       reinterpret_cast<uint64_t*>(ukeys.data_ptr<int64_t>()), L, (int32_t)n_grids,
                                       ws.data_ptr(), ws_bytes, s),
             "grid_sort_leaf_keys");
    const int64_t cap = next_pow2(2 * L + 2);
    at::Tensor hash = at::empty({cap, 2}, o64);
    check_rc(fvdbk_grid_hash_build(reinterpret_cast<const uint64_t*>(ukeys.data_ptr<int64_t>()), L,
                                   reinterpret_cast<fvdbk_hslot*>(hash.data_ptr<int64_t>()), cap, s),
             "grid_hash_build");
    at::Tensor masks = at::empty({L, 8}, o64);
    fill_masks(reinterpret_cast<const fvdbk_hslot*>(hash.data_ptr<int64_t>()), cap,
               reinterpret_cast<uint64_t*>(masks.data_ptr<int64_t>()), L);
    at::Tensor incl = at::empty({L}, o64), start = at::empty({L}, o64);
    if (L > 0) {
        const size_t scan_bytes = fvdbk_grid_leaf_scan_workspace(L);
        at::Tensor scan_ws = at::empty({(int64_t)std::max<size_t>(scan_bytes, 1)}, lkeys.options().dtype(at::kByte));
        check_rc(fvdbk_grid_leaf_scan(reinterpret_cast<const uint64_t*>(masks.data_ptr<int64_t>()), L,
                                      start.data_ptr<int64_t>(), incl.data_ptr<int64_t>(), scan_ws.data_ptr(),
                                      scan_bytes, s),
                 "grid_leaf_scan");
    }
    // grid offsets on the device; one read gives the total and the per-grid counts
    at::Tensor jo = at::empty({n_grids + 1}, o64);
    check_rc(fvdbk_grid_offsets(reinterpret_cast<const uint64_t*>(ukeys.data_ptr<int64_t>()), incl.data_ptr<int64_t>(),
                                L, (int32_t)n_grids, jo.data_ptr<int64_t>(), s),
             "grid_offsets");
    at::Tensor jo_host = jo.to(at::kCPU);
    const int64_t nv = jo_host.data_ptr<int64_t>()[n_grids];
    at::Tensor wbase = at::empty({L, 8}, o32);
    at::Tensor ijk_out = at::empty({nv, 3}, o32);
    at::Tensor jidx_out = at::empty({nv}, o32);
    check_rc(fvdbk_grid_finalize(reinterpret_cast<const uint64_t*>(ukeys.data_ptr<int64_t>()),
                                 reinterpret_cast<const uint64_t*>(masks.data_ptr<int64_t>()), start.data_ptr<int64_t>(),
                                 L, nv, wbase.data_ptr<int32_t>(), ijk_out.data_ptr<int32_t>(), jidx_out.data_ptr<int32_t>(),
                                 nullptr, (int32_t)n_grids, s),
             "grid_finalize");
    return {ijk_out, jidx_out, jo_host, jo, ukeys, masks, wbase, hash};
}

// Unique leaf keys (unsorted) of the given voxels; one device->host read (leaf count,
// out-of-range count).
std::pair<at::Tensor, int64_t> dedupe_leaves(const at::Tensor& ijk, const at::Tensor& jidx, bool has_jidx,
                                             int64_t n_grids, hipStream_t s) {
    const int64_t n = ijk.size(0);
    auto o32 = ijk.options().dtype(at::kInt);
    auto o64 = ijk.options().dtype(at::kLong);
    const int64_t set_cap = fvdbk_grid_dedupe_capacity(n);
    at::Tensor set = at::empty({set_cap}, o64);
    at::Tensor lkeys = at::empty({std::max<int64_t>(n, 1)}, o64);
    at::Tensor counters = at::zeros({2}, o32);
    check_rc(fvdbk_grid_leaf_dedupe(ijk.data_ptr<int32_t>(), has_jidx ? jidx.data_ptr<int32_t>() : nullptr, n,
                                    (int32_t)n_grids, reinterpret_cast<uint64_t*>(set.data_ptr<int64_t>()), set_cap,
                                    reinterpret_cast<uint64_t*>(lkeys.data_ptr<int64_t>()),
                                    reinterpret_cast<uint32_t*>(counters.data_ptr<int32_t>()), s),
             "grid_leaf_dedupe");
    at::Tensor hc = counters.to(at::kCPU);
    const int64_t L = hc.data_ptr<int32_t>()[0];
    const int nbad = hc.data_ptr<int32_t>()[1];
    TORCH_CHECK_VALUE(nbad == 0, nbad, " voxels have a grid index >= ", n_grids,
                      " or coordinates outside [-2^19, 2^19)");
    return {lkeys, L};
}

}  // namespace

// Build a grid batch index from voxel coordinates (duplicates merged, entries with a
// negative grid index skipped).  Two device->host reads: (leaf count, out-of-range
// count) and the grid offsets.  Returns [ijk, jidx, grid offsets (host int64 [B+1]),
// grid offsets (device), leaf_keys, leaf_masks, word_base, hash].
std::vector<at::Tensor> build_grid(at::Tensor ijk, at::Tensor jidx, int64_t n_grids) {
    require_hip(ijk, "ijk");
    TORCH_CHECK_VALUE(ijk.dim() == 2 && ijk.size(1) == 3, "ijk must have shape [N, 3]");
    TORCH_CHECK_TYPE(ijk.scalar_type() == at::kInt, "ijk must be int32");
    TORCH_CHECK_VALUE(n_grids >= 0 && n_grids <= 1024, "a grid batch holds at most 1024 grids");
    ijk = ijk.contiguous();
    const bool has_jidx = jidx.numel() > 0;
    if (has_jidx) {
        TORCH_CHECK_VALUE(jidx.numel() == ijk.size(0), "jidx must have one entry per voxel");
        jidx = jidx.to(at::kInt).contiguous();
    }
    c10::hip::HIPGuard guard(ijk.device().index());
    hipStream_t s = stream_of(ijk);
    const int64_t n = ijk.size(0);
    auto [lkeys, L] = dedupe_leaves(ijk, jidx, has_jidx, n_grids, s);
    return finish_grid(lkeys, L, n_grids, s, [&](const fvdbk_hslot* hash, int64_t cap, uint64_t* masks, int64_t nl) {
        check_rc(fvdbk_grid_leaf_masks(ijk.data_ptr<int32_t>(), has_jidx ? jidx.data_ptr<int32_t>() : nullptr, n,
                                       hash, cap, masks, nl, s),
                 "grid_leaf_masks");
    });
}

// Grid of floor(ijk / factor) over a built grid's voxels, for power-of-two factors
// (x, y, z), from the fine grid's leaf keys and masks: the de-duplication takes one
// voxel per fine leaf and the coarse masks are ORed from the fine masks.  Same outputs
// as build_grid over the coarsened voxels.
std::vector<at::Tensor> coarsen_grid(at::Tensor fine_keys, at::Tensor fine_masks, int64_t n_grids,
                                     std::vector<int64_t> factor) {
    require_hip(fine_keys, "leaf keys");
    auto f3 = vec3(factor, "coarsening_factor");
    TORCH_CHECK_VALUE(fine_keys.dim() == 1 && fine_masks.dim() == 2 && fine_masks.size(0) == fine_keys.size(0) &&
                          fine_masks.size(1) == 8,
                      "leaf keys [L] and masks [L, 8] expected");
    TORCH_CHECK_VALUE(n_grids >= 0 && n_grids <= 1024, "a grid batch holds at most 1024 grids");
    fine_keys = fine_keys.contiguous();
    fine_masks = fine_masks.contiguous();
    c10::hip::HIPGuard guard(fine_keys.device().index());
    hipStream_t s = stream_of(fine_keys);
    const int64_t Lf = fine_keys.size(0);
    at::Tensor rep = at::empty({Lf, 3}, fine_keys.options().dtype(at::kInt));
    at::Tensor rep_b = at::empty({Lf}, fine_keys.options().dtype(at::kInt));
    check_rc(fvdbk_grid_coarsen_reps(reinterpret_cast<const uint64_t*>(fine_keys.data_ptr<int64_t>()), Lf, f3.data(),
                                     rep.data_ptr<int32_t>(), rep_b.data_ptr<int32_t>(), s),
             "grid_coarsen_reps");
    auto [lkeys, L] = dedupe_leaves(rep, rep_b, true, n_grids, s);
    return finish_grid(lkeys, L, n_grids, s, [&](const fvdbk_hslot* hash, int64_t cap, uint64_t* masks, int64_t nl) {
        check_rc(fvdbk_grid_coarsen_masks(reinterpret_cast<const uint64_t*>(fine_keys.data_ptr<int64_t>()),
                                          reinterpret_cast<const uint64_t*>(fine_masks.data_ptr<int64_t>()), Lf,
                                          f3.data(), hash, cap, masks, nl, s),
                 "grid_coarsen_masks");
    });
}

at::Tensor ijk_to_index(at::Tensor g_ijk, at::Tensor g_keys, at::Tensor g_masks, at::Tensor g_wbase, at::Tensor g_hash,
                        int64_t n_grids, at::Tensor ijk, at::Tensor jidx) {
    require_hip(ijk, "ijk");
    TORCH_CHECK_VALUE(ijk.dim() == 2 && ijk.size(1) == 3, "ijk must have shape [N, 3]");
    ijk = ijk.to(at::kInt).contiguous();
    const bool has_jidx = jidx.numel() > 0;
    if (has_jidx) jidx = jidx.to(at::kInt).contiguous();
    c10::hip::HIPGuard guard(ijk.device().index());
    fvdbk_grid g = make_grid(g_ijk, g_keys, g_masks, g_wbase, g_hash, n_grids);
    at::Tensor out = at::empty({ijk.size(0)}, ijk.options().dtype(at::kInt));
    check_rc(fvdbk_grid_ijk_to_index(&g, ijk.data_ptr<int32_t>(), has_jidx ? jidx.data_ptr<int32_t>() : nullptr,
                                     ijk.size(0), out.data_ptr<int32_t>(), stream_of(ijk)),
             "ijk_to_index");
    return out;
}

// int64 indices; within each query's own grid when offsets (device int64 [B+1]) is not empty.
at::Tensor ijk_to_index64(at::Tensor g_ijk, at::Tensor g_keys, at::Tensor g_masks, at::Tensor g_wbase,
                          at::Tensor g_hash, int64_t n_grids, at::Tensor ijk, at::Tensor jidx, at::Tensor offsets) {
    require_hip(ijk, "ijk");
    TORCH_CHECK_VALUE(ijk.dim() == 2 && ijk.size(1) == 3, "ijk must have shape [N, 3]");
    ijk = ijk.to(at::kInt).contiguous();
    const bool has_jidx = jidx.numel() > 0, local = offsets.numel() > 0;
    if (has_jidx) jidx = jidx.to(at::kInt).contiguous();
    TORCH_CHECK_VALUE(!local || (offsets.numel() == n_grids + 1 && offsets.scalar_type() == at::kLong &&
                                 offsets.is_contiguous() && offsets.device() == ijk.device()),
                      "offsets must be contiguous device int64 [B+1]");
    c10::hip::HIPGuard guard(ijk.device().index());
    fvdbk_grid g = make_grid(g_ijk, g_keys, g_masks, g_wbase, g_hash, n_grids);
    at::Tensor out = at::empty({ijk.size(0)}, ijk.options().dtype(at::kLong));
    check_rc(fvdbk_grid_ijk_to_index64(&g, ijk.data_ptr<int32_t>(), has_jidx ? jidx.data_ptr<int32_t>() : nullptr,
                                       ijk.size(0), local ? offsets.data_ptr<int64_t>() : nullptr,
                                       out.data_ptr<int64_t>(), stream_of(ijk)),
             "ijk_to_index64");
    return out;
}

at::Tensor points_to_ijk(at::Tensor points, at::Tensor jidx, at::Tensor transforms) {
    require_hip(points, "points");
    TORCH_CHECK_VALUE(points.dim() == 2 && points.size(1) == 3, "points must have shape [N, 3]");
    TORCH_CHECK_TYPE(points.scalar_type() == at::kFloat || points.scalar_type() == at::kDouble,
                     "points must be float32 or float64");
    points = points.contiguous();
    const bool has_jidx = jidx.numel() > 0;
    if (has_jidx) jidx = jidx.to(at::kInt).contiguous();
    transforms = transforms.to(points.device(), at::kDouble).contiguous();
    c10::hip::HIPGuard guard(points.device().index());
    at::Tensor out = at::empty({points.size(0), 3}, points.options().dtype(at::kInt));
    check_rc(fvdbk_points_to_ijk(to_dtype(points.scalar_type()), points.data_ptr(),
                                 has_jidx ? jidx.data_ptr<int32_t>() : nullptr, points.size(0),
                                 transforms.data_ptr<double>(), out.data_ptr<int32_t>(), stream_of(points)),
             "points_to_ijk");
    return out;
}

std::vector<at::Tensor> conv_grid_candidates(at::Tensor ijk, at::Tensor jidx, std::vector<int64_t> ks,
                                             std::vector<int64_t> stride) {
    require_hip(ijk, "ijk");
    auto k3 = vec3(ks, "kernel_size");
    auto s3 = vec3(stride, "stride");
    ijk = ijk.contiguous();
    const bool has_jidx = jidx.numel() > 0;
    if (has_jidx) jidx = jidx.to(at::kInt).contiguous();
    c10::hip::HIPGuard guard(ijk.device().index());
    for (int d = 0; d < 3; ++d) {
        TORCH_CHECK_VALUE(k3[d] > 0, "kernel_size must be strictly positive.");
        TORCH_CHECK_VALUE(s3[d] > 0, "stride must be strictly positive.");
    }
    const int64_t per = fvdbk_conv_grid_candidates_per_voxel(k3.data(), s3.data());
    const int64_t n = ijk.size(0);
    at::Tensor cijk = at::empty({n * per, 3}, ijk.options());
    at::Tensor cj = at::empty({n * per}, ijk.options());
    check_rc(fvdbk_conv_grid_candidates(ijk.data_ptr<int32_t>(), has_jidx ? jidx.data_ptr<int32_t>() : nullptr, n,
                                        k3.data(), s3.data(), cijk.data_ptr<int32_t>(), cj.data_ptr<int32_t>(),
                                        stream_of(ijk)),
             "conv_grid_candidates");
    return {cijk, cj};
}

at::Tensor kernel_map(std::vector<at::Tensor> rows, std::vector<at::Tensor> probe, int64_t n_grids,
                      std::vector<int64_t> ks, std::vector<int64_t> stride, bool inverse, int64_t n_rows_padded) {
    TORCH_CHECK_VALUE(rows.size() == 5 && probe.size() == 5, "grid tensors: [ijk, keys, masks, word_base, hash]");
    require_hip(rows[0], "grid");
    auto k3 = vec3(ks, "kernel_size");
    auto s3 = vec3(stride, "stride");
    c10::hip::HIPGuard guard(rows[0].device().index());
    fvdbk_grid gr = make_grid(rows[0], rows[1], rows[2], rows[3], rows[4], n_grids);
    fvdbk_grid gp = make_grid(probe[0], probe[1], probe[2], probe[3], probe[4], n_grids);
    const int64_t K = (int64_t)k3[0] * k3[1] * k3[2];
    at::Tensor map = at::empty({n_rows_padded, K}, rows[0].options().dtype(at::kInt));
    check_rc(fvdbk_kmap_build(&gr, &gp, k3.data(), s3.data(), inverse ? 1 : 0, map.data_ptr<int32_t>(),
                              n_rows_padded, stream_of(rows[0])),
             "kmap_build");
    return map;
}

// Pooling window map: rows = coarse grid, probe = fine grid, window/stride in (x, y, z).
at::Tensor window_map(std::vector<at::Tensor> rows, std::vector<at::Tensor> probe, int64_t n_grids,
                      std::vector<int64_t> window, std::vector<int64_t> stride, int64_t n_rows_padded) {
    TORCH_CHECK_VALUE(rows.size() == 5 && probe.size() == 5, "grid tensors: [ijk, keys, masks, word_base, hash]");
    require_hip(rows[0], "grid");
    auto w3 = vec3(window, "pool_factor");
    auto s3 = vec3(stride, "stride");
    c10::hip::HIPGuard guard(rows[0].device().index());
    fvdbk_grid gr = make_grid(rows[0], rows[1], rows[2], rows[3], rows[4], n_grids);
    fvdbk_grid gp = make_grid(probe[0], probe[1], probe[2], probe[3], probe[4], n_grids);
    const int64_t K = (int64_t)w3[0] * w3[1] * w3[2];
    at::Tensor map = at::empty({n_rows_padded, K}, rows[0].options().dtype(at::kInt));
    check_rc(fvdbk_kmap_window(&gr, &gp, w3.data(), s3.data(), map.data_ptr<int32_t>(), n_rows_padded,
                               stream_of(rows[0])),
             "kmap_window");
    return map;
}

// Max pooling over a window map: (out [n_rows, C], argmax [n_rows, C] int32).
std::vector<at::Tensor> pool_max(at::Tensor in, at::Tensor map, int64_t n_rows, bool need_argmax) {
    require_hip(in, "features");
    TORCH_CHECK_VALUE(in.dim() == 2, "features must have shape [N, C]");
    TORCH_CHECK_VALUE(map.dim() == 2 && map.scalar_type() == at::kInt && map.size(0) >= n_rows,
                      "map must be int32 [>= rows, K]");
    in = in.contiguous();
    map = map.contiguous();
    c10::hip::HIPGuard guard(in.device().index());
    at::Tensor out = at::empty({n_rows, in.size(1)}, in.options());
    // the argmax (the voxel that receives the gradient) only when a backward will need it
    at::Tensor arg = at::empty({need_argmax ? n_rows : 0, in.size(1)}, in.options().dtype(at::kInt));
    if (n_rows > 0 && in.size(1) > 0)
        check_rc(fvdbk_pool_max(to_dtype(in.scalar_type()), in.data_ptr(), (int32_t)in.size(1), map.data_ptr<int32_t>(),
                                n_rows, (int32_t)map.size(1), out.data_ptr(),
                                need_argmax ? arg.data_ptr<int32_t>() : nullptr, stream_of(in)),
                 "pool_max");
    return {out, arg};
}

// Average pooling over a window map: T(sum) * T(scale).
at::Tensor pool_avg(at::Tensor in, at::Tensor map, int64_t n_rows, double scale) {
    require_hip(in, "features");
    TORCH_CHECK_VALUE(in.dim() == 2, "features must have shape [N, C]");
    TORCH_CHECK_VALUE(map.dim() == 2 && map.scalar_type() == at::kInt && map.size(0) >= n_rows,
                      "map must be int32 [>= rows, K]");
    in = in.contiguous();
    map = map.contiguous();
    c10::hip::HIPGuard guard(in.device().index());
    at::Tensor out = at::empty({n_rows, in.size(1)}, in.options());
    if (n_rows > 0 && in.size(1) > 0)
        check_rc(fvdbk_pool_avg(to_dtype(in.scalar_type()), in.data_ptr(), (int32_t)in.size(1), map.data_ptr<int32_t>(),
                                n_rows, (int32_t)map.size(1), scale, out.data_ptr(), stream_of(in)),
                 "pool_avg");
    return out;
}

// out[i] = src[idx[i]] where idx[i] is a row of src, else fill[i] (zeros without fill); out has
// idx.size(0) rows shaped like src's.
at::Tensor gather_rows(at::Tensor src, at::Tensor idx, c10::optional<at::Tensor> fill) {
    require_hip(src, "source rows");
    TORCH_CHECK_VALUE(src.dim() >= 1, "source rows must have at least one dimension");
    TORCH_CHECK_VALUE(idx.dim() == 1 && idx.scalar_type() == at::kInt && idx.device() == src.device(),
                      "indices must be int32 [n] on the rows' device");
    src = src.contiguous();
    idx = idx.contiguous();
    std::vector<int64_t> shape(src.sizes().begin(), src.sizes().end());
    shape[0] = idx.size(0);
    if (fill.has_value()) {
        TORCH_CHECK_VALUE(fill->sizes() == at::IntArrayRef(shape) && fill->scalar_type() == src.scalar_type() &&
                              fill->device() == src.device(),
                          "fill must have the output's shape, dtype and device");
        *fill = fill->contiguous();
    }
    c10::hip::HIPGuard guard(src.device().index());
    at::Tensor out = at::empty(shape, src.options());
    int64_t row_bytes = src.element_size();
    for (int64_t d = 1; d < src.dim(); ++d) row_bytes *= src.size(d);
    if (idx.size(0) > 0 && row_bytes > 0)
        check_rc(fvdbk_gather_rows(src.data_ptr(), src.size(0), row_bytes, idx.data_ptr<int32_t>(), idx.size(0),
                                   fill.has_value() ? fill->data_ptr() : nullptr, out.data_ptr(), stream_of(src)),
                 "gather_rows");
    return out;
}

// Rows <-> a dense frame: frame is a [B, X, Y, Z, C] view (any strides) of the dense tensor,
// rows [N, C] contiguous; write: frame <- rows of the voxels inside, else rows <- frame (zeros
// outside).  ijk int32 [N, 3], jidx int32 [N] or empty (one grid), origins int32 [B, 3].
void dense_rows(bool write, at::Tensor ijk, at::Tensor jidx, at::Tensor origins, at::Tensor frame, at::Tensor rows) {
    require_hip(frame, "dense frame");
    TORCH_CHECK_VALUE(frame.dim() == 5, "dense frame must be a [B, X, Y, Z, C] view");
    TORCH_CHECK_VALUE(rows.dim() == 2 && rows.is_contiguous() && rows.size(1) == frame.size(4) &&
                          rows.scalar_type() == frame.scalar_type() && rows.device() == frame.device(),
                      "rows must be contiguous [N, C] with the frame's dtype and device");
    TORCH_CHECK_VALUE(ijk.dim() == 2 && ijk.size(1) == 3 && ijk.size(0) == rows.size(0) &&
                          ijk.scalar_type() == at::kInt && ijk.is_contiguous() && ijk.device() == frame.device(),
                      "ijk must be contiguous int32 [N, 3] on the frame's device");
    TORCH_CHECK_VALUE(origins.dim() == 2 && origins.size(0) == frame.size(0) && origins.size(1) == 3 &&
                          origins.scalar_type() == at::kInt && origins.is_contiguous() &&
                          origins.device() == frame.device(),
                      "origins must be contiguous int32 [B, 3] on the frame's device");
    const bool batched = jidx.numel() > 0;
    TORCH_CHECK_VALUE(!batched || (jidx.dim() == 1 && jidx.size(0) == rows.size(0) && jidx.scalar_type() == at::kInt &&
                                   jidx.is_contiguous() && jidx.device() == frame.device()),
                      "jidx must be contiguous int32 [N] on the frame's device, or empty for one grid");
    TORCH_CHECK_VALUE(batched || frame.size(0) == 1, "a batch of several grids needs jidx");
    if (rows.size(0) == 0 || rows.size(1) == 0) return;
    c10::hip::HIPGuard guard(frame.device().index());
    const int64_t size[3] = {frame.size(1), frame.size(2), frame.size(3)};
    const int64_t strides[5] = {frame.stride(0), frame.stride(1), frame.stride(2), frame.stride(3), frame.stride(4)};
    check_rc(fvdbk_dense_rows(write ? 1 : 0, ijk.data_ptr<int32_t>(), batched ? jidx.data_ptr<int32_t>() : nullptr,
                              rows.size(0), (int32_t)rows.size(1), origins.data_ptr<int32_t>(), size, strides,
                              (int32_t)frame.element_size(), frame.data_ptr(), rows.data_ptr(), stream_of(frame)),
             "dense_rows");
}

// (pairs [P,2] (source index, target row) ordered by (k, row), nbsizes [K]) from a map.
std::vector<at::Tensor> neighbor_map(at::Tensor map, int64_t n_rows) {
    require_hip(map, "map");
    TORCH_CHECK_VALUE(map.dim() == 2 && map.scalar_type() == at::kInt, "map must be int32 [rows, K]");
    TORCH_CHECK_VALUE(map.size(0) >= n_rows, "map has fewer rows than requested");
    map = map.contiguous();
    c10::hip::HIPGuard guard(map.device().index());
    hipStream_t s = stream_of(map);
    const int64_t K = map.size(1);
    const int64_t nblk = (n_rows + FVDBK_NBMAP_ROWS - 1) / FVDBK_NBMAP_ROWS;
    auto o32 = map.options();
    if (nblk == 0) return {at::empty({0, 2}, o32), at::zeros({K}, o32)};
    at::Tensor counts = at::empty({nblk, K}, o32);
    check_rc(fvdbk_nbmap_block_counts(map.data_ptr<int32_t>(), n_rows, (int32_t)K, counts.data_ptr<int32_t>(), s),
             "nbmap_block_counts");
    at::Tensor nbsizes = counts.sum(0, false, at::kInt);
    // exclusive scan in (k, block) order, laid back out as [block][k]
    at::Tensor flat = counts.t().contiguous().view({-1}).to(at::kLong);
    at::Tensor incl = at::cumsum(flat, 0);
    const int64_t P = incl[incl.size(0) - 1].item<int64_t>();
    at::Tensor offs = (incl - flat).view({K, nblk}).t().contiguous();
    at::Tensor pairs = at::empty({P, 2}, o32);
    check_rc(fvdbk_nbmap_emit(map.data_ptr<int32_t>(), n_rows, (int32_t)K, offs.data_ptr<int64_t>(),
                              pairs.data_ptr<int32_t>(), s),
             "nbmap_emit");
    return {pairs, nbsizes};
}

// Row order by tap mask: stable descending sort of the per-row tap bitmask.
// Rows ordered by descending tap mask within consecutive windows of `window` rows
// (0 = one window over all rows); ties keep row order.  Windows keep the rows of a
// tile spatially close, so the gathers of a tile stay in a small region of the input.
at::Tensor row_order(at::Tensor map, int64_t n_rows, int64_t window) {
    require_hip(map, "map");
    TORCH_CHECK_VALUE(map.dim() == 2 && map.scalar_type() == at::kInt && map.size(0) >= n_rows,
                      "map must be int32 [>= rows, K]");
    TORCH_CHECK_VALUE(window >= 0, "window must be >= 0");
    map = map.contiguous();
    c10::hip::HIPGuard guard(map.device().index());
    if (window > 0) {
        // per-window order, one kernel
        at::Tensor order = at::empty({n_rows}, map.options());
        check_rc(fvdbk_kmap_row_order_windowed(map.data_ptr<int32_t>(), n_rows, (int32_t)map.size(1), (int32_t)window,
                                               order.data_ptr<int32_t>(), stream_of(map)),
                 "kmap_row_order_windowed");
        return order;
    }
    // one global order: stable descending sort of the 31-bit tap-mask keys
    at::Tensor keys = at::empty({n_rows}, map.options());
    check_rc(fvdbk_kmap_row_keys(map.data_ptr<int32_t>(), n_rows, (int32_t)map.size(1), keys.data_ptr<int32_t>(),
                                 stream_of(map)),
             "kmap_row_keys");
    at::Tensor order = std::get<1>(at::sort(keys, /*stable=*/true, /*dim=*/0, /*descending=*/true));
    return order.to(at::kInt).contiguous();
}

// out[r][m] = sum_k sum_c in[map[r][k]][c] * a[k][m][c]
at::Tensor implicit_gemm(at::Tensor in, at::Tensor a, at::Tensor map, int64_t n_rows,
                         c10::optional<at::Tensor> order, c10::optional<at::Tensor> bias, bool relu) {
    require_hip(in, "features");
    TORCH_CHECK_VALUE(in.dim() == 2, "features must have shape [N, C]");
    TORCH_CHECK_VALUE(a.dim() == 3, "weights must have shape [K, M, C]");
    TORCH_CHECK_TYPE(in.is_floating_point(), "features must have a floating point type");
    TORCH_CHECK_TYPE(a.scalar_type() == in.scalar_type(), "features and weights must have the same dtype (",
                     in.scalar_type(), " vs ", a.scalar_type(), ")");
    TORCH_CHECK_VALUE(a.size(2) == in.size(1), "weights expect ", a.size(2), " input channels, features have ",
                      in.size(1));
    TORCH_CHECK_VALUE(map.dim() == 2 && map.size(1) == a.size(0) && map.size(0) >= n_rows,
                      "kernel map must have shape [>= rows, K]");
    TORCH_CHECK_VALUE(map.scalar_type() == at::kInt, "kernel map must be int32");
    in = in.contiguous();
    a = a.contiguous();
    map = map.contiguous();
    c10::hip::HIPGuard guard(in.device().index());
    at::Tensor out = at::empty({n_rows, a.size(1)}, in.options());
    const bool has_bias = bias.has_value() && bias->defined();
    if (has_bias || relu) {
        // fused epilogue: act(conv + bias) rounded once
        at::Tensor b;
        if (has_bias) {
            b = bias->to(in.scalar_type()).contiguous();
            TORCH_CHECK_VALUE(b.dim() == 1 && b.size(0) == a.size(1), "bias must have shape [", a.size(1), "]");
            TORCH_CHECK_VALUE(b.device() == in.device(), "bias must be on the features' device");
        }
        const int32_t* ord = nullptr;
        at::Tensor o;
        if (order.has_value() && order->defined()) {
            o = order->contiguous();
            TORCH_CHECK_VALUE(o.scalar_type() == at::kInt && o.numel() == n_rows, "row order must be int32 [rows]");
            ord = o.data_ptr<int32_t>();
        }
        check_rc(fvdbk_conv_implicit_gemm_fused(to_dtype(in.scalar_type()), in.data_ptr(), in.size(0),
                                                (int32_t)in.size(1), a.data_ptr(), map.data_ptr<int32_t>(), ord, n_rows,
                                                (int32_t)a.size(0), (int32_t)a.size(1),
                                                has_bias ? b.data_ptr() : nullptr, relu ? FVDBK_CONV_RELU : 0u,
                                                out.data_ptr(), stream_of(in)),
                 "conv_implicit_gemm_fused");
        return out;
    }
    if (order.has_value() && order->defined()) {
        at::Tensor o = order->contiguous();
        TORCH_CHECK_VALUE(o.scalar_type() == at::kInt && o.numel() == n_rows, "row order must be int32 [rows]");
        check_rc(fvdbk_conv_implicit_gemm_sorted(to_dtype(in.scalar_type()), in.data_ptr(), in.size(0),
                                                 (int32_t)in.size(1), a.data_ptr(), map.data_ptr<int32_t>(),
                                                 o.data_ptr<int32_t>(), n_rows, (int32_t)a.size(0), (int32_t)a.size(1),
                                                 out.data_ptr(), 0u, stream_of(in)),
                 "conv_implicit_gemm_sorted");
        return out;
    }
    check_rc(fvdbk_conv_implicit_gemm(to_dtype(in.scalar_type()), in.data_ptr(), in.size(0), (int32_t)in.size(1),
                                      a.data_ptr(), map.data_ptr<int32_t>(), n_rows, (int32_t)a.size(0),
                                      (int32_t)a.size(1), out.data_ptr(), 0u, stream_of(in)),
             "conv_implicit_gemm");
    return out;
}

// implicit_gemm over the channel concatenation [in0 | in1] without forming it where the
// library has a two-source kernel for the shape (bit-identical to the concatenated call);
// otherwise the features are concatenated and the one-source path runs.
at::Tensor implicit_gemm_cat2(at::Tensor in0, at::Tensor in1, at::Tensor a, at::Tensor map, int64_t n_rows,
                              c10::optional<at::Tensor> order, c10::optional<at::Tensor> bias, bool relu) {
    require_hip(in0, "features");
    require_hip(in1, "features");
    TORCH_CHECK_VALUE(in0.dim() == 2 && in1.dim() == 2 && in0.size(0) == in1.size(0) && in0.size(1) == in1.size(1),
                      "the two feature arrays must have the same shape [N, C]");
    TORCH_CHECK_TYPE(in0.scalar_type() == in1.scalar_type(), "the two feature arrays must have the same dtype");
    TORCH_CHECK_VALUE(in0.device() == in1.device(), "the two feature arrays must be on one device");
    const bool fits = a.dim() == 3 && a.scalar_type() == in0.scalar_type() && a.size(2) == 2 * in0.size(1) &&
                      map.dim() == 2 && map.scalar_type() == at::kInt && map.size(1) == a.size(0) &&
                      map.size(0) >= n_rows && (!bias.has_value() || !bias->defined() ||
                                                (bias->dim() == 1 && bias->size(0) == a.size(1) &&
                                                 bias->device() == in0.device()));
    if (fits) {
        in0 = in0.contiguous();
        in1 = in1.contiguous();
        a = a.contiguous();
        map = map.contiguous();
        c10::hip::HIPGuard guard(in0.device().index());
        const int32_t* ord = nullptr;
        at::Tensor o;
        if (order.has_value() && order->defined()) {
            o = order->contiguous();
            TORCH_CHECK_VALUE(o.scalar_type() == at::kInt && o.numel() == n_rows, "row order must be int32 [rows]");
            ord = o.data_ptr<int32_t>();
        }
        at::Tensor b;
        if (bias.has_value() && bias->defined()) b = bias->to(in0.scalar_type()).contiguous();
        at::Tensor out = at::empty({n_rows, a.size(1)}, in0.options());
        const int rc = fvdbk_conv_implicit_gemm_cat2(
            to_dtype(in0.scalar_type()), in0.data_ptr(), in1.data_ptr(), in0.size(0), (int32_t)in0.size(1),
            a.data_ptr(), map.data_ptr<int32_t>(), ord, n_rows, (int32_t)a.size(0), (int32_t)a.size(1),
            b.defined() ? b.data_ptr() : nullptr, relu ? FVDBK_CONV_RELU : 0u, out.data_ptr(), stream_of(in0));
        if (rc != FVDBK_EUNSUPPORTED) {
            check_rc(rc, "conv_implicit_gemm_cat2");
            return out;
        }
    }
    return implicit_gemm(at::cat({in0, in1}, 1), a, map, n_rows, order, bias, relu);
}

// gw [K, Cin, Cout] (fp32, fp64 for fp64 inputs) from the neighbour-map pairs.
// nbsizes_cpu: int64 CPU tensor [K] (pair counts per offset, known on the host).
at::Tensor wgrad(at::Tensor in, at::Tensor gout, at::Tensor pairs, at::Tensor nbsizes_cpu, bool transposed,
                 int64_t chunk) {
    require_hip(in, "features");
    TORCH_CHECK_TYPE(in.scalar_type() == gout.scalar_type(), "features and gradients must have the same dtype");
    TORCH_CHECK_VALUE(in.dim() == 2 && gout.dim() == 2, "features and gradients must be 2-D");
    TORCH_CHECK_VALUE(pairs.dim() == 2 && pairs.size(1) == 2 && pairs.scalar_type() == at::kInt,
                      "pairs must be int32 [P, 2]");
    TORCH_CHECK_VALUE(!nbsizes_cpu.is_cuda(), "nbsizes must be a host tensor");
    in = in.contiguous();
    gout = gout.contiguous();
    pairs = pairs.contiguous();
    c10::hip::HIPGuard guard(in.device().index());
    at::Tensor sizes = nbsizes_cpu.to(at::kLong).contiguous();
    const int64_t K = sizes.size(0);
    const int64_t* sz = sizes.data_ptr<int64_t>();
    std::vector<int64_t> koff(K + 1, 0);
    std::vector<int32_t> cstart(K + 1, 0);
    for (int64_t k = 0; k < K; ++k) {
        koff[k + 1] = koff[k] + sz[k];
        cstart[k + 1] = cstart[k] + (int32_t)((sz[k] + chunk - 1) / chunk);
    }
    TORCH_CHECK_VALUE(koff[K] == pairs.size(0), "nbsizes do not add up to the number of pairs");
    const int64_t total = cstart[K];
    const fvdbk_dtype dt = to_dtype(in.scalar_type());
    auto dev = in.options();
    at::Tensor koff_t = at::from_blob(koff.data(), {K + 1}, at::kLong).to(in.device());
    at::Tensor cs_t = at::from_blob(cstart.data(), {K + 1}, at::kInt).to(in.device());
    const size_t ws_bytes = fvdbk_conv_wgrad_workspace_size(dt, total, (int32_t)in.size(1), (int32_t)gout.size(1));
    at::Tensor ws = at::empty({(int64_t)((ws_bytes + 7) / 8)}, dev.dtype(at::kLong));
    at::Tensor gw = at::empty({K, in.size(1), gout.size(1)}, dev.dtype(dt == FVDBK_F64 ? at::kDouble : at::kFloat));
    check_rc(fvdbk_conv_wgrad(dt, in.data_ptr(), (int32_t)in.size(1), gout.data_ptr(), (int32_t)gout.size(1),
                              pairs.data_ptr<int32_t>(), koff_t.data_ptr<int64_t>(), cs_t.data_ptr<int32_t>(),
                              (int32_t)K, (int32_t)chunk, total, transposed ? 1 : 0, ws.data_ptr(), ws_bytes,
                              gw.data_ptr(), stream_of(in)),
             "conv_wgrad");
    return gw;
}

// The same with the split-K tables already on the device (k_offsets int64 [K+1], chunk_start
// int32 [K+1], `total` chunks): no host work or host-to-device copy per call.
at::Tensor wgrad_tabled(at::Tensor in, at::Tensor gout, at::Tensor pairs, at::Tensor koff_t, at::Tensor cs_t,
                        int64_t total, bool transposed, int64_t chunk) {
    require_hip(in, "features");
    TORCH_CHECK_TYPE(in.scalar_type() == gout.scalar_type(), "features and gradients must have the same dtype");
    TORCH_CHECK_VALUE(in.dim() == 2 && gout.dim() == 2, "features and gradients must be 2-D");
    TORCH_CHECK_VALUE(pairs.dim() == 2 && pairs.size(1) == 2 && pairs.scalar_type() == at::kInt,
                      "pairs must be int32 [P, 2]");
    TORCH_CHECK_VALUE(koff_t.is_cuda() && koff_t.scalar_type() == at::kLong && cs_t.is_cuda() &&
                          cs_t.scalar_type() == at::kInt && koff_t.numel() == cs_t.numel() && koff_t.numel() >= 2,
                      "k_offsets must be device int64 [K+1] and chunk_start device int32 [K+1]");
    in = in.contiguous();
    gout = gout.contiguous();
    pairs = pairs.contiguous();
    c10::hip::HIPGuard guard(in.device().index());
    const int64_t K = koff_t.numel() - 1;
    const fvdbk_dtype dt = to_dtype(in.scalar_type());
    auto dev = in.options();
    const size_t ws_bytes = fvdbk_conv_wgrad_workspace_size(dt, total, (int32_t)in.size(1), (int32_t)gout.size(1));
    at::Tensor ws = at::empty({(int64_t)((ws_bytes + 7) / 8)}, dev.dtype(at::kLong));
    at::Tensor gw = at::empty({K, in.size(1), gout.size(1)}, dev.dtype(dt == FVDBK_F64 ? at::kDouble : at::kFloat));
    check_rc(fvdbk_conv_wgrad(dt, in.data_ptr(), (int32_t)in.size(1), gout.data_ptr(), (int32_t)gout.size(1),
                              pairs.data_ptr<int32_t>(), koff_t.data_ptr<int64_t>(), cs_t.data_ptr<int32_t>(),
                              (int32_t)K, (int32_t)chunk, total, transposed ? 1 : 0, ws.data_ptr(), ws_bytes,
                              gw.data_ptr(), stream_of(in)),
             "conv_wgrad");
    return gw;
}

// GroupNorm over each grid's voxels: (y, mean, rstd).  offsets: device int64 [B+1]; counts: host
// int64 [B] (the same grid sizes, known on the host); jidx: device int32 [n] or empty for B = 1.
std::vector<at::Tensor> group_norm_forward(at::Tensor x, at::Tensor offsets, at::Tensor counts, at::Tensor jidx,
                                           int64_t G, c10::optional<at::Tensor> gamma, c10::optional<at::Tensor> beta,
                                           double eps, bool relu) {
    require_hip(x, "features");
    TORCH_CHECK_VALUE(x.dim() == 2, "features must have shape [N, C]");
    TORCH_CHECK_TYPE(x.is_floating_point(), "features must have a floating point type");
    x = x.contiguous();
    counts = counts.to(at::kLong).contiguous();
    const int64_t B = counts.numel(), C = x.size(1), n = x.size(0);
    TORCH_CHECK_VALUE(offsets.numel() == B + 1 && offsets.is_cuda(), "offsets must be device int64 [B+1]");
    c10::hip::HIPGuard guard(x.device().index());
    const fvdbk_dtype dt = to_dtype(x.scalar_type());
    auto acc = x.options().dtype(dt == FVDBK_F64 ? at::kDouble : at::kFloat);
    at::Tensor g, b;
    if (gamma.has_value() && gamma->defined()) g = gamma->to(x.scalar_type()).contiguous();
    if (beta.has_value() && beta->defined()) b = beta->to(x.scalar_type()).contiguous();
    const int64_t nb = fvdbk_group_norm_blocks(counts.data_ptr<int64_t>(), (int32_t)B);
    const size_t wsb = fvdbk_group_norm_workspace_size(dt, nb, (int32_t)B, (int32_t)C, (int32_t)G);
    at::Tensor ws = at::empty({(int64_t)((wsb + 7) / 8)}, x.options().dtype(at::kLong));
    at::Tensor y = at::empty_like(x), mean = at::empty({B, G}, acc), rstd = at::empty({B, G}, acc);
    at::Tensor jx = jidx.numel() ? jidx.to(at::kInt).contiguous() : jidx;
    at::Tensor off = offsets.to(at::kLong).contiguous();
    check_rc(fvdbk_group_norm_forward(dt, x.data_ptr(), n, (int32_t)C, (int32_t)G, off.data_ptr<int64_t>(),
                                      jx.numel() ? jx.data_ptr<int32_t>() : nullptr, (int32_t)B, nb,
                                      g.defined() ? g.data_ptr() : nullptr, b.defined() ? b.data_ptr() : nullptr, eps,
                                      relu ? FVDBK_CONV_RELU : 0u, y.data_ptr(), mean.data_ptr(), rstd.data_ptr(),
                                      ws.data_ptr(), wsb, stream_of(x)),
             "group_norm_forward");
    return {y, mean, rstd};
}

// (dx, dgamma, dbeta); y is the forward output when the forward applied a ReLU (else empty).
std::vector<at::Tensor> group_norm_backward(at::Tensor x, at::Tensor dy, at::Tensor y, at::Tensor offsets,
                                            at::Tensor counts, at::Tensor jidx, int64_t G, at::Tensor mean,
                                            at::Tensor rstd, c10::optional<at::Tensor> gamma) {
    require_hip(x, "features");
    x = x.contiguous();
    dy = dy.to(x.scalar_type()).contiguous();
    counts = counts.to(at::kLong).contiguous();
    const int64_t B = counts.numel(), C = x.size(1), n = x.size(0);
    c10::hip::HIPGuard guard(x.device().index());
    const fvdbk_dtype dt = to_dtype(x.scalar_type());
    auto acc = x.options().dtype(dt == FVDBK_F64 ? at::kDouble : at::kFloat);
    at::Tensor g;
    if (gamma.has_value() && gamma->defined()) g = gamma->to(x.scalar_type()).contiguous();
    const int64_t nb = fvdbk_group_norm_blocks(counts.data_ptr<int64_t>(), (int32_t)B);
    const size_t wsb = fvdbk_group_norm_workspace_size(dt, nb, (int32_t)B, (int32_t)C, (int32_t)G);
    at::Tensor ws = at::empty({(int64_t)((wsb + 7) / 8)}, x.options().dtype(at::kLong));
    at::Tensor dx = at::empty_like(x), dg = at::empty({C}, acc), db = at::empty({C}, acc);
    at::Tensor jx = jidx.numel() ? jidx.to(at::kInt).contiguous() : jidx;
    at::Tensor off = offsets.to(at::kLong).contiguous();
    at::Tensor yy = y.numel() ? y.contiguous() : y;
    check_rc(fvdbk_group_norm_backward(dt, x.data_ptr(), dy.data_ptr(), yy.numel() ? yy.data_ptr() : nullptr, n,
                                       (int32_t)C, (int32_t)G, off.data_ptr<int64_t>(),
                                       jx.numel() ? jx.data_ptr<int32_t>() : nullptr, (int32_t)B, nb,
                                       mean.contiguous().data_ptr(), rstd.contiguous().data_ptr(),
                                       g.defined() ? g.data_ptr() : nullptr, dx.data_ptr(), dg.data_ptr(),
                                       db.data_ptr(), ws.data_ptr(), wsb, stream_of(x)),
             "group_norm_backward");
    return {dx, dg, db};
}

std::string version() { return fvdbk_version(); }

PYBIND11_MODULE(TORCH_EXTENSION_NAME, m) {
    m.doc() = "fvdb-core_amd tensor layer over the fvdbk C ABI";
    // Every entry point runs without the GIL (arguments are converted before, results after),
    // so a host thread blocked in one of them -- a grid build waits for its sizes -- does not
    // stall the other Python threads (the serving loop's ingest worker).
    const auto nogil = py::call_guard<py::gil_scoped_release>();
    m.def("version", &version);
    m.def("build_grid", &build_grid, nogil);
    m.def("coarsen_grid", &coarsen_grid, nogil);
    m.def("ijk_to_index", &ijk_to_index, nogil);
    m.def("ijk_to_index64", &ijk_to_index64, nogil);
    m.def("points_to_ijk", &points_to_ijk, nogil);
    m.def("conv_grid_candidates", &conv_grid_candidates, nogil);
    m.def("kernel_map", &kernel_map, nogil);
    m.def("neighbor_map", &neighbor_map, nogil);
    m.def("window_map", &window_map, nogil);
    m.def("pool_max", &pool_max, nogil, py::arg("in"), py::arg("map"), py::arg("n_rows"),
          py::arg("need_argmax") = true);
    m.def("pool_avg", &pool_avg, nogil);
    m.def("gather_rows", &gather_rows, nogil, py::arg("src"), py::arg("idx"), py::arg("fill") = py::none());
    m.def("dense_rows", &dense_rows, nogil, py::arg("write"), py::arg("ijk"), py::arg("jidx"), py::arg("origins"),
          py::arg("frame"), py::arg("rows"));
    m.def("implicit_gemm", &implicit_gemm, nogil, pybind11::arg("in"), pybind11::arg("a"), pybind11::arg("map"),
          pybind11::arg("n_rows"), pybind11::arg("order") = pybind11::none(), pybind11::arg("bias") = pybind11::none(),
          pybind11::arg("relu") = false);
    m.def("implicit_gemm_cat2", &implicit_gemm_cat2, nogil, py::arg("in0"), py::arg("in1"), py::arg("a"),
          py::arg("map"), py::arg("n_rows"), py::arg("order") = py::none(), py::arg("bias") = py::none(),
          py::arg("relu") = false);
    m.def("row_order", &row_order, nogil, py::arg("map"), py::arg("n_rows"), py::arg("window") = 0);
    m.def("wgrad", &wgrad, nogil);
    m.def("wgrad_tabled", &wgrad_tabled, nogil);
    m.def("group_norm_forward", &group_norm_forward, nogil);
    m.def("group_norm_backward", &group_norm_backward, nogil);
}
