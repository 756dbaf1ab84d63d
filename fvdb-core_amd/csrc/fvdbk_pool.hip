// Max and average pooling over a window map (fvdbk_kmap_window) for gfx950.
// One thread per (coarse voxel, channel), channels fastest, so the reads of a
// gathered fine row are coalesced; taps are visited in ascending order.
#include <hip/hip_runtime.h>

#include <math.h>

#include "fvdbk_common.h"

namespace fvdbk {

template <typename T> struct PoolAcc { using type = float; };
template <> struct PoolAcc<double> { using type = double; };

// out = max over the window's existing voxels (-inf when none); argmax = the first
// voxel holding it in tap order (strict comparison), -1 when none
// (src/fvdb/detail/ops/DownsampleGridMaxPool.cu:17-63 forward, :65-114 argmax).
template <typename T>
__global__ void k_pool_max(const T* __restrict__ in, int C, const int32_t* __restrict__ map, int64_t n_rows, int K,
                           T* __restrict__ out, int32_t* __restrict__ argmax) {
    using A = typename PoolAcc<T>::type;
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= n_rows * C) return;
    const int64_t r = idx / C;
    const int ch = (int)(idx - r * C);
    A best = -INFINITY;
    int32_t arg = -1;
    for (int k = 0; k < K; ++k) {
        const int32_t s = map[r * K + k];
        if (s < 0) continue;
        const A v = (A)in[(int64_t)s * C + ch];
        if (v > best) {
            best = v;
            arg = s;
        }
    }
    out[idx] = (T)best;
    if (argmax) argmax[idx] = arg;
}

// out = T(sum) * T(scale), the sum accumulated in fp32 (fp64 for fp64) in tap order
// (src/fvdb/detail/ops/DownsampleGridAvgPool.cu:17-62).
template <typename T>
__global__ void k_pool_avg(const T* __restrict__ in, int C, const int32_t* __restrict__ map, int64_t n_rows, int K,
                           T scale, T* __restrict__ out) {
    using A = typename PoolAcc<T>::type;
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= n_rows * C) return;
    const int64_t r = idx / C;
    const int ch = (int)(idx - r * C);
    A sum = 0;
    for (int k = 0; k < K; ++k) {
        const int32_t s = map[r * K + k];
        if (s >= 0) sum += (A)in[(int64_t)s * C + ch];
    }
    out[idx] = (T)((A)(T)sum * (A)scale);
}

template <typename T>
static int pool_launch(bool is_max, const void* in, int C, const int32_t* map, int64_t n_rows, int K, double scale,
                       void* out, int32_t* argmax, hipStream_t s) {
    const int64_t n = n_rows * C;
    if (n == 0) return FVDBK_OK;
    if (is_max) {
        hipLaunchKernelGGL(k_pool_max<T>, dim3(grid_1d(n, 256)), dim3(256), 0, s, (const T*)in, C, map, n_rows, K,
                           (T*)out, argmax);
        FVDBK_LAUNCH_CHECK("k_pool_max");
    } else {
        hipLaunchKernelGGL(k_pool_avg<T>, dim3(grid_1d(n, 256)), dim3(256), 0, s, (const T*)in, C, map, n_rows, K,
                           (T)scale, (T*)out);
        FVDBK_LAUNCH_CHECK("k_pool_avg");
    }
    return FVDBK_OK;
}

static int pool_entry(bool is_max, fvdbk_dtype dtype, const void* in, int32_t C, const int32_t* map, int64_t n_rows,
                      int32_t K, double scale, void* out, int32_t* argmax, hipStream_t s) {
    FVDBK_REQUIRE(C > 0 && K > 0 && n_rows >= 0, FVDBK_EINVAL, "pooling: bad sizes");
    FVDBK_REQUIRE(n_rows * (int64_t)C < ((int64_t)1 << 40), FVDBK_EINVAL, "pooling: output too large");
    FVDBK_REQUIRE(n_rows == 0 || (in && map && out), FVDBK_EINVAL, "pooling: null argument");
    switch (dtype) {
    case FVDBK_F32: return pool_launch<float>(is_max, in, C, map, n_rows, K, scale, out, argmax, s);
    case FVDBK_F64: return pool_launch<double>(is_max, in, C, map, n_rows, K, scale, out, argmax, s);
    case FVDBK_F16: return pool_launch<_Float16>(is_max, in, C, map, n_rows, K, scale, out, argmax, s);
    case FVDBK_BF16: return pool_launch<__bf16>(is_max, in, C, map, n_rows, K, scale, out, argmax, s);
    default: return set_error(FVDBK_ETYPE, "unsupported dtype %d", (int)dtype);
    }
}

}  // namespace fvdbk

using namespace fvdbk;

extern "C" {

int fvdbk_pool_max(fvdbk_dtype dtype, const void* in, int32_t channels, const int32_t* map, int64_t n_rows, int32_t K,
                   void* out, int32_t* argmax, hipStream_t stream) {
    return pool_entry(true, dtype, in, channels, map, n_rows, K, 0.0, out, argmax, stream);
}

int fvdbk_pool_avg(fvdbk_dtype dtype, const void* in, int32_t channels, const int32_t* map, int64_t n_rows, int32_t K,
                   double scale, void* out, hipStream_t stream) {
    return pool_entry(false, dtype, in, channels, map, n_rows, K, scale, out, nullptr, stream);
}

}  // extern "C"
