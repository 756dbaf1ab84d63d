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

// Vector form of k_pool_max / k_pool_avg: one thread per (row, 16-byte channel segment), for
// rows of whole 16-byte segments on 16-byte aligned arrays.  Each existing window voxel's
// segment is one 16-byte load; taps, comparisons and arithmetic as the scalar kernels, so
// the results are identical.
template <typename T>
union Seg16 {
    uint4 v;
    T e[16 / sizeof(T)];
};

template <typename T, bool MAX>
__global__ __launch_bounds__(256) void k_pool_vec(const T* __restrict__ in, int segs, const int32_t* __restrict__ map,
                                                  uint32_t total, int K, T scale, T* __restrict__ out,
                                                  int32_t* __restrict__ argmax) {
    using A = typename PoolAcc<T>::type;
    constexpr int E = 16 / (int)sizeof(T);
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= total) return;
    const uint32_t r = t / (uint32_t)segs, sg = t - r * (uint32_t)segs;
    const int64_t row_elems = (int64_t)segs * E;
    A acc[E];
    int32_t arg[E];
#pragma unroll
    for (int e = 0; e < E; ++e) {
        acc[e] = MAX ? (A)-INFINITY : (A)0;
        arg[e] = -1;
    }
    const int32_t* __restrict__ m = map + (int64_t)r * K;
    for (int k = 0; k < K; ++k) {
        const int32_t s = m[k];
        if (s < 0) continue;
        Seg16<T> v;
        v.v = *reinterpret_cast<const uint4*>(in + (int64_t)s * row_elems + (int64_t)sg * E);
#pragma unroll
        for (int e = 0; e < E; ++e) {
            const A x = (A)v.e[e];
            if constexpr (MAX) {
                if (x > acc[e]) {
                    acc[e] = x;
                    arg[e] = s;
                }
            } else {
                acc[e] += x;
            }
        }
    }
    Seg16<T> o;
#pragma unroll
    for (int e = 0; e < E; ++e) o.e[e] = MAX ? (T)acc[e] : (T)((A)(T)acc[e] * (A)scale);
    const int64_t base = (int64_t)r * row_elems + (int64_t)sg * E;
    *reinterpret_cast<uint4*>(out + base) = o.v;
    if (MAX && argmax != nullptr) {
#pragma unroll
        for (int q = 0; q < E / 4; ++q)
            *reinterpret_cast<int4*>(argmax + base + 4 * q) = make_int4(arg[4 * q], arg[4 * q + 1], arg[4 * q + 2],
                                                                        arg[4 * q + 3]);
        if constexpr (E == 2) *reinterpret_cast<int2*>(argmax + base) = make_int2(arg[0], arg[1]);
    }
}

template <typename T>
static int pool_launch(bool is_max, const void* in, int C, const int32_t* map, int64_t n_rows, int K, double scale,
                       void* out, int32_t* argmax, hipStream_t s) {
    const int64_t n = n_rows * C;
    if (n == 0) return FVDBK_OK;
    constexpr int E = 16 / (int)sizeof(T);
    const bool vec = C % E == 0 && n / E < ((int64_t)1 << 31) &&
                     (((uintptr_t)in | (uintptr_t)out | (uintptr_t)argmax) & 15u) == 0;
    if (vec) {
        const uint32_t total = (uint32_t)(n / E);
        if (is_max)
            hipLaunchKernelGGL((k_pool_vec<T, true>), dim3(grid_1d(total, 256)), dim3(256), 0, s, (const T*)in, C / E,
                               map, total, K, (T)0, (T*)out, argmax);
        else
            hipLaunchKernelGGL((k_pool_vec<T, false>), dim3(grid_1d(total, 256)), dim3(256), 0, s, (const T*)in,
                               C / E, map, total, K, (T)scale, (T*)out, nullptr);
        FVDBK_LAUNCH_CHECK("k_pool_vec");
        return FVDBK_OK;
    }
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

// out[i] = src[idx[i]] where 0 <= idx[i] < n_src, else fill[i] (zeros when fill is null),
// one thread per (row, U-sized unit of the row).
template <typename U>
__global__ __launch_bounds__(256) void k_gather_rows(const U* __restrict__ src, int64_t n_src, int64_t units,
                                                     const int32_t* __restrict__ idx, int64_t total,
                                                     const U* __restrict__ fill, U* __restrict__ out) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= total) return;
    const int64_t r = t / units, u = t - r * units;
    const int32_t s = idx[r];
    U v;
    if (s >= 0 && s < n_src) v = src[(int64_t)s * units + u];
    else if (fill != nullptr) v = fill[t];
    else v = U{};
    out[t] = v;
}

// Rows <-> dense frame.  Voxel i of grid b = jidx[i] (0 without jidx) sits at p = ijk[i] -
// origin[b]; inside when 0 <= p < (dx, dy, dz).  Unit c of its row (one element, or 16 bytes
// of channels when channels are contiguous in both) <-> dense[b sb + p.x sx + p.y sy + p.z sz +
// c sc], strides in units.  WRITE stores the rows of voxels inside into the frame; otherwise the
// rows read the frame, zeros for voxels outside.  CFAST: consecutive threads take consecutive
// units of a row (channels contiguous in the frame), else consecutive voxels of one unit.
struct DenseFrame {
    int64_t sb, sx, sy, sz, sc;
    int64_t dx, dy, dz;
};

template <typename U, bool WRITE, bool CFAST>
__global__ __launch_bounds__(256) void k_dense_rows(const int32_t* __restrict__ ijk, const int32_t* __restrict__ jidx,
                                                    int64_t n, int units, const int32_t* __restrict__ origin,
                                                    DenseFrame f, U* __restrict__ dense, U* __restrict__ rows) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n * units) return;
    int64_t i;
    int c;
    if (CFAST) {
        i = t / units;
        c = (int)(t - i * units);
    } else {
        c = (int)(t / n);
        i = t - (int64_t)c * n;
    }
    const int b = jidx ? jidx[i] : 0;
    const int64_t x = (int64_t)ijk[3 * i] - origin[3 * b], y = (int64_t)ijk[3 * i + 1] - origin[3 * b + 1],
                  z = (int64_t)ijk[3 * i + 2] - origin[3 * b + 2];
    const bool inside = x >= 0 && x < f.dx && y >= 0 && y < f.dy && z >= 0 && z < f.dz;
    const int64_t d = b * f.sb + x * f.sx + y * f.sy + z * f.sz + c * f.sc;
    U* row = rows + i * units + c;
    if constexpr (WRITE) {
        if (inside) dense[d] = *row;
    } else {
        *row = inside ? dense[d] : U{};
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

int fvdbk_gather_rows(const void* src, int64_t n_src, int64_t row_bytes, const int32_t* idx, int64_t n,
                      const void* fill, void* out, hipStream_t stream) {
    FVDBK_REQUIRE(row_bytes > 0 && n >= 0 && n_src >= 0, FVDBK_EINVAL, "row gather: bad sizes");
    if (n == 0) return FVDBK_OK;
    FVDBK_REQUIRE(idx && out && (src || n_src == 0), FVDBK_EINVAL, "row gather: null argument");
    const uintptr_t al = (uintptr_t)src | (uintptr_t)out | (uintptr_t)fill;
    auto go = [&](auto unit) {
        using U = decltype(unit);
        const int64_t units = row_bytes / (int64_t)sizeof(U), total = n * units;
        hipLaunchKernelGGL((k_gather_rows<U>), dim3(grid_1d(total, 256)), dim3(256), 0, stream, (const U*)src, n_src,
                           units, idx, total, (const U*)fill, (U*)out);
    };
    if (row_bytes % 16 == 0 && (al & 15u) == 0) go(uint4{});
    else if (row_bytes % 4 == 0 && (al & 3u) == 0) go(uint32_t{});
    else go(uint8_t{});
    FVDBK_LAUNCH_CHECK("k_gather_rows");
    return FVDBK_OK;
}

int fvdbk_dense_rows(int write, const int32_t* ijk, const int32_t* jidx, int64_t n, int32_t channels,
                     const int32_t* origins, const int64_t size[3], const int64_t strides[5], int32_t elem_bytes,
                     void* dense, void* rows, hipStream_t stream) {
    FVDBK_REQUIRE(n >= 0 && channels > 0 && size && strides, FVDBK_EINVAL, "dense rows: bad sizes");
    FVDBK_REQUIRE(elem_bytes == 1 || elem_bytes == 2 || elem_bytes == 4 || elem_bytes == 8, FVDBK_EINVAL,
                  "dense rows: element size %d not 1, 2, 4 or 8", elem_bytes);
    FVDBK_REQUIRE(size[0] >= 0 && size[1] >= 0 && size[2] >= 0, FVDBK_EINVAL, "dense rows: negative frame size");
    if (n == 0) return FVDBK_OK;
    FVDBK_REQUIRE(ijk && origins && dense && rows, FVDBK_EINVAL, "dense rows: null argument");
    const int64_t eb = elem_bytes;
    const bool cfast = strides[4] == 1;
    const bool vec = cfast && (channels * eb) % 16 == 0 && (((uintptr_t)dense | (uintptr_t)rows) & 15u) == 0 &&
                     (strides[0] * eb) % 16 == 0 && (strides[1] * eb) % 16 == 0 && (strides[2] * eb) % 16 == 0 &&
                     (strides[3] * eb) % 16 == 0;
    auto go = [&](auto unit) {
        using U = decltype(unit);
        const int64_t per = (int64_t)sizeof(U) / eb;
        const DenseFrame f{strides[0] / per, strides[1] / per, strides[2] / per, strides[3] / per,
                           cfast ? 1 : strides[4], size[0], size[1], size[2]};
        const int units = (int)(channels / per);
        const dim3 grid(grid_1d(n * units, 256));
        U* d = (U*)dense;
        U* r = (U*)rows;
        if (write && cfast)
            hipLaunchKernelGGL((k_dense_rows<U, true, true>), grid, dim3(256), 0, stream, ijk, jidx, n, units, origins,
                               f, d, r);
        else if (write)
            hipLaunchKernelGGL((k_dense_rows<U, true, false>), grid, dim3(256), 0, stream, ijk, jidx, n, units,
                               origins, f, d, r);
        else if (cfast)
            hipLaunchKernelGGL((k_dense_rows<U, false, true>), grid, dim3(256), 0, stream, ijk, jidx, n, units,
                               origins, f, d, r);
        else
            hipLaunchKernelGGL((k_dense_rows<U, false, false>), grid, dim3(256), 0, stream, ijk, jidx, n, units,
                               origins, f, d, r);
    };
    if (vec) go(uint4{});
    else if (eb == 1) go(uint8_t{});
    else if (eb == 2) go(uint16_t{});
    else if (eb == 4) go(uint32_t{});
    else go(uint64_t{});
    FVDBK_LAUNCH_CHECK("k_dense_rows");
    return FVDBK_OK;
}

}  // extern "C"
