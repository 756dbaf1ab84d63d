// Group normalization over the voxels of each grid of a jagged feature batch, for gfx950.
//
// x [N][C] row-major, grids contiguous (rows offsets[b] .. offsets[b+1]); channels split into
// G groups of C/G.  Statistics per (grid, group) over its rows x channels, the biased
// variance, as torch.nn.GroupNorm computes them on each grid's [1, C, n] view
// (fvdb/nn/modules.py:407-432 in the reference runs exactly that per grid).
//
// Every reduction has a fixed order, so results are deterministic and independent of the
// other grids in the batch:
//   1. a partial pass over blocks of GN_ROWS rows of one grid: per channel, Welford (count,
//      mean, M2) of the block's rows, threads of a channel merged in thread order;
//   2. one wavefront per (grid, group) merges its blocks' partials (lanes stride over
//      (block, channel) items, then a fixed butterfly) with Chan's pairwise update;
//   3. y = (x - mean) * rstd * gamma + beta, optionally ReLU, one thread per element.
// The backward pass reduces sum(dy) and sum(dy * xhat) per (grid, channel) the same way and
// forms dx = rstd * (gamma dy - mean_g(gamma dy) - xhat mean_g(gamma dy xhat)).
// Accumulation is fp32 (fp64 for fp64 data).
#include <hip/hip_runtime.h>

#include <math.h>

#include "fvdbk_common.h"

namespace fvdbk {

constexpr int GN_THREADS = 256;

template <typename T> struct NormAcc { using type = float; };
template <> struct NormAcc<double> { using type = double; };

template <typename A> struct Welford {
    A n, mean, m2;
};

template <typename A>
__device__ inline Welford<A> wmerge(Welford<A> a, Welford<A> b) {
    if (b.n == (A)0) return a;
    if (a.n == (A)0) return b;
    const A n = a.n + b.n;
    const A d = b.mean - a.mean;
    const A f = b.n / n;
    return Welford<A>{n, a.mean + d * f, a.m2 + b.m2 + d * d * a.n * f};
}

// prefix[b] = number of GN_ROWS-row blocks of grids 0..b-1; prefix[B] = all blocks.
__global__ void k_gn_block_prefix(const int64_t* __restrict__ off, int B, int32_t* __restrict__ prefix) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    int32_t acc = 0;
    for (int b = 0; b < B; ++b) {
        prefix[b] = acc;
        acc += (int32_t)((off[b + 1] - off[b] + FVDBK_GN_ROWS - 1) / FVDBK_GN_ROWS);
    }
    prefix[B] = acc;
}

// The grid of block `blk`: the last b with prefix[b] <= blk (grids without rows have no blocks).
__device__ inline int gn_grid_of(const int32_t* __restrict__ prefix, int B, int blk) {
    int lo = 0, hi = B;  // prefix[lo] <= blk < prefix[hi]
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (prefix[mid] <= blk) lo = mid;
        else hi = mid;
    }
    return lo;
}

// Rows [r0, r1) of block blk.
__device__ inline void gn_block_rows(const int64_t* __restrict__ off, const int32_t* __restrict__ prefix, int B,
                                     int blk, int& b, int64_t& r0, int64_t& r1) {
    b = gn_grid_of(prefix, B, blk);
    r0 = off[b] + (int64_t)(blk - prefix[b]) * FVDBK_GN_ROWS;
    r1 = min(r0 + (int64_t)FVDBK_GN_ROWS, off[b + 1]);
}

// Pass 1: per block and channel, Welford over the block's rows -> part[blk][C].
template <typename T>
__global__ __launch_bounds__(GN_THREADS) void k_gn_stats_partial(const T* __restrict__ x, int C,
                                                                 const int64_t* __restrict__ off,
                                                                 const int32_t* __restrict__ prefix, int B,
                                                                 Welford<typename NormAcc<T>::type>* __restrict__ part) {
    using A = typename NormAcc<T>::type;
    __shared__ Welford<A> red[GN_THREADS];
    const int blk = blockIdx.x, tid = threadIdx.x;
    int b;
    int64_t r0, r1;
    gn_block_rows(off, prefix, B, blk, b, r0, r1);
    for (int c0 = 0; c0 < C; c0 += GN_THREADS) {
        const int cc = min(GN_THREADS, C - c0);
        const int R = GN_THREADS / cc;  // rows visited in parallel
        Welford<A> w{(A)0, (A)0, (A)0};
        if (tid < R * cc) {
            const int c = c0 + tid % cc;
            for (int64_t r = r0 + tid / cc; r < r1; r += R) {
                const A v = (A)x[r * C + c];
                w.n += (A)1;
                const A d = v - w.mean;
                w.mean += d / w.n;
                w.m2 += d * (v - w.mean);
            }
        }
        red[tid] = w;
        __syncthreads();
        if (tid < cc) {
            Welford<A> acc = red[tid];
            for (int q = 1; q < R; ++q) acc = wmerge(acc, red[q * cc + tid]);
            part[(int64_t)blk * C + c0 + tid] = acc;
        }
        __syncthreads();
    }
}

template <typename A>
__device__ inline A shfl_xor(A v, int m) {
    return __shfl_xor(v, m, 64);
}

// Pass 2: one wavefront per (grid, group): mean and 1/sqrt(var + eps).
template <typename A>
__global__ void k_gn_stats_final(const Welford<A>* __restrict__ part, int C, int G, const int32_t* __restrict__ prefix,
                                 int B, double eps, A* __restrict__ mean, A* __restrict__ rstd) {
    const int wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6, lane = threadIdx.x & 63;
    if (wave >= B * G) return;
    const int b = wave / G, g = wave % G, Cg = C / G;
    const int nb = prefix[b + 1] - prefix[b];
    const int items = nb * Cg;
    Welford<A> w{(A)0, (A)0, (A)0};
    for (int it = lane; it < items; it += 64) {
        const int j = it / Cg, c = g * Cg + it % Cg;
        w = wmerge(w, part[(int64_t)(prefix[b] + j) * C + c]);
    }
#pragma unroll
    for (int m = 1; m < 64; m <<= 1) {
        Welford<A> o{shfl_xor(w.n, m), shfl_xor(w.mean, m), shfl_xor(w.m2, m)};
        // the lower lane's value first, so both lanes of a pair hold the same merge
        w = (lane & m) ? wmerge(o, w) : wmerge(w, o);
    }
    if (lane == 0) {
        const A var = w.n > (A)0 ? w.m2 / w.n : (A)0;
        mean[wave] = w.mean;
        rstd[wave] = (A)1 / sqrt(var + (A)eps);
    }
}

// Pass 3: y = (x - mean) * rstd * gamma + beta (gamma/beta may be NULL), ReLU when asked.
template <typename T>
__global__ void k_gn_apply(const T* __restrict__ x, int64_t n, int C, int G, const int32_t* __restrict__ jidx,
                           const typename NormAcc<T>::type* __restrict__ mean,
                           const typename NormAcc<T>::type* __restrict__ rstd, const T* __restrict__ gamma,
                           const T* __restrict__ beta, int relu, T* __restrict__ y) {
    using A = typename NormAcc<T>::type;
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= n * C) return;
    const int64_t r = idx / C;
    const int c = (int)(idx - r * C);
    const int b = jidx ? jidx[r] : 0;
    const int s = b * G + c / (C / G);
    A v = ((A)x[idx] - mean[s]) * rstd[s];
    if (gamma) v = v * (A)gamma[c];
    if (beta) v = v + (A)beta[c];
    if (relu && !(v > (A)0) && v == v) v = (A)0;
    y[idx] = (T)v;
}

// Backward pass 1: per block and channel, sum(dy) and sum(dy * xhat) (dy masked by y > 0
// after a fused ReLU) -> part[blk][C][2].
template <typename T>
__global__ __launch_bounds__(GN_THREADS) void k_gn_bwd_partial(const T* __restrict__ x, const T* __restrict__ dy,
                                                               const T* __restrict__ y, int C, int G,
                                                               const int64_t* __restrict__ off,
                                                               const int32_t* __restrict__ prefix, int B,
                                                               const typename NormAcc<T>::type* __restrict__ mean,
                                                               const typename NormAcc<T>::type* __restrict__ rstd,
                                                               typename NormAcc<T>::type* __restrict__ part) {
    using A = typename NormAcc<T>::type;
    __shared__ A red[2][GN_THREADS];
    const int blk = blockIdx.x, tid = threadIdx.x;
    int b;
    int64_t r0, r1;
    gn_block_rows(off, prefix, B, blk, b, r0, r1);
    for (int c0 = 0; c0 < C; c0 += GN_THREADS) {
        const int cc = min(GN_THREADS, C - c0);
        const int R = GN_THREADS / cc;
        A s1 = (A)0, s2 = (A)0;
        if (tid < R * cc) {
            const int c = c0 + tid % cc;
            const int s = b * G + c / (C / G);
            const A mu = mean[s], rs = rstd[s];
            for (int64_t r = r0 + tid / cc; r < r1; r += R) {
                A d = (A)dy[r * C + c];
                if (y && !((A)y[r * C + c] > (A)0)) d = (A)0;
                s1 += d;
                s2 += d * (((A)x[r * C + c] - mu) * rs);
            }
        }
        red[0][tid] = s1;
        red[1][tid] = s2;
        __syncthreads();
        if (tid < cc) {
            A a = red[0][tid], bb = red[1][tid];
            for (int q = 1; q < R; ++q) {
                a += red[0][q * cc + tid];
                bb += red[1][q * cc + tid];
            }
            part[((int64_t)blk * C + c0 + tid) * 2] = a;
            part[((int64_t)blk * C + c0 + tid) * 2 + 1] = bb;
        }
        __syncthreads();
    }
}

// Backward pass 2: one wavefront per (grid, channel): per-grid channel sums sums[b][c][2].
template <typename A>
__global__ void k_gn_bwd_grid_sums(const A* __restrict__ part, int C, const int32_t* __restrict__ prefix, int B,
                                   A* __restrict__ sums) {
    const int wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6, lane = threadIdx.x & 63;
    if (wave >= B * C) return;
    const int b = wave / C, c = wave % C;
    A s1 = (A)0, s2 = (A)0;
    for (int j = prefix[b] + lane; j < prefix[b + 1]; j += 64) {
        s1 += part[((int64_t)j * C + c) * 2];
        s2 += part[((int64_t)j * C + c) * 2 + 1];
    }
#pragma unroll
    for (int m = 1; m < 64; m <<= 1) {
        const A o1 = shfl_xor(s1, m), o2 = shfl_xor(s2, m);
        s1 = (lane & m) ? o1 + s1 : s1 + o1;
        s2 = (lane & m) ? o2 + s2 : s2 + o2;
    }
    if (lane == 0) {
        sums[(int64_t)wave * 2] = s1;
        sums[(int64_t)wave * 2 + 1] = s2;
    }
}

// Backward pass 3: per (grid, group) the gamma-weighted sums; per channel dgamma / dbeta
// (grids in order).  One thread per (grid, group) and one per channel.
template <typename T>
__global__ void k_gn_bwd_group_sums(const typename NormAcc<T>::type* __restrict__ sums, int C, int G, int B,
                                    const T* __restrict__ gamma, typename NormAcc<T>::type* __restrict__ gsums,
                                    float* __restrict__ dgamma_f, float* __restrict__ dbeta_f, double* __restrict__ dgamma_d,
                                    double* __restrict__ dbeta_d) {
    using A = typename NormAcc<T>::type;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int Cg = C / G;
    if (t < B * G) {
        const int b = t / G, g = t % G;
        A a = (A)0, bb = (A)0;
        for (int c = g * Cg; c < (g + 1) * Cg; ++c) {
            const A w = gamma ? (A)gamma[c] : (A)1;
            a += w * sums[((int64_t)b * C + c) * 2];
            bb += w * sums[((int64_t)b * C + c) * 2 + 1];
        }
        gsums[(int64_t)t * 2] = a;
        gsums[(int64_t)t * 2 + 1] = bb;
    } else if (t < B * G + C) {
        const int c = t - B * G;
        A a = (A)0, bb = (A)0;
        for (int b = 0; b < B; ++b) {
            a += sums[((int64_t)b * C + c) * 2];
            bb += sums[((int64_t)b * C + c) * 2 + 1];
        }
        if (dbeta_f) dbeta_f[c] = (float)a;
        if (dgamma_f) dgamma_f[c] = (float)bb;
        if (dbeta_d) dbeta_d[c] = (double)a;
        if (dgamma_d) dgamma_d[c] = (double)bb;
    }
}

// Backward pass 4: dx = rstd * (gamma dy - S1 / M - xhat * S2 / M), M = rows of the grid x C/G.
template <typename T>
__global__ void k_gn_bwd_dx(const T* __restrict__ x, const T* __restrict__ dy, const T* __restrict__ y, int64_t n,
                            int C, int G, const int64_t* __restrict__ off, const int32_t* __restrict__ jidx,
                            const typename NormAcc<T>::type* __restrict__ mean,
                            const typename NormAcc<T>::type* __restrict__ rstd, const T* __restrict__ gamma,
                            const typename NormAcc<T>::type* __restrict__ gsums, T* __restrict__ dx) {
    using A = typename NormAcc<T>::type;
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= n * C) return;
    const int64_t r = idx / C;
    const int c = (int)(idx - r * C);
    const int b = jidx ? jidx[r] : 0;
    const int s = b * G + c / (C / G);
    const A M = (A)(off[b + 1] - off[b]) * (A)(C / G);
    A d = (A)dy[idx];
    if (y && !((A)y[idx] > (A)0)) d = (A)0;
    const A w = gamma ? (A)gamma[c] : (A)1;
    const A xh = ((A)x[idx] - mean[s]) * rstd[s];
    dx[idx] = (T)(rstd[s] * (w * d - gsums[2 * s] / M - xh * gsums[2 * s + 1] / M));
}

template <typename T>
static int gn_forward(const T* x, int64_t n, int C, int G, const int64_t* off, const int32_t* jidx, int B,
                      int64_t nb, const T* gamma, const T* beta, double eps, int relu, T* y, void* mean, void* rstd,
                      void* ws, hipStream_t s) {
    using A = typename NormAcc<T>::type;
    int32_t* prefix = reinterpret_cast<int32_t*>(ws);
    Welford<A>* part = reinterpret_cast<Welford<A>*>(static_cast<char*>(ws) + 256 * ((4 * (B + 1) + 255) / 256));
    hipLaunchKernelGGL(k_gn_block_prefix, dim3(1), dim3(1), 0, s, off, B, prefix);
    FVDBK_LAUNCH_CHECK("k_gn_block_prefix");
    if (nb > 0) {
        hipLaunchKernelGGL((k_gn_stats_partial<T>), dim3((unsigned)nb), dim3(GN_THREADS), 0, s, x, C, off, prefix, B,
                           part);
        FVDBK_LAUNCH_CHECK("k_gn_stats_partial");
    }
    const int64_t waves = (int64_t)B * G;
    hipLaunchKernelGGL((k_gn_stats_final<A>), dim3(grid_1d(waves * 64, 256)), dim3(256), 0, s, part, C, G, prefix, B,
                       eps, (A*)mean, (A*)rstd);
    FVDBK_LAUNCH_CHECK("k_gn_stats_final");
    if (n > 0) {
        hipLaunchKernelGGL((k_gn_apply<T>), dim3(grid_1d(n * C, 256)), dim3(256), 0, s, x, n, C, G, jidx,
                           (const A*)mean, (const A*)rstd, gamma, beta, relu, y);
        FVDBK_LAUNCH_CHECK("k_gn_apply");
    }
    return FVDBK_OK;
}

template <typename T>
static int gn_backward(const T* x, const T* dy, const T* y, int64_t n, int C, int G, const int64_t* off,
                       const int32_t* jidx, int B, int64_t nb, const void* mean, const void* rstd, const T* gamma,
                       T* dx, void* dgamma, void* dbeta, void* ws, hipStream_t s) {
    using A = typename NormAcc<T>::type;
    char* p = static_cast<char*>(ws);
    int32_t* prefix = reinterpret_cast<int32_t*>(p);
    p += 256 * ((4 * (B + 1) + 255) / 256);
    A* part = reinterpret_cast<A*>(p);
    p += 256 * ((sizeof(A) * 2 * nb * C + 255) / 256);
    A* sums = reinterpret_cast<A*>(p);
    p += 256 * ((sizeof(A) * 2 * (int64_t)B * C + 255) / 256);
    A* gsums = reinterpret_cast<A*>(p);
    hipLaunchKernelGGL(k_gn_block_prefix, dim3(1), dim3(1), 0, s, off, B, prefix);
    FVDBK_LAUNCH_CHECK("k_gn_block_prefix");
    if (nb > 0) {
        hipLaunchKernelGGL((k_gn_bwd_partial<T>), dim3((unsigned)nb), dim3(GN_THREADS), 0, s, x, dy, y, C, G, off,
                           prefix, B, (const A*)mean, (const A*)rstd, part);
        FVDBK_LAUNCH_CHECK("k_gn_bwd_partial");
    }
    hipLaunchKernelGGL((k_gn_bwd_grid_sums<A>), dim3(grid_1d((int64_t)B * C * 64, 256)), dim3(256), 0, s, part, C,
                       prefix, B, sums);
    FVDBK_LAUNCH_CHECK("k_gn_bwd_grid_sums");
    const bool dbl = sizeof(A) == 8;
    hipLaunchKernelGGL((k_gn_bwd_group_sums<T>), dim3(grid_1d((int64_t)B * G + C, 256)), dim3(256), 0, s, sums, C, G,
                       B, gamma, gsums, dbl ? nullptr : (float*)dgamma, dbl ? nullptr : (float*)dbeta,
                       dbl ? (double*)dgamma : nullptr, dbl ? (double*)dbeta : nullptr);
    FVDBK_LAUNCH_CHECK("k_gn_bwd_group_sums");
    if (n > 0) {
        hipLaunchKernelGGL((k_gn_bwd_dx<T>), dim3(grid_1d(n * C, 256)), dim3(256), 0, s, x, dy, y, n, C, G, off, jidx,
                           (const A*)mean, (const A*)rstd, gamma, gsums, dx);
        FVDBK_LAUNCH_CHECK("k_gn_bwd_dx");
    }
    return FVDBK_OK;
}

static size_t round256(size_t v) { return (v + 255) / 256 * 256; }

}  // namespace fvdbk

using namespace fvdbk;

extern "C" {

int64_t fvdbk_group_norm_blocks(const int64_t* counts, int32_t n_grids) {
    int64_t nb = 0;
    for (int32_t b = 0; b < n_grids; ++b) nb += (counts[b] + FVDBK_GN_ROWS - 1) / FVDBK_GN_ROWS;
    return nb;
}

size_t fvdbk_group_norm_workspace_size(fvdbk_dtype dtype, int64_t n_blocks, int32_t n_grids, int32_t channels,
                                       int32_t groups) {
    const size_t a = dtype == FVDBK_F64 ? 8 : 4;
    const size_t pre = round256(4 * (size_t)(n_grids + 1));
    const size_t fwd = pre + round256(3 * a * (size_t)n_blocks * channels);
    const size_t bwd = pre + round256(2 * a * (size_t)n_blocks * channels) + round256(2 * a * (size_t)n_grids * channels) +
                       round256(2 * a * (size_t)n_grids * groups);
    return fwd > bwd ? fwd : bwd;
}

#define FVDBK_GN_CHECKS()                                                                                          \
    FVDBK_REQUIRE(C > 0 && G > 0 && C % G == 0, FVDBK_EINVAL, "num_channels must be divisible by num_groups");   \
    FVDBK_REQUIRE(n >= 0 && n_grids > 0 && n_blocks >= 0, FVDBK_EINVAL, "bad row, grid or block count");         \
    FVDBK_REQUIRE(n_grids == 1 || jidx || n == 0, FVDBK_EINVAL, "a batch of several grids needs jidx");          \
    FVDBK_REQUIRE(ws_bytes >= fvdbk_group_norm_workspace_size(dtype, n_blocks, n_grids, C, G), FVDBK_ENOSPC,     \
                  "group-norm workspace too small")

int fvdbk_group_norm_forward(fvdbk_dtype dtype, const void* x, int64_t n, int32_t C, int32_t G, const int64_t* offsets,
                             const int32_t* jidx, int32_t n_grids, int64_t n_blocks, const void* gamma,
                             const void* beta, double eps, uint32_t flags, void* y, void* mean, void* rstd, void* ws,
                             size_t ws_bytes, hipStream_t stream) {
    FVDBK_GN_CHECKS();
    FVDBK_REQUIRE((flags & ~(uint32_t)FVDBK_CONV_RELU) == 0, FVDBK_EINVAL, "unknown flags 0x%x", flags);
    const int relu = (flags & FVDBK_CONV_RELU) ? 1 : 0;
    const int32_t* jx = n_grids == 1 ? nullptr : jidx;
    switch (dtype) {
    case FVDBK_F32:
        return gn_forward<float>((const float*)x, n, C, G, offsets, jx, n_grids, n_blocks, (const float*)gamma,
                                 (const float*)beta, eps, relu, (float*)y, mean, rstd, ws, stream);
    case FVDBK_F64:
        return gn_forward<double>((const double*)x, n, C, G, offsets, jx, n_grids, n_blocks, (const double*)gamma,
                                  (const double*)beta, eps, relu, (double*)y, mean, rstd, ws, stream);
    case FVDBK_F16:
        return gn_forward<_Float16>((const _Float16*)x, n, C, G, offsets, jx, n_grids, n_blocks,
                                    (const _Float16*)gamma, (const _Float16*)beta, eps, relu, (_Float16*)y, mean, rstd,
                                    ws, stream);
    case FVDBK_BF16:
        return gn_forward<__bf16>((const __bf16*)x, n, C, G, offsets, jx, n_grids, n_blocks, (const __bf16*)gamma,
                                  (const __bf16*)beta, eps, relu, (__bf16*)y, mean, rstd, ws, stream);
    default:
        return set_error(FVDBK_ETYPE, "unsupported dtype %d", (int)dtype);
    }
}

int fvdbk_group_norm_backward(fvdbk_dtype dtype, const void* x, const void* dy, const void* y, int64_t n, int32_t C,
                              int32_t G, const int64_t* offsets, const int32_t* jidx, int32_t n_grids,
                              int64_t n_blocks, const void* mean, const void* rstd, const void* gamma, void* dx,
                              void* dgamma, void* dbeta, void* ws, size_t ws_bytes, hipStream_t stream) {
    FVDBK_GN_CHECKS();
    const int32_t* jx = n_grids == 1 ? nullptr : jidx;
    switch (dtype) {
    case FVDBK_F32:
        return gn_backward<float>((const float*)x, (const float*)dy, (const float*)y, n, C, G, offsets, jx, n_grids,
                                  n_blocks, mean, rstd, (const float*)gamma, (float*)dx, dgamma, dbeta, ws, stream);
    case FVDBK_F64:
        return gn_backward<double>((const double*)x, (const double*)dy, (const double*)y, n, C, G, offsets, jx,
                                   n_grids, n_blocks, mean, rstd, (const double*)gamma, (double*)dx, dgamma, dbeta, ws,
                                   stream);
    case FVDBK_F16:
        return gn_backward<_Float16>((const _Float16*)x, (const _Float16*)dy, (const _Float16*)y, n, C, G, offsets,
                                     jx, n_grids, n_blocks, mean, rstd, (const _Float16*)gamma, (_Float16*)dx, dgamma,
                                     dbeta, ws, stream);
    case FVDBK_BF16:
        return gn_backward<__bf16>((const __bf16*)x, (const __bf16*)dy, (const __bf16*)y, n, C, G, offsets, jx,
                                   n_grids, n_blocks, mean, rstd, (const __bf16*)gamma, (__bf16*)dx, dgamma, dbeta, ws,
                                   stream);
    default:
        return set_error(FVDBK_ETYPE, "unsupported dtype %d", (int)dtype);
    }
}

}  // extern "C"
