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
//      mean, M2) of the rows each thread visits, the threads of a channel merged in thread
//      order (pairwise, q with q + 1, q + 2, q + 4 ..., in the 16-byte vector form);
//   2. one wavefront per (grid, group) merges its blocks' partials (lanes stride over
//      (block, channel) items, then a fixed butterfly) with Chan's pairwise update;
//   3. y = (x - mean) * rstd * gamma + beta, optionally ReLU, one thread per element (per
//      16-byte channel segment of GN_RPT rows in the vector form).
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

// ---- Vector forms: C a multiple of E = 16 / sizeof(T) and row data 16-byte aligned.  A thread
// owns one 16-byte segment (E channels) of a row and walks rows; a block covers up to
// GN_THREADS segments of a row (segment chunk s0 .. s0 + ss) and R = GN_THREADS / ss rows at a time.

template <typename T> union GnSeg {
    uint4 v;
    T e[16 / sizeof(T)];
};

constexpr int GN_RPT = 8;  // rows per thread in the element-wise kernels

// Pass 1 (vector form), same partition and output as k_gn_stats_partial: thread (q, s) runs a
// Welford update over rows r0 + q, r0 + q + R, ... of segment s, one count and one reciprocal
// per row shared by the E channels; the R partials of a segment then merge pairwise in LDS
// (q with q + h for h = 1, 2, 4, ...).  The order depends only on the block's own rows.
template <typename T>
__global__ __launch_bounds__(GN_THREADS) void k_gn_stats_partial_vec(const T* __restrict__ x, int C,
                                                                     const int64_t* __restrict__ off,
                                                                     const int32_t* __restrict__ prefix, int B,
                                                                     Welford<typename NormAcc<T>::type>* __restrict__ part) {
    using A = typename NormAcc<T>::type;
    constexpr int E = 16 / sizeof(T);
    __shared__ A s_mean[GN_THREADS * E], s_m2[GN_THREADS * E], s_n[GN_THREADS];
    const int blk = blockIdx.x, tid = threadIdx.x;
    int b;
    int64_t r0, r1;
    gn_block_rows(off, prefix, B, blk, b, r0, r1);
    const int S = C / E, nrows = (int)(r1 - r0);
    for (int s0 = 0; s0 < S; s0 += GN_THREADS) {
        const int ss = min(GN_THREADS, S - s0), R = GN_THREADS / ss, cc = ss * E;
        const int q = tid / ss, s = tid % ss;
        A n = (A)0, mu[E], m2[E];
#pragma unroll
        for (int e = 0; e < E; ++e) mu[e] = m2[e] = (A)0;
        if (q < R) {
            const uint4* xs = reinterpret_cast<const uint4*>(x + r0 * C) + s0 + s;
            auto update = [&](const GnSeg<T>& v) {
                n += (A)1;
                const A inv = (A)1 / n;
#pragma unroll
                for (int e = 0; e < E; ++e) {
                    const A xv = (A)v.e[e];
                    const A d = xv - mu[e];
                    mu[e] += d * inv;
                    m2[e] += d * (xv - mu[e]);
                }
            };
            int r = q;
            for (; r + 3 * R < nrows; r += 4 * R) {
                GnSeg<T> v[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) v[u].v = xs[(int64_t)(r + u * R) * S];
#pragma unroll
                for (int u = 0; u < 4; ++u) update(v[u]);
            }
            for (; r < nrows; r += R) {
                GnSeg<T> v;
                v.v = xs[(int64_t)r * S];
                update(v);
            }
#pragma unroll
            for (int e = 0; e < E; ++e) {
                s_mean[q * cc + s * E + e] = mu[e];
                s_m2[q * cc + s * E + e] = m2[e];
            }
            if (s == 0) s_n[q] = n;
        }
        for (int h = 1; h < R; h <<= 1) {
            __syncthreads();
            const int pairs = (R + 2 * h - 1) / (2 * h);
            for (int it = tid; it < pairs * cc; it += GN_THREADS) {
                const int qa = (it / cc) * 2 * h, qb = qa + h, c = it % cc;
                if (qb >= R) continue;
                const A na = s_n[qa], nb = s_n[qb];
                if (nb == (A)0) continue;
                if (na == (A)0) {
                    s_mean[qa * cc + c] = s_mean[qb * cc + c];
                    s_m2[qa * cc + c] = s_m2[qb * cc + c];
                    continue;
                }
                const A f = nb / (na + nb);
                const A d = s_mean[qb * cc + c] - s_mean[qa * cc + c];
                s_mean[qa * cc + c] += d * f;
                s_m2[qa * cc + c] += s_m2[qb * cc + c] + d * d * na * f;
            }
            __syncthreads();
            if (tid < pairs && tid * 2 * h + h < R) s_n[tid * 2 * h] += s_n[tid * 2 * h + h];
        }
        __syncthreads();
        for (int c = tid; c < cc; c += GN_THREADS)
            part[(int64_t)blk * C + s0 * E + c] = Welford<A>{s_n[0], s_mean[c], s_m2[c]};
        __syncthreads();
    }
}

// Rows of thread (q, s) in the element-wise vector kernels: rb + q + k R for k < GN_RPT, with
// rb = blockIdx.x * R * GN_RPT; segment s0 + s with s0 = blockIdx.y * GN_THREADS.  The grid of
// the first row comes from jidx (all rows are grid 0 without it); later rows advance it along
// the offsets, since grids are contiguous and in order.

// Pass 3 (vector form): y = (x - mean) * rstd * gamma + beta, the same arithmetic as k_gn_apply.
template <typename T>
__global__ __launch_bounds__(GN_THREADS) void k_gn_apply_vec(const T* __restrict__ x, int64_t n, int C, int G,
                                                             const int64_t* __restrict__ off,
                                                             const int32_t* __restrict__ jidx,
                                                             const typename NormAcc<T>::type* __restrict__ mean,
                                                             const typename NormAcc<T>::type* __restrict__ rstd,
                                                             const T* __restrict__ gamma, const T* __restrict__ beta,
                                                             int relu, T* __restrict__ y) {
    using A = typename NormAcc<T>::type;
    constexpr int E = 16 / sizeof(T);
    const int S = C / E, s0 = blockIdx.y * GN_THREADS;
    const int ss = min(GN_THREADS, S - s0), R = GN_THREADS / ss;
    const int q = threadIdx.x / ss, s = s0 + threadIdx.x % ss;
    const int64_t rq = (int64_t)blockIdx.x * R * GN_RPT + q;
    if (q >= R || rq >= n) return;
    const int c0 = s * E, Cg = C / G;
    A w[E], bt[E];
    int grp[E];
#pragma unroll
    for (int e = 0; e < E; ++e) {
        w[e] = gamma ? (A)gamma[c0 + e] : (A)1;
        bt[e] = beta ? (A)beta[c0 + e] : (A)0;
        grp[e] = (c0 + e) / Cg;
    }
    const uint4* xs = reinterpret_cast<const uint4*>(x) + s;
    uint4* ys = reinterpret_cast<uint4*>(y) + s;
    GnSeg<T> v[GN_RPT];
#pragma unroll
    for (int k = 0; k < GN_RPT; ++k)
        if (rq + k * R < n) v[k].v = xs[(rq + k * R) * S];
    int b = jidx ? jidx[rq] : 0;
    int64_t bend = off[b + 1];
    A mu[E], rs[E];
#pragma unroll
    for (int e = 0; e < E; ++e) {
        mu[e] = mean[b * G + grp[e]];
        rs[e] = rstd[b * G + grp[e]];
    }
#pragma unroll
    for (int k = 0; k < GN_RPT; ++k) {
        const int64_t r = rq + k * R;
        if (r >= n) break;
        if (r >= bend) {
            do bend = off[++b + 1];
            while (r >= bend);
#pragma unroll
            for (int e = 0; e < E; ++e) {
                mu[e] = mean[b * G + grp[e]];
                rs[e] = rstd[b * G + grp[e]];
            }
        }
        GnSeg<T> o;
#pragma unroll
        for (int e = 0; e < E; ++e) {
            A t = ((A)v[k].e[e] - mu[e]) * rs[e];
            if (gamma) t = t * w[e];
            if (beta) t = t + bt[e];
            if (relu && !(t > (A)0) && t == t) t = (A)0;
            o.e[e] = (T)t;
        }
        ys[r * S] = o.v;
    }
}

// Backward pass 1 (vector form), same partition and output as k_gn_bwd_partial: per thread
// sequential sums over its rows, then the R partials of a segment summed pairwise in LDS.
template <typename T>
__global__ __launch_bounds__(GN_THREADS) void k_gn_bwd_partial_vec(const T* __restrict__ x, const T* __restrict__ dy,
                                                                   const T* __restrict__ y, int C, int G,
                                                                   const int64_t* __restrict__ off,
                                                                   const int32_t* __restrict__ prefix, int B,
                                                                   const typename NormAcc<T>::type* __restrict__ mean,
                                                                   const typename NormAcc<T>::type* __restrict__ rstd,
                                                                   typename NormAcc<T>::type* __restrict__ part) {
    using A = typename NormAcc<T>::type;
    constexpr int E = 16 / sizeof(T);
    __shared__ A s_1[GN_THREADS * E], s_2[GN_THREADS * E];
    const int blk = blockIdx.x, tid = threadIdx.x;
    int b;
    int64_t r0, r1;
    gn_block_rows(off, prefix, B, blk, b, r0, r1);
    const int S = C / E, nrows = (int)(r1 - r0), Cg = C / G;
    for (int s0 = 0; s0 < S; s0 += GN_THREADS) {
        const int ss = min(GN_THREADS, S - s0), R = GN_THREADS / ss, cc = ss * E;
        const int q = tid / ss, s = tid % ss;
        if (q < R) {
            A a1[E], a2[E], mu[E], rs[E];
#pragma unroll
            for (int e = 0; e < E; ++e) {
                a1[e] = a2[e] = (A)0;
                const int g = b * G + ((s0 + s) * E + e) / Cg;
                mu[e] = mean[g];
                rs[e] = rstd[g];
            }
            const int64_t base = r0 * S + s0 + s;
            const uint4* xs = reinterpret_cast<const uint4*>(x) + base;
            const uint4* ds = reinterpret_cast<const uint4*>(dy) + base;
            const uint4* ps = y ? reinterpret_cast<const uint4*>(y) + base : nullptr;
            auto update = [&](const GnSeg<T>& xv, const GnSeg<T>& dv, const GnSeg<T>& yv) {
#pragma unroll
                for (int e = 0; e < E; ++e) {
                    A d = (A)dv.e[e];
                    if (ps && !((A)yv.e[e] > (A)0)) d = (A)0;
                    a1[e] += d;
                    a2[e] += d * (((A)xv.e[e] - mu[e]) * rs[e]);
                }
            };
            int r = q;
            for (; r + 3 * R < nrows; r += 4 * R) {
                GnSeg<T> xv[4], dv[4], yv[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    xv[u].v = xs[(int64_t)(r + u * R) * S];
                    dv[u].v = ds[(int64_t)(r + u * R) * S];
                    yv[u].v = ps ? ps[(int64_t)(r + u * R) * S] : uint4{0, 0, 0, 0};
                }
#pragma unroll
                for (int u = 0; u < 4; ++u) update(xv[u], dv[u], yv[u]);
            }
            for (; r < nrows; r += R) {
                GnSeg<T> xv, dv, yv;
                xv.v = xs[(int64_t)r * S];
                dv.v = ds[(int64_t)r * S];
                yv.v = ps ? ps[(int64_t)r * S] : uint4{0, 0, 0, 0};
                update(xv, dv, yv);
            }
#pragma unroll
            for (int e = 0; e < E; ++e) {
                s_1[q * cc + s * E + e] = a1[e];
                s_2[q * cc + s * E + e] = a2[e];
            }
        }
        for (int h = 1; h < R; h <<= 1) {
            __syncthreads();
            const int pairs = (R + 2 * h - 1) / (2 * h);
            for (int it = tid; it < pairs * cc; it += GN_THREADS) {
                const int qa = (it / cc) * 2 * h, qb = qa + h, c = it % cc;
                if (qb >= R) continue;
                s_1[qa * cc + c] += s_1[qb * cc + c];
                s_2[qa * cc + c] += s_2[qb * cc + c];
            }
        }
        __syncthreads();
        for (int c = tid; c < cc; c += GN_THREADS) {
            part[((int64_t)blk * C + s0 * E + c) * 2] = s_1[c];
            part[((int64_t)blk * C + s0 * E + c) * 2 + 1] = s_2[c];
        }
        __syncthreads();
    }
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

// Backward pass 4 (vector form), the same arithmetic as k_gn_bwd_dx; rows as in k_gn_apply_vec.
template <typename T>
__global__ __launch_bounds__(GN_THREADS) void k_gn_bwd_dx_vec(const T* __restrict__ x, const T* __restrict__ dy,
                                                              const T* __restrict__ y, int64_t n, int C, int G,
                                                              const int64_t* __restrict__ off,
                                                              const int32_t* __restrict__ jidx,
                                                              const typename NormAcc<T>::type* __restrict__ mean,
                                                              const typename NormAcc<T>::type* __restrict__ rstd,
                                                              const T* __restrict__ gamma,
                                                              const typename NormAcc<T>::type* __restrict__ gsums,
                                                              T* __restrict__ dx) {
    using A = typename NormAcc<T>::type;
    constexpr int E = 16 / sizeof(T);
    const int S = C / E, s0 = blockIdx.y * GN_THREADS;
    const int ss = min(GN_THREADS, S - s0), R = GN_THREADS / ss;
    const int q = threadIdx.x / ss, s = s0 + threadIdx.x % ss;
    const int64_t rq = (int64_t)blockIdx.x * R * GN_RPT + q;
    if (q >= R || rq >= n) return;
    const int c0 = s * E, Cg = C / G;
    A w[E];
    int grp[E];
#pragma unroll
    for (int e = 0; e < E; ++e) {
        w[e] = gamma ? (A)gamma[c0 + e] : (A)1;
        grp[e] = (c0 + e) / Cg;
    }
    const uint4* xs = reinterpret_cast<const uint4*>(x) + s;
    const uint4* ds = reinterpret_cast<const uint4*>(dy) + s;
    const uint4* ps = y ? reinterpret_cast<const uint4*>(y) + s : nullptr;
    uint4* os = reinterpret_cast<uint4*>(dx) + s;
    GnSeg<T> xv[GN_RPT], dv[GN_RPT];
#pragma unroll
    for (int k = 0; k < GN_RPT; ++k)
        if (rq + k * R < n) {
            xv[k].v = xs[(rq + k * R) * S];
            dv[k].v = ds[(rq + k * R) * S];
        }
    int b = jidx ? jidx[rq] : 0;
    int64_t bend = off[b + 1];
    A mu[E], rs[E], t1[E], t2[E];
    auto load_grid = [&]() {
        const A M = (A)(off[b + 1] - off[b]) * (A)Cg;
#pragma unroll
        for (int e = 0; e < E; ++e) {
            const int g = b * G + grp[e];
            mu[e] = mean[g];
            rs[e] = rstd[g];
            t1[e] = gsums[2 * g] / M;
            t2[e] = gsums[2 * g + 1] / M;
        }
    };
    load_grid();
#pragma unroll
    for (int k = 0; k < GN_RPT; ++k) {
        const int64_t r = rq + k * R;
        if (r >= n) break;
        if (r >= bend) {
            do bend = off[++b + 1];
            while (r >= bend);
            load_grid();
        }
        GnSeg<T> yv, o;
        if (ps) yv.v = ps[r * S];
#pragma unroll
        for (int e = 0; e < E; ++e) {
            A d = (A)dv[k].e[e];
            if (ps && !((A)yv.e[e] > (A)0)) d = (A)0;
            const A xh = ((A)xv[k].e[e] - mu[e]) * rs[e];
            o.e[e] = (T)(rs[e] * (w[e] * d - t1[e] - xh * t2[e]));
        }
        os[r * S] = o.v;
    }
}

// The vector kernels apply when C is a multiple of 16 / sizeof(T) and every row array is
// 16-byte aligned.
template <typename T>
static bool gn_vec_ok(int C, const void* a, const void* b, const void* c, const void* d) {
    const uintptr_t m = (uintptr_t)a | (uintptr_t)b | (uintptr_t)c | (uintptr_t)d;
    return C % (16 / (int)sizeof(T)) == 0 && (m & 15) == 0;
}

// Launch grid of the element-wise vector kernels: x over blocks of R * GN_RPT rows, y over
// chunks of GN_THREADS segments (R of the first, widest chunk, so every chunk covers all rows).
template <typename T>
static dim3 gn_vec_grid(int64_t n, int C) {
    const int S = C / (16 / (int)sizeof(T));
    const int R = GN_THREADS / (S < GN_THREADS ? S : GN_THREADS);
    return dim3((unsigned)((n + (int64_t)R * GN_RPT - 1) / ((int64_t)R * GN_RPT)),
                (unsigned)((S + GN_THREADS - 1) / GN_THREADS));
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
    const bool vec = gn_vec_ok<T>(C, x, y, y, y);
    if (nb > 0) {
        if (vec)
            hipLaunchKernelGGL((k_gn_stats_partial_vec<T>), dim3((unsigned)nb), dim3(GN_THREADS), 0, s, x, C, off,
                               prefix, B, part);
        else
            hipLaunchKernelGGL((k_gn_stats_partial<T>), dim3((unsigned)nb), dim3(GN_THREADS), 0, s, x, C, off, prefix,
                               B, part);
        FVDBK_LAUNCH_CHECK("k_gn_stats_partial");
    }
    const int64_t waves = (int64_t)B * G;
    hipLaunchKernelGGL((k_gn_stats_final<A>), dim3(grid_1d(waves * 64, 256)), dim3(256), 0, s, part, C, G, prefix, B,
                       eps, (A*)mean, (A*)rstd);
    FVDBK_LAUNCH_CHECK("k_gn_stats_final");
    if (n > 0) {
        if (vec)
            hipLaunchKernelGGL((k_gn_apply_vec<T>), gn_vec_grid<T>(n, C), dim3(GN_THREADS), 0, s, x, n, C, G, off,
                               jidx, (const A*)mean, (const A*)rstd, gamma, beta, relu, y);
        else
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
    const bool vec = gn_vec_ok<T>(C, x, dy, y ? y : x, dx);
    if (nb > 0) {
        if (vec)
            hipLaunchKernelGGL((k_gn_bwd_partial_vec<T>), dim3((unsigned)nb), dim3(GN_THREADS), 0, s, x, dy, y, C, G,
                               off, prefix, B, (const A*)mean, (const A*)rstd, part);
        else
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
        if (vec)
            hipLaunchKernelGGL((k_gn_bwd_dx_vec<T>), gn_vec_grid<T>(n, C), dim3(GN_THREADS), 0, s, x, dy, y, n, C, G,
                               off, jidx, (const A*)mean, (const A*)rstd, gamma, gsums, dx);
        else
            hipLaunchKernelGGL((k_gn_bwd_dx<T>), dim3(grid_1d(n * C, 256)), dim3(256), 0, s, x, dy, y, n, C, G, off,
                               jidx, (const A*)mean, (const A*)rstd, gamma, gsums, dx);
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
