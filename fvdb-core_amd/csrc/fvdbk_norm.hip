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
//      16-byte channel segment of a row, in persistent blocks, in the vector form).
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
    // items lane, lane + 64, ... in order; UF loads in flight before they merge
    constexpr int UF = 8;
    for (int it0 = 0; it0 < items; it0 += 64 * UF) {
        Welford<A> v[UF];
#pragma unroll
        for (int u = 0; u < UF; ++u) {
            const int it = it0 + u * 64 + lane;
            v[u] = it < items ? part[(int64_t)(prefix[b] + it / Cg) * C + g * Cg + it % Cg]
                              : Welford<A>{(A)0, (A)0, (A)0};
        }
#pragma unroll
        for (int u = 0; u < UF; ++u) w = wmerge(w, v[u]);
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

constexpr int GN_RPT = 4;             // rows per thread and chunk in the element-wise kernels
constexpr int GN_STATS_THREADS = 128;  // threads per block of the vector statistics kernels

// Pass 1 (vector form), same partition and output as k_gn_stats_partial, one workgroup of NT
// threads per block (two wavefronts: the blocks of a large batch then all fit on the device at
// once): thread (q, s) sums d = x - K and d^2 over rows r0 + q, r0 + q + R, ... of segment s,
// K the channel's value in the thread's first row (shifted data: sum d^2 - (sum d)^2 / n stays
// exact to rounding while |mean - K| is a few standard deviations), turns them into (count,
// mean, M2) and the R partials of a segment then merge pairwise in LDS (q with q + h for
// h = 1, 2, 4, ...).  The order depends only on the block's own rows.
template <typename T, int NT>
__global__ __launch_bounds__(NT) void k_gn_stats_partial_vec(const T* __restrict__ x, int C,
                                                                     const int64_t* __restrict__ off,
                                                                     const int32_t* __restrict__ prefix, int B,
                                                                     Welford<typename NormAcc<T>::type>* __restrict__ part) {
    using A = typename NormAcc<T>::type;
    constexpr int E = 16 / sizeof(T);
    __shared__ A s_mean[NT * E], s_m2[NT * E], s_n[NT];
    const int blk = blockIdx.x, tid = threadIdx.x;
    int b;
    int64_t r0, r1;
    gn_block_rows(off, prefix, B, blk, b, r0, r1);
    const int S = C / E, nrows = (int)(r1 - r0);
    for (int s0 = 0; s0 < S; s0 += NT) {
        const int ss = min(NT, S - s0), R = NT / ss, cc = ss * E;
        const int q = tid / ss, s = tid % ss;
        if (q < R) {
            const uint4* xs = reinterpret_cast<const uint4*>(x + r0 * C) + s0 + s;
            // rows q + i R, i < mine, in batches of U; the next batch loads before this one is used
            const int mine = (nrows - q + R - 1) / R;
            constexpr int U = 8;
            GnSeg<T> cur[U], nxt[U];
            auto load = [&](int i0, GnSeg<T>* v) {
#pragma unroll
                for (int u = 0; u < U; ++u) v[u].v = xs[(int64_t)(q + min(i0 + u, mine - 1) * R) * S];
            };
            A K[E], sd[E], sq[E];
            if (mine > 0) load(0, cur);
            else cur[0].v = uint4{0, 0, 0, 0};
#pragma unroll
            for (int e = 0; e < E; ++e) {
                K[e] = (A)cur[0].e[e];
                sd[e] = sq[e] = (A)0;
            }
            for (int i0 = 0; i0 < mine; i0 += U) {
                if (i0 + U < mine) load(i0 + U, nxt);
#pragma unroll
                for (int u = 0; u < U; ++u)
                    if (i0 + u < mine) {
#pragma unroll
                        for (int e = 0; e < E; ++e) {
                            const A d = (A)cur[u].e[e] - K[e];
                            sd[e] += d;
                            sq[e] += d * d;
                        }
                    }
#pragma unroll
                for (int u = 0; u < U; ++u) cur[u] = nxt[u];
            }
            const A n = (A)max(mine, 0);
            const A inv = mine > 0 ? (A)1 / n : (A)0;
            A mu[E], m2[E];
#pragma unroll
            for (int e = 0; e < E; ++e) {
                const A md = sd[e] * inv;
                mu[e] = mine > 0 ? K[e] + md : (A)0;
                m2[e] = mine > 0 ? max(sq[e] - sd[e] * md, (A)0) : (A)0;
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
            for (int it = tid; it < pairs * cc; it += NT) {
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
        for (int c = tid; c < cc; c += NT)
            part[(int64_t)blk * C + s0 * E + c] = Welford<A>{s_n[0], s_mean[c], s_m2[c]};
        __syncthreads();
    }
}

// The element-wise vector kernels are persistent: block j handles the row chunks j,
// j + gridDim.x, ... of CH = R * GN_RPT rows (thread (q, s): rows chunk * CH + q + k R for
// k < GN_RPT, segment s0 + s with s0 = blockIdx.y * GN_THREADS) and loads the next chunk
// before it writes the current one.  A thread's rows only increase, so its grid index walks
// forward along the offsets from the grid of its first row (jidx; all rows are grid 0 without
// it), never past grid B - 1, and the statistics of a grid are loaded once per thread and grid.

// Pass 3 (vector form): y = (x - mean) * rstd * gamma + beta, the same arithmetic as k_gn_apply.
template <typename T>
__global__ __launch_bounds__(GN_THREADS) void k_gn_apply_vec(const T* __restrict__ x, int64_t n, int C, int G,
                                                             const int64_t* __restrict__ off, int B,
                                                             const int32_t* __restrict__ jidx,
                                                             const typename NormAcc<T>::type* __restrict__ mean,
                                                             const typename NormAcc<T>::type* __restrict__ rstd,
                                                             const T* __restrict__ gamma, const T* __restrict__ beta,
                                                             int relu, T* __restrict__ y) {
    using A = typename NormAcc<T>::type;
    constexpr int E = 16 / sizeof(T), P = GN_RPT;
    const int S = C / E, s0 = blockIdx.y * GN_THREADS;
    const int ss = min(GN_THREADS, S - s0), R = GN_THREADS / ss;
    const int q = threadIdx.x / ss, s = s0 + threadIdx.x % ss;
    const int64_t CH = (int64_t)R * P, nch = (n + CH - 1) / CH;
    int64_t ch = blockIdx.x;
    if (q >= R || ch >= nch) return;
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
    auto row = [&](int64_t c, int k) { return c * CH + q + (int64_t)k * R; };
    GnSeg<T> cur[P];
#pragma unroll
    for (int k = 0; k < P; ++k) cur[k].v = xs[min(row(ch, k), n - 1) * S];
    int b = jidx ? min(max(jidx[min(row(ch, 0), n - 1)], 0), B - 1) : 0;
    int64_t bend = off[b + 1];
    A mu[E], rs[E];
    auto load_grid = [&]() {
#pragma unroll
        for (int e = 0; e < E; ++e) {
            mu[e] = mean[b * G + grp[e]];
            rs[e] = rstd[b * G + grp[e]];
        }
    };
    load_grid();
    for (;;) {
        const int64_t nx = ch + gridDim.x;
        GnSeg<T> nxt[P];
        if (nx < nch) {
#pragma unroll
            for (int k = 0; k < P; ++k) nxt[k].v = xs[min(row(nx, k), n - 1) * S];
        }
#pragma unroll
        for (int k = 0; k < P; ++k) {
            const int64_t r = row(ch, k);
            if (r >= n) break;
            if (r >= bend && b + 1 < B) {
                do bend = off[++b + 1];
                while (r >= bend && b + 1 < B);
                load_grid();
            }
            GnSeg<T> o;
#pragma unroll
            for (int e = 0; e < E; ++e) {
                A t = ((A)cur[k].e[e] - mu[e]) * rs[e];
                if (gamma) t = t * w[e];
                if (beta) t = t + bt[e];
                if (relu && !(t > (A)0) && t == t) t = (A)0;
                o.e[e] = (T)t;
            }
            ys[r * S] = o.v;
        }
        if (nx >= nch) break;
        ch = nx;
#pragma unroll
        for (int k = 0; k < P; ++k) cur[k] = nxt[k];
    }
}

// Backward pass 1 (vector form), same partition and output as k_gn_bwd_partial: per thread
// sequential sums over its rows, then the R partials of a segment summed pairwise in LDS.
// NT threads per workgroup as in k_gn_stats_partial_vec; HAS_Y: the forward applied a ReLU
// and y masks dy.
template <typename T, bool HAS_Y, int NT>
__global__ __launch_bounds__(NT) void k_gn_bwd_partial_vec(const T* __restrict__ x, const T* __restrict__ dy,
                                                                   const T* __restrict__ y, int C, int G,
                                                                   const int64_t* __restrict__ off,
                                                                   const int32_t* __restrict__ prefix, int B,
                                                                   const typename NormAcc<T>::type* __restrict__ mean,
                                                                   const typename NormAcc<T>::type* __restrict__ rstd,
                                                                   typename NormAcc<T>::type* __restrict__ part) {
    using A = typename NormAcc<T>::type;
    constexpr int E = 16 / sizeof(T);
    __shared__ A s_1[NT * E], s_2[NT * E];
    const int blk = blockIdx.x, tid = threadIdx.x;
    int b;
    int64_t r0, r1;
    gn_block_rows(off, prefix, B, blk, b, r0, r1);
    const int S = C / E, nrows = (int)(r1 - r0), Cg = C / G;
    for (int s0 = 0; s0 < S; s0 += NT) {
        const int ss = min(NT, S - s0), R = NT / ss, cc = ss * E;
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
            const uint4* ps = reinterpret_cast<const uint4*>(y) + base;
            auto update = [&](const GnSeg<T>& xv, const GnSeg<T>& dv, const GnSeg<T>& yv) {
#pragma unroll
                for (int e = 0; e < E; ++e) {
                    A d = (A)dv.e[e];
                    if constexpr (HAS_Y)
                        if (!((A)yv.e[e] > (A)0)) d = (A)0;
                    a1[e] += d;
                    a2[e] += d * (((A)xv.e[e] - mu[e]) * rs[e]);
                }
            };
            const int mine = (nrows - q + R - 1) / R;
            constexpr int U = 4, NY = HAS_Y ? U : 1;
            GnSeg<T> cx[U], cd[U], cy[NY], nx_[U], nd[U], ny[NY];
            auto load = [&](int i0, GnSeg<T>* xv, GnSeg<T>* dv, GnSeg<T>* yv) {
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    const int64_t o = (int64_t)(q + min(i0 + u, mine - 1) * R) * S;
                    xv[u].v = xs[o];
                    dv[u].v = ds[o];
                    if constexpr (HAS_Y) yv[u].v = ps[o];
                }
            };
            if (mine > 0) load(0, cx, cd, cy);
            for (int i0 = 0; i0 < mine; i0 += U) {
                if (i0 + U < mine) load(i0 + U, nx_, nd, ny);
#pragma unroll
                for (int u = 0; u < U; ++u)
                    if (i0 + u < mine) update(cx[u], cd[u], cy[HAS_Y ? u : 0]);
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    cx[u] = nx_[u];
                    cd[u] = nd[u];
                    if constexpr (HAS_Y) cy[u] = ny[u];
                }
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
            for (int it = tid; it < pairs * cc; it += NT) {
                const int qa = (it / cc) * 2 * h, qb = qa + h, c = it % cc;
                if (qb >= R) continue;
                s_1[qa * cc + c] += s_1[qb * cc + c];
                s_2[qa * cc + c] += s_2[qb * cc + c];
            }
        }
        __syncthreads();
        for (int c = tid; c < cc; c += NT) {
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

// Backward pass 4 (vector form), the same arithmetic as k_gn_bwd_dx, persistent as
// k_gn_apply_vec; HAS_Y: the forward applied a ReLU and y masks dy.
template <typename T, bool HAS_Y>
__global__ __launch_bounds__(GN_THREADS) void k_gn_bwd_dx_vec(const T* __restrict__ x, const T* __restrict__ dy,
                                                              const T* __restrict__ y, int64_t n, int C, int G,
                                                              const int64_t* __restrict__ off, int B,
                                                              const int32_t* __restrict__ jidx,
                                                              const typename NormAcc<T>::type* __restrict__ mean,
                                                              const typename NormAcc<T>::type* __restrict__ rstd,
                                                              const T* __restrict__ gamma,
                                                              const typename NormAcc<T>::type* __restrict__ gsums,
                                                              T* __restrict__ dx) {
    using A = typename NormAcc<T>::type;
    constexpr int E = 16 / sizeof(T), P = GN_RPT, NY = HAS_Y ? P : 1;
    const int S = C / E, s0 = blockIdx.y * GN_THREADS;
    const int ss = min(GN_THREADS, S - s0), R = GN_THREADS / ss;
    const int q = threadIdx.x / ss, s = s0 + threadIdx.x % ss;
    const int64_t CH = (int64_t)R * P, nch = (n + CH - 1) / CH;
    int64_t ch = blockIdx.x;
    if (q >= R || ch >= nch) return;
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
    const uint4* ps = reinterpret_cast<const uint4*>(y) + s;
    uint4* os = reinterpret_cast<uint4*>(dx) + s;
    auto row = [&](int64_t c, int k) { return c * CH + q + (int64_t)k * R; };
    GnSeg<T> cx[P], cd[P], cy[NY];
    auto load_rows = [&](int64_t c, GnSeg<T>* xv, GnSeg<T>* dv, GnSeg<T>* yv) {
#pragma unroll
        for (int k = 0; k < P; ++k) {
            const int64_t r = min(row(c, k), n - 1) * S;
            xv[k].v = xs[r];
            dv[k].v = ds[r];
            if constexpr (HAS_Y) yv[k].v = ps[r];
        }
    };
    load_rows(ch, cx, cd, cy);
    int b = jidx ? min(max(jidx[min(row(ch, 0), n - 1)], 0), B - 1) : 0;
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
    for (;;) {
        const int64_t nx = ch + gridDim.x;
        GnSeg<T> nxx[P], nxd[P], nxy[NY];
        if (nx < nch) load_rows(nx, nxx, nxd, nxy);
#pragma unroll
        for (int k = 0; k < P; ++k) {
            const int64_t r = row(ch, k);
            if (r >= n) break;
            if (r >= bend && b + 1 < B) {
                do bend = off[++b + 1];
                while (r >= bend && b + 1 < B);
                load_grid();
            }
            GnSeg<T> o;
#pragma unroll
            for (int e = 0; e < E; ++e) {
                A d = (A)cd[k].e[e];
                if constexpr (HAS_Y)
                    if (!((A)cy[k].e[e] > (A)0)) d = (A)0;
                const A xh = ((A)cx[k].e[e] - mu[e]) * rs[e];
                o.e[e] = (T)(rs[e] * (w[e] * d - t1[e] - xh * t2[e]));
            }
            os[r * S] = o.v;
        }
        if (nx >= nch) break;
        ch = nx;
#pragma unroll
        for (int k = 0; k < P; ++k) {
            cx[k] = nxx[k];
            cd[k] = nxd[k];
            if constexpr (HAS_Y) cy[k] = nxy[k];
        }
    }
}

// The vector kernels apply when C is a multiple of 16 / sizeof(T) and every row array is
// 16-byte aligned.
template <typename T>
static bool gn_vec_ok(int C, const void* a, const void* b, const void* c, const void* d) {
    const uintptr_t m = (uintptr_t)a | (uintptr_t)b | (uintptr_t)c | (uintptr_t)d;
    return C % (16 / (int)sizeof(T)) == 0 && (m & 15) == 0;
}

// Launch grid of the persistent element-wise vector kernels: y over chunks of GN_THREADS
// segments; x the row chunks of R * GN_RPT rows (R of the first, widest segment chunk, so every
// segment chunk covers all rows), at most the blocks resident on the device at once.
template <typename T, typename K>
static dim3 gn_vec_grid(K kernel, int64_t n, int C) {
    static int resident[64] = {0};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) dev = 0;
    if (resident[dev] == 0) {
        int per_cu = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, GN_THREADS, 0) != hipSuccess || per_cu <= 0)
            per_cu = 1;
        resident[dev] = per_cu * device_cu_count();
    }
    const int S = C / (16 / (int)sizeof(T));
    const int gy = (S + GN_THREADS - 1) / GN_THREADS;
    const int R = GN_THREADS / (S < GN_THREADS ? S : GN_THREADS);
    const int64_t nch = (n + (int64_t)R * GN_RPT - 1) / ((int64_t)R * GN_RPT);
    const int64_t cap = resident[dev] / gy > 0 ? resident[dev] / gy : 1;
    return dim3((unsigned)(nch < cap ? nch : cap), (unsigned)gy);
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
            hipLaunchKernelGGL((k_gn_stats_partial_vec<T, GN_STATS_THREADS>), dim3((unsigned)nb),
                               dim3(GN_STATS_THREADS), 0, s, x, C, off, prefix, B, part);
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
            hipLaunchKernelGGL((k_gn_apply_vec<T>), gn_vec_grid<T>(k_gn_apply_vec<T>, n, C), dim3(GN_THREADS), 0, s,
                               x, n, C, G, off, B, jidx, (const A*)mean, (const A*)rstd, gamma, beta, relu, y);
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
        if (vec && y)
            hipLaunchKernelGGL((k_gn_bwd_partial_vec<T, true, GN_STATS_THREADS>), dim3((unsigned)nb),
                               dim3(GN_STATS_THREADS), 0, s, x, dy, y, C, G, off, prefix, B, (const A*)mean,
                               (const A*)rstd, part);
        else if (vec)
            hipLaunchKernelGGL((k_gn_bwd_partial_vec<T, false, GN_STATS_THREADS>), dim3((unsigned)nb),
                               dim3(GN_STATS_THREADS), 0, s, x, dy, y, C, G, off, prefix, B, (const A*)mean,
                               (const A*)rstd, part);
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
        if (vec && y)
            hipLaunchKernelGGL((k_gn_bwd_dx_vec<T, true>), gn_vec_grid<T>(k_gn_bwd_dx_vec<T, true>, n, C),
                               dim3(GN_THREADS), 0, s, x, dy, y, n, C, G, off, B, jidx, (const A*)mean,
                               (const A*)rstd, gamma, gsums, dx);
        else if (vec)
            hipLaunchKernelGGL((k_gn_bwd_dx_vec<T, false>), gn_vec_grid<T>(k_gn_bwd_dx_vec<T, false>, n, C),
                               dim3(GN_THREADS), 0, s, x, dy, y, n, C, G, off, B, jidx, (const A*)mean,
                               (const A*)rstd, gamma, gsums, dx);
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
