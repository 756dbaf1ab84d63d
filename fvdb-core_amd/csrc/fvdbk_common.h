// Internal helpers shared by the fvdbk HIP sources (not part of the C ABI).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "fvdbk.h"

namespace fvdbk {

// Records a thread-local message and returns `code`.
int set_error(int code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));

// Returns FVDBK_EHIP with the launch error of the last kernel, or FVDBK_OK.
int check_launch(const char* what);

#define FVDBK_REQUIRE(cond, code, ...)                      \
    do {                                                    \
        if (!(cond)) return ::fvdbk::set_error((code), __VA_ARGS__); \
    } while (0)

#define FVDBK_LAUNCH_CHECK(what)                   \
    do {                                           \
        int _rc = ::fvdbk::check_launch(what);     \
        if (_rc != FVDBK_OK) return _rc;           \
    } while (0)

constexpr int kCoordBits = 19;                       // coordinates in [-2^19, 2^19)
constexpr int32_t kCoordMin = -(1 << kCoordBits);
constexpr int32_t kCoordMax = (1 << kCoordBits) - 1;
constexpr int kMaxGrids = 1024;
constexpr uint64_t kSkipKey = 0x7FFFFFFFFFFFFFFFull;  // sorts after every valid key (signed or unsigned)

// Leaf sort key.  Sorting keys ascending orders leaves by (grid, root tile, upper-node
// child, lower-node child) with each level compared x, then y, then z: the order in
// which a NanoVDB OnIndex grid numbers its leaves.  The root tile digit is the low 8
// bits of uint32(c) >> 12, which keeps the unsigned root-key order for coordinates
// in [-2^19, 2^19) (negative tiles after positive ones).
__host__ __device__ inline uint64_t leaf_key(int32_t b, int32_t x, int32_t y, int32_t z) {
    const uint64_t tx = ((uint32_t)x >> 12) & 255u, ty = ((uint32_t)y >> 12) & 255u, tz = ((uint32_t)z >> 12) & 255u;
    const uint64_t ux = (uint32_t)(x >> 7) & 31u, uy = (uint32_t)(y >> 7) & 31u, uz = (uint32_t)(z >> 7) & 31u;
    const uint64_t lx = (uint32_t)(x >> 3) & 15u, ly = (uint32_t)(y >> 3) & 15u, lz = (uint32_t)(z >> 3) & 15u;
    return ((uint64_t)(uint32_t)b << 51) | (tx << 43) | (ty << 35) | (tz << 27) | (ux << 22) | (uy << 17) |
           (uz << 12) | (lx << 8) | (ly << 4) | lz;
}

// Inverse of leaf_key: grid index and the leaf's origin (multiple of 8).
__host__ __device__ inline void decode_leaf_key(uint64_t key, int32_t& b, int32_t& x, int32_t& y, int32_t& z) {
    b = (int32_t)(key >> 51);
    int32_t t[3], u[3], l[3];
    t[0] = (int32_t)((key >> 43) & 255u);
    t[1] = (int32_t)((key >> 35) & 255u);
    t[2] = (int32_t)((key >> 27) & 255u);
    u[0] = (int32_t)((key >> 22) & 31u);
    u[1] = (int32_t)((key >> 17) & 31u);
    u[2] = (int32_t)((key >> 12) & 31u);
    l[0] = (int32_t)((key >> 8) & 15u);
    l[1] = (int32_t)((key >> 4) & 15u);
    l[2] = (int32_t)(key & 15u);
    int32_t c[3];
    for (int d = 0; d < 3; ++d) {
        const int32_t tile = t[d] >= 128 ? t[d] - 256 : t[d];
        c[d] = tile * 4096 + u[d] * 128 + l[d] * 8;
    }
    x = c[0];
    y = c[1];
    z = c[2];
}

__host__ __device__ inline uint64_t mix64(uint64_t k) {
    k ^= k >> 33;
    k *= 0xff51afd7ed558ccdull;
    k ^= k >> 33;
    k *= 0xc4ceb9fe1a85ec53ull;
    k ^= k >> 33;
    return k;
}

// Leaf index of `key` in an open-addressing table, or -1.
__device__ inline int32_t hash_find(const fvdbk_hslot* __restrict__ h, int64_t cap, uint64_t key) {
    uint64_t s = mix64(key) & (uint64_t)(cap - 1);
    for (;;) {
        const uint4 v = *reinterpret_cast<const uint4*>(h + s);
        const int32_t leaf = (int32_t)v.z;
        if (leaf < 0) return -1;
        const uint64_t k = (uint64_t)v.x | ((uint64_t)v.y << 32);
        if (k == key) return leaf;
        s = (s + 1) & (uint64_t)(cap - 1);
    }
}

__device__ inline int32_t floordiv(int32_t a, int32_t b) {
    int32_t q = a / b;
    if ((a % b != 0) && ((a < 0) != (b < 0))) --q;
    return q;
}

inline unsigned grid_1d(int64_t n, int block) { return (unsigned)((n + block - 1) / block); }

// Compute units of the current device (cached per device); persistent kernels size their
// grids by it.
static inline int device_cu_count() {
    static int cus[64] = {0};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
    if (cus[dev] == 0) {
        int n = 0;
        if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n <= 0) n = 256;
        cus[dev] = n;
    }
    return cus[dev];
}

}  // namespace fvdbk
