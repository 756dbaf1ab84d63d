// MFMA operand fragments and instruction wrappers shared by the gfx950 contraction
// kernels (bf16/f16: v_mfma_f32_16x16x32_*, fp32: v_mfma_f32_16x16x4f32).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace fvdbk {

typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));

// Per element type: a lane's MFMA operand fragment is 16 bytes (EPL elements) and
// one "chunk" of channels (CH = 4 lane groups x EPL) is consumed per fragment.
template <typename T> struct Frag;
template <> struct Frag<__bf16> {
    using vec = bf16x8;
    static constexpr int EPL = 8;
    static constexpr int CH = 32;
};
template <> struct Frag<_Float16> {
    using vec = f16x8;
    static constexpr int EPL = 8;
    static constexpr int CH = 32;
};
template <> struct Frag<float> {
    using vec = floatx4;
    static constexpr int EPL = 4;
    static constexpr int CH = 16;
};

__device__ inline void mfma(const bf16x8& a, const bf16x8& b, floatx4& c) {
    c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}
__device__ inline void mfma(const f16x8& a, const f16x8& b, floatx4& c) {
    c = __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
}
// fp32: the lane's 4 elements feed 4 MFMAs; MFMA j reduces over element j of every
// lane group, so the 4 instructions together cover the 16 channels of the chunk.
__device__ inline void mfma(const floatx4& a, const floatx4& b, floatx4& c) {
    c = __builtin_amdgcn_mfma_f32_16x16x4f32(a[0], b[0], c, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f32_16x16x4f32(a[1], b[1], c, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f32_16x16x4f32(a[2], b[2], c, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f32_16x16x4f32(a[3], b[3], c, 0, 0, 0);
}

template <typename T> __device__ inline T from_f32(float v) { return (T)v; }

// Output epilogue of the contraction: optional per-channel bias and ReLU applied to
// the fp32 accumulator, then one rounding to T (NaN passes the ReLU unchanged).
template <typename T>
__device__ inline T epilogue(float v, const T* __restrict__ bias, int m, int relu) {
    if (bias) v += (float)bias[m];
    if (relu && !(v > 0.f) && v == v) v = 0.f;
    return from_f32<T>(v);
}

template <typename T, bool VEC>
__device__ inline typename Frag<T>::vec load_frag(const T* __restrict__ p, int c, int cin) {
    using V = typename Frag<T>::vec;
    V v;
    if (VEC) {
        if (c < cin) return *reinterpret_cast<const V*>(p + c);
#pragma unroll
        for (int e = 0; e < Frag<T>::EPL; ++e) v[e] = (T)0;
        return v;
    }
#pragma unroll
    for (int e = 0; e < Frag<T>::EPL; ++e) v[e] = (c + e < cin) ? p[c + e] : (T)0;
    return v;
}

// LDS transpose read of 16-bit data (gfx950 ds_read_b64_tr_b16).  Within each group of
// 16 lanes, lane s = 4j + q supplies the address of 4 consecutive elements and lane
// i = 4q + p receives element p of lane (4j + q)'s elements as its element j.  So when
// lane s points at row (r0 + s/4), columns (c0 + 4(s%4)) .. +3 of a row-major tile,
// lane i receives column c0 + i of rows r0 .. r0+3: one read per 4 rows instead of a
// 2-byte read per row.
typedef short shortx4 __attribute__((ext_vector_type(4)));
typedef short shortx8 __attribute__((ext_vector_type(8)));
typedef __attribute__((address_space(3))) shortx4 lds_shortx4;
typedef __attribute__((address_space(3))) short lds_short;

__device__ inline shortx4 lds_read_tr16(const void* p) {
    return __builtin_amdgcn_ds_read_tr16_b64_v4i16((lds_shortx4*)(lds_short*)(const short*)p);
}

// An 8-element MFMA operand whose element e is column (c0 + lane%16) of tile row
// (r0 + e), where `p` is this lane's address of row (r0 + (lane%16)/4), column
// (c0 + 4*(lane%4)) and rows are `row_stride` elements apart.
template <typename V>
__device__ inline V lds_frag_tr16(const void* p, int row_stride) {
    const shortx4 lo = lds_read_tr16(p);
    const shortx4 hi = lds_read_tr16((const short*)p + 4 * row_stride);
    const shortx8 v = __builtin_shufflevector(lo, hi, 0, 1, 2, 3, 4, 5, 6, 7);
    return __builtin_bit_cast(V, v);
}

template <typename T>
__device__ inline typename Frag<T>::vec zero_frag() {
    typename Frag<T>::vec v;
#pragma unroll
    for (int e = 0; e < Frag<T>::EPL; ++e) v[e] = (T)0;
    return v;
}

}  // namespace fvdbk
