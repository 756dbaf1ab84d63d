// Instances of the LDS-tiled implicit-GEMM contraction for __bf16 features.
#include "fvdbk_conv_tile.h"

namespace fvdbk {

bool conv_tile_dispatch(const __bf16* X, int64_t n_in, int cin, const __bf16* A, const int32_t* map, const int32_t* perm,
                        int64_t n_rows, int K, int M, __bf16* Y, const __bf16* bias, int relu, hipStream_t s) {
    return conv_tile_dispatch_t<__bf16>(X, n_in, cin, A, map, perm, n_rows, K, M, Y, bias, relu, s);
}

bool conv_tile_dispatch2(const __bf16* X, const __bf16* X1, int64_t n_in, int cin, const __bf16* A,
                         const int32_t* map, const int32_t* perm, int64_t n_rows, int K, int M, __bf16* Y,
                         const __bf16* bias, int relu, hipStream_t s) {
    return conv_tile_dispatch2_t<__bf16>(X, X1, n_in, cin, A, map, perm, n_rows, K, M, Y, bias, relu, s);
}
}  // namespace fvdbk
