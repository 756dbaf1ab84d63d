// Instances of the LDS-tiled implicit-GEMM contraction for float features.
#include "fvdbk_conv_tile.h"

namespace fvdbk {

bool conv_tile_dispatch(const float* X, int64_t n_in, int cin, const float* A, const int32_t* map, const int32_t* perm,
                        int64_t n_rows, int K, int M, float* Y, const float* bias, int relu, hipStream_t s) {
    return conv_tile_dispatch_t<float>(X, n_in, cin, A, map, perm, n_rows, K, M, Y, bias, relu, s);
}

}  // namespace fvdbk
