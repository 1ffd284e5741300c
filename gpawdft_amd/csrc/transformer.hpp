// Internal interface of the restriction / interpolation kernels.
#pragma once

#include "common.hpp"

struct gpd_transformer {
    int k = 2;
    bool interpolate = false;
    int n_in[3] = {0, 0, 0};
    int n_out[3] = {0, 0, 0};
    int pad[3][2] = {};
    int skip[3][2] = {};
    int pbc[3] = {0, 0, 0};
};

namespace gpd {

enum : int {
    TEPI_PLAIN = 0,          // out = T(in)
    TEPI_SCALED_COPY = 1,    // restriction: out = T(in), out2 = alpha * out
    TEPI_SUBTRACT_FROM = 2,  // interpolation: out = out - T(in)
};

template <typename T>
int transformer_apply_typed(const gpd_transformer* t, int epi, bool neg_in, long nb, const T* in,
                            T* out, T* out2, T alpha, hipStream_t st);

}  // namespace gpd
