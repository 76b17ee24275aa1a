// Field MLP (32 -> 64 -> 64 -> n_out, SELU) on MFMA.
#include "common.h"

extern "C" int rm_mlp_fwd(const float*, int64_t, const float*, const float*, const float*, const float*,
                          const float*, const float*, int32_t, int32_t, float*, rm_stream_t) {
    rmesh::set_error("rm_mlp_fwd: not built yet");
    return RM_ERR_INVALID_ARGUMENT;
}

extern "C" int rm_mlp_bwd(const float*, int64_t, const float*, const float*, const float*, const float*,
                          const float*, const float*, int32_t, int32_t, const float*, float*, float*, float*,
                          float*, float*, float*, float*, rm_stream_t) {
    rmesh::set_error("rm_mlp_bwd: not built yet");
    return RM_ERR_INVALID_ARGUMENT;
}
