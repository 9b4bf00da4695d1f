// Pole-pair overlap quantity: forward and backward kernels (jg_overlap_*).
#include <hip/hip_runtime.h>

#include "../../include/jagua_amd.h"
#include "jg_internal.hpp"

using namespace jg;

extern "C" jg_status jg_overlap_fwd(const float*, const float*, const uint32_t*, uint32_t, const uint32_t*, uint64_t,
                                    const float*, uint32_t, float*, void*) {
    return fail(JG_ERR_INTERNAL, "overlap kernels are not built yet");
}
extern "C" jg_status jg_overlap_bwd(const float*, const float*, const uint32_t*, uint32_t, const uint32_t*, uint64_t,
                                    const float*, uint32_t, float*, void*) {
    return fail(JG_ERR_INTERNAL, "overlap kernels are not built yet");
}
