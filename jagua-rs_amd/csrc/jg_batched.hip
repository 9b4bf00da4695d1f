// Batched persistent query form (jg_plan_*).
#include <hip/hip_runtime.h>

#include "../../include/jagua_amd.h"
#include "jg_internal.hpp"

using namespace jg;

extern "C" jg_status jg_plan_create(const jg_query_set*, uint32_t, const jg_plan_config*, jg_plan**) {
    return fail(JG_ERR_INTERNAL, "batched plans are not built yet");
}
extern "C" void jg_plan_free(jg_plan*) {}
extern "C" jg_status jg_plan_detect(jg_plan*, const float*, uint8_t*, void*) {
    return fail(JG_ERR_INTERNAL, "batched plans are not built yet");
}
extern "C" jg_status jg_plan_detect_graph(jg_plan*, const float*, uint8_t*, uint32_t, void*) {
    return fail(JG_ERR_INTERNAL, "batched plans are not built yet");
}
