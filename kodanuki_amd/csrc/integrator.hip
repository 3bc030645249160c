// The wavefront integrator (placeholder until the shading kernels land).
#include <hip/hip_runtime.h>

#include "ctx.h"

namespace kdn {

bool renderPass(Ctx& c, const kdn_pass_desc& d, kdn_pass_stats* stats) {
    (void)d;
    (void)stats;
    c.error = "kdn_render_pass: the integrator is not built into this library";
    return false;
}

bool uploadTables(Ctx& c) {
    (void)c;
    return true;
}

}  // namespace kdn
