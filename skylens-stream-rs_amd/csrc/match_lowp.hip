// Low-precision (BF16 / FP8) ratio_match: placeholder until the MFMA kernel lands.
#include "skylens_hip.h"
#include "skh_ctx.hpp"

namespace skh {
int Ctx::ratio_match_lowp(const float*, size_t, const float*, size_t, float, int, int, uint32_t*,
                          uint32_t*) {
    return fail(SKH_EINVAL, "low-precision ratio_match not built yet");
}
}  // namespace skh
