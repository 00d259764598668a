// splat.hip -- placeholder, replaced below.
#include "internal.h"
extern "C" int aarbf_render_splat(aarbf_context *ctx, const aarbf_raygen_data *, const aarbf_push_constants *, int, int) {
  return aarbf::set_error(ctx, AARBF_ERR_UNSUPPORTED, "splat not built yet");
}
