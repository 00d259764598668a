// tf_marks.cpp -- the reference's transfer-function editing pipeline: a
// gradient of (position, RGBA) marks sampled at 64 texel positions i / 63
// and written as the RGBA8 bytes of its R8G8B8A8_SRGB textures
// (hostCode.cpp:416-436 marks, 524-572 texel writes).  Host code only.
//
// The gradient between marks is linear in RGBA, clamped to the first and
// last mark outside them.  The colormap's RGB is gamma-encoded with
// pow(c, 1 / 2.2) before quantisation (the texture's sRGB decode then
// approximately undoes it) and alpha is stored linearly; the radius and
// density maps write only the red byte and leave the other three at 0.
// Quantisation is make_8bit(f) = clamp(int(f * 256), 0, 255).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <vector>

#include "aarbf.h"

namespace {

uint8_t make_8bit(float f) { return (uint8_t)std::min(255, std::max(0, (int)(f * 256.0f))); }

// the gradient at `pos` over marks sorted by position
void gradient_at(const std::vector<aarbf_tf_mark> &m, float pos, float out[4]) {
  if (pos <= m.front().position) {
    for (int c = 0; c < 4; ++c) out[c] = m.front().rgba[c];
    return;
  }
  for (size_t k = 1; k < m.size(); ++k) {
    if (pos <= m[k].position) {
      const aarbf_tf_mark &a = m[k - 1], &b = m[k];
      const float span = b.position - a.position;
      const float t = span > 0.0f ? (pos - a.position) / span : 1.0f;
      for (int c = 0; c < 4; ++c) out[c] = a.rgba[c] + t * (b.rgba[c] - a.rgba[c]);
      return;
    }
  }
  for (int c = 0; c < 4; ++c) out[c] = m.back().rgba[c];
}

} // namespace

extern "C" int aarbf_tf_texels_from_marks(const aarbf_tf_mark *marks, uint32_t n, int kind, uint8_t *out) {
  if (!marks || n == 0 || !out || (kind != AARBF_TF_COLORMAP && kind != AARBF_TF_SCALAR))
    return AARBF_ERR_INVALID_ARGUMENT;
  std::vector<aarbf_tf_mark> m(marks, marks + n);
  for (const aarbf_tf_mark &k : m)
    if (!(k.position >= 0.0f && k.position <= 1.0f)) return AARBF_ERR_INVALID_ARGUMENT;
  std::stable_sort(m.begin(), m.end(),
                   [](const aarbf_tf_mark &a, const aarbf_tf_mark &b) { return a.position < b.position; });
  for (uint32_t i = 0; i < AARBF_TF_TEXELS; ++i) {
    float c[4];
    gradient_at(m, (float)i / 63.0f, c);
    uint8_t *t = out + 4 * i;
    if (kind == AARBF_TF_COLORMAP) {
      for (int k = 0; k < 3; ++k) t[k] = make_8bit(std::pow(c[k], 1.0f / 2.2f));
      t[3] = make_8bit(c[3]);
    } else {
      t[0] = make_8bit(c[0]);
      t[1] = t[2] = t[3] = 0;
    }
  }
  return AARBF_OK;
}
