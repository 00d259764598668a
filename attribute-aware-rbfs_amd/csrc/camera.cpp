// camera.cpp -- host-side camera and light helpers of the reference's host
// program, in binary32 arithmetic as the reference evaluates them:
//
//   aarbf_light_direction  getLightDirection (sharedCode.h:127-132)
//   aarbf_camera_basis     the camera basis written into the push constants
//                          (hostCode.cpp:669-682 and 712-719)
//   aarbf_camera_update    one frame of the interactive controller
//                          (hostCode.cpp:593-731): up-axis keys, orbit about
//                          the look-at point, zoom, pan
//
// Host code only (no device calls), compiled with -ffp-contract=off so every
// product and sum is rounded on its own as in the reference.  Vector helpers
// follow linalg.h's definitions (dot = x*x' + y*y' + z*z' summed left to
// right, normalize = v / length(v), cross in its usual component order).
// The orbit's rotation matrices are built from the quaternion
// {axis * sin(angle / 2), cos(angle / 2)} in the homogeneous form
// (w^2 + x^2 - y^2 - z^2 on the diagonal), which for a non-unit axis -- the
// reference's lookRight is not normalised -- scales as well as rotates.
#include <cmath>
#include <cstring>

#include "aarbf.h"

namespace {

struct V3 {
  float x, y, z;
};

V3 v3(const float *p) { return V3{p[0], p[1], p[2]}; }
V3 v3(aarbf_float3 p) { return V3{p.x, p.y, p.z}; }
aarbf_float3 f3(V3 v) { return aarbf_float3{v.x, v.y, v.z}; }
V3 add(V3 a, V3 b) { return V3{a.x + b.x, a.y + b.y, a.z + b.z}; }
V3 sub(V3 a, V3 b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }
V3 scale(float s, V3 v) { return V3{s * v.x, s * v.y, s * v.z}; }
V3 cross(V3 a, V3 b) { return V3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
float length(V3 v) { return std::sqrt(v.x * v.x + v.y * v.y + v.z * v.z); }
V3 normalize(V3 v) {
  float l = length(v);
  return V3{v.x / l, v.y / l, v.z / l};
}

// position' = R(q) * (position - pivot) + pivot with q = rotation_quat(axis, angle)
V3 rotate_about(V3 axis, float angle, V3 position, V3 pivot) {
  float s = std::sin(angle / 2.0f), c = std::cos(angle / 2.0f);
  float qx = axis.x * s, qy = axis.y * s, qz = axis.z * s, qw = c;
  // columns of the rotation matrix
  V3 cx{qw * qw + qx * qx - qy * qy - qz * qz, (qx * qy + qz * qw) * 2.0f, (qz * qx - qy * qw) * 2.0f};
  V3 cy{(qx * qy - qz * qw) * 2.0f, qw * qw - qx * qx + qy * qy - qz * qz, (qy * qz + qx * qw) * 2.0f};
  V3 cz{(qz * qx + qy * qw) * 2.0f, (qy * qz - qx * qw) * 2.0f, qw * qw - qx * qx - qy * qy + qz * qz};
  V3 d = sub(position, pivot);
  // mul(float4x4, float4) with d.w = 0: the fourth column drops out
  V3 r{cx.x * d.x + cy.x * d.y + cz.x * d.z, cx.y * d.x + cy.y * d.y + cz.y * d.z,
       cx.z * d.x + cy.z * d.y + cz.z * d.z};
  return add(r, pivot);
}

void basis(V3 from, V3 at, V3 up, float cos_fovy, int width, int height, aarbf_push_constants *pc) {
  V3 d00 = normalize(sub(at, from));
  float aspect = (float)width / (float)height;
  V3 ddu = scale(cos_fovy * aspect, normalize(cross(d00, up)));
  V3 ddv = scale(cos_fovy, normalize(cross(ddu, d00)));
  d00 = sub(d00, scale(0.5f, ddu));
  d00 = sub(d00, scale(0.5f, ddv));
  pc->camera.pos = f3(from);
  pc->camera.dir_00 = f3(d00);
  pc->camera.dir_du = f3(ddu);
  pc->camera.dir_dv = f3(ddv);
}

} // namespace

extern "C" int aarbf_light_direction(float azimuth, float elevation, float out[3]) {
  if (!out) return AARBF_ERR_INVALID_ARGUMENT;
  float a = elevation * 3.14f - 0.5f * 3.14f;
  float b = azimuth * 2.0f * 3.14f;
  out[0] = std::sin(a) * std::cos(b);
  out[1] = std::sin(a) * std::sin(b);
  out[2] = std::cos(a);
  return AARBF_OK;
}

extern "C" int aarbf_camera_basis(const float look_from[3], const float look_at[3], const float look_up[3],
                                  float cos_fovy, int width, int height, aarbf_push_constants *pc) {
  if (!look_from || !look_at || !look_up || !pc || width <= 0 || height <= 0) return AARBF_ERR_INVALID_ARGUMENT;
  basis(v3(look_from), v3(look_at), v3(look_up), cos_fovy, width, height, pc);
  return AARBF_OK;
}

extern "C" int aarbf_camera_update(aarbf_camera_state *st, const aarbf_camera_input *in, int width, int height,
                                   float diagonal, aarbf_push_constants *pc, int *changed_out) {
  if (!st || !in || !pc || width <= 0 || height <= 0) return AARBF_ERR_INVALID_ARGUMENT;
  int changed = 0;
  // the first frame sees no cursor movement (hostCode.cpp:596-599)
  float dx = st->firstFrame ? 0.0f : in->dx;
  float dy = st->firstFrame ? 0.0f : in->dy;
  V3 from = v3(st->lookFrom), at = v3(st->lookAt), up = v3(st->lookUp);

  // up-axis keys (:623-640); a later key overrides an earlier one
  const int keys[3] = {in->key_x, in->key_y, in->key_z};
  for (int k = 0; k < 3; ++k)
    if (keys[k]) {
      up = V3{k == 0 ? 1.0f : 0.0f, k == 1 ? 1.0f : 0.0f, k == 2 ? 1.0f : 0.0f};
      if (in->left_shift) up = scale(-1.0f, up);
    }

  // orbit (:643-683)
  if (in->left || in->key_x || in->key_y || in->key_z || st->firstFrame) {
    st->firstFrame = 0;
    float delta_x = (float)(2.0 * M_PI / width);
    float delta_y = (float)(M_PI / height);
    float x_angle = -dx * delta_x;
    float y_angle = -dy * delta_y;
    V3 position = rotate_about(up, x_angle, from, at);
    V3 right = cross(up, normalize(sub(at, position)));
    from = rotate_about(right, y_angle, position, at);
    basis(from, at, up, st->cosFovy, width, height, pc);
    pc->accumID = 1;
    changed = 1;
  }

  // zoom (:685-702): only the camera position moves, the directions stay
  if (in->right) {
    V3 view = sub(from, at);
    if (dy > 0.0f) {
      view = V3{(float)(view.x * 0.95), (float)(view.y * 0.95), (float)(view.z * 0.95)};
    } else if (dy < 0.0f) {
      view = V3{(float)(view.x * 1.05), (float)(view.y * 1.05), (float)(view.z * 1.05)};
    }
    from = add(at, view);
    pc->camera.pos = f3(from);
    pc->accumID = 1;
    changed = 1;
  }

  // pan (:704-728)
  if (in->middle) {
    V3 right = cross(up, normalize(sub(at, from)));
    V3 t = add(scale(dx, right), scale(-dy, up));
    t = scale(diagonal, scale(0.001f, t));
    from = add(from, t);
    at = add(at, t);
    basis(from, at, up, st->cosFovy, width, height, pc);
    pc->accumID = 1;
    changed = 1;
  }

  st->lookFrom = f3(from);
  st->lookAt = f3(at);
  st->lookUp = f3(up);
  if (changed_out) *changed_out = changed;
  return AARBF_OK;
}
