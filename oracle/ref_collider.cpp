// ref_collider.cpp - TEST INFRASTRUCTURE, never linked into the product.
//
// The reference's collider matrices, computed the way the reference computes
// them: with its vendored glm 0.9.9 (external/glm, header-only, float32).
// Restates src/Scene.h:28-38 (Collider(trans, rot, scale): T * Rz * Ry * Rx * S,
// inverse, transpose) and the collider part of src/Scene.cpp:110-136
// (Scene::translateSphere: transform * translate(t), inverse, transpose).
// Built by oracle/Makefile from the reference tree into oracle/_ref/ when the
// reference is present; tests compare hs_collider_from_trs /
// hs_collider_translate (binary64 on the host) against it.
#include <cstring>

#include <glm/glm.hpp>
#define GLM_ENABLE_EXPERIMENTAL
#include <glm/gtx/transform.hpp>

namespace {

void put(float out[48], const glm::mat4& t, const glm::mat4& inv, const glm::mat4& inv_t)
{
    std::memcpy(out, &t[0][0], 64);
    std::memcpy(out + 16, &inv[0][0], 64);
    std::memcpy(out + 32, &inv_t[0][0], 64);
}

}  // namespace

extern "C" void ref_collider_from_trs(float out[48], const float tr[3], const float rot[3], const float sc[3])
{
    const double deg_to_rad = 0.01745329251;  // DEG_TO_RAD of Scene.h:12
    const glm::vec3 t(tr[0], tr[1], tr[2]), r(rot[0], rot[1], rot[2]), s(sc[0], sc[1], sc[2]);
    const glm::mat4 tm = glm::translate(t);
    const glm::mat4 rx = glm::rotate(r.x * (float)deg_to_rad, glm::vec3(1.0, 0.0, 0.0));
    const glm::mat4 ry = glm::rotate(r.y * (float)deg_to_rad, glm::vec3(0.0, 1.0, 0.0));
    const glm::mat4 rz = glm::rotate(r.z * (float)deg_to_rad, glm::vec3(0.0, 0.0, 1.0));
    const glm::mat4 sm = glm::scale(s);
    const glm::mat4 m = tm * rz * ry * rx * sm * glm::mat4(1.0);
    const glm::mat4 inv = glm::inverse(m);
    put(out, m, inv, glm::transpose(inv));
}

extern "C" void ref_collider_translate(float io[48], const float tr[3])
{
    glm::mat4 cur;
    std::memcpy(&cur[0][0], io, 64);
    const glm::mat4 m = glm::translate(cur, glm::vec3(tr[0], tr[1], tr[2]));
    const glm::mat4 inv = glm::inverse(m);
    put(io, m, inv, glm::transpose(inv));
}
