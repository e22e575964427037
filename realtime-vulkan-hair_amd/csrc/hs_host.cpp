// hs_host.cpp - host-side collider math of the reference's Scene.
//
// The reference builds collider matrices with glm (src/Scene.h:28-38):
//   transform = translate(t) * rotate(rz, Z) * rotate(ry, Y) * rotate(rx, X) * scale(s)
//   inv = inverse(transform), invTrans = transpose(inv)
// and moves collider 0 each frame with Scene::translateSphere
// (src/Scene.cpp:110-136): transform = transform * translate(t).
// This file restates that with its own small column-major mat4 code (no glm).
#include <cmath>
#include <cstring>

#include "../../include/hairsim.h"

namespace {

struct M4 {
    double m[16];  // column-major, m[c*4+r]
};

M4 identity()
{
    M4 r;
    std::memset(r.m, 0, sizeof(r.m));
    r.m[0] = r.m[5] = r.m[10] = r.m[15] = 1.0;
    return r;
}

M4 mul(const M4& a, const M4& b)
{
    M4 r;
    for (int c = 0; c < 4; ++c)
        for (int row = 0; row < 4; ++row) {
            double s = 0.0;
            for (int k = 0; k < 4; ++k) s += a.m[k * 4 + row] * b.m[c * 4 + k];
            r.m[c * 4 + row] = s;
        }
    return r;
}

M4 translate(const double t[3])
{
    M4 r = identity();
    r.m[12] = t[0];
    r.m[13] = t[1];
    r.m[14] = t[2];
    return r;
}

M4 scale(const double s[3])
{
    M4 r = identity();
    r.m[0] = s[0];
    r.m[5] = s[1];
    r.m[10] = s[2];
    return r;
}

// Right-handed rotation by `rad` about coordinate axis `axis` (0=x, 1=y, 2=z).
M4 rotate_axis(double rad, int axis)
{
    M4 r = identity();
    const double c = std::cos(rad), s = std::sin(rad);
    const int i = (axis + 1) % 3, j = (axis + 2) % 3;
    r.m[i * 4 + i] = c;
    r.m[i * 4 + j] = s;
    r.m[j * 4 + i] = -s;
    r.m[j * 4 + j] = c;
    return r;
}

// General 4x4 inverse by Gauss-Jordan elimination with partial pivoting.
M4 inverse(const M4& a)
{
    double w[4][8];
    for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 4; ++c) {
            w[r][c] = a.m[c * 4 + r];
            w[r][4 + c] = r == c ? 1.0 : 0.0;
        }
    for (int col = 0; col < 4; ++col) {
        int piv = col;
        for (int r = col + 1; r < 4; ++r)
            if (std::fabs(w[r][col]) > std::fabs(w[piv][col])) piv = r;
        if (piv != col)
            for (int c = 0; c < 8; ++c) std::swap(w[piv][c], w[col][c]);
        const double d = w[col][col];
        if (d == 0.0) return identity();  // singular: the reference's glm returns inf/nan; keep finite
        for (int c = 0; c < 8; ++c) w[col][c] /= d;
        for (int r = 0; r < 4; ++r) {
            if (r == col) continue;
            const double f = w[r][col];
            if (f == 0.0) continue;
            for (int c = 0; c < 8; ++c) w[r][c] -= f * w[col][c];
        }
    }
    M4 r;
    for (int row = 0; row < 4; ++row)
        for (int c = 0; c < 4; ++c) r.m[c * 4 + row] = w[row][4 + c];
    return r;
}

void store(hs_collider* out, const M4& t)
{
    const M4 inv = inverse(t);
    for (int k = 0; k < 16; ++k) {
        out->transform[k] = (float)t.m[k];
        out->inv[k] = (float)inv.m[k];
    }
    for (int c = 0; c < 4; ++c)
        for (int r = 0; r < 4; ++r) out->invTrans[c * 4 + r] = (float)inv.m[r * 4 + c];
}

}  // namespace

extern "C" {

void hs_collider_from_trs(hs_collider* out, const float tr[3], const float rot_deg[3], const float sc[3])
{
    if (!out || !tr || !rot_deg || !sc) return;
    // the angle as Scene.h:30-32 forms it: rot * (float)DEG_TO_RAD in float
    // (DEG_TO_RAD 0.01745329251, Scene.h:12); cos/sin and products in binary64
    const float kDegToRad = (float)0.01745329251;
    const double t[3] = {tr[0], tr[1], tr[2]};
    const double s[3] = {sc[0], sc[1], sc[2]};
    M4 m = translate(t);
    m = mul(m, rotate_axis((double)(rot_deg[2] * kDegToRad), 2));
    m = mul(m, rotate_axis((double)(rot_deg[1] * kDegToRad), 1));
    m = mul(m, rotate_axis((double)(rot_deg[0] * kDegToRad), 0));
    m = mul(m, scale(s));
    store(out, m);
}

void hs_collider_translate(hs_collider* c, const float tr[3])
{
    if (!c || !tr) return;
    M4 m;
    for (int k = 0; k < 16; ++k) m.m[k] = c->transform[k];
    const double t[3] = {tr[0], tr[1], tr[2]};
    store(c, mul(m, translate(t)));
}

}  // extern "C"
