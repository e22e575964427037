// hs_forcing.hpp - the reference's optional wind forcing (the two lines
// compute.comp:150-151 carries commented out), for host and device.
//
// Mode 1 (compute.comp:150):
//   force += 10.0 * vec3(2.0*sin(t*2.0) * cos(y*10.0) * sin((y+5.0)*15.0),
//                        0.0, -clamp(y*2.0, 0.2, 2.0))
// Mode 2 (compute.comp:151):
//   force += 7.0 * fbm(vec2(sin(t), cos(t))) *
//            vec3(2.0*sin(t*2.0) * cos(y*10.0) * sin((y+5.0)*15.0),
//                 4.0*sin(z*5.0 + t*3.0), -0.6*(y+3.0))
// with random / noise / fbm of compute.comp:81-120.  GLSL leaves the
// precision of sin and cos to the implementation; here they are one fixed
// binary32 algorithm (below), evaluated with the same operations on the CPU
// and the GPU, so a run is reproducible bit for bit on either.
//
// sin/cos: x = k*(pi/2) + r with k = floor(x*(2/pi) + 0.5) and r reduced by a
// three-part Cody-Waite split of pi/2 (8 + 8 significant bits + a float
// remainder: k*P1 and k*P2 are exact for |k| < 2^16); then odd / even
// polynomials in r on [-pi/4, pi/4] (coefficients fitted by weighted least
// squares to the Taylor-series remainders, rounded to float) and the
// quadrant of k.  Measured against binary64 sin/cos: absolute error below
// 1e-7 for |x| <= 1e5, below 1.6 ulp for |x| <= 10.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

namespace hs {

__host__ __device__ inline float fr_floor(float x)
{
#ifdef __HIP_DEVICE_COMPILE__
    return floorf(x);
#else
    return std::floor(x);
#endif
}

// sin and cos of x (one range reduction for both).
__host__ __device__ inline void fr_sincos(float x, float* sn, float* cs)
{
    const float kTwoOverPi = 0.636619772367581343f;
    const float P1 = 1.5703125f, P2 = 4.84466552734375e-4f, P3 = -6.397578431460715e-7f;
    const float k = fr_floor(x * kTwoOverPi + 0.5f);
    const float r = ((x - k * P1) - k * P2) - k * P3;
    const float r2 = r * r;
    float p = 2.7181240511708893e-6f;
    p = p * r2 + -1.9839312881231308e-4f;
    p = p * r2 + 8.333329111337662e-3f;
    p = p * r2 + -0.1666666716337204f;
    const float s = r + (r * r2) * p;
    float q = -2.7207511266169604e-7f;
    q = q * r2 + 2.4799483071547e-5f;
    q = q * r2 + -1.3888883404433727e-3f;
    q = q * r2 + 4.16666679084301e-2f;
    const float c = (1.0f - 0.5f * r2) + (r2 * r2) * q;
    // quadrant k mod 4, in float (exact for |k| < 2^24; NaN and inf give 0)
    const float m = k - 4.0f * fr_floor(k * 0.25f);
    if (m == 1.0f) {
        *sn = c;
        *cs = -s;
    } else if (m == 2.0f) {
        *sn = -s;
        *cs = -c;
    } else if (m == 3.0f) {
        *sn = -c;
        *cs = s;
    } else {
        *sn = s;
        *cs = c;
    }
}

__host__ __device__ inline float fr_sin(float x)
{
    float s, c;
    fr_sincos(x, &s, &c);
    return s;
}

__host__ __device__ inline float fr_cos(float x)
{
    float s, c;
    fr_sincos(x, &s, &c);
    return c;
}

// GLSL fract: x - floor(x)
__host__ __device__ inline float fr_fract(float x) { return x - fr_floor(x); }

// random(p) = fract(sin(dot(p, vec2(12.9898, 78.233))) * 43758.5453123)  (compute.comp:81-83)
__host__ __device__ inline float fr_random(float px, float py)
{
    return fr_fract(fr_sin(px * 12.9898f + py * 78.233f) * 43758.5453123f);
}

// noise(p), compute.comp:88-103; mix(a, b, t) = a*(1-t) + b*t
__host__ __device__ inline float fr_noise(float px, float py)
{
    const float ix = fr_floor(px), iy = fr_floor(py);
    const float fx = fr_fract(px), fy = fr_fract(py);
    const float a = fr_random(ix, iy);
    const float b = fr_random(ix + 1.0f, iy + 0.0f);
    const float c = fr_random(ix + 0.0f, iy + 1.0f);
    const float d = fr_random(ix + 1.0f, iy + 1.0f);
    const float ux = (fx * fx) * (3.0f - 2.0f * fx);
    const float uy = (fy * fy) * (3.0f - 2.0f * fy);
    const float mix = a * (1.0f - ux) + b * ux;
    return (mix + ((c - a) * uy) * (1.0f - ux)) + ((d - b) * ux) * uy;
}

// fbm(p), six octaves, compute.comp:108-120
__host__ __device__ inline float fr_fbm(float px, float py)
{
    float value = 0.0f, amplitude = 0.5f;
    for (int i = 0; i < 6; ++i) {
        value = value + amplitude * fr_noise(px, py);
        px = px * 2.0f;
        py = py * 2.0f;
        amplitude = amplitude * 0.5f;
    }
    return value;
}

// Per-step constants of a forcing mode at totalTime t.
struct Forcing {
    unsigned mode;  // 0 none, 1 compute.comp:150, 2 compute.comp:151
    float amp;      // 10 (mode 1) or 7*fbm(vec2(sin t, cos t)) (mode 2)
    float a2;       // 2*sin(t*2)
    float t3;       // t*3
};

__host__ __device__ inline Forcing forcing_at(unsigned mode, float t)
{
    Forcing f{mode, 0.0f, 0.0f, 0.0f};
    if (mode == 0) return f;
    f.a2 = 2.0f * fr_sin(t * 2.0f);
    f.t3 = t * 3.0f;
    if (mode == 1) {
        f.amp = 10.0f;
    } else {
        float s, c;
        fr_sincos(t, &s, &c);
        f.amp = 7.0f * fr_fbm(s, c);
    }
    return f;
}

// The forcing added to the force at position (x, y, z): amp * vec3(...)
// (scalar on the left, as the shader writes it).
__host__ __device__ inline void forcing_term(const Forcing& f, float y, float z, float* out)
{
    const float sx = (f.a2 * fr_cos(y * 10.0f)) * fr_sin((y + 5.0f) * 15.0f);
    if (f.mode == 1) {
        const float cl = fminf(fmaxf(y * 2.0f, 0.2f), 2.0f);
        out[0] = f.amp * sx;
        out[1] = f.amp * 0.0f;
        out[2] = f.amp * -cl;
    } else {
        out[0] = f.amp * sx;
        out[1] = f.amp * (4.0f * fr_sin(z * 5.0f + f.t3));
        out[2] = f.amp * (-0.6f * (y + 3.0f));
    }
}

}  // namespace hs
