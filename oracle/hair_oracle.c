/*
 * hair_oracle.c - CPU oracle (TEST INFRASTRUCTURE, never the product).
 * Instantiates the restatement of src/shaders/compute.comp:124-303 in
 * binary32 ("ref32") and binary64 ("oracle64").  Must be compiled with
 * -ffp-contract=off and without -ffast-math; see hair_oracle.h.
 */
#include "hair_oracle.h"

#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#ifdef __FAST_MATH__
#error "the oracle must not be built with -ffast-math"
#endif

/* sin and cos for the forcing (compute.comp:150-151 use GLSL sin/cos, whose
 * precision the language leaves open): k = nearest multiple of pi/2 as
 * floor(x*(2/pi) + 0.5); r = ((x - k*P1) - k*P2) - k*P3 with P1 = 201/128
 * and P2 of 8 significant bits (k*P1, k*P2 exact for |k| < 2^16) and P3 the
 * float remainder of pi/2; sin r = r + r^3 (S1 + r^2 (S2 + r^2 (S3 + r^2 S4))),
 * cos r = (1 - r^2/2) + r^4 (C1 + r^2 (C2 + r^2 (C3 + r^2 C4))), coefficients
 * fitted to the Taylor remainders and rounded to float; then the quadrant
 * k mod 4 (computed in float: NaN and inf land in quadrant 0). */
static void ho_sincosf(float x, float* sn, float* cs)
{
    const float k = floorf(x * 0.636619772367581343f + 0.5f);
    const float r = ((x - k * 1.5703125f) - k * 4.84466552734375e-4f) - k * -6.397578431460715e-7f;
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
    const float m = k - 4.0f * floorf(k * 0.25f);
    if (m == 1.0f) { *sn = c; *cs = -s; }
    else if (m == 2.0f) { *sn = -s; *cs = -c; }
    else if (m == 3.0f) { *sn = -c; *cs = s; }
    else { *sn = s; *cs = c; }
}

/* mix(a, b, t) = a*(1-t) + b*t per component (GLSL) */
static void ho_mix3(const float* a, const float* b, float t, float* o)
{
    const float u = 1.0f - t;
    for (int k = 0; k < 3; ++k) o[k] = a[k] * u + b[k] * t;
}

void ho_resample_f32(const float* strands, int S, int N, int M, float* out)
{
    const float third = 1.0f / 3.0f;
    for (int s = 0; s < S; ++s) {
        const float* P = strands + (size_t)s * 12 * N; /* curvePoints[N] (vec4) */
        for (int j = 0; j < M; ++j) {
            const float v = (float)j / (float)(M - 1);
            const float x = v * (float)(N - 1);
            const float fl = floorf(x);
            int seg = (int)fl;
            float t = x - fl;
            if (seg >= N - 1) { seg = N - 2; t = 1.0f; }
            float v0[3], v1[3], v2[3], v3[3], b0[3], b1[3], b2[3], b3[3];
            for (int k = 0; k < 3; ++k) {
                v1[k] = P[4 * seg + k];
                v2[k] = P[4 * (seg + 1) + k];
            }
            for (int k = 0; k < 3; ++k) {
                v0[k] = seg == 0 ? v1[k] + (v1[k] - v2[k]) : P[4 * (seg - 1) + k];
                v3[k] = seg + 2 == N ? v2[k] + (v2[k] - v1[k]) : P[4 * (seg + 2) + k];
            }
            for (int k = 0; k < 3; ++k) {
                b0[k] = v1[k];
                b1[k] = v1[k] + third * ((v2[k] - v0[k]) / 2.0f);
                b2[k] = v2[k] - third * ((v3[k] - v1[k]) / 2.0f);
                b3[k] = v2[k];
            }
            float b01[3], b11[3], b21[3], b02[3], b12[3], c[3];
            ho_mix3(b0, b1, t, b01);
            ho_mix3(b1, b2, t, b11);
            ho_mix3(b2, b3, t, b21);
            ho_mix3(b01, b11, t, b02);
            ho_mix3(b11, b21, t, b12);
            ho_mix3(b02, b12, t, c);
            float* o = out + ((size_t)s * M + j) * 4;
            o[0] = c[0];
            o[1] = c[1];
            o[2] = c[2];
            o[3] = 1.0f;
        }
    }
}

float ho_sinf(float x) { float s, c; ho_sincosf(x, &s, &c); return s; }
float ho_cosf(float x) { float s, c; ho_sincosf(x, &s, &c); return c; }

#define R float
#define SFX _f32
#define FORCING_FBM(t) fbm_f32(ho_sinf(t), ho_cosf(t))
#define SQRT sqrtf
#define SIN ho_sinf
#define COS ho_cosf
#define FLOOR floorf
#define FABS fabsf
#define FMIN fminf
#define FMAX fmaxf
#include "hair_oracle_body.h"
#undef R
#undef SFX
#undef SQRT
#undef SIN
#undef COS
#undef FLOOR
#undef FABS
#undef FMIN
#undef FMAX

float ho_fbmf(float x, float y) { return fbm_f32(x, y); }

#undef FORCING_FBM
#define R double
#define SFX _f64
#define FORCING_FBM(t) fbm_f32(ho_sinf((float)(t)), ho_cosf((float)(t)))
#define SQRT sqrt
#define SIN sin
#define COS cos
#define FLOOR floor
#define FABS fabs
#define FMIN fmin
#define FMAX fmax
#include "hair_oracle_body.h"
#undef R
#undef SFX
#undef SQRT
#undef FLOOR
#undef FABS
#undef FMIN
#undef FMAX

void ho_step_f32(const ho_params* p, float* strands, int S, int N, int32_t* grid)
{
    step_impl_f32(p, strands, S, N, grid, 1);
}

void ho_step_f64(const ho_params* p, double* strands, int S, int N, int32_t* grid)
{
    step_impl_f64(p, strands, S, N, grid, 1);
}

void ho_step_f32_mt(const ho_params* p, float* strands, int S, int N, int32_t* grid, int threads)
{
    step_impl_f32(p, strands, S, N, grid, threads);
}

/* ---- mesh preparation (binary64 normals, rounded to float) ---- */
typedef struct { double x, y, z; } d3;
static d3 d3_sub(d3 a, d3 b) { d3 r = {a.x - b.x, a.y - b.y, a.z - b.z}; return r; }
static d3 d3_add(d3 a, d3 b) { d3 r = {a.x + b.x, a.y + b.y, a.z + b.z}; return r; }
static d3 d3_mul(d3 a, double s) { d3 r = {a.x * s, a.y * s, a.z * s}; return r; }
static double d3_dot(d3 a, d3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static d3 d3_cross(d3 a, d3 b)
{
    d3 r = {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
    return r;
}
static d3 d3_unit(d3 a)
{
    double l = sqrt(d3_dot(a, a));
    if (l > 0) return d3_mul(a, 1.0 / l);
    d3 z = {0, 0, 0};
    return z;
}

typedef struct { uint32_t lo, hi; int face; } edge_rec;
static int edge_cmp(const void* pa, const void* pb)
{
    const edge_rec* a = (const edge_rec*)pa;
    const edge_rec* b = (const edge_rec*)pb;
    if (a->lo != b->lo) return a->lo < b->lo ? -1 : 1;
    if (a->hi != b->hi) return a->hi < b->hi ? -1 : 1;
    return (a->face > b->face) - (a->face < b->face);
}

int ho_mesh_prepare(const float* v, int nv, const uint32_t* t, int nt, float* out)
{
    if (nv < 4 || nt < 4) return -1;
    for (long k = 0; k < 3L * nt; ++k)
        if (t[k] >= (uint32_t)nv) return -1;
    d3* fn = (d3*)malloc(sizeof(d3) * (size_t)nt);
    d3* vn = (d3*)calloc((size_t)nv, sizeof(d3));
    edge_rec* e = (edge_rec*)malloc(sizeof(edge_rec) * 3 * (size_t)nt);
    int* other = (int*)malloc(sizeof(int) * 3 * (size_t)nt); /* [face][k]: the face across edge k */
    int rc = 0;
    for (int f = 0; f < nt; ++f) {
        d3 p[3];
        for (int k = 0; k < 3; ++k) {
            const float* q = v + 3 * t[3 * f + k];
            d3 w = {q[0], q[1], q[2]};
            p[k] = w;
        }
        fn[f] = d3_unit(d3_cross(d3_sub(p[1], p[0]), d3_sub(p[2], p[0])));
        for (int k = 0; k < 3; ++k) {
            d3 e1 = d3_unit(d3_sub(p[(k + 1) % 3], p[k]));
            d3 e2 = d3_unit(d3_sub(p[(k + 2) % 3], p[k]));
            double c = d3_dot(e1, e2);
            double ang = acos(c < -1.0 ? -1.0 : (c > 1.0 ? 1.0 : c));
            uint32_t vi = t[3 * f + k];
            vn[vi] = d3_add(vn[vi], d3_mul(fn[f], ang));
            uint32_t a = t[3 * f + k], b = t[3 * f + (k + 1) % 3];
            edge_rec r = {a < b ? a : b, a < b ? b : a, 3 * f + k};
            e[3 * f + k] = r;
        }
    }
    qsort(e, 3 * (size_t)nt, sizeof(edge_rec), edge_cmp);
    for (long i = 0; i < 3L * nt && rc == 0; i += 2) {
        if (i + 1 >= 3L * nt || e[i].lo != e[i + 1].lo || e[i].hi != e[i + 1].hi) { rc = -1; break; }
        if (i + 2 < 3L * nt && e[i + 2].lo == e[i].lo && e[i + 2].hi == e[i].hi) { rc = -1; break; }
        other[e[i].face] = e[i + 1].face / 3;
        other[e[i + 1].face] = e[i].face / 3;
    }
    for (int f = 0; f < nt && rc == 0; ++f) {
        float* r = out + (size_t)f * HO_MESH_TRI_FLOATS;
        memset(r, 0, sizeof(float) * HO_MESH_TRI_FLOATS);
        for (int k = 0; k < 3; ++k) {
            const float* q = v + 3 * t[3 * f + k];
            r[4 * k] = q[0];
            r[4 * k + 1] = q[1];
            r[4 * k + 2] = q[2];
        }
        r[12] = (float)fn[f].x;
        r[13] = (float)fn[f].y;
        r[14] = (float)fn[f].z;
        for (int k = 0; k < 3; ++k) {
            int g = other[3 * f + k];
            /* the two adjacent faces, lower index first */
            int f0 = f < g ? f : g, f1 = f < g ? g : f;
            d3 en = d3_unit(d3_add(fn[f0], fn[f1]));
            r[16 + 4 * k] = (float)en.x;
            r[16 + 4 * k + 1] = (float)en.y;
            r[16 + 4 * k + 2] = (float)en.z;
            d3 w = d3_unit(vn[t[3 * f + k]]);
            r[28 + 4 * k] = (float)w.x;
            r[28 + 4 * k + 1] = (float)w.y;
            r[28 + 4 * k + 2] = (float)w.z;
        }
    }
    free(fn);
    free(vn);
    free(e);
    free(other);
    return rc;
}
