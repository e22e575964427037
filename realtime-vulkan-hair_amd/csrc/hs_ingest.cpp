// hs_ingest.cpp - groom ingestion of the reference demo (host code): the
// triangles of an OBJ mesh, hair roots sampled on them, initial strands.
//
// Reference: Strand.cpp:10-24 (random helpers), 26-146 (GeneratePointsOnMesh),
// 149-191 (Hair::Hair, initial strands); the OBJ is read by tinyobjloader
// 2.0 (src/tiny_obj_loader.h, vendored by the reference), whose number parsing
// and polygon triangulation decide the bits of every root, so the steps of
// both are kept here (in this file's own code):
//  * numbers: integer digits accumulated as m = m*10 + d in binary64,
//    fraction digits as m += d * 10^-k (10^-k from a table for k < 8, else
//    pow), exponent applied as ldexp(m * 5^e, e), then rounded to float;
//  * polygons: ear clipping in the plane of the two axes picked by the
//    first non-degenerate corner's normal, reflex corners (cross * area < 0)
//    and corners containing another vertex (crossing-number test) skipped;
//    a convex quad gives (0,1,2), (0,2,3).
#include <cmath>
#include <cfloat>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/hairsim.h"

namespace {

bool is_digit(char c) { return c >= '0' && c <= '9'; }

// One number token [s, e); false if it is not a number.
bool parse_number(const char* s, const char* e, double* out)
{
    if (s >= e) return false;
    double m = 0.0;
    int ex = 0;
    bool neg = false, lead_dot = false;
    const char* p = s;
    if (*p == '+' || *p == '-') {
        neg = *p == '-';
        ++p;
        if (p != e && *p == '.') lead_dot = true;
    } else if (*p == '.') {
        lead_dot = true;
    } else if (!is_digit(*p)) {
        return false;
    }
    if (!lead_dot) {
        int n = 0;
        while (p != e && is_digit(*p)) {
            m = m * 10;
            m = m + (double)(*p - '0');
            ++p;
            ++n;
        }
        if (n == 0) return false;
    }
    if (p != e && *p == '.') {
        ++p;
        static const double kTenths[8] = {1.0, 0.1, 0.01, 0.001, 0.0001, 0.00001, 0.000001, 0.0000001};
        int k = 1;
        while (p != e && is_digit(*p)) {
            m = m + (double)(*p - '0') * (k < 8 ? kTenths[k] : std::pow(10.0, -k));
            ++k;
            ++p;
        }
    }
    if (p != e && (*p == 'e' || *p == 'E')) {
        ++p;
        bool eneg = false;
        if (p != e && (*p == '+' || *p == '-')) {
            eneg = *p == '-';
            ++p;
        } else if (p == e || !is_digit(*p)) {
            return false;
        }
        int n = 0;
        while (p != e && is_digit(*p)) {
            if (ex < 100000) ex = ex * 10 + (*p - '0');  // beyond |e| = 1e5 the value is 0 or inf either way
            ++p;
            ++n;
        }
        if (n == 0) return false;
        if (eneg) ex = -ex;
    }
    *out = (neg ? -1.0 : 1.0) * (ex ? std::ldexp(m * std::pow(5.0, ex), ex) : m);
    return true;
}

// Next whitespace-separated token of the line.
bool next_token(const char*& p, const char* end, const char** ts, const char** te)
{
    while (p < end && (*p == ' ' || *p == '\t')) ++p;
    if (p >= end) return false;
    *ts = p;
    while (p < end && *p != ' ' && *p != '\t' && *p != '\r') ++p;
    *te = p;
    return true;
}

float read_real(const char*& p, const char* end)
{
    const char *s, *e;
    double v = 0.0;
    if (next_token(p, end, &s, &e)) parse_number(s, e, &v);
    return (float)v;
}

// OBJ index (1-based, negative: relative to the end) -> 0-based; false if invalid.
bool fix_index(long idx, size_t n, long* out)
{
    if (idx > 0) {
        *out = idx - 1;
        return true;
    }
    if (idx == 0) return false;
    *out = (long)n + idx;
    return *out >= 0;
}

struct Corner {
    long v, vn;
};

// Crossing-number point-in-polygon test (W. R. Franklin).
bool inside(int n, const float* vx, const float* vy, float tx, float ty)
{
    bool c = false;
    for (int i = 0, j = n - 1; i < n; j = i++)
        if (((vy[i] > ty) != (vy[j] > ty)) && (tx < (vx[j] - vx[i]) * (ty - vy[i]) / (vy[j] - vy[i]) + vx[i]))
            c = !c;
    return c;
}

// Triangles of one polygon by ear clipping (steps described at the top).
void triangulate(const std::vector<Corner>& poly, const std::vector<float>& V, std::vector<Corner>* out)
{
    const size_t n = poly.size();
    if (n < 3) return;
    auto coord = [&](long v, size_t axis) { return (size_t)v * 3 + axis < V.size() ? V[(size_t)v * 3 + axis] : 0.0f; };
    if (n == 3) {
        out->insert(out->end(), poly.begin(), poly.end());
        return;
    }
    size_t ax[2] = {1, 2};
    for (size_t k = 0; k < n; ++k) {
        const long a = poly[k].v, b = poly[(k + 1) % n].v, c = poly[(k + 2) % n].v;
        const float e0x = coord(b, 0) - coord(a, 0), e0y = coord(b, 1) - coord(a, 1), e0z = coord(b, 2) - coord(a, 2);
        const float e1x = coord(c, 0) - coord(b, 0), e1y = coord(c, 1) - coord(b, 1), e1z = coord(c, 2) - coord(b, 2);
        const float cx = std::fabs(e0y * e1z - e0z * e1y);
        const float cy = std::fabs(e0z * e1x - e0x * e1z);
        const float cz = std::fabs(e0x * e1y - e0y * e1x);
        if (cx > FLT_EPSILON || cy > FLT_EPSILON || cz > FLT_EPSILON) {
            if (!(cx > cy && cx > cz)) {
                ax[0] = 0;
                if (cz > cx && cz > cy) ax[1] = 1;
            }
            break;
        }
    }
    float area = 0.0f;
    for (size_t k = 0; k < n; ++k) {
        const long a = poly[k].v, b = poly[(k + 1) % n].v;
        area += (coord(a, ax[0]) * coord(b, ax[1]) - coord(a, ax[1]) * coord(b, ax[0])) * 0.5f;
    }
    std::vector<Corner> rem = poly;
    size_t guess = 0, iters = n, prev = n;
    float vx[3], vy[3];
    Corner tri[3];
    while (rem.size() > 3 && iters > 0) {
        const size_t np = rem.size();
        if (guess >= np) guess -= np;
        if (prev != np) {
            prev = np;
            iters = np;
        } else {
            --iters;
        }
        for (int k = 0; k < 3; ++k) {
            tri[k] = rem[(guess + k) % np];
            vx[k] = coord(tri[k].v, ax[0]);
            vy[k] = coord(tri[k].v, ax[1]);
        }
        const float cr = (vx[1] - vx[0]) * (vy[2] - vy[1]) - (vy[1] - vy[0]) * (vx[2] - vx[1]);
        if (cr * area < 0.0f) {
            ++guess;
            continue;
        }
        bool overlap = false;
        for (size_t o = 3; o < np; ++o) {
            const long ov = rem[(guess + o) % np].v;
            if (inside(3, vx, vy, coord(ov, ax[0]), coord(ov, ax[1]))) {
                overlap = true;
                break;
            }
        }
        if (overlap) {
            ++guess;
            continue;
        }
        out->insert(out->end(), tri, tri + 3);
        rem.erase(rem.begin() + (long)((guess + 1) % np));
    }
    if (rem.size() == 3) out->insert(out->end(), rem.begin(), rem.end());
}

// MSVC's rand(): x = x*214013 + 2531011 (mod 2^32), result bits 16..30.
struct MsvcRand {
    uint32_t x;
    int next()
    {
        x = x * 214013u + 2531011u;
        return (int)((x >> 16) & 0x7fffu);
    }
};
constexpr float kRandMax = 32767.0f;

}  // namespace

extern "C" {

int hs_obj_triangles(const char* path, float* positions, float* normals, uint32_t capacity, uint32_t* num_triangles)
{
    if (!path || !num_triangles) return HS_ERR_INVALID;
    FILE* f = std::fopen(path, "rb");
    if (!f) return HS_ERR_NOT_FOUND;
    std::string text;
    {
        char buf[1 << 16];
        size_t got;
        while ((got = std::fread(buf, 1, sizeof(buf), f)) > 0) text.append(buf, got);
    }
    std::fclose(f);
    std::vector<float> V, VN;
    std::vector<Corner> tris;
    bool bad = false;
    size_t at = 0;
    while (at < text.size() && !bad) {
        size_t nl = text.find('\n', at);
        if (nl == std::string::npos) nl = text.size();
        const char* p = text.data() + at;
        const char* end = text.data() + nl;
        at = nl + 1;
        while (end > p && end[-1] == '\r') --end;
        while (p < end && (*p == ' ' || *p == '\t')) ++p;
        if (end - p >= 2 && p[0] == 'v' && (p[1] == ' ' || p[1] == '\t')) {
            p += 2;
            for (int k = 0; k < 3; ++k) V.push_back(read_real(p, end));
        } else if (end - p >= 3 && p[0] == 'v' && p[1] == 'n' && (p[2] == ' ' || p[2] == '\t')) {
            p += 3;
            for (int k = 0; k < 3; ++k) VN.push_back(read_real(p, end));
        } else if (end - p >= 2 && p[0] == 'f' && (p[1] == ' ' || p[1] == '\t')) {
            p += 2;
            std::vector<Corner> poly;
            const char *ts, *te;
            while (next_token(p, end, &ts, &te)) {
                // v, v/vt, v//vn or v/vt/vn
                const std::string tok(ts, te);
                long parts[3] = {0, 0, 0};
                int np = 0;
                size_t q = 0;
                while (np < 3) {
                    const size_t slash = tok.find('/', q);
                    const std::string part = tok.substr(q, slash == std::string::npos ? std::string::npos : slash - q);
                    parts[np++] = part.empty() ? 0 : std::strtol(part.c_str(), nullptr, 10);
                    if (slash == std::string::npos) break;
                    q = slash + 1;
                }
                Corner c{-1, -1};
                if (!fix_index(parts[0], V.size() / 3, &c.v) || (np == 3 && parts[2] != 0 &&
                                                                  !fix_index(parts[2], VN.size() / 3, &c.vn))) {
                    bad = true;
                    break;
                }
                poly.push_back(c);
            }
            if (bad) break;
            triangulate(poly, V, &tris);
        }
    }
    if (bad) return HS_ERR_INVALID;
    const size_t nt = tris.size() / 3;
    for (const Corner& c : tris)
        if (c.v < 0 || (size_t)c.v >= V.size() / 3 || c.vn < 0 || (size_t)c.vn >= VN.size() / 3)
            return HS_ERR_INVALID;  // every corner needs a position and a normal
    if (nt > UINT32_MAX) return HS_ERR_INVALID;
    *num_triangles = (uint32_t)nt;
    if (!positions || !normals || capacity < nt) return HS_OK;
    for (size_t i = 0; i < tris.size(); ++i)
        for (int k = 0; k < 3; ++k) {
            positions[3 * i + k] = V[(size_t)tris[i].v * 3 + k];
            normals[3 * i + k] = VN[(size_t)tris[i].vn * 3 + k];
        }
    return HS_OK;
}

int hs_sample_roots(const float* positions, const float* normals, uint32_t num_triangles, uint32_t num_strands,
                    uint32_t seed, float* roots, float* root_normals)
{
    if (!positions || !normals || !roots || !root_normals || num_triangles == 0 || num_triangles > INT32_MAX)
        return HS_ERR_INVALID;
    MsvcRand r{seed};  // srand(seed) (Strand.cpp:88 uses 8)
    for (uint32_t i = 0; i < num_strands; ++i) {
        // Strand.cpp:89-107: triangle rand() % T, barycentrics rand()/RAND_MAX folded into the triangle
        const int t = r.next() % (int)num_triangles;
        float u = (float)r.next() / kRandMax;
        float v = (float)r.next() / kRandMax;
        if (u + v >= 1.0f) {
            u = 1.0f - u;
            v = 1.0f - v;
        }
        const float* p = positions + (size_t)t * 9;
        const float w = (1.0f - u) - v;
        for (int k = 0; k < 3; ++k) {
            roots[3 * (size_t)i + k] = (p[k] * u + p[3 + k] * v) + p[6 + k] * w;
            root_normals[3 * (size_t)i + k] = normals[(size_t)t * 9 + k];  // the first corner's normal
        }
    }
    return HS_OK;
}

int hs_initial_strands(const float* roots, const float* root_normals, uint32_t num_strands, uint32_t num_points,
                       float length, float* strands_aos)
{
    if (!roots || !root_normals || !strands_aos || num_points < 2) return HS_ERR_INVALID;
    const uint32_t N = num_points;
    // Strand.cpp:160-172: points along the root normal bent by (+0.05, +5, -2),
    // spaced (float)(length / (N - 1.0)) times that direction (computed in
    // binary64 as the C++ source does); velocities (0, 0, -1); no corrections
    const float step = (float)((double)length / ((double)N - 1.0));
    for (uint32_t s = 0; s < num_strands; ++s) {
        float* o = strands_aos + (size_t)s * 12 * N;
        float cp[3] = {roots[3 * (size_t)s], roots[3 * (size_t)s + 1], roots[3 * (size_t)s + 2]};
        float dir[3] = {root_normals[3 * (size_t)s], root_normals[3 * (size_t)s + 1], root_normals[3 * (size_t)s + 2]};
        dir[2] = dir[2] - 2.0f;
        dir[1] = dir[1] + 5.0f;
        dir[0] = dir[0] + 0.05f;
        for (uint32_t j = 0; j < N; ++j) {
            float* P = o + 4 * j;
            float* Vv = o + 4 * (N + j);
            float* Cc = o + 4 * (2 * N + j);
            P[0] = cp[0];
            P[1] = cp[1];
            P[2] = cp[2];
            P[3] = 1.0f;
            Vv[0] = 0.0f;
            Vv[1] = 0.0f;
            Vv[2] = -1.0f;
            Vv[3] = 0.0f;
            Cc[0] = Cc[1] = Cc[2] = Cc[3] = 0.0f;
            for (int k = 0; k < 3; ++k) cp[k] = cp[k] + step * dir[k];
        }
    }
    return HS_OK;
}

}  // extern "C"
