// hs_mesh_build.h - host preprocessing of a closed triangle mesh collider:
// per-triangle records (vertices, face/edge/vertex pseudo-normals) and a
// uniform grid of triangle lists, uploaded by hs_mesh_create.
//
// The reference has no mesh collider (compute.comp:155-184 handles a sphere
// and ellipsoids only); the contact rule this serves is stated with the
// kernel (hs_kernels.hip, mesh_contact).
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <utility>
#include <vector>

namespace hsmesh {

// One triangle: vertices a, b, c; the face normal; the pseudo-normals of
// edges ab, bc, ca; the pseudo-normals of vertices a, b, c (angle-weighted,
// Baerentzen & Aanaes 2005).  10 float4 (w unused) = 160 B.
constexpr int kTriFloats = 40;

// One entry of a cell's candidate list, 48 B: (a, key), (b, triangle id as
// float bits), (c, bounding radius).  The vertices are copies of the triangle
// record's, so the kernel's loads of a round do not wait on an id; the
// bounding sphere's centre is not stored: the kernel forms it from the
// vertices with the same binary32 operations as sphere_centre below.
constexpr int kRecFloats = 12;

// Centre of a triangle's bounding sphere: ((a + b) + c) * (1/3) per axis in
// binary32, the expression the mesh kernel evaluates (mesh_contact), so host
// and device hold the same bits.
inline float sphere_centre(float a, float b, float c)
{
    const float third = 1.0f / 3.0f;
    const float s = a + b;
    return (s + c) * third;
}

struct Built {
    std::vector<float> tris;          // [T][kTriFloats]
    std::vector<float> spheres;       // [T][4]: centroid, radius (rounded outward) of each triangle
    std::vector<uint32_t> cell_start; // [cells + 1]
    std::vector<float> cell_recs;     // [entries][kRecFloats], per cell nearest bounding sphere first;
                                      // key = |cell centre - sphere centre| - radius, rounded down
    float half_diag = 0.0f;           // half the cell diagonal, rounded up
    float origin[3] = {0, 0, 0};
    float cell[3] = {1, 1, 1};
    float inv_cell[3] = {1, 1, 1};
    uint32_t dim[3] = {1, 1, 1};
    float margin = 0.0f;
};

struct D3 {
    double x, y, z;
};
inline D3 sub(D3 a, D3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline D3 add(D3 a, D3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline D3 mul(D3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
inline double dot(D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline D3 cross(D3 a, D3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline D3 unit(D3 a)
{
    const double l = std::sqrt(dot(a, a));
    return l > 0 ? mul(a, 1.0 / l) : D3{0, 0, 0};
}

// Returns false for a mesh that is not a closed 2-manifold (an edge not
// shared by exactly two triangles), has an index out of range or a
// non-finite vertex coordinate.
inline bool build(const float* v, uint32_t nv, const uint32_t* t, uint32_t nt, const uint32_t dims[3], float margin,
                  Built* out)
{
    if (nv < 4 || nt < 4) return false;
    for (uint64_t k = 0; k < 3ull * nt; ++k)
        if (t[k] >= nv) return false;
    for (uint64_t k = 0; k < 3ull * nv; ++k)
        if (!std::isfinite(v[k])) return false;
    auto V = [&](uint32_t i) { return D3{v[3 * i], v[3 * i + 1], v[3 * i + 2]}; };
    std::vector<D3> fn(nt), vn(nv, D3{0, 0, 0});
    std::map<std::pair<uint32_t, uint32_t>, std::vector<uint32_t>> edges;
    for (uint32_t f = 0; f < nt; ++f) {
        const uint32_t i[3] = {t[3 * f], t[3 * f + 1], t[3 * f + 2]};
        const D3 p[3] = {V(i[0]), V(i[1]), V(i[2])};
        fn[f] = unit(cross(sub(p[1], p[0]), sub(p[2], p[0])));
        for (int k = 0; k < 3; ++k) {
            const D3 e1 = unit(sub(p[(k + 1) % 3], p[k]));
            const D3 e2 = unit(sub(p[(k + 2) % 3], p[k]));
            const double ang = std::acos(std::max(-1.0, std::min(1.0, dot(e1, e2))));
            vn[i[k]] = add(vn[i[k]], mul(fn[f], ang));
            const uint32_t a = i[k], b = i[(k + 1) % 3];
            edges[{std::min(a, b), std::max(a, b)}].push_back(f);
        }
    }
    for (auto& e : edges)
        if (e.second.size() != 2) return false;
    Built& B = *out;
    B.tris.assign((size_t)nt * kTriFloats, 0.0f);
    D3 lo{1e300, 1e300, 1e300}, hi{-1e300, -1e300, -1e300};
    for (uint32_t f = 0; f < nt; ++f) {
        const uint32_t i[3] = {t[3 * f], t[3 * f + 1], t[3 * f + 2]};
        float* r = &B.tris[(size_t)f * kTriFloats];
        auto put = [&](int slot, D3 x) {
            r[4 * slot] = (float)x.x;
            r[4 * slot + 1] = (float)x.y;
            r[4 * slot + 2] = (float)x.z;
        };
        for (int k = 0; k < 3; ++k) {
            const D3 p = V(i[k]);
            put(k, p);
            lo = {std::min(lo.x, p.x), std::min(lo.y, p.y), std::min(lo.z, p.z)};
            hi = {std::max(hi.x, p.x), std::max(hi.y, p.y), std::max(hi.z, p.z)};
        }
        put(3, fn[f]);
        for (int k = 0; k < 3; ++k) {
            const uint32_t a = i[k], b = i[(k + 1) % 3];
            const auto& adj = edges[{std::min(a, b), std::max(a, b)}];
            put(4 + k, unit(add(fn[adj[0]], fn[adj[1]])));
            put(7 + k, unit(vn[i[k]]));
        }
    }
    // bounding sphere of each triangle about its centroid; the radius is
    // padded so the float sphere contains the float triangle
    B.spheres.assign((size_t)nt * 4, 0.0f);
    for (uint32_t f = 0; f < nt; ++f) {
        const float* r = &B.tris[(size_t)f * kTriFloats];
        const D3 a{r[0], r[1], r[2]}, b{r[4], r[5], r[6]}, c{r[8], r[9], r[10]};
        const float cf[3] = {sphere_centre(r[0], r[4], r[8]), sphere_centre(r[1], r[5], r[9]),
                             sphere_centre(r[2], r[6], r[10])};
        const D3 cq{cf[0], cf[1], cf[2]};
        const double rad = std::max({std::sqrt(dot(sub(a, cq), sub(a, cq))), std::sqrt(dot(sub(b, cq), sub(b, cq))),
                                     std::sqrt(dot(sub(c, cq), sub(c, cq)))});
        const double mag = std::max({std::fabs(cq.x), std::fabs(cq.y), std::fabs(cq.z), 1.0});
        float* sp = &B.spheres[(size_t)f * 4];
        sp[0] = cf[0];
        sp[1] = cf[1];
        sp[2] = cf[2];
        sp[3] = (float)(rad * (1.0 + 1e-5) + 1e-6 * mag);
    }
    // grid over the mesh bounds grown by the margin
    const double m = margin;
    lo = sub(lo, D3{m, m, m});
    hi = add(hi, D3{m, m, m});
    const double ext[3] = {hi.x - lo.x, hi.y - lo.y, hi.z - lo.z};
    for (int k = 0; k < 3; ++k) {
        B.dim[k] = std::max<uint32_t>(1, dims[k]);
        B.cell[k] = (float)(ext[k] / B.dim[k]);
        B.inv_cell[k] = 1.0f / B.cell[k];
    }
    B.origin[0] = (float)lo.x;
    B.origin[1] = (float)lo.y;
    B.origin[2] = (float)lo.z;
    B.margin = margin;
    const uint64_t ncell = (uint64_t)B.dim[0] * B.dim[1] * B.dim[2];
    std::vector<std::vector<uint32_t>> lists(ncell);
    // a cell lists every triangle whose bounds grown by a bit more than the
    // margin reach it (the extra absorbs float rounding of the point's cell)
    const double reach = m * 1.001 + 1e-6 + 1e-5 * std::max({ext[0], ext[1], ext[2]}) / 1024.0;
    for (uint32_t f = 0; f < nt; ++f) {
        const float* r = &B.tris[(size_t)f * kTriFloats];
        int c0[3], c1[3];
        for (int k = 0; k < 3; ++k) {
            const double tmin = std::min({(double)r[k], (double)r[4 + k], (double)r[8 + k]}) - reach;
            const double tmax = std::max({(double)r[k], (double)r[4 + k], (double)r[8 + k]}) + reach;
            c0[k] = std::max(0, (int)std::floor((tmin - B.origin[k]) / B.cell[k]));
            c1[k] = std::min((int)B.dim[k] - 1, (int)std::floor((tmax - B.origin[k]) / B.cell[k]));
        }
        for (int z = c0[2]; z <= c1[2]; ++z)
            for (int y = c0[1]; y <= c1[1]; ++y)
                for (int x = c0[0]; x <= c1[0]; ++x)
                    lists[((uint64_t)z * B.dim[1] + y) * B.dim[0] + x].push_back(f);
    }
    // Cells entirely outside the mesh need no list: the contact rule requires
    // the point to be inside.  A cell is "surface" when a triangle's bounds,
    // grown by a small epsilon, reach it; cells joined to the grid's border by
    // non-surface cells hold no surface point and lie outside (crossing into
    // the mesh means crossing its surface).  The epsilon keeps every point of
    // such a cell well away from the surface, where the float inside test is
    // unambiguous.
    {
        std::vector<uint8_t> surf(ncell, 0), outside(ncell, 0);
        for (uint32_t f = 0; f < nt; ++f) {
            const float* r = &B.tris[(size_t)f * kTriFloats];
            int c0[3], c1[3];
            for (int k = 0; k < 3; ++k) {
                const double lo = std::min({(double)r[k], (double)r[4 + k], (double)r[8 + k]});
                const double hi = std::max({(double)r[k], (double)r[4 + k], (double)r[8 + k]});
                const double eps = 1e-4 * (1.0 + std::max(std::fabs(lo), std::fabs(hi))) + 0.01 * B.cell[k];
                c0[k] = std::max(0, (int)std::floor((lo - eps - B.origin[k]) / B.cell[k]));
                c1[k] = std::min((int)B.dim[k] - 1, (int)std::floor((hi + eps - B.origin[k]) / B.cell[k]));
            }
            for (int z = c0[2]; z <= c1[2]; ++z)
                for (int y = c0[1]; y <= c1[1]; ++y)
                    for (int x = c0[0]; x <= c1[0]; ++x) surf[((uint64_t)z * B.dim[1] + y) * B.dim[0] + x] = 1;
        }
        std::vector<uint64_t> stack;
        auto seed = [&](uint64_t c) {
            if (!surf[c] && !outside[c]) {
                outside[c] = 1;
                stack.push_back(c);
            }
        };
        const uint32_t X = B.dim[0], Y = B.dim[1], Z = B.dim[2];
        for (uint32_t z = 0; z < Z; ++z)
            for (uint32_t y = 0; y < Y; ++y)
                for (uint32_t x = 0; x < X; ++x)
                    if (x == 0 || y == 0 || z == 0 || x == X - 1 || y == Y - 1 || z == Z - 1)
                        seed(((uint64_t)z * Y + y) * X + x);
        while (!stack.empty()) {
            const uint64_t c = stack.back();
            stack.pop_back();
            const uint32_t x = (uint32_t)(c % X), y = (uint32_t)((c / X) % Y), z = (uint32_t)(c / ((uint64_t)X * Y));
            if (x > 0) seed(c - 1);
            if (x + 1 < X) seed(c + 1);
            if (y > 0) seed(c - X);
            if (y + 1 < Y) seed(c + X);
            if (z > 0) seed(c - (uint64_t)X * Y);
            if (z + 1 < Z) seed(c + (uint64_t)X * Y);
        }
        for (uint64_t c = 0; c < ncell; ++c)
            if (outside[c]) lists[c].clear();
    }
    // each cell's list sorted by a lower bound of the distance from the cell to
    // the triangle (centre distance minus radius): the kernel stops once the
    // bound minus the cell's half diagonal passes the best distance so far
    const double hd = 0.5 * std::sqrt((double)B.cell[0] * B.cell[0] + (double)B.cell[1] * B.cell[1] +
                                      (double)B.cell[2] * B.cell[2]);
    B.half_diag = (float)(hd * (1.0 + 1e-5) + 1e-6);
    std::vector<std::pair<float, uint32_t>> keyed;
    B.cell_start.assign(ncell + 1, 0);
    B.cell_recs.clear();
    uint32_t entries = 0;
    for (uint64_t cidx = 0; cidx < ncell; ++cidx) {
        const uint64_t x = cidx % B.dim[0], y = (cidx / B.dim[0]) % B.dim[1], z = cidx / ((uint64_t)B.dim[0] * B.dim[1]);
        const D3 cc{B.origin[0] + (x + 0.5) * B.cell[0], B.origin[1] + (y + 0.5) * B.cell[1],
                    B.origin[2] + (z + 0.5) * B.cell[2]};
        keyed.clear();
        for (uint32_t f : lists[cidx]) {
            const float* sp = &B.spheres[(size_t)f * 4];
            const D3 d{cc.x - sp[0], cc.y - sp[1], cc.z - sp[2]};
            const double key = std::sqrt(dot(d, d)) - sp[3];
            // rounded down (and a little more) so the float key never exceeds the true bound
            keyed.emplace_back((float)(key - 1e-6 * (1.0 + std::fabs(key))), f);
        }
        std::sort(keyed.begin(), keyed.end());
        for (auto& kf : keyed) {
            const float* r = &B.tris[(size_t)kf.second * kTriFloats];
            const float* sp = &B.spheres[(size_t)kf.second * 4];
            float id;
            std::memcpy(&id, &kf.second, 4);
            const float rec[kRecFloats] = {r[0], r[1], r[2], kf.first, r[4], r[5], r[6], id, r[8], r[9], r[10], sp[3]};
            B.cell_recs.insert(B.cell_recs.end(), rec, rec + kRecFloats);
        }
        entries += (uint32_t)keyed.size();
        B.cell_start[cidx + 1] = entries;
    }
    return true;
}

}  // namespace hsmesh
