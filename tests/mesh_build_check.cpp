// Property check of the host mesh preprocessing (realtime-vulkan-hair_amd/csrc/hs_mesh_build.h),
// built with g++ by tests/test_mesh_build.py.  Reads a mesh from a binary file
// (u32 nv, u32 nt, u32 dims[3], f32 margin, f32 verts[nv][3], u32 tris[nt][3]),
// builds it, and checks what the mesh kernel relies on:
//   * each candidate record holds bit copies of its triangle's vertices and a
//     valid triangle id;
//   * the record's radius bounds every vertex's distance from the centre the
//     kernel forms, ((a + b) + c) * (1/3) in binary32 (the sphere test is a
//     conservative reject);
//   * keys ascend within a cell and never exceed the distance from the cell's
//     centre to the bounding sphere (the walk's early exit is conservative);
//   * every triangle within the margin of a sampled point is in the point's
//     cell list, unless that cell lies outside the mesh (lists cleared);
//   * the record count matches the cell starts.
// Prints "ok <entries> <sampled points>" and exits 0, or prints the first
// failures and exits 1.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <vector>

#include "hs_mesh_build.h"

namespace {

struct V3 {
    double x, y, z;
};
V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
V3 operator*(V3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

// Squared distance from p to triangle abc (closest point by region tests).
double tri_dist2(V3 p, V3 a, V3 b, V3 c)
{
    const V3 ab = b - a, ac = c - a, ap = p - a;
    const double d1 = dot(ab, ap), d2 = dot(ac, ap);
    if (d1 <= 0 && d2 <= 0) return dot(ap, ap);
    const V3 bp = p - b;
    const double d3 = dot(ab, bp), d4 = dot(ac, bp);
    if (d3 >= 0 && d4 <= d3) return dot(bp, bp);
    const double vc = d1 * d4 - d3 * d2;
    if (vc <= 0 && d1 >= 0 && d3 <= 0) {
        const V3 q = a + ab * (d1 / (d1 - d3));
        return dot(p - q, p - q);
    }
    const V3 cp = p - c;
    const double d5 = dot(ab, cp), d6 = dot(ac, cp);
    if (d6 >= 0 && d5 <= d6) return dot(cp, cp);
    const double vb = d5 * d2 - d1 * d6;
    if (vb <= 0 && d2 >= 0 && d6 <= 0) {
        const V3 q = a + ac * (d2 / (d2 - d6));
        return dot(p - q, p - q);
    }
    const double va = d3 * d6 - d5 * d4;
    if (va <= 0 && (d4 - d3) >= 0 && (d5 - d6) >= 0) {
        const V3 q = b + (c - b) * ((d4 - d3) / ((d4 - d3) + (d5 - d6)));
        return dot(p - q, p - q);
    }
    const double den = 1.0 / (va + vb + vc);
    const V3 q = a + ab * (vb * den) + ac * (vc * den);
    return dot(p - q, p - q);
}

int fails = 0;
void fail(const char* what, long a, long b)
{
    if (fails++ < 10) std::printf("FAIL %s (%ld, %ld)\n", what, a, b);
}

}  // namespace

int main(int argc, char** argv)
{
    if (argc != 3) return 2;
    FILE* f = std::fopen(argv[1], "rb");
    if (!f) return 2;
    uint32_t nv = 0, nt = 0, dims[3];
    float margin = 0;
    bool rd = std::fread(&nv, 4, 1, f) == 1 && std::fread(&nt, 4, 1, f) == 1 && std::fread(dims, 4, 3, f) == 3 &&
              std::fread(&margin, 4, 1, f) == 1;
    std::vector<float> verts(3ull * nv);
    std::vector<uint32_t> tris(3ull * nt);
    rd = rd && std::fread(verts.data(), 4, verts.size(), f) == verts.size() &&
         std::fread(tris.data(), 4, tris.size(), f) == tris.size();
    std::fclose(f);
    if (!rd) return 2;
    const int samples = std::atoi(argv[2]);

    hsmesh::Built B;
    if (!hsmesh::build(verts.data(), nv, tris.data(), nt, dims, margin, &B)) {
        std::printf("FAIL build rejected the mesh\n");
        return 1;
    }
    const uint64_t ncell = (uint64_t)B.dim[0] * B.dim[1] * B.dim[2];
    const uint64_t entries = B.cell_start[ncell];
    if (B.cell_recs.size() != entries * hsmesh::kRecFloats) fail("record count", (long)B.cell_recs.size(), (long)entries);

    for (uint64_t c = 0; c < ncell; ++c) {
        const uint64_t x = c % B.dim[0], y = (c / B.dim[0]) % B.dim[1], z = c / ((uint64_t)B.dim[0] * B.dim[1]);
        const V3 cc{B.origin[0] + (x + 0.5) * (double)B.cell[0], B.origin[1] + (y + 0.5) * (double)B.cell[1],
                    B.origin[2] + (z + 0.5) * (double)B.cell[2]};
        for (uint32_t e = B.cell_start[c]; e < B.cell_start[c + 1]; ++e) {
            const float* r = &B.cell_recs[(size_t)e * hsmesh::kRecFloats];
            uint32_t id;
            std::memcpy(&id, &r[7], 4);
            if (id >= nt) {
                fail("triangle id", (long)e, (long)id);
                continue;
            }
            const float* t = &B.tris[(size_t)id * hsmesh::kTriFloats];
            for (int v = 0; v < 3; ++v)
                for (int k = 0; k < 3; ++k)
                    if (std::memcmp(&r[4 * v + k], &t[4 * v + k], 4) != 0) fail("vertex copy", (long)e, v * 3 + k);
            // the centre as the kernel forms it
            const float third = 1.0f / 3.0f;
            float cen[3];
            for (int k = 0; k < 3; ++k) {
                volatile float s = r[k] + r[4 + k];
                volatile float s2 = s + r[8 + k];
                cen[k] = s2 * third;
            }
            const V3 cv{cen[0], cen[1], cen[2]};
            const double rad = r[11];
            for (int v = 0; v < 3; ++v) {
                const V3 p{r[4 * v], r[4 * v + 1], r[4 * v + 2]};
                if (std::sqrt(dot(p - cv, p - cv)) > rad) fail("vertex outside its bounding sphere", (long)e, v);
            }
            const double bound = std::sqrt(dot(cc - cv, cc - cv)) - rad;
            if ((double)r[3] > bound) fail("key above the cell-centre bound", (long)e, (long)c);
            if (e > B.cell_start[c] && r[3] < B.cell_recs[(size_t)(e - 1) * hsmesh::kRecFloats + 3])
                fail("keys not ascending", (long)e, (long)c);
        }
    }

    // completeness at sampled points (cells whose list was cleared lie outside
    // the mesh and are skipped)
    std::mt19937 rng(1234);
    std::uniform_real_distribution<double> u(0.0, 1.0);
    const double m2 = (double)margin * margin;
    int checked = 0;
    for (int s = 0; s < samples; ++s) {
        // points near the surface: a random vertex plus an offset within 2 margins
        const uint32_t vi = (uint32_t)(u(rng) * nv) % nv;
        const float px = verts[3 * vi] + (float)((u(rng) * 2 - 1) * 2 * margin);
        const float py = verts[3 * vi + 1] + (float)((u(rng) * 2 - 1) * 2 * margin);
        const float pz = verts[3 * vi + 2] + (float)((u(rng) * 2 - 1) * 2 * margin);
        // the kernel's cell arithmetic
        const float gx = (px - B.origin[0]) * B.inv_cell[0], gy = (py - B.origin[1]) * B.inv_cell[1],
                    gz = (pz - B.origin[2]) * B.inv_cell[2];
        if (!(gx >= 0 && gx < (float)B.dim[0] && gy >= 0 && gy < (float)B.dim[1] && gz >= 0 && gz < (float)B.dim[2]))
            continue;
        const uint32_t cx = std::min((uint32_t)gx, B.dim[0] - 1), cy = std::min((uint32_t)gy, B.dim[1] - 1),
                       cz = std::min((uint32_t)gz, B.dim[2] - 1);
        const uint64_t c = ((uint64_t)cz * B.dim[1] + cy) * B.dim[0] + cx;
        const uint32_t i0 = B.cell_start[c], i1 = B.cell_start[c + 1];
        if (i1 == i0) continue;
        ++checked;
        std::vector<char> listed(nt, 0);
        for (uint32_t e = i0; e < i1; ++e) {
            uint32_t id;
            std::memcpy(&id, &B.cell_recs[(size_t)e * hsmesh::kRecFloats + 7], 4);
            if (id < nt) listed[id] = 1;
        }
        const V3 p{px, py, pz};
        for (uint32_t t = 0; t < nt; ++t) {
            const float* r = &B.tris[(size_t)t * hsmesh::kTriFloats];
            const double d2 = tri_dist2(p, V3{r[0], r[1], r[2]}, V3{r[4], r[5], r[6]}, V3{r[8], r[9], r[10]});
            if (d2 < m2 && !listed[t]) fail("triangle within the margin not listed", (long)s, (long)t);
        }
    }
    if (fails) return 1;
    std::printf("ok %llu %d\n", (unsigned long long)entries, checked);
    return 0;
}
