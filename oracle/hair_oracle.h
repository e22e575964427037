/*
 * hair_oracle.h - CPU oracle for the strand-physics pass (TEST INFRASTRUCTURE).
 *
 * This is the parity reference, never the product: only tests/, the smoke check and
 * bench.py's CPU-baseline leg may load it.  It restates the reference's
 * compute shader src/shaders/compute.comp:124-303 line by line on the
 * reference's own AoS `Strand` layout (compute.comp:44-48), once in binary32
 * (ho_step_f32, "ref32") and once in binary64 (ho_step_f64, "oracle64"), with
 * the racy grid phases of the shader (workgroup-only barrier() at
 * compute.comp:209 and :255) resolved as a true global split: every strand
 * splats before any strand gathers.
 *
 * Arithmetic contract (both precisions; the HIP kernels follow the same one):
 *  - every +,-,*,/ and sqrt is one correctly rounded IEEE operation, no fused
 *    multiply-add (built with -ffp-contract=off), left-to-right evaluation as
 *    the GLSL expression is written;
 *  - dot(a,b) = (ax*bx + ay*by) + az*bz; length(v) = sqrt(dot(v,v));
 *    distance(a,b) = length(a-b);
 *  - normalize(v) = v * (1/length(v))  (one reciprocal, three products: the
 *    usual v*rsq(dot) lowering with rsq taken as a rounded 1/sqrt);
 *  - division by a dispatch-uniform scalar (deltaTime at compute.comp:195 and
 *    :214, grid cell size h at :219-221 and :260-262) is a product with the
 *    rounded reciprocal computed once (1/dt, 1/h); every other division is a
 *    true division.  GLSL 4.50 allows 2.5 ulp for `/`, so both readings are
 *    the shader's;
 *  - mat4 * vec4 (column-major m[c*4+r]):
 *      r_row = ((m0*x + m1*y) + m2*z) + m3*w;
 *  - clamp(x,0,1) = min(max(x,0),1) with IEEE maxNum/minNum (NaN -> 0);
 *  - int(f): truncation, saturating to [INT32_MIN, INT32_MAX], NaN -> 0
 *    (the hardware cvt; the shader's out-of-range case is undefined);
 *  - grid atomicAdd wraps modulo 2^32.
 */
#ifndef HAIR_ORACLE_H
#define HAIR_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HO_MAX_COLLIDERS 8

typedef struct ho_params {
    float dt;                 /* Time.deltaTime (compute.comp:142)          */
    float gravity[3];         /* compute.comp:150                           */
    float wind[3];            /* uniform wind (added to gravity)            */
    float damping;            /* DAMPING (compute.comp:7)                   */
    float spring_k;           /* compute.comp:163,173                       */
    float sphere_radius;      /* compute.comp:161                           */
    float vmax;               /* compute.comp:198-200                       */
    float friction;           /* compute.comp:296                           */
    int hair_hair;            /* 0: skip splat and gather                   */
    int grid_dim;             /* GRID_DIM (compute.comp:9)                  */
    float grid_origin[3];     /* compute.comp:206                           */
    float grid_h;             /* compute.comp:205                           */
    int num_colliders;        /* NUM_COLLIDERS (compute.comp:8)             */
    const float* colliders;   /* num_colliders x 48 floats: transform, inv,
                                 invTrans, each column-major (Scene.h:23-39) */
    const float* seg_length;  /* S floats: segment rest length per strand   */
    const float* wind_noise;  /* S floats or NULL                           */
    float noise_dir[3];
    /* Triangle-mesh collider (no reference counterpart; see ho_mesh_prepare):
     * mesh_ntris records of HO_MESH_TRI_FLOATS floats, or 0 for none. */
    const float* mesh_tris;
    int mesh_ntris;
    float mesh_margin;
    /* Wind forcing carried commented out in the shader (compute.comp:150-151):
     * 0 none, 1 line 150, 2 line 151 (fbm-modulated), at Time.totalTime. */
    float total_time;
    int forcing;
} ho_params;

/* Mesh collider.  The reference has no mesh collider (compute.comp:155-184
 * handles a sphere and ellipsoids); the rule restated here is the library's:
 *   q = the point of the mesh nearest to x, over every triangle (closest point
 *       on a triangle: Ericson, Real-Time Collision Detection, 2005, 5.1.5),
 *       the lowest triangle index winning ties;
 *   contact when |x - q|^2 < margin^2 and x is inside: dot(x - q, n) < 0 with
 *       n the angle-weighted pseudo-normal of the feature q lies on (face,
 *       edge or vertex; Baerentzen & Aanaes, IEEE TVCG 11(3), 2005);
 *   then, as one more collider after the UBO ones: acc += (k*d) * normalize(q - x),
 *       d = |x - q|, hits += 1.
 * ho_mesh_prepare writes, per triangle, 10 groups of 4 floats (w = 0):
 * a, b, c, face normal, edge normals ab bc ca, vertex normals a b c; normals
 * computed in binary64 and rounded.  Returns 0, or -1 for an index out of
 * range or an edge not shared by exactly two triangles. */
#define HO_MESH_TRI_FLOATS 40
int ho_mesh_prepare(const float* verts, int nv, const uint32_t* tris, int nt, float* out);

/* One step over S strands of N points, in place on `strands` (reference AoS
 * layout, 12*N values per strand).  When hair_hair is set, `grid` (G^3 cells
 * x 4 int32: vx, vy, vz, density) is cleared, splatted by every strand, then
 * gathered by every strand; on return it holds this step's splat, as the
 * shader's grid SSBO does after the dispatch. */
void ho_step_f32(const ho_params* p, float* strands, int S, int N, int32_t* grid);
void ho_step_f64(const ho_params* p, double* strands, int S, int N, int32_t* grid);

/* The splat of one point alone into `grid` (compute.comp:217-252), and the
 * gathered grid velocity at one point (compute.comp:258-294); exposed for the
 * brute-force tests. */
void ho_splat_point_f32(const ho_params* p, const float pos[3], const float vel[3], int32_t* grid);
void ho_gather_point_f32(const ho_params* p, const float pos[3], const int32_t* grid, float out[3]);
void ho_splat_point_f64(const ho_params* p, const double pos[3], const double vel[3], int32_t* grid);
void ho_gather_point_f64(const ho_params* p, const double pos[3], const int32_t* grid, double out[3]);

/* Penalty force of the collider loop at one position (compute.comp:155-184):
 * returns the number of colliders hit and writes the total force
 * (gravity + wind + noise*dir + sum/count) into out. */
int ho_force_f32(const ho_params* p, const float pos[3], float noise, float out[3]);
int ho_force_f64(const ho_params* p, const double pos[3], double noise, double out[3]);

/* Saturating float->int32 (the int() of the contract above). */
int32_t ho_sat_i32_f32(float f);
/* binary32 sin / cos / fbm of the forcing: one fixed algorithm (range
 * reduction k = floor(x*2/pi + 0.5), three-part Cody-Waite split of pi/2,
 * degree-9 odd and degree-10 even polynomials on [-pi/4, pi/4]) whose every
 * operation is a single rounded binary32 add, sub or mul, so any IEEE
 * implementation of the same steps returns the same bits. */
float ho_sinf(float x);
float ho_cosf(float x);
float ho_fbmf(float x, float y);
/* The curve of the hair tessellation shader (hair.tese:34-80, `func`, one
 * guide strand) through each strand of the AoS buffer `strands` (S strands of
 * N points, Strand layout): M samples per strand at v = j/(M-1), written as
 * xyz1 into out[(s*M + j)*4 ..].  v = 1 evaluates the last segment at t = 1
 * (the shader's index there runs one past the strand). */
void ho_resample_f32(const float* strands, int S, int N, int M, float* out);
int32_t ho_sat_i32_f64(double f);

/* Multithreaded FP32 step (OpenMP over strands, same results as ho_step_f32);
 * used as the timed CPU baseline.  threads <= 0 uses the OpenMP default. */
void ho_step_f32_mt(const ho_params* p, float* strands, int S, int N, int32_t* grid, int threads);

#ifdef __cplusplus
}
#endif
#endif
