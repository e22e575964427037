/*
 * hair_oracle_body.h - body of the oracle, included twice by hair_oracle.c
 * with R = float (suffix _f32) and R = double (suffix _f64).
 * TEST INFRASTRUCTURE: see hair_oracle.h for the arithmetic contract.
 *
 * Every function cites the lines of src/shaders/compute.comp it restates.
 */

#define CAT2(a, b) a##b
#define CAT(a, b) CAT2(a, b)
#define F(name) CAT(name, SFX)

typedef struct { R x, y, z; } F(v3_t);
#define V3 F(v3_t)

static inline V3 F(v3)(R x, R y, R z) { V3 v; v.x = x; v.y = y; v.z = z; return v; }
static inline V3 F(vadd)(V3 a, V3 b) { return F(v3)(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline V3 F(vsub)(V3 a, V3 b) { return F(v3)(a.x - b.x, a.y - b.y, a.z - b.z); }
/* scalar * vector, scalar on the left as the GLSL source writes it */
static inline V3 F(smul)(R s, V3 a) { return F(v3)(s * a.x, s * a.y, s * a.z); }
/* vector * scalar */
static inline V3 F(muls)(V3 a, R s) { return F(v3)(a.x * s, a.y * s, a.z * s); }
static inline R F(dot)(V3 a, V3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
static inline R F(len)(V3 a) { return SQRT(F(dot)(a, a)); }
/* normalize(v) = v * (1/length(v)) */
static inline V3 F(nrm)(V3 a) { R inv = (R)1 / F(len)(a); return F(muls)(a, inv); }

/* mat4 (column-major, m[c*4+r]) * vec4(v, w) -> xyz */
static inline V3 F(mat_vec)(const float* m, V3 v, R w)
{
    R o[3];
    for (int r = 0; r < 3; ++r)
        o[r] = (((R)m[0 * 4 + r] * v.x + (R)m[1 * 4 + r] * v.y) + (R)m[2 * 4 + r] * v.z) + (R)m[3 * 4 + r] * w;
    return F(v3)(o[0], o[1], o[2]);
}

static inline R F(clamp01)(R x) { return FMIN(FMAX(x, (R)0), (R)1); }

int32_t F(ho_sat_i32)(R f)
{
    if (f != f) return 0;
    if (f >= (R)2147483648.0) return INT32_MAX;
    if (f < (R)-2147483648.0) return INT32_MIN;
    return (int32_t)f;
}

/* Integer range of the cells one grid coordinate touches:
 * [max(int(floor q), 0), min(int(floor q) + 1, G - 1)] (compute.comp:224-229). */
static inline void F(cell_range)(R q, int G, int64_t* lo, int64_t* hi)
{
    int64_t f = (int64_t)F(ho_sat_i32)(FLOOR(q));
    *lo = f > 0 ? f : 0;
    *hi = (f + 1) < (int64_t)(G - 1) ? (f + 1) : (int64_t)(G - 1);
}

/* Grid-space coordinate of a position: (p - origin) / h, with /h taken as *(1/h)
 * (compute.comp:219-221, 260-262). */
static inline V3 F(grid_coord)(const ho_params* p, V3 pos)
{
    R inv_h = (R)1 / (R)p->grid_h;
    V3 o = F(v3)((R)p->grid_origin[0], (R)p->grid_origin[1], (R)p->grid_origin[2]);
    return F(muls)(F(vsub)(pos, o), inv_h);
}

/* Trilinear splat of one point into the fixed-point grid, compute.comp:217-252. */
static void F(splat)(const ho_params* p, V3 pos, V3 vel, int32_t* grid, int atomic_adds)
{
    const int G = p->grid_dim;
    const R SCALE = (R)1000000.0;
    V3 q = F(grid_coord)(p, pos);
    int64_t x0, x1, y0, y1, z0, z1;
    F(cell_range)(q.x, G, &x0, &x1);
    F(cell_range)(q.y, G, &y0, &y1);
    F(cell_range)(q.z, G, &z0, &z1);
    for (int64_t a = x0; a <= x1; ++a)
        for (int64_t b = y0; b <= y1; ++b)
            for (int64_t c = z0; c <= z1; ++c) {
                int64_t idx = a + b * G + c * (int64_t)G * G;
                R wx = F(clamp01)((R)1 - FABS(q.x - (R)a));
                R wy = F(clamp01)((R)1 - FABS(q.y - (R)b));
                R wz = F(clamp01)((R)1 - FABS(q.z - (R)c));
                R w = (wx * wy) * wz;
                V3 wv = F(smul)(w, vel);
                uint32_t add[4] = {(uint32_t)F(ho_sat_i32)(SCALE * wv.x),
                                   (uint32_t)F(ho_sat_i32)(SCALE * wv.y),
                                   (uint32_t)F(ho_sat_i32)(SCALE * wv.z),
                                   (uint32_t)F(ho_sat_i32)(SCALE * w)};
                uint32_t* cell = (uint32_t*)(grid + 4 * idx);
                for (int k = 0; k < 4; ++k) {
                    if (atomic_adds)
                        __atomic_fetch_add(&cell[k], add[k], __ATOMIC_RELAXED);
                    else
                        cell[k] += add[k]; /* wraps mod 2^32 like atomicAdd */
                }
            }
}

/* Grid velocity at one point, compute.comp:258-294: sum over the touched cells
 * with density > 0 of (w * (1/float(density))) * vec3(velocity). */
static V3 F(gather)(const ho_params* p, V3 pos, const int32_t* grid)
{
    const int G = p->grid_dim;
    V3 q = F(grid_coord)(p, pos);
    int64_t x0, x1, y0, y1, z0, z1;
    F(cell_range)(q.x, G, &x0, &x1);
    F(cell_range)(q.y, G, &y0, &y1);
    F(cell_range)(q.z, G, &z0, &z1);
    V3 gv = F(v3)(0, 0, 0);
    for (int64_t a = x0; a <= x1; ++a)
        for (int64_t b = y0; b <= y1; ++b)
            for (int64_t c = z0; c <= z1; ++c) {
                const int32_t* cell = grid + 4 * (a + b * G + c * (int64_t)G * G);
                int32_t d = cell[3];
                if (d > 0) {
                    R wx = F(clamp01)((R)1 - FABS(q.x - (R)a));
                    R wy = F(clamp01)((R)1 - FABS(q.y - (R)b));
                    R wz = F(clamp01)((R)1 - FABS(q.z - (R)c));
                    R w = (wx * wy) * wz;
                    R t = w * ((R)1 / (R)d);
                    gv = F(vadd)(gv, F(smul)(t, F(v3)((R)cell[0], (R)cell[1], (R)cell[2])));
                }
            }
    return gv;
}

/* Closest point on triangle abc to x (Ericson 2005, 5.1.5) and the feature it
 * lies on: 0..2 vertex a/b/c, 3 edge ab, 4 edge bc, 5 edge ca, 6 face. */
static V3 F(closest_on_tri)(V3 x, V3 a, V3 b, V3 c, int* feature)
{
    V3 ab = F(vsub)(b, a), ac = F(vsub)(c, a), ap = F(vsub)(x, a);
    R d1 = F(dot)(ab, ap), d2 = F(dot)(ac, ap);
    if (d1 <= 0 && d2 <= 0) { *feature = 0; return a; }
    V3 bp = F(vsub)(x, b);
    R d3 = F(dot)(ab, bp), d4 = F(dot)(ac, bp);
    if (d3 >= 0 && d4 <= d3) { *feature = 1; return b; }
    R vc = d1 * d4 - d3 * d2;
    if (vc <= 0 && d1 >= 0 && d3 <= 0) {
        R v = d1 / (d1 - d3);
        *feature = 3;
        return F(vadd)(a, F(smul)(v, ab));
    }
    V3 cp = F(vsub)(x, c);
    R d5 = F(dot)(ab, cp), d6 = F(dot)(ac, cp);
    if (d6 >= 0 && d5 <= d6) { *feature = 2; return c; }
    R vb = d5 * d2 - d1 * d6;
    if (vb <= 0 && d2 >= 0 && d6 <= 0) {
        R w = d2 / (d2 - d6);
        *feature = 5;
        return F(vadd)(a, F(smul)(w, ac));
    }
    R va = d3 * d6 - d5 * d4;
    if (va <= 0 && (d4 - d3) >= 0 && (d5 - d6) >= 0) {
        R w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
        *feature = 4;
        return F(vadd)(b, F(smul)(w, F(vsub)(c, b)));
    }
    R denom = (R)1 / ((va + vb) + vc);
    R v = vb * denom, w = vc * denom;
    *feature = 6;
    return F(vadd)(F(vadd)(a, F(smul)(v, ab)), F(smul)(w, ac));
}

/* Mesh contact of the rule in hair_oracle.h, by brute force over every
 * triangle; adds the penalty to *acc and returns 1 on contact. */
static int F(mesh_contact)(const ho_params* p, V3 x, R k, V3* acc)
{
    static const int slot_of_feature[7] = {7, 8, 9, 4, 5, 6, 3};
    const R m = (R)p->mesh_margin;
    const R m2 = m * m;
    /* triangles whose bounds grown by a bit more than the margin miss x are
       farther than the margin and cannot give a contact */
    const R reach = m * (R)1.001 + (R)1e-6;
    R best = m2;
    int best_f = -1, best_feat = 0;
    V3 best_q = F(v3)(0, 0, 0);
    for (int f = 0; f < p->mesh_ntris; ++f) {
        const float* t = p->mesh_tris + (size_t)f * HO_MESH_TRI_FLOATS;
        int skip = 0;
        for (int c = 0; c < 3 && !skip; ++c) {
            R lo = FMIN(FMIN((R)t[c], (R)t[4 + c]), (R)t[8 + c]) - reach;
            R hi = FMAX(FMAX((R)t[c], (R)t[4 + c]), (R)t[8 + c]) + reach;
            R xc = c == 0 ? x.x : (c == 1 ? x.y : x.z);
            skip = xc < lo || xc > hi;
        }
        if (skip) continue;
        int feat;
        V3 q = F(closest_on_tri)(x, F(v3)(t[0], t[1], t[2]), F(v3)(t[4], t[5], t[6]), F(v3)(t[8], t[9], t[10]),
                                 &feat);
        V3 dq = F(vsub)(x, q);
        R d2 = F(dot)(dq, dq);
        if (d2 < best) {
            best = d2;
            best_f = f;
            best_feat = feat;
            best_q = q;
        }
    }
    if (best_f < 0) return 0;
    const float* t = p->mesh_tris + (size_t)best_f * HO_MESH_TRI_FLOATS;
    const float* n = t + 4 * slot_of_feature[best_feat];
    V3 dq = F(vsub)(x, best_q);
    if (!(F(dot)(dq, F(v3)(n[0], n[1], n[2])) < 0)) return 0;
    R d = SQRT(best);
    V3 dir = F(nrm)(F(vsub)(best_q, x));
    *acc = F(vadd)(*acc, F(smul)(k * d, dir));
    return 1;
}

/* random / noise / fbm of compute.comp:81-120 (mix(a,b,t) = a*(1-t) + b*t,
 * fract(x) = x - floor(x)). */
static inline R F(fract)(R x) { return x - FLOOR(x); }
static inline R F(random)(R px, R py)
{
    return F(fract)(SIN(px * (R)12.9898f + py * (R)78.233f) * (R)43758.5453123f);
}
static R F(noise2)(R px, R py)
{
    const R ix = FLOOR(px), iy = FLOOR(py);
    const R fx = F(fract)(px), fy = F(fract)(py);
    const R a = F(random)(ix, iy);
    const R b = F(random)(ix + (R)1, iy + (R)0);
    const R c = F(random)(ix + (R)0, iy + (R)1);
    const R d = F(random)(ix + (R)1, iy + (R)1);
    const R ux = (fx * fx) * ((R)3 - (R)2 * fx);
    const R uy = (fy * fy) * ((R)3 - (R)2 * fy);
    const R mix = a * ((R)1 - ux) + b * ux;
    return (mix + ((c - a) * uy) * ((R)1 - ux)) + ((d - b) * ux) * uy;
}
__attribute__((unused)) static R F(fbm)(R px, R py)
{
    R value = 0, amplitude = (R)0.5;
    for (int i = 0; i < 6; ++i) {
        value = value + amplitude * F(noise2)(px, py);
        px = px * (R)2;
        py = py * (R)2;
        amplitude = amplitude * (R)0.5;
    }
    return value;
}

/* The forcing of compute.comp:150 (mode 1) or :151 (mode 2) at position x and
 * Time.totalTime t, the scalar factor on the left as the shader writes it
 * (FORCING_FBM(t) is fbm(vec2(sin(t), cos(t))) evaluated in binary32 in both
 * instantiations: the shader's hash fract(sin(.) * 43758.5453) multiplies a
 * one-ulp difference of sin by 4.4e4, so a binary64 fbm would be another
 * pseudo-random number, not a more accurate one):
 *   1: 10.0 * vec3(2.0*sin(t*2.0)*cos(y*10.0)*sin((y+5.0)*15.0), 0.0, -clamp(y*2.0, 0.2, 2.0))
 *   2: 7.0*fbm(vec2(sin(t), cos(t))) *
 *        vec3(2.0*sin(t*2.0)*cos(y*10.0)*sin((y+5.0)*15.0), 4.0*sin(z*5.0 + t*3.0), -0.6*(y+3.0)) */
static V3 F(forcing)(const ho_params* p, V3 x)
{
    const R t = (R)p->total_time;
    const R sx = (((R)2 * SIN(t * (R)2)) * COS(x.y * (R)10)) * SIN((x.y + (R)5) * (R)15);
    if (p->forcing == 1) {
        const R cl = FMIN(FMAX(x.y * (R)2, (R)0.2f), (R)2);
        return F(smul)((R)10, F(v3)(sx, (R)0, -cl));
    }
    const R amp = (R)7 * (R)FORCING_FBM(t);
    return F(smul)(amp, F(v3)(sx, (R)4 * SIN(x.z * (R)5 + t * (R)3), (R)-0.6f * (x.y + (R)3)));
}

/* Force at a start-of-step position: gravity (+ wind, + per-strand noise along
 * noise_dir, + the forcing) then the collider penalty loop, compute.comp:150-184. */
static int F(force)(const ho_params* p, V3 x, R noise, V3* out)
{
    V3 f = F(vadd)(F(v3)((R)p->gravity[0], (R)p->gravity[1], (R)p->gravity[2]),
                   F(v3)((R)p->wind[0], (R)p->wind[1], (R)p->wind[2]));
    if (p->wind_noise)
        f = F(vadd)(f, F(smul)(noise, F(v3)((R)p->noise_dir[0], (R)p->noise_dir[1], (R)p->noise_dir[2])));
    if (p->forcing) f = F(vadd)(f, F(forcing)(p, x));
    const R k = (R)p->spring_k;
    V3 acc = F(v3)(0, 0, 0);
    int hits = 0;
    for (int j = 0; j < p->num_colliders; ++j) {
        const float* T = p->colliders + 48 * j;
        const float* Inv = T + 16;
        const float* InvT = T + 32;
        if (j == 0) {
            /* sphere: compute.comp:160-169 */
            V3 c = F(v3)((R)T[12], (R)T[13], (R)T[14]);
            V3 dv = F(vsub)(x, c);
            R dist = F(len)(dv);
            R rad = (R)p->sphere_radius;
            if (dist < rad) {
                R d = rad - dist;
                V3 n = F(muls)(dv, (R)1 / dist);
                acc = F(vadd)(acc, F(smul)(k * d, n));
                hits += 1;
            }
        } else {
            /* ellipsoid: EllipsoidCollision / GetPointOnEllipsoid /
             * GetEllipsoidNormal, compute.comp:64-79 and 170-179 */
            V3 tp = F(mat_vec)(Inv, x, (R)1);
            if (F(len)(tp) <= (R)1) {
                V3 onunit = F(nrm)(tp);
                V3 surf = F(mat_vec)(T, onunit, (R)1);
                R d = F(len)(F(vsub)(surf, x));
                V3 n = F(nrm)(F(nrm)(F(mat_vec)(InvT, tp, (R)0)));
                acc = F(vadd)(acc, F(smul)(k * d, n));
                hits += 1;
            }
        }
    }
    if (p->mesh_ntris > 0 && F(mesh_contact)(p, x, k, &acc)) hits += 1;
    if (hits > 0) {
        R m = (R)hits;
        f = F(vadd)(f, F(v3)(acc.x / m, acc.y / m, acc.z / m));
    }
    *out = f;
    return hits;
}

/* Phase 1 (compute.comp:144-202) then phase 2 (211-253) of one strand. */
static void F(strand_phase12)(const ho_params* p, R* st, int N, int s, int32_t* grid, int atomic_adds)
{
    R* P = st;
    R* Vl = st + 4 * N;
    R* C = st + 8 * N;
    const R dt = (R)p->dt;
    const R inv_dt = (R)1 / dt;
    const R dt2 = dt * dt;
    const R D = (R)p->damping;
    const R vmax = (R)p->vmax;
    const R r = (R)p->seg_length[s];
    const R noise = p->wind_noise ? (R)p->wind_noise[s] : (R)0;

    for (int i = 1; i < N; ++i) {
        V3 x = F(v3)(P[4 * i], P[4 * i + 1], P[4 * i + 2]);
        V3 u = F(v3)(Vl[4 * i], Vl[4 * i + 1], Vl[4 * i + 2]);
        V3 par = F(v3)(P[4 * (i - 1)], P[4 * (i - 1) + 1], P[4 * (i - 1) + 2]);
        V3 f;
        F(force)(p, x, noise, &f);
        /* predictedPos = currentPos + dt*currentVel + dt*dt*force  (:187) */
        V3 pred = F(vadd)(F(vadd)(x, F(smul)(dt, u)), F(smul)(dt2, f));
        /* follow the leader (:191-192) */
        V3 dir = F(nrm)(F(vsub)(pred, par));
        V3 pn = F(vadd)(par, F(smul)(r, dir));
        /* damped velocity + clamp + correction (:195-201) */
        V3 vel = F(smul)(D, F(muls)(F(vsub)(pn, x), inv_dt));
        if (F(len)(vel) > vmax)
            vel = F(muls)(F(nrm)(vel), vmax);
        V3 corr = F(smul)(D, F(vsub)(pn, pred));
        P[4 * i] = pn.x; P[4 * i + 1] = pn.y; P[4 * i + 2] = pn.z; P[4 * i + 3] = (R)1;
        Vl[4 * i] = vel.x; Vl[4 * i + 1] = vel.y; Vl[4 * i + 2] = vel.z; Vl[4 * i + 3] = (R)0;
        C[4 * i] = corr.x; C[4 * i + 1] = corr.y; C[4 * i + 2] = corr.z; C[4 * i + 3] = (R)0;
    }
    for (int i = 1; i < N; ++i) {
        /* DFTL velocity correction (:213-215) */
        if (i != N - 1) {
            V3 c1 = F(v3)(C[4 * (i + 1)], C[4 * (i + 1) + 1], C[4 * (i + 1) + 2]);
            V3 v = F(v3)(Vl[4 * i], Vl[4 * i + 1], Vl[4 * i + 2]);
            v = F(vsub)(v, F(muls)(c1, inv_dt));
            Vl[4 * i] = v.x; Vl[4 * i + 1] = v.y; Vl[4 * i + 2] = v.z;
        }
        if (p->hair_hair) {
            V3 pos = F(v3)(P[4 * i], P[4 * i + 1], P[4 * i + 2]);
            V3 v = F(v3)(Vl[4 * i], Vl[4 * i + 1], Vl[4 * i + 2]);
            F(splat)(p, pos, v, grid, atomic_adds);
        }
    }
}

/* Phase 3 (compute.comp:257-298) of one strand: friction blend with the grid. */
static void F(strand_phase3)(const ho_params* p, R* st, int N, const int32_t* grid)
{
    R* P = st;
    R* Vl = st + 4 * N;
    const R mu = (R)p->friction;
    const R omu = (R)1 - mu;
    for (int i = 1; i < N; ++i) {
        V3 pos = F(v3)(P[4 * i], P[4 * i + 1], P[4 * i + 2]);
        V3 gv = F(gather)(p, pos, grid);
        V3 v = F(v3)(Vl[4 * i], Vl[4 * i + 1], Vl[4 * i + 2]);
        v = F(vadd)(F(smul)(omu, v), F(smul)(mu, gv));
        Vl[4 * i] = v.x; Vl[4 * i + 1] = v.y; Vl[4 * i + 2] = v.z; Vl[4 * i + 3] = (R)0;
    }
}

static void F(step_impl)(const ho_params* p, R* strands, int S, int N, int32_t* grid, int threads)
{
    const size_t stride = 12u * (size_t)N;
    const int hh = p->hair_hair && p->grid_dim > 0;
    ho_params q = *p;
    q.hair_hair = hh;
    if (hh) {
        size_t cells = (size_t)p->grid_dim * p->grid_dim * p->grid_dim;
        memset(grid, 0, cells * 4 * sizeof(int32_t)); /* vkCmdFillBuffer, Renderer.cpp:2063 */
    }
    if (threads == 1) {
        for (int s = 0; s < S; ++s)
            F(strand_phase12)(&q, strands + s * stride, N, s, grid, 0);
        /* global barrier: the split the shader's workgroup barrier() misses */
        if (hh)
            for (int s = 0; s < S; ++s)
                F(strand_phase3)(&q, strands + s * stride, N, grid);
    } else {
#ifdef _OPENMP
        if (threads > 0) omp_set_num_threads(threads);
#endif
#pragma omp parallel for schedule(static)
        for (int s = 0; s < S; ++s)
            F(strand_phase12)(&q, strands + s * stride, N, s, grid, 1);
        if (hh) {
#pragma omp parallel for schedule(static)
            for (int s = 0; s < S; ++s)
                F(strand_phase3)(&q, strands + s * stride, N, grid);
        }
    }
}

void F(ho_splat_point)(const ho_params* p, const R pos[3], const R vel[3], int32_t* grid)
{
    F(splat)(p, F(v3)(pos[0], pos[1], pos[2]), F(v3)(vel[0], vel[1], vel[2]), grid, 0);
}

void F(ho_gather_point)(const ho_params* p, const R pos[3], const int32_t* grid, R out[3])
{
    V3 g = F(gather)(p, F(v3)(pos[0], pos[1], pos[2]), grid);
    out[0] = g.x; out[1] = g.y; out[2] = g.z;
}

int F(ho_force)(const ho_params* p, const R pos[3], R noise, R out[3])
{
    V3 f;
    int hits = F(force)(p, F(v3)(pos[0], pos[1], pos[2]), noise, &f);
    out[0] = f.x; out[1] = f.y; out[2] = f.z;
    return hits;
}

#undef V3
#undef F
#undef CAT
#undef CAT2
