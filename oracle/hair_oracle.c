/*
 * hair_oracle.c - CPU oracle (TEST INFRASTRUCTURE, never the product).
 * Instantiates the restatement of src/shaders/compute.comp:124-303 in
 * binary32 ("ref32") and binary64 ("oracle64").  Must be compiled with
 * -ffp-contract=off and without -ffast-math; see hair_oracle.h.
 */
#include "hair_oracle.h"

#include <math.h>
#include <stdint.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#ifdef __FAST_MATH__
#error "the oracle must not be built with -ffast-math"
#endif

#define R float
#define SFX _f32
#define SQRT sqrtf
#define FLOOR floorf
#define FABS fabsf
#define FMIN fminf
#define FMAX fmaxf
#include "hair_oracle_body.h"
#undef R
#undef SFX
#undef SQRT
#undef FLOOR
#undef FABS
#undef FMIN
#undef FMAX

#define R double
#define SFX _f64
#define SQRT sqrt
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
