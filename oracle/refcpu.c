/*
 * refcpu.c -- float, OpenMP restatement of the reference CPU loop shape.
 *
 * TEST INFRASTRUCTURE ONLY: this is the CPU baseline that bench.py times
 * beside the GPU (cpu_baseline kind "port") and that tests compare with the
 * FP64 oracle.  It keeps the reference's execution structure:
 *   - per-facet records holding temp[NX+1], stemp[2][NTAU], illumination and
 *     the dependent-facet list (the thermal fields of AstFacet,
 *     T175706/thermal.cpp:362-553);
 *   - for every time step two OpenMP parallel-for regions over facets with
 *     schedule(dynamic,100) (T175706/thermal.cpp:1911-1975): the first copies
 *     stemp(1,step) to stemp(0,step) (it replaces the racy `operated` flag by
 *     a separate pass, so neighbours always read the previous rotation), the
 *     second runs FTCS, the self-heating gather, the FP32 Newton SBC at
 *     TACC/10 and the bottom boundary;
 *   - shape_converged after rev > 5, cap 301 (:2422-2454, :3590-3607);
 *   - float arithmetic with powf as in the reference.
 * The crater path has no CPU version in the reference; rc_crater_solve is a
 * labelled restatement of the CUDA kernels finite_difference /
 * boundary_condition / converged (T175706/thermal_GPU.cu:430-692) as OpenMP
 * loops over parent facets, one parallel region per time step.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define RC_PI 3.14159265358979323846

typedef struct {
    float *temp;      /* [nx+1]        */
    float *stemp0;    /* [ntau]        */
    float *stemp1;    /* [ntau]        */
    float *illum;     /* [ntau]        */
    float albedo, inert, numrevs;
    int ndep;
    const int *dep;
    const float *depvf;
    int converged;
} rc_facet;

typedef struct {
    int ntau, nx, max_rev, check_after;
    float rot_period, eps, sigma, ir_abs, tacc;
} rc_params;

static void rc_set_threads(int nthreads)
{
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
}

int rc_max_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* Smooth solve.  Arrays follow the oracle layouts; stemp_out[nf][ntau]
 * receives the final surface temperatures.  With warm = 0 it starts from the
 * reference initial state; with warm = 1 from temp_io[nf][nx+1] and the
 * surface series already in stemp_out (the TI sweep's carried state,
 * T175706/thermal.cpp:3676-3700).  temp_io (may be NULL) receives the final
 * depth profiles.  Returns the number of rotations.                         */
int rc_smooth_solve(int nf, int ntau, int nx, float rot_period, float eps,
                    float sigma, float ir_abs, float tacc, int max_rev,
                    int check_after, const float *F, const float *albedo,
                    const float *ti, const float *numrevs, const int *rowptr,
                    const int *colidx, const float *vf, float *stemp_out,
                    int nthreads, long *column_steps, float *temp_io, int warm)
{
    rc_set_threads(nthreads);
    rc_facet *A = (rc_facet *)calloc((size_t)nf, sizeof(rc_facet));
    for (int f = 0; f < nf; ++f) {
        rc_facet *C = A + f;
        C->temp = (float *)malloc(sizeof(float) * (size_t)(nx + 1));
        C->stemp0 = (float *)malloc(sizeof(float) * (size_t)ntau);
        C->stemp1 = (float *)malloc(sizeof(float) * (size_t)ntau);
        C->illum = (float *)malloc(sizeof(float) * (size_t)ntau);
        memcpy(C->illum, F + (long)f * ntau, sizeof(float) * (size_t)ntau);
        C->albedo = albedo[f];
        C->inert = ti[f];
        C->numrevs = numrevs[f];
        C->ndep = rowptr ? rowptr[f + 1] - rowptr[f] : 0;
        C->dep = rowptr ? colidx + rowptr[f] : NULL;
        C->depvf = rowptr ? vf + rowptr[f] : NULL;
        C->converged = 0;
    }
    const float EPSSIGMA = eps * sigma;

    /* initialisetemperatures, T175706/thermal.cpp:1861-1889 */
#pragma omp parallel for schedule(static)
    for (int f = 0; f < nf; ++f) {
        rc_facet *C = A + f;
        if (warm) {
            memcpy(C->temp, temp_io + (long)f * (nx + 1), sizeof(float) * (size_t)(nx + 1));
            memcpy(C->stemp1, stemp_out + (long)f * ntau, sizeof(float) * (size_t)ntau);
            memcpy(C->stemp0, C->stemp1, sizeof(float) * (size_t)ntau);
            continue;
        }
        float t = 0.0f;
        for (int s = 0; s < ntau; ++s) t = t + ((1.0f / ntau) * C->illum[s]);
        t = powf((1.0f - C->albedo) * t / (eps * sigma), 0.25f);
        for (int i = 0; i <= nx; ++i) C->temp[i] = t;
        for (int s = 0; s < ntau; ++s) C->stemp0[s] = C->stemp1[s] = t;
    }

    const float DZ = 8.0f / (float)nx;
    long steps = 0;
    int rev = 0, done = 0;
    while (!done && rev < max_rev) {
        for (int s = 0; s < ntau; ++s) {
#pragma omp parallel for schedule(dynamic, 100)
            for (int f = 0; f < nf; ++f) A[f].stemp0[s] = A[f].stemp1[s];

#pragma omp parallel for schedule(dynamic, 100) reduction(+ : steps)
            for (int f = 0; f < nf; ++f) {
                rc_facet *C = A + f;
                if (C->converged) continue;
                float DTAU = 2.0f * (float)RC_PI * C->numrevs / (float)ntau;
                float tprev = C->temp[0], tcur;
                for (int i = 1; i != nx; ++i) {
                    tcur = C->temp[i];
                    C->temp[i] = C->temp[i] + ((DTAU / (DZ * DZ)) *
                                 (C->temp[i + 1] - (2.0f * C->temp[i]) + tprev));
                    tprev = tcur;
                }
                float incident = 0.0f;
                for (int k = 0; k < C->ndep; ++k)
                    incident = incident + (ir_abs * eps * sigma *
                               powf(A[C->dep[k]].stemp0[s], 4.0f) * C->depvf[k]);
                float m1 = C->inert * sqrtf(2.0f * (float)RC_PI * C->numrevs / rot_period) / DZ;
                float DT = 1.0f, TR0 = C->temp[0], TR1 = TR0;
                int it = 0;
                while (DT >= tacc / 10.0 && it < 200) {
                    float SBC = ((1.0f - C->albedo) * C->illum[s]) + incident +
                                (m1 * (C->temp[1] - TR0)) - (EPSSIGMA * powf(TR0, 4.0f));
                    float SBCD = m1 + (4.0f * EPSSIGMA * powf(TR0, 3.0f));
                    TR1 = TR0 + (SBC / SBCD);
                    DT = fabsf(TR1 - TR0);
                    TR0 = TR1;
                    ++it;
                }
                C->temp[0] = TR1;
                C->stemp1[s] = TR1;
                C->temp[nx] = C->temp[nx - 1];
                steps++;
            }
        }
        if (rev > check_after) {
            int all = 1;
            for (int f = 0; f < nf; ++f) {
                rc_facet *E = A + f;
                int ok = 1;
                for (int s = 0; s < ntau; ++s)
                    if (fabsf(E->stemp1[s] - E->stemp0[s]) >= tacc) ok = 0;
                if (ok) E->converged = 1; else all = 0;
            }
            done = all;
        }
        ++rev;
    }
    for (int f = 0; f < nf; ++f) {
        memcpy(stemp_out + (long)f * ntau, A[f].stemp1, sizeof(float) * (size_t)ntau);
        if (temp_io) memcpy(temp_io + (long)f * (nx + 1), A[f].temp, sizeof(float) * (size_t)(nx + 1));
        free(A[f].temp);
        free(A[f].stemp0);
        free(A[f].stemp1);
        free(A[f].illum);
    }
    free(A);
    if (column_steps) *column_steps = steps;
    return rev;
}

/* Crater solve: nfac parent facets x ns sub-facets, layouts of the oracle
 * ([f][s][j], vf[i*ns+j] = VF(j->i)).  stemp_out[nfac][ntau][ns].  With
 * warm = 0 the initial state is the two-pass
 * roughness_initialise_temperatures; with warm = 1 it is temp_io
 * [nfac][nx+1][ns] and the series already in stemp_out (the carried state of
 * the TI sweep).  temp_io (may be NULL) receives the final profiles.        */
int rc_crater_solve(int nfac, int ns, int ntau, int nx, float rot_period,
                    float eps, float sigma, float ir_abs, float tacc,
                    int max_rev, int check_after, const float *vf,
                    const float *albedo, const float *ti, const float *numrevs,
                    const float *illum, const float *rgh, float *stemp_out,
                    int nthreads, long *column_steps, float *temp_io, int warm)
{
    rc_set_threads(nthreads);
    const float EPSSIGMA = eps * sigma;
    const size_t per_f = (size_t)ntau * ns;
    const size_t ntemp = (size_t)nfac * (nx + 1) * ns;
    float *temp = (float *)malloc(sizeof(float) * ntemp);
    float *stemp0 = (float *)malloc(sizeof(float) * (size_t)nfac * per_f);
    float *stemp1 = stemp_out;
    int *conv = (int *)calloc((size_t)nfac, sizeof(int));

    if (warm) {
        memcpy(temp, temp_io, sizeof(float) * ntemp);
        memcpy(stemp0, stemp1, sizeof(float) * (size_t)nfac * per_f);
    }
    /* roughness_initialise_temperatures, T175706/thermal_GPU.cu:374-423 */
#pragma omp parallel for schedule(dynamic)
    for (int f = 0; f < nfac; ++f) {
        if (warm) continue;
        float *st0 = (float *)malloc(sizeof(float) * (size_t)ns);
        for (int j = 0; j < ns; ++j) {
            float t = 0.0f;
            for (int s = 0; s < ntau; ++s) {
                size_t b = (size_t)f * per_f + (size_t)s * ns + j;
                t = t + ((1.0f / ntau) * (((1.0f - albedo[f]) * illum[b]) + rgh[b]));
            }
            st0[j] = powf(t / EPSSIGMA, 0.25f);
        }
        for (int j = 0; j < ns; ++j) {
            float self = 0.0f;
            for (int i = 0; i < ns; ++i)
                self = self + (ir_abs * EPSSIGMA * powf(st0[i], 4.0f) * vf[(size_t)i * ns + j]);
            float t = 0.0f;
            for (int s = 0; s < ntau; ++s) {
                size_t b = (size_t)f * per_f + (size_t)s * ns + j;
                t = t + ((1.0f / ntau) * (((1.0f - albedo[f]) * illum[b]) + rgh[b] + self));
            }
            t = powf(t / EPSSIGMA, 0.25f);
            for (int i = 0; i <= nx; ++i) temp[((size_t)f * (nx + 1) + i) * ns + j] = t;
            for (int s = 0; s < ntau; ++s) {
                size_t b = (size_t)f * per_f + (size_t)s * ns + j;
                stemp0[b] = stemp1[b] = t;
            }
        }
        free(st0);
    }

    const float DZ = 8.0f / (float)nx;
    const float TACC2 = tacc / 10.0f;
    long steps = 0;
    int rev = 0, done = 0;
    while (!done && rev < max_rev) {
        for (int s = 0; s < ntau; ++s) {
#pragma omp parallel for schedule(dynamic) reduction(+ : steps)
            for (int f = 0; f < nfac; ++f) {
                if (conv[f]) continue;
                float e[128];
                const float MULTI0 = (2.0f * (float)RC_PI * numrevs[f] / (float)ntau) / (DZ * DZ);
                const float MULTI1 = ti[f] * sqrtf(2.0f * (float)RC_PI * numrevs[f] / rot_period) / DZ;
                float *T = temp + (size_t)f * (nx + 1) * ns;
                size_t b = (size_t)f * per_f + (size_t)s * ns;
                for (int j = 0; j < ns; ++j) {      /* finite_difference */
                    float t0 = T[j], t1 = T[ns + j], t2 = T[2 * ns + j];
                    for (int i = 1; i != nx; ++i) {
                        T[i * ns + j] = t1 + (MULTI0 * (t2 - (2.0f * t1) + t0));
                        t0 = t1;
                        t1 = t2;
                        if (i + 2 <= nx) t2 = T[(i + 2) * ns + j];
                    }
                    T[nx * ns + j] = T[(nx - 1) * ns + j];
                }
                for (int j = 0; j < ns; ++j) {      /* boundary_condition, load */
                    stemp0[b + j] = stemp1[b + j];
                    e[j] = ir_abs * EPSSIGMA * powf(stemp1[b + j], 4.0f);
                }
                for (int j = 0; j < ns; ++j) {
                    float incident = rgh[b + j];
                    for (int i = 0; i < ns; ++i) incident = incident + (e[i] * vf[(size_t)i * ns + j]);
                    float MULTI2 = ((1.0f - albedo[f]) * illum[b + j]) + incident;
                    float DT = 1.0f, TR0 = T[j], TR1 = TR0, temp1 = T[ns + j];
                    int it = 0;
                    while (DT >= TACC2 && it < 200) {
                        float SBC = MULTI2 + (MULTI1 * (temp1 - TR0)) - (EPSSIGMA * TR0 * TR0 * TR0 * TR0);
                        float SBCD = MULTI1 + (4.0f * EPSSIGMA * TR0 * TR0 * TR0);
                        TR1 = TR0 + (SBC / SBCD);
                        DT = fabsf(TR1 - TR0);
                        TR0 = TR1;
                        ++it;
                    }
                    T[j] = TR1;
                    stemp1[b + j] = TR1;
                }
                steps += ns;
            }
        }
        if (rev > check_after) {
            int all = 1;
#pragma omp parallel for schedule(dynamic) reduction(& : all)
            for (int f = 0; f < nfac; ++f) {
                if (conv[f]) continue;
                int ok = 1;
                for (size_t k = 0; k < per_f; ++k) {
                    size_t b = (size_t)f * per_f + k;
                    if (fabsf(stemp0[b] - stemp1[b]) >= tacc) { ok = 0; break; }
                }
                conv[f] = ok;
                all = all & ok;
            }
            done = all;
        }
        ++rev;
    }
    if (temp_io) memcpy(temp_io, temp, sizeof(float) * ntemp);
    free(temp);
    free(stemp0);
    free(conv);
    if (column_steps) *column_steps = steps;
    return rev;
}
