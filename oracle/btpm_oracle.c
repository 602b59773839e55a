/*
 * btpm_oracle.c -- FP64 CPU restatement of the Thermal-stage hot path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker of the HIP library; see the header).
 * Loops are written for clarity.  OpenMP (when compiled with -fopenmp)
 * parallelises only over independent columns inside one time step, so the
 * results do not depend on the thread count.
 */
#include "btpm_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define ORC_PI 3.14159265358979323846

void orc_default_params(orc_params *p, int ntau, double rot_period, double tacc)
{
    /* constants of T175706/thermal.cpp:14,33,77,81 */
    p->ntau = ntau;
    p->nx = 56;
    p->rot_period = rot_period;
    p->eps = 0.95;
    p->sigma = 5.67e-8;
    p->ir_abs = 1.0;
    p->tacc = tacc;
    p->newton_tol = tacc / 10.0;    /* T175706/thermal.cpp:1936 */
    p->max_rev = 301;               /* :3592 */
    p->check_after = 5;             /* :3600 */
    p->newton_cap = 200;
}

/* Surface energy balance at step s (Appendix B; T175706/thermal.cpp:1933-1942,
 * thermal_GPU.cu:548-557):
 *   g(T) = src + m1 (t1 - T) - es T^4,  g'(T) = -(m1 + 4 es T^3)
 * Newton  T <- T + g / (m1 + 4 es T^3)  from t0 until |dT| < tol.           */
double orc_newton_sbc(double src, double m1, double t1, double es, double t0,
                      double tol, int cap, int *iters)
{
    double x = t0, xn = t0, d;
    int it = 0;
    do {
        double x3 = x * x * x;
        double g = src + m1 * (t1 - x) - es * x3 * x;
        double gp = m1 + 4.0 * es * x3;
        xn = x + g / gp;
        d = fabs(xn - x);
        x = xn;
        ++it;
    } while (d >= tol && it < cap);
    if (iters) *iters = it;
    return xn;
}

/* FTCS update of the interior nodes 1..nx-1 with the old neighbours, then
 * the zero-flux bottom T[nx] = T[nx-1]
 * (T175706/thermal.cpp:1925-1931 and :1951; thermal_GPU.cu:457-476).       */
static void ftcs(double *T, int nx, double r, long stride)
{
    double before = T[0];
    for (int i = 1; i < nx; ++i) {
        double cur = T[i * stride];
        T[i * stride] = cur + r * (T[(i + 1) * stride] - 2.0 * cur + before);
        before = cur;
    }
    T[nx * stride] = T[(nx - 1) * stride];
}

void orc_smooth_init(const orc_params *p, int nf, const double *F,
                     const double *albedo, double *temp, double *stemp)
{
    const int nt = p->ntau, nx = p->nx;
    const double es = p->eps * p->sigma;
    for (int f = 0; f < nf; ++f) {
        double mean = 0.0;
        for (int s = 0; s < nt; ++s) mean += F[(long)f * nt + s];
        mean /= nt;
        double t = pow((1.0 - albedo[f]) * mean / es, 0.25);
        for (int i = 0; i <= nx; ++i) temp[(long)f * (nx + 1) + i] = t;
        for (int s = 0; s < nt; ++s) stemp[(long)f * nt + s] = t;
    }
}

void orc_smooth_rotation(const orc_params *p, int nf, const double *F,
                         const double *albedo, const double *ti,
                         const double *numrevs, const int *rowptr,
                         const int *colidx, const double *vf,
                         const unsigned char *frozen, double *temp,
                         double *stemp, double *dmax, orc_stats *st)
{
    const int nt = p->ntau, nx = p->nx;
    const double dz = 8.0 / nx;
    const double es = p->eps * p->sigma;
    const double esia = p->ir_abs * es;
    double *emis = NULL;
    long iters = 0, capped = 0, steps = 0;

    if (rowptr) emis = (double *)malloc(sizeof(double) * (size_t)nf);
    for (int f = 0; f < nf; ++f)
        if (!frozen || !frozen[f]) dmax[f] = 0.0;

    for (int s = 0; s < nt; ++s) {
        /* previous-rotation emission of every facet at this step (the
         * stemp(0)/stemp(1) pair of T175706/thermal.cpp:1919,1939-1946) */
        if (emis) {
            for (int g = 0; g < nf; ++g) {
                double t = stemp[(long)g * nt + s];
                emis[g] = esia * t * t * t * t;
            }
        }
#pragma omp parallel for schedule(static) reduction(+ : iters, capped, steps)
        for (int f = 0; f < nf; ++f) {
            if (frozen && frozen[f]) continue;
            double *T = temp + (long)f * (nx + 1);
            const double prev = stemp[(long)f * nt + s];
            const double r = (2.0 * ORC_PI * numrevs[f] / nt) / (dz * dz);
            const double m1 = ti[f] * sqrt(2.0 * ORC_PI * numrevs[f] / p->rot_period) / dz;
            double q = 0.0;
            int it;
            ftcs(T, nx, r, 1);
            if (emis)
                for (int k = rowptr[f]; k < rowptr[f + 1]; ++k) q += vf[k] * emis[colidx[k]];
            double src = (1.0 - albedo[f]) * F[(long)f * nt + s] + q;
            double t0 = orc_newton_sbc(src, m1, T[1], es, T[0], p->newton_tol,
                                       p->newton_cap, &it);
            iters += it;
            if (it >= p->newton_cap) capped++;
            steps++;
            T[0] = t0;
            stemp[(long)f * nt + s] = t0;
            double d = fabs(t0 - prev);
            if (d > dmax[f]) dmax[f] = d;
        }
    }
    if (st) {
        st->newton_iters += iters;
        st->newton_capped += capped;
        st->column_steps += steps;
    }
    free(emis);
}

void orc_inst_rotation(const orc_params *p, int nf, const double *F,
                       const double *albedo, const int *rowptr,
                       const int *colidx, const double *vf, double *stemp,
                       double *dmax)
{
    const int nt = p->ntau;
    const double es = p->eps * p->sigma;
    const double esia = p->ir_abs * es;
    double *emis = (double *)malloc(sizeof(double) * (size_t)nf);
    double *snew = (double *)malloc(sizeof(double) * (size_t)nf);
    for (int f = 0; f < nf; ++f) dmax[f] = 0.0;
    for (int s = 0; s < nt; ++s) {
        for (int g = 0; g < nf; ++g) {
            double t = stemp[(long)g * nt + s];
            emis[g] = esia * t * t * t * t;
        }
        for (int f = 0; f < nf; ++f) {
            double q = 0.0;
            if (rowptr)
                for (int k = rowptr[f]; k < rowptr[f + 1]; ++k) q += vf[k] * emis[colidx[k]];
            snew[f] = pow(((1.0 - albedo[f]) * F[(long)f * nt + s] + q) / es, 0.25);
        }
        for (int f = 0; f < nf; ++f) {
            double d = fabs(snew[f] - stemp[(long)f * nt + s]);
            if (d > dmax[f]) dmax[f] = d;
            stemp[(long)f * nt + s] = snew[f];
        }
    }
    free(emis);
    free(snew);
}

static int all_zero(const double *x, int n)
{
    for (int i = 0; i < n; ++i)
        if (x[i] != 0.0) return 0;
    return 1;
}

int orc_smooth_solve(const orc_params *p, int nf, const double *F,
                     const double *albedo, const double *ti,
                     const double *numrevs, const int *rowptr,
                     const int *colidx, const double *vf, double *temp,
                     double *stemp, int *frozen_rev, orc_stats *st)
{
    unsigned char *frozen = (unsigned char *)calloc((size_t)nf, 1);
    double *dmax = (double *)calloc((size_t)nf, sizeof(double));
    const int inst = all_zero(ti, nf);
    int rev = 0, done = 0;
    if (frozen_rev)
        for (int f = 0; f < nf; ++f) frozen_rev[f] = -1;
    while (!done && rev < p->max_rev) {
        if (inst)
            orc_inst_rotation(p, nf, F, albedo, rowptr, colidx, vf, stemp, dmax);
        else
            orc_smooth_rotation(p, nf, F, albedo, ti, numrevs, rowptr, colidx, vf,
                                frozen, temp, stemp, dmax, st);
        if (rev > p->check_after) {
            /* shape_converged, T175706/thermal.cpp:2422-2454 */
            int all = 1;
            for (int f = 0; f < nf; ++f) {
                if (inst) {
                    if (dmax[f] >= p->tacc) all = 0;
                    continue;
                }
                if (!frozen[f] && dmax[f] < p->tacc) {
                    frozen[f] = 1;
                    if (frozen_rev) frozen_rev[f] = rev;
                }
                if (!frozen[f]) all = 0;
            }
            done = all;
        }
        ++rev;
    }
    if (st) {
        st->rotations = rev;
        st->converged = done;
    }
    free(frozen);
    free(dmax);
    return rev;
}

void orc_facet_scatter_single(int nf, int ntau, const double *direct,
                              const double *albedo, const int *rowptr,
                              const int *colidx, const double *vf, double *F)
{
    for (int f = 0; f < nf; ++f)
        for (int s = 0; s < ntau; ++s) {
            double v = direct[(long)f * ntau + s];
            if (rowptr)
                for (int k = rowptr[f]; k < rowptr[f + 1]; ++k) {
                    int g = colidx[k];
                    v += vf[k] * albedo[g] * direct[(long)g * ntau + s];
                }
            F[(long)f * ntau + s] = v;
        }
}

void orc_facet_scatter_multiple(int nf, int ntau, const double *direct,
                                const double *albedo, const int *rowptr,
                                const int *colidx, const double *vf,
                                double acc, double *F)
{
    double *p = (double *)malloc(sizeof(double) * (size_t)nf);
    for (int s = 0; s < ntau; ++s) {
        for (int f = 0; f < nf; ++f) p[f] = albedo[f] * F[(long)f * ntau + s];
        for (int sweep = 0; sweep < 100000; ++sweep) {
            int changed = 0;
            /* Gauss-Seidel in facet order as T175706/thermal.cpp:1800-1826 */
            for (int f = 0; f < nf; ++f) {
                if (!rowptr || rowptr[f + 1] == rowptr[f]) continue;
                double sum = 0.0;
                for (int k = rowptr[f]; k < rowptr[f + 1]; ++k) sum += vf[k] * p[colidx[k]];
                double pn = albedo[f] * (direct[(long)f * ntau + s] + sum);
                if (fabs(pn - p[f]) > acc) changed = 1;
                p[f] = pn;
            }
            if (!changed) break;
        }
        for (int f = 0; f < nf; ++f)
            F[(long)f * ntau + s] = albedo[f] > 0.0 ? p[f] / albedo[f] : direct[(long)f * ntau + s];
    }
    free(p);
}

static void rg_sum(int nfac, int ntau, int ns, const int *rowptr, const int *colidx,
                   const double *drf, const double *w /* [ngl][ntau] */, double *out)
{
#pragma omp parallel for schedule(static)
    for (int f = 0; f < nfac; ++f) {
        for (int s = 0; s < ntau; ++s) {
            double *o = out + ((long)f * ntau + s) * ns;
            for (int j = 0; j < ns; ++j) o[j] = 0.0;
            for (int k = rowptr[f]; k < rowptr[f + 1]; ++k) {
                double wg = w[(long)colidx[k] * ntau + s];
                const double *d = drf + (long)k * ns;
                for (int j = 0; j < ns; ++j) o[j] += d[j] * wg;
            }
        }
    }
}

void orc_rgi(int nfac, int ntau, int ns, const int *rowptr, const int *colidx,
             const double *drf, const double *F, const double *albedo,
             double *rgi)
{
    int ngl = 0;
    for (int k = 0; k < rowptr[nfac]; ++k)
        if (colidx[k] + 1 > ngl) ngl = colidx[k] + 1;
    double *w = (double *)malloc(sizeof(double) * (size_t)ngl * ntau + 8);
    for (int g = 0; g < ngl; ++g)
        for (int s = 0; s < ntau; ++s) w[(long)g * ntau + s] = F[(long)g * ntau + s] * albedo[g];
    rg_sum(nfac, ntau, ns, rowptr, colidx, drf, w, rgi);
    free(w);
}

void orc_rgh(int nfac, int ntau, int ns, const int *rowptr, const int *colidx,
             const double *drf, const double *stemp_smooth, double eps_sigma,
             double *rgh)
{
    int ngl = 0;
    for (int k = 0; k < rowptr[nfac]; ++k)
        if (colidx[k] + 1 > ngl) ngl = colidx[k] + 1;
    double *w = (double *)malloc(sizeof(double) * (size_t)ngl * ntau + 8);
    for (long i = 0; i < (long)ngl * ntau; ++i) {
        double t = stemp_smooth[i];
        w[i] = eps_sigma * t * t * t * t;
    }
    rg_sum(nfac, ntau, ns, rowptr, colidx, drf, w, rgh);
    free(w);
}

void orc_crater_scatter_single(int nfac, int ntau, int ns, const double *vf,
                               const double *albedo, const double *shadow,
                               const double *cosang, const double *rgi,
                               double solar, double *illum)
{
    double *il = (double *)malloc(sizeof(double) * (size_t)ns);
    for (int f = 0; f < nfac; ++f)
        for (int s = 0; s < ntau; ++s) {
            long b = ((long)f * ntau + s) * ns;
            for (int j = 0; j < ns; ++j) il[j] = solar * shadow[b + j] * cosang[b + j] + rgi[b + j];
            for (int j = 0; j < ns; ++j) {
                double t = il[j];
                for (int i = 0; i < ns; ++i) t += vf[(long)i * ns + j] * albedo[f] * il[i];
                illum[b + j] = t > 0.0 ? t : 0.0;
            }
        }
    free(il);
}

void orc_crater_scatter_multiple(int nfac, int ntau, int ns, const double *vf,
                                 const double *albedo, const double *shadow,
                                 const double *cosang, const double *rgi,
                                 double solar, double acc, double *illum)
{
    double *sp = (double *)malloc(sizeof(double) * (size_t)ns);
    double *po = (double *)malloc(sizeof(double) * (size_t)ns);
    double *pn = (double *)malloc(sizeof(double) * (size_t)ns);
    for (int f = 0; f < nfac; ++f) {
        const double a = albedo[f];
        for (int s = 0; s < ntau; ++s) {
            long b = ((long)f * ntau + s) * ns;
            for (int j = 0; j < ns; ++j) {
                sp[j] = solar * cosang[b + j] * shadow[b + j] + rgi[b + j];
                po[j] = a * illum[b + j];
            }
            if (a <= 0.0) {
                /* no scattering without albedo: the A -> 0 limit of p/A */
                for (int j = 0; j < ns; ++j) illum[b + j] = sp[j];
                continue;
            }
            for (int sweep = 0; sweep < 100000; ++sweep) {
                int changed = 0;
                for (int j = 0; j < ns; ++j) {
                    double sum = 0.0;
                    for (int i = 0; i < ns; ++i) sum += vf[(long)i * ns + j] * po[i];
                    pn[j] = a * (sp[j] + sum);
                }
                for (int j = 0; j < ns; ++j) {
                    if (fabs(pn[j] - po[j]) > acc) changed = 1;
                    po[j] = pn[j];
                }
                if (!changed) break;
            }
            for (int j = 0; j < ns; ++j) illum[b + j] = pn[j] / a;
        }
    }
    free(sp);
    free(po);
    free(pn);
}

void orc_crater_init(const orc_params *p, int nfac, int ns, const double *vf,
                     const double *albedo, const double *illum,
                     const double *rgh, double *temp, double *stemp)
{
    const int nt = p->ntau, nx = p->nx;
    const double es = p->eps * p->sigma;
    double *t1 = (double *)malloc(sizeof(double) * (size_t)ns);
    for (int f = 0; f < nfac; ++f) {
        const double a = albedo[f];
        for (int j = 0; j < ns; ++j) {
            double m = 0.0;
            for (int s = 0; s < nt; ++s) {
                long b = ((long)f * nt + s) * ns + j;
                m += (1.0 - a) * illum[b] + rgh[b];
            }
            t1[j] = pow(m / nt / es, 0.25);
        }
        for (int j = 0; j < ns; ++j) {
            double self = 0.0;
            for (int i = 0; i < ns; ++i) {
                double t = t1[i];
                self += p->ir_abs * es * t * t * t * t * vf[(long)i * ns + j];
            }
            double m = 0.0;
            for (int s = 0; s < nt; ++s) {
                long b = ((long)f * nt + s) * ns + j;
                m += (1.0 - a) * illum[b] + rgh[b] + self;
            }
            double t = pow(m / nt / es, 0.25);
            for (int i = 0; i <= nx; ++i) temp[((long)f * (nx + 1) + i) * ns + j] = t;
            for (int s = 0; s < nt; ++s) stemp[((long)f * nt + s) * ns + j] = t;
        }
    }
    free(t1);
}

void orc_crater_rotation(const orc_params *p, int nfac, int ns,
                         const double *vf, const double *albedo,
                         const double *ti, const double *numrevs,
                         const double *illum, const double *rgh,
                         const unsigned char *frozen, double *temp,
                         double *stemp, double *dmax, orc_stats *st)
{
    const int nt = p->ntau, nx = p->nx;
    const double dz = 8.0 / nx;
    const double es = p->eps * p->sigma;
    const double esia = p->ir_abs * es;
    long iters = 0, capped = 0, steps = 0;

#pragma omp parallel for schedule(dynamic) reduction(+ : iters, capped, steps)
    for (int f = 0; f < nfac; ++f) {
        if (frozen && frozen[f]) continue;
        double *e = (double *)malloc(sizeof(double) * (size_t)ns);
        const double r = (2.0 * ORC_PI * numrevs[f] / nt) / (dz * dz);
        const double m1 = ti[f] * sqrt(2.0 * ORC_PI * numrevs[f] / p->rot_period) / dz;
        double dm = 0.0;
        for (int s = 0; s < nt; ++s) {
            long b = ((long)f * nt + s) * ns;
            for (int j = 0; j < ns; ++j) {
                ftcs(temp + (long)f * (nx + 1) * ns + j, nx, r, ns);
                double t = stemp[b + j];
                e[j] = esia * t * t * t * t;
            }
            for (int j = 0; j < ns; ++j) {
                double *T = temp + (long)f * (nx + 1) * ns + j;
                double q = rgh[b + j];
                for (int i = 0; i < ns; ++i) q += vf[(long)i * ns + j] * e[i];
                double src = (1.0 - albedo[f]) * illum[b + j] + q;
                int it;
                double t0 = orc_newton_sbc(src, m1, T[ns], es, T[0], p->newton_tol,
                                           p->newton_cap, &it);
                iters += it;
                if (it >= p->newton_cap) capped++;
                steps++;
                double d = fabs(t0 - stemp[b + j]);
                if (d > dm) dm = d;
                T[0] = t0;
                stemp[b + j] = t0;
            }
        }
        dmax[f] = dm;
        free(e);
    }
    if (st) {
        st->newton_iters += iters;
        st->newton_capped += capped;
        st->column_steps += steps;
    }
}

void orc_crater_inst_rotation(const orc_params *p, int nfac, int ns,
                              const double *vf, const double *albedo,
                              const double *illum, const double *rgh,
                              const unsigned char *frozen, double *stemp,
                              double *dmax)
{
    const int nt = p->ntau;
    const double es = p->eps * p->sigma;
    const double esia = p->ir_abs * es;
    double *e = (double *)malloc(sizeof(double) * (size_t)ns);
    for (int f = 0; f < nfac; ++f) {
        if (frozen && frozen[f]) continue;
        double dm = 0.0;
        for (int s = 0; s < nt; ++s) {
            long b = ((long)f * nt + s) * ns;
            for (int i = 0; i < ns; ++i) {
                double t = stemp[b + i];
                e[i] = esia * t * t * t * t;
            }
            for (int j = 0; j < ns; ++j) {
                double q = rgh[b + j];
                for (int i = 0; i < ns; ++i) q += vf[(long)i * ns + j] * e[i];
                double t = pow(((1.0 - albedo[f]) * illum[b + j] + q) / es, 0.25);
                double d = fabs(t - stemp[b + j]);
                if (d > dm) dm = d;
                stemp[b + j] = t;
            }
        }
        dmax[f] = dm;
    }
    free(e);
}

int orc_crater_solve(const orc_params *p, int nfac, int ns, const double *vf,
                     const double *albedo, const double *ti,
                     const double *numrevs, const double *illum,
                     const double *rgh, double *temp, double *stemp,
                     int *frozen_rev, orc_stats *st)
{
    unsigned char *frozen = (unsigned char *)calloc((size_t)nfac, 1);
    double *dmax = (double *)calloc((size_t)nfac, sizeof(double));
    const int inst = all_zero(ti, nfac);
    int rev = 0, done = 0;
    if (frozen_rev)
        for (int f = 0; f < nfac; ++f) frozen_rev[f] = -1;
    while (!done && rev < p->max_rev) {
        if (inst)
            orc_crater_inst_rotation(p, nfac, ns, vf, albedo, illum, rgh, frozen, stemp, dmax);
        else
            orc_crater_rotation(p, nfac, ns, vf, albedo, ti, numrevs, illum, rgh, frozen,
                                temp, stemp, dmax, st);
        if (rev > p->check_after) {
            /* converged kernel + roughness_converged,
             * T175706/thermal_GPU.cu:648-692, thermal.cpp:2460-2490 */
            int all = 1;
            for (int f = 0; f < nfac; ++f) {
                if (!frozen[f] && dmax[f] < p->tacc) {
                    frozen[f] = 1;
                    if (frozen_rev) frozen_rev[f] = rev;
                }
                if (!frozen[f]) all = 0;
            }
            done = all;
        }
        ++rev;
    }
    if (st) {
        st->rotations = rev;
        st->converged = done;
    }
    free(frozen);
    free(dmax);
    return rev;
}
