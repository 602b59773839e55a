/*
 * observe.c -- FP64 CPU restatement of the Observe stage's flux synthesis:
 * the thermal-infrared flux of the rough and smooth model at the observer,
 * per time step and wavelength, from the Thermal stage's temperatures and
 * the Visibility stage's masks.
 *
 * TEST INFRASTRUCTURE ONLY: the parity checker for the library's flux
 * kernels; imported by tests/ and the cpu_baseline leg of bench.py, never by
 * the product path.  Built with -ffp-contract=off (operation order as the
 * reference's).
 *
 * Follows 175706/Run-04/observe_generator.cpp:
 *   asteroidgeometry :530-698  vertices turned per step (primary and secondary
 *                              revolutions), placed at the heliocentric
 *                              position in metres (AU = 149.6e9, :9); per
 *                              facet the observer vector from the observer's
 *                              position, unit normal, area (the value of the
 *                              last step is the one the flux uses), the
 *                              view vector in the facet frame and the view
 *                              angle; facets seen beyond 90 degrees are
 *                              invisible
 *   surfacegeometry  :465-529  crater sub-facet areas, unit normals, the
 *                              projected crater area S = sum A_j n_jz, and
 *                              per step and facet the view angle of every
 *                              sub-facet
 *   observe          :699-836  per step: Planck intensity x emissivity x
 *                              area x visibility x cos(view angle) / d^2,
 *                              summed over the crater sub-facets of every
 *                              visible facet (scaled by A_facet / S) and over
 *                              the facets; a colour temperature 100..566 K by
 *                              least squares on the band ratios picks the
 *                              colour correction of each band.  The smooth
 *                              fit's second and third terms are
 *                              `((b-m)/m, 2.0)` (comma operator: 2.0 each,
 *                              :782), so it fits band 1 alone; kept.
 */
#include <math.h>
#include <stdlib.h>

#include "btpm_oracle.h"

static double planck(double pre, double hc, double lk, double T)
{
    /* observe_generator.cpp:737/757: (2hc^2/lambda^5) (1/(exp(hc/(lambda k T)) - 1)) */
    return pre * (1.0 / (exp(hc / (lk * T)) - 1.0));
}

void orc_observe(const orc_observe_in *in, double *smooth_out, double *rough_out)
{
    const double PI = 4.0 * atan(1.0);
    const double hcon = 6.626068963e-34, cvel = 299792458.0, kcon = 1.380650424e-23;
    const int nv = in->nv, nf = in->nf, ntau = in->ntau, ncf = in->ncf, nwl = in->nwl;
    /* :531-537 */
    double acen[3], ocen[3];
    acen[0] = in->au_m * in->hel[0] * cos(in->hel[1]) * cos(in->hel[2]);
    acen[1] = in->au_m * in->hel[0] * cos(in->hel[1]) * sin(in->hel[2]);
    acen[2] = in->au_m * in->hel[0] * sin(in->hel[1]);
    ocen[0] = acen[0] - (in->au_m * in->obs[0] * cos(in->obs[1]) * cos(in->obs[2]));
    ocen[1] = acen[1] - (in->au_m * in->obs[0] * cos(in->obs[1]) * sin(in->obs[2]));
    ocen[2] = acen[2] - (in->au_m * in->obs[0] * sin(in->obs[1]));
    /* :539-549 */
    const double r0 = in->rot[0], r1 = in->rot[1];
    const double gamma = acos(cos(r0) * sin(fabs(r1)));
    double tm[3][3];
    tm[0][0] = sin(r0) / sin(gamma);
    tm[0][1] = -1.0 * cos(r0) * cos(r0) * sin(r1) * cos(r1) / sin(gamma);
    tm[0][2] = cos(r0) * cos(r1);
    tm[1][0] = 0.0;
    tm[1][1] = ((sin(r0) * sin(r0)) + (cos(r0) * cos(r0) * cos(r1) * cos(r1))) / sin(gamma);
    tm[1][2] = cos(r0) * sin(r1);
    tm[2][0] = -1.0 * cos(r0) * cos(r1) / sin(gamma);
    tm[2][1] = -1.0 * sin(r0) * cos(r0) * sin(r1) / sin(gamma);
    tm[2][2] = sin(r0);

    /* crater geometry (surfacegeometry :466-492) */
    double *cn = (double *)malloc(sizeof(double) * 3 * (size_t)(ncf > 0 ? ncf : 1));
    double *ca = (double *)malloc(sizeof(double) * (size_t)(ncf > 0 ? ncf : 1));
    double stot = 0.0;
    for (int j = 0; j < ncf; ++j) {
        double p[3][3], a[3], b[3], n[3];
        for (int v = 0; v < 3; ++v)
            for (int k = 0; k < 3; ++k) p[v][k] = (double)in->cvtx[3 * in->cfv[3 * j + v] + k];
        for (int k = 0; k < 3; ++k) {
            a[k] = p[1][k] - p[0][k];
            b[k] = p[2][k] - p[0][k];
        }
        n[0] = (a[1] * b[2]) - (a[2] * b[1]);
        n[1] = (a[2] * b[0]) - (a[0] * b[2]);
        n[2] = (a[0] * b[1]) - (a[1] * b[0]);
        const double l = sqrt((n[0] * n[0]) + (n[1] * n[1]) + (n[2] * n[2]));
        ca[j] = l / 2.0;
        for (int k = 0; k < 3; ++k) cn[3 * j + k] = n[k] / l;
        stot = stot + (ca[j] * cn[3 * j + 2]);
    }

    /* per step: facet geometry (asteroidgeometry :557-690) into these */
    double *hv = (double *)malloc(sizeof(double) * 3 * (size_t)nv);
    double *vv = (double *)malloc(sizeof(double) * 3 * (size_t)nf);
    double *vang = (double *)malloc(sizeof(double) * (size_t)nf);
    int *vis = (int *)malloc(sizeof(int) * (size_t)nf);
    double *area = (double *)malloc(sizeof(double) * (size_t)nf);
    for (int pass = 0; pass < 2; ++pass) {
        /* pass 0: the last step, for the areas the flux uses (the reference
         * keeps the area of its last geometry step); pass 1: the steps */
        const int s_lo = pass == 0 ? ntau - 1 : in->s_begin;
        const int s_hi = pass == 0 ? ntau : in->s_begin + in->s_count;
        for (int s = s_lo; s < s_hi; ++s) {
            for (int v = 0; v < nv; ++v) {
                const double rev = v < in->nv_primary ? in->rev_primary : in->rev_secondary;
                const double ang = 2.0 * PI * s * rev / ntau;
                const double *e = in->equpos + 3 * (size_t)v;
                double q[3], ec[3];
                q[0] = (e[0] * cos(ang)) - (e[1] * sin(ang));
                q[1] = (e[0] * sin(ang)) + (e[1] * cos(ang));
                q[2] = e[2];
                for (int k = 0; k < 3; ++k) {
                    ec[k] = (tm[k][0] * q[0]) + (tm[k][1] * q[1]) + (tm[k][2] * q[2]);
                    hv[3 * (size_t)v + k] = acen[k] + ec[k];
                }
            }
            for (int f = 0; f < nf; ++f) {
                const double *p0 = hv + 3 * (size_t)in->fv[3 * f];
                const double *p1 = hv + 3 * (size_t)in->fv[3 * f + 1];
                const double *p2 = hv + 3 * (size_t)in->fv[3 * f + 2];
                double mid[3], ov[3], a[3], b[3], n[3], y[3];
                for (int k = 0; k < 3; ++k) {
                    mid[k] = (p0[k] + p1[k] + p2[k]) / 3.0;
                    ov[k] = mid[k] - ocen[k];
                    a[k] = p1[k] - p0[k];
                    b[k] = p2[k] - p0[k];
                }
                const double o = sqrt((ov[0] * ov[0]) + (ov[1] * ov[1]) + (ov[2] * ov[2]));
                n[0] = (a[1] * b[2]) - (a[2] * b[1]);
                n[1] = (a[2] * b[0]) - (a[0] * b[2]);
                n[2] = (a[0] * b[1]) - (a[1] * b[0]);
                const double nl = sqrt((n[0] * n[0]) + (n[1] * n[1]) + (n[2] * n[2]));
                if (pass == 0) {
                    area[f] = nl / 2.0;
                    continue;
                }
                for (int k = 0; k < 3; ++k) n[k] = n[k] / nl;
                const double x = sqrt((a[0] * a[0]) + (a[1] * a[1]) + (a[2] * a[2]));
                for (int k = 0; k < 3; ++k) a[k] = a[k] / x;
                y[0] = (n[1] * a[2]) - (n[2] * a[1]);
                y[1] = (n[2] * a[0]) - (n[0] * a[2]);
                y[2] = (n[0] * a[1]) - (n[1] * a[0]);
                vv[3 * f] = -1.0 * ((a[0] * ov[0]) + (a[1] * ov[1]) + (a[2] * ov[2]));
                vv[3 * f + 1] = -1.0 * ((y[0] * ov[0]) + (y[1] * ov[1]) + (y[2] * ov[2]));
                vv[3 * f + 2] = -1.0 * ((n[0] * ov[0]) + (n[1] * ov[1]) + (n[2] * ov[2]));
                vang[f] = acos(vv[3 * f + 2] / o);
                vis[f] = in->visible[(size_t)f * ntau + s];
                if (vang[f] > (PI / 2.0)) vis[f] = 0;
            }
            if (pass == 0) continue;
            /* observe (:715-796) */
            double sm[8] = {0}, ro[8] = {0};
            for (int f = 0; f < nf; ++f) {
                if (vis[f] == 0) continue;
                const double *w = vv + 3 * f;
                const double d = sqrt((w[0] * w[0]) + (w[1] * w[1]) + (w[2] * w[2]));
                if (in->rough) {
                    double sf[8] = {0};
                    for (int j = 0; j < ncf; ++j) {
                        const size_t o = ((size_t)f * ntau + s) * ncf + j;
                        const double cv = (double)in->cvis[o];
                        if (!(cv > 0.0)) continue;
                        const double va = acos(((w[0] * cn[3 * j]) + (w[1] * cn[3 * j + 1]) +
                                                (w[2] * cn[3 * j + 2])) / d);
                        for (int i = 0; i < nwl; ++i) {
                            const double lam = in->wavelength_um[i] * 1.0e-6;
                            double it = planck((2.0 * hcon * cvel * cvel) / pow(lam, 5.0), hcon * cvel,
                                               lam * kcon, (double)in->ctemp[o]);
                            double mu = (in->eps * ca[j] * cv * cos(va)) / (d * d);
                            if (it != it) it = 0.0;
                            if (mu != mu) mu = 0.0;
                            sf[i] = sf[i] + (it * mu);
                        }
                    }
                    for (int i = 0; i < nwl; ++i) ro[i] = ro[i] + (sf[i] * area[f] / stot);
                }
                for (int i = 0; i < nwl; ++i) {
                    const double lam = in->wavelength_um[i] * 1.0e-6;
                    double it = planck((2.0 * hcon * cvel * cvel) / pow(lam, 5.0), hcon * cvel,
                                       lam * kcon, (double)in->stemp[(size_t)f * ntau + s]);
                    double mu = (in->eps * area[f] * vis[f] * cos(vang[f])) / (d * d);
                    if (it != it) it = 0.0;
                    if (mu != mu) mu = 0.0;
                    sm[i] = sm[i] + (it * mu);
                }
            }
            /* colour temperature and corrections (:765-795); three bands */
            double cs_s = 100.0, cs_r = 100.0;
            int bt_s = in->tmin, bt_r = in->tmin;
            for (int T = in->tmin; T != in->tmin + in->tcount; ++T) {
                double b[3];
                for (int i = 0; i < 3; ++i) {
                    const double lam = in->wavelength_um[i] * 1.0e-6;
                    b[i] = planck((2.0 * hcon * cvel * cvel) / pow(lam, 5.0), hcon * cvel, lam * kcon,
                                  (double)T);
                }
                const double ab = (b[0] + b[1] + b[2]) / 3.0;
                for (int i = 0; i < 3; ++i) b[i] = b[i] / ab;
                double am = (sm[0] + sm[1] + sm[2]) / 3.0;
                double m2 = sm[0] / am;
                double chi = pow((b[0] - m2) / m2, 2.0) + 2.0 + 2.0;
                if (chi < cs_s) {
                    cs_s = chi;
                    bt_s = T;
                }
                am = (ro[0] + ro[1] + ro[2]) / 3.0;
                m2 = ro[0] / am;
                const double m3 = ro[1] / am, m4 = ro[2] / am;
                chi = pow((b[0] - m2) / m2, 2.0) + pow((b[1] - m3) / m3, 2.0) + pow((b[2] - m4) / m4, 2.0);
                if (chi < cs_r) {
                    cs_r = chi;
                    bt_r = T;
                }
            }
            const size_t so = (size_t)(s - in->s_begin) * nwl;
            for (int i = 0; i < nwl; ++i) {
                smooth_out[so + i] = sm[i] * in->colour[(size_t)i * in->tcount + (bt_s - in->tmin)];
                rough_out[so + i] = ro[i] * in->colour[(size_t)i * in->tcount + (bt_r - in->tmin)];
            }
        }
    }
    free(area);
    free(vis);
    free(vang);
    free(vv);
    free(hv);
    free(ca);
    free(cn);
}
