/*
 * geometry.c -- FP64 CPU restatement of the illumination geometry and shadow
 * producer upstream of the Thermal stage.
 *
 * TEST INFRASTRUCTURE ONLY: the parity checker for the library's geometry
 * kernels; imported by tests/ and the cpu_baseline leg of bench.py, never by
 * the product path.  Built with -ffp-contract=off: every product and sum is
 * rounded on its own, in the reference's operation order, so the shadow
 * flags (integers decided by signs of FP64 expressions) are reproducible
 * bit for bit.
 *
 * Follows T175706/thermal.cpp (= 175706/Run-04/thermal.cpp):
 *   asteroidgeometry :1272-1471  vertex rotation to heliocentric positions per
 *                                step (primary and secondary at their own
 *                                revolutions per cycle), facet centroids,
 *                                unit normals, local frames, illumination
 *                                angle; back-facing facets unlit (:1448-1449)
 *   astshadowtest    :1612-1667  the Sun->centroid ray of the farther facet
 *                                against the plane of the nearer one, then
 *                                three edge tests against its normal
 *   astshadow        :1677-1707  every facet pair, the farther facet tested
 *   singlescattering :1757       direct term solar cos(angle) shadow (float)
 * with PI = 4 atan(1) (:12) and the position in polar form (:1274-1277).
 */
#include <math.h>
#include <stdlib.h>

#include "btpm_oracle.h"

typedef struct {
    double mid[3];     /* heliocentric centroid */
    double n[3];       /* unit normal           */
    double v[3][3];    /* heliocentric vertices */
    double dist;       /* |mid|                 */
} orc_facet_geom;

/* thermal.cpp:1612-1667: 0 if the ray from the Sun (origin) through G's
 * centroid meets H inside H's triangle, else 1 */
static int shadow_test(const orc_facet_geom *G, const orc_facet_geom *H)
{
    const double tcalc2 = G->mid[0] * H->n[0] + G->mid[1] * H->n[1] + G->mid[2] * H->n[2];
    const double tcalc1 = H->mid[0] * H->n[0] + H->mid[1] * H->n[1] + H->mid[2] * H->n[2];
    const double t = tcalc1 / tcalc2;
    double x[3], e1[3], e2[3], e3[3], x1[3], x2[3], x3[3], c1[3], c2[3], c3[3];
    for (int k = 0; k < 3; ++k) {
        x[k] = t * G->mid[k];
        e1[k] = H->v[1][k] - H->v[0][k];
        e2[k] = H->v[0][k] - H->v[2][k];
        e3[k] = H->v[2][k] - H->v[1][k];
        x1[k] = x[k] - H->v[0][k];
        x3[k] = x[k] - H->v[1][k];
        x2[k] = x[k] - H->v[2][k];
    }
    c1[0] = e1[1] * x1[2] - e1[2] * x1[1];
    c1[1] = e1[2] * x1[0] - e1[0] * x1[2];
    c1[2] = e1[0] * x1[1] - e1[1] * x1[0];
    c2[0] = e2[1] * x2[2] - e2[2] * x2[1];
    c2[1] = e2[2] * x2[0] - e2[0] * x2[2];
    c2[2] = e2[0] * x2[1] - e2[1] * x2[0];
    c3[0] = e3[1] * x3[2] - e3[2] * x3[1];
    c3[1] = e3[2] * x3[0] - e3[0] * x3[2];
    c3[2] = e3[0] * x3[1] - e3[1] * x3[0];
    const double k1 = c1[0] * H->n[0] + c1[1] * H->n[1] + c1[2] * H->n[2];
    const double k2 = c2[0] * H->n[0] + c2[1] * H->n[1] + c2[2] * H->n[2];
    const double k3 = c3[0] * H->n[0] + c3[1] * H->n[1] + c3[2] * H->n[2];
    return (k1 >= 0.0 && k2 >= 0.0 && k3 >= 0.0) ? 0 : 1;
}

void orc_geometry(const orc_geometry_in *in, int *shadow, double *cosang, float *illum,
                  float *sun_facet, float *solar_out)
{
    const double PI = 4.0 * atan(1.0);
    const int nv = in->nv, nf = in->nf, ntau = in->ntau, ns = in->s_count;
    /* thermal.cpp:1274-1277, 1282 */
    double cen[3];
    cen[0] = in->au_km * in->pol[0] * cos(in->pol[1]) * cos(in->pol[2]);
    cen[1] = in->au_km * in->pol[0] * cos(in->pol[1]) * sin(in->pol[2]);
    cen[2] = in->au_km * in->pol[0] * sin(in->pol[1]);
    const float solar = (float)(1367.0 / (in->pol[0] * in->pol[0]));
    if (solar_out) *solar_out = solar;
    /* thermal.cpp:1286-1296: ecliptic <- equatorial */
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

    double *hv = (double *)malloc(sizeof(double) * 3 * (size_t)nv);
    orc_facet_geom *fg = (orc_facet_geom *)malloc(sizeof(orc_facet_geom) * (size_t)nf);
    for (int si = 0; si < ns; ++si) {
        const int s = in->s_begin + si;
        /* thermal.cpp:1349-1393: rotate each body's vertices and place them */
        for (int v = 0; v < nv; ++v) {
            const double rev = v < in->nv_primary ? in->rev_primary : in->rev_secondary;
            const double ang = (2.0 * PI * s * rev / ntau) + 0.0;
            const double *e = in->equpos + 3 * (size_t)v;
            const double q0 = (e[0] * cos(ang)) - (e[1] * sin(ang));
            const double q1 = (e[0] * sin(ang)) + (e[1] * cos(ang));
            const double q2 = e[2];
            for (int k = 0; k < 3; ++k)
                hv[3 * (size_t)v + k] = cen[k] + ((tm[k][0] * q0) + (tm[k][1] * q1) + (tm[k][2] * q2));
        }
        /* thermal.cpp:1395-1451 */
        for (int f = 0; f < nf; ++f) {
            orc_facet_geom *g = &fg[f];
            const double *p0 = hv + 3 * (size_t)in->fv[3 * f];
            const double *p1 = hv + 3 * (size_t)in->fv[3 * f + 1];
            const double *p2 = hv + 3 * (size_t)in->fv[3 * f + 2];
            for (int k = 0; k < 3; ++k) {
                g->mid[k] = (p0[k] + p1[k] + p2[k]) / 3.0;
                g->v[0][k] = p0[k];
                g->v[1][k] = p1[k];
                g->v[2][k] = p2[k];
            }
            const double t = sqrt((g->mid[0] * g->mid[0]) + (g->mid[1] * g->mid[1]) +
                                  (g->mid[2] * g->mid[2]));
            g->dist = t;
            double a[3], b[3], y[3];
            for (int k = 0; k < 3; ++k) {
                a[k] = p1[k] - p0[k];
                b[k] = p2[k] - p0[k];
            }
            g->n[0] = (a[1] * b[2]) - (a[2] * b[1]);
            g->n[1] = (a[2] * b[0]) - (a[0] * b[2]);
            g->n[2] = (a[0] * b[1]) - (a[1] * b[0]);
            const double nn = sqrt((g->n[0] * g->n[0]) + (g->n[1] * g->n[1]) + (g->n[2] * g->n[2]));
            for (int k = 0; k < 3; ++k) g->n[k] = g->n[k] / nn;
            const double xl = sqrt((a[0] * a[0]) + (a[1] * a[1]) + (a[2] * a[2]));
            for (int k = 0; k < 3; ++k) a[k] = a[k] / xl;
            y[0] = (g->n[1] * a[2]) - (g->n[2] * a[1]);
            y[1] = (g->n[2] * a[0]) - (g->n[0] * a[2]);
            y[2] = (g->n[0] * a[1]) - (g->n[1] * a[0]);
            /* illumination vector in the facet frame (rows a, y, n) */
            const double iv0 = -1.0 * ((a[0] * g->mid[0]) + (a[1] * g->mid[1]) + (a[2] * g->mid[2]));
            const double iv1 = -1.0 * ((y[0] * g->mid[0]) + (y[1] * g->mid[1]) + (y[2] * g->mid[2]));
            const double iv2 = -1.0 * ((g->n[0] * g->mid[0]) + (g->n[1] * g->mid[1]) +
                                       (g->n[2] * g->mid[2]));
            const double ang = acos(iv2 / t);
            const size_t o = (size_t)f * ns + si;
            shadow[o] = ang > (PI / 2.0) ? 0 : 1;
            if (cosang) cosang[o] = cos(ang);
            if (sun_facet && in->sun_scale > 0.0) {  /* thermal.cpp:1453-1455 */
                sun_facet[3 * o] = (float)(in->sun_scale * iv0 / t);
                sun_facet[3 * o + 1] = (float)(in->sun_scale * iv1 / t);
                sun_facet[3 * o + 2] = (float)(in->sun_scale * iv2 / t);
            } else if (sun_facet) {  /* the same, scaled to unit length */
                sun_facet[3 * o] = (float)(iv0 / t);
                sun_facet[3 * o + 1] = (float)(iv1 / t);
                sun_facet[3 * o + 2] = (float)(iv2 / t);
            }
        }
        /* thermal.cpp:1677-1707 */
        for (int G = 0; G < nf; ++G)
            for (int H = G + 1; H < nf; ++H) {
                const size_t oG = (size_t)G * ns + si, oH = (size_t)H * ns + si;
                if (fg[G].dist > fg[H].dist) {
                    if (shadow[oG] == 1 && shadow_test(&fg[G], &fg[H]) == 0) shadow[oG] = 0;
                } else {
                    if (shadow[oH] == 1 && shadow_test(&fg[H], &fg[G]) == 0) shadow[oH] = 0;
                }
            }
        /* thermal.cpp:1757 (direct term; the facet's own angle recomputed
         * as cos(acos(.)) in double, as the reference stores the angle) */
        if (illum)
            for (int f = 0; f < nf; ++f) {
                const size_t o = (size_t)f * ns + si;
                const double ang = acos(-1.0 * ((fg[f].n[0] * fg[f].mid[0]) + (fg[f].n[1] * fg[f].mid[1]) +
                                                (fg[f].n[2] * fg[f].mid[2])) / fg[f].dist);
                illum[o] = solar * (float)cos(ang) * (float)shadow[o];
            }
    }
    free(fg);
    free(hv);
}
