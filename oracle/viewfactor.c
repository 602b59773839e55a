/*
 * viewfactor.c -- float CPU restatement of the facet view-factor producer
 * (Shape_Selfheating) and of the shape->crater view factors, the two inputs
 * of the Thermal stage's self-heating terms.
 *
 * TEST INFRASTRUCTURE ONLY: the parity checker for the library's view-factor
 * kernels; imported by tests/ and the cpu_baseline leg of bench.py, never by
 * the product path.  Built with -ffp-contract=off: every float product and
 * sum is rounded on its own, in the reference's operation order, so list
 * membership (decided by signs and comparisons of float expressions) and the
 * sub-facet sums are reproducible bit for bit.
 *
 * Follows SH/main.cpp (= 3905/Shape_Selfheating/main.cpp):
 *   initialise        :750-809  centroids ((v0+v1)+v2)/3, unnormalised normal
 *                               (v1-v0)x(v2-v0), area |n|/2, edge vectors
 *                               normcalc1 = v1-v0, normcalc2 = v0-v2,
 *                               normcalc3 = v2-v1
 *   selfheating       :628-744  every ordered pair I != J: facing when
 *                               acos(cos) <= PI/2 on both facets (float acos,
 *                               PI = 4 atan(1) in float); centroid view factor
 *                               cos_i cos_j A_J / (PI d^2) clamped at 0 (:709);
 *                               VF vector A_I VF d/|d|; J ascending
 *   hidden            :187-368  pair (G,Z) dropped when a facet S among G's
 *                               facing partners, nearer to Z than G is, with
 *                               t = (s.n_S)/(g.n_S) > 0, holds the point
 *                               Z + t g inside its three edges
 *   betterviewfactors :374-622  pairs with VF >= 0.05: both triangles split
 *                               into n*n sub-triangles (vertex and facet
 *                               tables :466-539); VF = sum_j (a_i/A_I)
 *                               sum_k cos cos a_k/(PI d^2), each sub-pair
 *                               term clamped at 0
 *   main              :817-835  total VF and photon momentum vector per facet
 * and T3905/main.cpp (= 3905/Thermal/NominalPole/main.cpp)
 *   surfacegeometry   :816-860  crater sub-facet normals and trivectors
 *   calculate_surface_global_viewfactors :1054-1181
 *                               drf[k][j] = VF_k vis_j cos_j / dir_z with a
 *                               m*m sample rim visibility count, capped at 1
 * and the crater sunlight of the same crater model: surfacegeometry
 * (T3905/main.cpp:863-886) and sfcpartialshadow (T3905/GPU_code.cu:152-210),
 * restated as written (the reference compiles the latter with nvcc, whose
 * default contraction into fused multiply-adds would round differently).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "btpm_oracle.h"

typedef struct {
    float mid[3], n[3], area, nlen;
    float v[3][3];            /* vertices                      */
    float nc1[3], nc2[3], nc3[3];
} vf_facet;

static void facet_prep(const orc_vf_in *in, vf_facet *fc)
{
    for (int f = 0; f < in->nf; ++f) {
        vf_facet *c = &fc[f];
        const float *p[3];
        for (int a = 0; a < 3; ++a) p[a] = in->vtx + 3 * (size_t)in->fv[3 * f + a];
        for (int k = 0; k < 3; ++k) {
            c->mid[k] = (p[0][k] + p[1][k] + p[2][k]) / 3.0f;
            for (int a = 0; a < 3; ++a) c->v[a][k] = p[a][k];
        }
        float e0[3], e1[3];
        for (int k = 0; k < 3; ++k) {
            e0[k] = p[1][k] - p[0][k];
            e1[k] = p[2][k] - p[0][k];
        }
        c->n[0] = (e0[1] * e1[2]) - (e0[2] * e1[1]);
        c->n[1] = (e0[2] * e1[0]) - (e0[0] * e1[2]);
        c->n[2] = (e0[0] * e1[1]) - (e0[1] * e1[0]);
        c->nlen = sqrtf((c->n[0] * c->n[0]) + (c->n[1] * c->n[1]) + (c->n[2] * c->n[2]));
        c->area = c->nlen / 2.0f;
        for (int k = 0; k < 3; ++k) {
            c->nc1[k] = e0[k];
            c->nc2[k] = -1.0f * e1[k];
            c->nc3[k] = p[2][k] - p[1][k];
        }
    }
}

/* SH/main.cpp:652-662 / :689-706: the pair test and the two cosines */
static int facing(const vf_facet *I, const vf_facet *J, float PI, float *ti, float *tj,
                  float sep[3], float *sd)
{
    for (int k = 0; k < 3; ++k) sep[k] = J->mid[k] - I->mid[k];
    *sd = sqrtf((sep[0] * sep[0]) + (sep[1] * sep[1]) + (sep[2] * sep[2]));
    *ti = acosf(((I->n[0] * sep[0]) + (I->n[1] * sep[1]) + (I->n[2] * sep[2])) / (I->nlen * *sd));
    *tj = acosf(((-1.0f * J->n[0] * sep[0]) + (-1.0f * J->n[1] * sep[1]) + (-1.0f * J->n[2] * sep[2])) /
                (J->nlen * *sd));
    return *ti <= (PI / 2.0f) && *tj <= (PI / 2.0f);
}

/* SH/main.cpp:218-330: 1 if some S among G's partners hides Z from G */
static int is_hidden(const vf_facet *fc, const int *dep, int ndep, int G, int Z)
{
    const vf_facet *g_ = &fc[G], *z_ = &fc[Z];
    float gv[3];
    for (int k = 0; k < 3; ++k) gv[k] = g_->mid[k] - z_->mid[k];
    for (int j = 0; j < ndep; ++j) {
        const int Sx = dep[j];
        if (Sx == G || Sx == Z) continue;
        const vf_facet *S = &fc[Sx];
        float sv[3];
        for (int k = 0; k < 3; ++k) sv[k] = S->mid[k] - z_->mid[k];
        const float g = sqrtf((gv[0] * gv[0]) + (gv[1] * gv[1]) + (gv[2] * gv[2]));
        const float s = sqrtf((sv[0] * sv[0]) + (sv[1] * sv[1]) + (sv[2] * sv[2]));
        if (!(g > s)) continue;
        const float tc1 = (sv[0] * S->n[0]) + (sv[1] * S->n[1]) + (sv[2] * S->n[2]);
        const float tc2 = (gv[0] * S->n[0]) + (gv[1] * S->n[1]) + (gv[2] * S->n[2]);
        const float t = tc1 / tc2;
        if (!(t > 0.0f)) continue;
        float x[3], x1[3], x2[3], x3[3], c1[3], c2[3], c3[3];
        for (int k = 0; k < 3; ++k) {
            x[k] = (t * gv[k]) + z_->mid[k];
            x1[k] = x[k] - S->v[0][k];
            x3[k] = x[k] - S->v[1][k];
            x2[k] = x[k] - S->v[2][k];
        }
        c1[0] = (S->nc1[1] * x1[2]) - (S->nc1[2] * x1[1]);
        c1[1] = (S->nc1[2] * x1[0]) - (S->nc1[0] * x1[2]);
        c1[2] = (S->nc1[0] * x1[1]) - (S->nc1[1] * x1[0]);
        c2[0] = (S->nc2[1] * x2[2]) - (S->nc2[2] * x2[1]);
        c2[1] = (S->nc2[2] * x2[0]) - (S->nc2[0] * x2[2]);
        c2[2] = (S->nc2[0] * x2[1]) - (S->nc2[1] * x2[0]);
        c3[0] = (S->nc3[1] * x3[2]) - (S->nc3[2] * x3[1]);
        c3[1] = (S->nc3[2] * x3[0]) - (S->nc3[0] * x3[2]);
        c3[2] = (S->nc3[0] * x3[1]) - (S->nc3[1] * x3[0]);
        const float k1 = (c1[0] * S->n[0]) + (c1[1] * S->n[1]) + (c1[2] * S->n[2]);
        const float k2 = (c2[0] * S->n[0]) + (c2[1] * S->n[1]) + (c2[2] * S->n[2]);
        const float k3 = (c3[0] * S->n[0]) + (c3[1] * S->n[1]) + (c3[2] * S->n[2]);
        if (k1 >= 0.0f && k2 >= 0.0f && k3 >= 0.0f) return 1;
    }
    return 0;
}

/* SH/main.cpp:466-539: sub-vertex positions and sub-facet centroids of a
 * triangle split n x n; rows j = 1..n+1 of j vertices from vertex 2 */
static void subdivide(const vf_facet *c, int n, float *vpos, int *tri, float *mid)
{
    float t1[3], t2[3];
    for (int k = 0; k < 3; ++k) {
        t1[k] = (c->v[1][k] - c->v[0][k]) / (float)n;
        t2[k] = (c->v[2][k] - c->v[0][k]) / (float)n;
    }
    int l = 0;
    for (int j = 1; j != n + 2; j++)
        for (int k = 1; k != j + 1; k++, l++)
            for (int q = 0; q < 3; ++q)
                vpos[3 * l + q] = c->v[2][q] + ((float)(k - 1) * t1[q]) - ((float)(j - 1) * t2[q]);
    l = 0;
    for (int j = 1; j != n + 1; j++) {
        if (j != 1)
            for (int k = 1; k != j; k++) {
                tri[3 * l] = ((j * (j - 1)) / 2) + k;
                tri[3 * l + 1] = ((j * (j - 1)) / 2) + k - 1;
                tri[3 * l + 2] = ((j * (j + 1)) / 2) + k - 1;
                l++;
                tri[3 * l] = ((j * (j + 1)) / 2) + k - 1;
                tri[3 * l + 1] = ((j * (j + 1)) / 2) + k;
                tri[3 * l + 2] = ((j * (j - 1)) / 2) + k;
                l++;
            }
        tri[3 * l] = (((j + 1) * (j + 2)) / 2) - 2;
        tri[3 * l + 1] = (((j + 1) * (j + 2)) / 2) - 1;
        tri[3 * l + 2] = ((j * (j + 1)) / 2) - 1;
        l++;
    }
    for (int j = 0; j < n * n; ++j)
        for (int q = 0; q < 3; ++q)
            mid[3 * j + q] = (vpos[3 * tri[3 * j] + q] + vpos[3 * tri[3 * j + 1] + q] +
                              vpos[3 * tri[3 * j + 2] + q]) / 3.0f;
}

void orc_vf_subdivide(const float *v9, int n, float *vpos, int *tri, float *mid)
{
    vf_facet c;
    memcpy(c.v, v9, sizeof(c.v));
    subdivide(&c, n, vpos, tri, mid);
}

/* SH/main.cpp:544-615: the refined view factor and vector of pair (I,J) */
static float refine_pair(const vf_facet *I, const vf_facet *J, int n, float PI,
                         float *m1, float *m2, float *scratch_v, int *scratch_t,
                         float *tot, float *totv, float vec[3])
{
    const int nsf = n * n;
    subdivide(I, n, scratch_v, scratch_t, m1);
    subdivide(J, n, scratch_v, scratch_t, m2);
    const float a1 = I->area / (float)nsf, a2 = J->area / (float)nsf;
    for (int j = 0; j < nsf; ++j) {
        tot[j] = 0.0f;
        totv[3 * j] = totv[3 * j + 1] = totv[3 * j + 2] = 0.0f;
        for (int k = 0; k < nsf; ++k) {
            float sep[3];
            for (int q = 0; q < 3; ++q) sep[q] = m2[3 * k + q] - m1[3 * j + q];
            const float sd = sqrtf((sep[0] * sep[0]) + (sep[1] * sep[1]) + (sep[2] * sep[2]));
            const float ci = ((I->n[0] * sep[0]) + (I->n[1] * sep[1]) + (I->n[2] * sep[2])) / (I->nlen * sd);
            const float cj = ((-1.0f * J->n[0] * sep[0]) + (-1.0f * J->n[1] * sep[1]) +
                              (-1.0f * J->n[2] * sep[2])) / (J->nlen * sd);
            float vf = (ci * cj * a2) / (PI * sd * sd);
            if (vf < 0.0f) vf = 0.0f;
            tot[j] = tot[j] + vf;
            for (int q = 0; q < 3; ++q) sep[q] = sep[q] / sd;
            for (int q = 0; q < 3; ++q) totv[3 * j + q] = totv[3 * j + q] + (vf * sep[q]);
        }
    }
    float eff = 0.0f;
    vec[0] = vec[1] = vec[2] = 0.0f;
    for (int j = 0; j < nsf; ++j) {
        eff = eff + ((a1 / I->area) * tot[j]);
        for (int q = 0; q < 3; ++q) vec[q] = vec[q] + (a1 * totv[3 * j + q]);
    }
    return eff;
}

int orc_viewfactors(const orc_vf_in *in, orc_vf_out *out)
{
    const float PI = 4.0f * atanf(1.0f);
    const int nf = in->nf, n = in->nsub;
    memset(out, 0, sizeof(*out));
    vf_facet *fc = (vf_facet *)malloc(sizeof(vf_facet) * (size_t)nf);
    facet_prep(in, fc);

    /* selfheating: counts, then the ordered lists (SH/main.cpp:641-735) */
    int *rp1 = (int *)malloc(sizeof(int) * (size_t)(nf + 1));
    rp1[0] = 0;
#pragma omp parallel for schedule(dynamic, 16)
    for (int I = 0; I < nf; ++I) {
        int c = 0;
        for (int J = 0; J < nf; ++J) {
            if (J == I) continue;
            float ti, tj, sep[3], sd;
            c += facing(&fc[I], &fc[J], PI, &ti, &tj, sep, &sd);
        }
        rp1[I + 1] = c;
    }
    for (int I = 0; I < nf; ++I) rp1[I + 1] += rp1[I];
    const long long n1 = rp1[nf];
    int *col1 = (int *)malloc(sizeof(int) * (size_t)(n1 + 1));
    float *vf1 = (float *)malloc(sizeof(float) * (size_t)(n1 + 1));
    float *vec1 = (float *)malloc(sizeof(float) * 3 * (size_t)(n1 + 1));
#pragma omp parallel for schedule(dynamic, 16)
    for (int I = 0; I < nf; ++I) {
        int e = rp1[I];
        for (int J = 0; J < nf; ++J) {
            if (J == I) continue;
            float ti, tj, sep[3], sd;
            if (!facing(&fc[I], &fc[J], PI, &ti, &tj, sep, &sd)) continue;
            float v = (cosf(ti) * cosf(tj) * fc[J].area) / (PI * sd * sd);
            if (v < 0.0f) v = 0.0f;
            for (int k = 0; k < 3; ++k) sep[k] = sep[k] / sd;
            col1[e] = J;
            vf1[e] = v;
            for (int k = 0; k < 3; ++k) vec1[3 * (size_t)e + k] = fc[I].area * v * sep[k];
            ++e;
        }
    }

    /* hidden (SH/main.cpp:208-360) */
    unsigned char *keep = (unsigned char *)malloc((size_t)(n1 + 1));
#pragma omp parallel for schedule(dynamic, 4)
    for (int G = 0; G < nf; ++G) {
        const int *dep = col1 + rp1[G];
        const int nd = rp1[G + 1] - rp1[G];
        for (int i = 0; i < nd; ++i)
            keep[rp1[G] + i] = in->cull ? (unsigned char)!is_hidden(fc, dep, nd, G, dep[i]) : 1;
    }
    int *rp = (int *)malloc(sizeof(int) * (size_t)(nf + 1));
    rp[0] = 0;
    for (int G = 0; G < nf; ++G) {
        int c = 0;
        for (int e = rp1[G]; e < rp1[G + 1]; ++e) c += keep[e];
        rp[G + 1] = rp[G] + c;
    }
    const long long nnz = rp[nf];
    out->n_facing = n1;
    out->nnz = nnz;
    out->rowptr = rp;
    out->col = (int *)malloc(sizeof(int) * (size_t)(nnz + 1));
    out->vf = (float *)malloc(sizeof(float) * (size_t)(nnz + 1));
    out->vec = (float *)malloc(sizeof(float) * 3 * (size_t)(nnz + 1));
    out->dir = (float *)malloc(sizeof(float) * 3 * (size_t)(nnz + 1));
    out->refined = (unsigned char *)calloc((size_t)(nnz + 1), 1);
    out->total = (float *)malloc(sizeof(float) * (size_t)nf);
    out->photon = (float *)malloc(sizeof(float) * 3 * (size_t)nf);
    for (int G = 0; G < nf; ++G) {
        int o = rp[G];
        for (int e = rp1[G]; e < rp1[G + 1]; ++e)
            if (keep[e]) {
                out->col[o] = col1[e];
                out->vf[o] = vf1[e];
                memcpy(out->vec + 3 * (size_t)o, vec1 + 3 * (size_t)e, 3 * sizeof(float));
                ++o;
            }
    }
    free(keep);
    free(vec1);
    free(vf1);
    free(col1);
    free(rp1);

    /* betterviewfactors (SH/main.cpp:401-620) */
    long long nref = 0;
    if (n > 0) {
#pragma omp parallel reduction(+ : nref)
        {
            const int nsf = n * n, nsv = (n + 1) * (n + 2) / 2;
            float *m1 = (float *)malloc(sizeof(float) * 3 * (size_t)nsf);
            float *m2 = (float *)malloc(sizeof(float) * 3 * (size_t)nsf);
            float *sv = (float *)malloc(sizeof(float) * 3 * (size_t)nsv);
            int *st = (int *)malloc(sizeof(int) * 3 * (size_t)nsf);
            float *tot = (float *)malloc(sizeof(float) * (size_t)nsf);
            float *totv = (float *)malloc(sizeof(float) * 3 * (size_t)nsf);
#pragma omp for schedule(dynamic, 4)
            for (int I = 0; I < nf; ++I)
                for (int e = rp[I]; e < rp[I + 1]; ++e) {
                    if (!(out->vf[e] >= in->refine_min)) continue;
                    out->vf[e] = refine_pair(&fc[I], &fc[out->col[e]], n, PI, m1, m2, sv, st, tot,
                                             totv, out->vec + 3 * (size_t)e);
                    out->refined[e] = 1;
                    ++nref;
                }
            free(totv);
            free(tot);
            free(st);
            free(sv);
            free(m2);
            free(m1);
        }
    }
    out->n_refined = nref;

    /* main (SH/main.cpp:822-835) and the vector-list value (:874-878) */
    for (int I = 0; I < nf; ++I) {
        float t = 0.0f, p[3] = {0.0f, 0.0f, 0.0f};
        for (int e = rp[I]; e < rp[I + 1]; ++e) {
            t = t + out->vf[e];
            for (int k = 0; k < 3; ++k) p[k] = p[k] + (out->vec[3 * (size_t)e + k] / fc[I].area);
            for (int k = 0; k < 3; ++k)
                out->dir[3 * (size_t)e + k] = out->vec[3 * (size_t)e + k] / (out->vf[e] * fc[I].area);
        }
        out->total[I] = rp[I + 1] > rp[I] ? t : 0.0f;
        for (int k = 0; k < 3; ++k) out->photon[3 * I + k] = p[k];
    }
    free(fc);
    return 0;
}

void orc_vf_free(orc_vf_out *out)
{
    free(out->rowptr);
    free(out->col);
    free(out->vf);
    free(out->vec);
    free(out->dir);
    free(out->refined);
    free(out->total);
    free(out->photon);
    memset(out, 0, sizeof(*out));
}

/* T3905/main.cpp:1054-1181.  The facet frame (rows: unit first edge, n x e,
 * unit normal) in double from the double vertices; the rest in float. */
void orc_crater_drf(const orc_drf_in *in, float *drf)
{
    const int m = in->m, ncf = in->ncf;
    /* surfacegeometry (T3905/main.cpp:819-860): crater unit normals and
     * trivectors */
    float *cn = (float *)malloc(sizeof(float) * 3 * (size_t)ncf);
    float *ctv = (float *)malloc(sizeof(float) * 6 * (size_t)ncf);
    float *cv2 = (float *)malloc(sizeof(float) * 3 * (size_t)ncf);
    for (int j = 0; j < ncf; ++j) {
        const float *p0 = in->cvtx + 3 * (size_t)in->cfv[3 * j];
        const float *p1 = in->cvtx + 3 * (size_t)in->cfv[3 * j + 1];
        const float *p2 = in->cvtx + 3 * (size_t)in->cfv[3 * j + 2];
        float a[3], b[3], nn[3];
        for (int k = 0; k < 3; ++k) {
            a[k] = p1[k] - p0[k];
            b[k] = p2[k] - p0[k];
            cv2[3 * j + k] = p2[k];
        }
        nn[0] = (a[1] * b[2]) - (a[2] * b[1]);
        nn[1] = (a[2] * b[0]) - (a[0] * b[2]);
        nn[2] = (a[0] * b[1]) - (a[1] * b[0]);
        const float l = sqrtf((nn[0] * nn[0]) + (nn[1] * nn[1]) + (nn[2] * nn[2]));
        for (int k = 0; k < 3; ++k) {
            cn[3 * j + k] = nn[k] / l;
            ctv[6 * j + k] = -1.0f * b[k] / (float)m;
            ctv[6 * j + 3 + k] = a[k] / (float)m;
        }
    }
    /* trivector multipliers (T3905/main.cpp:1101-1114) */
    float *mul = (float *)malloc(sizeof(float) * 2 * (size_t)m * m);
    for (int k = 1; k != m + 1; ++k)
        for (int l = 1; l != k + 1; ++l) {
            const int p = ((k - 1) * (k - 1)) + (2 * (l - 1));
            mul[2 * p] = (float)k - (1.0f / 3.0f);
            mul[2 * p + 1] = (float)l - (2.0f / 3.0f);
        }
    for (int k = 2; k != m + 1; ++k)
        for (int l = 1; l != k; ++l) {
            const int p = ((k - 1) * (k - 1)) + (2 * (l - 1)) + 1;
            mul[2 * p] = (float)k - (2.0f / 3.0f);
            mul[2 * p + 1] = (float)l - (1.0f / 3.0f);
        }
    const float step = 1.0f / (float)(m * m);
    const float R = in->crater_radius;
#pragma omp parallel for schedule(dynamic, 16)
    for (int f = 0; f < in->nf; ++f) {
        if (in->rowptr[f + 1] == in->rowptr[f]) continue;
        const double *p0 = in->vtx + 3 * (size_t)in->fv[3 * f];
        const double *p1 = in->vtx + 3 * (size_t)in->fv[3 * f + 1];
        const double *p2 = in->vtx + 3 * (size_t)in->fv[3 * f + 2];
        double e[3], b[3], nrm[3], y[3];
        for (int k = 0; k < 3; ++k) {
            e[k] = p1[k] - p0[k];
            b[k] = p2[k] - p0[k];
        }
        nrm[0] = (e[1] * b[2]) - (e[2] * b[1]);
        nrm[1] = (e[2] * b[0]) - (e[0] * b[2]);
        nrm[2] = (e[0] * b[1]) - (e[1] * b[0]);
        const double nl = sqrt((nrm[0] * nrm[0]) + (nrm[1] * nrm[1]) + (nrm[2] * nrm[2]));
        for (int k = 0; k < 3; ++k) nrm[k] = nrm[k] / nl;
        const double x = sqrt((e[0] * e[0]) + (e[1] * e[1]) + (e[2] * e[2]));
        for (int k = 0; k < 3; ++k) e[k] = e[k] / x;
        y[0] = (nrm[1] * e[2]) - (nrm[2] * e[1]);
        y[1] = (nrm[2] * e[0]) - (nrm[0] * e[2]);
        y[2] = (nrm[0] * e[1]) - (nrm[1] * e[0]);
        const double *M[3] = {e, y, nrm};
        for (int q = in->rowptr[f]; q < in->rowptr[f + 1]; ++q) {
            const float *d = in->dir + 3 * (size_t)q;
            float w[3];
            for (int r = 0; r < 3; ++r)
                w[r] = (float)((M[r][0] * d[0]) + (M[r][1] * d[1]) + (M[r][2] * d[2]));
            const float a = fabsf(w[2]) / sqrtf((w[0] * w[0]) + (w[1] * w[1]));
            for (int j = 0; j < ncf; ++j) {
                float vis = 1.0f;
                const float ang = (w[0] * cn[3 * j]) + (w[1] * cn[3 * j + 1]) + (w[2] * cn[3 * j + 2]);
                if (ang < 0.0f) vis = 0.0f;
                for (int p = 0; p != m * m; ++p) {
                    float mp[3];
                    for (int k = 0; k < 3; ++k)
                        mp[k] = cv2[3 * j + k] + (mul[2 * p] * ctv[6 * j + k]) +
                                (mul[2 * p + 1] * ctv[6 * j + 3 + k]);
                    const float tc1 = (-1.0f * w[1] * mp[0]) + (w[0] * mp[1]);
                    const float sr = sqrtf((R * R) - (tc1 * tc1));
                    const float tc0 = (w[0] * mp[0]) + (w[1] * mp[1]) - sr;
                    if (mp[2] <= (a * tc0) && vis != 0.0f) vis = vis - step;
                    if (vis < 0.0f) vis = 0.0f;
                }
                float v = (in->vf[q] * vis * ang) / w[2];
                if (v > 1.0f) v = 1.0f;
                drf[(size_t)q * ncf + j] = v;
            }
        }
    }
    free(mul);
    free(cv2);
    free(ctv);
    free(cn);
}

/* T3905/main.cpp:863-886 and T3905/GPU_code.cu:152-210 */
void orc_crater_sun(const orc_crater_sun_in *in, float *shadow, float *cosang)
{
    const int m = in->m, ncf = in->ncf, ns = m * m;
    float *cn = (float *)malloc(sizeof(float) * 3 * (size_t)ncf);
    float *mp = (float *)malloc(sizeof(float) * 3 * (size_t)ncf * ns);
    float *mul = (float *)malloc(sizeof(float) * 2 * (size_t)ns);
    for (int k = 1; k != m + 1; ++k)
        for (int l = 1; l != k + 1; ++l) {
            const int p = ((k - 1) * (k - 1)) + (2 * (l - 1));
            mul[2 * p] = (float)k - (1.0f / 3.0f);
            mul[2 * p + 1] = (float)l - (2.0f / 3.0f);
        }
    for (int k = 2; k != m + 1; ++k)
        for (int l = 1; l != k; ++l) {
            const int p = ((k - 1) * (k - 1)) + (2 * (l - 1)) + 1;
            mul[2 * p] = (float)k - (2.0f / 3.0f);
            mul[2 * p + 1] = (float)l - (1.0f / 3.0f);
        }
    for (int j = 0; j < ncf; ++j) {
        const float *p0 = in->cvtx + 3 * (size_t)in->cfv[3 * j];
        const float *p1 = in->cvtx + 3 * (size_t)in->cfv[3 * j + 1];
        const float *p2 = in->cvtx + 3 * (size_t)in->cfv[3 * j + 2];
        float a[3], b[3], nn[3], t0[3], t1[3];
        for (int k = 0; k < 3; ++k) {
            a[k] = p1[k] - p0[k];
            b[k] = p2[k] - p0[k];
        }
        nn[0] = (a[1] * b[2]) - (a[2] * b[1]);
        nn[1] = (a[2] * b[0]) - (a[0] * b[2]);
        nn[2] = (a[0] * b[1]) - (a[1] * b[0]);
        const float l = sqrtf((nn[0] * nn[0]) + (nn[1] * nn[1]) + (nn[2] * nn[2]));
        for (int k = 0; k < 3; ++k) {
            cn[3 * j + k] = nn[k] / l;
            t0[k] = -1.0f * b[k] / (float)m;
            t1[k] = a[k] / (float)m;
        }
        for (int p = 0; p < ns; ++p)
            for (int k = 0; k < 3; ++k)
                mp[(3 * (size_t)j * ns) + 3 * p + k] =
                    p2[k] + (mul[2 * p] * t0[k]) + (mul[2 * p + 1] * t1[k]);
    }
    const float R = in->crater_radius, step = 1.0f / (float)(m * m);
#pragma omp parallel for schedule(static)
    for (long long q = 0; q < (long long)in->nf * in->ntau; ++q) {
        const float *sv = in->sun + 3 * (size_t)q;
        const int lit = in->shadow_in[q];
        const float t = sqrtf((sv[0] * sv[0]) + (sv[1] * sv[1]) + (sv[2] * sv[2]));
        float w[3], a = 0.0f;
        if (lit == 1) {
            const float s = sqrtf((sv[0] * sv[0]) + (sv[1] * sv[1]));
            for (int k = 0; k < 3; ++k) w[k] = sv[k] / s;
            a = fabsf(w[2]) / sqrtf((w[0] * w[0]) + (w[1] * w[1]));
        }
        for (int j = 0; j < ncf; ++j) {
            const size_t o = (size_t)q * ncf + j;
            const float ill = ((sv[0] * cn[3 * j]) + (sv[1] * cn[3 * j + 1]) + (sv[2] * cn[3 * j + 2])) / t;
            float sh = 1.0f;
            if (lit == 0) sh = 0.0f;
            if (ill <= 0.0f) sh = 0.0f;
            if (lit == 1)
                for (int p = 0; p != ns; ++p) {
                    const float *x = mp + (3 * (size_t)j * ns) + 3 * p;
                    const float tc1 = (-1.0f * w[1] * x[0]) + (w[0] * x[1]);
                    const float sr = sqrtf((R * R) - (tc1 * tc1));
                    const float tc0 = (w[0] * x[0]) + (w[1] * x[1]) - sr;
                    if (x[2] <= (a * tc0) && sh != 0.0f) sh = sh - step;
                    if (sh < 0.0f) sh = 0.0f;
                }
            cosang[o] = ill;
            shadow[o] = sh;
        }
    }
    free(mul);
    free(mp);
    free(cn);
}
