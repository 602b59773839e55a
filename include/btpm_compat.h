/*
 * btpm_compat.h -- the reference's extern "C" crater-path host API, symbol for
 * symbol, implemented by libbtpm on gfx950.
 *
 * Replaces 175706/Run-04/thermal_GPU.cu:698-1279 (the SoA variant with
 * numrevs and IR_ABS); the calling translation unit declares these same
 * functions at 175706/Run-04/thermal.cpp:2379, 2460, 2496, 2579, 2693, 2769,
 * 2840, 2920, 3003, 3073, 3443-3459.  Linking libbtpm.so in place of
 * thermal_GPU.cu keeps thermal.cpp unchanged except for one added call to
 * btpm_compat_configure() before the first initialise_* call (the reference
 * fixes the dimensions at compile time, thermal_GPU.cu:4-10).
 *
 * Array layouts are the reference's: per-facet arrays [NF], crater arrays
 * [NF][NTAU][100], the crater profile [NF][NX+1][100], viewfactors[i][j] =
 * VF(j -> i).  Like the reference, every call is synchronous and returns
 * nothing; a failure is recorded and reported by btpm_compat_status().
 */
#ifndef BTPM_COMPAT_H
#define BTPM_COMPAT_H

#include <stdbool.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- additions (no counterpart in the reference) ------------------------- */

/* Dimensions the reference compiles in (ASTNUMFACE per part, NTAU, NX; the
 * crater has 100 sub-facets).  nx must be 56.  Returns 0 or a BTPM_ERR_*.  */
int btpm_compat_configure(int nfacets, int ntau, int nx);
/* 0 while every compat call succeeded, else the first error code; the text
 * is in btpm_last_error().                                                  */
int btpm_compat_status(void);

/* ---- reference types (thermal_GPU.cu:36-62) ------------------------------ */
struct GPU_SfcVertex {
    float pos[3];
};

struct GPU_SfcFacet {
    int sfcvertices[3];
    float midpos[3];
    float trivector[2][3];
    float normcalc[3][3];
    float normal[3];
};

struct GPU_Surface {
    int NV, NF;
    struct GPU_SfcVertex *SV;
    struct GPU_SfcFacet *SF;
};

/* ---- reference entry points (thermal_GPU.cu line of each definition) ----- */
void set_cuda_device(void);                                              /* :699  */
void initialise_GPU_constants(float rot_period_H, float viewfactors_H[100][100],
                              struct GPU_Surface S_H);                   /* :710  */
void initialise_GPU_facet_properties_arrays(void);                       /* :795  */
void initialise_GPU_shadow_arrays(void);                                 /* :817  */
void initialise_GPU_illumination_arrays(void);                           /* :840  */
void initialise_GPU_globalillumination_arrays(void);                     /* :857  */
void initialise_GPU_globalheating_arrays(void);                          /* :874  */
void initialise_GPU_temperature_arrays(void);                            /* :891  */
void freeup_GPU_shadow_arrays(void);                                     /* :911  */
void freeup_GPU_globalillumination_arrays(void);                         /* :927  */
void freeup_GPU_globalheating_arrays(void);                              /* :940  */
void freeup_GPU_temperature_arrays(void);                                /* :953  */
void set_GPU_facet_properties(bool *sfc_converged_H, float *albedo_H,
                              float *thermal_inertia_H, float *numrevs_H); /* :970 */
void load_GPU_globalillumination(float *sfc_globalillumination_H);       /* :994  */
void load_GPU_globalheating(float *sfc_globalheating_H);                 /* :1013 */
void obtain_GPU_facet_properties(bool *sfc_converged_H, float *albedo_H,
                                 float *thermal_inertia_H, float *numrevs_H); /* :1032 */
void obtain_GPU_roughness_shadows(int *shadow_H, float *sfc_shadow_H,
                                  float *sfc_illumvector_H, float *sfc_illumangle_H); /* :1056 */
void obtain_GPU_roughness_temperatures(float *sfc_stemp0_H, float *sfc_stemp1_H,
                                       float *sfc_temp_H);               /* :1081 */
/* Copies the global-illumination array, as the reference does (:1108-1111). */
void obtain_GPU_roughness_illumination(float *sfc_globalillumination_H); /* :1103 */
void GPU_sfcpartialshadow(int *shadow_H, float *sfc_shadow_H, float *sfc_illumvector_H,
                          float *sfc_illumangle_H);                      /* :1122 */
void GPU_single_scattering(float solar);                                 /* :1155 */
void GPU_multiple_scattering(float solar, float scattering_accuracy);    /* :1174 */
void GPU_zero_temperatures(void);                                        /* :1193 */
void GPU_initialise_temperatures(float EPS, float SIGMA, float IR_ABS);   /* :1208 */
void GPU_inertia_difference(float EPS, float SIGMA, float IR_ABS, float TACC); /* :1226 */
void GPU_inst_equilibrium(float EPS, float SIGMA, float IR_ABS);          /* :1250 */
void GPU_converged(float TACC);                                          /* :1272 */

#ifdef __cplusplus
}
#endif
#endif
