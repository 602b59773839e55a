// btpm_kernels.h -- prototypes of every kernel shared between translation
// units.  The defining files include this header too, so a definition that
// drifts from its prototype is a compile error (conflicting extern "C"
// declarations) rather than a silent argument mismatch at launch.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "btpm_internal.h"

// conduct.hip
// smooth conduction, W = 256, 128 or 64 columns per workgroup
template <int W>
__global__ void btpm_conduct_smooth(const btpm_tile* tiles, const btpm_set_dev* sets,
                                    const btpm_status_dev* status);
extern template __global__ void btpm_conduct_smooth<256>(const btpm_tile*, const btpm_set_dev*,
                                                         const btpm_status_dev*);
extern template __global__ void btpm_conduct_smooth<128>(const btpm_tile*, const btpm_set_dev*,
                                                         const btpm_status_dev*);
extern template __global__ void btpm_conduct_smooth<64>(const btpm_tile*, const btpm_set_dev*,
                                                        const btpm_status_dev*);
extern "C" __global__ void btpm_conduct_crater(const btpm_tile* tiles, const btpm_set_dev* sets,
                                               const btpm_status_dev* status);
extern "C" __global__ void btpm_contract_kernel(const _Float16* vfA, const float* e, float* q,
                                                int ncol, float unscale_v,
                                                const float* unscale_p);

// aux.hip
extern "C" __global__ void btpm_pack_blocked(const float* raw, float* blk, const float* scale,
                                             int ncols, int ntau, int nb);
extern "C" __global__ void btpm_unpack_blocked(const float* blk, float* raw, int ncols, int ntau,
                                               int nb);
extern "C" __global__ void btpm_pack_crater(const float* raw, float* blk, int nfac, int ntau,
                                            int nb);
extern "C" __global__ void btpm_unpack_crater(const float* blk, float* raw, int nfac, int ntau,
                                              int nb);
extern "C" __global__ void btpm_init_smooth(btpm_set_dev S);
extern "C" __global__ void btpm_warm_smooth(btpm_set_dev S);
extern "C" __global__ void btpm_init_crater(btpm_set_dev S, const float* vf32);
extern "C" __global__ void btpm_pack_drf(const int* rowptr, const float* drf, const int* cbase,
                                         float* drf_scale, _Float16* drfA);
extern "C" __global__ void btpm_shape_to_crater(const int* rowptr, const int* colidx,
                                                const _Float16* drfA, const int* cbase,
                                                const float* drf_scale, const float* w, int nfac,
                                                int ntau, int nb, const float* base,
                                                const float* base_scale, float base_gain,
                                                float* dst);
extern "C" __global__ void btpm_weights_fm(const float* x, const float* scale, float gain,
                                           int mode, float es, int ncols, int nb, float* out);
extern "C" __global__ void btpm_crater_scatter(const float* vf32, const float* albedo,
                                               const float* shadow, const float* cosang,
                                               const float* sun, const float* vis_table,
                                               int n_el, int n_az, const float* normals,
                                               const float* rgi, float solar, int multiple,
                                               float acc, int max_iter, int nfac, int ntau,
                                               int nb, float* illum_out);
extern "C" __global__ void btpm_spmm_blocked(const int* rowptr, const int* colidx,
                                             const float* vals, const float* x, const float* x0,
                                             const float* alpha, float* y, int nrows, int nb,
                                             int bpt, const btpm_status_dev* st);
extern "C" __global__ void btpm_selfheat_tiled(const int* cptr, const int2* ent, int nf,
                                               int nchunk, const int* slots,
                                               const btpm_set_dev* sets,
                                               const btpm_status_dev* status);
extern "C" __global__ void btpm_selfheat_ell(const int2* wseg, const int2* ent, int nf,
                                             int nchunk, const int* slots,
                                             const btpm_set_dev* sets,
                                             const btpm_status_dev* status);
extern "C" __global__ void btpm_selfheat_ell_h32(const int2* wseg, const int2* ent, int nf,
                                                 int nchunk, const int* slots,
                                                 const btpm_set_dev* sets,
                                                 const btpm_status_dev* status);
extern "C" __global__ void btpm_selfheat_grp(const int2* gseg, const int4* gent, const int* gperm,
                                             int nf, int nchunk, const int* slots,
                                             const btpm_set_dev* sets,
                                             const btpm_status_dev* status);
extern "C" __global__ void btpm_selfheat_rows(const int* cptr, const int2* ent, int nf,
                                              int nchunk, const int* slots,
                                              const btpm_set_dev* sets,
                                              const btpm_status_dev* status);
extern "C" __global__ void btpm_scale_cols(float* y, const float* x, const float* scale,
                                           float gain, int ncols, size_t n);
extern "C" __global__ void btpm_delta_max(float* p, const float* y, float* part, size_t n);
extern "C" __global__ void btpm_finish_scatter(float* F, const float* p, const float* d,
                                               const float* albedo, int ncols, size_t n);
extern "C" __global__ void btpm_status_kernel(const int* ids, const btpm_set_dev* sets,
                                              btpm_status_dev* status);

// geometry.hip
extern "C" __global__ void btpm_geom_facets(const double* equpos, const int* fv, int nf,
                                            int nv_primary, const double* trig,
                                            const double* frame, int s0, int nsteps, double* rec,
                                            int* lit, float* cosang, float* sun, double sun_scale);
extern "C" __global__ void btpm_geom_shadow(const double* rec, int nf, int nsteps, int* lit);
extern "C" __global__ void btpm_geom_finish(const int* lit, const float* cosang, const float* sun,
                                            int nf, int nsteps, int s0, int ntau, float solar,
                                            int* shadow, float* illum, float* sun_out);

extern "C" __global__ void btpm_unpack_crater_rows(const float* in, float* out, int ncols, int c0,
                                                   int nc, int ntau, int nb);

// viewfactor.hip
extern "C" __global__ void btpm_vf_prep(const float* vtx, const int* fv, int nf, float* rec);
extern "C" __global__ void btpm_vf_facing(const float* rec, int nf, const int* rowptr, int* count,
                                          int* col, float* vf, float* vec);
extern "C" __global__ void btpm_vf_hidden(const float* rec, const int* rowptr1, const int* col1,
                                          const int2* work, unsigned char* keep);
extern "C" __global__ void btpm_vf_kept(const int* rowptr1, const unsigned char* keep, int* count);
extern "C" __global__ void btpm_vf_compact(const int* rowptr1, const int* col1, const float* vf1,
                                           const float* vec1, const unsigned char* keep,
                                           const int* rowptr, int* col, float* vf, float* vec,
                                           int refine, float refine_min, int* nref,
                                           int2* reflist);
extern "C" __global__ void btpm_vf_refine(const float* rec, const int* col, const int2* reflist,
                                          int nsub, float* vf, float* vec);
extern "C" __global__ void btpm_vf_finish(const int* rowptr, const float* vf, const float* vec,
                                          const float* rec, float* dir, float* total,
                                          float* photon);
extern "C" __global__ void btpm_crater_sun(const int* lit_in, const float* sun, const float4* csub,
                                           const float4* samp, const float* vtab, int ncf,
                                           int nsamp, float radius, long long n, float* shadow,
                                           float* cosang);
extern "C" __global__ void btpm_drf_frames(const double* vtx, const int* fv, int nf, double* frame);
extern "C" __global__ void btpm_drf_rows(const int* rowptr, int* row);
extern "C" __global__ void btpm_drf(const double* frame, const int* row, const float* vfv,
                                    const float* dir, const float4* csub, const float4* samp,
                                    const float* vtab, int ncf, int nsamp, float radius,
                                    long long nnz, float* drf);

// observe.hip
extern "C" __global__ void btpm_obs_geom(const double* equpos, const int* fv, int nf,
                                         int nv_primary, const double* trig, const double* frame,
                                         int s0, int nsteps, int ntau, const int* visible,
                                         double* rec, double* area);
extern "C" __global__ void btpm_obs_flux(const double* rec, const double* area,
                                         const float* stemp, const float* cvis,
                                         const float* ctemp, const double4* csub, int ncf,
                                         double stot, int rough, const double* pl, double eps,
                                         int nf, int ntau, int s0, int nsteps, double* out);
extern "C" __global__ void btpm_obs_sum(const double* terms, int nf, int nsteps, const double* bn,
                                        const double* colour, int tmin, int tcount, double* smooth,
                                        double* rough);

// stepwise.hip
extern "C" __global__ void btpm_step_profile(float* temp, float* prof, int ncols, int to_records);
extern "C" __global__ void btpm_step_fd(btpm_set_dev S, float* temp);
extern "C" __global__ void btpm_step_bc_smooth(btpm_set_dev S, float* temp, const float* Tp, int s,
                                               const int* rowptr, const int* colidx,
                                               const float* vf, float* dm);
extern "C" __global__ void btpm_step_bc_crater(btpm_set_dev S, float* temp, const float* Tp, int s,
                                               const float* vf32, float* dm);
extern "C" __global__ void btpm_step_converged(btpm_set_dev S, const float* dm, int check);
