// compat.hip -- the reference's extern "C" crater-path host API (declared in
// include/btpm_compat.h) over libbtpm's kernels.
//
// State lives in one module object, as the reference keeps module-global
// device pointers (thermal_GPU.cu:70-165); every call is synchronous.  The
// crater conduction runs through btpm_conduct_crater with a single resident
// set whose freeze test is left to GPU_converged, so the host-visible control
// flow (one GPU_inertia_difference per rotation, GPU_converged after
// rev > 5, obtain_GPU_facet_properties to read the flags) is the reference's.
#include <hip/hip_runtime.h>

#include <climits>
#include <cmath>
#include <memory>
#include <vector>

#include "../../include/btpm.h"
#include "../../include/btpm_compat.h"
#include "btpm_internal.h"
#include "btpm_kernels.h"
#include "runtime.h"

using namespace btpm;


namespace {

constexpr int NS = BTPM_NS;
constexpr int kM = 10;  // partial-shadow subdivision: m*m samples per sub-facet

// src[b][c][k] = (1 - A_f) illum[b][c][k] + rgh[f][s][j]  (rgh in the
// reference layout; null = 0)
__global__ void compat_source(const float* __restrict__ illum, const float* __restrict__ rgh,
                              const float* __restrict__ oma, float* __restrict__ src, int nf,
                              int ntau, int nb) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int ncols = nf * NS;
  if (i >= (size_t)ncols * nb) return;
  const int b = (int)(i / ncols);
  const int c = (int)(i - (size_t)b * ncols);
  const int f = c / NS, j = c - f * NS;
  const float a = oma[f];
  for (int k = 0; k < BTPM_BLK; ++k) {
    const int s = min(b * BTPM_BLK + k, ntau - 1);
    const size_t o = i * BTPM_BLK + k;
    src[o] = a * illum[o] + (rgh ? rgh[((size_t)f * ntau + s) * NS + j] : 0.0f);
  }
}

// stemp0 <- stemp1 for facets not yet converged (boundary_condition's copy,
// thermal_GPU.cu:522-523), on blocked arrays
__global__ void compat_keep_previous(const float* __restrict__ T, float* __restrict__ Tprev,
                                     const uint8_t* __restrict__ frozen, int ncols, size_t n) {
  // grid-stride: a crater series can hold more than 2^32 elements
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const int c = (int)((i / BTPM_BLK) % ncols);
    if (!frozen[c / NS]) Tprev[i] = T[i];
  }
}

// converged kernel (thermal_GPU.cu:648-692) from the rotation's max change
__global__ void compat_converged(uint8_t* __restrict__ frozen, const float* __restrict__ dmax,
                                 float tacc, int nf) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f < nf && !frozen[f] && dmax[f] < tacc) frozen[f] = 1;
}

__global__ void compat_profile_out(const float* __restrict__ prof, float* __restrict__ out,
                                   int nf, int nx) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;  // over [f][i][j]
  if (i >= (size_t)nf * (nx + 1) * NS) return;
  const int j = (int)(i % NS);
  const int layer = (int)((i / NS) % (nx + 1));
  const int f = (int)(i / ((size_t)NS * (nx + 1)));
  out[i] = prof[((size_t)f * NS + j) * BTPM_PSTRIDE + layer];
}

__global__ void compat_profile_in(const float* __restrict__ in, float* __restrict__ prof, int nf,
                                  int nx) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (size_t)nf * (nx + 1) * NS) return;
  const int j = (int)(i % NS);
  const int layer = (int)((i / NS) % (nx + 1));
  const int f = (int)(i / ((size_t)NS * (nx + 1)));
  prof[((size_t)f * NS + j) * BTPM_PSTRIDE + layer] = in[i];
}

// sfcpartialshadow (thermal_GPU.cu:172-234): the fraction of m*m sample
// points of each crater sub-facet from which the Sun is not hidden by the
// hemispherical rim of radius R.  A point at height z below the rim plane,
// at coordinate a along the Sun's azimuth and b across it, sees the Sun iff
// the ray leaves through the rim circle: z > tan(e) (a - sqrt(R^2 - b^2)).
// Starts from the sub-facet's input value and removes 1/m^2 per hidden
// sample, clamped at 0; facets with shadow != 1 are left untouched.
__global__ __launch_bounds__(128) void compat_partial_shadow(
    const int* __restrict__ shadow, float* __restrict__ sfc_shadow,
    const float* __restrict__ sunvec, const float* __restrict__ base,  // [100][3]
    const float* __restrict__ tri,                                     // [100][2][3]
    const float* __restrict__ mult,                                    // [m*m][2]
    float radius, int ntau) {
  const int s = blockIdx.x, f = blockIdx.y, j = threadIdx.x;
  if (j >= NS) return;
  if (shadow[(size_t)f * ntau + s] != 1) return;
  const float* u = sunvec + ((size_t)f * ntau + s) * 3;
  const float sh = sqrtf(u[0] * u[0] + u[1] * u[1]);
  if (!(sh > 0.0f)) return;
  const float ux = u[0] / sh, uy = u[1] / sh, tan_e = fabsf(u[2] / sh);
  const size_t o = ((size_t)f * ntau + s) * NS + j;
  float v = sfc_shadow[o];
  for (int p = 0; p < kM * kM; ++p) {
    const float mx = base[j * 3 + 0] + mult[2 * p] * tri[j * 6 + 0] + mult[2 * p + 1] * tri[j * 6 + 3];
    const float my = base[j * 3 + 1] + mult[2 * p] * tri[j * 6 + 1] + mult[2 * p + 1] * tri[j * 6 + 4];
    const float mz = base[j * 3 + 2] + mult[2 * p] * tri[j * 6 + 2] + mult[2 * p + 1] * tri[j * 6 + 5];
    const float across = -uy * mx + ux * my;
    const float along = ux * mx + uy * my - sqrtf(radius * radius - across * across);
    if (mz <= tan_e * along && v != 0.0f) v -= 1.0f / (float)(kM * kM);
    if (v < 0.0f) v = 0.0f;
  }
  sfc_shadow[o] = v;
}

struct Compat {
  int configured = 0;
  int err = BTPM_OK;
  int nf = 0, ntau = 0, nb = 0;
  int r_bad = -1;         // first facet with dtau/dz^2 outside the stable range
  float r_bad_val = 0.0f;
  btpm_ctx* ctx = nullptr;
  float rot_period = 1.0f;
  std::vector<float> vf;  // [100][100], vf[i*100+j] = VF(j->i)
  float vf_unscale = 1.0f;
  DBuf vfA, vf32;
  // crater geometry for the partial shadows
  DBuf g_base, g_tri, g_mult;
  float radius = 10.0f;
  // per facet
  std::vector<uint8_t> h_frozen;
  std::vector<float> h_albedo, h_ti, h_numrevs;
  DBuf frozen, order, oma, albedo, r, m1, dmax;
  // reference-layout inputs
  DBuf shadow, sfc_shadow, illumvec, illumangle, rgi, rgh;
  // blocked crater state
  DBuf illum, src, T, Tprev, prof, tile_active, tile_work;
  bool src_dirty = true;
  DBuf d_set, d_status, d_tiles;
  int ntiles = 0;
};

Compat& C() {
  static Compat c;
  return c;
}

// run f() with the module's error handling; returns false if skipped/failed
template <class F>
bool guarded(const char* name, F&& f) {
  Compat& c = C();
  if (c.err != BTPM_OK) return false;
  if (!c.configured) {
    c.err = fail(BTPM_ERR_STATE, "%s: btpm_compat_configure() was not called", name);
    return false;
  }
  try {
    if (!c.ctx) {
      btpm_ctx* x = nullptr;
      const int rc = btpm_ctx_create(0, &x);
      if (rc != BTPM_OK) {
        c.err = rc;
        return false;
      }
      c.ctx = x;
    }
    HIPCHK(hipSetDevice(c.ctx->device));
    StreamScope scope(c.ctx->stream);  // buffers released in f() stay on this stream
    f();
    HIPCHK(hipStreamSynchronize(c.ctx->stream));
    return true;
  } catch (const HipError& e) {
    c.err = e.code;
    return false;
  } catch (const std::bad_alloc&) {
    c.err = fail(BTPM_ERR_NOMEM, "%s: host allocation failed", name);
    return false;
  } catch (const std::exception& e) {
    c.err = fail(BTPM_ERR_STATE, "%s: internal error: %s", name, e.what());
    return false;
  }
}

size_t ser() { return (size_t)C().nf * C().ntau * NS; }
size_t blk() { return (size_t)C().nb * C().nf * NS * BTPM_BLK; }

void coefficients() {
  Compat& c = C();
  std::vector<float> r(c.nf), m1(c.nf), oma(c.nf);
  const double dz = 8.0 / BTPM_NX;
  c.r_bad = -1;
  for (int f = 0; f < c.nf; ++f) {
    const double n = c.h_numrevs[f];
    r[f] = (float)((2.0 * 3.14159265358979323846 * n / c.ntau) / (dz * dz));
    if (!(r[f] >= 0.0f && r[f] <= (float)BTPM_R_MAX) && c.r_bad < 0) {
      c.r_bad = f;  // reported when conduction is asked for (shadows need no r)
      c.r_bad_val = r[f];
    }
    m1[f] = (float)(c.h_ti[f] * std::sqrt(2.0 * 3.14159265358979323846 * n / c.rot_period) / dz);
    oma[f] = 1.0f - c.h_albedo[f];
  }
  hipStream_t st = c.ctx->stream;
  upload(c.r, r.data(), r.size(), st);
  upload(c.m1, m1.data(), m1.size(), st);
  upload(c.oma, oma.data(), oma.size(), st);
  upload(c.albedo, c.h_albedo.data(), c.h_albedo.size(), st);
}

btpm_set_dev make_set(float es, float esia, float tol, int inst) {
  Compat& c = C();
  btpm_set_dev d{};
  d.kind = BTPM_SET_CRATER;
  d.ncols = c.nf * NS;
  d.ngroups = c.nf;
  d.ntau = c.ntau;
  d.nb = c.nb;
  d.max_rev = INT_MAX;
  d.check_after = INT_MAX;  // freezing is GPU_converged's job
  d.inst = inst;
  d.es = es;
  d.esia = esia;
  d.tacc = 0.0f;
  d.newton_tol = tol;
  d.newton_cap = 200;
  d.vf_unscale = c.vf_unscale;
  d.src = c.src.as<float>();
  d.T = c.T.as<float>();
  d.prof = c.prof.as<float>();
  d.r = c.r.as<float>();
  d.m1 = c.m1.as<float>();
  d.frozen = c.frozen.as<uint8_t>();
  d.order = c.order.as<int>();
  d.dmax = c.dmax.as<float>();
  d.vfA = reinterpret_cast<const float*>(c.vfA.p);
  d.tile_active = c.tile_active.as<int>();
  d.tile_work = c.tile_work.as<int>();
  d.ntiles = c.ntiles;
  d.tile_cols = BTPM_CRATERS_WG;
  return d;
}

void ensure_source() {
  Compat& c = C();
  if (!c.src_dirty) return;
  if (!c.illum.p) throw HipError{fail(BTPM_ERR_STATE, "crater illumination not computed")};
  if (!c.src.p) c.src.alloc(blk() * 4);
  hipLaunchKernelGGL(compat_source, dim3(grid1((size_t)c.nf * NS * c.nb, 256)), dim3(256), 0,
                     c.ctx->stream, c.illum.as<float>(), c.rgh.as<float>(), c.oma.as<float>(),
                     c.src.as<float>(), c.nf, c.ntau, c.nb);
  c.src_dirty = false;
}

void rotation(float EPS, float SIGMA, float IR_ABS, float tol, int inst) {
  Compat& c = C();
  hipStream_t st = c.ctx->stream;
  if (!inst && c.r_bad >= 0)
    throw HipError{fail(BTPM_ERR_ARG, "facet %d: dtau/dz^2 = %.6g outside [0, %g] (FTCS stability "
                        "limit)", c.r_bad, (double)c.r_bad_val, BTPM_R_MAX)};
  ensure_source();
  const size_t n = blk();
  hipLaunchKernelGGL(compat_keep_previous, dim3(grid_stride(n, 256)), dim3(256), 0, st,
                     c.T.as<float>(), c.Tprev.as<float>(), c.frozen.as<uint8_t>(), c.nf * NS, n);
  btpm_set_dev d = make_set(EPS * SIGMA, IR_ABS * EPS * SIGMA, tol, inst);
  HIPCHK(hipMemcpyAsync(c.d_set.p, &d, sizeof d, hipMemcpyHostToDevice, st));
  hipLaunchKernelGGL(btpm_conduct_crater, dim3(c.ntiles), dim3(BTPM_CRATER_THREADS), 0, st,
                     c.d_tiles.as<btpm_tile>(), c.d_set.as<btpm_set_dev>(),
                     c.d_status.as<btpm_status_dev>());
  HIPCHK(hipGetLastError());
}

}  // namespace

extern "C" {

int btpm_compat_configure(int nfacets, int ntau, int nx) {
  if (nfacets < 1 || ntau < 1) return fail(BTPM_ERR_ARG, "empty compat dimensions");
  if (nx != BTPM_NX) return fail(BTPM_ERR_ARG, "nx must be %d (got %d)", BTPM_NX, nx);
  Compat& c = C();
  c.nf = nfacets;
  c.ntau = ntau;
  c.nb = nblocks(ntau);
  c.configured = 1;
  c.err = BTPM_OK;
  c.h_frozen.assign(nfacets, 0);
  c.h_albedo.assign(nfacets, 0.0f);
  c.h_ti.assign(nfacets, 0.0f);
  c.h_numrevs.assign(nfacets, 1.0f);
  return BTPM_OK;
}

int btpm_compat_status(void) { return C().err; }

void set_cuda_device(void) {
  guarded("set_cuda_device", [] {});
}

void initialise_GPU_constants(float rot_period_H, float viewfactors_H[100][100],
                              struct GPU_Surface S_H) {
  guarded("initialise_GPU_constants", [&] {
    Compat& c = C();
    hipStream_t st = c.ctx->stream;
    c.rot_period = rot_period_H;
    c.vf.assign(&viewfactors_H[0][0], &viewfactors_H[0][0] + NS * NS);
    for (int i = 0; i < NS; ++i) c.vf[i * NS + i] = 0.0f;
    std::vector<_Float16> a;
    c.vf_unscale = vf_tiles(c.vf.data(), a);
    upload(c.vfA, a.data(), a.size(), st);
    upload(c.vf32, c.vf.data(), c.vf.size(), st);
    // partial-shadow geometry: sample base vertex (the facet's third vertex)
    // and the two edge steps, as initialise_GPU_constants stores them
    if (S_H.SV && S_H.SF && S_H.NF >= NS) {
      std::vector<float> base(NS * 3), tri(NS * 6);
      float rmax = 0.0f;
      for (int v = 0; v < S_H.NV; ++v) {
        const float* p = S_H.SV[v].pos;
        rmax = std::max(rmax, std::sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]));
      }
      for (int j = 0; j < NS; ++j) {
        const float* p = S_H.SV[S_H.SF[j].sfcvertices[2]].pos;
        for (int k = 0; k < 3; ++k) {
          base[j * 3 + k] = p[k];
          tri[j * 6 + k] = S_H.SF[j].trivector[0][k];
          tri[j * 6 + 3 + k] = S_H.SF[j].trivector[1][k];
        }
      }
      // sample multipliers of the m x m subdivision (thermal_GPU.cu:768-783)
      std::vector<float> mult(kM * kM * 2);
      for (int k = 1; k <= kM; ++k)
        for (int l = 1; l <= k; ++l) {
          const int p = (k - 1) * (k - 1) + 2 * (l - 1);
          mult[2 * p] = (float)k - 1.0f / 3.0f;
          mult[2 * p + 1] = (float)l - 2.0f / 3.0f;
        }
      for (int k = 2; k <= kM; ++k)
        for (int l = 1; l < k; ++l) {
          const int p = (k - 1) * (k - 1) + 2 * (l - 1) + 1;
          mult[2 * p] = (float)k - 2.0f / 3.0f;
          mult[2 * p + 1] = (float)l - 1.0f / 3.0f;
        }
      c.radius = rmax;
      upload(c.g_base, base.data(), base.size(), st);
      upload(c.g_tri, tri.data(), tri.size(), st);
      upload(c.g_mult, mult.data(), mult.size(), st);
    }
  });
}

void initialise_GPU_facet_properties_arrays(void) {
  guarded("initialise_GPU_facet_properties_arrays", [] {
    Compat& c = C();
    c.frozen.alloc(c.nf);
    c.dmax.alloc((size_t)c.nf * 4);
    HIPCHK(hipMemsetAsync(c.frozen.p, 0, c.nf, c.ctx->stream));
    HIPCHK(hipMemsetAsync(c.dmax.p, 0, (size_t)c.nf * 4, c.ctx->stream));
    c.ntiles = (c.nf + BTPM_CRATERS_WG - 1) / BTPM_CRATERS_WG;
    std::vector<btpm_tile> t(c.ntiles);
    for (int i = 0; i < c.ntiles; ++i) t[i] = btpm_tile{0, i};
    upload(c.d_tiles, t.data(), t.size(), c.ctx->stream);
    c.tile_active.alloc((size_t)c.ntiles * 4);
    c.tile_work.alloc((size_t)c.ntiles * 4);
    c.d_set.alloc(sizeof(btpm_set_dev));
    // the crater kernel integrates order[0..norder): here every facet, in
    // order; frozen facets are skipped by their flag (GPU_converged sets it)
    std::vector<int> id((size_t)c.nf);
    for (int f = 0; f < c.nf; ++f) id[f] = f;
    upload(c.order, id.data(), id.size(), c.ctx->stream);
    btpm_status_dev z{};
    z.norder = c.nf;
    upload(c.d_status, &z, 1, c.ctx->stream);
    coefficients();
  });
}

void initialise_GPU_shadow_arrays(void) {
  guarded("initialise_GPU_shadow_arrays", [] {
    Compat& c = C();
    c.shadow.alloc((size_t)c.nf * c.ntau * 4);
    c.sfc_shadow.alloc(ser() * 4);
    c.illumvec.alloc((size_t)c.nf * c.ntau * 3 * 4);
    c.illumangle.alloc(ser() * 4);
  });
}

void initialise_GPU_illumination_arrays(void) {
  guarded("initialise_GPU_illumination_arrays", [] { C().illum.alloc(blk() * 4); });
}

void initialise_GPU_globalillumination_arrays(void) {
  guarded("initialise_GPU_globalillumination_arrays", [] { C().rgi.alloc(ser() * 4); });
}

void initialise_GPU_globalheating_arrays(void) {
  guarded("initialise_GPU_globalheating_arrays", [] {
    C().rgh.alloc(ser() * 4);
    C().src_dirty = true;
  });
}

void initialise_GPU_temperature_arrays(void) {
  guarded("initialise_GPU_temperature_arrays", [] {
    Compat& c = C();
    c.T.alloc(blk() * 4);
    c.Tprev.alloc(blk() * 4);
    c.prof.alloc((size_t)c.nf * NS * BTPM_PSTRIDE * 4);
  });
}

void freeup_GPU_shadow_arrays(void) {
  guarded("freeup_GPU_shadow_arrays", [] {
    Compat& c = C();
    c.shadow.release();
    c.sfc_shadow.release();
    c.illumvec.release();
    c.illumangle.release();
  });
}

void freeup_GPU_globalillumination_arrays(void) {
  guarded("freeup_GPU_globalillumination_arrays", [] { C().rgi.release(); });
}

void freeup_GPU_globalheating_arrays(void) {
  guarded("freeup_GPU_globalheating_arrays", [] {
    C().rgh.release();
    C().src_dirty = true;
  });
}

void freeup_GPU_temperature_arrays(void) {
  guarded("freeup_GPU_temperature_arrays", [] {
    Compat& c = C();
    c.T.release();
    c.Tprev.release();
    c.prof.release();
    c.src.release();
    c.src_dirty = true;
  });
}

void set_GPU_facet_properties(bool* sfc_converged_H, float* albedo_H, float* thermal_inertia_H,
                              float* numrevs_H) {
  guarded("set_GPU_facet_properties", [&] {
    Compat& c = C();
    for (int f = 0; f < c.nf; ++f) {
      c.h_frozen[f] = sfc_converged_H[f] ? 1 : 0;
      c.h_albedo[f] = albedo_H[f];
      c.h_ti[f] = thermal_inertia_H[f];
      c.h_numrevs[f] = numrevs_H[f];
    }
    if (!c.frozen.p) c.frozen.alloc(c.nf);
    HIPCHK(hipMemcpyAsync(c.frozen.p, c.h_frozen.data(), c.nf, hipMemcpyHostToDevice,
                          c.ctx->stream));
    coefficients();
    c.src_dirty = true;
  });
}

void load_GPU_globalillumination(float* sfc_globalillumination_H) {
  guarded("load_GPU_globalillumination", [&] {
    Compat& c = C();
    if (!c.rgi.p) c.rgi.alloc(ser() * 4);
    HIPCHK(hipMemcpyAsync(c.rgi.p, sfc_globalillumination_H, ser() * 4, hipMemcpyHostToDevice,
                          c.ctx->stream));
  });
}

void load_GPU_globalheating(float* sfc_globalheating_H) {
  guarded("load_GPU_globalheating", [&] {
    Compat& c = C();
    if (!c.rgh.p) c.rgh.alloc(ser() * 4);
    HIPCHK(hipMemcpyAsync(c.rgh.p, sfc_globalheating_H, ser() * 4, hipMemcpyHostToDevice,
                          c.ctx->stream));
    c.src_dirty = true;
  });
}

void obtain_GPU_facet_properties(bool* sfc_converged_H, float* albedo_H, float* thermal_inertia_H,
                                 float* numrevs_H) {
  guarded("obtain_GPU_facet_properties", [&] {
    Compat& c = C();
    HIPCHK(hipMemcpyAsync(c.h_frozen.data(), c.frozen.p, c.nf, hipMemcpyDeviceToHost,
                          c.ctx->stream));
    HIPCHK(hipStreamSynchronize(c.ctx->stream));
    for (int f = 0; f < c.nf; ++f) {
      sfc_converged_H[f] = c.h_frozen[f] != 0;
      albedo_H[f] = c.h_albedo[f];
      thermal_inertia_H[f] = c.h_ti[f];
      numrevs_H[f] = c.h_numrevs[f];
    }
  });
}

void obtain_GPU_roughness_shadows(int* shadow_H, float* sfc_shadow_H, float* sfc_illumvector_H,
                                  float* sfc_illumangle_H) {
  guarded("obtain_GPU_roughness_shadows", [&] {
    Compat& c = C();
    hipStream_t st = c.ctx->stream;
    HIPCHK(hipMemcpyAsync(shadow_H, c.shadow.p, (size_t)c.nf * c.ntau * 4, hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(sfc_shadow_H, c.sfc_shadow.p, ser() * 4, hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(sfc_illumvector_H, c.illumvec.p, (size_t)c.nf * c.ntau * 12,
                          hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(sfc_illumangle_H, c.illumangle.p, ser() * 4, hipMemcpyDeviceToHost, st));
  });
}

void obtain_GPU_roughness_temperatures(float* sfc_stemp0_H, float* sfc_stemp1_H,
                                       float* sfc_temp_H) {
  guarded("obtain_GPU_roughness_temperatures", [&] {
    Compat& c = C();
    hipStream_t st = c.ctx->stream;
    DBuf raw;
    raw.alloc(ser() * 4);
    const unsigned g = grid1((size_t)c.nf * NS * c.nb, 256);
    hipLaunchKernelGGL(btpm_unpack_crater, dim3(g), dim3(256), 0, st, c.Tprev.as<float>(),
                       raw.as<float>(), c.nf, c.ntau, c.nb);
    HIPCHK(hipMemcpyAsync(sfc_stemp0_H, raw.p, raw.n, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    hipLaunchKernelGGL(btpm_unpack_crater, dim3(g), dim3(256), 0, st, c.T.as<float>(),
                       raw.as<float>(), c.nf, c.ntau, c.nb);
    HIPCHK(hipMemcpyAsync(sfc_stemp1_H, raw.p, raw.n, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    const size_t np = (size_t)c.nf * (BTPM_NX + 1) * NS;
    DBuf pr;
    pr.alloc(np * 4);
    hipLaunchKernelGGL(compat_profile_out, dim3(grid1(np, 256)), dim3(256), 0, st,
                       c.prof.as<float>(), pr.as<float>(), c.nf, BTPM_NX);
    HIPCHK(hipMemcpyAsync(sfc_temp_H, pr.p, np * 4, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));  // the copy has left pr before it is released
  });
}

void obtain_GPU_roughness_illumination(float* sfc_globalillumination_H) {
  guarded("obtain_GPU_roughness_illumination", [&] {
    Compat& c = C();
    HIPCHK(hipMemcpyAsync(sfc_globalillumination_H, c.rgi.p, ser() * 4, hipMemcpyDeviceToHost,
                          c.ctx->stream));
  });
}

void GPU_sfcpartialshadow(int* shadow_H, float* sfc_shadow_H, float* sfc_illumvector_H,
                          float* sfc_illumangle_H) {
  guarded("GPU_sfcpartialshadow", [&] {
    Compat& c = C();
    hipStream_t st = c.ctx->stream;
    if (!c.shadow.p) initialise_GPU_shadow_arrays();
    if (!c.g_base.p) throw HipError{fail(BTPM_ERR_STATE, "no crater geometry")};
    HIPCHK(hipMemcpyAsync(c.shadow.p, shadow_H, (size_t)c.nf * c.ntau * 4, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(c.sfc_shadow.p, sfc_shadow_H, ser() * 4, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(c.illumvec.p, sfc_illumvector_H, (size_t)c.nf * c.ntau * 12,
                          hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(c.illumangle.p, sfc_illumangle_H, ser() * 4, hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(compat_partial_shadow, dim3(c.ntau, c.nf), dim3(128), 0, st,
                       c.shadow.as<int>(), c.sfc_shadow.as<float>(), c.illumvec.as<float>(),
                       c.g_base.as<float>(), c.g_tri.as<float>(), c.g_mult.as<float>(), c.radius,
                       c.ntau);
    HIPCHK(hipGetLastError());
  });
}

static void scatter(float solar, int multiple, float acc) {
  Compat& c = C();
  if (!c.illum.p) c.illum.alloc(blk() * 4);
  hipLaunchKernelGGL(btpm_crater_scatter, dim3(c.nb, c.nf), dim3(128), 0, c.ctx->stream,
                     c.vf32.as<float>(), c.albedo.as<float>(), c.sfc_shadow.as<float>(),
                     c.illumangle.as<float>(), (const float*)nullptr, (const float*)nullptr, 0, 0,
                     (const float*)nullptr, c.rgi.as<float>(), solar, multiple, acc, 100000,
                     c.nf, c.ntau, c.nb, c.illum.as<float>());
  HIPCHK(hipGetLastError());
  c.src_dirty = true;
}

void GPU_single_scattering(float solar) {
  guarded("GPU_single_scattering", [&] { scatter(solar, 0, 0.0f); });
}

void GPU_multiple_scattering(float solar, float scattering_accuracy) {
  guarded("GPU_multiple_scattering", [&] { scatter(solar, 1, scattering_accuracy); });
}

void GPU_zero_temperatures(void) {
  guarded("GPU_zero_temperatures", [] {
    Compat& c = C();
    hipStream_t st = c.ctx->stream;
    HIPCHK(hipMemsetAsync(c.T.p, 0, c.T.n, st));
    HIPCHK(hipMemsetAsync(c.Tprev.p, 0, c.Tprev.n, st));
    HIPCHK(hipMemsetAsync(c.prof.p, 0, c.prof.n, st));
  });
}

void GPU_initialise_temperatures(float EPS, float SIGMA, float IR_ABS) {
  guarded("GPU_initialise_temperatures", [&] {
    Compat& c = C();
    hipStream_t st = c.ctx->stream;
    ensure_source();
    btpm_set_dev d = make_set(EPS * SIGMA, IR_ABS * EPS * SIGMA, 0.0f, 0);
    // init_crater clears frozen/dmax; the flags the host set are restored
    hipLaunchKernelGGL(btpm_init_crater, dim3(c.nf), dim3(128), 0, st, d, c.vf32.as<float>());
    HIPCHK(hipMemcpyAsync(c.Tprev.p, c.T.p, c.T.n, hipMemcpyDeviceToDevice, st));
    HIPCHK(hipMemcpyAsync(c.frozen.p, c.h_frozen.data(), c.nf, hipMemcpyHostToDevice, st));
    HIPCHK(hipGetLastError());
  });
}

void GPU_inertia_difference(float EPS, float SIGMA, float IR_ABS, float TACC) {
  // Newton to TACC/10, as the reference (thermal_GPU.cu:1231-1232)
  guarded("GPU_inertia_difference", [&] { rotation(EPS, SIGMA, IR_ABS, TACC / 10.0f, 0); });
}

void GPU_inst_equilibrium(float EPS, float SIGMA, float IR_ABS) {
  guarded("GPU_inst_equilibrium", [&] { rotation(EPS, SIGMA, IR_ABS, 1e-3f, 1); });
}

void GPU_converged(float TACC) {
  guarded("GPU_converged", [&] {
    Compat& c = C();
    hipLaunchKernelGGL(compat_converged, dim3(grid1(c.nf, 256)), dim3(256), 0, c.ctx->stream,
                       c.frozen.as<uint8_t>(), c.dmax.as<float>(), TACC, c.nf);
    HIPCHK(hipGetLastError());
  });
}

}  // extern "C"
