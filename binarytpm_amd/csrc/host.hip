// host.hip -- C++ runtime of libbtpm: contexts, device-resident shapes and
// solves, the batched rotation scheduler (hipGraph-captured launches,
// HBM-budgeted admission, on-device convergence status) and the
// continuous-batching server thread.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstdint>
#include <cstring>
#include <deque>
#include <functional>
#include <map>
#include <set>
#include <memory>
#include <mutex>
#include <string>
#include <system_error>
#include <thread>
#include <vector>

#include "../../include/btpm.h"
#include "btpm_internal.h"
#include "btpm_kernels.h"
#include "runtime.h"

// ---- kernels (conduct.hip, aux.hip) ----------------------------------------

using namespace btpm;

namespace btpm {
namespace {
struct Cached {
  void* p;
  hipStream_t tag;  // stream that may still use the block (null: idle)
};
struct Pool {
  std::mutex mu;
  std::map<int, std::multimap<size_t, Cached>> cache;  // device -> capacity -> block
  std::map<void*, std::pair<int, size_t>> live;        // block -> (device, capacity)
  std::map<int, size_t> cached;
  std::map<int, size_t> live_bytes;
  std::map<int, size_t> limit;                          // 0 / absent: none
};
Pool& the_pool() {
  static Pool* p = new Pool;  // never destroyed: DBufs of static objects may outlive main
  return *p;
}

}  // namespace

// Test hook: BTPM_INJECT_NOMEM=k (read when a batch or server starts) makes
// every k-th column-set build (smooth or crater) fail as if the device were
// full while another solve runs, so the scheduler's deferral path runs on
// demand.
int inject_period() {
  const char* e = std::getenv("BTPM_INJECT_NOMEM");
  return e ? std::max(0, std::atoi(e)) : 0;
}

hipStream_t& scope_stream() {
  thread_local hipStream_t s = nullptr;
  return s;
}

void* pool_alloc(size_t bytes) {
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  const size_t want = (bytes + 255) & ~(size_t)255;
  const hipStream_t tag = scope_stream();
  Pool& P = the_pool();
  size_t lim = 0;
  {
    std::lock_guard<std::mutex> lk(P.mu);
    auto& c = P.cache[dev];
    for (auto it = c.lower_bound(want); it != c.end() && it->first <= want + want / 4; ++it) {
      if (it->second.tag != nullptr && it->second.tag != tag) continue;
      void* p = it->second.p;
      const size_t cap = it->first;
      c.erase(it);
      P.cached[dev] -= cap;
      P.live[p] = {dev, cap};
      P.live_bytes[dev] += cap;
      return p;
    }
    lim = P.limit.count(dev) ? P.limit[dev] : 0;
  }
  // the device limit counts live and cached blocks: beyond it, the cache is
  // returned first, and the allocation fails as the driver's would
  auto over_limit = [&] {
    if (lim == 0) return false;
    std::lock_guard<std::mutex> lk(P.mu);
    return P.live_bytes[dev] + P.cached[dev] + want > lim;
  };
  void* p = nullptr;
  hipError_t e = over_limit() ? hipErrorOutOfMemory : hipMalloc(&p, want);
  size_t fr0 = 0, fr1 = 0, tot = 0, trimmed = 0;
  if (e != hipSuccess) {
    // return cached blocks, largest first, until the driver can hold the
    // request (an eighth spare); only if that fails, the whole cache: each
    // returned block costs a driver call, and after a large batch the cache
    // can hold a hundred of them
    (void)hipGetLastError();
    (void)hipMemGetInfo(&fr0, &tot);
    const size_t need = want + want / 8;
    trimmed = pool_trim_bytes(dev, need > fr0 ? need - fr0 : 0);
    e = over_limit() ? hipErrorOutOfMemory : hipMalloc(&p, want);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      trimmed += pool_cached(dev);
      pool_trim(dev);
      e = over_limit() ? hipErrorOutOfMemory : hipMalloc(&p, want);
    }
    (void)hipMemGetInfo(&fr1, &tot);
  }
  if (e != hipSuccess) {
    (void)hipGetLastError();
    const int code = e == hipErrorOutOfMemory ? BTPM_ERR_NOMEM : BTPM_ERR_HIP;
    size_t live = 0;
    {
      std::lock_guard<std::mutex> lk(P.mu);
      live = P.live_bytes[dev];
    }
    fail(code, "hipMalloc(%zu): %s (device free %.1f GB, %.1f GB after returning %.1f GB of "
         "cached blocks, of %.1f GB; the library holds %.1f GB%s)", want, hipGetErrorString(e),
         fr0 / 1e9, fr1 / 1e9, trimmed / 1e9, tot / 1e9, live / 1e9,
         lim ? ", at its device limit" : "");
    throw HipError{code};
  }
  std::lock_guard<std::mutex> lk(P.mu);
  P.live[p] = {dev, want};
  P.live_bytes[dev] += want;
  return p;
}

void pool_free(void* p) {
  Pool& P = the_pool();
  std::lock_guard<std::mutex> lk(P.mu);
  auto it = P.live.find(p);
  if (it == P.live.end()) {
    (void)hipFree(p);
    return;
  }
  const int dev = it->second.first;
  const size_t cap = it->second.second;
  P.live.erase(it);
  P.live_bytes[dev] -= cap;
  P.cache[dev].emplace(cap, Cached{p, scope_stream()});
  P.cached[dev] += cap;
}

size_t pool_cached(int device) {
  Pool& P = the_pool();
  std::lock_guard<std::mutex> lk(P.mu);
  auto it = P.cached.find(device);
  return it == P.cached.end() ? 0 : it->second;
}

size_t pool_live(int device) {
  Pool& P = the_pool();
  std::lock_guard<std::mutex> lk(P.mu);
  auto it = P.live_bytes.find(device);
  return it == P.live_bytes.end() ? 0 : it->second;
}

void pool_set_limit(int device, size_t bytes) {
  Pool& P = the_pool();
  std::lock_guard<std::mutex> lk(P.mu);
  P.limit[device] = bytes;
}

size_t pool_limit(int device) {
  Pool& P = the_pool();
  std::lock_guard<std::mutex> lk(P.mu);
  auto it = P.limit.find(device);
  return it == P.limit.end() ? 0 : it->second;
}

// hipFree waits for the device's work on the block, so trimming needs no tag
// return cached blocks of the device to the driver, largest first, until at
// least `bytes` are returned; the bytes returned
size_t pool_trim_bytes(int device, size_t bytes) {
  Pool& P = the_pool();
  std::vector<void*> blocks;
  size_t got = 0;
  {
    std::lock_guard<std::mutex> lk(P.mu);
    auto& c = P.cache[device];
    while (got < bytes && !c.empty()) {
      auto it = std::prev(c.end());
      blocks.push_back(it->second.p);
      got += it->first;
      P.cached[device] -= it->first;
      c.erase(it);
    }
  }
  int cur = 0;
  (void)hipGetDevice(&cur);
  (void)hipSetDevice(device);
  for (void* b : blocks) (void)hipFree(b);
  (void)hipSetDevice(cur);
  return got;
}

void pool_trim(int device) {
  Pool& P = the_pool();
  std::vector<void*> blocks;
  {
    std::lock_guard<std::mutex> lk(P.mu);
    for (auto& kv : P.cache[device]) blocks.push_back(kv.second.p);
    P.cache[device].clear();
    P.cached[device] = 0;
  }
  int cur = 0;
  (void)hipGetDevice(&cur);
  (void)hipSetDevice(device);
  for (void* b : blocks) (void)hipFree(b);
  (void)hipSetDevice(cur);
}

std::string& last_error() {
  thread_local std::string e;
  return e;
}

void noalloc_setup(int device);  // noalloc.hip

float vf_tiles(const float* vf, std::vector<_Float16>& a) {
  float vmax = 0.0f;
  for (int e = 0; e < BTPM_NS * BTPM_NS; ++e) vmax = std::max(vmax, std::fabs(vf[e]));
  int s = 0;
  if (vmax > 0.0f) {
    int ex = 0;
    std::frexp(vmax, &ex);  // vmax = m 2^ex, m in [0.5, 1)
    s = std::max(-14, std::min(24, -ex));
  }
  const float sc = std::ldexp(1.0f, s);
  a.assign(2 * 112 * 112, (_Float16)0.0f);
  for (int jj = 0; jj < BTPM_NS; ++jj)
    for (int ii = 0; ii < BTPM_NS; ++ii) {
      const float v = vf[ii * BTPM_NS + jj] * sc;
      const _Float16 hi = (_Float16)v;
      a[jj * 112 + ii] = hi;
      a[112 * 112 + jj * 112 + ii] = (_Float16)(v - (float)hi);
    }
  return std::ldexp(1.0f, -s);
}
}  // namespace btpm

namespace {
constexpr int kMaxSets = 1024;
const double kPi = 3.14159265358979323846;
}  // namespace

struct btpm_shape {
  btpm_ctx* ctx = nullptr;
  btpm_params p{};
  int nf = 0, nb = 0, nnz = 0;
  int has_crater = 0;
  DBuf F;          // [nb][nf][BLK] total facet illumination
  DBuf albedo;     // [nf]
  DBuf one_minus_a;
  DBuf numrevs;    // [nf]
  DBuf rowptr, colidx, vf;
  // self-heating SpMM layout: rows' entries sorted by emitter, and per
  // emitter chunk of BTPM_SH_CHUNK the first entry of every row, cptr[c][f]
  DBuf sent, cptr;  // sent: (emitter, float bits of VF) pairs
  int nchunk = 0;
  // the same entries lane-interleaved per (receiver block, chunk, wave) for
  // btpm_selfheat_ell: wseg = (first entry, entries per lane), ell = (emitter
  // within the chunk, float bits of VF); used while the padding
  // stays within BTPM_SH_ELL_MAX_PAD times the view-factor count
  DBuf wseg, ell;
  bool use_ell = false;
  // the same entries per (receiver block, chunk, wave, pass of 8 receivers)
  // for btpm_selfheat_grp: gseg = (first pair, entries rounded up to even),
  // gent = int4 pairs of (emitter within the chunk, VF bits), gperm = the
  // block's receivers by decreasing entry count (nf for an empty slot)
  DBuf gseg, gent, gperm;
  bool use_grp = false;
  size_t grp_slots = 0;  // padded entries of the grouped layout
  // BTPM_SH_LAYOUT forcing one self-heating kernel (no tuning): the kernel
  // code (see heat_kernel_code) or -1 for tuned
  int force_heat = -1;
  DBuf cvf32;      // [100][100] crater VF (fp32, vf[i*100+j] = VF(j->i))
  DBuf cvfA;       // [2][112][112] fp16 hi/lo of VF^T, zero padded
  DBuf cillum;     // [nb][nf*100][BLK] crater illumination (after scattering)
  // shape->crater view factors ([nnz][100] on input) packed per facet in
  // chunks of 32 entries as fp16 hi/lo tiles of drf * drf_scale (btpm_pack_drf)
  DBuf drfA;       // [chunks][2][112][32] halves
  DBuf cbase;      // [nf] first chunk of each facet
  DBuf drf_scale;  // [nf] power of two putting the facet's largest |drf| in [0.5, 1)
  // [nf][nb*BLK] facet-major weights of btpm_shape_to_crater (RGI at shape
  // creation, then each solve's RGH; every use is on the context's stream)
  DBuf wfm;
  float vf_unscale = 1.0f;
  std::vector<float> h_albedo, h_numrevs;
};

namespace {

struct SetHost {
  btpm_set_dev d{};
  int slot = -1;
  DBuf src, T, prof, r, m1, frozen, order, dmax, qext, emis, tile_active, tile_work;
  int ntiles = 0;
  int listed = 0;  // tiles in the current launch lists
};

// order[] starts as the identity: every group integrates
void alloc_order(SetHost* S, int ngroups, hipStream_t st) {
  std::vector<int> id((size_t)ngroups);
  for (int g = 0; g < ngroups; ++g) id[g] = g;
  upload(S->order, id.data(), id.size(), st);
  S->d.order = S->order.as<int>();
}

}  // namespace

struct btpm_solve {
  btpm_ctx* ctx = nullptr;
  btpm_shape* shape = nullptr;
  float ti = 0.0f;
  std::vector<float> ti_facet;
  int with_craters = 0;
  float flux_scale = 1.0f;  // multiplies every absorbed flux of the shape
  bool queued = false;      // handed to a server (not yet admitted)
  bool mem_waited = false;  // admission waited for device memory
  // 0 created, 1 smooth running, 2 crater running, 3 done,
  // 4 smooth done and the crater set waiting for device memory
  std::atomic<int> phase{0};
  std::unique_ptr<SetHost> smooth, crater;
  // warm start (the TI sweep's carried state): this solve's sets start from
  // the final state of `warm`; `dependents` counts solves that will still
  // copy this one's state (it cannot be destroyed before they have)
  btpm_solve* warm = nullptr;
  bool warm_pending = false;  // a copy from `warm` is still to be taken
  std::atomic<int> dependents{0};
  btpm_solve_info info{};
  std::chrono::steady_clock::time_point t_submit;
  std::mutex mu;
  std::condition_variable cv;
};

// ---------------------------------------------------------------------------
namespace {

void fill_set_common(btpm_set_dev& d, const btpm_params& p) {
  d.ntau = p.ntau;
  d.nb = nblocks(p.ntau);
  d.max_rev = p.max_rev;
  d.check_after = p.check_after;
  d.es = p.eps * p.sigma;
  d.esia = p.ir_abs * p.eps * p.sigma;
  d.tacc = p.tacc;
  d.newton_tol = p.newton_tol;
  d.newton_cap = p.newton_cap;
}

// per-group coefficients r = dtau/dz^2 and m1 = Gamma sqrt(2 pi n / P)/dz
void coefficients(const btpm_shape* sh, const btpm_solve* s, std::vector<float>& r,
                  std::vector<float>& m1, int* inst) {
  const int nf = sh->nf;
  const double dz = 8.0 / BTPM_NX;
  r.resize(nf);
  m1.resize(nf);
  int all_zero = 1;
  for (int f = 0; f < nf; ++f) {
    const double n = sh->h_numrevs[f];
    const double ti = s->ti_facet.empty() ? s->ti : s->ti_facet[f];
    if (ti != 0.0) all_zero = 0;
    r[f] = (float)((2.0 * kPi * n / sh->p.ntau) / (dz * dz));
    m1[f] = (float)(ti * std::sqrt(2.0 * kPi * n / sh->p.rot_period) / dz);
  }
  *inst = all_zero;
}

// Warm start.  A predecessor whose smooth set ran at TI 0 everywhere carries
// no depth profile (instantaneous equilibrium never integrates one): the
// successor then starts from the reference initial state, as the reference
// initialises temperatures at the first TI > 0 of its sweep
// (T3905/main.cpp:2503-2515).
const SetHost* warm_smooth_source(const btpm_solve* s) {
  const btpm_solve* w = s->warm;
  if (!s->warm_pending || !w || !w->smooth || !w->smooth->T.p || !w->smooth->prof.p) return nullptr;
  if (w->smooth->d.inst) return nullptr;
  return w->smooth.get();
}
// the crater set takes the predecessor's crater state iff both carry craters
bool crater_takes_warm(const btpm_solve* s) {
  return s->warm_pending && s->warm && s->with_craters && s->warm->with_craters;
}
const SetHost* warm_crater_source(const btpm_solve* s) {
  if (!crater_takes_warm(s)) return nullptr;
  const btpm_solve* w = s->warm;
  if (!w->crater || !w->crater->T.p || !w->crater->prof.p || w->crater->d.inst) return nullptr;
  return w->crater.get();
}
// the copies are enqueued on the context's stream before the predecessor
// can be destroyed (its destroy synchronises that stream first)
void release_warm(btpm_solve* s, hipStream_t) {
  s->warm_pending = false;
  if (s->warm) s->warm->dependents--;
}

// allocate and initialise the smooth column set of a solve
void build_smooth(btpm_solve* s, hipStream_t st) {
  btpm_shape* sh = s->shape;
  auto S = std::make_unique<SetHost>();
  btpm_set_dev& d = S->d;
  fill_set_common(d, sh->p);
  d.kind = BTPM_SET_SMOOTH;
  d.ncols = sh->nf;
  d.ngroups = sh->nf;
  const size_t nser = (size_t)d.nb * d.ncols * BTPM_BLK;
  std::vector<float> r, m1;
  coefficients(sh, s, r, m1, &d.inst);
  S->src.alloc(nser * sizeof(float));
  S->T.alloc(nser * sizeof(float));
  S->prof.alloc((size_t)BTPM_PSTRIDE * d.ncols * sizeof(float));
  upload(S->r, r.data(), r.size(), st);
  upload(S->m1, m1.data(), m1.size(), st);
  S->frozen.alloc(d.ncols);
  alloc_order(S.get(), d.ncols, st);
  S->dmax.alloc(d.ncols * sizeof(float));
  if (sh->nnz > 0) {
    S->qext.alloc(nser * sizeof(float));
    S->emis.alloc(nser * sizeof(float));
  }
  // tile arrays for the narrowest tuned width (64 columns per workgroup)
  S->ntiles = (d.ncols + 63) / 64;
  S->tile_active.alloc(S->ntiles * sizeof(int));
  S->tile_work.alloc(S->ntiles * sizeof(int));
  d.src = S->src.as<float>();
  d.T = S->T.as<float>();
  d.prof = S->prof.as<float>();
  d.r = S->r.as<float>();
  d.m1 = S->m1.as<float>();
  d.frozen = S->frozen.as<uint8_t>();
  d.dmax = S->dmax.as<float>();
  d.qext = S->qext.as<float>();
  d.emis = S->emis.as<float>();
  d.vfA = nullptr;
  d.tile_active = S->tile_active.as<int>();
  d.tile_work = S->tile_work.as<int>();
  d.ntiles = S->ntiles;
  d.tile_cols = BTPM_SMOOTH_WG;
  // src = (1-A) F
  hipLaunchKernelGGL(btpm_scale_cols, dim3(grid_stride(nser, 256)), dim3(256), 0, st, S->src.as<float>(),
                     sh->F.as<float>(), sh->one_minus_a.as<float>(), s->flux_scale, d.ncols, nser);
  const SetHost* W = warm_smooth_source(s);
  if (W) {
    // the previous solve's final depth profiles and last-rotation surface
    // temperatures; emission and freeze state derived as at initialisation
    HIPCHK(hipMemcpyAsync(S->T.p, W->T.p, nser * sizeof(float), hipMemcpyDeviceToDevice, st));
    HIPCHK(hipMemcpyAsync(S->prof.p, W->prof.p, (size_t)BTPM_PSTRIDE * d.ncols * sizeof(float),
                          hipMemcpyDeviceToDevice, st));
    hipLaunchKernelGGL(btpm_warm_smooth, dim3(grid1(d.ncols, 256)), dim3(256), 0, st, d);
  } else {
    hipLaunchKernelGGL(btpm_init_smooth, dim3(grid1(d.ncols, 256)), dim3(256), 0, st, d);
  }
  HIPCHK(hipGetLastError());
  s->smooth = std::move(S);
  if (s->warm_pending && !crater_takes_warm(s)) release_warm(s, st);
}

// RGH from the converged smooth temperatures, crater source and initial state
void build_crater(btpm_solve* s, hipStream_t st) {
  btpm_shape* sh = s->shape;
  auto S = std::make_unique<SetHost>();
  btpm_set_dev& d = S->d;
  fill_set_common(d, sh->p);
  d.kind = BTPM_SET_CRATER;
  d.ngroups = sh->nf;
  d.ncols = sh->nf * BTPM_NS;
  const size_t nser = (size_t)d.nb * d.ncols * BTPM_BLK;
  std::vector<float> r, m1;
  coefficients(sh, s, r, m1, &d.inst);
  S->src.alloc(nser * sizeof(float));
  S->T.alloc(nser * sizeof(float));
  S->prof.alloc((size_t)BTPM_PSTRIDE * d.ncols * sizeof(float));
  upload(S->r, r.data(), r.size(), st);
  upload(S->m1, m1.data(), m1.size(), st);
  S->frozen.alloc(d.ngroups);
  alloc_order(S.get(), d.ngroups, st);
  S->dmax.alloc(d.ngroups * sizeof(float));
  S->ntiles = (d.ngroups + BTPM_CRATERS_WG - 1) / BTPM_CRATERS_WG;
  S->tile_active.alloc(S->ntiles * sizeof(int));
  S->tile_work.alloc(S->ntiles * sizeof(int));
  d.src = S->src.as<float>();
  d.T = S->T.as<float>();
  d.prof = S->prof.as<float>();
  d.r = S->r.as<float>();
  d.m1 = S->m1.as<float>();
  d.frozen = S->frozen.as<uint8_t>();
  d.dmax = S->dmax.as<float>();
  d.qext = nullptr;
  d.emis = nullptr;
  d.vfA = reinterpret_cast<const float*>(sh->cvfA.p);
  d.vf_unscale = sh->vf_unscale;
  d.tile_active = S->tile_active.as<int>();
  d.tile_work = S->tile_work.as<int>();
  d.ntiles = S->ntiles;
  d.tile_cols = BTPM_CRATERS_WG;
  // src = (1-A) illum + RGH (RGH from the converged smooth T)
  if (sh->nnz > 0 && sh->drfA.p) {
    hipLaunchKernelGGL(btpm_weights_fm, dim3(grid1((size_t)d.nb * sh->nf, 256)), dim3(256), 0,
                       st, s->smooth->T.as<float>(), (const float*)nullptr, 1.0f, 1,
                       sh->p.eps * sh->p.sigma, sh->nf, d.nb, sh->wfm.as<float>());
    dim3 g(8 * grid1(sh->nf, 8) * grid1((size_t)d.nb * BTPM_BLK, 64));  // XCD-grouped facets
    hipLaunchKernelGGL(btpm_shape_to_crater, g, dim3(256), 0, st, sh->rowptr.as<int>(),
                       sh->colidx.as<int>(), sh->drfA.as<_Float16>(), sh->cbase.as<int>(),
                       sh->drf_scale.as<float>(), sh->wfm.as<float>(), sh->nf, sh->p.ntau, d.nb,
                       sh->cillum.as<float>(), sh->one_minus_a.as<float>(), s->flux_scale,
                       S->src.as<float>());
  } else {
    // (1-A_f) per crater column: scale with a per-column expansion
    std::vector<float> oma((size_t)d.ncols);
    for (int f = 0; f < sh->nf; ++f)
      for (int j = 0; j < BTPM_NS; ++j) oma[(size_t)f * BTPM_NS + j] = 1.0f - sh->h_albedo[f];
    DBuf tmp;
    upload(tmp, oma.data(), oma.size(), st);
    hipLaunchKernelGGL(btpm_scale_cols, dim3(grid_stride(nser, 256)), dim3(256), 0, st,
                       S->src.as<float>(), sh->cillum.as<float>(), tmp.as<float>(), s->flux_scale,
                       d.ncols, nser);
    HIPCHK(hipStreamSynchronize(st));
  }
  const SetHost* W = warm_crater_source(s);
  if (W) {
    // the previous solve's final crater profiles and surface temperatures
    HIPCHK(hipMemcpyAsync(S->T.p, W->T.p, nser * sizeof(float), hipMemcpyDeviceToDevice, st));
    HIPCHK(hipMemcpyAsync(S->prof.p, W->prof.p, (size_t)BTPM_PSTRIDE * d.ncols * sizeof(float),
                          hipMemcpyDeviceToDevice, st));
    HIPCHK(hipMemsetAsync(S->frozen.p, 0, S->frozen.n, st));
    HIPCHK(hipMemsetAsync(S->dmax.p, 0, S->dmax.n, st));
  } else if (d.inst) {
    // TI = 0 starts from zero temperatures (roughness_zero_temperatures,
    // T175706/thermal_GPU.cu:348-367, chosen at thermal.cpp:3567-3574)
    HIPCHK(hipMemsetAsync(S->T.p, 0, S->T.n, st));
    HIPCHK(hipMemsetAsync(S->prof.p, 0, S->prof.n, st));
    HIPCHK(hipMemsetAsync(S->frozen.p, 0, S->frozen.n, st));
    HIPCHK(hipMemsetAsync(S->dmax.p, 0, S->dmax.n, st));
  } else {
    hipLaunchKernelGGL(btpm_init_crater, dim3(d.ngroups), dim3(128), 0, st, d,
                       sh->cvf32.as<float>());
  }
  HIPCHK(hipGetLastError());
  s->crater = std::move(S);
  if (crater_takes_warm(s)) release_warm(s, st);
}

// smooth set no longer needed for stepping: keep T (results) and prof
void shrink_smooth(SetHost* S) {
  S->src.release();
  S->qext.release();
  S->emis.release();
}

// device bytes a solve allocates: its smooth set at admission, its crater
// set at the smooth -> crater transition
size_t smooth_bytes(const btpm_shape* sh) {
  const size_t nb = nblocks(sh->p.ntau);
  return (size_t)sh->nf * (nb * BTPM_BLK * (sh->nnz ? 4 : 2) + BTPM_PSTRIDE + 16) * 4 + (1 << 20);
}
size_t crater_bytes(const btpm_shape* sh) {
  const size_t nb = nblocks(sh->p.ntau);
  return (size_t)sh->nf * BTPM_NS * (nb * BTPM_BLK * 2 + BTPM_PSTRIDE + 4) * 4 + (1 << 20);
}

// ---------------------------------------------------------------------------
// scheduler
// ---------------------------------------------------------------------------
// Scheduler trace on stderr (admissions, crater-set builds, completions with
// the memory figures behind them) when the environment sets BTPM_TRACE.
bool trace_on() {
  static const bool on = std::getenv("BTPM_TRACE") != nullptr;
  return on;
}

struct Sched {
  btpm_ctx* ctx;
  btpm_batch_opts opt;
  std::vector<btpm_solve*> active;
  std::deque<btpm_solve*> pending;
  // per-slot owner (set host) for status readback
  std::map<int, std::pair<btpm_solve*, SetHost*>> slots;
  bool dirty = true;
  hipGraphExec_t gexec = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t phase_exec[4] = {nullptr, nullptr, nullptr, nullptr};
  hipGraph_t phase_graph[4] = {nullptr, nullptr, nullptr, nullptr};
  DBuf d_tiles_c, d_ids;
  // smooth tiles, one launch per workgroup width (the set on trial, if any,
  // in a launch of its own)
  struct SmoothList {
    int width;
    int n;
    bool trial;
    std::unique_ptr<DBuf> tiles;
  };
  std::vector<SmoothList> smooth_lists;
  // smooth sets with facet self-heating, grouped by shape: one tiled SpMM
  // launch per group per rotation
  struct HeatGroup {
    btpm_shape* shape;
    int nsets;
    std::unique_ptr<DBuf> slots;
    // kernel code: kHeatRows = btpm_selfheat_rows (time along the lanes);
    // 64 / 32 = steps per workgroup of the lane-per-receiver kernel (ELL, or
    // the per-chunk CSR btpm_selfheat_tiled when the shape has no ELL
    // layout); kHeatGrp + 32 = the receiver-grouped kernel
    int code = 0;
    bool trial = false;
  };
  static constexpr int kHeatRows = 0, kHeatGrp = 1000;
  static int heat_steps(int code) {
    return code == kHeatRows ? BTPM_SH_STEPS_T : code % kHeatGrp;
  }
  // self-heating kernels a shape's layouts allow, for the tuner
  static std::vector<int> heat_candidates(const btpm_shape* sh) {
    std::vector<int> c{kHeatRows, 64};
    if (sh->use_ell) c.push_back(32);
    if (sh->use_grp) c.push_back(kHeatGrp + 32);
    return c;
  }
  // untuned: the grouped kernel while its padding stays within
  // BTPM_SH_ELL_MAX_PAD times the entry count, else ELL under the same
  // bound, else the time-lane kernel (skewed rows: a few receivers see most
  // of the shape)
  static int heat_default(const btpm_shape* sh) {
    if (sh->use_grp && sh->grp_slots <= (size_t)BTPM_SH_ELL_MAX_PAD * sh->nnz)
      return kHeatGrp + 32;
    return sh->use_ell ? BTPM_SH_STEPS : kHeatRows;
  }
  std::vector<HeatGroup> heat;

  // ---- autotuner ----------------------------------------------------------
  // Tile choices per size class, measured on the device:
  //   kind 0: smooth conduction columns per workgroup {256, 128, 64}, per
  //           column-count class (ceil log2 of the set's columns);
  //   kind 1: facet self-heating kernel {time-lane (0), per-chunk CSR or ELL
  //           64, ELL 32, receiver-grouped} as the shape's layouts allow, per
  //           (log2 nnz, log2 facets, step blocks) class.
  // A candidate replaces the first one only if it is kTuneMargin cheaper.
  // Every candidate computes bit-identical results (only the launch shape
  // differs), so a trial is an ordinary rotation of the running solves:
  // while a class is untuned the scheduler syncs after every rotation, runs
  // that class's sets in a launch of their own between two events and
  // charges the time to the candidate per unfrozen column.  Candidates run in
  // palindromic order (A B C C B A), which cancels a linear drift of the
  // per-rotation cost as the solves converge.  The cheapest is kept in the
  // context for every later run, and the graphs are re-captured with it.
  static constexpr double kTuneMargin = 0.02;
  using TuneKey = std::pair<int, long long>;
  struct Trial {
    TuneKey key;
    std::vector<int> cand, seq;
    size_t next = 0;
    std::vector<double> cost, n;
    double units = 0.0;  // unfrozen columns of the rotation being timed
    bool timed = false;  // the events bracket a launch of this chunk
  };
  std::deque<Trial> trials;
  hipEvent_t tev[2] = {nullptr, nullptr};

  static long long log2_class(long long v) {
    long long c = 0;
    while ((1LL << c) < v) ++c;
    return c;
  }
  static TuneKey smooth_key(const SetHost* S) { return {0, log2_class(S->d.ncols)}; }
  static TuneKey heat_key(const btpm_shape* sh) {
    return {1, (log2_class(sh->nnz) << 40) | (log2_class(sh->nf) << 24) | sh->nb};
  }
  bool tuning() const { return opt.autotune && !trials.empty(); }
  bool on_trial(const TuneKey& k) const { return tuning() && trials.front().key == k; }
  int choice(const TuneKey& k, int dflt) const {
    if (on_trial(k)) {
      const Trial& t = trials.front();
      return t.cand[t.seq[t.next]];
    }
    auto it = ctx->tuned.find(k);
    return it == ctx->tuned.end() ? dflt : it->second;
  }
  void want_trial(const TuneKey& k, const std::vector<int>& cand) {
    if (ctx->tuned.count(k)) return;
    for (const Trial& t : trials)
      if (t.key == k) return;
    Trial t;
    t.key = k;
    t.cand = cand;
    for (size_t i = 0; i < cand.size(); ++i) t.seq.push_back((int)i);
    for (size_t i = cand.size(); i-- > 0;) t.seq.push_back((int)i);
    t.cost.assign(cand.size(), 0.0);
    t.n.assign(cand.size(), 0.0);
    trials.push_back(std::move(t));
  }
  // after the sync of a trial rotation: charge its time, decide when done
  void trial_sample() {
    if (trials.empty()) return;
    Trial& t = trials.front();
    if (t.timed && t.units > 0.0) {
      float ms = 0.0f;
      HIPCHK(hipEventElapsedTime(&ms, tev[0], tev[1]));
      const int c = t.seq[t.next];
      t.cost[c] += ms / t.units;
      t.n[c] += 1.0;
      t.next++;
      stats.tune_rotations++;
      if (trace_on())
        std::fprintf(stderr, "[btpm] tune (%d, %lld) candidate %d: %.3f ms for %.0f columns\n",
                     t.key.first, t.key.second, t.cand[c], ms, t.units);
    }
    t.timed = false;
    if (t.next == t.seq.size()) {
      size_t best = 0;
      for (size_t i = 1; i < t.cand.size(); ++i)
        if (t.cost[i] / t.n[i] < t.cost[best] / t.n[best]) best = i;
      // the first candidate (the untuned width of the smooth conduction) is
      // kept unless another is at least kTuneMargin cheaper: near-ties are
      // timing noise, and a decision flipped by noise carries over to every
      // later batch and server on the context
      if (best != 0 && t.cost[best] / t.n[best] > (1.0 - kTuneMargin) * (t.cost[0] / t.n[0]))
        best = 0;
      ctx->tuned[t.key] = t.cand[best];
      stats.tune_decisions++;
      if (trace_on())
        std::fprintf(stderr, "[btpm] tune (%d, %lld) -> %d\n", t.key.first, t.key.second,
                     t.cand[best]);
      trials.pop_front();
    }
    dirty = true;
  }
  int n_tiles_s = 0, n_tiles_c = 0, n_ids = 0;
  // Fast lane: sets at most 1/8 the size (columns) of the largest resident
  // set run opt.fast_extra more rotations after every rotation of all sets,
  // from lists of their own, so a small solve does not advance in lock-step
  // with the big sets beside it (sets are independent: results are the same).
  // Off in tuning trials; in profiled runs the fast rotations are launched
  // after each timed rotation, outside its events.
  std::vector<SmoothList> fsmooth;
  std::vector<HeatGroup> fheat;
  DBuf fd_tiles_c, fd_ids;
  int fn_tiles_c = 0, fn_ids = 0;
  bool fast_active = false;
  btpm_batch_stats stats{};
  // launch timing (opts.profile): events around each conduction launch
  std::vector<hipEvent_t> ev_s, ev_c, ev_h;  // pairs (start, end) per rotation
  // Admission works on the device's live free memory, so results that
  // completed solves keep until they are destroyed are accounted for:
  size_t headroom = 0;  // kept free: (1 - mem_fraction) of what was free at start
  size_t reserved = 0;  // crater bytes promised to admitted solves not yet allocated
  int craters_active = 0;
  std::chrono::steady_clock::time_point t0;
  std::function<void()> on_complete;  // server: wakes btpm_server_wait_any

  // device memory free for new sets: the driver's free memory plus the
  // library's cached blocks
  size_t free_bytes() const {
    size_t fr = 0, tot = 0;
    HIPCHK(hipMemGetInfo(&fr, &tot));
    return fr + pool_cached(ctx->device);
  }

  explicit Sched(btpm_ctx* c, const btpm_batch_opts& o) : ctx(c), opt(o) {
    const size_t fr = free_bytes();
    const double frac = opt.mem_fraction > 0 ? std::min((double)opt.mem_fraction, 1.0) : 0.8;
    headroom = (size_t)((double)fr * (1.0 - frac));
    t0 = std::chrono::steady_clock::now();
    HIPCHK(hipEventCreate(&tev[0]));
    HIPCHK(hipEventCreate(&tev[1]));
  }
  ~Sched() {
    drop_graph();
    for (hipEvent_t e : tev)
      if (e) (void)hipEventDestroy(e);
    for (hipEvent_t e : ev_s) (void)hipEventDestroy(e);
    for (hipEvent_t e : ev_c) (void)hipEventDestroy(e);
    for (hipEvent_t e : ev_h) (void)hipEventDestroy(e);
  }

  void make_events(int R) {
    auto grow = [](std::vector<hipEvent_t>& v, size_t n) {
      while (v.size() < n) {
        hipEvent_t e;
        HIPCHK(hipEventCreate(&e));
        v.push_back(e);
      }
    };
    grow(ev_s, 2 * (size_t)R);
    grow(ev_c, 2 * (size_t)R);
    grow(ev_h, 2 * (size_t)R);
  }

  void drop_graph() {
    if (gexec) (void)hipGraphExecDestroy(gexec);
    if (graph) (void)hipGraphDestroy(graph);
    gexec = nullptr;
    graph = nullptr;
    for (int ph = 0; ph < 4; ++ph) {
      if (phase_exec[ph]) (void)hipGraphExecDestroy(phase_exec[ph]);
      if (phase_graph[ph]) (void)hipGraphDestroy(phase_graph[ph]);
      phase_exec[ph] = nullptr;
      phase_graph[ph] = nullptr;
    }
  }

  int take_slot() {
    if (ctx->free_slots.empty()) throw HipError{fail(BTPM_ERR_STATE, "too many resident sets")};
    int s = ctx->free_slots.back();
    ctx->free_slots.pop_back();
    return s;
  }

  void add_set(btpm_solve* s, SetHost* S) {
    S->slot = take_slot();
    btpm_status_dev z{};
    z.norder = S->d.ngroups;
    ctx->h_status[S->slot] = z;
    HIPCHK(hipMemcpyAsync(ctx->d_sets + S->slot, &S->d, sizeof(btpm_set_dev),
                          hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipMemcpyAsync(ctx->d_status + S->slot, &z, sizeof z, hipMemcpyHostToDevice,
                          ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    slots[S->slot] = {s, S};
    dirty = true;
  }

  void remove_set(SetHost* S) {
    if (S->slot < 0) return;
    slots.erase(S->slot);
    ctx->free_slots.push_back(S->slot);
    S->slot = -1;
    dirty = true;
  }

  bool can_admit(btpm_solve* s) {
    if (s->with_craters && opt.max_active_craters > 0 && craters_active >= opt.max_active_craters)
      return false;
    if (active.empty()) return true;  // always let one in
    const size_t need = smooth_bytes(s->shape) + (s->with_craters ? crater_bytes(s->shape) : 0);
    const size_t fr = free_bytes();
    if (fr > headroom + reserved && need <= fr - headroom - reserved) return true;
    if (!s->mem_waited) {
      s->mem_waited = true;
      stats.mem_waits++;
    }
    return false;
  }

  // A warm-started solve waits for its predecessor: the smooth set for the
  // predecessor's smooth phase, the crater set for the whole predecessor.
  static bool smooth_ready(const btpm_solve* s) {
    if (!s->warm_pending || !s->warm) return true;
    const int ph = s->warm->phase;
    return ph == 2 || ph == 3 || ph == 4;
  }
  static bool crater_ready(const btpm_solve* s) {
    return !crater_takes_warm(s) || s->warm->phase == 3;
  }

  // the injected out-of-memory of the BTPM_INJECT_NOMEM test hook (only
  // while another solve runs, which guarantees a retry)
  int builds = 0;
  const int inject_k = inject_period();
  void maybe_inject(bool running) {
    const int k = inject_k;
    if (k > 0 && running && ++builds % k == 0)
      throw HipError{fail(BTPM_ERR_NOMEM, "injected out-of-memory (BTPM_INJECT_NOMEM=%d)", k)};
  }

  void admit() {
    // first come, first served among the solves that are ready; a ready
    // solve that does not fit stops admission until memory frees up
    for (auto it = pending.begin(); it != pending.end();) {
      btpm_solve* s = *it;
      if (!smooth_ready(s)) {
        ++it;
        continue;
      }
      if (!can_admit(s)) break;
      try {
        maybe_inject(!active.empty());
        build_smooth(s, ctx->stream);
      } catch (const HipError& e) {
        // the footprint estimate was short: wait for running solves to finish
        if (e.code != BTPM_ERR_NOMEM || active.empty()) throw;
        stats.deferrals++;
        break;
      }
      it = pending.erase(it);
      if (trace_on())
        std::fprintf(stderr, "[btpm] admit %p nf %d craters %d: free %.1f GB, reserved %.1f GB, "
                     "headroom %.1f GB, crater need %.1f GB, active %zu\n", (void*)s,
                     s->shape->nf, s->with_craters, free_bytes() / 1e9, reserved / 1e9,
                     headroom / 1e9, s->with_craters ? crater_bytes(s->shape) / 1e9 : 0.0,
                     active.size());
      if (s->with_craters) reserved += crater_bytes(s->shape);
      if (s->with_craters) craters_active++;
      s->phase = 1;
      add_set(s, s->smooth.get());
      active.push_back(s);
    }
    stats.max_resident = std::max(stats.max_resident, (int)active.size());
  }

  // Builds the crater set of a solve whose smooth phase is done (phase 4).
  // Out of memory while other solves still run is not an error: the solve
  // waits and is retried after the next chunk.
  bool try_build_crater(btpm_solve* s) {
    if (!crater_ready(s)) return false;
    bool running = false;
    for (btpm_solve* o : active) running = running || o->phase == 1 || o->phase == 2;
    try {
      maybe_inject(running);
      build_crater(s, ctx->stream);
    } catch (const HipError& e) {
      if (e.code == BTPM_ERR_NOMEM && !running)
        throw HipError{fail(BTPM_ERR_NOMEM, "%s; no other solve is running -- finished solves keep "
                            "their results on the device until they are destroyed",
                            last_error().c_str())};
      if (e.code != BTPM_ERR_NOMEM) throw;
      stats.deferrals++;
      if (trace_on())
        std::fprintf(stderr, "[btpm] crater set of %p deferred: free %.1f GB, reserved %.1f GB\n",
                     (void*)s, free_bytes() / 1e9, reserved / 1e9);
      return false;
    }
    if (trace_on())
      std::fprintf(stderr, "[btpm] crater set of %p built: free %.1f GB, reserved %.1f GB\n",
                   (void*)s, free_bytes() / 1e9, reserved / 1e9);
    reserved -= std::min(reserved, crater_bytes(s->shape));
    s->phase = 2;
    add_set(s, s->crater.get());
    return true;
  }

  int needed_tiles(const SetHost* S) const {
    const int n = ctx->h_status[S->slot].norder;
    return std::min(S->ntiles, (n + S->d.tile_cols - 1) / S->d.tile_cols);
  }

  void rebuild_lists() {
    // tuning: classes present without a decision get a trial; a trial whose
    // sets have all left is dropped (it restarts if the class returns)
    std::map<btpm_shape*, std::vector<int>> by_shape;
    for (auto& kv : slots) {
      SetHost* S = kv.second.second;
      if (S->d.kind == BTPM_SET_SMOOTH && S->d.qext)
        by_shape[kv.second.first->shape].push_back(S->slot);
    }
    if (opt.autotune) {
      std::set<TuneKey> present;
      for (auto& kv : slots)
        if (kv.second.second->d.kind == BTPM_SET_SMOOTH) {
          present.insert(smooth_key(kv.second.second));
          want_trial(smooth_key(kv.second.second), {256, 128, 64});
        }
      for (auto& kv : by_shape)
        if (kv.first->force_heat < 0) {
          present.insert(heat_key(kv.first));
          want_trial(heat_key(kv.first), heat_candidates(kv.first));
        }
      while (!trials.empty() && !present.count(trials.front().key)) trials.pop_front();
      if (!trials.empty()) trials.front().units = 0.0;
    }
    std::map<int, std::vector<btpm_tile>> ts;
    std::vector<btpm_tile> ts_trial, tc;
    std::vector<int> ids;
    for (auto& kv : slots) {
      SetHost* S = kv.second.second;
      ids.push_back(S->slot);
      const int unfrozen = ctx->h_status[S->slot].norder;
      if (S->d.kind == BTPM_SET_SMOOTH) {
        const TuneKey k = smooth_key(S);
        const int w = choice(k, BTPM_SMOOTH_WG);
        if (S->d.tile_cols != w) {  // the status kernel reads the width too
          S->d.tile_cols = w;
          HIPCHK(hipMemcpyAsync(ctx->d_sets + S->slot, &S->d, sizeof(btpm_set_dev),
                                hipMemcpyHostToDevice, ctx->stream));
        }
        S->listed = needed_tiles(S);
        const bool tr = on_trial(k);
        if (tr) trials.front().units += unfrozen;
        for (int t = 0; t < S->listed; ++t) (tr ? ts_trial : ts[w]).push_back(btpm_tile{S->slot, t});
      } else {
        // only the tiles the compacted order[] can fill (it only shrinks)
        S->listed = needed_tiles(S);
        for (int t = 0; t < S->listed; ++t) tc.push_back(btpm_tile{S->slot, t});
      }
    }
    smooth_lists.clear();
    n_tiles_s = 0;
    auto add_list = [&](int w, const std::vector<btpm_tile>& v, bool trial) {
      if (v.empty()) return;
      SmoothList l{w, (int)v.size(), trial, std::make_unique<DBuf>()};
      upload(*l.tiles, v.data(), v.size(), ctx->stream);
      n_tiles_s += l.n;
      smooth_lists.push_back(std::move(l));
    };
    if (!ts_trial.empty()) add_list(choice(trials.front().key, BTPM_SMOOTH_WG), ts_trial, true);
    for (auto& kv : ts) add_list(kv.first, kv.second, false);
    n_tiles_c = (int)tc.size();
    n_ids = (int)ids.size();
    heat.clear();
    for (auto& kv : by_shape) {
      HeatGroup g{kv.first, (int)kv.second.size(), std::make_unique<DBuf>()};
      upload(*g.slots, kv.second.data(), kv.second.size(), ctx->stream);
      if (kv.first->force_heat >= 0) {
        g.code = kv.first->force_heat;
      } else {
        const TuneKey k = heat_key(kv.first);
        g.code = choice(k, heat_default(kv.first));
        // a choice tuned on another shape of the class that this shape's
        // layouts cannot run
        const std::vector<int> ok = heat_candidates(kv.first);
        if (std::find(ok.begin(), ok.end(), g.code) == ok.end()) g.code = heat_default(kv.first);
        g.trial = on_trial(k);
        if (g.trial)
          for (int sl : kv.second) trials.front().units += ctx->h_status[sl].norder;
      }
      heat.push_back(std::move(g));
    }
    upload(d_tiles_c, tc.data(), tc.size(), ctx->stream);
    upload(d_ids, ids.data(), ids.size(), ctx->stream);
    rebuild_fast();
  }

  void rebuild_fast() {
    fsmooth.clear();
    fheat.clear();
    fn_tiles_c = fn_ids = 0;
    fast_active = false;
    if (opt.fast_extra <= 0 || tuning()) return;
    int max_cols = 0;
    for (auto& kv : slots) max_cols = std::max(max_cols, kv.second.second->d.ncols);
    auto fast = [&](const SetHost* S) { return 8LL * S->d.ncols <= max_cols; };
    std::map<int, std::vector<btpm_tile>> ts;
    std::vector<btpm_tile> tc;
    std::vector<int> ids;
    std::map<btpm_shape*, std::vector<int>> by_shape;
    for (auto& kv : slots) {
      const SetHost* S = kv.second.second;
      if (!fast(S)) continue;
      ids.push_back(S->slot);
      auto& v = S->d.kind == BTPM_SET_SMOOTH ? ts[S->d.tile_cols] : tc;
      for (int t = 0; t < S->listed; ++t) v.push_back(btpm_tile{S->slot, t});
      if (S->d.kind == BTPM_SET_SMOOTH && S->d.qext)
        by_shape[kv.second.first->shape].push_back(S->slot);
    }
    if (ids.empty()) return;
    for (auto& kv : ts) {
      SmoothList l{kv.first, (int)kv.second.size(), false, std::make_unique<DBuf>()};
      upload(*l.tiles, kv.second.data(), kv.second.size(), ctx->stream);
      fsmooth.push_back(std::move(l));
    }
    for (auto& kv : by_shape) {
      HeatGroup g{kv.first, (int)kv.second.size(), std::make_unique<DBuf>()};
      upload(*g.slots, kv.second.data(), kv.second.size(), ctx->stream);
      for (const HeatGroup& m : heat)
        if (m.shape == kv.first) g.code = m.code;
      fheat.push_back(std::move(g));
    }
    fn_tiles_c = (int)tc.size();
    fn_ids = (int)ids.size();
    upload(fd_tiles_c, tc.data(), tc.size(), ctx->stream);
    upload(fd_ids, ids.data(), ids.size(), ctx->stream);
    fast_active = true;
  }

  // the events around the launch on trial (only in tuning chunks, which
  // are never captured)
  void trial_event(hipStream_t st, int i) {
    HIPCHK(hipEventRecord(tev[i], st));
    if (i == 1) trials.front().timed = true;
  }

  // one rotation of every resident set, in four phases: facet self-heating
  // SpMMs, smooth conduction, crater conduction, status
  void enqueue_phase(hipStream_t st, int ph, bool fast = false) {
    if (ph == 0) {
      // qext = src + VF * emis: the total absorbed flux of this rotation, for
      // every smooth set of a shape in one launch
      for (const HeatGroup& g : fast ? fheat : heat) {
        btpm_shape* sh = g.shape;
        const dim3 grid(grid1(sh->nf, BTPM_SH_ROWS),
                        grid1((size_t)sh->nb * BTPM_BLK, heat_steps(g.code)),
                        g.nsets);  // the lane-per-receiver and grouped kernels
        if (g.trial) trial_event(st, 0);
        if (g.code == kHeatRows)
          hipLaunchKernelGGL(btpm_selfheat_rows,
                             dim3(grid1(sh->nf, BTPM_SH_ROWS_T),
                                  grid1((size_t)sh->nb * BTPM_BLK, BTPM_SH_STEPS_T), g.nsets),
                             dim3(BTPM_SH_ROWS_T / 32 * 64), 0, st, sh->cptr.as<int>(),
                             sh->sent.as<int2>(), sh->nf, sh->nchunk, g.slots->as<int>(),
                             ctx->d_sets, ctx->d_status);
        else if (g.code == kHeatGrp + 32)
          hipLaunchKernelGGL(btpm_selfheat_grp, grid, dim3(BTPM_SH_ROWS), 0, st,
                             sh->gseg.as<int2>(), sh->gent.as<int4>(), sh->gperm.as<int>(),
                             sh->nf, sh->nchunk, g.slots->as<int>(), ctx->d_sets, ctx->d_status);
        else if (sh->use_ell && g.code == 32)
          hipLaunchKernelGGL(btpm_selfheat_ell_h32, grid, dim3(BTPM_SH_ROWS), 0, st,
                             sh->wseg.as<int2>(), sh->ell.as<int2>(), sh->nf, sh->nchunk,
                             g.slots->as<int>(), ctx->d_sets, ctx->d_status);
        else if (sh->use_ell)
          hipLaunchKernelGGL(btpm_selfheat_ell, grid, dim3(BTPM_SH_ROWS), 0, st,
                             sh->wseg.as<int2>(), sh->ell.as<int2>(), sh->nf, sh->nchunk,
                             g.slots->as<int>(), ctx->d_sets, ctx->d_status);
        else
          hipLaunchKernelGGL(btpm_selfheat_tiled, grid, dim3(BTPM_SH_ROWS), 0, st,
                             sh->cptr.as<int>(), sh->sent.as<int2>(), sh->nf, sh->nchunk,
                             g.slots->as<int>(), ctx->d_sets, ctx->d_status);
        if (g.trial) trial_event(st, 1);
      }
    } else if (ph == 1) {
      for (const SmoothList& l : fast ? fsmooth : smooth_lists) {
        const btpm_tile* t = l.tiles->as<btpm_tile>();
        if (l.trial) trial_event(st, 0);
        if (l.width == 64)
          hipLaunchKernelGGL(btpm_conduct_smooth<64>, dim3(l.n), dim3(64), 0, st, t,
                             ctx->d_sets, ctx->d_status);
        else if (l.width == 128)
          hipLaunchKernelGGL(btpm_conduct_smooth<128>, dim3(l.n), dim3(128), 0, st, t,
                             ctx->d_sets, ctx->d_status);
        else
          hipLaunchKernelGGL(btpm_conduct_smooth<256>, dim3(l.n), dim3(256), 0, st, t,
                             ctx->d_sets, ctx->d_status);
        if (l.trial) trial_event(st, 1);
      }
    } else if (ph == 2) {
      const int n = fast ? fn_tiles_c : n_tiles_c;
      if (n)
        hipLaunchKernelGGL(btpm_conduct_crater, dim3(n), dim3(BTPM_CRATER_THREADS), 0, st,
                           (fast ? fd_tiles_c : d_tiles_c).as<btpm_tile>(), ctx->d_sets,
                           ctx->d_status);
    } else {
      const int n = fast ? fn_ids : n_ids;
      if (n)
        hipLaunchKernelGGL(btpm_status_kernel, dim3(n), dim3(256), 0, st,
                           (fast ? fd_ids : d_ids).as<int>(), ctx->d_sets, ctx->d_status);
    }
  }

  void enqueue_rotation(hipStream_t st) {
    for (int ph = 0; ph < 4; ++ph) enqueue_phase(st, ph);
    if (fast_active)
      for (int i = 0; i < opt.fast_extra; ++i)
        for (int ph = 0; ph < 4; ++ph) enqueue_phase(st, ph, true);
  }

  bool phase_empty(int ph) const {
    if (ph == 0) {
      for (auto& kv : slots)
        if (kv.second.second->d.kind == BTPM_SET_SMOOTH && kv.second.second->d.qext) return false;
      return true;
    }
    if (ph == 1) return n_tiles_s == 0;
    if (ph == 2) return n_tiles_c == 0;
    return n_ids == 0;
  }

  // profiled rotation: each conduction phase bracketed by stream events
  // (around a launch of the phase's own graph when graphs are in use and the
  // chunk is not a tuning trial)
  void profiled_rotation(int i, bool direct) {
    hipStream_t st = ctx->stream;
    for (int ph = 0; ph < 4; ++ph) {
      hipEvent_t* ev = ph == 0 ? &ev_h[2 * i] : ph == 1 ? &ev_s[2 * i] : &ev_c[2 * i];
      if (ph == 3) ev = nullptr;
      const bool empty = phase_empty(ph);
      if (empty) continue;
      if (ev) HIPCHK(hipEventRecord(ev[0], st));
      if (opt.use_graph && !direct) {
        HIPCHK(hipGraphLaunch(phase_exec[ph], st));
        stats.graph_launches++;
      } else {
        enqueue_phase(st, ph);
      }
      if (ev) HIPCHK(hipEventRecord(ev[1], st));
    }
    if (fast_active)  // the small sets' extra rotations, not timed
      for (int k = 0; k < opt.fast_extra; ++k)
        for (int ph = 0; ph < 4; ++ph) enqueue_phase(st, ph, true);
  }

  void capture(hipGraph_t* g, hipGraphExec_t* x, int R, int phase) {
    HIPCHK(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    if (phase < 0)
      for (int i = 0; i < R; ++i) enqueue_rotation(ctx->stream);
    else
      enqueue_phase(ctx->stream, phase);
    HIPCHK(hipStreamEndCapture(ctx->stream, g));
    HIPCHK(hipGraphInstantiate(x, *g, nullptr, nullptr, 0));
  }

  // Rotations per host sync: opt.rotations_per_sync, or with 0 the
  // scheduler's own choice: the fewest rotations that keep a chunk at least
  // kChunkMs long (a sync costs tens of microseconds; fewer rotations per
  // chunk retire finished solves and admit new ones sooner), from the
  // measured wall time per rotation; it moves only by 2x steps (each move
  // re-captures the graph).
  static constexpr double kChunkMs = 2.0;
  int auto_R = 1;
  double rot_ms = -1.0;  // smoothed wall time per rotation
  int rotations_per_chunk() const {
    return opt.rotations_per_sync > 0 ? opt.rotations_per_sync : auto_R;
  }
  void schedule_sample(double chunk_ms, int R) {
    if (opt.rotations_per_sync > 0 || R <= 0) return;
    const double t = chunk_ms / R;
    rot_ms = rot_ms < 0.0 ? t : 0.7 * rot_ms + 0.3 * t;
    const int want = std::max(1, std::min(16, (int)std::ceil(kChunkMs / std::max(rot_ms, 1e-3))));
    if (want >= 2 * auto_R || 2 * want <= auto_R) {
      if (trace_on())
        std::fprintf(stderr, "[btpm] rotations per sync %d -> %d (%.3f ms per rotation)\n",
                     auto_R, want, rot_ms);
      auto_R = want;
      dirty = true;
    }
  }

  void run_chunk() {
    // a tuning trial is one rotation, enqueued directly and synced
    const bool tune = tuning();
    const int R = tune ? 1 : std::max(1, rotations_per_chunk());
    if (opt.profile) make_events(R);
    if (dirty || tune) {
      drop_graph();
      rebuild_lists();
      if (opt.use_graph && !tune) {
        if (opt.profile) {
          for (int ph = 0; ph < 4; ++ph)
            if (!phase_empty(ph)) capture(&phase_graph[ph], &phase_exec[ph], 1, ph);
        }
        else
          capture(&graph, &gexec, R, -1);
      }
      dirty = false;
    }
    const auto c0 = std::chrono::steady_clock::now();  // the rotations alone
    if (opt.profile) {
      for (int i = 0; i < R; ++i) profiled_rotation(i, tune);
    } else if (tune) {
      enqueue_rotation(ctx->stream);
    } else if (opt.use_graph) {
      HIPCHK(hipGraphLaunch(gexec, ctx->stream));
      stats.graph_launches++;
    } else {
      for (int i = 0; i < R; ++i) enqueue_rotation(ctx->stream);
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(ctx->h_status, ctx->d_status, sizeof(btpm_status_dev) * kMaxSets,
                          hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    stats.syncs++;
    if (fast_active) stats.fast_rotations += R * opt.fast_extra;
    schedule_sample(
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - c0).count(),
        R);
    if (tune) trial_sample();
    if (opt.profile) {
      for (int i = 0; i < R; ++i) {
        float ms = 0.0f;
        if (n_tiles_s) {
          HIPCHK(hipEventElapsedTime(&ms, ev_s[2 * i], ev_s[2 * i + 1]));
          stats.smooth_kernel_s += ms * 1e-3;
          stats.smooth_launches++;
        }
        if (!phase_empty(0)) {
          HIPCHK(hipEventElapsedTime(&ms, ev_h[2 * i], ev_h[2 * i + 1]));
          stats.heat_kernel_s += ms * 1e-3;
          stats.heat_launches += (int)heat.size();
          for (const HeatGroup& g : heat) stats.heat_set_rotations += g.nsets;
        }
        if (n_tiles_c) {
          HIPCHK(hipEventElapsedTime(&ms, ev_c[2 * i], ev_c[2 * i + 1]));
          stats.crater_kernel_s += ms * 1e-3;
          stats.crater_launches++;
        }
      }
    }
  }

  void finish_set(btpm_solve* s, SetHost* S) {
    const btpm_status_dev& st = ctx->h_status[S->slot];
    if (S->d.kind == BTPM_SET_SMOOTH) {
      s->info.rotations = st.rev;
      s->info.converged = st.converged;
      s->info.column_steps = st.colsteps;
    } else {
      s->info.crater_rotations = st.rev;
      s->info.crater_converged = st.converged;
      s->info.crater_column_steps = st.colsteps;
    }
    stats.column_steps += st.colsteps;
    if (S->d.kind == BTPM_SET_SMOOTH) stats.smooth_column_steps += st.colsteps;
    else stats.crater_column_steps += st.colsteps;
    stats.column_steps_all += (long long)st.rev * S->d.ncols * S->d.ntau;
    remove_set(S);
  }

  void complete(btpm_solve* s) {
    s->info.latency_s =
        std::chrono::duration<double>(std::chrono::steady_clock::now() - s->t_submit).count();
    if (s->with_craters) craters_active--;
    stats.solves++;
    if (s->info.converged && (!s->with_craters || s->info.crater_converged)) stats.converged++;
    if (trace_on())
      std::fprintf(stderr, "[btpm] complete %p after %.2f s: free %.1f GB\n", (void*)s,
                   s->info.latency_s, free_bytes() / 1e9);
    {
      std::lock_guard<std::mutex> lk(s->mu);
      s->phase = 3;
    }
    s->cv.notify_all();
    if (on_complete) on_complete();
  }

  // inspect statuses, advance phases, retire finished solves
  void advance() {
    std::vector<btpm_solve*> keep;
    for (btpm_solve* s : active) {
      bool alive = true;
      if (s->phase == 1) {
        SetHost* S = s->smooth.get();
        if (ctx->h_status[S->slot].done) {
          finish_set(s, S);
          shrink_smooth(S);
          if (s->with_craters) {
            s->phase = 4;
            try_build_crater(s);
          } else {
            complete(s);
            alive = false;
          }
        }
      } else if (s->phase == 4) {
        try_build_crater(s);
      } else if (s->phase == 2) {
        SetHost* S = s->crater.get();
        if (ctx->h_status[S->slot].done) {
          finish_set(s, S);
          S->src.release();
          complete(s);
          alive = false;
        }
      }
      if (alive) keep.push_back(s);
    }
    active.swap(keep);
    // re-tile (and re-capture) once compaction has emptied a quarter of a
    // set's listed tiles
    for (auto& kv : slots) {
      const SetHost* S = kv.second.second;
      if (4 * needed_tiles(S) < 3 * S->listed) dirty = true;
    }
  }

  void step() {
    admit();
    if (active.empty()) return;
    run_chunk();
    advance();
  }
};

}  // namespace

// ===========================================================================
// C ABI
// ===========================================================================
namespace btpm {
namespace {
constexpr size_t kStageSlice = 256u << 20;

void stage_buffers(btpm_ctx* ctx) {
  for (int k = 0; k < 2; ++k)
    if (!ctx->stage[k]) {
      HIPCHK(hipHostMalloc((void**)&ctx->stage[k], kStageSlice, hipHostMallocDefault));
      HIPCHK(hipEventCreateWithFlags(&ctx->stage_ev[k], hipEventDisableTiming));
    }
}

// memcpy split over up to 8 host threads (pageable <-> pinned slices)
void par_memcpy(char* d, const char* s, size_t n) {
  const int nt = (int)std::min(8u, std::max(1u, std::thread::hardware_concurrency()));
  const size_t part = (n + nt - 1) / nt;
  std::vector<std::thread> pool;
  for (int t = 1; t < nt; ++t) {
    const size_t a = std::min(n, (size_t)t * part), b = std::min(n, a + part);
    if (a >= b) continue;
    try {
      pool.emplace_back([=] { std::memcpy(d + a, s + a, b - a); });
    } catch (const std::system_error&) {
      std::memcpy(d + a, s + a, b - a);  // no thread to spare: copy here
    }
  }
  std::memcpy(d, s, std::min(n, part));
  for (auto& th : pool) th.join();
}
}  // namespace

void upload_staged(btpm_ctx* ctx, void* dst, const void* src, size_t bytes) {
  hipStream_t st = ctx->stream;
  if (bytes < 2 * kStageSlice) {
    if (bytes) HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, st));
    HIPCHK(hipStreamSynchronize(st));
    return;
  }
  stage_buffers(ctx);
  const char* s = static_cast<const char*>(src);
  char* d = static_cast<char*>(dst);
  size_t i = 0;
  for (size_t off = 0; off < bytes; off += kStageSlice, ++i) {
    const size_t n = std::min(kStageSlice, bytes - off);
    const int k = (int)(i & 1);
    // this buffer's previous slice has left for the device
    if (i >= 2) HIPCHK(hipEventSynchronize(ctx->stage_ev[k]));
    par_memcpy(ctx->stage[k], s + off, n);
    HIPCHK(hipMemcpyAsync(d + off, ctx->stage[k], n, hipMemcpyHostToDevice, st));
    HIPCHK(hipEventRecord(ctx->stage_ev[k], st));
  }
  HIPCHK(hipStreamSynchronize(st));
}

void download_staged(btpm_ctx* ctx, void* dst, const void* src, size_t bytes) {
  hipStream_t st = ctx->stream;
  if (bytes < 2 * kStageSlice) {
    if (bytes) HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    return;
  }
  stage_buffers(ctx);
  const char* s = static_cast<const char*>(src);
  char* d = static_cast<char*>(dst);
  const size_t ns = (bytes + kStageSlice - 1) / kStageSlice;
  auto issue = [&](size_t j) {
    const size_t off = j * kStageSlice, n = std::min(kStageSlice, bytes - off);
    HIPCHK(hipMemcpyAsync(ctx->stage[j & 1], s + off, n, hipMemcpyDeviceToHost, st));
    HIPCHK(hipEventRecord(ctx->stage_ev[j & 1], st));
  };
  issue(0);
  for (size_t j = 0; j < ns; ++j) {
    // the other buffer's slice was copied out in the previous pass
    if (j + 1 < ns) issue(j + 1);
    HIPCHK(hipEventSynchronize(ctx->stage_ev[j & 1]));
    const size_t off = j * kStageSlice;
    par_memcpy(d + off, ctx->stage[j & 1], std::min(kStageSlice, bytes - off));
  }
}
}  // namespace btpm

extern "C" {

int btpm_version(void) { return 100; }

const char* btpm_last_error(void) { return last_error().c_str(); }

void btpm_params_default(btpm_params* p, int ntau, float rot_period, float tacc) {
  p->ntau = ntau;
  p->nx = BTPM_NX;
  p->rot_period = rot_period;
  p->eps = 0.95f;
  p->sigma = 5.67e-8f;
  p->ir_abs = 1.0f;
  p->tacc = tacc;
  p->newton_tol = tacc / 10.0f;
  p->max_rev = 301;
  p->check_after = 5;
  p->newton_cap = 200;
}

void btpm_batch_opts_default(btpm_batch_opts* o) {
  o->rotations_per_sync = 4;
  o->use_graph = 1;
  o->max_active_craters = 0;
  o->mem_fraction = 0.8;
  o->profile = 0;
  o->autotune = 1;
  o->fast_extra = 3;
}

int btpm_ctx_create(int device, btpm_ctx** out) {
  if (!out) return fail(BTPM_ERR_ARG, "null out");
  API_BEGIN
  int n = 0;
  HIPCHK(hipGetDeviceCount(&n));
  if (device < 0 || device >= n) return fail(BTPM_ERR_ARG, "device %d of %d", device, n);
  HIPCHK(hipSetDevice(device));
  noalloc_setup(device);  // this device's scratch of the allocation-free entry point
  auto c = std::make_unique<btpm_ctx>();
  c->device = device;
  HIPCHK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  HIPCHK(hipMalloc(&c->d_sets, sizeof(btpm_set_dev) * kMaxSets));
  HIPCHK(hipMalloc(&c->d_status, sizeof(btpm_status_dev) * kMaxSets));
  HIPCHK(hipMemset(c->d_status, 0, sizeof(btpm_status_dev) * kMaxSets));
  HIPCHK(hipHostMalloc(&c->h_status, sizeof(btpm_status_dev) * kMaxSets, hipHostMallocDefault));
  for (int i = kMaxSets - 1; i >= 0; --i) c->free_slots.push_back(i);
  *out = c.release();
  API_END
}


int btpm_ctx_destroy(btpm_ctx* c) {
  if (!c) return BTPM_OK;
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->stream);
  pool_trim(c->device);  // cached blocks are idle: hand them back to the driver
  (void)hipFree(c->d_sets);
  (void)hipFree(c->d_status);
  (void)hipHostFree(c->h_status);
  for (int k = 0; k < 2; ++k) {
    if (c->stage[k]) (void)hipHostFree(c->stage[k]);
    if (c->stage_ev[k]) (void)hipEventDestroy(c->stage_ev[k]);
  }
  (void)hipStreamDestroy(c->stream);
  delete c;
  return BTPM_OK;
}

int btpm_ctx_mem_info(btpm_ctx* c, size_t* fr, size_t* tot) {
  if (!c) return fail(BTPM_ERR_ARG, "null ctx");
  API_BEGIN
  HIPCHK(hipSetDevice(c->device));
  HIPCHK(hipMemGetInfo(fr, tot));
  *fr += pool_cached(c->device);
  API_END
}

int btpm_shape_create(btpm_ctx* ctx, const btpm_params* p, const btpm_body* body,
                      const btpm_crater* crater, btpm_shape** out) {
  if (!ctx || !p || !body || !out) return fail(BTPM_ERR_ARG, "null argument");
  if (p->nx != BTPM_NX) return fail(BTPM_ERR_ARG, "nx must be %d (got %d)", BTPM_NX, p->nx);
  if (p->ntau < 1 || body->nf < 1 || !body->illum || !body->albedo)
    return fail(BTPM_ERR_ARG, "empty body or ntau");
  if ((body->nnz > 0) != (body->rowptr != nullptr))
    return fail(BTPM_ERR_ARG, "nnz and rowptr disagree");
  if (crater && !crater->vf) return fail(BTPM_ERR_ARG, "crater without view factors");
  if (crater && !crater->illum && (!crater->shadow || !crater->cosang) &&
      (!crater->sun_local || !crater->vis_table || !crater->normals || crater->n_el < 2 ||
       crater->n_az < 1))
    return fail(BTPM_ERR_ARG, "crater needs illum, shadow+cosang, or sun_local+vis_table+normals");
  if (p->max_rev < 1 || p->newton_cap < 1 || !(p->tacc > 0) || !(p->newton_tol > 0))
    return fail(BTPM_ERR_ARG, "bad convergence parameters");
  CtxLock lk(ctx);
  API_BEGIN
  HIPCHK(hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  auto sh = std::make_unique<btpm_shape>();
  sh->ctx = ctx;
  sh->p = *p;
  const int nf = sh->nf = body->nf;
  const int nt = p->ntau;
  const int nb = sh->nb = nblocks(nt);
  sh->nnz = body->nnz;
  sh->h_albedo.assign(body->albedo, body->albedo + nf);
  sh->h_numrevs.assign(nf, 1.0f);
  if (body->numrevs) sh->h_numrevs.assign(body->numrevs, body->numrevs + nf);
  for (int f = 0; f < nf; ++f) {  // explicit-scheme stability, as the kernels assume
    const double dz = 8.0 / BTPM_NX;
    const float r = (float)((2.0 * kPi * sh->h_numrevs[f] / nt) / (dz * dz));
    if (!(r >= 0.0f && r <= (float)BTPM_R_MAX))
      return fail(BTPM_ERR_ARG, "facet %d: dtau/dz^2 = %.6g outside [0, %g] (FTCS stability "
                  "limit; raise ntau)", f, (double)r, BTPM_R_MAX);
  }
  std::vector<float> oma(nf);
  for (int f = 0; f < nf; ++f) oma[f] = 1.0f - sh->h_albedo[f];
  upload(sh->albedo, sh->h_albedo.data(), nf, st);
  upload(sh->one_minus_a, oma.data(), nf, st);
  upload(sh->numrevs, sh->h_numrevs.data(), nf, st);
  if (body->nnz > 0) {
    for (int f = 0; f < nf; ++f)
      if (body->rowptr[f] > body->rowptr[f + 1]) return fail(BTPM_ERR_ARG, "rowptr not monotone");
    if (body->rowptr[0] != 0 || body->rowptr[nf] != body->nnz)
      return fail(BTPM_ERR_ARG, "rowptr does not span nnz");
    for (int k = 0; k < body->nnz; ++k)
      if (body->colidx[k] < 0 || body->colidx[k] >= nf)
        return fail(BTPM_ERR_ARG, "colidx[%d]=%d out of range", k, body->colidx[k]);
    upload(sh->rowptr, body->rowptr, nf + 1, st);
    upload(sh->colidx, body->colidx, body->nnz, st);
    upload(sh->vf, body->vf, body->nnz, st);
    {
      constexpr int kChunk = BTPM_SH_CHUNK;  // emitter chunk of the self-heating kernels
      const int nc = (nf + kChunk - 1) / kChunk;
      std::vector<int> se(2 * (size_t)body->nnz), cp((size_t)(nc + 1) * nf);
      std::vector<std::pair<int, float>> row;
      for (int f = 0; f < nf; ++f) {
        const int k0 = body->rowptr[f], k1 = body->rowptr[f + 1];
        row.clear();
        for (int k = k0; k < k1; ++k) row.emplace_back(body->colidx[k], body->vf[k]);
        std::stable_sort(row.begin(), row.end(),
                         [](const std::pair<int, float>& a, const std::pair<int, float>& b) {
                           return a.first < b.first;
                         });
        int k = k0;
        for (int c = 0; c <= nc; ++c) {
          while (k < k1 && row[k - k0].first < c * kChunk) ++k;
          cp[(size_t)c * nf + f] = c == nc ? k1 : k;
        }
        for (int i = 0; i < k1 - k0; ++i) {
          se[2 * (size_t)(k0 + i)] = row[i].first;
          std::memcpy(&se[2 * (size_t)(k0 + i) + 1], &row[i].second, 4);
        }
      }
      sh->nchunk = nc;
      // lane-interleaved copy: one segment per (receiver block, chunk, wave)
      const int nrb = (nf + BTPM_SH_ROWS - 1) / BTPM_SH_ROWS;
      std::vector<int> ws(2 * (size_t)nrb * nc * BTPM_SH_WAVES);
      size_t total = 0;
      for (int rb = 0; rb < nrb; ++rb)
        for (int c = 0; c < nc; ++c)
          for (int w = 0; w < BTPM_SH_WAVES; ++w) {
            int n = 0;
            for (int l = 0; l < 64; ++l) {
              const int f = rb * BTPM_SH_ROWS + w * 64 + l;
              if (f < nf)
                n = std::max(n, cp[(size_t)(c + 1) * nf + f] - cp[(size_t)c * nf + f]);
            }
            const size_t i = ((size_t)rb * nc + c) * BTPM_SH_WAVES + w;
            ws[2 * i] = (int)total;
            ws[2 * i + 1] = n;
            total += (size_t)n * 64;
          }
      // receiver-grouped copy: per block, the receivers in slots by
      // decreasing entry count; per (block, chunk, wave, pass of 8 slots) the
      // longest entry count rounded up to even, entries in int4 pairs
      constexpr int kPass = 8;
      std::vector<int> gp((size_t)nrb * BTPM_SH_ROWS, nf);
      for (int rb = 0; rb < nrb; ++rb) {
        const int f0 = rb * BTPM_SH_ROWS, f1 = std::min(nf, f0 + BTPM_SH_ROWS);
        int* b = gp.data() + f0;
        for (int f = f0; f < f1; ++f) b[f - f0] = f;
        std::stable_sort(b, b + (f1 - f0), [&](int x, int y) {
          return body->rowptr[x + 1] - body->rowptr[x] > body->rowptr[y + 1] - body->rowptr[y];
        });
      }
      std::vector<int> gs(2 * (size_t)nrb * nc * BTPM_SH_WAVES * kPass);
      size_t gpairs = 0;
      for (int rb = 0; rb < nrb; ++rb)
        for (int c = 0; c < nc; ++c)
          for (int w = 0; w < BTPM_SH_WAVES; ++w)
            for (int p = 0; p < kPass; ++p) {
              int n = 0;
              for (int g = 0; g < 8; ++g) {
                const int f = gp[(size_t)rb * BTPM_SH_ROWS + 64 * w + 8 * p + g];
                if (f < nf)
                  n = std::max(n, cp[(size_t)(c + 1) * nf + f] - cp[(size_t)c * nf + f]);
              }
              n = (n + 1) & ~1;
              const size_t i = (((size_t)rb * nc + c) * BTPM_SH_WAVES + w) * kPass + p;
              gs[2 * i] = (int)std::min(gpairs, (size_t)INT32_MAX);
              gs[2 * i + 1] = n;
              gpairs += (size_t)(n / 2) * 8;
            }
      // BTPM_SH_LAYOUT forces one kernel (diagnostics and tests): csr (the
      // per-chunk CSR, lane per receiver), ell / ell32 (ELL, 64 / 32 steps per
      // workgroup), grp (grouped, 32 steps), rows (time-lane)
      const char* lay = std::getenv("BTPM_SH_LAYOUT");
      const bool fits = total < (size_t)INT32_MAX;
      const bool gfits = gpairs < (size_t)INT32_MAX;
      sh->grp_slots = 2 * gpairs;
      // the grouped layout is kept while it stays within 16 entries per view
      // factor (the tuner times it; the untuned choice needs the ELL bound)
      sh->use_grp = gfits && sh->grp_slots <= 16 * (size_t)body->nnz;
      if (lay && std::strcmp(lay, "csr") == 0) {
        sh->use_ell = false;
        sh->force_heat = 64;
      } else if (lay && (std::strcmp(lay, "ell") == 0 || std::strcmp(lay, "ell32") == 0)) {
        sh->use_ell = fits;
        sh->force_heat = lay[3] ? 32 : 64;
        if (!fits) sh->force_heat = 64;  // the per-chunk CSR
      } else {
        sh->use_ell = fits && total <= (size_t)BTPM_SH_ELL_MAX_PAD * body->nnz;
        if (lay && std::strcmp(lay, "rows") == 0) sh->force_heat = 0;
        if (lay && std::strcmp(lay, "grp") == 0) {
          if (!gfits) return fail(BTPM_ERR_ARG, "BTPM_SH_LAYOUT=grp: grouped layout too large");
          sh->use_grp = true;
          sh->force_heat = 1032;  // kHeatGrp + 32
        }
      }
      if (sh->use_grp) {
        std::vector<int> ge(4 * gpairs, 0);
        for (int rb = 0; rb < nrb; ++rb)
          for (int c = 0; c < nc; ++c)
            for (int w = 0; w < BTPM_SH_WAVES; ++w)
              for (int p = 0; p < kPass; ++p) {
                const size_t i = (((size_t)rb * nc + c) * BTPM_SH_WAVES + w) * kPass + p;
                const size_t first = (size_t)gs[2 * i];
                for (int g = 0; g < 8; ++g) {
                  const int f = gp[(size_t)rb * BTPM_SH_ROWS + 64 * w + 8 * p + g];
                  if (f >= nf) continue;
                  const int k0 = cp[(size_t)c * nf + f], k1 = cp[(size_t)(c + 1) * nf + f];
                  for (int k = k0; k < k1; ++k) {
                    const size_t o = first + (size_t)((k - k0) / 2) * 8 + g;
                    const int h = 2 * ((k - k0) & 1);
                    ge[4 * o + h] = se[2 * (size_t)k] - c * kChunk;  // emitter within the chunk
                    ge[4 * o + h + 1] = se[2 * (size_t)k + 1];
                  }
                }
              }
        upload(sh->gperm, gp.data(), gp.size(), st);
        upload(sh->gseg, gs.data(), gs.size(), st);
        upload(sh->gent, ge.data(), ge.size(), st);
      }
      // the per-chunk CSR serves the time-lane kernel and btpm_selfheat_tiled
      upload(sh->sent, se.data(), se.size(), st);
      upload(sh->cptr, cp.data(), cp.size(), st);
      if (sh->use_ell) {
        std::vector<int> el(2 * total, 0);
        for (int rb = 0; rb < nrb; ++rb)
          for (int c = 0; c < nc; ++c)
            for (int w = 0; w < BTPM_SH_WAVES; ++w) {
              const size_t i = ((size_t)rb * nc + c) * BTPM_SH_WAVES + w;
              const size_t first = (size_t)ws[2 * i];
              for (int l = 0; l < 64; ++l) {
                const int f = rb * BTPM_SH_ROWS + w * 64 + l;
                if (f >= nf) continue;
                const int k0 = cp[(size_t)c * nf + f], k1 = cp[(size_t)(c + 1) * nf + f];
                for (int k = k0; k < k1; ++k) {
                  const size_t o = first + (size_t)(k - k0) * 64 + l;
                  el[2 * o] = se[2 * (size_t)k] - c * kChunk;  // emitter within the chunk
                  el[2 * o + 1] = se[2 * (size_t)k + 1];
                }
              }
            }
        upload(sh->wseg, ws.data(), ws.size(), st);
        upload(sh->ell, el.data(), el.size(), st);
      }
    }
  }
  // facet illumination -> blocked
  const size_t nser = (size_t)nb * nf * BTPM_BLK;
  {
    DBuf raw;
    upload(raw, body->illum, (size_t)nf * nt, st);
    sh->F.alloc(nser * sizeof(float));
    hipLaunchKernelGGL(btpm_pack_blocked, dim3(grid1((size_t)nf * nb, 256)), dim3(256), 0, st,
                       raw.as<float>(), sh->F.as<float>(), (const float*)nullptr, nf, nt, nb);
    HIPCHK(hipStreamSynchronize(st));
  }
  if (body->scatter && body->nnz > 0) {
    // singlescattering: F = D + VF (A D); multiplescattering: p = A(D + VF p)
    DBuf D, AD, p, y, part;
    D.alloc(nser * 4);
    AD.alloc(nser * 4);
    p.alloc(nser * 4);
    y.alloc(nser * 4);
    const unsigned nblk = grid1(nser, 256);
    part.alloc((size_t)nblk * 4);
    HIPCHK(hipMemcpyAsync(D.p, sh->F.p, nser * 4, hipMemcpyDeviceToDevice, st));
    hipLaunchKernelGGL(btpm_scale_cols, dim3(nblk), dim3(256), 0, st, AD.as<float>(),
                       D.as<float>(), sh->albedo.as<float>(), 1.0f, nf, nser);
    const int bpt = 16;
    dim3 g(grid1(nf, 256), grid1(nb, bpt));
    hipLaunchKernelGGL(btpm_spmm_blocked, g, dim3(256), 0, st, sh->rowptr.as<int>(),
                       sh->colidx.as<int>(), sh->vf.as<float>(), AD.as<float>(), D.as<float>(),
                       (const float*)nullptr, sh->F.as<float>(), nf, nb, bpt,
                       (const btpm_status_dev*)nullptr);
    if (body->scatter >= 2) {
      const float acc = body->scatter_acc > 0 ? body->scatter_acc : 1e-4f;
      hipLaunchKernelGGL(btpm_scale_cols, dim3(nblk), dim3(256), 0, st, p.as<float>(),
                         sh->F.as<float>(), sh->albedo.as<float>(), 1.0f, nf, nser);
      std::vector<float> hpart(nblk);
      for (int it = 0; it < 1000; ++it) {
        hipLaunchKernelGGL(btpm_spmm_blocked, g, dim3(256), 0, st, sh->rowptr.as<int>(),
                           sh->colidx.as<int>(), sh->vf.as<float>(), p.as<float>(), D.as<float>(),
                           sh->albedo.as<float>(), y.as<float>(), nf, nb, bpt,
                           (const btpm_status_dev*)nullptr);
        hipLaunchKernelGGL(btpm_delta_max, dim3(nblk), dim3(256), 0, st, p.as<float>(),
                           y.as<float>(), part.as<float>(), nser);
        HIPCHK(hipMemcpyAsync(hpart.data(), part.p, nblk * 4, hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        float m = 0.0f;
        for (float v : hpart) m = std::max(m, v);
        if (!(m > acc)) break;
      }
      hipLaunchKernelGGL(btpm_finish_scatter, dim3(nblk), dim3(256), 0, st, sh->F.as<float>(),
                         p.as<float>(), D.as<float>(), sh->albedo.as<float>(), nf, nser);
    }
    HIPCHK(hipStreamSynchronize(st));
  }
  if (crater) {
    sh->has_crater = 1;
    upload(sh->cvf32, crater->vf, (size_t)BTPM_NS * BTPM_NS, st);
    std::vector<_Float16> a;
    sh->vf_unscale = vf_tiles(crater->vf, a);
    upload(sh->cvfA, a.data(), a.size(), st);
    const size_t ncser = (size_t)nb * nf * BTPM_NS * BTPM_BLK;
    sh->cillum.alloc(ncser * sizeof(float));
    if (crater->drf && body->nnz > 0) {
      std::vector<int> cb(nf);
      size_t nchunks = 0;
      for (int f = 0; f < nf; ++f) {
        cb[f] = (int)nchunks;
        nchunks += (size_t)(body->rowptr[f + 1] - body->rowptr[f] + 31) / 32;
      }
      DBuf raw;
      upload_big(ctx, raw, crater->drf, (size_t)body->nnz * BTPM_NS);
      upload(sh->cbase, cb.data(), cb.size(), st);
      sh->drf_scale.alloc((size_t)nf * sizeof(float));
      sh->drfA.alloc(std::max<size_t>(nchunks, 1) * BTPM_DRF_TILE * sizeof(_Float16));
      hipLaunchKernelGGL(btpm_pack_drf, dim3(nf), dim3(256), 0, st, sh->rowptr.as<int>(),
                         raw.as<float>(), sh->cbase.as<int>(), sh->drf_scale.as<float>(),
                         sh->drfA.as<_Float16>());
      sh->wfm.alloc((size_t)nf * nb * BTPM_BLK * sizeof(float));
      HIPCHK(hipGetLastError());
      HIPCHK(hipStreamSynchronize(st));
    }
    if (crater->illum) {
      DBuf raw;
      upload_big(ctx, raw, crater->illum, (size_t)nf * nt * BTPM_NS);
      hipLaunchKernelGGL(btpm_pack_crater, dim3(grid1((size_t)nf * BTPM_NS * nb, 256)), dim3(256),
                         0, st, raw.as<float>(), sh->cillum.as<float>(), nf, nt, nb);
      HIPCHK(hipStreamSynchronize(st));
    } else {
      DBuf dsh, dcos, dsun, dvis, dnrm, rgi_blk, rgi;
      const bool by_sun = crater->sun_local != nullptr;
      if (by_sun) {  // shadow and cosine evaluated in the scattering kernel
        upload(dsun, crater->sun_local, (size_t)nf * nt * 3, st);
        upload(dvis, crater->vis_table, (size_t)crater->n_el * crater->n_az * BTPM_NS, st);
        upload(dnrm, crater->normals, (size_t)BTPM_NS * 3, st);
      } else {
        upload_big(ctx, dsh, crater->shadow, (size_t)nf * nt * BTPM_NS);
        upload_big(ctx, dcos, crater->cosang, (size_t)nf * nt * BTPM_NS);
      }
      const float* rgip = nullptr;
      if (sh->drfA.p) {
        // RGI = sum_g drf * A_g F_g(s): blocked over crater columns, then
        // unpacked to the reference layout the scattering kernel reads
        hipLaunchKernelGGL(btpm_weights_fm, dim3(grid1((size_t)nb * nf, 256)), dim3(256), 0, st,
                           sh->F.as<float>(), sh->albedo.as<float>(), 1.0f, 0, 0.0f, nf, nb,
                           sh->wfm.as<float>());
        rgi_blk.alloc(ncser * 4);
        dim3 g(8 * grid1(nf, 8) * grid1((size_t)nb * BTPM_BLK, 64));  // XCD-grouped facets
        hipLaunchKernelGGL(btpm_shape_to_crater, g, dim3(256), 0, st, sh->rowptr.as<int>(),
                           sh->colidx.as<int>(), sh->drfA.as<_Float16>(), sh->cbase.as<int>(),
                           sh->drf_scale.as<float>(), sh->wfm.as<float>(), nf, nt, nb,
                           (const float*)nullptr, (const float*)nullptr, 1.0f,
                           rgi_blk.as<float>());
        rgi.alloc((size_t)nf * nt * BTPM_NS * 4);
        hipLaunchKernelGGL(btpm_unpack_crater, dim3(grid1((size_t)nf * BTPM_NS * nb, 256)),
                           dim3(256), 0, st, rgi_blk.as<float>(), rgi.as<float>(), nf, nt, nb);
        HIPCHK(hipStreamSynchronize(st));
        rgi_blk.release();
        rgip = rgi.as<float>();
      }
      const float acc = crater->scatter_acc > 0 ? crater->scatter_acc : 1e-4f;
      hipLaunchKernelGGL(btpm_crater_scatter, dim3(nb, nf), dim3(128), 0, st,
                         sh->cvf32.as<float>(), sh->albedo.as<float>(), dsh.as<float>(),
                         dcos.as<float>(), dsun.as<float>(), dvis.as<float>(),
                         by_sun ? crater->n_el : 0, by_sun ? crater->n_az : 0, dnrm.as<float>(),
                         rgip, crater->solar, crater->multiple, acc, 10000, nf, nt, nb,
                         sh->cillum.as<float>());
      HIPCHK(hipStreamSynchronize(st));
    }
  }
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(st));
  *out = sh.release();
  API_END
}

int btpm_shape_destroy(btpm_shape* sh) {
  if (!sh) return BTPM_OK;
  CtxLock lk(sh->ctx);
  (void)hipSetDevice(sh->ctx->device);
  (void)hipStreamSynchronize(sh->ctx->stream);
  delete sh;
  return BTPM_OK;
}

int btpm_shape_facet_illum(btpm_shape* sh, float* out) {
  if (!sh || !out) return fail(BTPM_ERR_ARG, "null argument");
  CtxLock lk(sh->ctx);
  API_BEGIN
  btpm_ctx* c = sh->ctx;
  HIPCHK(hipSetDevice(c->device));
  DBuf raw;
  raw.alloc((size_t)sh->nf * sh->p.ntau * 4);
  hipLaunchKernelGGL(btpm_unpack_blocked, dim3(grid1((size_t)sh->nf * sh->nb, 256)), dim3(256), 0,
                     c->stream, sh->F.as<float>(), raw.as<float>(), sh->nf, sh->p.ntau, sh->nb);
  HIPCHK(hipMemcpyAsync(out, raw.p, raw.n, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  API_END
}

int btpm_shape_crater_illum(btpm_shape* sh, float* out) {
  if (!sh || !out) return fail(BTPM_ERR_ARG, "null argument");
  if (!sh->has_crater) return fail(BTPM_ERR_STATE, "shape has no crater model");
  CtxLock lk(sh->ctx);
  API_BEGIN
  btpm_ctx* c = sh->ctx;
  HIPCHK(hipSetDevice(c->device));
  DBuf raw;
  raw.alloc((size_t)sh->nf * sh->p.ntau * BTPM_NS * 4);
  hipLaunchKernelGGL(btpm_unpack_crater, dim3(grid1((size_t)sh->nf * BTPM_NS * sh->nb, 256)),
                     dim3(256), 0, c->stream, sh->cillum.as<float>(), raw.as<float>(), sh->nf,
                     sh->p.ntau, sh->nb);
  HIPCHK(hipMemcpyAsync(out, raw.p, raw.n, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  API_END
}

int btpm_solve_create(btpm_ctx* ctx, btpm_shape* shape, float ti, const float* ti_per_facet,
                      int with_craters, btpm_solve** out) {
  if (!ctx || !shape || !out) return fail(BTPM_ERR_ARG, "null argument");
  if (with_craters && !shape->has_crater) return fail(BTPM_ERR_ARG, "shape has no crater model");
  if (!(ti >= 0.0f)) return fail(BTPM_ERR_ARG, "negative thermal inertia");
  auto s = std::make_unique<btpm_solve>();
  s->ctx = ctx;
  s->shape = shape;
  s->ti = ti;
  if (ti_per_facet) s->ti_facet.assign(ti_per_facet, ti_per_facet + shape->nf);
  s->with_craters = with_craters;
  *out = s.release();
  return BTPM_OK;
}

int btpm_solve_set_flux_scale(btpm_solve* s, float scale) {
  if (!s) return fail(BTPM_ERR_ARG, "null argument");
  if (!(scale > 0.0f) || !std::isfinite(scale))
    return fail(BTPM_ERR_ARG, "flux scale must be positive and finite (got %g)", (double)scale);
  if (s->phase != 0 || s->queued) return fail(BTPM_ERR_STATE, "solve already started");
  s->flux_scale = scale;
  return BTPM_OK;
}

int btpm_solve_destroy(btpm_solve* s) {
  if (!s) return BTPM_OK;
  CtxLock lk(s->ctx);
  if (s->dependents > 0)
    return fail(BTPM_ERR_STATE, "the solve is the warm start of %d solve(s) that have not taken "
                "its state yet", s->dependents.load());
  (void)hipSetDevice(s->ctx->device);
  (void)hipStreamSynchronize(s->ctx->stream);
  if (s->warm_pending && s->warm) s->warm->dependents--;
  delete s;
  return BTPM_OK;
}

int btpm_solve_warm_start(btpm_solve* s, btpm_solve* from) {
  if (!s || !from) return fail(BTPM_ERR_ARG, "null argument");
  if (s == from) return fail(BTPM_ERR_ARG, "a solve cannot start from itself");
  if (s->shape != from->shape)
    return fail(BTPM_ERR_ARG, "warm start needs the same shape (state layouts differ)");
  CtxLock lk(s->ctx);
  if (s->phase != 0 || s->queued) return fail(BTPM_ERR_STATE, "solve already started");
  for (const btpm_solve* w = from; w; w = w->warm_pending ? w->warm : nullptr)
    if (w == s) return fail(BTPM_ERR_ARG, "warm-start chain would form a cycle");
  if (s->warm_pending && s->warm) s->warm->dependents--;
  s->warm = from;
  s->warm_pending = true;
  from->dependents++;
  return BTPM_OK;
}

int btpm_ctx_set_mem_limit(btpm_ctx* ctx, size_t bytes) {
  if (!ctx) return fail(BTPM_ERR_ARG, "null ctx");
  CtxLock lk(ctx);
  API_BEGIN
  HIPCHK(hipSetDevice(ctx->device));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  pool_trim(ctx->device);  // the cap counts cached blocks too: start from none
  pool_set_limit(ctx->device, bytes);
  API_END
}

int btpm_batch_run(btpm_ctx* ctx, btpm_solve** solves, int n, const btpm_batch_opts* opts,
                   btpm_batch_stats* stats) {
  if (!ctx || (n > 0 && !solves)) return fail(BTPM_ERR_ARG, "null argument");
  for (int i = 0; i < n; ++i)
    if (!solves[i] || solves[i]->phase != 0 || solves[i]->queued)
      return fail(BTPM_ERR_STATE, "solve %d not fresh", i);
  {
    // every warm-start predecessor has finished or runs in this batch
    std::set<const btpm_solve*> in(solves, solves + n);
    for (int i = 0; i < n; ++i) {
      const btpm_solve* w = solves[i]->warm_pending ? solves[i]->warm : nullptr;
      if (w && w->phase != 3 && !in.count(w))
        return fail(BTPM_ERR_STATE, "solve %d starts from a solve that neither finished nor runs "
                    "in this batch", i);
    }
  }
  CtxLock lk(ctx);
  if (ctx->servers > 0) return fail(BTPM_ERR_STATE, "a server is running on this context");
  API_BEGIN
  HIPCHK(hipSetDevice(ctx->device));
  btpm_batch_opts o;
  btpm_batch_opts_default(&o);
  if (opts) o = *opts;
  Sched sc(ctx, o);
  for (int i = 0; i < n; ++i) {
    solves[i]->t_submit = sc.t0;
    sc.pending.push_back(solves[i]);
  }
  while (!sc.pending.empty() || !sc.active.empty()) sc.step();
  sc.stats.wall_s =
      std::chrono::duration<double>(std::chrono::steady_clock::now() - sc.t0).count();
  if (stats) *stats = sc.stats;
  API_END
}

int btpm_solve_run(btpm_solve* s) {
  if (!s) return fail(BTPM_ERR_ARG, "null solve");
  btpm_batch_opts o;
  btpm_batch_opts_default(&o);
  o.rotations_per_sync = 1;
  o.use_graph = 0;
  return btpm_batch_run(s->ctx, &s, 1, &o, nullptr);
}

namespace {
// One set through the stepwise kernels until every group is frozen, the
// rotation cap of the parameters, or `cap` rotations (cap > 0).
void run_stepwise_set(btpm_solve* s, SetHost* S, int cap, int* rev_out, int* conv_out,
                      long long* colsteps_out) {
  btpm_shape* sh = s->shape;
  hipStream_t st = s->ctx->stream;
  const btpm_set_dev& d = S->d;
  const bool crater = d.kind == BTPM_SET_CRATER;
  const int nc = d.ncols, ng = d.ngroups;
  DBuf temp, Tp, dm;
  temp.alloc((size_t)(BTPM_NX + 1) * nc * 4);
  Tp.alloc(S->T.n);
  dm.alloc((size_t)nc * 4);
  hipLaunchKernelGGL(btpm_step_profile, dim3(grid1(nc, 256)), dim3(256), 0, st, temp.as<float>(),
                     S->prof.as<float>(), nc, 0);
  std::vector<uint8_t> fr((size_t)ng);
  HIPCHK(hipMemcpyAsync(fr.data(), S->frozen.p, ng, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  const bool csr = !crater && sh->nnz > 0;
  int rev = 0;
  bool done = false;
  long long cs = 0;
  while (!done && rev < d.max_rev && (cap <= 0 || rev < cap)) {
    // stemp0 <- stemp1 for the whole rotation, then the steps
    HIPCHK(hipMemcpyAsync(Tp.p, S->T.p, S->T.n, hipMemcpyDeviceToDevice, st));
    HIPCHK(hipMemsetAsync(dm.p, 0, dm.n, st));
    long long active = 0;
    for (uint8_t x : fr) active += x ? 0 : 1;
    for (int k = 0; k < d.ntau; ++k) {
      hipLaunchKernelGGL(btpm_step_fd, dim3(grid1(nc, 256)), dim3(256), 0, st, d,
                         temp.as<float>());
      if (crater)
        hipLaunchKernelGGL(btpm_step_bc_crater, dim3(ng), dim3(128), 0, st, d, temp.as<float>(),
                           (const float*)Tp.as<float>(), k, sh->cvf32.as<float>(),
                           dm.as<float>());
      else
        hipLaunchKernelGGL(btpm_step_bc_smooth, dim3(grid1(nc, 256)), dim3(256), 0, st, d,
                           temp.as<float>(), (const float*)Tp.as<float>(), k,
                           csr ? sh->rowptr.as<int>() : nullptr,
                           csr ? sh->colidx.as<int>() : nullptr,
                           csr ? sh->vf.as<float>() : nullptr, dm.as<float>());
      HIPCHK(hipStreamSynchronize(st));
    }
    cs += active * (crater ? BTPM_NS : 1) * d.ntau;
    hipLaunchKernelGGL(btpm_step_converged, dim3(grid1(ng, 256)), dim3(256), 0, st, d,
                       (const float*)dm.as<float>(), rev > d.check_after ? 1 : 0);
    HIPCHK(hipMemcpyAsync(fr.data(), S->frozen.p, ng, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    ++rev;
    done = true;
    for (uint8_t x : fr) done = done && x;
  }
  hipLaunchKernelGGL(btpm_step_profile, dim3(grid1(nc, 256)), dim3(256), 0, st, temp.as<float>(),
                     S->prof.as<float>(), nc, 1);
  HIPCHK(hipStreamSynchronize(st));
  *rev_out = rev;
  *conv_out = done ? 1 : 0;
  *colsteps_out = cs;
}
}  // namespace

int btpm_solve_run_stepwise(btpm_solve* s, int rotation_cap) {
  if (!s) return fail(BTPM_ERR_ARG, "null solve");
  if (s->phase != 0 || s->queued) return fail(BTPM_ERR_STATE, "solve not fresh");
  if (s->warm_pending && s->warm && s->warm->phase != 3)
    return fail(BTPM_ERR_STATE, "the warm-start predecessor has not finished");
  btpm_ctx* ctx = s->ctx;
  CtxLock lk(ctx);
  if (ctx->servers > 0) return fail(BTPM_ERR_STATE, "a server is running on this context");
  API_BEGIN
  HIPCHK(hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  s->t_submit = std::chrono::steady_clock::now();
  build_smooth(s, st);
  s->phase = 1;
  run_stepwise_set(s, s->smooth.get(), rotation_cap, &s->info.rotations, &s->info.converged,
                   &s->info.column_steps);
  shrink_smooth(s->smooth.get());
  if (s->with_craters) {
    build_crater(s, st);
    s->phase = 2;
    run_stepwise_set(s, s->crater.get(), rotation_cap, &s->info.crater_rotations,
                     &s->info.crater_converged, &s->info.crater_column_steps);
    s->crater->src.release();
  }
  s->info.latency_s =
      std::chrono::duration<double>(std::chrono::steady_clock::now() - s->t_submit).count();
  s->phase = 3;
  API_END
}

int btpm_solve_get_info(btpm_solve* s, btpm_solve_info* info) {
  if (!s || !info) return fail(BTPM_ERR_ARG, "null argument");
  *info = s->info;
  return BTPM_OK;
}

static int fetch_series(btpm_solve* s, SetHost* S, float* out, int crater) {
  btpm_ctx* c = s->ctx;
  CtxLock lk(c);
  API_BEGIN
  HIPCHK(hipSetDevice(c->device));
  const int nt = s->shape->p.ntau, nb = S->d.nb, nf = s->shape->nf;
  DBuf raw;
  raw.alloc((size_t)S->d.ncols * nt * 4);
  if (crater)
    hipLaunchKernelGGL(btpm_unpack_crater, dim3(grid1((size_t)S->d.ncols * nb, 256)), dim3(256),
                       0, c->stream, S->T.as<float>(), raw.as<float>(), nf, nt, nb);
  else
    hipLaunchKernelGGL(btpm_unpack_blocked, dim3(grid1((size_t)S->d.ncols * nb, 256)), dim3(256),
                       0, c->stream, S->T.as<float>(), raw.as<float>(), S->d.ncols, nt, nb);
  // large results leave through the context's pinned slices (host threads
  // copy one slice out while the DMA fills the other)
  download_staged(c, out, raw.p, raw.n);
  HIPCHK(hipStreamSynchronize(c->stream));
  API_END
}

extern "C++" {
namespace btpm {
// A finished solve's temperatures unpacked on the device for the Observe
// synthesis: surface [nf][ntau] into stemp and, unless ctemp is null, crater
// [nf][ntau][100] into ctemp, floored at +0 K like every readback.  The caller
// holds the context's lock and runs this inside its API_BEGIN (HIP failures
// throw); the kernels go on the context's stream.
int solve_temps_on_device(btpm_solve* s, btpm_ctx* ctx, int nf, int ntau, float* stemp,
                          float* ctemp) {
  if (s->ctx != ctx) return fail(BTPM_ERR_ARG, "the solve belongs to another context");
  if (s->phase != 3 || !s->smooth) return fail(BTPM_ERR_STATE, "solve not finished");
  if (s->shape->nf != nf || s->shape->p.ntau != ntau)
    return fail(BTPM_ERR_ARG, "solve of %d facets x %d steps, observation of %d x %d",
                s->shape->nf, s->shape->p.ntau, nf, ntau);
  if (ctemp && !s->crater) return fail(BTPM_ERR_STATE, "no finished crater solve");
  const int nb = s->smooth->d.nb;
  hipLaunchKernelGGL(btpm_unpack_blocked, dim3(grid1((size_t)nf * nb, 256)), dim3(256), 0,
                     ctx->stream, s->smooth->T.as<float>(), stemp, nf, ntau, nb);
  if (ctemp)
    hipLaunchKernelGGL(btpm_unpack_crater, dim3(grid1((size_t)nf * BTPM_NS * nb, 256)),
                       dim3(256), 0, ctx->stream, s->crater->T.as<float>(), ctemp, nf, ntau, nb);
  HIPCHK(hipGetLastError());
  return BTPM_OK;
}
}  // namespace btpm
}  // extern "C++"

int btpm_solve_surface_temps(btpm_solve* s, float* out) {
  if (!s || !out) return fail(BTPM_ERR_ARG, "null argument");
  if (s->phase != 3 || !s->smooth) return fail(BTPM_ERR_STATE, "solve not finished");
  return fetch_series(s, s->smooth.get(), out, 0);
}

int btpm_solve_crater_temps(btpm_solve* s, float* out) {
  if (!s || !out) return fail(BTPM_ERR_ARG, "null argument");
  if (s->phase != 3 || !s->crater) return fail(BTPM_ERR_STATE, "no finished crater solve");
  return fetch_series(s, s->crater.get(), out, 1);
}

// binary_writesurfacetemperature (T3905/main.cpp:2314-2341) straight from the
// device: facet chunks unpacked into the file's [facet][sub-facet][step]
// order, copied through two pinned buffers while the previous chunk is
// written; host memory stays at two chunks of at most 128 MB.
int btpm_solve_write_crater_dat(btpm_solve* s, const char* path, const float* cvtx, int ncv,
                                const int* cfv, int ncf) {
  if (!s || !path || !cvtx || !cfv) return fail(BTPM_ERR_ARG, "null argument");
  if (ncf != BTPM_NS) return fail(BTPM_ERR_ARG, "the crater model has %d facets", BTPM_NS);
  if (s->phase != 3 || !s->crater) return fail(BTPM_ERR_STATE, "no finished crater solve");
  btpm_ctx* c = s->ctx;
  CtxLock lk(c);
  FILE* out = std::fopen(path, "wb");
  if (!out) return fail(BTPM_ERR_ARG, "cannot open %s for writing", path);
  int rc = BTPM_OK;
  float* host[2] = {nullptr, nullptr};
  hipEvent_t done[2] = {nullptr, nullptr};
  try {
    HIPCHK(hipSetDevice(c->device));
    SetHost* S = s->crater.get();
    const int nt = s->shape->p.ntau, nb = S->d.nb, nf = s->shape->nf, ncols = S->d.ncols;
    const size_t per_facet = (size_t)BTPM_NS * nt;
    const int chunk = (int)std::max<size_t>(1, std::min<size_t>((size_t)nf, (128u << 20) /
                                                                          (per_facet * 4)));
    DBuf dev[2];
    for (int k = 0; k < 2; ++k) {
      dev[k].alloc((size_t)chunk * per_facet * 4);
      HIPCHK(hipHostMalloc((void**)&host[k], (size_t)chunk * per_facet * 4, hipHostMallocDefault));
      HIPCHK(hipEventCreateWithFlags(&done[k], hipEventDisableTiming));
    }
    if (std::fwrite(cvtx, sizeof(float) * 3, (size_t)ncv, out) != (size_t)ncv ||
        std::fwrite(cfv, sizeof(int) * 3, (size_t)ncf, out) != (size_t)ncf)
      throw HipError{fail(BTPM_ERR_ARG, "write to %s failed", path)};
    const int nchunks = (nf + chunk - 1) / chunk;
    auto issue = [&](int k) {
      const int f0 = k * chunk, nfc = std::min(chunk, nf - f0);
      const int nc = nfc * BTPM_NS;
      hipLaunchKernelGGL(btpm_unpack_crater_rows, dim3(grid1((size_t)nc * nb, 256)), dim3(256), 0,
                         c->stream, S->T.as<float>(), dev[k & 1].as<float>(), ncols,
                         f0 * BTPM_NS, nc, nt, nb);
      HIPCHK(hipGetLastError());
      HIPCHK(hipMemcpyAsync(host[k & 1], dev[k & 1].p, (size_t)nc * nt * 4, hipMemcpyDeviceToHost,
                            c->stream));
      HIPCHK(hipEventRecord(done[k & 1], c->stream));
    };
    if (nchunks > 0) issue(0);
    for (int k = 0; k < nchunks; ++k) {
      HIPCHK(hipEventSynchronize(done[k & 1]));
      if (k + 1 < nchunks) issue(k + 1);   // overlaps the write below
      const size_t n = (size_t)std::min(chunk, nf - k * chunk) * per_facet;
      if (std::fwrite(host[k & 1], sizeof(float), n, out) != n)
        throw HipError{fail(BTPM_ERR_ARG, "write to %s failed", path)};
    }
    HIPCHK(hipStreamSynchronize(c->stream));
  } catch (const HipError& e) {
    rc = e.code;
  } catch (const std::bad_alloc&) {
    rc = fail(BTPM_ERR_NOMEM, "host allocation failed");
  }
  for (int k = 0; k < 2; ++k) {
    if (host[k]) (void)hipHostFree(host[k]);
    if (done[k]) (void)hipEventDestroy(done[k]);
  }
  if (std::fclose(out) != 0 && rc == BTPM_OK) rc = fail(BTPM_ERR_ARG, "closing %s failed", path);
  return rc;
}

static int fetch_profile(btpm_solve* s, SetHost* S, float* out, int crater) {
  btpm_ctx* c = s->ctx;
  CtxLock lk(c);
  API_BEGIN
  HIPCHK(hipSetDevice(c->device));
  const int nc = S->d.ncols, np = BTPM_NX + 1;
  std::vector<float> h((size_t)BTPM_PSTRIDE * nc);
  HIPCHK(hipMemcpyAsync(h.data(), S->prof.p, h.size() * 4, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  // device [i][c] -> smooth [f][i], crater [f][i][j]
  if (!crater) {
    for (int f = 0; f < nc; ++f)
      for (int i = 0; i < np; ++i) out[(size_t)f * np + i] = h[(size_t)f * BTPM_PSTRIDE + i];
  } else {
    const int nf = nc / BTPM_NS;
    for (int f = 0; f < nf; ++f)
      for (int i = 0; i < np; ++i)
        for (int j = 0; j < BTPM_NS; ++j)
          out[((size_t)f * np + i) * BTPM_NS + j] = h[((size_t)f * BTPM_NS + j) * BTPM_PSTRIDE + i];
  }
  API_END
}

int btpm_solve_profile(btpm_solve* s, float* out) {
  if (!s || !out) return fail(BTPM_ERR_ARG, "null argument");
  if (s->phase != 3 || !s->smooth) return fail(BTPM_ERR_STATE, "solve not finished");
  return fetch_profile(s, s->smooth.get(), out, 0);
}

int btpm_solve_crater_profile(btpm_solve* s, float* out) {
  if (!s || !out) return fail(BTPM_ERR_ARG, "null argument");
  if (s->phase != 3 || !s->crater) return fail(BTPM_ERR_STATE, "no finished crater solve");
  return fetch_profile(s, s->crater.get(), out, 1);
}

int btpm_crater_contract(btpm_ctx* ctx, const float* vf, const float* e, int ncol, float* q) {
  if (!ctx || !vf || !e || !q || ncol < 1) return fail(BTPM_ERR_ARG, "bad argument");
  CtxLock lk(ctx);
  API_BEGIN
  HIPCHK(hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  std::vector<_Float16> a;
  const float unscale = vf_tiles(vf, a);
  DBuf dA, dE, dQ;
  upload(dA, a.data(), a.size(), st);
  upload(dE, e, (size_t)ncol * BTPM_NS, st);
  dQ.alloc((size_t)ncol * BTPM_NS * 4);
  const int per = BTPM_CRATERS_WG * BTPM_BLK;  // product columns per workgroup
  hipLaunchKernelGGL(btpm_contract_kernel, dim3((ncol + per - 1) / per), dim3(BTPM_CRATER_THREADS),
                     0, st, dA.as<_Float16>(), dE.as<float>(), dQ.as<float>(), ncol, unscale,
                     (const float*)nullptr);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(q, dQ.p, dQ.n, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  API_END
}

}  // extern "C"

// ===========================================================================
// server: continuous batching on a host thread
// ===========================================================================
struct btpm_server {
  btpm_ctx* ctx = nullptr;
  btpm_batch_opts opt{};
  std::thread th;
  std::mutex mu;
  std::condition_variable cv;
  std::deque<btpm_solve*> inbox;
  bool stop = false;
  std::mutex done_mu;  // completions, for btpm_server_wait_any
  std::condition_variable done_cv;
  std::atomic<int> err{BTPM_OK};
  std::string errmsg;
  btpm_batch_stats stats{};
};

// Records the error that ended the server loop and wakes every waiter
// (btpm_server_wait's predicate sees the error).
static void server_failed(btpm_server* srv, Sched* sc, int code) {
  srv->errmsg = last_error();
  srv->err = code;  // after errmsg: a waiter that sees err reads a complete message
  std::vector<btpm_solve*> all;
  if (sc) {
    all.insert(all.end(), sc->active.begin(), sc->active.end());
    all.insert(all.end(), sc->pending.begin(), sc->pending.end());
  }
  {
    std::lock_guard<std::mutex> lk(srv->mu);
    all.insert(all.end(), srv->inbox.begin(), srv->inbox.end());
  }
  for (btpm_solve* s : all) {
    std::lock_guard<std::mutex> lk(s->mu);
    s->cv.notify_all();
  }
  std::lock_guard<std::mutex> lk(srv->done_mu);
  srv->done_cv.notify_all();
}

static void server_loop(btpm_server* srv) {
  btpm_ctx* ctx = srv->ctx;
  std::unique_ptr<Sched> sc;
  try {
    {
      std::lock_guard<std::mutex> glk(ctx->mu);
      StreamScope scope(ctx->stream);
      HIPCHK(hipSetDevice(ctx->device));
      sc = std::make_unique<Sched>(ctx, srv->opt);
      sc->on_complete = [srv] {
        std::lock_guard<std::mutex> lk(srv->done_mu);
        srv->done_cv.notify_all();
      };
    }
    for (;;) {
      {
        std::unique_lock<std::mutex> lk(srv->mu);
        if (sc->active.empty() && sc->pending.empty())
          srv->cv.wait(lk, [&] { return srv->stop || !srv->inbox.empty(); });
        while (!srv->inbox.empty()) {
          btpm_solve* s = srv->inbox.front();
          srv->inbox.pop_front();
          sc->pending.push_back(s);
        }
        if (srv->stop && sc->active.empty() && sc->pending.empty()) break;
      }
      // the stream is the server's for one scheduling step at a time; API
      // calls waiting for it (results, destroys) go first
      while (ctx->waiting.load() > 0) std::this_thread::yield();
      std::lock_guard<std::mutex> glk(ctx->mu);
      StreamScope scope(ctx->stream);
      sc->step();
    }
    sc->stats.wall_s =
        std::chrono::duration<double>(std::chrono::steady_clock::now() - sc->t0).count();
    srv->stats = sc->stats;
  } catch (const HipError& e) {
    server_failed(srv, sc.get(), e.code);
  } catch (const std::bad_alloc&) {
    server_failed(srv, sc.get(), fail(BTPM_ERR_NOMEM, "server: host allocation failed"));
  } catch (const std::exception& e) {
    server_failed(srv, sc.get(), fail(BTPM_ERR_STATE, "server: internal error: %s", e.what()));
  }
  std::lock_guard<std::mutex> glk(ctx->mu);
  StreamScope scope(ctx->stream);
  sc.reset();
  ctx->servers--;
}

extern "C" {

int btpm_server_start(btpm_ctx* ctx, const btpm_batch_opts* opts, btpm_server** out) {
  if (!ctx || !out) return fail(BTPM_ERR_ARG, "null argument");
  {
    CtxLock lk(ctx);
    if (ctx->servers > 0) return fail(BTPM_ERR_STATE, "a server is already running on this context");
    ctx->servers++;
    // the idle blocks an earlier batch left in the device cache go back to
    // the driver before the first request: admission then budgets real free
    // memory, and a request's first large allocation does not wait for the
    // cache to be returned
    pool_trim(ctx->device);
  }
  auto srv = std::make_unique<btpm_server>();
  srv->ctx = ctx;
  btpm_batch_opts_default(&srv->opt);
  if (opts) srv->opt = *opts;
  srv->th = std::thread(server_loop, srv.get());
  *out = srv.release();
  return BTPM_OK;
}

int btpm_server_submit(btpm_server* srv, btpm_solve* s) {
  if (!srv || !s) return fail(BTPM_ERR_ARG, "null argument");
  if (s->phase != 0 || s->queued) return fail(BTPM_ERR_STATE, "solve already used");
  if (s->warm_pending && s->warm && s->warm->phase != 3 && !s->warm->queued)
    return fail(BTPM_ERR_STATE, "submit the warm-start predecessor first");
  {
    std::lock_guard<std::mutex> lk(srv->mu);
    if (srv->stop) return fail(BTPM_ERR_STATE, "server stopping");
    s->queued = true;
    s->t_submit = std::chrono::steady_clock::now();
    srv->inbox.push_back(s);
  }
  srv->cv.notify_all();
  return BTPM_OK;
}

int btpm_server_wait(btpm_server* srv, btpm_solve* s, double timeout_s) {
  if (!srv || !s) return fail(BTPM_ERR_ARG, "null argument");
  std::unique_lock<std::mutex> lk(s->mu);
  const bool ok = s->cv.wait_for(lk, std::chrono::duration<double>(timeout_s),
                                 [&] { return s->phase == 3 || srv->err != BTPM_OK; });
  if (srv->err != BTPM_OK) return fail(srv->err, "server failed: %s", srv->errmsg.c_str());
  if (!ok) return fail(BTPM_ERR_STATE, "timeout");
  return BTPM_OK;
}

int btpm_server_wait_any(btpm_server* srv, btpm_solve* const* solves, int n, double timeout_s,
                         int* index) {
  if (!srv || !solves || n < 1 || !index) return fail(BTPM_ERR_ARG, "bad argument");
  for (int i = 0; i < n; ++i)
    if (!solves[i]) return fail(BTPM_ERR_ARG, "null solve %d", i);
  int found = -1;
  std::unique_lock<std::mutex> lk(srv->done_mu);
  const bool ok = srv->done_cv.wait_for(lk, std::chrono::duration<double>(timeout_s), [&] {
    for (int i = 0; i < n; ++i)
      if (solves[i]->phase == 3) {
        found = i;
        return true;
      }
    return srv->err != BTPM_OK;
  });
  if (srv->err != BTPM_OK && found < 0)
    return fail(srv->err, "server failed: %s", srv->errmsg.c_str());
  if (!ok) return fail(BTPM_ERR_STATE, "timeout");
  *index = found;
  return BTPM_OK;
}

int btpm_server_stop(btpm_server* srv, btpm_batch_stats* stats) {
  if (!srv) return BTPM_OK;
  {
    std::lock_guard<std::mutex> lk(srv->mu);
    srv->stop = true;
  }
  srv->cv.notify_all();
  if (srv->th.joinable()) srv->th.join();
  const int err = srv->err;
  if (stats) *stats = srv->stats;
  std::string msg = srv->errmsg;
  delete srv;
  if (err != BTPM_OK) return fail(err, "server failed: %s", msg.c_str());
  return BTPM_OK;
}

}  // extern "C"
