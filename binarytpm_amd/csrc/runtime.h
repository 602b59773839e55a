// runtime.h -- host-side helpers shared by libbtpm's translation units:
// error reporting, RAII device buffers and uploads.
#pragma once

#include <stdexcept>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <atomic>
#include <exception>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/btpm.h"
#include "btpm_internal.h"

namespace btpm {

// message of the last failure on this thread (btpm_last_error)
std::string& last_error();

// a finished solve's temperatures unpacked on the device (host.hip)
int solve_temps_on_device(btpm_solve* s, btpm_ctx* ctx, int nf, int ntau, float* stemp,
                          float* ctemp);
// crater lit / visible fractions computed into device memory (viewfactor.hip)
void crater_fractions_on_device(btpm_ctx* ctx, const int* lit, const float* src, long long n,
                                const float* cvtx, const int* cfv, int ncf, int m, float radius,
                                float* frac);

inline int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  last_error() = buf;
  return code;
}

struct HipError {
  int code;
};

#define HIPCHK(x)                                                                   \
  do {                                                                              \
    hipError_t e_ = (x);                                                            \
    if (e_ != hipSuccess) {                                                         \
      ::btpm::fail(e_ == hipErrorOutOfMemory ? BTPM_ERR_NOMEM : BTPM_ERR_HIP,       \
                   "%s: %s (%s:%d)", #x, hipGetErrorString(e_), __FILE__, __LINE__); \
      throw ::btpm::HipError{e_ == hipErrorOutOfMemory ? BTPM_ERR_NOMEM : BTPM_ERR_HIP}; \
    }                                                                               \
  } while (0)

// Every exception stops at the C ABI: HIP and argument failures keep their
// code, host allocation failures are BTPM_ERR_NOMEM and any other C++
// exception (e.g. a thread the runtime could not start) is BTPM_ERR_STATE.
#define API_BEGIN try {
#define API_END                                              \
  }                                                          \
  catch (const ::btpm::HipError& e) {                        \
    return e.code;                                           \
  }                                                          \
  catch (const std::bad_alloc&) {                            \
    return ::btpm::fail(BTPM_ERR_NOMEM, "host allocation failed"); \
  }                                                          \
  catch (const std::exception& e) {                          \
    return ::btpm::fail(BTPM_ERR_STATE, "internal error: %s", e.what()); \
  }                                                          \
  return BTPM_OK;

inline int nblocks(int ntau) { return (ntau + BTPM_BLK - 1) / BTPM_BLK; }
// Workgroups covering n items, b per workgroup.  A dispatch's grid size is a
// 32-bit count of work-items per dimension: a launch past it would be cut
// short without an error, so a count that could reach it is refused here
// (element-wise passes over whole series use grid_stride instead).
inline unsigned grid1(size_t n, unsigned b) {
  if (n > (size_t)UINT32_MAX - 1024u) throw std::length_error("launch over 2^32 work-items");
  return (unsigned)((n + b - 1) / b);
}
// Workgroups of a grid-stride launch over n items (b lanes per workgroup):
// at most 16384, which already gives every CU 64 workgroups.
inline unsigned grid_stride(size_t n, unsigned b) {
  const size_t g = (n + b - 1) / b;
  return (unsigned)(g < 16384 ? (g > 0 ? g : 1) : 16384);
}

// device buffer owned by RAII
// Device memory cache.  Blocks released by DBuf stay cached per device and
// are handed out again best-fit (capacity within 1.25x of the request):
// returning tens of GB to the driver and allocating them again can stall a
// single hipMalloc for seconds.  A failed hipMalloc empties the device's
// cache and retries once.
// A block released while work on its stream may still use it (a copy in
// flight, or an error unwinding an API call) is tagged with the stream of
// the releasing thread's scope (StreamScope; CtxLock sets one) and is handed
// out again only to that stream, where stream order puts the new user's
// work after the old.  Blocks released outside any scope (after a
// synchronisation) are idle and go to any stream.
// An optional per-device limit caps the bytes the library holds (live plus
// cached): an allocation beyond it fails as if the device were full.
void* pool_alloc(size_t bytes);   // throws HipError
void pool_free(void* p);
size_t pool_cached(int device);   // bytes held in the cache
size_t pool_live(int device);     // bytes handed out and not yet released
void pool_trim(int device);       // return the device's cache to the driver
size_t pool_trim_bytes(int device, size_t bytes);  // at least `bytes` of it, largest first
void pool_set_limit(int device, size_t bytes);  // 0 = no limit
size_t pool_limit(int device);

// the stream that the calling thread's library work is enqueued on (null
// outside any scope)
hipStream_t& scope_stream();
struct StreamScope {
  hipStream_t prev;
  explicit StreamScope(hipStream_t s) : prev(scope_stream()) { scope_stream() = s; }
  ~StreamScope() { scope_stream() = prev; }
  StreamScope(const StreamScope&) = delete;
  StreamScope& operator=(const StreamScope&) = delete;
};

struct DBuf {
  void* p = nullptr;
  size_t n = 0;
  DBuf() = default;
  DBuf(const DBuf&) = delete;
  DBuf& operator=(const DBuf&) = delete;
  ~DBuf() { release(); }
  void release() {
    if (p) pool_free(p);
    p = nullptr;
    n = 0;
  }
  void alloc(size_t bytes) {
    release();
    if (bytes == 0) return;
    p = pool_alloc(bytes);
    n = bytes;
  }
  template <class T>
  T* as() const {
    return static_cast<T*>(p);
  }
};

template <class T>
void upload(DBuf& b, const T* h, size_t count, hipStream_t st) {
  b.alloc(count * sizeof(T));
  if (count) {
    HIPCHK(hipMemcpyAsync(b.p, h, count * sizeof(T), hipMemcpyHostToDevice, st));
    HIPCHK(hipStreamSynchronize(st));
  }
}

// VF^T as fp16 hi/lo tiles [2][112][112] (A[m=j][k=i] = vf[i*100+j]), zero
// padded, scaled by 2^s so the largest entry lies in [0.5, 1): the low
// halves then stay normal fp16 numbers.  Returns 2^-s.
float vf_tiles(const float* vf, std::vector<_Float16>& a);

}  // namespace btpm

// one device: its stream and the resident-set tables of the scheduler
struct btpm_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  btpm_set_dev* d_sets = nullptr;
  btpm_status_dev* d_status = nullptr;
  btpm_status_dev* h_status = nullptr;  // pinned mirror
  std::vector<int> free_slots;
  // Serialises every use of `stream` across threads (a server thread
  // capturing a graph on it must not see another thread's work or syncs):
  // the server holds it per scheduling step, the other entry points per call.
  std::mutex mu;
  std::atomic<int> waiting{0};  // callers blocked on mu: a server yields to them
  int servers = 0;              // running servers (at most one; guarded by mu)
  // autotuner decisions of this context (guarded by mu): (tunable, size
  // class) -> chosen value; kept for every later batch run or server
  std::map<std::pair<int, long long>, int> tuned;
  // two pinned staging slices for large uploads (allocated on first use,
  // freed with the context; used under mu)
  char* stage[2] = {nullptr, nullptr};
  hipEvent_t stage_ev[2] = {nullptr, nullptr};
};

namespace btpm {
// Host-to-device copy of a large pageable buffer, complete on return: slices
// are copied by host threads into one of the context's two pinned staging
// buffers while the DMA of the other slice runs on the context's stream.
// Small copies go straight to hipMemcpyAsync.  Caller holds ctx->mu.
void upload_staged(btpm_ctx* ctx, void* dst, const void* src, size_t bytes);
// The device-to-host counterpart: the DMA of the next slice runs while host
// threads copy the previous one out of its pinned buffer.
void download_staged(btpm_ctx* ctx, void* dst, const void* src, size_t bytes);
template <class T>
void upload_big(btpm_ctx* ctx, DBuf& b, const T* h, size_t count) {
  b.alloc(count * sizeof(T));
  upload_staged(ctx, b.p, h, count * sizeof(T));
}
}  // namespace btpm

// Takes ctx->mu for one API call, announcing itself so that a running
// server lets it in between two scheduling steps (std::mutex is not fair).
// Buffers released inside the call are tagged with the context's stream.
struct CtxLock {
  btpm_ctx* c;
  hipStream_t prev;
  explicit CtxLock(btpm_ctx* ctx) : c(ctx) {
    c->waiting++;
    c->mu.lock();
    c->waiting--;
    prev = btpm::scope_stream();
    btpm::scope_stream() = c->stream;
  }
  ~CtxLock() {
    btpm::scope_stream() = prev;
    c->mu.unlock();
  }
  CtxLock(const CtxLock&) = delete;
  CtxLock& operator=(const CtxLock&) = delete;
};

