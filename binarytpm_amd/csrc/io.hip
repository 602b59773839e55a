// io.cpp -- the Thermal stage's output files at scale (host code).
//
//   btpm_write_temperature_txt  writeglobaltemperature (T3905/main.cpp:
//       2290-2308): vertex lines, facet lines, then one surface temperature
//       per line, facet-major, printed as the default ostream prints (%g, 6
//       significant digits).  Formatting is split over threads by facet
//       ranges and written in order: the text, not the disk, is what makes
//       the reference's writer slow at NF x NTAU lines.
//   btpm_write_crater_dat       binary_writesurfacetemperature
//       (T3905/main.cpp:2314-2341, 175706/Run-04/thermal.cpp:3351-3381) from
//       host temperatures [facet][step][sub-facet], transposed one facet at
//       a time into the file's [facet][sub-facet][step] order (the streaming
//       writer straight from a solve's device memory is
//       btpm_solve_write_crater_dat, next to the other solve accessors).
//   btpm_stitch_crater_parts    stitch_temperatures (175706/Run-04/
//       thermal.cpp:104-244): part files of consecutive facet ranges joined
//       into one file, streamed (the reference holds every part in memory).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "btpm_internal.h"
#include "runtime.h"

using namespace btpm;

namespace {

void append_g(std::string& s, double x) {
  char buf[32];
  const int n = std::snprintf(buf, sizeof buf, "%g", x);
  s.append(buf, (size_t)n);
}

struct File {
  FILE* f = nullptr;
  explicit File(const char* path, const char* mode) : f(std::fopen(path, mode)) {}
  ~File() {
    if (f) std::fclose(f);
  }
  bool write(const void* p, size_t n) { return n == 0 || std::fwrite(p, 1, n, f) == n; }
};

}  // namespace

extern "C" int btpm_write_temperature_txt(const char* path, const double* vtx, int nv,
                                          const int* facets, int nf, const float* stemp, int ntau,
                                          int threads) {
  if (!path || !vtx || !facets || !stemp) return fail(BTPM_ERR_ARG, "null argument");
  if (nv < 0 || nf < 0 || ntau < 0) return fail(BTPM_ERR_ARG, "bad sizes");
  File out(path, "wb");
  if (!out.f) return fail(BTPM_ERR_ARG, "cannot open %s for writing", path);
  std::string head;
  head.reserve((size_t)nv * 40 + (size_t)nf * 20);
  for (int v = 0; v < nv; ++v) {
    for (int k = 0; k < 3; ++k) {
      append_g(head, vtx[3 * (size_t)v + k]);
      head.push_back(k < 2 ? '\t' : '\n');
    }
  }
  for (int f = 0; f < nf; ++f) {
    char buf[48];
    const int n = std::snprintf(buf, sizeof buf, "%d\t%d\t%d\n", facets[3 * f], facets[3 * f + 1],
                                facets[3 * f + 2]);
    head.append(buf, (size_t)n);
  }
  if (!out.write(head.data(), head.size())) return fail(BTPM_ERR_ARG, "write to %s failed", path);
  // the temperatures: facet ranges formatted in parallel, written in order
  int nt = threads > 0 ? threads : (int)std::max(1u, std::thread::hardware_concurrency());
  nt = std::max(1, std::min({nt, 64, std::max(1, nf)}));
  const int per = (nf + nt - 1) / std::max(nt, 1);
  const int rounds = 4;   // bounded memory: each round formats nt ranges of per/rounds facets
  const int sub = std::max(1, (per + rounds - 1) / rounds);
  std::vector<std::string> buf(nt);
  for (int f0 = 0; f0 < nf; f0 += sub * nt) {
    std::vector<std::thread> pool;
    for (int t = 0; t < nt; ++t) {
      const int a = f0 + t * sub, b = std::min(nf, a + sub);
      buf[t].clear();
      if (a >= b) continue;
      pool.emplace_back([&, t, a, b] {
        std::string& s = buf[t];
        s.reserve((size_t)(b - a) * ntau * 8);
        for (size_t i = (size_t)a * ntau; i < (size_t)b * ntau; ++i) {
          append_g(s, (double)stemp[i]);
          s.push_back('\n');
        }
      });
    }
    for (auto& th : pool) th.join();
    for (int t = 0; t < nt; ++t)
      if (!out.write(buf[t].data(), buf[t].size()))
        return fail(BTPM_ERR_ARG, "write to %s failed", path);
  }
  return BTPM_OK;
}

extern "C" int btpm_write_crater_dat(const char* path, const float* cvtx, int ncv, const int* cfv,
                                     int ncf, const float* temps, int nf, int ntau) {
  if (!path || !cvtx || !cfv || !temps) return fail(BTPM_ERR_ARG, "null argument");
  if (ncv < 0 || ncf < 1 || nf < 0 || ntau < 1) return fail(BTPM_ERR_ARG, "bad sizes");
  File out(path, "wb");
  if (!out.f) return fail(BTPM_ERR_ARG, "cannot open %s for writing", path);
  if (!out.write(cvtx, sizeof(float) * 3 * (size_t)ncv) ||
      !out.write(cfv, sizeof(int) * 3 * (size_t)ncf))
    return fail(BTPM_ERR_ARG, "write to %s failed", path);
  // temps [nf][ntau][ncf] -> file [nf][ncf][ntau], one facet at a time
  std::vector<float> row((size_t)ncf * ntau);
  for (int f = 0; f < nf; ++f) {
    const float* t = temps + (size_t)f * ntau * ncf;
    for (int s = 0; s < ntau; ++s)
      for (int j = 0; j < ncf; ++j) row[(size_t)j * ntau + s] = t[(size_t)s * ncf + j];
    if (!out.write(row.data(), row.size() * sizeof(float)))
      return fail(BTPM_ERR_ARG, "write to %s failed", path);
  }
  return BTPM_OK;
}

extern "C" int btpm_stitch_crater_parts(const char* out_path, const char* const* parts, int nparts,
                                        int ncv, int ncf, int ntau) {
  if (!out_path || !parts || nparts < 1 || ncv < 0 || ncf < 1 || ntau < 1)
    return fail(BTPM_ERR_ARG, "bad arguments");
  const size_t head = sizeof(float) * 3 * (size_t)ncv + sizeof(int) * 3 * (size_t)ncf;
  const size_t rec = sizeof(float) * (size_t)ncf * ntau;
  std::vector<char> hbuf(head);
  // the header of the last part, as the reference keeps the last one read
  {
    File last(parts[nparts - 1], "rb");
    if (!last.f || std::fread(hbuf.data(), 1, head, last.f) != head)
      return fail(BTPM_ERR_ARG, "cannot read the header of %s", parts[nparts - 1]);
  }
  File out(out_path, "wb");
  if (!out.f) return fail(BTPM_ERR_ARG, "cannot open %s for writing", out_path);
  if (!out.write(hbuf.data(), head)) return fail(BTPM_ERR_ARG, "write to %s failed", out_path);
  std::vector<char> buf(64u << 20);
  for (int k = 0; k < nparts; ++k) {
    File in(parts[k], "rb");
    if (!in.f) return fail(BTPM_ERR_ARG, "cannot open %s", parts[k]);
    std::fseek(in.f, 0, SEEK_END);
    const long size = std::ftell(in.f);
    if (size < (long)head || (size - (long)head) % (long)rec != 0)
      return fail(BTPM_ERR_ARG, "%s: %ld bytes is not a header plus whole facets", parts[k], size);
    std::fseek(in.f, (long)head, SEEK_SET);
    size_t left = (size_t)size - head;
    while (left > 0) {
      const size_t n = std::min(left, buf.size());
      if (std::fread(buf.data(), 1, n, in.f) != n) return fail(BTPM_ERR_ARG, "read of %s failed", parts[k]);
      if (!out.write(buf.data(), n)) return fail(BTPM_ERR_ARG, "write to %s failed", out_path);
      left -= n;
    }
  }
  return BTPM_OK;
}
