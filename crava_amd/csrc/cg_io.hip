// cg_io.hip -- CRAVA's binary grid file (FFTGrid::writeCravaFile / readCravaFile,
// fftgrid.cpp:2183-2271): a text tag line, the survey geometry (big-endian doubles and ints),
// the padded dimensions rnxp, nyp, nzp, then the rnxp * nyp * nzp floats of the in-place real
// layout, big-endian (NRLib's binary default, fileio.hpp:208-280).
//
// The byte order is swapped on the device (one pass over the grid in HBM); the file moves
// through a pinned host buffer in chunks.
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "cg_internal.h"

// named (not anonymous) so that profilers show the kernel's name
namespace cgio {

constexpr size_t kChunk = 64u << 20;  // bytes per file read / write
const char kTag[] = "crava_fftgrid_binary";

// Big-endian <-> host order of n 32-bit words, two per lane (n is even: rnxp is); src may
// equal dst.
__global__ void __launch_bounds__(cg::kThreads) k_bswap32(const uint2 *src, uint2 *dst, size_t pairs)
{
  for (size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x; p < pairs; p += (size_t)gridDim.x * blockDim.x) {
    const uint2 v = src[p];
    dst[p] = make_uint2(__builtin_bswap32(v.x), __builtin_bswap32(v.y));
  }
}

int bswap_grid(cg_ctx *ctx, const float *src, float *dst, size_t n)
{
  const size_t pairs = n / 2;
  size_t blocks = (pairs + cg::kThreads - 1) / cg::kThreads;
  if (blocks > 65536) blocks = 65536;
  if (blocks == 0) blocks = 1;
  hipEvent_t ev;
  cg::launch_begin(ctx, CG_K_ELEMENTWISE, &ev);
  k_bswap32<<<(unsigned)blocks, cg::kThreads, 0, ctx->stream>>>(reinterpret_cast<const uint2 *>(src),
                                                                reinterpret_cast<uint2 *>(dst), pairs);
  cg::launch_end(ctx, CG_K_ELEMENTWISE, ev);
  return cg::hip_fail(ctx, hipGetLastError(), "k_bswap32");
}

void put_be(std::vector<unsigned char> &b, const void *v, size_t n)
{
  const unsigned char *p = static_cast<const unsigned char *>(v);
  for (size_t i = 0; i < n; ++i) b.push_back(p[n - 1 - i]);  // host is little-endian
}

template <typename T>
T get_be(const unsigned char *&p)
{
  unsigned char r[sizeof(T)];
  for (size_t i = 0; i < sizeof(T); ++i) r[i] = p[sizeof(T) - 1 - i];
  p += sizeof(T);
  T v;
  std::memcpy(&v, r, sizeof(T));
  return v;
}

// A pinned host buffer for one call.
struct Pinned {
  void *p = nullptr;
  ~Pinned()
  {
    if (p) (void)hipHostFree(p);
  }
};

struct File {
  FILE *f = nullptr;
  ~File()
  {
    if (f) fclose(f);
  }
};

}  // namespace cgio

using namespace cgio;

extern "C" {

int cg_grid_write_crava(cg_grid *g, const char *path, const cg_crava_geometry *geo)
{
  if (!g || !path || !geo) return CG_EINVAL;
  cg_ctx *ctx = g->ctx;
  (void)hipSetDevice(ctx->device);
  std::vector<unsigned char> head(kTag, kTag + sizeof(kTag) - 1);
  head.push_back('\n');
  const double d1[4] = {geo->x0, geo->y0, geo->dx, geo->dy};
  for (double v : d1) put_be(head, &v, 8);
  put_be(head, &geo->nx, 4);
  put_be(head, &geo->ny, 4);
  const double d2[7] = {geo->il0, geo->xl0, geo->il_step_x, geo->il_step_y, geo->xl_step_x, geo->xl_step_y,
                        geo->angle};
  for (double v : d2) put_be(head, &v, 8);
  const int dims[3] = {g->rnxp, g->nyp, g->nzp};
  for (int v : dims) put_be(head, &v, 4);
  File file;
  file.f = fopen(path, "wb");
  if (!file.f) return cg::fail(ctx, CG_EIO, std::string("write_crava: cannot open '") + path + "'");
  if (fwrite(head.data(), 1, head.size(), file.f) != head.size())
    return cg::fail(ctx, CG_EIO, std::string("write_crava: cannot write '") + path + "'");
  const size_t bytes = g->rsize * sizeof(float);
  void *swapped = nullptr;
  hipError_t e = cg::dev_alloc(ctx, &swapped, bytes);
  if (e != hipSuccess) return cg::hip_fail(ctx, e, "write_crava: staging buffer");
  int rc = bswap_grid(ctx, g->d, static_cast<float *>(swapped), g->rsize);
  Pinned host;
  if (rc == CG_OK) {
    e = hipHostMalloc(&host.p, kChunk, hipHostMallocDefault);
    rc = cg::hip_fail(ctx, e, "write_crava: pinned buffer");
  }
  for (size_t off = 0; rc == CG_OK && off < bytes; off += kChunk) {
    const size_t n = std::min(kChunk, bytes - off);
    e = hipMemcpyAsync(host.p, static_cast<char *>(swapped) + off, n, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    rc = cg::hip_fail(ctx, e, "write_crava: copy out");
    if (rc == CG_OK && fwrite(host.p, 1, n, file.f) != n)
      rc = cg::fail(ctx, CG_EIO, std::string("write_crava: cannot write '") + path + "'");
  }
  cg::dev_free(ctx, swapped);
  if (rc == CG_OK && fflush(file.f) != 0)
    rc = cg::fail(ctx, CG_EIO, std::string("write_crava: cannot write '") + path + "'");
  return rc;
}

int cg_grid_read_crava(cg_grid *g, const char *path, cg_crava_geometry *geo_out)
{
  if (!g || !path) return CG_EINVAL;
  cg_ctx *ctx = g->ctx;
  (void)hipSetDevice(ctx->device);
  File file;
  file.f = fopen(path, "rb");
  if (!file.f) return cg::fail(ctx, CG_EIO, std::string("read_crava: cannot open '") + path + "'");
  // the tag line is read and not checked (getline in readCravaFile)
  int c;
  while ((c = fgetc(file.f)) != EOF && c != '\n') {
  }
  unsigned char raw[11 * 8 + 5 * 4];
  if (c == EOF || fread(raw, 1, sizeof(raw), file.f) != sizeof(raw))
    return cg::fail(ctx, CG_EIO, std::string("read_crava: '") + path + "' ends inside the header");
  const unsigned char *p = raw;
  cg_crava_geometry geo;
  geo.x0 = get_be<double>(p);
  geo.y0 = get_be<double>(p);
  geo.dx = get_be<double>(p);
  geo.dy = get_be<double>(p);
  geo.nx = get_be<int>(p);
  geo.ny = get_be<int>(p);
  geo.il0 = get_be<double>(p);
  geo.xl0 = get_be<double>(p);
  geo.il_step_x = get_be<double>(p);
  geo.il_step_y = get_be<double>(p);
  geo.xl_step_x = get_be<double>(p);
  geo.xl_step_y = get_be<double>(p);
  geo.angle = get_be<double>(p);
  const int rnxp = get_be<int>(p), nyp = get_be<int>(p), nzp = get_be<int>(p);
  if (rnxp != g->rnxp || nyp != g->nyp || nzp != g->nzp)
    return cg::fail(ctx, CG_EINVAL,
                    "read_crava: Grid dimension is wrong for file '" + std::string(path) + "': file " +
                        std::to_string(rnxp) + " x " + std::to_string(nyp) + " x " + std::to_string(nzp) +
                        ", grid " + std::to_string(g->rnxp) + " x " + std::to_string(g->nyp) + " x " +
                        std::to_string(g->nzp));
  const size_t bytes = g->rsize * sizeof(float);
  Pinned host;
  hipError_t e = hipHostMalloc(&host.p, kChunk, hipHostMallocDefault);
  int rc = cg::hip_fail(ctx, e, "read_crava: pinned buffer");
  for (size_t off = 0; rc == CG_OK && off < bytes; off += kChunk) {
    const size_t n = std::min(kChunk, bytes - off);
    if (fread(host.p, 1, n, file.f) != n) {
      rc = cg::fail(ctx, CG_EIO, std::string("read_crava: '") + path + "' ends inside the grid");
      break;
    }
    e = hipMemcpyAsync(reinterpret_cast<char *>(g->d) + off, host.p, n, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    rc = cg::hip_fail(ctx, e, "read_crava: copy in");
  }
  if (rc == CG_OK) rc = bswap_grid(ctx, g->d, g->d, g->rsize);
  if (rc == CG_OK) {
    g->transformed = 0;
    if (geo_out) *geo_out = geo;
  }
  return rc;
}

}  // extern "C"
