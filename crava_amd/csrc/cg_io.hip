// cg_io.hip -- CRAVA's binary grid file (FFTGrid::writeCravaFile / readCravaFile,
// fftgrid.cpp:2183-2271): a text tag line, the survey geometry (big-endian doubles and ints),
// the padded dimensions rnxp, nyp, nzp, then the rnxp * nyp * nzp floats of the in-place real
// layout, big-endian (NRLib's binary default, fileio.hpp:208-280).
//
// The byte order is swapped on the device (one pass over the grid in HBM); the file moves
// through a pinned host buffer in chunks.
#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdlib>
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

// Big-endian copy of the box [0, nx) x [0, ny) x [0, nz) of a real grid (row pitch rnxp, nyp rows
// per plane) into a contiguous nz * ny * nx array, i fastest (the STORM file order), and back.
__global__ void __launch_bounds__(cg::kThreads)
k_box_to_be(const float *rv, unsigned *out, int nx, int ny, int nz, int rnxp, int nyp)
{
  const size_t n = (size_t)nx * ny * nz;
  for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (size_t)gridDim.x * blockDim.x) {
    const size_t i = t % nx, jk = t / nx, j = jk % ny, k = jk / ny;
    out[t] = __builtin_bswap32(__float_as_uint(rv[i + (size_t)rnxp * (j + (size_t)nyp * k)]));
  }
}

__global__ void __launch_bounds__(cg::kThreads) k_be_to_host32(const unsigned *in, float *out, size_t n)
{
  for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (size_t)gridDim.x * blockDim.x)
    out[t] = __uint_as_float(__builtin_bswap32(in[t]));
}

unsigned stream_blocks(size_t n)
{
  size_t b = (n + cg::kThreads - 1) / cg::kThreads;
  if (b > 65536) b = 65536;
  return (unsigned)(b ? b : 1);
}

std::string fixed6(double v)
{
  char buf[64];
  snprintf(buf, sizeof(buf), "%.6f", v);  // NRLib::ToString(v, 6): std::fixed, precision 6
  return buf;
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

// ---------------------------------------------------------------- SEG-Y
// One output trace of FFTGrid::writeSegyFile: grid column (i, j), the samples [first, end) taken
// from the grid (zero elsewhere), the trace's top time and cell thickness.
struct SegyTrace {
  int i, j, first, end;
  double top, gdz;
};

// IEEE single -> IBM single (hexadecimal) float by its definition, truncating the fraction:
// value = (-1)^s 0.F 16^(E - 64) with a 24-bit F in [1/16, 1).  Both zeros give 0.  (Normal floats
// only are exact: an IEEE denormal becomes an unnormalised IBM fraction.)
__host__ __device__ inline unsigned ieee_to_ibm(float f)
{
  const unsigned u = __builtin_bit_cast(unsigned, f);
  if ((u & 0x7fffffffu) == 0) return 0u;
  const unsigned e = (u >> 23) & 0xffu;
  const unsigned mant = e ? ((u & 0x7fffffu) | 0x800000u) : (u & 0x7fffffu);  // f = mant 2^(p - 24)
  const int p = (e ? (int)e : 1) - 126;
  const int q = (p + 3) >> 2;            // ceil(p / 4): f = (mant >> (4q - p)) 2^-24 16^q
  const unsigned frac = mant >> (unsigned)(4 * q - p);
  return (u & 0x80000000u) | ((unsigned)(q + 64) << 24) | frac;
}

// Samples of the sorted output traces: FFTGrid::getRegularZInterpolatedRealValue
// (fftgrid.cpp:2275-2306) for k in [first, end), with its float/double roundings, the truncation
// of the position to an index and the clamps at both ends; getRealValue's RMISSING outside the
// box (fftgrid.cpp:1237-1262).  Written as big-endian IBM floats, trace after trace.
__global__ void __launch_bounds__(cg::kThreads)
k_segy_traces(const float *__restrict__ g, int rnxp, int nyp, int nz, const SegyTrace *__restrict__ tr, int ns,
              double z0, double dz, unsigned *__restrict__ out)
{
#pragma clang fp contract(off)
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= ns) return;
  const SegyTrace t = tr[blockIdx.y];
  float v = 0.f;
  if (k >= t.first && k < t.end) {
    const float z = (float)(z0 + dz * (double)k);
    float tt = (float)(((double)z - t.top) / t.gdz);
    int kg = (int)tt;
    tt -= (float)kg;
    bool missing = false;
    if (kg < 0) {
      if (kg < -1) missing = true;
      else {
        kg = 0;
        tt = 0.f;
      }
    } else if (kg > nz - 2) {
      if (kg > nz - 1) missing = true;
      else {
        kg = nz - 2;
        tt = 1.f;
      }
    }
    if (missing) {
      v = -99999.f;
    } else {
      const size_t col = (size_t)t.i + (size_t)rnxp * (size_t)t.j, slab = (size_t)rnxp * nyp;
      const float v1 = kg >= 0 && kg < nz ? g[col + slab * kg] : -99999.f;
      const float v2 = kg + 1 >= 0 && kg + 1 < nz ? g[col + slab * (kg + 1)] : -99999.f;
      v = (1.f - tt) * v1 + tt * v2;
    }
  }
  out[(size_t)blockIdx.y * ns + k] = __builtin_bswap32(ieee_to_ibm(v));
}

void put_be16(unsigned char *b, int v)
{
  b[0] = (unsigned char)((v >> 8) & 0xff);
  b[1] = (unsigned char)(v & 0xff);
}

void put_be32(unsigned char *b, int v)
{
  const unsigned u = (unsigned)v;
  b[0] = (unsigned char)(u >> 24);
  b[1] = (unsigned char)(u >> 16);
  b[2] = (unsigned char)(u >> 8);
  b[3] = (unsigned char)u;
}

// ASCII -> EBCDIC (code page 037) for the printable characters a textual header uses; others
// become EBCDIC spaces.
unsigned char ebcdic(char c)
{
  static const char *lower = "abcdefghijklmnopqrstuvwxyz", *upper = "ABCDEFGHIJKLMNOPQRSTUVWXYZ";
  static const unsigned char lo[26] = {0x81, 0x82, 0x83, 0x84, 0x85, 0x86, 0x87, 0x88, 0x89, 0x91, 0x92, 0x93, 0x94,
                                       0x95, 0x96, 0x97, 0x98, 0x99, 0xa2, 0xa3, 0xa4, 0xa5, 0xa6, 0xa7, 0xa8, 0xa9};
  for (int i = 0; i < 26; ++i) {
    if (c == lower[i]) return lo[i];
    if (c == upper[i]) return (unsigned char)(lo[i] + 0x40);
  }
  if (c >= '0' && c <= '9') return (unsigned char)(0xf0 + (c - '0'));
  switch (c) {
    case '.': return 0x4b;
    case '<': return 0x4c;
    case '(': return 0x4d;
    case '+': return 0x4e;
    case '&': return 0x50;
    case '*': return 0x5c;
    case ')': return 0x5d;
    case ';': return 0x5e;
    case '-': return 0x60;
    case '/': return 0x61;
    case ',': return 0x6b;
    case '%': return 0x6c;
    case '_': return 0x6d;
    case '>': return 0x6e;
    case '?': return 0x6f;
    case ':': return 0x7a;
    case '#': return 0x7b;
    case '@': return 0x7c;
    case '\'': return 0x7d;
    case '=': return 0x7e;
    case '"': return 0x7f;
  }
  return 0x40;
}

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

int cg_grid_write_storm(cg_grid *g, const char *path, const cg_storm_geometry *geo, int padding)
{
  if (!g || !path || !geo) return CG_EINVAL;
  cg_ctx *ctx = g->ctx;
  if (g->transformed) return cg::fail(ctx, CG_ESTATE, "write_storm: grid is in the Fourier domain");
  (void)hipSetDevice(ctx->device);
  const int nx = padding ? g->nxp : g->nx, ny = padding ? g->nyp : g->ny, nz = padding ? g->nzp : g->nz;
  // Simbox::getStormHeader (simbox.cpp:1012-1034), binary form
  std::string head = "storm_petro_binary\n";
  head += "0 " + std::to_string(g->cubetype) + " " + fixed6(-99999.0) + "\n";
  head += "FFTGrid\n";
  if (geo->top_name && geo->bot_name)
    head += fixed6(geo->x_min) + " " + fixed6(geo->lx) + " " + fixed6(geo->y_min) + " " + fixed6(geo->ly) + " " +
            geo->top_name + " " + geo->bot_name + " 0.0 0.0\n";
  else
    head += fixed6(geo->x_min) + " " + fixed6(geo->lx) + " " + fixed6(geo->y_min) + " " + fixed6(geo->ly) +
            " 0.0 " + fixed6(geo->lz) + " 0.0 0.0\n";
  head += fixed6(geo->lz) + " " + fixed6(geo->angle * 180.0 / M_PI) + "\n\n";
  head += std::to_string(nx) + " " + std::to_string(ny) + " " + std::to_string(nz) + "\n";
  File file;
  file.f = fopen(path, "wb");
  if (!file.f) return cg::fail(ctx, CG_EIO, std::string("write_storm: cannot open '") + path + "'");
  if (fwrite(head.data(), 1, head.size(), file.f) != head.size())
    return cg::fail(ctx, CG_EIO, std::string("write_storm: cannot write '") + path + "'");
  const size_t n = (size_t)nx * ny * nz, bytes = n * sizeof(float);
  void *packed = nullptr;
  hipError_t e = cg::dev_alloc(ctx, &packed, std::max<size_t>(bytes, 4));
  if (e != hipSuccess) return cg::hip_fail(ctx, e, "write_storm: staging buffer");
  hipEvent_t ev;
  cg::launch_begin(ctx, CG_K_ELEMENTWISE, &ev);
  k_box_to_be<<<stream_blocks(n), cg::kThreads, 0, ctx->stream>>>(g->d, static_cast<unsigned *>(packed), nx, ny, nz,
                                                                   g->rnxp, g->nyp);
  cg::launch_end(ctx, CG_K_ELEMENTWISE, ev);
  int rc = cg::hip_fail(ctx, hipGetLastError(), "k_box_to_be");
  Pinned host;
  if (rc == CG_OK) {
    e = hipHostMalloc(&host.p, kChunk, hipHostMallocDefault);
    rc = cg::hip_fail(ctx, e, "write_storm: pinned buffer");
  }
  for (size_t off = 0; rc == CG_OK && off < bytes; off += kChunk) {
    const size_t m = std::min(kChunk, bytes - off);
    e = hipMemcpyAsync(host.p, static_cast<char *>(packed) + off, m, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    rc = cg::hip_fail(ctx, e, "write_storm: copy out");
    if (rc == CG_OK && fwrite(host.p, 1, m, file.f) != m)
      rc = cg::fail(ctx, CG_EIO, std::string("write_storm: cannot write '") + path + "'");
  }
  cg::dev_free(ctx, packed);
  if (rc == CG_OK && (fputs("0\n", file.f) < 0 || fflush(file.f) != 0))
    rc = cg::fail(ctx, CG_EIO, std::string("write_storm: cannot write '") + path + "'");
  return rc;
}

// The next whitespace-separated token of the header (NRLib::ReadNext), or "" at the end.
static std::string next_token(FILE *f)
{
  std::string t;
  int c;
  while ((c = fgetc(f)) != EOF && isspace(c)) {
  }
  while (c != EOF && !isspace(c)) {
    t.push_back((char)c);
    c = fgetc(f);
  }
  if (c != EOF) ungetc(c, f);
  return t;
}

int cg_grid_read_storm(cg_grid *g, const char *path, cg_storm_geometry *geo_out)
{
  if (!g || !path) return CG_EINVAL;
  cg_ctx *ctx = g->ctx;
  (void)hipSetDevice(ctx->device);
  File file;
  file.f = fopen(path, "rb");
  if (!file.f) return cg::fail(ctx, CG_EIO, std::string("read_storm: cannot open '") + path + "'");
  // StormContGrid::ReadFromFile (stormcontgrid.cpp:82-170): format, zone number, model file name,
  // missing code, variable name, the volume (x0 lx y0 ly top base erosion-top erosion-base, lz angle),
  // nx ny nz, the rest of that line, then the big-endian floats and the barrier count
  const std::string fmt = next_token(file.f);
  if (fmt != "storm_petro_binary")
    return cg::fail(ctx, CG_EINVAL, "read_storm: '" + std::string(path) + "' is not a STORM binary grid ('" + fmt + "')");
  std::vector<std::string> tok;
  for (int t = 0; t < 4 + 8 + 2 + 3; ++t) tok.push_back(next_token(file.f));
  if (tok.back().empty()) return cg::fail(ctx, CG_EIO, std::string("read_storm: '") + path + "' ends inside the header");
  const int nx = std::atoi(tok[14].c_str()), ny = std::atoi(tok[15].c_str()), nz = std::atoi(tok[16].c_str());
  if (nx != g->nx || ny != g->ny || nz != g->nz)
    return cg::fail(ctx, CG_EINVAL, "read_storm: the file's grid " + std::to_string(nx) + " x " + std::to_string(ny) +
                                        " x " + std::to_string(nz) + " is not the grid's box " + std::to_string(g->nx) +
                                        " x " + std::to_string(g->ny) + " x " + std::to_string(g->nz));
  int c;
  while ((c = fgetc(file.f)) != EOF && c != '\n') {  // DiscardRestOfLine
  }
  const size_t n = (size_t)nx * ny * nz, bytes = n * sizeof(float);
  void *cube = nullptr;
  hipError_t e = cg::dev_alloc(ctx, &cube, std::max<size_t>(bytes, 4));
  if (e != hipSuccess) return cg::hip_fail(ctx, e, "read_storm: staging buffer");
  Pinned host;
  e = hipHostMalloc(&host.p, kChunk, hipHostMallocDefault);
  int rc = cg::hip_fail(ctx, e, "read_storm: pinned buffer");
  for (size_t off = 0; rc == CG_OK && off < bytes; off += kChunk) {
    const size_t m = std::min(kChunk, bytes - off);
    if (fread(host.p, 1, m, file.f) != m) {
      rc = cg::fail(ctx, CG_EIO, std::string("read_storm: '") + path + "' ends inside the grid");
      break;
    }
    e = hipMemcpyAsync(static_cast<char *>(cube) + off, host.p, m, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    rc = cg::hip_fail(ctx, e, "read_storm: copy in");
  }
  if (rc == CG_OK) {
    k_be_to_host32<<<stream_blocks(n), cg::kThreads, 0, ctx->stream>>>(static_cast<const unsigned *>(cube),
                                                                        static_cast<float *>(cube), n);
    rc = cg::hip_fail(ctx, hipGetLastError(), "k_be_to_host32");
  }
  // FFTGrid(const NRLib::Grid<float> *, nxp, nyp, nzp): the cube and its padding fill
  if (rc == CG_OK) rc = cg_grid_fill_padded(g, static_cast<const float *>(cube), 1);
  cg::dev_free(ctx, cube);
  if (rc == CG_OK && geo_out) {
    geo_out->x_min = std::atof(tok[4].c_str());
    geo_out->lx = std::atof(tok[5].c_str());
    geo_out->y_min = std::atof(tok[6].c_str());
    geo_out->ly = std::atof(tok[7].c_str());
    geo_out->lz = std::atof(tok[12].c_str());
    geo_out->angle = std::atof(tok[13].c_str()) * M_PI / 180.0;
    geo_out->top_name = nullptr;
    geo_out->bot_name = nullptr;
  }
  return rc;
}

}  // extern "C"

extern "C" {

int cg_grid_write_segy(cg_grid *g, const char *path, const cg_segy_geometry *geo)
{
  if (!g || !path || !geo) return CG_EINVAL;
  cg_ctx *ctx = g->ctx;
  if (g->transformed) return cg::fail(ctx, CG_ESTATE, "write_segy: grid is in the Fourier domain");
  if (geo->nx != g->nx || geo->ny != g->ny || !geo->top || !geo->rel_thick || geo->ntext < 0 || geo->ntext > 38 ||
      (geo->ntext > 0 && !geo->text))
    return cg::fail(ctx, CG_EINVAL, "write_segy: geometry of another box, or missing tables");
  (void)hipSetDevice(ctx->device);
  // FFTGrid::writeSegyFile (fftgrid.cpp:1960-2050): sample interval rounded to whole ms, enough
  // samples from z0 to the simbox's bottom
  float dzf = (float)std::floor(geo->dz + 0.5);
  if (dzf == 0.f) dzf = 1.f;
  const int ns = (int)std::ceil((geo->zmax - (double)geo->z0) / (double)dzf);
  if (ns < 1 || ns > 32767) return cg::fail(ctx, CG_EINVAL, "write_segy: sample count outside 1..32767");
  // traces: cell centres (Simbox::getXYCoord), inline / crossline of the float coordinates
  // (SegyGeometry::FindILXL), sorted by inline then crossline as SegY::WriteAllTracesToFile does;
  // traces with a missing top are not written
  struct Out {
    SegyTrace t;
    int il, xl, ux, uy;
  };
  std::vector<Out> tr;
  tr.reserve((size_t)g->nx * g->ny);
  const double cr = std::cos(geo->angle), sr = std::sin(geo->angle);
  for (int j = 0; j < g->ny; ++j)
    for (int i = 0; i < g->nx; ++i) {
      const double top = geo->top[(size_t)j * g->nx + i];
      if (top == -99999.0 || top == -88888.0) continue;  // RMISSING, WELLMISSING
      const double rx = ((double)i + 0.5) * geo->dx, ry = ((double)j + 0.5) * geo->dy;
      const double x = rx * cr - ry * sr + geo->x0, y = rx * sr + ry * cr + geo->y0;
      const float xf = (float)x, yf = (float)y;
      Out o;
      o.t.i = i;
      o.t.j = j;
      o.t.top = top;
      o.t.gdz = geo->dz * geo->rel_thick[(size_t)j * g->nx + i];
      o.t.first = (int)std::floor(0.5 + (top - (double)geo->z0) / (double)dzf);
      o.t.end = std::min(ns, (int)std::floor(0.5 + ((top - (double)geo->z0) + g->nz * o.t.gdz) / (double)dzf));
      o.il = (int)(0.5 + geo->il0 + ((double)xf - geo->x0) * geo->il_step_x + ((double)yf - geo->y0) * geo->il_step_y);
      o.xl = (int)(0.5 + geo->xl0 + ((double)xf - geo->x0) * geo->xl_step_x + ((double)yf - geo->y0) * geo->xl_step_y);
      o.ux = (int)(double)xf;  // TraceHeader::SetUtmx with scalco 1, written as int
      o.uy = (int)(double)yf;
      tr.push_back(o);
    }
  std::stable_sort(tr.begin(), tr.end(),
                   [](const Out &a, const Out &b) { return a.il < b.il || (a.il == b.il && a.xl < b.xl); });
  File file;
  file.f = fopen(path, "wb");
  if (!file.f) return cg::fail(ctx, CG_EIO, std::string("write_segy: cannot open '") + path + "'");
  // textual header: 40 lines of 80 EBCDIC characters, "C" + line number, text from column 5.
  // With the caller's lines (from the second line on, 75 characters each:
  // TextualHeader::SetLine(i + 1, text), as writeSegyFile places them) on NRLib's empty header,
  // whose last line "C40 END EBCDIC" is 79 characters and whose 3200th byte is the string's
  // terminating zero (commonheaders.cpp:60-104, 152-213); with none, TextualHeader::standardHeader
  // (:106-163): the writing program's credits on lines 12-28 and a full last line.
  static const struct {
    int line;
    const char *text;
  } kStandard[] = {{12, "       THIS SEGY FILE WAS WRITTEN BY"},
                   {14, "       THE CRAVA PROGRAM"},
                   {16, "       DEVELOPED BY"},
                   {18, "       RAGNAR HAUGE"},
                   {20, "       ANNE RANDI SYVERSVEEN           AND"},
                   {22, "       ODD KOLBJORNSEN"},
                   {24, "       AT"},
                   {26, "       THE NORWEGIAN COMPUTING CENTER"},
                   {28, "       2003"}};
  std::vector<unsigned char> head(3600, 0);
  for (int l = 0; l < 40; ++l) {
    char line[96];
    const char *text = "";
    if (geo->ntext > 0) {
      if (l >= 1 && l - 1 < geo->ntext && geo->text[l - 1]) text = geo->text[l - 1];
    } else {
      for (const auto &st : kStandard)
        if (st.line == l + 1) text = st.text;
    }
    if (l == 39) snprintf(line, sizeof(line), "C40 END EBCDIC%66s", "");
    else snprintf(line, sizeof(line), "C%2d  %-75.75s", l + 1, text);
    for (int c = 0; c < 80; ++c) head[80 * l + c] = ebcdic(line[c]);
  }
  if (geo->ntext > 0) head[3199] = 0;  // the empty header's string ends one character short
  // binary header (BinaryHeader::Write, commonheaders.cpp:234-270)
  unsigned char *bh = head.data() + 3200;
  put_be32(bh + 0, 9999);
  put_be32(bh + 4, 1);
  put_be32(bh + 8, 1);
  put_be16(bh + 12, (short)g->nx);
  put_be16(bh + 16, (short)(1000 * dzf));
  put_be16(bh + 20, (short)ns);
  put_be16(bh + 24, 1);  // IBM floats
  put_be16(bh + 26, 1);
  put_be16(bh + 28, 4);
  put_be16(bh + 54, 1);
  put_be16(bh + 302, 1);  // fixed trace length
  if (fwrite(head.data(), 1, head.size(), file.f) != head.size())
    return cg::fail(ctx, CG_EIO, std::string("write_segy: cannot write '") + path + "'");
  if (tr.empty()) return fflush(file.f) == 0 ? CG_OK : cg::fail(ctx, CG_EIO, "write_segy: flush");
  // samples on the device, traces in chunks of at most kChunk bytes
  const size_t per = (size_t)ns * 4;
  const size_t chunk = std::max<size_t>(1, std::min(tr.size(), kChunk / per));
  void *dtr = nullptr, *dout = nullptr;
  hipError_t e = cg::dev_alloc(ctx, &dtr, sizeof(SegyTrace) * chunk);
  if (e == hipSuccess) e = cg::dev_alloc(ctx, &dout, per * chunk);
  int rc = cg::hip_fail(ctx, e, "write_segy: staging buffers");
  Pinned host;
  if (rc == CG_OK) rc = cg::hip_fail(ctx, hipHostMalloc(&host.p, sizeof(SegyTrace) * chunk + per * chunk, hipHostMallocDefault),
                                     "write_segy: pinned buffer");
  std::vector<unsigned char> th(240);
  for (size_t c0 = 0; rc == CG_OK && c0 < tr.size(); c0 += chunk) {
    const size_t m = std::min(chunk, tr.size() - c0);
    SegyTrace *ht = static_cast<SegyTrace *>(host.p);
    unsigned char *hs = static_cast<unsigned char *>(host.p) + sizeof(SegyTrace) * chunk;
    for (size_t t = 0; t < m; ++t) ht[t] = tr[c0 + t].t;
    e = hipMemcpyAsync(dtr, ht, sizeof(SegyTrace) * m, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) {
      hipEvent_t ev;
      cg::launch_begin(ctx, CG_K_ELEMENTWISE, &ev);
      k_segy_traces<<<dim3((unsigned)((ns + cg::kThreads - 1) / cg::kThreads), (unsigned)m), cg::kThreads, 0,
                      ctx->stream>>>(g->d, g->rnxp, g->nyp, g->nz, static_cast<const SegyTrace *>(dtr), ns,
                                     (double)geo->z0, (double)dzf, static_cast<unsigned *>(dout));
      cg::launch_end(ctx, CG_K_ELEMENTWISE, ev);
      e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(hs, dout, per * m, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    rc = cg::hip_fail(ctx, e, "write_segy: samples");
    for (size_t t = 0; rc == CG_OK && t < m; ++t) {
      // trace header (TraceHeader::Write, SeisWorks positions, traceheader.cpp:527-586)
      const Out &o = tr[c0 + t];
      std::fill(th.begin(), th.end(), 0);
      put_be32(&th[8], o.il);
      put_be32(&th[20], o.xl);
      put_be32(&th[36], 0);  // offset
      put_be16(&th[70], 1);  // scalco
      put_be32(&th[72], o.ux);
      put_be32(&th[76], o.uy);
      put_be16(&th[108], (short)geo->z0);
      put_be16(&th[114], ns);
      put_be16(&th[116], (unsigned short)(dzf * 1000));
      if (fwrite(th.data(), 1, 240, file.f) != 240 || fwrite(hs + per * t, 1, per, file.f) != per)
        rc = cg::fail(ctx, CG_EIO, std::string("write_segy: cannot write '") + path + "'");
    }
  }
  if (dtr) cg::dev_free(ctx, dtr);
  if (dout) cg::dev_free(ctx, dout);
  if (rc == CG_OK && fflush(file.f) != 0) rc = cg::fail(ctx, CG_EIO, "write_segy: flush");
  return rc;
}


// NRLib SegY reading (libs/nrlib/segy: the binary header's interval, sample count and format;
// each trace's header fields and samples), for SegY::GetNearestTrace's callers: the traces in
// file order, IBM (format 1) or IEEE (format 5) big-endian samples converted to floats (an IBM
// float's 24-bit fraction fits a float's significand: the conversion is exact inside float's
// range).  Two calls: the first (null buffers) returns the counts, the second fills what the
// caller asks for.
static int segy_field(const unsigned char *h, int pos, int bytes)
{
  const unsigned char *p = h + pos - 1;
  return bytes == 2 ? (int)(short)(((unsigned)p[0] << 8) | p[1])
                    : (int)(((unsigned)p[0] << 24) | ((unsigned)p[1] << 16) | ((unsigned)p[2] << 8) | p[3]);
}

static float ibm_to_ieee(uint32_t w)
{
  const uint32_t frac = w & 0x00ffffffu;
  if (frac == 0) return (w >> 31) ? -0.0f : 0.0f;
  const int e16 = (int)((w >> 24) & 0x7f) - 64;  // value = 0.frac (base 16) * 16^e16
  const double v = std::ldexp((double)frac, 4 * e16 - 24);
  return (float)((w >> 31) ? -v : v);
}

int cg_segy_read(const char *path, const cg_segy_format *fmt, cg_segy_traces *out)
{
  if (!path || !out) return CG_EINVAL;
  const int il_pos = fmt && fmt->il_pos ? fmt->il_pos : 9, xl_pos = fmt && fmt->xl_pos ? fmt->xl_pos : 21;
  const int x_pos = fmt && fmt->x_pos ? fmt->x_pos : 73, y_pos = fmt && fmt->y_pos ? fmt->y_pos : 77;
  const int sc_pos = fmt && fmt->scalco_pos ? fmt->scalco_pos : 71;
  const int t0_pos = fmt && fmt->start_time_pos ? fmt->start_time_pos : 109;
  for (int pos : {il_pos, xl_pos, x_pos, y_pos})
    if (pos < 1 || pos > 237) return CG_EINVAL;
  for (int pos : {sc_pos, t0_pos})
    if (pos < 1 || pos > 239) return CG_EINVAL;
  FILE *f = std::fopen(path, "rb");
  if (!f) return CG_EIO;
  struct Closer {
    FILE *f;
    ~Closer() { std::fclose(f); }
  } closer{f};
  unsigned char bh[400];
  if (std::fseek(f, 3200, SEEK_SET) != 0 || std::fread(bh, 1, 400, f) != 400) return CG_EIO;
  const int dt_us = segy_field(bh, 17, 2), ns = segy_field(bh, 21, 2), code = segy_field(bh, 25, 2);
  if (ns <= 0 || dt_us <= 0 || (code != 1 && code != 5)) return CG_EINVAL;
  if (std::fseek(f, 0, SEEK_END) != 0) return CG_EIO;
  const long size = std::ftell(f);
  const long tb = 240L + 4L * ns;
  if (size < 3600 || (size - 3600) % tb != 0) return CG_EINVAL;
  const long ntr = (size - 3600) / tb;
  if (ntr > INT32_MAX) return CG_EINVAL;
  const bool fill = out->samples || out->il || out->xl || out->x || out->y || out->start_time;
  if (fill && (out->ntraces != (int)ntr || out->nsamples != ns)) return CG_EINVAL;
  out->ntraces = (int)ntr;
  out->nsamples = ns;
  out->dz = (float)dt_us / 1000.0f;
  out->format_code = code;
  if (!fill) return CG_OK;
  if (std::fseek(f, 3600, SEEK_SET) != 0) return CG_EIO;
  std::vector<unsigned char> buf((size_t)tb);
  for (long t = 0; t < ntr; ++t) {
    if (std::fread(buf.data(), 1, (size_t)tb, f) != (size_t)tb) return CG_EIO;
    const unsigned char *h = buf.data();
    if (out->il) out->il[t] = segy_field(h, il_pos, 4);
    if (out->xl) out->xl[t] = segy_field(h, xl_pos, 4);
    // TraceHeader::Read (traceheader.cpp:499-518): scalco -10000..-10 and 10..10000 in powers of
    // ten, as a float factor; any other value leaves the factor 1 (GetUtmx: factor * utmx)
    const int sc = segy_field(h, sc_pos, 2);
    float factor = 1.0f;
    switch (sc) {
      case -10000: factor = 0.0001f; break;
      case -1000: factor = 0.001f; break;
      case -100: factor = 0.01f; break;
      case -10: factor = 0.1f; break;
      case 10: case 100: case 1000: case 10000: factor = (float)sc; break;
      default: break;
    }
    const double scale = factor;
    if (out->x) out->x[t] = segy_field(h, x_pos, 4) * scale;
    if (out->y) out->y[t] = segy_field(h, y_pos, 4) * scale;
    if (out->start_time) out->start_time[t] = (float)segy_field(h, t0_pos, 2);
    if (out->samples) {
      float *dst = out->samples + t * (long)ns;
      for (int k = 0; k < ns; ++k) {
        const unsigned char *q = h + 240 + 4 * k;
        const uint32_t w = ((uint32_t)q[0] << 24) | ((uint32_t)q[1] << 16) | ((uint32_t)q[2] << 8) | q[3];
        if (code == 1) {
          dst[k] = ibm_to_ieee(w);
        } else {
          float v;
          std::memcpy(&v, &w, 4);
          dst[k] = v;
        }
      }
    }
  }
  return CG_OK;
}

}  // extern "C"
