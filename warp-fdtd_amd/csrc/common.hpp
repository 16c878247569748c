// Shared device helpers and C-ABI error plumbing for libwfdtd (gfx950).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>

#include "../../include/wfdtd.h"

namespace wfd {

// Thread-local last-error text returned by wfd_last_error().
void set_error(const char* fmt, ...);

// Argument-check helpers: every C entry returns an int status and never throws.
#define WFD_REQUIRE(cond, ...)                  \
  do {                                          \
    if (!(cond)) {                              \
      ::wfd::set_error(__VA_ARGS__);            \
      return WFD_ERR_ARGUMENT;                  \
    }                                           \
  } while (0)

#define WFD_LAUNCH_CHECK(what)                                              \
  do {                                                                      \
    hipError_t err_ = hipGetLastError();                                    \
    if (err_ != hipSuccess) {                                               \
      ::wfd::set_error("%s: %s", what, hipGetErrorString(err_));            \
      return WFD_ERR_HIP;                                                   \
    }                                                                       \
  } while (0)

#define WFD_HIP(call)                                                       \
  do {                                                                      \
    hipError_t err_ = (call);                                               \
    if (err_ != hipSuccess) {                                               \
      ::wfd::set_error("%s: %s", #call, hipGetErrorString(err_));           \
      return WFD_ERR_HIP;                                                   \
    }                                                                       \
  } while (0)

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

// Largest flat index a 32-bit kernel index may address.
constexpr int64_t kMaxCells = (int64_t(1) << 31) - 1;

inline int64_t cells(int nx, int ny, int nz) { return (int64_t)nx * ny * nz; }

struct F3 {
  float x, y, z;
};

typedef __amdgpu_buffer_rsrc_t Rsrc;

// A raw buffer over `bytes` bytes at p: loads past the end read 0 and stores
// past it are dropped.  Addressed by a 32-bit lane offset plus a uniform
// offset (an SGPR) instead of a 64-bit address per lane.
__device__ __forceinline__ Rsrc buffer(const float* p, unsigned bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<float*>(p), 0, (int)bytes, 0x00020000);
}
// (the b32 builtins move 32-bit integers: the floats travel as their bits)
__device__ __forceinline__ float ld(Rsrc r, int voff, int soff) {
  return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
}
__device__ __forceinline__ void st(Rsrc r, float v, int voff, int soff) {
  __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(v), r, voff, soff, 0);
}

// One float array of components of n cells each (E, H: 3; psi: 6; 1/eps: 1)
// through a buffer resource: every access of a cell shares its 32-bit byte
// offset, the component adds a uniform offset, so a lane holds one offset
// register per cell instead of a 64-bit address per load.  The array must
// span fewer than 2^31 bytes (checked on the host).
struct Field {
  Rsrc r;
  int cb;  // bytes per component
  __device__ __forceinline__ Field(const float* p, int comps, int n)
      : r(buffer(p, (unsigned)comps * (unsigned)n * 4u)), cb(n * 4) {}
  __device__ __forceinline__ float operator()(int c, int i) const { return ld(r, i * 4, c * cb); }
  __device__ __forceinline__ void put(int c, int i, float v) const { st(r, v, i * 4, c * cb); }
  // by byte offset: kAway reads 0 and drops the store, so a masked access
  // needs no branch
  __device__ __forceinline__ float at(int c, int voff) const { return ld(r, voff, c * cb); }
  __device__ __forceinline__ void set(int c, int voff, float v) const { st(r, v, voff, c * cb); }
};

// A byte offset past the end of every Field (hosts keep the arrays a kernel
// addresses this way below 0x7ff00000 bytes, so even with a component offset
// added it stays out of range without wrapping).
constexpr int64_t kFieldBytes = 0x7ff00000;
constexpr int kAway = 0x7ffffff0;

// A buffer's base address as a value the compiler cannot see through: a
// Field built from it inside a loop is rebuilt there (a few scalar moves)
// instead of being hoisted, so that many live buffer descriptors do not
// spill the scalar registers.
__device__ __forceinline__ const float* opaque(const float* p) {
  asm volatile("" : "+s"(p));
  return p;
}

// a where the mask is all ones, else b (bitwise, no lane-mask register)
__device__ __forceinline__ float pick(int m, float a, float b) {
  return __int_as_float((__float_as_int(a) & m) | (__float_as_int(b) & ~m));
}
__device__ __forceinline__ int pick(int m, int a, int b) { return (a & m) | (b & ~m); }

// Device copy of a wfd_stage footprint (box plus its cell count).
struct StageBox {
  float* buf;
  int x0, x1, y0, y1, z0, z1;
  int count;
};

inline int to_stage_box(const wfd_stage& s, int nx, int ny, int nz, StageBox& out) {
  const int* b = s.box;
  if (!s.buf || b[0] < 0 || b[2] < 0 || b[4] < 0 || b[1] > nx || b[3] > ny || b[5] > nz ||
      b[1] <= b[0] || b[3] <= b[2] || b[5] <= b[4])
    return WFD_ERR_ARGUMENT;
  out.buf = s.buf;
  out.x0 = b[0];
  out.x1 = b[1];
  out.y0 = b[2];
  out.y1 = b[3];
  out.z0 = b[4];
  out.z1 = b[5];
  out.count = (b[1] - b[0]) * (b[3] - b[2]) * (b[5] - b[4]);
  return WFD_OK;
}

// kernels.py:8-26 — zero-padded (or periodic) SoA fetch at (x*ny+y)*nz+z.
__device__ __forceinline__ F3 sample(const float* __restrict__ f, int n, int x, int y,
                                     int z, int nx, int ny, int nz, int periodic) {
  if (periodic) {
    x = (x + nx) % nx;
    y = (y + ny) % ny;
    z = (z + nz) % nz;
  }
  if (x < 0 || y < 0 || z < 0 || x >= nx || y >= ny || z >= nz) return F3{0.f, 0.f, 0.f};
  const int i = (x * ny + y) * nz + z;
  return F3{f[i], f[n + i], f[2 * n + i]};
}

// Exact integer division o / d for 0 <= o < 2^24 through a float reciprocal
// plus one correction step (the float quotient is within one of the truth).
__device__ __forceinline__ int div_small(int o, int d, float inv_d) {
  int q = __float2int_rz((float)o * inv_d);
  int r = o - q * d;
  if (r < 0) q -= 1;
  else if (r >= d) q += 1;
  return q;
}

// One CPML partial (kernels.py:29-47): row c = (a, b, 1/kappa); psi updated
// in place when a != 0.  Kept as separate multiplies and adds (no FMA) to
// round exactly as the reference arithmetic.
__device__ __forceinline__ float cpml_term(float d, float a, float b, float ik,
                                           float* __restrict__ psi_at) {
  if (a != 0.f) {
    const float p = b * (*psi_at) + a * d;
    *psi_at = p;
    return ik * d + p;
  }
  return ik * d;
}

}  // namespace wfd
