// cg_mem.h -- explicitly global (address space 1) loads and stores and constant (address space
// 4) views for device code whose pointers come from descriptors in memory.  The compiler treats
// a pointer loaded from memory as generic: its accesses become flat operations (which also
// count against the LDS counter that barriers wait for), and wavefront-uniform reads of
// read-only tables become per-lane vector loads instead of scalar ones.  (float2's member
// operators take generic pointers, hence the plain vector type for it.)
#pragma once

#include <hip/hip_runtime.h>

namespace cg {

typedef float fvec2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ float2 gld(const float2 *p)
{
  const fvec2 v = *(const __attribute__((address_space(1))) fvec2 *)p;
  return make_float2(v.x, v.y);
}
__device__ __forceinline__ void gst(float2 *p, float2 v)
{
  fvec2 w;
  w.x = v.x;
  w.y = v.y;
  *(__attribute__((address_space(1))) fvec2 *)p = w;
}
template <typename T>
__device__ __forceinline__ T gld(const T *p)
{
  return *(const __attribute__((address_space(1))) T *)p;
}
template <typename T>
__device__ __forceinline__ void gst(T *p, T v)
{
  *(__attribute__((address_space(1))) T *)p = v;
}
// Element i of a read-only array as a constant-address-space reference (scalar loads where the
// index is wavefront-uniform).
template <typename T>
__device__ __forceinline__ const __attribute__((address_space(4))) T &kat(const T *p, long i)
{
  return ((const __attribute__((address_space(4))) T *)p)[i];
}

}  // namespace cg
