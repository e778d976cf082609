// Device helpers shared by the forward and adjoint kernels.
#pragma once
#include <hip/hip_runtime.h>

#include "solar_tables.hpp"

namespace solar {

// The per-plane radial rows seen through the constant address space: a
// kernel indexes them with a block-uniform plane, so the compiler reads them
// with scalar loads (the tables are never written while kernels run).
using CRadialRows = const __attribute__((address_space(4))) RadialRow*;
__device__ __forceinline__ CRadialRows const_rows(const RadialRow* p) { return (CRadialRows)p; }

// A pointer read from memory (an ensemble member's buffers) carries no
// address space, so its accesses would compile to flat instructions, which
// count against the scalar-load wait counter too and so tie the scalar loads
// to every vector load in flight.  Kernels hold such buffers as pointers into
// the global address space instead.
template <typename T>
using global_ptr = __attribute__((address_space(1))) T*;
template <typename T>
__device__ __forceinline__ global_ptr<T> to_global(T* p) {
  return (global_ptr<T>)p;
}

// Block index remap so that consecutive logical blocks (neighbouring tiles,
// whose halos overlap) run on the same XCD and share its L2.  Hardware
// dispatches block b to XCD b % 8; XCD x receives nb / 8 blocks, one more
// when x < nb % 8, and is given the contiguous logical range that starts at
// x * (nb / 8) + min(x, nb % 8).  A bijection on [0, nb) for every nb.
__device__ __forceinline__ int xcd_remap(int b, int nb) {
  const int q = nb >> 3, r = nb & 7, x = b & 7;
  return x * q + min(x, r) + (b >> 3);
}

}  // namespace solar
