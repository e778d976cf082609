// Device helpers shared by the forward and adjoint kernels.
#pragma once
#include <hip/hip_runtime.h>

namespace solar {

// Block index remap so that consecutive logical blocks (neighbouring tiles,
// whose halos overlap) run on the same XCD and share its L2.  Hardware
// dispatches block b to XCD b % 8.
__device__ __forceinline__ int xcd_remap(int b, int nb) {
  if (nb % 8 != 0) return b;
  return (b % 8) * (nb / 8) + b / 8;
}

}  // namespace solar
