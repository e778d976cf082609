// Interface of the adjoint kernels (solar_adjoint.hip).
#pragma once
#include <hip/hip_runtime_api.h>

#include "solar_kernels.hpp"

namespace solar {

constexpr int kAdjRhsOnly = 1;  // Yout = L^T W (no Yin, no W update, dt = 1)
constexpr int kAdjWriteW = 2;   // write Wout = alpha*W + beta*Yout
constexpr int kAdjGrad = 4;     // accumulate the c^2 gradient

struct AdjStageArgs {
  const double* W = nullptr;  // nullptr: W = wscale * Yin (Yout must not alias Yin)
  const double* Yin = nullptr;
  double* Yout = nullptr;
  double* Wout = nullptr;
  const double* divs = nullptr;
  double* gc2 = nullptr;
  const double* c2f = nullptr;
  double *gxa_p = nullptr, *gya_p = nullptr, *gxa_v = nullptr, *gya_v = nullptr;
  double dt = 1.0, alpha = 0.0, beta = 0.0, wscale = 1.0;
  int mode = 0;
  int div32 = 0;  // divs holds float values
  int chunks = 0;  // radial chunks per tile column (0: automatic)
};

hipError_t launch_adj_stage(const DevTables& T, const AdjStageArgs& a, hipStream_t st);
// Transposed panel gather.  vec: receiver-basis vector ghosts (fields 1, 2)
// plus the scalar field sfield; else nfields scalar fields from sfield.
// radial_filter: the ghost adjoints are per-plane values still to be taken
// through FILTER_R^T (the fused filter adjoint leaves them so).
hipError_t launch_adj_scatter(const DevTables& T, bool vec, const double* gxs, const double* gys, const double* gxv,
                              const double* gyv, double* Y, double* Wn, double beta, int sfield, int nfields,
                              hipStream_t st, bool radial_filter = false);
// out = FILTER_EXP^T(in) without the ghost contributions, which go to
// fgxa/fgya per plane (to be scattered with radial_filter = true).
// chunks: radial chunks per tile column (0: automatic).
hipError_t launch_adj_filter(const DevTables& T, const double* in, double* out, double* fgxa, double* fgya,
                             hipStream_t st, int chunks = 0);
hipError_t launch_sample(const double* q, const long long* idx, int nrec, double* out, hipStream_t st);
hipError_t launch_inject(double* lam, const long long* idx, int nrec, const double* seis, const double* data,
                         hipStream_t st);

}  // namespace solar
