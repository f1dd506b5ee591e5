// Kernel dispatch of the element-block FR kernels: (dimension, order, mode,
// viscosity law, metric storage) -> one template instantiation.  The FP64 and
// the mixed-precision instantiations live in separate translation units
// (kernels_fp64.hip, kernels_mixed.hip) that include fr_launch_impl.hpp.
#pragma once
#include <hip/hip_runtime.h>

#include "fr_tensor_kernels.hpp"
#include "fr_tet_mfma.hpp"

namespace hifr {

// vis: 0 Euler, 1 Navier-Stokes with constant viscosity, 2 Sutherland viscosity;
// aff: metrics per element / per face; les: 0 none, 1 Smagorinsky, 2 WALE (vis 1),
// 3 learned eddy viscosity with WALE fallback (vis 1, hexes P = 2, 3)
hipError_t launch_fp64(int nd, int P, int mode, int vis, bool aff, int les, const StageArgsT<double>& a,
                       hipStream_t s);
hipError_t launch_mixed(int nd, int P, int mode, int vis, bool aff, int les, const StageArgsT<float>& a,
                        hipStream_t s);
// tetrahedra, P = 1..3 (dense operators on MFMA)
hipError_t launch_tet_fp64(int P, int mode, int vis, bool aff, int les, const StageArgsT<double>& a,
                           hipStream_t s);
hipError_t launch_tet_mixed(int P, int mode, int vis, bool aff, int les, const StageArgsT<float>& a,
                            hipStream_t s);

// Roe interface flux (riemann_solve_type 2): FP64, no LES, flux modes (RESID,
// ADVANCE) only; quads/hexes and tets
hipError_t launch_roe_fp64(int nd, int P, int mode, int vis, bool aff, const StageArgsT<double>& a,
                           hipStream_t s);
hipError_t launch_tet_roe_fp64(int P, int mode, int vis, bool aff, const StageArgsT<double>& a, hipStream_t s);

}  // namespace hifr
