// Host-side check, run before every stage-kernel launch, that each device array
// the selected kernel variant reads or writes has been set.  A kernel handed a
// null array would fault the device on its first access; the check turns that
// into an error return instead.
//
// Plain C++ (no HIP): the CPU tests compile it on its own.
#pragma once

namespace hifr {

// The device arrays of one stage launch (null = not set).
struct LaunchPtrs {
  const void *u, *res, *div, *uf_in, *uf_out, *fv_in, *fv_out;
  const void *detj, *jg, *nbr, *fgeo;
  const void *tI, *tG, *tD;                                 // tet operator tiles
  const void *mlp_w1, *mlp_w2, *mlp_w3, *mlp_w4, *mlp_p;   // learned-closure weights
  const void* nut;                                          // learned closure: eddy viscosity per point
};

// The variant: mode (0 traces of u, 1 gradient traces, 2 residual, 3 advance, as
// hifr::MODE_*), viscous flux, LES model (0 none, 1 Smagorinsky, 2 WALE,
// 3 learned), tetrahedra (dense MFMA kernels) or tensor-product elements.
struct LaunchKind {
  int mode;
  bool viscous;
  int les;
  bool tet;
};

// Name of the first array the variant needs that is null, or nullptr when all are
// set.  Optional outputs (the gradient traces of mode 1, div of mode 3) are not
// required.
inline const char* missing_input(const LaunchPtrs& p, const LaunchKind& k) {
  struct Need {
    const void* ptr;
    const char* name;
    bool needed;
  };
  const bool flux = k.mode == 2 || k.mode == 3;
  const bool grad = k.mode == 1;
  // the gradient-trace kernel evaluates the network and stores the eddy viscosity,
  // the flux kernels read it
  const bool learned = k.les == 3 && k.mode != 0;
  const bool net = k.les == 3 && k.mode == 1;
  const Need needs[] = {
      {p.u, "state u", true},
      {p.uf_out, "output state traces", k.mode == 0 || k.mode == 3},
      {p.uf_in, "input state traces", k.mode != 0},
      {p.nbr, "neighbour table", k.mode != 0},
      {p.fgeo, "face geometry", k.mode != 0},
      {p.detj, "detJ", k.mode != 0},
      {p.jg, "JGinv", k.mode != 0},
      {p.fv_out, "output viscous flux traces", grad},
      {p.fv_in, "input viscous flux traces", flux && k.viscous},
      {p.div, "divergence output", k.mode == 2},
      {p.res, "RK register", k.mode == 3},
      {p.tI, "tet interpolation tiles", k.tet},
      {p.tG, "tet gradient tiles", k.tet && k.mode != 0},
      {p.tD, "tet divergence tiles", k.tet && k.mode != 0},
      {p.mlp_w1, "network W1", net},
      {p.mlp_w2, "network W2", net},
      {p.mlp_w3, "network W3", net},
      {p.mlp_w4, "network w4", net},
      {p.mlp_p, "network parameters", net},
      {p.nut, "eddy viscosity per point", learned},
  };
  for (const Need& n : needs)
    if (n.needed && n.ptr == nullptr) return n.name;
  return nullptr;
}

}  // namespace hifr
