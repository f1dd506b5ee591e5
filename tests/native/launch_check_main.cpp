// CPU test of the pre-launch check of the stage kernels (fr_launch_check.hpp):
// prints one line per case and exits non-zero on the first wrong answer.
#include <cstdio>
#include <cstring>

#include "fr_launch_check.hpp"

using hifr::LaunchKind;
using hifr::LaunchPtrs;
using hifr::missing_input;

static int fails = 0;

static void expect(const char* what, const char* got, const char* want) {
  const bool ok = (got == nullptr && want == nullptr) || (got && want && std::strcmp(got, want) == 0);
  std::printf("%-60s -> %-28s %s\n", what, got ? got : "(all set)", ok ? "ok" : "WRONG");
  if (!ok) fails++;
}

int main() {
  static char mem[64];
  const void* x = mem;
  LaunchPtrs all{x, x, x, x, x, x, x, x, x, x, x, x, x, x, x, x, x, x, x, x};
  // every variant runs with everything set
  for (int mode = 0; mode < 4; mode++)
    for (int visc = 0; visc < 2; visc++)
      for (int les = 0; les < 4; les++)
        for (int tet = 0; tet < 2; tet++) {
          if (les && !visc) continue;
          char name[96];
          std::snprintf(name, sizeof name, "all set: mode %d viscous %d les %d tet %d", mode, visc, les, tet);
          expect(name, missing_input(all, LaunchKind{mode, visc != 0, les, tet != 0}), nullptr);
        }
  // the learned closure with no network weights (a context that had never been
  // given a network kept null weight pointers): the gradient-trace kernel, which
  // evaluates the network, is refused; the flux kernels read only the stored eddy
  // viscosity, and the state traces of mode 0 never read the network
  LaunchPtrs nonet = all;
  nonet.mlp_w1 = nonet.mlp_w2 = nonet.mlp_w3 = nonet.mlp_w4 = nonet.mlp_p = nullptr;
  expect("learned LES, no weights, gradient traces", missing_input(nonet, LaunchKind{1, true, 3, false}),
         "network W1");
  expect("learned LES, no weights, residual", missing_input(nonet, LaunchKind{2, true, 3, false}), nullptr);
  expect("learned LES, no weights, advance", missing_input(nonet, LaunchKind{3, true, 3, false}), nullptr);
  expect("learned LES, no weights, traces of u", missing_input(nonet, LaunchKind{0, true, 3, false}), nullptr);
  expect("WALE, no network weights, advance", missing_input(nonet, LaunchKind{3, true, 2, false}), nullptr);
  LaunchPtrs nob = all;
  nob.mlp_p = nullptr;
  expect("learned LES, no parameter vector, gradient traces", missing_input(nob, LaunchKind{1, true, 3, false}),
         "network parameters");
  // the stored eddy viscosity: written by the gradient traces, read by the flux passes
  LaunchPtrs nonu = all;
  nonu.nut = nullptr;
  expect("learned LES, no eddy-viscosity array, gradient traces",
         missing_input(nonu, LaunchKind{1, true, 3, false}), "eddy viscosity per point");
  expect("learned LES, no eddy-viscosity array, residual", missing_input(nonu, LaunchKind{2, true, 3, false}),
         "eddy viscosity per point");
  expect("WALE, no eddy-viscosity array, advance", missing_input(nonu, LaunchKind{3, true, 2, false}), nullptr);
  // arrays of one mode only
  LaunchPtrs p = all;
  p.res = nullptr;
  expect("advance without RK register", missing_input(p, LaunchKind{3, true, 0, false}), "RK register");
  expect("residual without RK register", missing_input(p, LaunchKind{2, true, 0, false}), nullptr);
  p = all;
  p.div = nullptr;
  expect("residual without divergence output", missing_input(p, LaunchKind{2, false, 0, false}),
         "divergence output");
  expect("advance without divergence output (optional)", missing_input(p, LaunchKind{3, false, 0, false}), nullptr);
  p = all;
  p.fv_out = nullptr;
  expect("gradient traces without their output", missing_input(p, LaunchKind{1, true, 0, false}),
         "output viscous flux traces");
  p = all;
  p.fv_in = nullptr;
  expect("viscous advance without viscous traces", missing_input(p, LaunchKind{3, true, 0, false}),
         "input viscous flux traces");
  expect("Euler advance without viscous traces", missing_input(p, LaunchKind{3, false, 0, false}), nullptr);
  p = all;
  p.uf_in = p.nbr = p.fgeo = p.detj = p.jg = nullptr;
  expect("traces of u without metrics or faces", missing_input(p, LaunchKind{0, false, 0, false}), nullptr);
  expect("residual without input traces", missing_input(p, LaunchKind{2, false, 0, false}), "input state traces");
  p = all;
  p.uf_out = nullptr;
  expect("traces of u without their output", missing_input(p, LaunchKind{0, false, 0, false}),
         "output state traces");
  p = all;
  p.tD = nullptr;
  expect("tet advance without divergence tiles", missing_input(p, LaunchKind{3, true, 0, true}),
         "tet divergence tiles");
  expect("tet traces of u without divergence tiles", missing_input(p, LaunchKind{0, true, 0, true}), nullptr);
  expect("hex advance without tet tiles", missing_input(p, LaunchKind{3, true, 0, false}), nullptr);
  p = all;
  p.u = nullptr;
  expect("any mode without the state", missing_input(p, LaunchKind{0, false, 0, false}), "state u");
  std::printf("%d wrong\n", fails);
  return fails ? 1 : 0;
}
