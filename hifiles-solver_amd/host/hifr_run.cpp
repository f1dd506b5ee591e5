// hifr_run <input_file>: the solver executable (src/HiFiLES.cpp:49-266 main loop):
// reads the deck and mesh, runs n_steps with the deck's time integrator on the GPU
// and prints the residual norms every monitor_res_freq steps in the reference's
// column layout; writes restart files every restart_dump_freq steps and the
// convergence history history.plt (HistoryOutput, src/output.cpp:1525-1616) with the
// deck's integral_quantities.  Forces are not computed: their columns hold 0 on
// meshes without wall faces (the reference's value there) and nan otherwise.
#include <chrono>
#include <cmath>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/hifr_solver.h"
#include "input.h"

int main(int argc, char** argv) {
  if (argc < 2) {
    std::printf("ERROR: No input file specified ... \n");
    return 0;
  }
  hifr_host::input in;
  const std::string e = in.read_input_file(argv[1]);
  if (!e.empty()) {
    std::fprintf(stderr, "Fatal error '%s'\n", e.c_str());
    return 1;
  }
  char err[512] = {0};
  hifr_solver* s = nullptr;
  if (hifr_solver_from_deck(argv[1], &s, err, sizeof err) != 0) {
    std::fprintf(stderr, "Fatal error '%s'\n", err);
    return 1;
  }
  int sz[6];
  hifr_solver_sizes(s, sz);
  const int nf = sz[2], nd = sz[3];
  std::printf("%d elements, %d solution points each, %d interfaces\n", sz[4], sz[0], sz[5]);
  if (nd == 2) std::printf("\n  Iter       Res[Rho]   Res[RhoVelx]   Res[RhoVely]      Res[RhoE]\n");
  else std::printf("\n  Iter       Res[Rho]   Res[RhoVelx]   Res[RhoVely]   Res[RhoVelz]      Res[RhoE]\n");
  const int freq = in.monitor_res_freq > 0 ? in.monitor_res_freq : in.n_steps + 1;
  const int nq = hifr_solver_n_integrals(s);
  const double force = hifr_solver_n_boundary_points(s) > 0 ? NAN : 0.0;
  FILE* hist = std::fopen("history.plt", "w");
  if (hist) {
    std::fprintf(hist, "TITLE = \"HiFiLES simulation\"\nVARIABLES = \"Iteration\"");
    const char* res2[] = {"<greek>r</greek>", "<greek>r</greek>v<sub>x</sub>", "<greek>r</greek>v<sub>y</sub>",
                          "<greek>r</greek>E"};
    const char* res3[] = {"<greek>r</greek>", "<greek>r</greek>v<sub>x</sub>", "<greek>r</greek>v<sub>y</sub>",
                          "<greek>r</greek>v<sub>z</sub>", "<greek>r</greek>E"};
    for (int k = 0; k < nf; k++) std::fprintf(hist, ",\"log<sub>10</sub>(Res[%s])\"", nd == 2 ? res2[k] : res3[k]);
    const char* fx[] = {"x", "y", "z"};
    for (int d = 0; d < nd; d++) std::fprintf(hist, ",\"F<sub>%s</sub>(Total)\"", fx[d]);
    std::fprintf(hist, ",\"CL</sub>(Total)\",\"CD</sub>(Total)\"");
    for (int q = 0; q < nq; q++) std::fprintf(hist, ",\"Diagnostics[%d]\"", q);
    std::fprintf(hist, ",\"Time<sub>Physical</sub>\",\"Time<sub>Comp</sub>(m)\"\nZONE T= \"Convergence history\"\n");
  }
  const auto t0 = std::chrono::steady_clock::now();
  for (int it = 1; it <= in.n_steps; it++) {
    if (hifr_solver_run(s, 1) != 0) {
      std::fprintf(stderr, "Fatal error: step %d failed\n", it);
      hifr_solver_destroy(s);
      return 1;
    }
    if (it == 1 || it % freq == 0) {
      double r[5];
      if (hifr_solver_norms(s, in.res_norm_type, r) != 0) {
        std::fprintf(stderr, "Fatal error 'NaN residual encountered. Exiting'\n");
        hifr_solver_destroy(s);
        return 1;
      }
      std::printf("%6d", in.restart_iter + it);
      for (int k = 0; k < nf; k++) std::printf("%15.8f", r[k]);
      std::printf("\n");
      std::vector<double> q(nq > 0 ? nq : 1);
      if (nq > 0 && hifr_solver_integrals(s, q.data()) != 0) {
        std::fprintf(stderr, "Fatal error: integral diagnostics failed\n");
        hifr_solver_destroy(s);
        return 1;
      }
      if (hist) {
        std::fprintf(hist, "%d", in.restart_iter + it);
        for (int k = 0; k < nf; k++) std::fprintf(hist, ", %.15g", std::log10(r[k]));
        for (int d = 0; d < nd + 2; d++) std::fprintf(hist, ", %.15g", force);
        for (int k = 0; k < nq; k++) std::fprintf(hist, ", %.15g", q[k]);
        const double mins = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count() / 60.0;
        std::fprintf(hist, ", %.15g, %.15g\n", hifr_solver_time(s), mins);
        std::fflush(hist);
      }
    }
    if (in.restart_dump_freq > 0 && it % in.restart_dump_freq == 0) {
      char name[64];
      std::snprintf(name, sizeof name, "Rest_%09d_p%04d.dat", in.restart_iter + it, 0);
      hifr_solver_write_restart(s, name);
    }
  }
  const double sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  std::printf("Execution time= %f s\n", sec);
  if (hist) std::fclose(hist);
  hifr_solver_destroy(s);
  return 0;
}
