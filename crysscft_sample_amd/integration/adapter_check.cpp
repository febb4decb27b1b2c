// Drives CrysPSSolverGPU through a CrysPSSolverBase& exactly as the
// reference's driver does (crysSCFT_sample.cpp:373-396, :466-471): set the
// cell and the contour step, write the field into get_iomat_for_diffusion(),
// call diffusion() `reps` times, read the result back.
//
//   adapter_check KIND NX NY NZ L DS REPS IN OUT
//
// KIND is P1, 3m or PmmmDCT; IN and OUT hold get_space_size_phy() raw
// doubles.  Prints the getters; exit status 2 when the library reports an
// error (no device, bad grid), 1 on a usage or file error.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <vector>

#include "solver_gpu.h"

int main(int argc, char** argv) {
  if (argc != 10) {
    std::fprintf(stderr, "usage: %s KIND NX NY NZ L DS REPS IN OUT\n", argv[0]);
    return 1;
  }
  const int kind = std::strcmp(argv[1], "P1") == 0   ? CRYS_P1
                   : std::strcmp(argv[1], "3m") == 0 ? CRYS_3M
                                                     : CRYS_PMMM_DCT;
  const c_array<size_t, 3> grid = {(size_t)std::atol(argv[2]), (size_t)std::atol(argv[3]),
                                   (size_t)std::atol(argv[4])};
  const double L = std::atof(argv[5]), ds = std::atof(argv[6]);
  const int reps = std::atoi(argv[7]);
  try {
    CrysPSSolverGPU gpu(kind, grid);
    CrysPSSolverBase& solver = gpu;  // the driver only holds the base class
    solver.set_cell_para({L, L, L, 0, 0, 0});
    solver.set_contour_step(ds);
    const size_t n = solver.get_space_size_phy();
    std::vector<double> buf(n);
    FILE* f = std::fopen(argv[8], "rb");
    if (!f || std::fread(buf.data(), sizeof(double), n, f) != n) {
      std::fprintf(stderr, "cannot read %zu doubles from %s\n", n, argv[8]);
      if (f) std::fclose(f);
      return 1;
    }
    std::fclose(f);
    double* io = solver.get_iomat_for_diffusion();
    for (size_t i = 0; i < n; ++i) io[i] = buf[i];
    for (int r = 0; r < reps; ++r) solver.diffusion();
    const auto phy = solver.get_grid_size_phy();
    const auto cell = solver.get_cell_para();
    std::printf("grid_size_phy %zu %zu %zu space_size_phy %zu space_size_log %zu ds %.17g "
                "cell %.17g %.17g %.17g\n",
                phy[0], phy[1], phy[2], n, solver.get_space_size_log(), solver.get_ds(), cell[0],
                cell[1], cell[2]);
    f = std::fopen(argv[9], "wb");
    if (!f || std::fwrite(io, sizeof(double), n, f) != n) {
      std::fprintf(stderr, "cannot write %s\n", argv[9]);
      if (f) std::fclose(f);
      return 1;
    }
    std::fclose(f);
  } catch (const std::exception& e) {
    std::fprintf(stderr, "libcrys_solver: %s\n", e.what());
    return 2;
  }
  return 0;
}
