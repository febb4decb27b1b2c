// solver_gpu.h -- a CrysPSSolverBase (solver_base.h:8-51) backed by
// libcrys_solver, for the reference program.
//
// A maintainer puts this header next to solver_p1.h, includes it from
// crysSCFT_sample.cpp and links -lcrys_solver.  The reference's driver keeps
// its per-point loops on host memory, so each diffusion() stages the host
// iomat through the device: call for call the reference's behaviour, PCIe
// bound.  The device-resident route is crys_scft_create / crys_scft_run.
#pragma once

#include <stdexcept>
#include <vector>

#include "crys_solver.h"
#include "solver_base.h"

class CrysPSSolverGPU : public CrysPSSolverBase {
 public:
  // kind: CRYS_P1, CRYS_3M or CRYS_PMMM_DCT (the three constructors,
  // solver_p1.h:16, solver_3m.h:17, solver_Pmmm.h:16).
  CrysPSSolverGPU(int kind, const c_array<size_t, 3>& grid,
                  const c_array<double, 9> g = {0, 0, 0, 0, 0, 0, 0, 0, 0},
                  int precision = CRYS_FP64)
      : CrysPSSolverBase(grid) {
    const size_t gl[3] = {grid[0], grid[1], grid[2]};
    check(crys_solver_create(kind, gl, g.begin(), precision, 0, &h_));
    size_t phy[3];
    check(crys_get_grid_size_phy(h_, phy));
    grid_size_phy_ = {phy[0], phy[1], phy[2]};
    check(crys_get_space_size_phy(h_, &space_size_phy_));
    host_.assign(space_size_phy_, 0.0);
    iomat_for_diffusion_ = host_.data();  // the driver reads and writes here
  }
  ~CrysPSSolverGPU() override { crys_solver_destroy(h_); }
  CrysPSSolverGPU(const CrysPSSolverGPU&) = delete;
  CrysPSSolverGPU& operator=(const CrysPSSolverGPU&) = delete;

  void set_cell_para(const c_array<double, 6>& c) override {
    check(crys_set_cell_para(h_, c.begin()));
    double out[6];
    check(crys_get_cell_para(h_, out));
    cell_para_ = {out[0], out[1], out[2], out[3], out[4], out[5]};
    check(crys_get_ds(h_, &ds_current_));
  }
  void set_contour_step(double ds) override {
    check(crys_set_contour_step(h_, ds));
    ds_current_ = ds;
  }
  void diffusion() const override {
    check(crys_iomat_upload(h_, iomat_for_diffusion_));
    check(crys_diffusion(h_));
    check(crys_iomat_download(h_, iomat_for_diffusion_));
  }

  // the library handle, for the device-resident entry points
  crys_solver_t* handle() const { return h_; }

 private:
  static void check(int st) {
    if (st != CRYS_OK) throw std::runtime_error(crys_last_error());
  }
  crys_solver_t* h_ = nullptr;
  std::vector<double> host_;
};
