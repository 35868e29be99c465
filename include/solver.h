// solver.h - GaussNewtonSolver with the reference's interface
// (include/solver.h:10-89), backed by nrf_solver (deterministic assembly,
// persistent FP64-dot block-Jacobi PCG on gfx950).
#pragma once
#include "types.h"

class GaussNewtonSolver {
 public:
  struct Params {
    int gn_iterations = 3;
    int pcg_iterations = 10;
    float pcg_tolerance = 1e-4f;
    float lambda_smooth = 1.0f;
    float lambda_rot = 1.0f;
    float lambda_damping = 1e-2f;
    bool debug = false;
  };

  explicit GaussNewtonSolver(const Params &p, int max_nodes);
  ~GaussNewtonSolver();
  GaussNewtonSolver(const GaussNewtonSolver &) = delete;
  GaussNewtonSolver &operator=(const GaussNewtonSolver &) = delete;

  // kernels/solver_kernels.cu:543-570: one linearisation with the stored
  // corrs[i].src; delta_x is resized to 6 * num_nodes and overwritten.
  void solve(const DeviceArray<Correspondence> &corrs, int num_corrs, const DeformNode *d_nodes,
             const DualQuat *d_transforms, int num_nodes, DeviceArray<float> &delta_x);

  // Copies of the assembled system of the last solve (refreshed on access).
  const BSRMatrix &system_matrix() const;
  const DeviceArray<float> &rhs() const;
  // PCG outcome of the last solve (the reference discards it).
  const nrf_pcg_result &last_result() const { return result_; }

 private:
  Params params_;
  int max_nodes_;
  nrf_solver *solver_ = nullptr;
  nrf_pcg_result result_{};
  int last_num_nodes_ = 0;
  mutable BSRMatrix A_;
  mutable DeviceArray<float> b_;
};
