// Solver entry shared between the C ABI and the TRL translation unit.
#pragma once
#include "../../include/smgpu.h"
#include "sm_device.h"

namespace smgpu {

using SolveParams = smgpu_solve_params;
using SolveStats = smgpu_solve_stats;

// Execution context of a solve: the stream it runs on and its (grow-only) Lanczos
// workspaces, one slot per precision.  A serving worker keeps one for its lifetime.
struct SolveCtx {
    hipStream_t stream = nullptr;
    mutable std::shared_ptr<void> ws[2];   // filled on first use
};

void trl_solve(DeviceHam& h, const SolveParams& p, double* evals, double* evecs, SolveStats& st,
               const double* v0 = nullptr, const SolveCtx* ctx = nullptr);
void restart_small_device(int device, const double* alpha, const double* beta, int m, int ls, double* S,
                          double* alpha_out, double* beta_out);
void lanczos_steps(DeviceHam& h, const SolveParams& p, int nsteps, const double* v0, double* alpha, double* beta);

}  // namespace smgpu
