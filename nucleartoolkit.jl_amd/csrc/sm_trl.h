// Solver entry shared between the C ABI and the TRL translation unit.
#pragma once
#include "../../include/smgpu.h"
#include "sm_device.h"

namespace smgpu {

using SolveParams = smgpu_solve_params;
using SolveStats = smgpu_solve_stats;

void trl_solve(DeviceHam& h, const SolveParams& p, double* evals, double* evecs, SolveStats& st);

}  // namespace smgpu
