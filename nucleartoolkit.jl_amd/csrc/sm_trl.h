// Solver entry shared between the C ABI and the TRL translation unit.
#pragma once
#include "../../include/smgpu.h"
#include "sm_device.h"

#include <atomic>
#include <functional>
#include <stdexcept>
#include <string>

namespace smgpu {

using SolveParams = smgpu_solve_params;
using SolveStats = smgpu_solve_stats;

// Execution context of a solve: the stream it runs on and its (grow-only) Lanczos
// workspaces, one slot per precision.  A serving worker keeps one for its lifetime.
struct SolveCtx {
    hipStream_t stream = nullptr;
    mutable std::shared_ptr<void> ws[2];   // filled on first use
    const std::atomic<bool>* cancel = nullptr;   // polled once per Lanczos iteration
};

// thrown by a solve whose cancel flag was raised
struct Canceled : std::runtime_error {
    Canceled() : std::runtime_error("canceled") {}
};

// HBM a solve allocates in its workspace (Krylov basis, work vectors, small problems)
int64_t solve_workspace_bytes(int64_t mdim, const SolveParams& p);
int64_t workspace_capacity(const SolveCtx& ctx);
// bytes by which ctx's workspace grows to run such a solve (buffers are grow-only)
int64_t workspace_growth(const SolveCtx& ctx, int64_t mdim, const SolveParams& p);
void release_workspace(SolveCtx& ctx);

// J projection of the Lanczos vectors (TRL's doubleLanczos, lanczos_methods.jl:1-91, 114-151):
// j2 applies J^2 on the same basis, which has 2M = twoJ; twoJmax is the largest 2J of
// the space.  A vector whose <J^2> differs from J(J+1) by more than tol is replaced
// by its projection prod_{J'} (J^2 - J'(J'+1)) / (J(J+1) - J'(J'+1)), J' = J+1..Jmax,
// re-orthogonalised against the basis and normalised.
struct JProj {
    const DeviceHam* j2 = nullptr;
    int twoJ = 0, twoJmax = 0;
    double tol = 1e-7;
};

void trl_solve(DeviceHam& h, const SolveParams& p, double* evals, double* evecs, SolveStats& st,
               const double* v0 = nullptr, const SolveCtx* ctx = nullptr, const JProj* jp = nullptr);
// Lockstep TRL solves on one Hamiltonian with shared H·v launches on hv_stream.
// Per member slot b < cap (ctx[b]: its own stream and workspace): params p[b],
// start vector v0[b] (v0 or the entry may be null), evals[b], evecs[b] (evecs or
// the entry may be null), stats st[b], status[b] (SMGPU_* code) and msgs[b] (error
// text; msgs may be null).  Slots [0, nb) are filled at the call; a member's
// failure does not stop the others.  on_done(b) runs when slot b's member has
// finished (results written) or failed; afterwards, and between rounds, join(b)
// may fill io at a free slot b with a new member (same precision) and return true.
// hams (optional): member b solves *hams[b], a Hamiltonian on h's index tables
// (interaction variants on one basis; null: every member solves h); the shared
// launches then give each vector its member's coefficients.
struct BatchIO {
    DeviceHam* const* hams = nullptr;
    SolveParams* p = nullptr;
    const double* const* v0 = nullptr;
    double* const* evals = nullptr;
    double* const* evecs = nullptr;
    SolveStats* st = nullptr;
    int32_t* status = nullptr;
    std::string* msgs = nullptr;
};
void trl_solve_batch(DeviceHam& h, int cap, int nb, const BatchIO& io, const SolveCtx* const* ctx,
                     hipStream_t hv_stream, const std::function<bool(int)>& join = {},
                     const std::function<void(int)>& on_done = {});
// thick-restart block Lanczos with block size q (TRBL); v0: q start vectors (q*mdim) or null
void trbl_solve(DeviceHam& h, int q, const SolveParams& p, double* evals, double* evecs, SolveStats& st,
                const double* v0 = nullptr, const JProj* jp = nullptr);
void restart_small_device(int device, const double* alpha, const double* beta, int m, int ls, double* S,
                          double* alpha_out, double* beta_out);
void lanczos_steps(DeviceHam& h, const SolveParams& p, int nsteps, const double* v0, double* alpha, double* beta);
// lanczos_steps with every device buffer in the library's static scratch array (no
// hipMalloc / hipFree / host-pinned allocation; one stream, the handle's).  Calls on
// one device are serialized by a host mutex; sizes whose buffers exceed
// lanczos_scratch_max() bytes are rejected with std::invalid_argument.
void lanczos_steps_noalloc(DeviceHam& h, const SolveParams& p, int nsteps, const double* v0, double* alpha,
                           double* beta);
int64_t lanczos_scratch_max();

}  // namespace smgpu
