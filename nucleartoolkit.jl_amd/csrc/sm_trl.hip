// Thick-restart Lanczos on the device.
//
// Reference algorithm: TRL (src/ShellModel/lanczos_methods.jl:1-91) with
// diagonalize_T! (:250-258), ReORTH (src/ShellModel/shellmodel_main.jl:1005-1012),
// ThickRestart (:260-307) and the Ritz-vector assembly of main_sm (:208-224).
//
// Per iteration (it = index of the current basis vector v = V[it]):
//   w  = H v                                  (FP32 or FP64 H·v kernel)
//   classical Gram-Schmidt twice against V[0..it] (FP64 sums): h1 = V^T w,
//   w1 = w - V h1, h2 = V^T w1, w = w1 - V h2, in three streaming passes over
//   the basis (the middle pass computes w1 and the h2 partials from one read
//   of an LDS-staged slab).  One pass is not enough: it multiplies existing
//   non-orthogonality by ~|h|/beta every step;
//   alpha[it] = h[it] (= v.Hv, the reference's dot before ReORTH),
//   beta[it] = |w|,  V[it+1] = w / beta
//   Ritz values of the tridiagonal T[0..it] (multisection bisection in FP64,
//   bracketed by the previous step's values through Cauchy interlacing) and the
//   reference's convergence rule all(en_prev - en_new < tol).
// The host enqueues iteration it+1 before it reads the flag of iteration it,
// so the device never waits for the host; a speculative iteration after
// convergence only writes entries that are never read.
// Restart keeps T tridiagonal: the kept Ritz vectors are rotated within their
// span (tri_restart) so the projected matrix stays tridiagonal; its eigenvalues
// equal those of the reference's arrowhead Tmat.
// Precision 0: basis, H·v and all vectors FP64 (the plain FP64 path).
// Precision 1: FP32 H·v and FP32 basis storage; dot products, alpha/beta, T,
// the small eigen-problems and the restart rotation coefficients are FP64;
// the final Ritz vectors are assembled in FP64 and their eigenvalues refined by
// an FP64 Rayleigh quotient (one FP64 H·v per eigenpair).
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <functional>
#include <memory>
#include <random>
#include <stdexcept>
#include <string>
#include <thread>

#include <mutex>

#include "dense_eig.h"
#include "sm_device.h"
#include "sm_status.h"
#include "sm_trl.h"
#include "tridiag.h"

// Device scratch of lanczos_steps_noalloc: the Krylov basis, work vectors, partial
// sums, flags and the small problems of one call, carved from this array.  Each
// device the library runs on holds its own copy (LANCZOS_SCRATCH_BYTES) until the
// process ends.  Host code never uses the array's name as an address: it uses the
// pointer hipGetSymbolAddress returns for the current device.
static constexpr size_t LANCZOS_SCRATCH_BYTES = size_t(64) << 20;
static __device__ __align__(256) unsigned char scratch[LANCZOS_SCRATCH_BYTES];

namespace smgpu {

static constexpr int TB = 256;
// Largest vector (amplitudes) whose mixed-precision Lanczos step (partial
// re-orthogonalisation) runs as the one-workgroup k_step_small
static constexpr int64_t SMALL_STEP_N = 8192;
// solo solves up to this many amplitudes replay captured steps by default (SMGPU_STEPS_AUTO)
static constexpr int64_t GRAPH_AUTO_N = 0;
static constexpr int EPL = 8;                 // elements per lane in the streaming passes
static constexpr int CHUNK = 64 * EPL;        // x per block in k_dotsw
static constexpr int UCHUNK = TB * 4;         // x per block in k_updatew
// a J-filtered Lanczos vector that keeps less than this of its norm was noise
// outside the J space: the J-space Krylov subspace is exhausted
static constexpr double J_EXHAUSTED = 1e-3;

// Where one Lanczos iteration stands, kept in device memory so that one captured
// graph replays every step: the step kernels' *_s forms read the iteration and the
// ring slots from here, and the last node (k_step_next) moves it to the next step.
struct StepRitz {
    int it;      // T[0..it] of the step
    int nn;      // min(num_ev, it + 1): Ritz values of the step
    int nprev;   // values of the step before usable as interlacing brackets
    int ring;    // (step + 1) & 3: the step's slot of the Ritz-value and flag rings
    int pring;   // step & 3: the slot of the step before
};
struct StepDev {
    int it;      // basis index of the H·v input v_it
    int nv;      // it + 1: vectors of the projection pass
    int store;   // it + 1 < lm: the step stores v_{it+1}
    StepRitz cur;    // this step's Ritz problem
    StepRitz prev;   // the previous step's, solved beside this step's H·v
    int prev_valid;  // 0 for the first step of a sweep
};

__device__ inline void step_at(StepDev* sd, int it, int ring, int nprev, int num_ev, int lm) {
    sd->it = it;
    sd->nv = it + 1;
    sd->store = it + 1 < lm ? 1 : 0;
    sd->cur.it = it;
    sd->cur.nn = min(num_ev, it + 1);
    sd->cur.nprev = nprev;
    sd->cur.ring = ring;
    sd->cur.pring = (ring + 3) & 3;
}

__global__ void k_step_set(StepDev* sd, int it, int step, int nprev, int num_ev, int lm) {
    step_at(sd, it, (step + 1) & 3, nprev, num_ev, lm);
    sd->prev_valid = 0;
}

__global__ void k_step_next(StepDev* sd, int num_ev, int lm) {
    sd->prev = sd->cur;
    sd->prev_valid = 1;
    const int it = sd->it + 1;
    step_at(sd, it, (sd->cur.ring + 1) & 3, min(num_ev, it), num_ev, lm);
}

__device__ inline double wave_sum(double v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__device__ inline double block_sum(double v, double* red) {
    v = wave_sum(v);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) red[wv] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0)
        for (int i = 0; i < (int)(blockDim.x >> 6); i++) s += red[i];
    return s;   // valid in thread 0
}

// part[blk*ldp + j] = sum over the block's CHUNK of B[j][x] w[x]  (j < nv);
// part[blk*ldp + nv] = sum of w^2 when with_norm.  One wave per j, no barriers.
template <typename TBv, typename TW>
__global__ __launch_bounds__(TB) void k_dotsw(const TBv* __restrict__ B, int64_t ldb, int nv,
                                              const TW* __restrict__ w, int64_t n, double* __restrict__ part,
                                              int ldp, int with_norm, const int* __restrict__ gate) {
    if (gate && *gate == 0) return;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int64_t x0 = (int64_t)blockIdx.x * CHUNK + lane;
    double wr[EPL];
#pragma unroll
    for (int e = 0; e < EPL; e++) {
        const int64_t x = x0 + 64 * e;
        wr[e] = x < n ? (double)w[x] : 0.0;
    }
    const int nj = nv + (with_norm ? 1 : 0);
    for (int j = wv; j < nj; j += TB / 64) {
        double s = 0.0;
        if (j < nv) {
            const TBv* b = B + (int64_t)j * ldb;
#pragma unroll
            for (int e = 0; e < EPL; e++) {
                const int64_t x = x0 + 64 * e;
                if (x < n) s += (double)b[x] * wr[e];
            }
        } else {
#pragma unroll
            for (int e = 0; e < EPL; e++) s += wr[e] * wr[e];
        }
        s = wave_sum(s);
        if (lane == 0) part[(int64_t)blockIdx.x * ldp + j] = s;
    }
}

// part[blk * ldp] = sum over the block's grid-stride range of a[x] b[x]
template <typename TA, typename TB2>
__device__ __forceinline__ void dot2_body(const TA* __restrict__ a, const TB2* __restrict__ b, int64_t n,
                                             double* __restrict__ part, int ldp) {
    __shared__ double red[TB / 64];
    double s = 0.0;
    for (int64_t x = (int64_t)blockIdx.x * TB + threadIdx.x; x < n; x += (int64_t)gridDim.x * TB)
        s += (double)a[x] * (double)b[x];
    s = block_sum(s, red);
    if (threadIdx.x == 0) part[(int64_t)blockIdx.x * ldp] = s;
}
template <typename TA, typename TB2>
__global__ __launch_bounds__(TB) void k_dot2(const TA* __restrict__ a, const TB2* __restrict__ b, int64_t n,
                                             double* __restrict__ part, int ldp) {
    dot2_body<TA, TB2>(a, b, n, part, ldp);
}
// step-state form: a = base + it * lda
template <typename TA, typename TB2>
__global__ __launch_bounds__(TB) void k_dot2_s(const StepDev* __restrict__ sd, const TA* __restrict__ base, int64_t lda,
                                               const TB2* __restrict__ b, int64_t n, double* __restrict__ part, int ldp) {
    dot2_body<TA, TB2>(base + (int64_t)sd->it * lda, b, n, part, ldp);
}


// out[j] = sum_g part[g*ldp + j], j < nj
__device__ __forceinline__ void reduce_body(const double* __restrict__ part, int G, int ldp,
                                               double* __restrict__ out, const int* __restrict__ gate) {
    if (gate && *gate == 0) return;
    __shared__ double red[TB / 64];
    const int j = blockIdx.x;
    double s = 0.0;
    for (int g = threadIdx.x; g < G; g += TB) s += part[(int64_t)g * ldp + j];
    s = block_sum(s, red);
    if (threadIdx.x == 0) out[j] = s;
}
__global__ __launch_bounds__(TB) void k_reduce(const double* __restrict__ part, int G, int ldp,
                                               double* __restrict__ out, const int* __restrict__ gate) {
    reduce_body(part, G, ldp, out, gate);
}
// step-state form over nv = it + 1 outputs: launched with lm blocks, those past nv exit
__global__ __launch_bounds__(TB) void k_reduce_s(const StepDev* __restrict__ sd, const double* __restrict__ part, int G,
                                                 int ldp, double* __restrict__ out, const int* __restrict__ gate) {
    if ((int)blockIdx.x >= sd->nv) return;
    reduce_body(part, G, ldp, out, gate);
}


// wout[x] = win[x] - sum_j h[j] B[j][x];  part[blk] = sum wout^2 over the block
template <typename TBv, typename TW>
__global__ __launch_bounds__(TB) void k_updatew(const TBv* __restrict__ B, int64_t ldb, int nv,
                                                const double* __restrict__ h, const TW* __restrict__ win,
                                                double* __restrict__ wout, int64_t n, double* __restrict__ part,
                                                const int* __restrict__ gate) {
    if (gate && *gate == 0) return;
    __shared__ double hs[1024];
    __shared__ double red[TB / 64];
    for (int j = threadIdx.x; j < nv; j += TB) hs[j] = h[j];
    __syncthreads();
    const int64_t x0 = (int64_t)blockIdx.x * UCHUNK + threadIdx.x;
    double a[4];
#pragma unroll
    for (int e = 0; e < 4; e++) {
        const int64_t x = x0 + TB * e;
        a[e] = x < n ? (double)win[x] : 0.0;
    }
    for (int j = 0; j < nv; j++) {
        const TBv* b = B + (int64_t)j * ldb;
        const double c = hs[j];
#pragma unroll
        for (int e = 0; e < 4; e++) {
            const int64_t x = x0 + TB * e;
            if (x < n) a[e] -= c * (double)b[x];
        }
    }
    double nr = 0.0;
#pragma unroll
    for (int e = 0; e < 4; e++) {
        const int64_t x = x0 + TB * e;
        if (x < n) { wout[x] = a[e]; nr += a[e] * a[e]; }
    }
    nr = block_sum(nr, red);
    if (threadIdx.x == 0) part[blockIdx.x] = nr;
}

// ---- CGS2 streaming passes.  Basis rows are read with 16-byte loads: a lane
// owns 4 consecutive elements (FP32) or 2 (FP64) per load, a wave instruction
// covers 1 KiB of one basis vector.  A block covers SEG elements of every basis
// vector; its waves take different vectors (j = wave, wave + 4, ...).
template <typename T> struct Ld16;
template <> struct Ld16<float> {
    static constexpr int W = 4;
    __device__ static inline void get(const float* p, double* o) {
        const float4 v = *reinterpret_cast<const float4*>(p);
        o[0] = v.x; o[1] = v.y; o[2] = v.z; o[3] = v.w;
    }
};
template <> struct Ld16<double> {
    static constexpr int W = 2;
    __device__ static inline void get(const double* p, double* o) {
        const double2 v = *reinterpret_cast<const double2*>(p);
        o[0] = v.x; o[1] = v.y;
    }
};

// elements per lane per basis vector and per-block segment length
template <typename TBv> struct GsShape {
    static constexpr int Q = 4;                              // 16-byte loads per lane per vector
    static constexpr int E = Q * Ld16<TBv>::W;               // elements per lane
    static constexpr int SEG = 64 * E;                       // elements per block (one wave width)
};

// element index of (lane, q, i) within the segment: lane-contiguous 16-byte pieces
template <typename TBv> __device__ inline int64_t gs_x(int64_t seg0, int lane, int q, int i) {
    constexpr int W = Ld16<TBv>::W;
    return seg0 + (int64_t)q * 64 * W + (int64_t)lane * W + i;
}

// part[blk*ldp + j] = sum_{x in segment} B[j][x] w'[x]   (ld of B is a multiple of 64)
// LOCAL = false: w' = w.
// LOCAL = true (the Lanczos local step): w' = w - alpha v_c - beta v_p with
// alpha = *alpha, beta = *bprev (v_p and bprev may be null), formed on the fly;
// w' is also written to wout (FP64) and |w'|^2 partials to npart[blk].
// gate: skip the launch when *gate == 0 (conditional second pass).
template <typename TBv, typename TW, bool LOCAL>
__device__ __forceinline__ void gs_dots_body(const TBv* __restrict__ B, int64_t ldb, int nv,
                                                const TW* __restrict__ w, int64_t n, double* __restrict__ part,
                                                int ldp, const int* __restrict__ gate,
                                                const TBv* __restrict__ vc, const TBv* __restrict__ vp,
                                                const double* __restrict__ alpha, const double* __restrict__ bprev,
                                                double* __restrict__ wout, double* __restrict__ npart) {
    if (gate && *gate == 0) return;
    using S = GsShape<TBv>;
    constexpr int W = Ld16<TBv>::W;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int64_t seg0 = (int64_t)blockIdx.x * S::SEG;
    double a[S::E];
#pragma unroll
    for (int q = 0; q < S::Q; q++)
#pragma unroll
        for (int i = 0; i < W; i++) {
            const int64_t x = gs_x<TBv>(seg0, lane, q, i);
            a[q * W + i] = x < n ? (double)w[x] : 0.0;
        }
    if (LOCAL) {
        const double al = *alpha;
        const double be = (vp && bprev) ? *bprev : 0.0;
        double nr = 0.0;
#pragma unroll
        for (int q = 0; q < S::Q; q++)
#pragma unroll
            for (int i = 0; i < W; i++) {
                const int64_t x = gs_x<TBv>(seg0, lane, q, i);
                if (x < n) {
                    double t = a[q * W + i] - al * (double)vc[x];
                    if (vp) t -= be * (double)vp[x];
                    a[q * W + i] = t;
                    nr += t * t;
                    if (wv == 0) wout[x] = t;
                }
            }
        if (wv == 0) {
            nr = wave_sum(nr);
            if (lane == 0) npart[blockIdx.x] = nr;
        }
    }
    const bool full = seg0 + S::SEG <= n;
#pragma unroll 2
    for (int j = wv; j < nv; j += TB / 64) {
        const TBv* b = B + (int64_t)j * ldb;
        double s = 0.0;
#pragma unroll
        for (int q = 0; q < S::Q; q++) {
            const int64_t x = gs_x<TBv>(seg0, lane, q, 0);
            double v[W];
            if (full) {
                Ld16<TBv>::get(b + x, v);
            } else {
                // tail segment: elements past n (padding of the basis rows) count as 0
                if (x < ldb) Ld16<TBv>::get(b + x, v);
                for (int i = 0; i < W; i++) if (!(x < ldb) || x + i >= n) v[i] = 0.0;
            }
#pragma unroll
            for (int i = 0; i < W; i++) s += v[i] * a[q * W + i];
        }
        s = wave_sum(s);
        if (lane == 0) part[(int64_t)blockIdx.x * ldp + j] = s;
    }
}
template <typename TBv, typename TW, bool LOCAL>
__global__ __launch_bounds__(TB) void k_gs_dots(const TBv* __restrict__ B, int64_t ldb, int nv,
                                                const TW* __restrict__ w, int64_t n, double* __restrict__ part,
                                                int ldp, const int* __restrict__ gate,
                                                const TBv* __restrict__ vc, const TBv* __restrict__ vp,
                                                const double* __restrict__ alpha, const double* __restrict__ bprev,
                                                double* __restrict__ wout, double* __restrict__ npart) {
    gs_dots_body<TBv, TW, LOCAL>(B, ldb, nv, w, n, part, ldp, gate, vc, vp, alpha, bprev, wout, npart);
}
// step-state form: nv = it + 1; with LOCAL, v_c = B_it, v_p = B_{it-1} and bprev = beta + it - 1
template <typename TBv, typename TW, bool LOCAL>
__global__ __launch_bounds__(TB) void k_gs_dots_s(const StepDev* __restrict__ sd, const TBv* __restrict__ B, int64_t ldb,
                                                  const TW* __restrict__ w, int64_t n, double* __restrict__ part,
                                                  int ldp, const int* __restrict__ gate,
                                                  const double* __restrict__ alpha, const double* __restrict__ beta,
                                                  double* __restrict__ wout, double* __restrict__ npart) {
    const int it = sd->it;
    gs_dots_body<TBv, TW, LOCAL>(B, ldb, it + 1, w, n, part, ldp, gate, B + (int64_t)it * ldb,
                                 it > 0 ? B + (int64_t)(it - 1) * ldb : nullptr, alpha,
                                 it > 0 ? beta + it - 1 : nullptr, wout, npart);
}


// wout = win - sum_j h[j] B[j]  and, if dots, part = partial B^T wout; npart = |wout|^2 partial.
// Each wave owns a quarter of the block's segment for the update and then takes
// vectors j = wave, wave+4, ... for the dot products (its segment rows were
// just streamed by the block, so this second read hits the caches).
template <typename TBv, typename TW, bool DOTS>
__device__ __forceinline__ void gs_upd_body(const TBv* __restrict__ B, int64_t ldb, int nv,
                                               const double* __restrict__ h, const TW* __restrict__ win,
                                               double* __restrict__ wout, int64_t n, double* __restrict__ part,
                                               int ldp, double* __restrict__ npart, const int* __restrict__ gate) {
    if (gate && *gate == 0) return;
    using S = GsShape<TBv>;
    constexpr int W = Ld16<TBv>::W;
    constexpr int QW = S::Q / 4 > 0 ? S::Q / 4 : 1;            // 16-byte pieces per lane in the update
    __shared__ double hs[1024];
    __shared__ double ws[S::SEG];
    __shared__ double red[TB / 64];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    for (int j = threadIdx.x; j < nv; j += TB) hs[j] = h[j];
    __syncthreads();
    const int64_t seg0 = (int64_t)blockIdx.x * S::SEG;
    // update: thread t owns elements t*W .. t*W+W-1 of each of QW quarter-rows
    double nr = 0.0;
#pragma unroll
    for (int qw = 0; qw < QW; qw++) {
        const int loc = (qw * TB + threadIdx.x) * W;         // position inside the segment
        const int64_t x = seg0 + loc;
        double a[W];
#pragma unroll
        for (int i = 0; i < W; i++) a[i] = (x + i < n) ? (double)win[x + i] : 0.0;
        if (x < ldb) {
#pragma unroll 4
            for (int j = 0; j < nv; j++) {
                double v[W];
                Ld16<TBv>::get(B + (int64_t)j * ldb + x, v);
                const double c = hs[j];
#pragma unroll
                for (int i = 0; i < W; i++) a[i] -= c * v[i];
            }
        }
#pragma unroll
        for (int i = 0; i < W; i++) {
            if (x + i < n) { wout[x + i] = a[i]; nr += a[i] * a[i]; }
            else a[i] = 0.0;
            ws[loc + i] = a[i];
        }
    }
    nr = block_sum(nr, red);
    if (threadIdx.x == 0) npart[blockIdx.x] = nr;
    if (!DOTS) return;
    __syncthreads();
    double wr[S::E];
#pragma unroll
    for (int q = 0; q < S::Q; q++)
#pragma unroll
        for (int i = 0; i < W; i++) wr[q * W + i] = ws[q * 64 * W + lane * W + i];
    const bool full = seg0 + S::SEG <= n;
#pragma unroll 2
    for (int j = wv; j < nv; j += TB / 64) {
        const TBv* b = B + (int64_t)j * ldb;
        double s = 0.0;
#pragma unroll
        for (int q = 0; q < S::Q; q++) {
            const int64_t x = gs_x<TBv>(seg0, lane, q, 0);
            double v[W];
            if (full) {
                Ld16<TBv>::get(b + x, v);
            } else {
                if (x < ldb) Ld16<TBv>::get(b + x, v);
                for (int i = 0; i < W; i++) if (!(x < ldb) || x + i >= n) v[i] = 0.0;
            }
#pragma unroll
            for (int i = 0; i < W; i++) s += v[i] * wr[q * W + i];
        }
        s = wave_sum(s);
        if (lane == 0) part[(int64_t)blockIdx.x * ldp + j] = s;
    }
}
template <typename TBv, typename TW, bool DOTS>
__global__ __launch_bounds__(TB) void k_gs_upd(const TBv* __restrict__ B, int64_t ldb, int nv,
                                               const double* __restrict__ h, const TW* __restrict__ win,
                                               double* __restrict__ wout, int64_t n, double* __restrict__ part,
                                               int ldp, double* __restrict__ npart, const int* __restrict__ gate) {
    gs_upd_body<TBv, TW, DOTS>(B, ldb, nv, h, win, wout, n, part, ldp, npart, gate);
}
// step-state form: nv = it + 1
template <typename TBv, typename TW, bool DOTS>
__global__ __launch_bounds__(TB) void k_gs_upd_s(const StepDev* __restrict__ sd, const TBv* __restrict__ B, int64_t ldb,
                                                 const double* __restrict__ h, const TW* __restrict__ win,
                                                 double* __restrict__ wout, int64_t n, double* __restrict__ part,
                                                 int ldp, double* __restrict__ npart, const int* __restrict__ gate) {
    gs_upd_body<TBv, TW, DOTS>(B, ldb, sd->nv, h, win, wout, n, part, ldp, npart, gate);
}


// alpha[it] = v.Hv (before any orthogonalisation), beta[it] = |w| after CGS2
__global__ void k_finish(const double* h1, int it, const double* npart, int G, double* alpha, double* beta) {
    __shared__ double red[TB / 64];
    double s = 0.0;
    for (int g = threadIdx.x; g < G; g += TB) s += npart[g];
    s = block_sum(s, red);
    if (threadIdx.x == 0) {
        alpha[it] = h1[it];
        beta[it] = sqrt(s);
    }
}

// Daniel-Gragg-Kaufman-Stewart test after one CGS pass of the local-step vector:
// loc[1] = |w'|^2, loc[2] = |w''|^2; a second pass is needed (gate = 1) when the
// pass removed more than half of |w'|^2 (|w''| < |w'| / sqrt(2)).  first: the gate
// of the first pass (null: it always ran); no first pass, no second.
__global__ void k_dgks(const double* npart_after, const double* npart_before, int G, double* loc, int* gate,
                       const int* first) {
    if (first && *first == 0) {
        if (threadIdx.x == 0) { gate[0] = 0; loc[4] = 0.0; }
        return;
    }
    __shared__ double red[TB / 64];
    double a = 0.0, b = 0.0;
    for (int g = threadIdx.x; g < G; g += TB) { a += npart_after[g]; b += npart_before[g]; }
    a = block_sum(a, red);
    __syncthreads();
    b = block_sum(b, red);
    if (threadIdx.x == 0) {
        loc[1] = b;
        loc[2] = a;
        gate[0] = (a < 0.5 * b) ? 1 : 0;
        loc[4] = gate[0];
    }
}

// alpha[it] = v_it . H v_it (the local-step coefficient), beta[it] = |w| at the end
__device__ __forceinline__ void finish_local_body(double* loc, int it, const double* npart, int G, double* alpha, double* beta) {
    __shared__ double red[TB / 64];
    double s = 0.0;
    for (int g = threadIdx.x; g < G; g += TB) s += npart[g];
    s = block_sum(s, red);
    if (threadIdx.x == 0) {
        alpha[it] = loc[0];
        beta[it] = sqrt(s);
        loc[3] += loc[4];          // second CGS passes so far (loc[4] = this step's gate)
    }
}
__global__ void k_finish_local(double* loc, int it, const double* npart, int G, double* alpha, double* beta) {
    finish_local_body(loc, it, npart, G, alpha, beta);
}
__global__ void k_finish_local_s(const StepDev* __restrict__ sd, double* loc, const double* npart, int G, double* alpha,
                                 double* beta) {
    finish_local_body(loc, sd->it, npart, G, alpha, beta);
}


__global__ void k_sqrt1(const double* a, double* b) { b[0] = sqrt(a[0]); }

// Local Lanczos step without a projection pass: w = win - alpha v_c - beta v_p (FP64)
// with |w|^2 partials per segment of GsShape<TBv>::SEG elements into npart and npart2
// (the same partition as the Gram-Schmidt kernels, so k_finish_local reads either).
template <typename TBv, typename TW>
__device__ __forceinline__ void local_step_body(const TW* __restrict__ win, const TBv* __restrict__ vc,
                                                   const TBv* __restrict__ vp, const double* __restrict__ alpha,
                                                   const double* __restrict__ bprev, double* __restrict__ wout,
                                                   int64_t n, double* __restrict__ npart, double* __restrict__ npart2) {
    __shared__ double red[TB / 64];
    const int64_t seg0 = (int64_t)blockIdx.x * GsShape<TBv>::SEG;
    const int64_t seg1 = min(seg0 + (int64_t)GsShape<TBv>::SEG, n);
    const double al = *alpha;
    const double be = (vp && bprev) ? *bprev : 0.0;
    double nr = 0.0;
    for (int64_t x = seg0 + threadIdx.x; x < seg1; x += TB) {
        double t = (double)win[x] - al * (double)vc[x];
        if (vp) t -= be * (double)vp[x];
        wout[x] = t;
        nr += t * t;
    }
    nr = block_sum(nr, red);
    if (threadIdx.x == 0) { npart[blockIdx.x] = nr; npart2[blockIdx.x] = nr; }
}
template <typename TBv, typename TW>
__global__ __launch_bounds__(TB) void k_local_step(const TW* __restrict__ win, const TBv* __restrict__ vc,
                                                   const TBv* __restrict__ vp, const double* __restrict__ alpha,
                                                   const double* __restrict__ bprev, double* __restrict__ wout,
                                                   int64_t n, double* __restrict__ npart, double* __restrict__ npart2) {
    local_step_body<TBv, TW>(win, vc, vp, alpha, bprev, wout, n, npart, npart2);
}
// step-state form: v_c = B_it, v_p = B_{it-1}, bprev = beta + it - 1
template <typename TBv, typename TW>
__global__ __launch_bounds__(TB) void k_local_step_s(const StepDev* __restrict__ sd, const TW* __restrict__ win,
                                                     const TBv* __restrict__ B, int64_t ldb,
                                                     const double* __restrict__ alpha, const double* __restrict__ beta,
                                                     double* __restrict__ wout, int64_t n, double* __restrict__ npart,
                                                     double* __restrict__ npart2) {
    const int it = sd->it;
    local_step_body<TBv, TW>(win, B + (int64_t)it * ldb, it > 0 ? B + (int64_t)(it - 1) * ldb : nullptr, alpha,
                             it > 0 ? beta + it - 1 : nullptr, wout, n, npart, npart2);
}


// parameters of the omega recurrence: eps = unit of the basis precision (the value
// a reorthogonalised row restarts from), eps_r = rounding level of one Lanczos step
// (the recurrence's error term), thr = gate threshold of the projection pass
struct OmegaParams {
    double eps, eps_r, thr;
};

// Partial re-orthogonalisation (Simon's omega recurrence) for step j = it: with
// beta' = |w'| from the partials, a_j = loc[0] and the stored alpha/beta of steps < j,
//   beta' om_{j+1,k} = b_k om_{j,k+1} + (a_k - a_j) om_{j,k} + b_{k-1} om_{j,k-1}
//                      - b_{j-1} om_{j-1,k} +- eps (b_k + beta') * 3        (k < j)
// (rounding term taken with the sign that grows |om|), om_{j+1,j} = eps, om_{j+1,j+1} = 1.
// Rows om_{j-1}, om_j, om_{j+1} rotate through three rows of om (stride ldo).
// gate[1] = 1 runs the projection pass of this step: when max_k |om_{j+1,k}| >
// sqrt(eps) (and then also for the next step, gate[2] counts forced steps), after
// which om_{j+1,k} = eps.  loc[5] counts the steps that ran it.  A step without the
// pass also has no DGKS second pass: gate[0] = 0, loc[4] = 0.
// (one workgroup of TB threads; returns the gate of this step's pass in every thread)
__device__ int omega_block(const double* __restrict__ alpha, const double* __restrict__ beta, int j,
                           const double* __restrict__ npart, int G, double* __restrict__ om, int ldo,
                           double* __restrict__ loc, int* __restrict__ gate, const OmegaParams op,
                           double* bj_out = nullptr) {
    const double eps = op.eps;
    __shared__ double red[TB / 64];
    __shared__ double mx[TB / 64];
    const int forced = gate[2];        // read before the barriers below; thread 0 rewrites it last
    double s = 0.0;
    for (int g = threadIdx.x; g < G; g += TB) s += npart[g];
    s = block_sum(s, red);
    __shared__ double bsh;
    if (threadIdx.x == 0) bsh = sqrt(s);
    __syncthreads();
    const double bj = bsh;              // |w'| in every thread
    const double* cur = om + (int64_t)(j % 3) * ldo;
    const double* prev = om + (int64_t)((j + 2) % 3) * ldo;
    double* next = om + (int64_t)((j + 1) % 3) * ldo;
    const double aj = loc[0];
    const double bjm1 = j > 0 ? beta[j - 1] : 0.0;
    double m = 0.0;
    if (bj > 0.0) {
        for (int k = threadIdx.x; k < j; k += TB) {
            double t = beta[k] * cur[k + 1] + (alpha[k] - aj) * cur[k] - bjm1 * prev[k];
            if (k > 0) t += beta[k - 1] * cur[k - 1];
            const double th = 3.0 * op.eps_r * (beta[k] + bj);
            t += t >= 0.0 ? th : -th;
            t /= bj;
            next[k] = t;
            m = fmax(m, fabs(t));
        }
    }
    // block max
    for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o));
    if ((threadIdx.x & 63) == 0) mx[threadIdx.x >> 6] = m;
    __syncthreads();
    m = fmax(fmax(mx[0], mx[1]), fmax(mx[2], mx[3]));
    const int run = (bj == 0.0 || forced > 0 || m > op.thr) ? 1 : 0;
    if (run) for (int k = threadIdx.x; k < j; k += TB) next[k] = eps;
    if (threadIdx.x == 0) {
        next[j] = eps;
        next[j + 1] = 1.0;
        gate[1] = run;
        gate[2] = forced > 0 ? forced - 1 : (run ? 1 : 0);
        loc[5] += run;
        if (!run) { gate[0] = 0; loc[4] = 0.0; }
        if (bj_out) *bj_out = bj;
    }
    return run;
}

// ... and, when hgate is set, the gate also to the host-visible *hgate (coherent host
// memory): the host then launches the projection pass only for steps that run it.  A
// step without the pass also gets alpha[j] = loc[0] and beta[j] = |w'| here, the
// values k_finish_local would write (the same sum of the same partials).
__global__ void k_omega(double* __restrict__ alpha, double* __restrict__ beta, int j,
                        const double* __restrict__ npart, int G, double* __restrict__ om, int ldo,
                        double* __restrict__ loc, int* __restrict__ gate, OmegaParams op, int* hgate) {
    __shared__ double bj;
    const int run = omega_block(alpha, beta, j, npart, G, om, ldo, loc, gate, op, &bj);
    if (threadIdx.x == 0) {
        if (!run) { alpha[j] = loc[0]; beta[j] = bj; }
        if (hgate) *hgate = run;
    }
}
__global__ void k_omega_s(const StepDev* __restrict__ sd, double* __restrict__ alpha, double* __restrict__ beta,
                          const double* __restrict__ npart, int G, double* __restrict__ om, int ldo,
                          double* __restrict__ loc, int* __restrict__ gate, OmegaParams op) {
    __shared__ double bj;
    const int j = sd->it;
    const int run = omega_block(alpha, beta, j, npart, G, om, ldo, loc, gate, op, &bj);
    if (threadIdx.x == 0 && !run) { alpha[j] = loc[0]; beta[j] = bj; }
}


// One projection pass of w (FP64, in place) against the basis B[0..nv) by one
// workgroup: h[j] = B_j . w (wave j mod 4, lanes along the vector), then
// w -= sum_j h[j] B_j; returns |w|^2 after the pass (every thread).
template <typename TBv>
__device__ double small_pass(const TBv* __restrict__ B, int64_t ldb, int nv, double* __restrict__ w, int64_t n,
                             double* __restrict__ h, double* hs, double* red) {
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    for (int j = wv; j < nv; j += TB / 64) {
        const TBv* b = B + (int64_t)j * ldb;
        double d = 0.0;
        #pragma unroll 4
        for (int64_t x = lane; x < n; x += 64) d += (double)b[x] * w[x];
        d = wave_sum(d);
        if (lane == 0) hs[j] = d;
    }
    __syncthreads();
    for (int j = threadIdx.x; j < nv; j += TB) h[j] = hs[j];
    double nr = 0.0;
    #pragma unroll 4
    for (int64_t x = threadIdx.x; x < n; x += TB) {
        double t = w[x];
        for (int j = 0; j < nv; j++) t -= hs[j] * (double)B[(int64_t)j * ldb + x];
        w[x] = t;
        nr += t * t;
    }
    nr = block_sum(nr, red);
    __shared__ double nsh;
    if (threadIdx.x == 0) nsh = nr;
    __syncthreads();
    return nsh;
}

// The whole Lanczos step after H·v for a small vector (n <= SMALL_STEP_N), in one
// workgroup, with partial re-orthogonalisation: alpha = v_it . Hv, the local step
// w = Hv - alpha v_it - beta_{it-1} v_{it-1} (FP64), the omega recurrence and its
// gated projection pass with the DGKS second pass, alpha[it] / beta[it], and
// vnext = w / beta[it] (when set).  The same operations as the multi-kernel step
// (k_dot2 .. k_store), with one launch instead of about a dozen dependent ones.
template <typename TBv, typename TW>
__device__ __forceinline__ void step_small_body(const TW* __restrict__ win, const TBv* __restrict__ B,
                                                   int64_t ldb, int it, int64_t n, double* __restrict__ w,
                                                   double* __restrict__ loc, int* __restrict__ gate,
                                                   double* __restrict__ alpha, double* __restrict__ beta,
                                                   double* __restrict__ om, int ldo, OmegaParams op,
                                                   double* __restrict__ h1, double* __restrict__ h2,
                                                   TBv* __restrict__ vnext) {
    __shared__ double red[TB / 64];
    __shared__ double hs[1024];
    __shared__ double sh[2];
    const TBv* vc = B + (int64_t)it * ldb;
    const TBv* vp = it > 0 ? B + (int64_t)(it - 1) * ldb : nullptr;
    double s = 0.0;
    #pragma unroll 4
    for (int64_t x = threadIdx.x; x < n; x += TB) s += (double)vc[x] * (double)win[x];
    s = block_sum(s, red);
    if (threadIdx.x == 0) sh[0] = s;
    __syncthreads();
    const double al = sh[0];
    const double be = vp ? beta[it - 1] : 0.0;
    double nr = 0.0;
    #pragma unroll 4
    for (int64_t x = threadIdx.x; x < n; x += TB) {
        double t = (double)win[x] - al * (double)vc[x];
        if (vp) t -= be * (double)vp[x];
        w[x] = t;
        nr += t * t;
    }
    nr = block_sum(nr, red);
    if (threadIdx.x == 0) { sh[1] = nr; loc[0] = al; }
    __syncthreads();
    const double nb = sh[1];                 // |w'|^2
    const int run = omega_block(alpha, beta, it, sh + 1, 1, om, ldo, loc, gate, op);
    double nf = nb;
    int second = 0;
    if (run) {
        const double na = small_pass<TBv>(B, ldb, it + 1, w, n, h1, hs, red);
        second = na < 0.5 * nb ? 1 : 0;      // DGKS: the pass removed more than half of |w'|^2
        if (threadIdx.x == 0) { loc[1] = nb; loc[2] = na; }
        nf = second ? small_pass<TBv>(B, ldb, it + 1, w, n, h2, hs, red) : na;
    }
    if (threadIdx.x == 0) {
        gate[0] = second;
        loc[4] = second;
        loc[3] += second;
        alpha[it] = al;
        beta[it] = sqrt(nf);
    }
    if (!vnext) return;
    const double inv = 1.0 / sqrt(nf);
    #pragma unroll 4
    for (int64_t x = threadIdx.x; x < n; x += TB) vnext[x] = (TBv)(w[x] * inv);
}
template <typename TBv, typename TW>
__global__ __launch_bounds__(TB) void k_step_small(const TW* __restrict__ win, const TBv* __restrict__ B,
                                                   int64_t ldb, int it, int64_t n, double* __restrict__ w,
                                                   double* __restrict__ loc, int* __restrict__ gate,
                                                   double* __restrict__ alpha, double* __restrict__ beta,
                                                   double* __restrict__ om, int ldo, OmegaParams op,
                                                   double* __restrict__ h1, double* __restrict__ h2,
                                                   TBv* __restrict__ vnext) {
    step_small_body<TBv, TW>(win, B, ldb, it, n, w, loc, gate, alpha, beta, om, ldo, op, h1, h2, vnext);
}
// step-state form: it from sd, vnext = B_{it+1} while it + 1 < lm
template <typename TBv, typename TW>
__global__ __launch_bounds__(TB) void k_step_small_s(const StepDev* __restrict__ sd, const TW* __restrict__ win,
                                                     TBv* __restrict__ B, int64_t ldb, int64_t n, double* __restrict__ w,
                                                     double* __restrict__ loc, int* __restrict__ gate,
                                                     double* __restrict__ alpha, double* __restrict__ beta,
                                                     double* __restrict__ om, int ldo, OmegaParams op,
                                                     double* __restrict__ h1, double* __restrict__ h2) {
    const int it = sd->it;
    step_small_body<TBv, TW>(win, B, ldb, it, n, w, loc, gate, alpha, beta, om, ldo, op, h1, h2,
                             sd->store ? B + (int64_t)(it + 1) * ldb : nullptr);
}


// omega rows after a thick restart at step j: the kept vectors and v_j are
// orthonormalised, so om_{j,k} = om_{j-1,k} = eps off the diagonal; the next two
// steps run their projection pass.
__global__ void k_omega_reset(double* __restrict__ om, int ldo, int j, double eps, int* __restrict__ gate) {
    double* cur = om + (int64_t)(j % 3) * ldo;
    double* prev = om + (int64_t)((j + 2) % 3) * ldo;
    for (int k = threadIdx.x; k <= j; k += TB) {
        cur[k] = k == j ? 1.0 : eps;
        prev[k] = k == j - 1 ? 1.0 : (k < j - 1 ? eps : 0.0);
    }
    if (threadIdx.x == 0) gate[2] = 2;
}

// one factor of the J projector: y = (z - c y) * inv, with z = J^2 y
__global__ __launch_bounds__(TB) void k_jfactor(double* __restrict__ y, const double* __restrict__ z, int64_t n,
                                                double c, double inv) {
    for (int64_t x = (int64_t)blockIdx.x * TB + threadIdx.x; x < n; x += (int64_t)gridDim.x * TB)
        y[x] = (z[x] - c * y[x]) * inv;
}

// norm of a vector into out[0] (as |x|)
__global__ void k_norm_from_parts(const double* npart, int G, double* out) {
    __shared__ double red[TB / 64];
    double s = 0.0;
    for (int g = threadIdx.x; g < G; g += TB) s += npart[g];
    s = block_sum(s, red);
    if (threadIdx.x == 0) out[0] = sqrt(s);
}

template <typename TBv>
__global__ __launch_bounds__(TB) void k_store(const double* __restrict__ w, const double* __restrict__ den,
                                              TBv* __restrict__ dst, int64_t n) {
    const double inv = 1.0 / den[0];
    for (int64_t x = (int64_t)blockIdx.x * TB + threadIdx.x; x < n; x += (int64_t)gridDim.x * TB)
        dst[x] = (TBv)(w[x] * inv);
}
// step-state form: V_{it+1} = w / beta[it] while it + 1 < lm
template <typename TBv>
__global__ __launch_bounds__(TB) void k_store_s(const StepDev* __restrict__ sd, const double* __restrict__ w,
                                                const double* __restrict__ beta, TBv* __restrict__ B, int64_t ldb,
                                                int64_t n) {
    if (!sd->store) return;
    const int it = sd->it;
    const double inv = 1.0 / beta[it];
    TBv* dst = B + (int64_t)(it + 1) * ldb;
    for (int64_t x = (int64_t)blockIdx.x * TB + threadIdx.x; x < n; x += (int64_t)gridDim.x * TB)
        dst[x] = (TBv)(w[x] * inv);
}


template <typename TA, typename TBv>
__global__ __launch_bounds__(TB) void k_convert(const TA* __restrict__ a, TBv* __restrict__ b, int64_t n, double scale) {
    for (int64_t x = (int64_t)blockIdx.x * TB + threadIdx.x; x < n; x += (int64_t)gridDim.x * TB)
        b[x] = (TBv)((double)a[x] * scale);
}

// Ritz values: block j computes eigenvalue j of T[0..n-1] by 256-way multisection.
// prev (nprev values) are the eigenvalues of the leading (n-1)x(n-1) block when
// nprev > 0: lambda_j(T_n) lies in [prev[j-1], prev[j]] (Cauchy interlacing).
// With a convergence slot (flag != null) the last block to finish also runs the
// test of diagonalize_T! (lanczos_methods.jl:250-258: all(prev - out < tol) over
// num_ev values; entries past the blocks keep prev) and writes its result to
// *flag and to the host-visible *hflag (ticket: zero between launches).
__device__ __forceinline__ void ritz_body(const double* __restrict__ alpha, const double* __restrict__ beta,
                                          int n, const double* __restrict__ prev, int nprev,
                                          double* __restrict__ out, int num_ev, double tol,
                                          int* __restrict__ flag, int* hflag, int* __restrict__ ticket, int nblk) {
    __shared__ double sa[1024], sb[1024];
    __shared__ int firstbig;
    __shared__ double slo, shi;
    __shared__ double rl[TB / 64], rh[TB / 64], rm[TB / 64];
    __shared__ int ok[2];
    const int j = blockIdx.x;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    for (int i = threadIdx.x; i < n; i += TB) { sa[i] = alpha[i]; sb[i] = i < n - 1 ? beta[i] : 0.0; }
    __syncthreads();
    // Gershgorin bounds and the pivot floor (tri_bounds, tri_pivmin) as block reductions
    double l = 1e300, h = -1e300, m = 1.0;
    for (int i = threadIdx.x; i < n; i += TB) {
        const double r = (i > 0 ? fabs(sb[i - 1]) : 0.0) + (i < n - 1 ? fabs(sb[i]) : 0.0);
        l = fmin(l, sa[i] - r);
        h = fmax(h, sa[i] + r);
        if (i < n - 1) m = fmax(m, sb[i] * sb[i]);
    }
    for (int o = 32; o > 0; o >>= 1) {
        l = fmin(l, __shfl_xor(l, o));
        h = fmax(h, __shfl_xor(h, o));
        m = fmax(m, __shfl_xor(m, o));
    }
    if (lane == 0) { rl[wv] = l; rh[wv] = h; rm[wv] = m; }
    __syncthreads();
    l = rl[0]; h = rh[0]; m = rm[0];
    for (int w = 1; w < TB / 64; w++) { l = fmin(l, rl[w]); h = fmax(h, rh[w]); m = fmax(m, rm[w]); }
    const double bpad = 1e-12 * fmax(1.0, fmax(fabs(l), fabs(h)));
    const double lo = l - bpad, hi = h + bpad;
    const double pm = 2.2250738585072014e-308 / 2.220446049250313e-16 * m;
    // Cauchy interlacing bracket from the previous step's values, kept only if the
    // Sturm counts at its ends (two threads) confirm it
    double l2 = lo, h2 = hi;
    bool geo = false;
    if (nprev > 0) {
        const double pad = 1e-9 * fmax(1.0, fmax(fabs(lo), fabs(hi)));
        if (j >= 1 && j - 1 < nprev) l2 = fmax(lo, prev[j - 1] - pad);
        if (j < nprev) h2 = fmin(hi, prev[j] + pad);
        if (threadIdx.x == 0) ok[0] = l2 < h2 && tri_count(sa, sb, n, l2, pm) <= j;
        if (threadIdx.x == 1) ok[1] = l2 < h2 && tri_count(sa, sb, n, h2, pm) > j;
        __syncthreads();
        if (!(ok[0] && ok[1])) { l2 = lo; h2 = hi; }
        else geo = j < nprev && h2 - l2 > 2.0 * pad;
    }
    if (threadIdx.x == 0) { slo = l2; shi = h2; }
    __syncthreads();
    // First round over the interlacing bracket [l2, h2 = prev_j + pad]: the eigenvalue
    // is at or below prev_j and close to it once the solve converges, so the points
    // are spaced geometrically in their distance from h2 (from the bracket width down
    // to pad / 2); later rounds are uniform multisection.
    const double gpad = 0.5e-9 * fmax(1.0, fmax(fabs(lo), fabs(hi)));
    const double gr = geo ? pow(gpad / (h2 - l2), 1.0 / (double)(TB - 1)) : 0.0;
    for (int round = 0; round < 12; round++) {
        const double rlo = slo, rhi = shi;
        if (rhi - rlo <= 1e-12 * fmax(1.0, fmax(fabs(rlo), fabs(rhi)))) break;
        const bool g = geo && round == 0;
        if (threadIdx.x == 0) firstbig = TB;
        __syncthreads();
        const double x = g ? rhi - (rhi - rlo) * pow(gr, (double)threadIdx.x)
                           : rlo + (rhi - rlo) * (double)(threadIdx.x + 1) / (double)(TB + 1);
        const int c = tri_count(sa, sb, n, x, pm);
        if (c > j) atomicMin(&firstbig, (int)threadIdx.x);
        __syncthreads();
        if (threadIdx.x == 0) {
            const int t = firstbig;
            double nlo, nhi;
            if (g) {
                nlo = t == 0 ? rlo : rhi - (rhi - rlo) * pow(gr, (double)(t - 1));
                nhi = t == TB ? rhi : rhi - (rhi - rlo) * pow(gr, (double)t);
            } else {
                nlo = t == 0 ? rlo : rlo + (rhi - rlo) * (double)t / (double)(TB + 1);
                nhi = t == TB ? rhi : rlo + (rhi - rlo) * (double)(t + 1) / (double)(TB + 1);
            }
            slo = nlo; shi = nhi;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) out[j] = 0.5 * (slo + shi);
    if (!flag) return;
    __shared__ int is_last;
    if (threadIdx.x == 0) {
        __threadfence();
        is_last = atomicAdd(ticket, 1) == nblk - 1;
    }
    __syncthreads();
    if (!is_last || threadIdx.x != 0) return;
    __threadfence();
    for (int k = nblk; k < num_ev; k++) out[k] = prev[k];
    // converged once all num_ev values stagnate; while T holds fewer than num_ev
    // (num_ev above ls, right after a restart) the stale ones do not count as converged
    int okc = nblk >= num_ev ? 1 : 0;
    for (int k = 0; k < num_ev; k++) okc &= (prev[k] - out[k] < tol) ? 1 : 0;
    *flag = okc;
    *hflag = okc;
    __threadfence_system();
    *ticket = 0;
}
__global__ __launch_bounds__(TB) void k_ritz(const double* __restrict__ alpha, const double* __restrict__ beta,
                                             int n, const double* __restrict__ prev, int nprev,
                                             double* __restrict__ out, int num_ev, double tol,
                                             int* __restrict__ flag, int* hflag, int* __restrict__ ticket) {
    ritz_body(alpha, beta, n, prev, nprev, out, num_ev, tol, flag, hflag, ticket, (int)gridDim.x);
}
// step-state form: the Ritz problem of the previous step (T[0..it], its ring slots),
// nothing for the first step of a sweep; launched with num_ev blocks, those past
// min(num_ev, it + 1) exit
__global__ __launch_bounds__(TB) void k_ritz_s(const StepDev* __restrict__ sd, const double* __restrict__ alpha,
                                               const double* __restrict__ beta, double* __restrict__ enring,
                                               int num_ev, double tol, int* __restrict__ flags, int* hflag,
                                               int* __restrict__ ticket) {
    if (!sd->prev_valid) return;
    const StepRitz r = sd->prev;
    if ((int)blockIdx.x >= r.nn) return;
    ritz_body(alpha, beta, r.it + 1, enring + (int64_t)r.pring * 64, r.nprev, enring + (int64_t)r.ring * 64,
              num_ev, tol, flags + r.ring, hflag + r.ring, ticket, r.nn);
}


// diagonalize_T! bookkeeping: entries n..num_ev-1 keep the previous values,
// flag = all(prev - cur < tol)

// Eigenvectors of T[0..n-1] for the k eigenvalues vals (ascending) by inverse
// iteration: one thread per cluster of close eigenvalues (chained gaps below
// 1e-3 |T|); members of a cluster are orthogonalised against the earlier ones
// exactly as tri_lowest does.  work: k * 5n doubles.
__global__ void k_tri_vecs(const double* __restrict__ alpha, const double* __restrict__ beta, int n,
                           const double* __restrict__ vals, int k, double* __restrict__ Y, double* __restrict__ work) {
    const int t = threadIdx.x;
    if (t >= k) return;
    const double tn = tri_norm(alpha, beta, n);
    const double ortol = 1e-3 * tn;
    if (t > 0 && fabs(vals[t] - vals[t - 1]) < ortol) return;       // not a cluster leader
    int c1 = t + 1;
    while (c1 < k && fabs(vals[c1] - vals[c1 - 1]) < ortol) c1++;
    for (int j = t; j < c1; j++) {
        const int first = tri_cluster_first(vals, j, ortol);
        tri_eigvec(alpha, beta, n, vals[j], Y + (long)j * n, work + (long)t * 5 * n, Y + (long)first * n, j - first, tn);
    }
}

// restart, step 3: coupling vector sigma and the small Lanczos on diag(theta);
// overwrites alpha/beta[0..ls-1] with the new leading block
__global__ void k_restart_q(double* alpha, double* beta, int m, int ls, const double* __restrict__ theta,
                            const double* __restrict__ Y, double* __restrict__ Qrev, double* __restrict__ work) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double* sig = work;
    const double bcoup = beta[m - 1];
    for (int j = 0; j < ls; j++) sig[j] = bcoup * Y[(long)j * m + (m - 1)];
    tri_diag_lanczos(theta, sig, ls, Qrev, alpha, beta, work + ls);
}

// restart, step 4: S = Y Qrev (m x ls), column j at S + j*m
__global__ void k_restart_s(const double* __restrict__ Y, const double* __restrict__ Qrev, int m, int ls,
                            double* __restrict__ S) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    for (int j = 0; j < ls; j++) {
        double s = 0.0;
        for (int t = 0; t < ls; t++) s += Y[(long)t * m + i] * Qrev[(long)j * ls + t];
        S[(long)j * m + i] = s;
    }
}

// start vector: counter-based normal deviates (SplitMix64 + Box-Muller), so the
// seed alone fixes the vector for any launch shape
__device__ inline uint64_t splitmix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}
__global__ void k_randn(double* __restrict__ v, int64_t n, uint64_t seed) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const uint64_t a = splitmix64(seed ^ splitmix64(2 * (uint64_t)i));
        const uint64_t b = splitmix64(seed ^ splitmix64(2 * (uint64_t)i + 1));
        const double u1 = ((double)(a >> 11) + 0.5) * (1.0 / 9007199254740992.0);
        const double u2 = (double)(b >> 11) * (1.0 / 9007199254740992.0);
        v[i] = sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
    }
}

// U[k][x] = sum_j S[k*m + j] * B[j][x]   (k < nk <= 32)
template <typename TBv, typename TU>
__global__ __launch_bounds__(TB) void k_rotate(const TBv* __restrict__ B, int64_t ldb, int m,
                                               const double* __restrict__ S, int nk, TU* __restrict__ U,
                                               int64_t ldu, int64_t n) {
    extern __shared__ double Ss[];
    for (int i = threadIdx.x; i < nk * m; i += TB) Ss[i] = S[i];
    __syncthreads();
    for (int64_t x = (int64_t)blockIdx.x * TB + threadIdx.x; x < n; x += (int64_t)gridDim.x * TB) {
        double acc[32];
#pragma unroll
        for (int k = 0; k < 32; k++) acc[k] = 0.0;
        for (int j = 0; j < m; j++) {
            const double b = (double)B[(int64_t)j * ldb + x];
#pragma unroll
            for (int k = 0; k < 32; k++)
                if (k < nk) acc[k] += Ss[k * m + j] * b;
        }
#pragma unroll
        for (int k = 0; k < 32; k++)
            if (k < nk) U[(int64_t)k * ldu + x] = (TU)acc[k];
    }
}

// U[k] = sum_j S[k*m + j] B[j] for k < nk (any nk): k_rotate over chunks of outputs whose
// coefficients fit 64 KiB of LDS and 32 accumulators; each output keeps the summation
// order of one pass, so the result does not depend on the chunking.
template <typename TBv, typename TU>
static void rotate_chunks(const TBv* B, int64_t ldb, int m, const double* S, int nk, TU* U, int64_t ldu, int64_t n,
                          hipStream_t st) {
    const int kc = std::max(1, std::min(32, (64 * 1024 / 8) / std::max(1, m)));
    const int gb = (int)std::min<int64_t>(2048, (n + TB - 1) / TB);
    for (int k0 = 0; k0 < nk; k0 += kc) {
        const int kn = std::min(kc, nk - k0);
        hipLaunchKernelGGL((k_rotate<TBv, TU>), dim3(gb), dim3(TB), (size_t)kn * m * 8, st, B, ldb, m,
                           S + (int64_t)k0 * m, kn, U + (int64_t)k0 * ldu, ldu, n);
    }
}

// restart small problem on a stream: Ritz values, Ritz vectors, rotation to tridiagonal form
static void restart_small_on_stream(double* alpha, double* beta, int m, int ls, double* S, double* vals,
                                    double* Y, double* Qrev, double* work, hipStream_t st) {
    hipLaunchKernelGGL(k_ritz, dim3(ls), dim3(TB), 0, st, (const double*)alpha, (const double*)beta, m,
                       (const double*)nullptr, 0, vals, 0, 0.0, (int*)nullptr, (int*)nullptr, (int*)nullptr);
    hipLaunchKernelGGL(k_tri_vecs, dim3(1), dim3(64), 0, st, (const double*)alpha, (const double*)beta, m,
                       (const double*)vals, ls, Y, work);
    hipLaunchKernelGGL(k_restart_q, dim3(1), dim3(64), 0, st, alpha, beta, m, ls, (const double*)vals,
                       (const double*)Y, Qrev, work);
    hipLaunchKernelGGL(k_restart_s, dim3((m + 63) / 64), dim3(64), 0, st, (const double*)Y, (const double*)Qrev,
                       m, ls, S);
}

// ------------------------------------------------------------------ driver
namespace {

// thrown by a mixed-precision solve whose Ritz values disagree with the FP64
// Rayleigh quotients of their Ritz vectors; the solve is then redone with
// reorthogonalisation at every step
struct Inconsistent : std::runtime_error {
    Inconsistent() : std::runtime_error("Ritz values inconsistent with their Rayleigh quotients") {}
};

struct Workspace {
    DBuf basis, wH, w64, part, npart, npart2, h1, h2, alpha, beta, enring, flags, S, swork, U, scal, tvals, Y, Q,
        R, hv64, loc, gate, jy, jz, om, stepdev;
    int* hflag = nullptr;          // host ring of convergence flags (coherent; k_ritz writes it),
                                   // hflag[4]: the projection gate of the step (k_omega writes it)
    hipEvent_t ev[4] = {};         // side stream: a step's convergence flag is in hflag
    hipEvent_t evg = nullptr;      // main stream: k_omega has written hflag[4]
    hipEvent_t evm[4] = {};        // main stream: a step's alpha/beta are written
    hipStream_t side = nullptr;    // Ritz values and convergence test, off the H·v stream
    bool view = false;             // every buffer (hflag included) is a view of device scratch:
                                   // one stream, no events, nothing freed here
    void host_sync_objects() {
        if (hflag || view) return;
        // coherent: k_ritz writes the convergence flags here directly
        SM_HIP(hipHostMalloc((void**)&hflag, 8 * sizeof(int), hipHostMallocCoherent));
        SM_HIP(hipEventCreateWithFlags(&evg, hipEventDisableTiming));
        for (int q = 0; q < 4; q++) SM_HIP(hipEventCreateWithFlags(&ev[q], hipEventDisableTiming));
        for (int q = 0; q < 4; q++) SM_HIP(hipEventCreateWithFlags(&evm[q], hipEventDisableTiming));
        SM_HIP(hipStreamCreateWithFlags(&side, hipStreamNonBlocking));
    }
    int64_t capacity() const {
        int64_t c = 0;
        for (const DBuf* d : {&basis, &wH, &w64, &part, &npart, &npart2, &h1, &h2, &alpha, &beta, &enring, &flags, &S,
                              &swork, &U, &scal, &tvals, &Y, &Q, &R, &hv64, &loc, &gate, &jy, &jz, &om, &stepdev})
            c += (int64_t)d->cap;
        return c;
    }
    ~Workspace() {
        if (side) { (void)hipStreamSynchronize(side); (void)hipStreamDestroy(side); }
        for (int q = 0; q < 4; q++) if (evm[q]) (void)hipEventDestroy(evm[q]);
        for (int q = 0; q < 4; q++) if (ev[q]) (void)hipEventDestroy(ev[q]);
        if (evg) (void)hipEventDestroy(evg);
        if (hflag && !view) (void)hipHostFree(hflag);
    }
};

// Vectors kept at a thick restart: ls, raised to num_ev when ls is below it.  The
// reference keeps ls (lanczos_methods.jl:260-307), and with num_ev above ls the
// states past ls are lost at every restart: the Ritz values then never all settle
// (or, with the stale-value convergence test of diagonalize_T!, the final Ritz
// vectors are asked of a T smaller than num_ev).
static int64_t kept_ls(const SolveParams& p, int64_t n) {
    return std::max<int64_t>(p.ls, std::min<int64_t>(p.num_ev, n));
}

// Every device buffer of a solve over mdim amplitudes with parameters p, in the
// working precision of TBv: f(member, bytes).  setup() allocates exactly these.
template <typename TBv, typename F>
static void workspace_plan(int64_t n, const SolveParams& p, F&& f) {
    const int64_t ld = (n + 63) / 64 * 64;
    const int64_t lm = p.lm, ls = kept_ls(p, n), nev = std::max<int64_t>(1, std::min<int64_t>(p.num_ev, n));
    const int64_t ldp = lm + 2;
    const int64_t Gd = (n + CHUNK - 1) / CHUNK, Gu = (n + UCHUNK - 1) / UCHUNK;
    const int64_t Gx = (n + GsShape<TBv>::SEG - 1) / GsShape<TBv>::SEG;
    const bool mixed = sizeof(TBv) == 4;
    f(&Workspace::basis, lm * ld * (int64_t)sizeof(TBv));
    f(&Workspace::wH, ld * (int64_t)sizeof(TBv));
    f(&Workspace::w64, ld * 8);
    f(&Workspace::part, std::max(Gd, Gx) * ldp * 8);
    f(&Workspace::npart, std::max(Gu, Gx) * 8);
    f(&Workspace::npart2, Gx * 8);
    f(&Workspace::loc, 16 * 8);
    f(&Workspace::om, 3 * (lm + 2) * 8);
    f(&Workspace::gate, 4 * (int64_t)sizeof(int));
    f(&Workspace::h1, (lm + 2) * 8);
    f(&Workspace::h2, (lm + 2) * 8);
    f(&Workspace::alpha, (lm + 1) * 8);
    f(&Workspace::beta, (lm + 1) * 8);
    f(&Workspace::enring, 4 * 64 * 8);
    f(&Workspace::flags, 8 * (int64_t)sizeof(int));   // convergence flags [0,4), k_ritz's ticket [4]
    const int64_t lsw = std::max<int64_t>(32, std::max(ls, nev));   // restart and Ritz small problems
    f(&Workspace::S, lm * lsw * 8);
    f(&Workspace::swork, (lsw * 5 * lm + 4096) * 8);
    f(&Workspace::U, ls * ld * (int64_t)sizeof(TBv));
    f(&Workspace::scal, 64 * 8);
    f(&Workspace::tvals, 64 * 8);
    f(&Workspace::Y, lsw * lm * 8);
    f(&Workspace::Q, lsw * lsw * 8);
    f(&Workspace::R, nev * ld * 8);
    f(&Workspace::hv64, mixed && p.refine ? nev * ld * 8 : 0);   // FP64 H·v of the Ritz vectors
    f(&Workspace::jy, ld * 8);     // J projection scratch (FP64), small next to the basis
    f(&Workspace::jz, ld * 8);
    f(&Workspace::stepdev, (int64_t)sizeof(StepDev));   // captured steps: where the iteration stands
}

template <typename TBv>
struct Solver {
    DeviceHam& H;
    const SolveParams& P;
    SolveStats& st;
    hipStream_t s;
    int64_t n, ld;
    int lm, ls, nev;
    int prec_hv;   // 0 fp64, 1 fp32
    int ldp;       // partial-sum row stride
    int64_t Gd, Gu, Gx;
    int gstore;
    Workspace& W;
    const std::atomic<bool>* cancel = nullptr;
    bool reorth_cgs2 = false;   // SMGPU_REORTH=cgs2: unconditional CGS2 instead of the local step
    bool partial = false;       // partial re-orthogonalisation (omega recurrence)
    bool small = false;         // partial, n <= SMALL_STEP_N: the step is one k_step_small
    double eps_basis = 0.0;     // unit of the basis precision for the omega recurrence
    OmegaParams omega{};
    const JProj* jp = nullptr;  // J projection (doubleLanczos), or none

    static Workspace& cached(std::shared_ptr<void>* slots) {
        const int k = sizeof(TBv) == 4 ? 1 : 0;
        if (!slots[k]) slots[k] = std::make_shared<Workspace>();
        return *static_cast<Workspace*>(slots[k].get());
    }

    // ctx: stream + workspace slots of the caller (a serving worker), or the handle's own
    Solver(DeviceHam& h, const SolveParams& p, SolveStats& stats, const SolveCtx* ctx, const JProj* jproj = nullptr)
        : Solver(h, p, stats, cached(ctx ? ctx->ws : h.trl_ws), ctx ? ctx->stream : h.stream, jproj) {
        cancel = ctx ? ctx->cancel : nullptr;
    }
    // an explicit workspace and stream (a view workspace runs everything on `stream`)
    Solver(DeviceHam& h, const SolveParams& p, SolveStats& stats, Workspace& ws, hipStream_t stream,
           const JProj* jproj = nullptr)
        : H(h), P(p), st(stats), W(ws), jp(jproj) {
        s = stream;
        if (const char* e = std::getenv("SMGPU_REORTH")) reorth_cgs2 = std::string(e) == "cgs2";
        partial = !reorth_cgs2 && !jp &&
                  (P.reorth == SMGPU_REORTH_PARTIAL || (P.reorth == SMGPU_REORTH_DEFAULT && sizeof(TBv) == 4));
        eps_basis = sizeof(TBv) == 4 ? 1.1920928955078125e-07 : 2.220446049250313e-16;
        // The omega recurrence's error term models the rounding of one Lanczos step.
        // With an FP32 basis the step's H·v is FP32 too, and a row of it sums hundreds
        // of FP32 products (error ~ sqrt(terms) eps), so the term is 10 eps: with eps
        // alone the estimate fell behind the real loss of orthogonality on small and
        // medium bases (10 of 428 requests of the serving stream broke down after
        // ~100 steps, eigenvalues 15-35 MeV off); with 10 eps none did, for 44 instead
        // of 24 projection passes in a Cr48 solve.  The FP64 basis keeps eps.
        // SMGPU_OMEGA_ROUNDING overrides the FP32 multiplier (tests use 1 to drive
        // the Rayleigh-quotient check and its fallback).
        double kr = 10.0;
        if (const char* e = std::getenv("SMGPU_OMEGA_ROUNDING")) kr = std::atof(e);
        omega.eps = eps_basis;
        omega.eps_r = (sizeof(TBv) == 4 ? kr : 1.0) * eps_basis;
        omega.thr = std::sqrt(eps_basis);
        n = H.mdim;
        ld = (n + 63) / 64 * 64;
        lm = P.lm;
        ls = (int)kept_ls(P, n);
        nev = (int)std::min<int64_t>(P.num_ev, n);
        prec_hv = sizeof(TBv) == 4 ? 1 : 0;
        ldp = lm + 2;
        Gd = (n + CHUNK - 1) / CHUNK;
        Gu = (n + UCHUNK - 1) / UCHUNK;
        Gx = (n + GsShape<TBv>::SEG - 1) / GsShape<TBv>::SEG;
        gstore = (int)std::min<int64_t>(4096, (n + TB - 1) / TB);
        small = partial && n <= SMALL_STEP_N && lm + 1 <= 1024;
    }

    TBv* V(int j) { return W.basis.as<TBv>() + (int64_t)j * ld; }
    double* en(int step) { return W.enring.as<double>() + (int64_t)((step + 1) & 3) * 64; }
    int* flag(int step) { return W.flags.as<int>() + ((step + 1) & 3); }


    // CGS2 of `win` against V[0..nv-1] into w64 (classical Gram-Schmidt twice):
    //   h1 = V^T win;  w1 = win - V h1;  h2 = V^T w1;  w64 = w1 - V h2
    // in three streaming passes (the middle one computes w1 and the h2 partials).
    // W.h1 keeps the first-pass coefficients, W.npart the |w64|^2 partials.
    template <typename TW>
    void cgs2(int nv, const TW* win) {
        const TBv* B = V(0);
        double* part = W.part.as<double>();
        hipLaunchKernelGGL((k_gs_dots<TBv, TW, false>), dim3(Gx), dim3(TB), 0, s, B, ld, nv, win, n, part, ldp,
                           (const int*)nullptr, (const TBv*)nullptr, (const TBv*)nullptr, (const double*)nullptr,
                           (const double*)nullptr, (double*)nullptr, (double*)nullptr);
        hipLaunchKernelGGL(k_reduce, dim3(nv), dim3(TB), 0, s, part, (int)Gx, ldp, W.h1.as<double>(),
                           (const int*)nullptr);
        hipLaunchKernelGGL((k_gs_upd<TBv, TW, true>), dim3(Gx), dim3(TB), 0, s, B, ld, nv, W.h1.as<double>(), win,
                           W.w64.as<double>(), n, part, ldp, W.npart.as<double>(), (const int*)nullptr);
        hipLaunchKernelGGL(k_reduce, dim3(nv), dim3(TB), 0, s, part, (int)Gx, ldp, W.h2.as<double>(),
                           (const int*)nullptr);
        hipLaunchKernelGGL((k_gs_upd<TBv, double, false>), dim3(Gx), dim3(TB), 0, s, B, ld, nv, W.h2.as<double>(),
                           W.w64.as<double>(), W.w64.as<double>(), n, part, ldp, W.npart.as<double>(),
                           (const int*)nullptr);
    }

    // Reorthogonalisation of w = H v_it against V[0..it] (full, as ReORTH,
    // shellmodel_main.jl:1005-1012), in the form of a Lanczos step:
    //   alpha = v_it . w;  w' = w - alpha v_it - beta_{it-1} v_{it-1}   (local step, FP64)
    //   one CGS pass  w'' = w' - V (V^T w')
    //   a second pass only if |w''| < |w'|/sqrt(2) (decided on the device).
    // After the local step the projections are of rounding size, so one pass
    // leaves w'' orthogonal to the stored basis to second order; the test
    // catches the iterations where it is not (the first after a restart).
    // Two streaming reads of V in the common case (CGS2 needs three).
    // W.h1[0..it] = projections of the pass, alpha/beta written by k_finish_local.
    template <typename TW>
    void reorth_local(int it, const TW* win) {
        const int nv = it + 1;
        const TBv* B = V(0);
        double* part = W.part.as<double>();
        double* loc = W.loc.as<double>();
        int* gate = W.gate.as<int>();
        const TBv* vp = it > 0 ? (const TBv*)V(it - 1) : nullptr;
        const double* bp = it > 0 ? W.beta.as<double>() + (it - 1) : nullptr;
        const int G2 = (int)std::min<int64_t>(1024, Gd);
        hipLaunchKernelGGL((k_dot2<TBv, TW>), dim3(G2), dim3(TB), 0, s, (const TBv*)V(it), win, n, part, ldp);
        hipLaunchKernelGGL(k_reduce, dim3(1), dim3(TB), 0, s, part, G2, ldp, loc, (const int*)nullptr);
        hipLaunchKernelGGL((k_gs_dots<TBv, TW, true>), dim3(Gx), dim3(TB), 0, s, B, ld, nv, win, n, part, ldp,
                           (const int*)nullptr, (const TBv*)V(it), vp, (const double*)loc, bp, W.w64.as<double>(),
                           W.npart2.as<double>());
        hipLaunchKernelGGL(k_reduce, dim3(nv), dim3(TB), 0, s, part, (int)Gx, ldp, W.h1.as<double>(),
                           (const int*)nullptr);
        hipLaunchKernelGGL((k_gs_upd<TBv, double, false>), dim3(Gx), dim3(TB), 0, s, B, ld, nv, W.h1.as<double>(),
                           W.w64.as<double>(), W.w64.as<double>(), n, part, ldp, W.npart.as<double>(),
                           (const int*)nullptr);
        hipLaunchKernelGGL(k_dgks, dim3(1), dim3(TB), 0, s, (const double*)W.npart.as<double>(),
                           (const double*)W.npart2.as<double>(), (int)Gx, loc, gate, (const int*)nullptr);
        hipLaunchKernelGGL((k_gs_dots<TBv, double, false>), dim3(Gx), dim3(TB), 0, s, B, ld, nv,
                           (const double*)W.w64.as<double>(), n, part, ldp, (const int*)gate, (const TBv*)nullptr,
                           (const TBv*)nullptr, (const double*)nullptr, (const double*)nullptr, (double*)nullptr,
                           (double*)nullptr);
        hipLaunchKernelGGL(k_reduce, dim3(nv), dim3(TB), 0, s, part, (int)Gx, ldp, W.h2.as<double>(),
                           (const int*)gate);
        hipLaunchKernelGGL((k_gs_upd<TBv, double, false>), dim3(Gx), dim3(TB), 0, s, B, ld, nv, W.h2.as<double>(),
                           W.w64.as<double>(), W.w64.as<double>(), n, part, ldp, W.npart.as<double>(),
                           (const int*)gate);
        hipLaunchKernelGGL(k_finish_local, dim3(1), dim3(TB), 0, s, loc, it,
                           (const double*)W.npart.as<double>(), (int)Gx, W.alpha.as<double>(), W.beta.as<double>());
    }

    // Lanczos step with partial re-orthogonalisation: the local step, then the omega
    // recurrence decides on the device whether this step runs one CGS pass against
    // V[0..it] (plus the DGKS second pass when that removes more than half of |w|^2).
    // This half enqueues the local step and k_omega, which also hands its gate to the
    // host; reorth_partial_end enqueues the rest.  (Small bases: all of it in k_step_small.)
    template <typename TW>
    void reorth_partial_begin(int it, const TW* win) {
        const TBv* B = V(0);
        double* part = W.part.as<double>();
        double* loc = W.loc.as<double>();
        int* gate = W.gate.as<int>();
        const TBv* vp = it > 0 ? (const TBv*)V(it - 1) : nullptr;
        const double* bp = it > 0 ? W.beta.as<double>() + (it - 1) : nullptr;
        if (small) {
            hipLaunchKernelGGL((k_step_small<TBv, TW>), dim3(1), dim3(TB), 0, s, win, (const TBv*)B, ld, it, n,
                               W.w64.as<double>(), loc, gate, W.alpha.as<double>(), W.beta.as<double>(),
                               W.om.as<double>(), lm + 2, omega, W.h1.as<double>(), W.h2.as<double>(),
                               it + 1 < lm ? V(it + 1) : (TBv*)nullptr);
            return;
        }
        const int G2 = (int)std::min<int64_t>(1024, Gd);
        hipLaunchKernelGGL((k_dot2<TBv, TW>), dim3(G2), dim3(TB), 0, s, (const TBv*)V(it), win, n, part, ldp);
        hipLaunchKernelGGL(k_reduce, dim3(1), dim3(TB), 0, s, part, G2, ldp, loc, (const int*)nullptr);
        hipLaunchKernelGGL((k_local_step<TBv, TW>), dim3(Gx), dim3(TB), 0, s, win, (const TBv*)V(it), vp,
                           (const double*)loc, bp, W.w64.as<double>(), n, W.npart.as<double>(), W.npart2.as<double>());
        hipLaunchKernelGGL(k_omega, dim3(1), dim3(TB), 0, s, W.alpha.as<double>(),
                           W.beta.as<double>(), it, (const double*)W.npart.as<double>(), (int)Gx,
                           W.om.as<double>(), lm + 2, loc, gate, omega, d_hg ? W.hflag + 4 : nullptr);
        if (d_hg) SM_HIP(hipEventRecord(W.evg, s));
    }

    // The rest of the step: the projection pass and its DGKS second pass (both gated on
    // the device), then alpha/beta.  A batch member in a round of one Hamiltonian (d_hg)
    // first waits for the gate of k_omega and launches these kernels only on steps that
    // run the pass: on the other steps (about nine in ten) they do nothing, or rewrite
    // what k_omega wrote, so leaving them out changes no result and saves eight
    // dependent dispatches, each of which waits for a free CU slot behind the other
    // solves' H·v; the wait itself overlaps its mates' steps and the other H·v.  A solve
    // alone, and a round of interaction variants, keeps them queued behind k_omega
    // instead, so the host stays a step ahead of the device.
    void reorth_partial_end(int it) {
        if (small) return;
        const int nv = it + 1;
        const TBv* B = V(0);
        double* part = W.part.as<double>();
        double* loc = W.loc.as<double>();
        int* gate = W.gate.as<int>();
        bool run = true;
        if (d_hg) {
            SM_HIP(hipEventSynchronize(W.evg));
            run = static_cast<volatile int*>(W.hflag)[4] != 0;
        }
        if (run) {
            const int* g1 = gate + 1;
            hipLaunchKernelGGL((k_gs_dots<TBv, double, false>), dim3(Gx), dim3(TB), 0, s, B, ld, nv,
                               (const double*)W.w64.as<double>(), n, part, ldp, g1, (const TBv*)nullptr,
                               (const TBv*)nullptr, (const double*)nullptr, (const double*)nullptr, (double*)nullptr,
                               (double*)nullptr);
            hipLaunchKernelGGL(k_reduce, dim3(nv), dim3(TB), 0, s, part, (int)Gx, ldp, W.h1.as<double>(), g1);
            hipLaunchKernelGGL((k_gs_upd<TBv, double, false>), dim3(Gx), dim3(TB), 0, s, B, ld, nv,
                               W.h1.as<double>(), W.w64.as<double>(), W.w64.as<double>(), n, part, ldp,
                               W.npart.as<double>(), g1);
            hipLaunchKernelGGL(k_dgks, dim3(1), dim3(TB), 0, s, (const double*)W.npart.as<double>(),
                               (const double*)W.npart2.as<double>(), (int)Gx, loc, gate, g1);
            hipLaunchKernelGGL((k_gs_dots<TBv, double, false>), dim3(Gx), dim3(TB), 0, s, B, ld, nv,
                               (const double*)W.w64.as<double>(), n, part, ldp, (const int*)gate, (const TBv*)nullptr,
                               (const TBv*)nullptr, (const double*)nullptr, (const double*)nullptr, (double*)nullptr,
                               (double*)nullptr);
            hipLaunchKernelGGL(k_reduce, dim3(nv), dim3(TB), 0, s, part, (int)Gx, ldp, W.h2.as<double>(),
                               (const int*)gate);
            hipLaunchKernelGGL((k_gs_upd<TBv, double, false>), dim3(Gx), dim3(TB), 0, s, B, ld, nv,
                               W.h2.as<double>(), W.w64.as<double>(), W.w64.as<double>(), n, part, ldp,
                               W.npart.as<double>(), (const int*)gate);
            hipLaunchKernelGGL(k_finish_local, dim3(1), dim3(TB), 0, s, loc, it,
                               (const double*)W.npart.as<double>(), (int)Gx, W.alpha.as<double>(),
                               W.beta.as<double>());
        }
        // (no pass: k_omega has written alpha[it] and beta[it])
    }

    // the first half of the next advance() (partial path): a batch driver calls it for
    // every member before their advance() calls, so their k_omega waits overlap
    bool d_pre = false;
    bool d_hg = false;          // this step reads the gate on the host (set by pre_advance)
    void partial_begin(int it) {
        if (prec_hv == 0) reorth_partial_begin<double>(it, W.wH.as<double>());
        else reorth_partial_begin<float>(it, W.wH.as<float>());
    }
    void pre_advance(bool host_gate) {
        if (d_pre || reorth_cgs2 || !partial) return;
        d_hg = host_gate;
        partial_begin(d_it);
        d_pre = true;
    }

    void enqueue_iteration(int it, int step, int nprev) {
        launch_hv(H, prec_hv, V(it), W.wH.p, s);
        after_hv(it, step, nprev);
    }

    // the rest of a Lanczos step once w = H v_it is in W.wH
    void after_hv(int it, int step, int nprev) {
        st.n_hv++;
        if (reorth_cgs2) {
            const int nv = it + 1;
            if (prec_hv == 0) cgs2<double>(nv, W.wH.as<double>());
            else cgs2<float>(nv, W.wH.as<float>());
            hipLaunchKernelGGL(k_finish, dim3(1), dim3(TB), 0, s, W.h1.as<double>(), it, W.npart.as<double>(),
                               (int)Gx, W.alpha.as<double>(), W.beta.as<double>());
        } else if (partial) {
            if (!d_pre) { d_hg = false; partial_begin(it); }
            d_pre = false;
            reorth_partial_end(it);
        } else {
            if (prec_hv == 0) reorth_local<double>(it, W.wH.as<double>());
            else reorth_local<float>(it, W.wH.as<float>());
        }
        // Ritz values of T[0..it] and diagonalize_T!'s test run off the path to the next
        // H·v (later steps write other entries of alpha/beta; a restart starts after the
        // host has waited for the last test): alone, on the side stream once this step's
        // alpha/beta exist; in a batch, on the solve's own stream after step_ready, the
        // event the batch's next H·v waits for, so they overlap that H·v.
        const int nn = std::min(nev, it + 1);
        hipStream_t rs = W.side;
        if (step_ready) {
            if (it + 1 < lm && !small)      // (k_step_small stores v_{it+1} itself)
                hipLaunchKernelGGL((k_store<TBv>), dim3(gstore), dim3(TB), 0, s, W.w64.as<double>(),
                                   W.beta.as<double>() + it, V(it + 1), n);
            SM_HIP(hipEventRecord(step_ready, s));
            rs = s;
        } else if (W.view) {
            rs = s;                          // one stream: the test runs in order after the step
        } else {
            SM_HIP(hipEventRecord(W.evm[(step + 1) & 3], s));
            SM_HIP(hipStreamWaitEvent(W.side, W.evm[(step + 1) & 3], 0));
        }
        // Ritz values and the convergence test; the flag also lands in the host ring
        hipLaunchKernelGGL(k_ritz, dim3(nn), dim3(TB), 0, rs, W.alpha.as<double>(), W.beta.as<double>(), it + 1,
                           en(step - 1), nprev, en(step), nev, P.tol, flag(step), W.hflag + ((step + 1) & 3),
                           W.flags.as<int>() + 4);
        if (!step_ready && it + 1 < lm && !small)
            hipLaunchKernelGGL((k_store<TBv>), dim3(gstore), dim3(TB), 0, s, W.w64.as<double>(),
                               W.beta.as<double>() + it, V(it + 1), n);
    }

    void setup(const double* v0_in) {
        // grow-only buffers cached on the handle (no per-solve hipMalloc / memset)
        workspace_plan<TBv>(n, P, [&](DBuf Workspace::*m, int64_t b) { (W.*m).ensure((size_t)b); });
        W.host_sync_objects();
        if (W.side) SM_HIP(hipStreamSynchronize(W.side));     // a canceled solve may have left tests queued
        // host sources of the asynchronous copies below live until the stream
        // synchronisation at the end of setup()
        std::vector<double> e0(4 * 64, 1.0e4);
        const double one = 1.0;              // om_{0,0}
        SM_HIP(hipMemcpyAsync(W.enring.p, e0.data(), e0.size() * 8, hipMemcpyHostToDevice, s));
        SM_HIP(hipMemsetAsync(W.loc.p, 0, W.loc.bytes, s));
        SM_HIP(hipMemsetAsync(W.gate.p, 0, W.gate.bytes, s));
        SM_HIP(hipMemsetAsync(W.om.p, 0, W.om.bytes, s));
        SM_HIP(hipMemsetAsync(W.flags.p, 0, W.flags.bytes, s));     // flags and k_ritz's ticket
        SM_HIP(hipMemcpyAsync(W.om.p, &one, 8, hipMemcpyHostToDevice, s));
        // start vector: caller's, or seeded normal deviates generated on the device
        // (initialize_wf, :939-948); normalised in FP64
        if (v0_in) {
            SM_HIP(hipMemcpyAsync(W.w64.p, v0_in, n * 8, hipMemcpyHostToDevice, s));
        } else {
            hipLaunchKernelGGL(k_randn, dim3(gstore), dim3(TB), 0, s, W.w64.as<double>(), n, (uint64_t)P.seed);
        }
        hipLaunchKernelGGL((k_dotsw<double, double>), dim3(Gd), dim3(TB), 0, s, (const double*)nullptr, ld, 0,
                           (const double*)W.w64.as<double>(), n, W.part.as<double>(), ldp, 1, (const int*)nullptr);
        hipLaunchKernelGGL(k_reduce, dim3(1), dim3(TB), 0, s, W.part.as<double>(), (int)Gd, ldp,
                           W.scal.as<double>() + 1, (const int*)nullptr);
        hipLaunchKernelGGL(k_sqrt1, dim3(1), dim3(1), 0, s, W.scal.as<double>() + 1, W.scal.as<double>());
        hipLaunchKernelGGL((k_store<TBv>), dim3(gstore), dim3(TB), 0, s, W.w64.as<double>(), W.scal.as<double>(),
                           V(0), n);
        SM_HIP(hipStreamSynchronize(s));
        double nr = 0.0;
        SM_HIP(hipMemcpy(&nr, W.scal.p, 8, hipMemcpyDeviceToHost));
        if (!(nr > 0.0) || !std::isfinite(nr)) throw std::invalid_argument("start vector is zero or not finite");
    }

    // nsteps plain Lanczos steps (no convergence test, no restart); alpha/beta out
    void run_steps(int nsteps, const double* v0_in, double* alpha_out, double* beta_out) {
        if (nsteps < 1 || nsteps >= lm) throw std::invalid_argument("need 1 <= nsteps < lm");
        setup(v0_in);
        for (int it = 0; it < nsteps; it++) enqueue_iteration(it, it, it == 0 ? 0 : std::min(nev, it));
        SM_HIP(hipStreamSynchronize(s));
        if (W.side) SM_HIP(hipStreamSynchronize(W.side));
        SM_HIP(hipMemcpy(alpha_out, W.alpha.p, nsteps * 8, hipMemcpyDeviceToHost));
        SM_HIP(hipMemcpy(beta_out, W.beta.p, nsteps * 8, hipMemcpyDeviceToHost));
    }

    // ---- driver: start(), then per step H·v of hv_in() into hv_out() on the solve's
    // stream followed by advance(), until advance() reports the end; then end().
    // run() is the one-solve loop; a batch driver issues the H·v of several solves
    // on one Hamiltonian in shared launches between their advance() calls.
    int d_k = 0, d_it = 0, d_step = 0, d_first = 0, d_elit = 1, d_conv_step = 0;
    int64_t d_total = 0;
    bool d_conv = false, d_after_restart = false;
    std::chrono::steady_clock::time_point d_t0;

    void start(const double* v0_in) {
        d_t0 = std::chrono::steady_clock::now();
        if (lm < 3 || ls < 1 || ls > 64 || ls >= lm - 1 || nev < 1 || nev > 64)
            throw std::invalid_argument("need 1 <= num_ev <= 64, 1 <= max(ls, num_ev) <= 64 and max(ls, num_ev) < lm - 1");
        if (lm > 1000) throw std::invalid_argument("lm above 1000 is not supported");
        setup(v0_in);
        if (jp) jfilter(0, 0, true);
        d_k = 0; d_it = 0; d_step = 0; d_first = 0; d_elit = 1; d_conv_step = 0;
        d_total = 0; d_conv = false; d_after_restart = false;
        st.n_hv = 0; st.n_restart = 0;
        begin_step();
    }

    // checks before the H·v of step d_step (iteration d_it)
    void begin_step() {
        if (++d_total > P.max_iter) throw std::runtime_error("TRL: iteration limit reached");
        if (cancel && cancel->load(std::memory_order_relaxed)) throw Canceled();
    }

    // set by a batch driver: recorded on the solve's stream when the next H·v's input
    // vector is ready (the Ritz test then runs on the solve's stream behind it)
    hipEvent_t step_ready = nullptr;

    const TBv* hv_in() { return V(d_it); }
    void* hv_out() { return W.wH.p; }
    hipStream_t stream() const { return s; }

    bool check(int stp) {
        SM_HIP(hipEventSynchronize(W.ev[(stp + 1) & 3]));
        return W.hflag[(stp + 1) & 3] != 0;
    }

    // after the H·v of the current step: enqueue the rest of it, run the lag-1
    // convergence test, restart at the end of a sweep.  true when the solve is done.
    bool advance() {
        const int it = d_it, step = d_step, k = d_k;
        // interlacing brackets need T[0..it-1] to be the previous step's matrix
        const int nprev = (it == k && d_after_restart) || it == 0 ? 0 : std::min(nev, it);
        after_hv(it, step, nprev);
        hipStream_t rs = step_ready ? s : W.side;
        SM_HIP(hipEventRecord(W.ev[(step + 1) & 3], rs));       // k_ritz has written the flag
        bool done = false;
        bool moved = false;     // the next input vector changed after step_ready
        if (jp && it + 1 < lm && it + 1 < n) {
            // the J subspace is exhausted when beta vanishes (TRL: tbeta < Jtol stops)
            double b = 0.0;
            SM_HIP(hipMemcpyAsync(&b, W.beta.as<double>() + it, 8, hipMemcpyDeviceToHost, s));
            SM_HIP(hipStreamSynchronize(s));
            // ... or when the projection of the next vector keeps almost nothing of
            // it: the residual was rounding noise outside the (invariant) J space
            if (b < jp->tol || jfilter(it + 1, it + 1, false) < J_EXHAUSTED) {
                d_conv_step = step; d_elit = it + 1; done = true;
            }
            moved = true;
        }
        if (!done && it + 1 == n) {                         // Krylov space exhausted (it == mdim)
            if (step > d_first && check(step - 1)) { d_conv_step = step - 1; d_elit = it; }
            else { d_conv_step = step; d_elit = it + 1; }
            done = true;
        }
        if (!done && step > d_first && check(step - 1)) {
            d_conv_step = step - 1; d_elit = it; done = true;
        }
        d_it = it + 1;
        d_step = step + 1;
        if (!done && d_it == lm - 1) {
            // last iteration of the sweep: its flag decides between stop and restart
            if (check(d_step - 1)) {
                d_conv_step = d_step - 1; d_elit = lm - 1; done = true;
            } else {
                restart();
                st.n_restart++;
                d_k = ls;
                d_it = ls;
                d_first = d_step;
                d_after_restart = true;
                moved = true;
            }
        }
        d_conv = done;
        if (moved && step_ready && !done) SM_HIP(hipEventRecord(step_ready, s));
        if (!done) begin_step();
        return done;
    }

    void end(double* evals, double* evecs) {
        SM_HIP(hipStreamSynchronize(s));
        SM_HIP(hipStreamSynchronize(W.side));
        const int elit = d_elit, conv_step = d_conv_step;
        if (elit < nev) throw std::invalid_argument("the Krylov space holds fewer states than num_ev (J space too small)");
        st.iterations = elit;
        std::vector<double> en_last(nev);
        SM_HIP(hipMemcpy(en_last.data(), en(conv_step), nev * 8, hipMemcpyDeviceToHost));
        double lc[6] = {};
        SM_HIP(hipMemcpy(lc, W.loc.p, 6 * 8, hipMemcpyDeviceToHost));
        st.n_reorth2 = (int32_t)lc[3];
        st.n_reorth = partial ? (int32_t)lc[5] : st.n_hv;
        for (int i = 0; i < nev; i++) { evals[i] = en_last[i]; st.ritz[i] = en_last[i]; }
        finish(elit, conv_step, evals, evecs);
        st.wall_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - d_t0).count();
    }

    void run(double* evals, double* evecs, const double* v0_in = nullptr) {
        if (graph_steps()) {
            run_graph(evals, evecs, v0_in);
            return;
        }
        start(v0_in);
        do {
            launch_hv(H, prec_hv, hv_in(), hv_out(), s);
        } while (!advance());
        end(evals, evecs);
    }

    // ---- captured steps: one Lanczos step (H·v, the step after it, the Ritz values
    // and convergence test, the move to the next step) captured once as a hipGraph
    // and replayed for every step; the kernels read the iteration from W.stepdev.
    // The host keeps the lag-1 test and runs restarts eagerly, then resets the step
    // state.  Every replay runs the kernels of the launched step with the same
    // arguments, so the results are bit-identical to the launched steps.
    hipGraph_t g_graph = nullptr;
    hipGraphExec_t g_exec = nullptr;
    ~Solver() {
        if (g_exec) (void)hipGraphExecDestroy(g_exec);
        if (g_graph) (void)hipGraphDestroy(g_graph);
    }

    bool graph_steps() const {
        if (jp || reorth_cgs2 || step_ready) return false;
        int mode = P.steps;
        if (const char* e = std::getenv("SMGPU_STEPS")) {
            const std::string v(e);
            if (v == "graph") mode = SMGPU_STEPS_GRAPH;
            else if (v == "launch") mode = SMGPU_STEPS_LAUNCH;
        }
        if (mode == SMGPU_STEPS_GRAPH) return true;
        if (mode == SMGPU_STEPS_LAUNCH) return false;
        return n <= GRAPH_AUTO_N;
    }

    StepDev* sdev() { return W.stepdev.as<StepDev>(); }

    template <typename TW>
    void chain_s(const TW* win) {
        StepDev* sd = sdev();
        TBv* B = V(0);
        double* part = W.part.as<double>();
        double* loc = W.loc.as<double>();
        int* gate = W.gate.as<int>();
        double* al = W.alpha.as<double>();
        double* be = W.beta.as<double>();
        double* w64 = W.w64.as<double>();
        if (small) {
            hipLaunchKernelGGL((k_step_small_s<TBv, TW>), dim3(1), dim3(TB), 0, s, sd, win, B, ld, n, w64, loc, gate, al,
                               be, W.om.as<double>(), lm + 2, omega, W.h1.as<double>(), W.h2.as<double>());
            return;
        }
        const int G2 = (int)std::min<int64_t>(1024, Gd);
        hipLaunchKernelGGL((k_dot2_s<TBv, TW>), dim3(G2), dim3(TB), 0, s, sd, (const TBv*)B, ld, win, n, part, ldp);
        hipLaunchKernelGGL(k_reduce, dim3(1), dim3(TB), 0, s, part, G2, ldp, loc, (const int*)nullptr);
        const int* g1 = partial ? gate + 1 : nullptr;
        if (partial) {
            hipLaunchKernelGGL((k_local_step_s<TBv, TW>), dim3(Gx), dim3(TB), 0, s, sd, win, (const TBv*)B, ld,
                               (const double*)loc, (const double*)be, w64, n, W.npart.as<double>(),
                               W.npart2.as<double>());
            hipLaunchKernelGGL(k_omega_s, dim3(1), dim3(TB), 0, s, sd, al, be, (const double*)W.npart.as<double>(),
                               (int)Gx, W.om.as<double>(), lm + 2, loc, gate, omega);
            hipLaunchKernelGGL((k_gs_dots_s<TBv, double, false>), dim3(Gx), dim3(TB), 0, s, sd, (const TBv*)B, ld,
                               (const double*)w64, n, part, ldp, g1, (const double*)nullptr, (const double*)nullptr,
                               (double*)nullptr, (double*)nullptr);
        } else {
            hipLaunchKernelGGL((k_gs_dots_s<TBv, TW, true>), dim3(Gx), dim3(TB), 0, s, sd, (const TBv*)B, ld, win, n,
                               part, ldp, (const int*)nullptr, (const double*)loc, (const double*)be, w64,
                               W.npart2.as<double>());
        }
        hipLaunchKernelGGL(k_reduce_s, dim3(lm), dim3(TB), 0, s, sd, part, (int)Gx, ldp, W.h1.as<double>(), g1);
        hipLaunchKernelGGL((k_gs_upd_s<TBv, double, false>), dim3(Gx), dim3(TB), 0, s, sd, (const TBv*)B, ld,
                           (const double*)W.h1.as<double>(), (const double*)w64, w64, n, part, ldp,
                           W.npart.as<double>(), g1);
        hipLaunchKernelGGL(k_dgks, dim3(1), dim3(TB), 0, s, (const double*)W.npart.as<double>(),
                           (const double*)W.npart2.as<double>(), (int)Gx, loc, gate, g1);
        hipLaunchKernelGGL((k_gs_dots_s<TBv, double, false>), dim3(Gx), dim3(TB), 0, s, sd, (const TBv*)B, ld,
                           (const double*)w64, n, part, ldp, (const int*)gate, (const double*)nullptr,
                           (const double*)nullptr, (double*)nullptr, (double*)nullptr);
        hipLaunchKernelGGL(k_reduce_s, dim3(lm), dim3(TB), 0, s, sd, part, (int)Gx, ldp, W.h2.as<double>(),
                           (const int*)gate);
        hipLaunchKernelGGL((k_gs_upd_s<TBv, double, false>), dim3(Gx), dim3(TB), 0, s, sd, (const TBv*)B, ld,
                           (const double*)W.h2.as<double>(), (const double*)w64, w64, n, part, ldp,
                           W.npart.as<double>(), (const int*)gate);
        hipLaunchKernelGGL(k_finish_local_s, dim3(1), dim3(TB), 0, s, sd, loc, (const double*)W.npart.as<double>(),
                           (int)Gx, al, be);
    }

    // the Ritz values and convergence test of the step in sd->prev (on stream rs)
    void ritz_prev(hipStream_t rs) {
        hipLaunchKernelGGL(k_ritz_s, dim3(nev), dim3(TB), 0, rs, (const StepDev*)sdev(),
                           (const double*)W.alpha.as<double>(), (const double*)W.beta.as<double>(),
                           W.enring.as<double>(), nev, P.tol, W.flags.as<int>(), W.hflag, W.flags.as<int>() + 4);
    }

    // Graph of one step: the previous step's Ritz values and convergence test on a
    // branch beside this step's H·v (as launched steps run them on the side stream),
    // then the step after H·v, the store of v_{it+1}, and the move to the next step.
    void capture_step() {
        StepDev* sd = sdev();
        SM_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        try {
            SM_HIP(hipEventRecord(W.evm[0], s));
            SM_HIP(hipStreamWaitEvent(W.side, W.evm[0], 0));
            ritz_prev(W.side);
            SM_HIP(hipEventRecord(W.evm[1], W.side));
            launch_hv_column(H, prec_hv, V(0), ld, &sd->it, W.wH.p, s);
            if (prec_hv == 0) chain_s<double>(W.wH.as<double>());
            else chain_s<float>(W.wH.as<float>());
            if (!small)
                hipLaunchKernelGGL((k_store_s<TBv>), dim3(gstore), dim3(TB), 0, s, sd, (const double*)W.w64.as<double>(),
                                   (const double*)W.beta.as<double>(), V(0), ld, n);
            SM_HIP(hipStreamWaitEvent(s, W.evm[1], 0));
            hipLaunchKernelGGL(k_step_next, dim3(1), dim3(1), 0, s, sd, nev, lm);
        } catch (...) {
            hipGraph_t g = nullptr;
            (void)hipStreamEndCapture(s, &g);
            if (g) (void)hipGraphDestroy(g);
            throw;
        }
        SM_HIP(hipStreamEndCapture(s, &g_graph));
        SM_HIP(hipGraphInstantiate(&g_exec, g_graph, nullptr, nullptr, 0));
    }

    void step_set(int it, int step, int nprev) {
        hipLaunchKernelGGL(k_step_set, dim3(1), dim3(1), 0, s, sdev(), it, step, nprev, nev, lm);
    }

    // The graph of step k holds the convergence test of step k-1, so the host tests
    // step k-2 after launching step k (one step further behind than launched steps,
    // which test k-1 after enqueueing k): the device always has the next step
    // queued.  At the end of a sweep or of the Krylov space the tests of the last two
    // steps are taken there, with the same outcome as launched steps; a solve that
    // converges runs one more speculative step, which only writes entries past the
    // converged ones.
    void run_graph(double* evals, double* evecs, const double* v0_in) {
        start(v0_in);
        W.host_sync_objects();
        if (!g_exec) capture_step();
        step_set(d_it, d_step, 0);
        for (;;) {
            const int it = d_it, step = d_step;
            SM_HIP(hipGraphLaunch(g_exec, s));
            st.n_hv++;
            SM_HIP(hipEventRecord(W.ev[step & 3], s));          // the test of step - 1 is done
            bool done = false, restarted = false;
            if (step - 2 >= d_first && check(step - 2)) {
                d_conv_step = step - 2; d_elit = it - 1; done = true;
            }
            const bool exhausted = it + 1 == n, sweep_end = it + 2 == lm;
            if (!done && (exhausted || sweep_end)) {
                ritz_prev(s);                                     // this step's test, now
                SM_HIP(hipEventRecord(W.ev[(step + 1) & 3], s));
                if (step - 1 >= d_first && check(step - 1)) {
                    d_conv_step = step - 1; d_elit = it; done = true;
                } else if (exhausted) {
                    d_conv_step = step; d_elit = it + 1; done = true;
                } else if (check(step)) {
                    d_conv_step = step; d_elit = lm - 1; done = true;
                } else {
                    restart();
                    st.n_restart++;
                    d_k = ls;
                    d_first = step + 1;
                    d_after_restart = true;
                    restarted = true;
                    step_set(ls, step + 1, 0);
                }
            }
            d_conv = done;
            if (done) break;
            d_it = restarted ? ls : it + 1;
            d_step = step + 1;
            begin_step();
        }
        end(evals, evecs);
    }

    // <y|J^2|y> / <y|y> with z = J^2 y (synchronous)
    double j_expect(const double* y, double* z) {
        launch_hv(*jp->j2, 0, y, z, s);
        const int G2 = (int)std::min<int64_t>(1024, Gd);
        double* part = W.part.as<double>();
        double* sc = W.scal.as<double>();
        hipLaunchKernelGGL((k_dot2<double, double>), dim3(G2), dim3(TB), 0, s, y, (const double*)z, n, part, ldp);
        hipLaunchKernelGGL(k_reduce, dim3(1), dim3(TB), 0, s, part, G2, ldp, sc + 60, (const int*)nullptr);
        hipLaunchKernelGGL((k_dot2<double, double>), dim3(G2), dim3(TB), 0, s, y, y, n, part, ldp);
        hipLaunchKernelGGL(k_reduce, dim3(1), dim3(TB), 0, s, part, G2, ldp, sc + 61, (const int*)nullptr);
        double r[2];
        SM_HIP(hipMemcpyAsync(r, sc + 60, 16, hipMemcpyDeviceToHost, s));
        SM_HIP(hipStreamSynchronize(s));
        return r[0] / r[1];
    }

    // J filter of V(j) (the doubleLanczos step of TRL, lanczos_methods.jl:62-75): if
    // <J^2> is off J(J+1) by more than the tolerance (or `force`), V(j) becomes the
    // normalised, re-orthogonalised (against V[0..nv-1]) exact projection on J.
    // Returns the norm that projection and re-orthogonalisation kept of the unit
    // vector V(j) (1 when V(j) already has the J).
    double jfilter(int j, int nv, bool force) {
        double* y = W.jy.as<double>();
        double* z = W.jz.as<double>();
        hipLaunchKernelGGL((k_convert<TBv, double>), dim3(gstore), dim3(TB), 0, s, (const TBv*)V(j), y, n, 1.0);
        const double hJ = 0.5 * jp->twoJ, cJ = hJ * (hJ + 1.0);
        if (!force && std::fabs(j_expect(y, z) - cJ) <= jp->tol) return 1.0;
        // factors J' = J+1 .. Jmax, alternately from the top and the bottom (bounded norms)
        std::vector<int> two;
        for (int t = jp->twoJ + 2; t <= jp->twoJmax; t += 2) two.push_back(t);
        std::vector<int> order;
        for (size_t lo = 0, hi = two.size(); lo < hi;) {
            order.push_back(two[--hi]);
            if (lo < hi) order.push_back(two[lo++]);
        }
        for (int t : order) {
            const double hp = 0.5 * t, c = hp * (hp + 1.0);
            launch_hv(*jp->j2, 0, y, z, s);
            hipLaunchKernelGGL(k_jfactor, dim3(gstore), dim3(TB), 0, s, y, (const double*)z, n, c, 1.0 / (cJ - c));
        }
        st.n_jproj++;
        if (nv > 0) {
            cgs2<double>(nv, (const double*)y);          // -> w64, |w64|^2 partials in npart
            store_normalised(j);
        } else {
            SM_HIP(hipMemcpyAsync(W.w64.p, y, n * 8, hipMemcpyDeviceToDevice, s));
            hipLaunchKernelGGL((k_dotsw<double, double>), dim3(Gd), dim3(TB), 0, s, (const double*)nullptr, ld, 0,
                               (const double*)W.w64.as<double>(), n, W.part.as<double>(), ldp, 1,
                               (const int*)nullptr);
            hipLaunchKernelGGL(k_reduce, dim3(1), dim3(TB), 0, s, W.part.as<double>(), (int)Gd, ldp,
                               W.scal.as<double>() + 1, (const int*)nullptr);
            hipLaunchKernelGGL(k_sqrt1, dim3(1), dim3(1), 0, s, W.scal.as<double>() + 1, W.scal.as<double>());
            hipLaunchKernelGGL((k_store<TBv>), dim3(gstore), dim3(TB), 0, s, W.w64.as<double>(),
                               W.scal.as<double>(), V(j), n);
        }
        double kept = 0.0;
        SM_HIP(hipMemcpyAsync(&kept, W.scal.p, 8, hipMemcpyDeviceToHost, s));
        SM_HIP(hipStreamSynchronize(s));
        if (nv == 0 && !(kept > 1e-10))
            throw std::invalid_argument("start vector has no component with the requested J");
        return kept;
    }

    // V(j) = w64 / |w64|  (norm partials of w64 already in W.npart)
    void store_normalised(int j) {
        hipLaunchKernelGGL(k_norm_from_parts, dim3(1), dim3(TB), 0, s, W.npart.as<double>(), (int)Gx,
                           W.scal.as<double>());
        hipLaunchKernelGGL((k_store<TBv>), dim3(gstore), dim3(TB), 0, s, W.w64.as<double>(), W.scal.as<double>(),
                           V(j), n);
    }

    void restart() {
        int m = lm - 1;
        if (std::getenv("SMGPU_DEBUG_RESTART")) debug_sweep();
        restart_small_on_stream(W.alpha.as<double>(), W.beta.as<double>(), m, ls, W.S.as<double>(),
                                W.tvals.as<double>(), W.Y.as<double>(), W.Q.as<double>(), W.swork.as<double>(), s);
        rotate_chunks<TBv, TBv>((const TBv*)V(0), ld, m, W.S.as<double>(), ls, W.U.as<TBv>(), ld, n, s);
        // v_ls = v_{lm-1}; kept vectors from U
        SM_HIP(hipMemcpyAsync(V(ls), V(lm - 1), n * sizeof(TBv), hipMemcpyDeviceToDevice, s));
        SM_HIP(hipMemcpyAsync(V(0), W.U.p, (size_t)ls * ld * sizeof(TBv), hipMemcpyDeviceToDevice, s));
        // re-orthonormalise the kept vectors (the reference's make_sure pass)
        for (int j = 0; j < ls; j++) {
            if (j == 0) {
                hipLaunchKernelGGL((k_convert<TBv, double>), dim3(gstore), dim3(TB), 0, s, (const TBv*)V(0),
                                   W.w64.as<double>(), n, 1.0);
                hipLaunchKernelGGL((k_dotsw<double, double>), dim3(Gd), dim3(TB), 0, s, (const double*)nullptr,
                                   ld, 0, W.w64.as<double>(), n, W.part.as<double>(), ldp, 1, (const int*)nullptr);
                hipLaunchKernelGGL(k_reduce, dim3(1), dim3(TB), 0, s, W.part.as<double>(), (int)Gd, ldp,
                                   W.scal.as<double>() + 1, (const int*)nullptr);
                hipLaunchKernelGGL(k_sqrt1, dim3(1), dim3(1), 0, s, W.scal.as<double>() + 1, W.scal.as<double>());
                hipLaunchKernelGGL((k_store<TBv>), dim3(gstore), dim3(TB), 0, s, W.w64.as<double>(),
                                   W.scal.as<double>(), V(0), n);
            } else {
                cgs2<TBv>(j, (const TBv*)V(j));
                store_normalised(j);
            }
        }
        // v_ls against the kept vectors (no renormalisation, as the reference)
        cgs2<TBv>(ls, (const TBv*)V(ls));
        hipLaunchKernelGGL((k_convert<double, TBv>), dim3(gstore), dim3(TB), 0, s, W.w64.as<double>(), V(ls), n, 1.0);
        if (partial)
            hipLaunchKernelGGL(k_omega_reset, dim3(1), dim3(TB), 0, s, W.om.as<double>(), lm + 2, ls, eps_basis,
                               W.gate.as<int>());
        if (std::getenv("SMGPU_DEBUG_RESTART")) debug_restart();
    }

    // opt-in check of the sweep before a restart: orthonormality of V[0..lm-1] and
    // |H v_k - a_k v_k - b_{k-1} v_{k-1} - b_k v_{k+1}|
    void debug_sweep() {
        SM_HIP(hipStreamSynchronize(s));
        std::vector<TBv> hv((size_t)lm * ld);
        SM_HIP(hipMemcpy(hv.data(), V(0), hv.size() * sizeof(TBv), hipMemcpyDeviceToHost));
        std::vector<double> al(lm), be(lm);
        SM_HIP(hipMemcpy(al.data(), W.alpha.p, (lm - 1) * 8, hipMemcpyDeviceToHost));
        SM_HIP(hipMemcpy(be.data(), W.beta.p, (lm - 1) * 8, hipMemcpyDeviceToHost));
        for (int i = 0; i < lm; i++) {
            double worst = 0.0;
            for (int j = 0; j <= i; j++) {
                double d = 0.0;
                for (int64_t x = 0; x < n; x++) d += (double)hv[(size_t)i * ld + x] * (double)hv[(size_t)j * ld + x];
                worst = std::max(worst, std::fabs(d - (i == j ? 1.0 : 0.0)));
            }
            if (i % 4 == 0 || i == lm - 1) fprintf(stderr, "[sweep] v%d max |<v_i,v_j>-d_ij| over j<=i: %.3e\n", i, worst);
        }
        std::vector<TBv> hw(n);
        for (int k = 0; k < lm - 1; k += 4) {
            if (prec_hv == 0) launch_hv(H, 0, V(k), W.wH.p, s);
            else launch_hv(H, 1, V(k), W.wH.p, s);
            SM_HIP(hipStreamSynchronize(s));
            SM_HIP(hipMemcpy(hw.data(), W.wH.p, n * sizeof(TBv), hipMemcpyDeviceToHost));
            double r2 = 0.0;
            for (int64_t x = 0; x < n; x++) {
                double e = (double)hw[x] - al[k] * (double)hv[(size_t)k * ld + x];
                if (k > 0) e -= be[k - 1] * (double)hv[(size_t)(k - 1) * ld + x];
                e -= be[k] * (double)hv[(size_t)(k + 1) * ld + x];
                r2 += e * e;
            }
            fprintf(stderr, "[sweep] k=%d alpha=%.6f beta=%.6f |H v - T v|=%.3e\n", k, al[k], be[k], std::sqrt(r2));
        }
    }

    // opt-in consistency check of the state after a restart: orthonormality of
    // V[0..ls] and the tridiagonal relation H u_k = a_k u_k + b_{k-1} u_{k-1} + b_k u_{k+1}
    void debug_restart() {
        SM_HIP(hipStreamSynchronize(s));
        std::vector<TBv> hv((size_t)(ls + 1) * ld);
        SM_HIP(hipMemcpy(hv.data(), V(0), hv.size() * sizeof(TBv), hipMemcpyDeviceToHost));
        std::vector<double> al(ls + 1), be(ls + 1);
        SM_HIP(hipMemcpy(al.data(), W.alpha.p, ls * 8, hipMemcpyDeviceToHost));
        SM_HIP(hipMemcpy(be.data(), W.beta.p, ls * 8, hipMemcpyDeviceToHost));
        double worst = 0.0;
        for (int i = 0; i <= ls; i++)
            for (int j = 0; j <= i; j++) {
                double d = 0.0;
                for (int64_t x = 0; x < n; x++) d += (double)hv[(size_t)i * ld + x] * (double)hv[(size_t)j * ld + x];
                worst = std::max(worst, std::fabs(d - (i == j ? 1.0 : 0.0)));
            }
        fprintf(stderr, "[restart] n=%lld orthonormality error %.3e\n", (long long)n, worst);
        std::vector<TBv> hw(n);
        for (int k = 0; k < ls; k++) {
            if (prec_hv == 0) launch_hv(H, 0, V(k), W.wH.p, s);
            else launch_hv(H, 1, V(k), W.wH.p, s);
            SM_HIP(hipStreamSynchronize(s));
            SM_HIP(hipMemcpy(hw.data(), W.wH.p, n * sizeof(TBv), hipMemcpyDeviceToHost));
            double r2 = 0.0, rq = 0.0;
            for (int64_t x = 0; x < n; x++) {
                double e = (double)hw[x] - al[k] * (double)hv[(size_t)k * ld + x];
                if (k > 0) e -= be[k - 1] * (double)hv[(size_t)(k - 1) * ld + x];
                e -= be[k] * (double)hv[(size_t)(k + 1) * ld + x];
                r2 += e * e;
                rq += (double)hw[x] * (double)hv[(size_t)k * ld + x];
            }
            fprintf(stderr, "[restart] k=%d alpha=%.6f rq=%.6f beta=%.6f |H u - T u|=%.3e\n", k, al[k], rq, be[k],
                    std::sqrt(r2));
        }
    }

    void finish(int elit, int conv_step, double* evals, double* evecs) {
        bool need_vecs = evecs != nullptr || (prec_hv == 1 && P.refine);
        if (!need_vecs) return;
        // eigenvectors of T[0..elit-1] for the converged Ritz values
        // (inverse iteration on the host: a few thousand dependent operations, which
        // one GPU thread would spend milliseconds on in global-memory latency)
        double* Y = W.Y.as<double>();
        {
            std::vector<double> a(elit), b(elit), vals(nev), y((size_t)nev * elit), work(5 * (size_t)elit + 8);
            SM_HIP(hipMemcpyAsync(a.data(), W.alpha.p, elit * 8, hipMemcpyDeviceToHost, s));
            SM_HIP(hipMemcpyAsync(b.data(), W.beta.p, elit * 8, hipMemcpyDeviceToHost, s));
            SM_HIP(hipMemcpyAsync(vals.data(), en(conv_step), nev * 8, hipMemcpyDeviceToHost, s));
            SM_HIP(hipStreamSynchronize(s));
            const double tn = tri_norm(a.data(), b.data(), elit);
            const double ortol = 1e-3 * tn;
            for (int j = 0; j < nev; j++) {
                const int first = tri_cluster_first(vals.data(), j, ortol);
                tri_eigvec(a.data(), b.data(), elit, vals[j], y.data() + (size_t)j * elit, work.data(),
                           y.data() + (size_t)first * elit, j - first, tn);
            }
            SM_HIP(hipMemcpyAsync(Y, y.data(), y.size() * 8, hipMemcpyHostToDevice, s));
            SM_HIP(hipStreamSynchronize(s));
        }
        DBuf& R = W.R;
        rotate_chunks<TBv, double>((const TBv*)V(0), ld, elit, (const double*)Y, nev, R.as<double>(), ld, n, s);
        const bool refine = prec_hv == 1 && P.refine;
        DBuf& hv64 = W.hv64;
        double* sc = W.scal.as<double>();
        for (int i = 0; i < nev; i++) {
            double* r = R.as<double>() + (int64_t)i * ld;
            // |r|^2 -> sc[2i]; normalise (Rvec .*= 1/sqrt(dot))
            hipLaunchKernelGGL((k_dotsw<double, double>), dim3(Gd), dim3(TB), 0, s, (const double*)nullptr, ld, 0,
                               (const double*)r, n, W.part.as<double>(), ldp, 1, (const int*)nullptr);
            hipLaunchKernelGGL(k_reduce, dim3(1), dim3(TB), 0, s, W.part.as<double>(), (int)Gd, ldp, sc + 2 * i,
                               (const int*)nullptr);
            hipLaunchKernelGGL(k_sqrt1, dim3(1), dim3(1), 0, s, sc + 2 * i, sc + 2 * i + 1);
            hipLaunchKernelGGL((k_store<double>), dim3(gstore), dim3(TB), 0, s, (const double*)r, sc + 2 * i + 1,
                               r, n);
        }
        if (refine) {
            // FP64 Rayleigh quotients with the FP64 H·v, all Ritz vectors in shared launches
            launch_hv_multi(H, 0, nev, R.p, ld, hv64.p, ld, s);
            st.n_hv64 += nev;
            for (int i = 0; i < nev; i++) {
                const double* r = R.as<double>() + (int64_t)i * ld;
                hipLaunchKernelGGL((k_dotsw<double, double>), dim3(Gd), dim3(TB), 0, s, r, ld, 1,
                                   (const double*)(hv64.as<double>() + (int64_t)i * ld), n, W.part.as<double>(), ldp,
                                   0, (const int*)nullptr);
                hipLaunchKernelGGL(k_reduce, dim3(1), dim3(TB), 0, s, W.part.as<double>(), (int)Gd, ldp,
                                   W.h1.as<double>() + i, (const int*)nullptr);
            }
        }
        if (refine) {
            std::vector<double> rq(nev);
            SM_HIP(hipMemcpyAsync(rq.data(), W.h1.p, nev * 8, hipMemcpyDeviceToHost, s));
            SM_HIP(hipStreamSynchronize(s));
            // A Ritz value and the FP64 Rayleigh quotient of its Ritz vector agree to
            // the convergence level (~1e-6 MeV); a gap above 1e-3 |theta| means the
            // Lanczos basis lost its (semi-)orthogonality and T no longer projects H
            for (int i = 0; i < nev; i++)
                if (!(std::fabs(rq[i] - evals[i]) <= 1e-3 * std::max(1.0, std::fabs(evals[i])))) throw Inconsistent();
            for (int i = 0; i < nev; i++) evals[i] = rq[i];
            std::sort(evals, evals + nev);
        }
        if (evecs) {
            for (int i = 0; i < nev; i++)
                SM_HIP(hipMemcpyAsync(evecs + (int64_t)i * n, R.as<double>() + (int64_t)i * ld, n * 8,
                                      hipMemcpyDeviceToHost, s));
        }
        SM_HIP(hipStreamSynchronize(s));
    }
};

// ------------------------------------------------------------------ block TRL
__global__ __launch_bounds__(TB) void k_scale_h(double* __restrict__ x, int64_t n, double a) {
    for (int64_t i = (int64_t)blockIdx.x * TB + threadIdx.x; i < n; i += (int64_t)gridDim.x * TB) x[i] *= a;
}
// y += a x
__global__ __launch_bounds__(TB) void k_axpy_h(double* __restrict__ y, const double* __restrict__ x, int64_t n,
                                               double a) {
    for (int64_t i = (int64_t)blockIdx.x * TB + threadIdx.x; i < n; i += (int64_t)gridDim.x * TB) y[i] += a * x[i];
}

// Thick-restart block Lanczos with block size q (TRBL, lanczos_methods.jl:312-420;
// bl_ReORTH :422-433; bl_QR! shellmodel_main.jl:1049-1066; add_bl_T! :1201-1205;
// bl_ThickRestart lanczos_methods.jl:563-617).  One block step:
//   HV = H V_it  (q H·v);  R = V_it^T HV, lower triangle mirrored (diagonal block of T)
//   HV <- HV - V R, re-orthogonalised against the whole basis (CGS2 per vector;
//         the reference sweeps MGS over the earlier blocks)
//   HV = V_{it+1} B  (MGS QR), B and B^T the next off-diagonal blocks of T.
// T lives on the host in FP64; its lowest eigenvalues are checked after every
// block with diagonalize_T!'s rule.  After lm/q - 1 blocks without convergence,
// the basis restarts from the ls = q*(ls/q) lowest Ritz vectors and the last
// block (arrowhead coupling B Y_last).
// Rank loss: a QR column whose norm after re-orthogonalisation is at rounding
// level (the block Krylov space is exhausted: the basis spans the space, or an
// invariant subspace) is dropped; the remaining columns form a last, partial
// block, after which the solve stops with T the exact projection on the span.
// The reference normalises such columns as they are.
template <typename TBv>
struct BlockSolver {
    DeviceHam& H;
    const SolveParams& P;
    SolveStats& st;
    const JProj* jp;
    const std::atomic<bool>* cancel = nullptr;
    hipStream_t s;
    int q, nbl, lsub, lsv, nev, lmT, ldp, gstore, prec_hv;
    int64_t n, ld, Gd, Gx;
    DBuf basis, hv, w, part, npart, h1, h2, rd, U, Sd, R, hv64, jz;
    std::vector<double> T;      // lmT x lmT, row-major
    std::vector<double> Bq;     // q x q coupling of the last full QR
    DenseEig eig;

    BlockSolver(DeviceHam& h, const SolveParams& p, SolveStats& stats, int qq, const JProj* j)
        : H(h), P(p), st(stats), jp(j), q(qq) {
        s = H.stream;
        n = H.mdim;
        ld = (n + 63) / 64 * 64;
        nbl = P.lm / q;
        nev = (int)std::min<int64_t>(P.num_ev, n);
        lsub = P.ls / q;
        if (q * lsub < nev) lsub = (nev + q - 1) / q;     // keep at least num_ev (as kept_ls)
        lsv = q * lsub;
        lmT = q * nbl;
        ldp = lmT + 2;
        prec_hv = sizeof(TBv) == 4 ? 1 : 0;
        Gd = (n + CHUNK - 1) / CHUNK;
        Gx = (n + GsShape<TBv>::SEG - 1) / GsShape<TBv>::SEG;
        gstore = (int)std::min<int64_t>(4096, (n + TB - 1) / TB);
    }
    TBv* V(int j) { return basis.as<TBv>() + (int64_t)j * ld; }
    TBv* HVb(int b) { return hv.as<TBv>() + (int64_t)b * ld; }
    double* Wb(int b) { return w.as<double>() + (int64_t)b * ld; }
    double& t(int i, int j) { return T[(size_t)i * lmT + j]; }

    void check_args() {
        if (q < 1 || q > 16) throw std::invalid_argument("block size q must be in 1..16");
        if (lsub < 1 || lsv > 64) throw std::invalid_argument("need q <= ls and q*(ls/q) <= 64");
        if (nev < 1 || nev > 64) throw std::invalid_argument("need 1 <= num_ev <= 64");
        if (nbl < lsub + 3) throw std::invalid_argument("need lm/q >= ls/q + 3");
        if (lmT > 1000) throw std::invalid_argument("lm above 1000 is not supported");
        if (n < q) throw std::invalid_argument("block size larger than the basis");
    }

    void alloc() {
        basis.alloc((size_t)lmT * ld * sizeof(TBv));
        hv.alloc((size_t)q * ld * sizeof(TBv));
        w.alloc((size_t)q * ld * 8);
        part.alloc((size_t)std::max(Gd, Gx) * ldp * 8);
        npart.alloc((size_t)Gx * 8);
        h1.alloc((size_t)q * ldp * 8);
        h2.alloc((size_t)q * ldp * 8);
        rd.alloc(64 * 8);
        U.alloc((size_t)lsv * ld * sizeof(TBv));
        Sd.alloc((size_t)lmT * std::max(32, std::max(lsv, nev)) * 8);
        R.alloc((size_t)nev * ld * 8);
        if (prec_hv == 1 && P.refine) hv64.alloc((size_t)nev * ld * 8);
        if (jp) jz.alloc((size_t)ld * 8);
    }

    // out[0] = a . b  (device scalar)
    void dot_to(const double* a, const double* b, double* out) {
        const int G2 = (int)std::min<int64_t>(1024, Gd);
        hipLaunchKernelGGL((k_dot2<double, double>), dim3(G2), dim3(TB), 0, s, a, b, n, part.as<double>(), ldp);
        hipLaunchKernelGGL(k_reduce, dim3(1), dim3(TB), 0, s, part.as<double>(), G2, ldp, out, (const int*)nullptr);
    }

    // wout = win - V c1 - V c2 with c1 = V^T win, c2 = V^T (win - V c1) over V[0..nv-1]
    // (CGS2, three streaming passes); c1 -> hh1, c2 -> hh2
    template <typename TW>
    void cgs2_into(int nv, const TW* win, double* wout, double* hh1, double* hh2) {
        double* pt = part.as<double>();
        hipLaunchKernelGGL((k_gs_dots<TBv, TW, false>), dim3(Gx), dim3(TB), 0, s, (const TBv*)V(0), ld, nv, win, n,
                           pt, ldp, (const int*)nullptr, (const TBv*)nullptr, (const TBv*)nullptr,
                           (const double*)nullptr, (const double*)nullptr, (double*)nullptr, (double*)nullptr);
        hipLaunchKernelGGL(k_reduce, dim3(nv), dim3(TB), 0, s, pt, (int)Gx, ldp, hh1, (const int*)nullptr);
        hipLaunchKernelGGL((k_gs_upd<TBv, TW, true>), dim3(Gx), dim3(TB), 0, s, (const TBv*)V(0), ld, nv,
                           (const double*)hh1, win, wout, n, pt, ldp, npart.as<double>(), (const int*)nullptr);
        hipLaunchKernelGGL(k_reduce, dim3(nv), dim3(TB), 0, s, pt, (int)Gx, ldp, hh2, (const int*)nullptr);
        hipLaunchKernelGGL((k_gs_upd<TBv, double, false>), dim3(Gx), dim3(TB), 0, s, (const TBv*)V(0), ld, nv,
                           (const double*)hh2, (const double*)wout, wout, n, pt, ldp, npart.as<double>(),
                           (const int*)nullptr);
    }

    // MGS QR of Wb(0..qe-1) in place (bl_QR!): Rq (q x q, row-major) gets the
    // coefficients; a column whose norm is <= thr is dropped.  Returns the kept columns.
    std::vector<int> qr_block(int qe, double thr, std::vector<double>& Rq) {
        Rq.assign((size_t)q * q, 0.0);
        std::vector<int> ok;
        double* r = rd.as<double>();
        std::vector<double> d(q);
        for (int i = 0; i < qe; i++) {
            for (int k = i; k < qe; k++) dot_to(Wb(i), Wb(k), r + k);
            SM_HIP(hipMemcpyAsync(d.data() + i, r + i, (qe - i) * 8, hipMemcpyDeviceToHost, s));
            SM_HIP(hipStreamSynchronize(s));
            const double nr = std::sqrt(std::max(d[i], 0.0));
            if (!(nr > thr)) continue;          // dropped (also NaN)
            Rq[(size_t)i * q + i] = nr;
            hipLaunchKernelGGL(k_scale_h, dim3(gstore), dim3(TB), 0, s, Wb(i), n, 1.0 / nr);
            for (int k = i + 1; k < qe; k++) {
                const double c = d[k] / nr;
                Rq[(size_t)i * q + k] = c;
                hipLaunchKernelGGL(k_axpy_h, dim3(gstore), dim3(TB), 0, s, Wb(k), (const double*)Wb(i), n, -c);
            }
            ok.push_back(i);
        }
        return ok;
    }

    // J projection of the FP64 vector y in place (the exact form of bl_JJ_Lanczos'
    // filter, lanczos_methods.jl:704-760): skipped when <J^2> is within tol of J(J+1)
    // unless forced.  Returns whether it projected.
    bool jproject(double* y, bool force) {
        const double hJ = 0.5 * jp->twoJ, cJ = hJ * (hJ + 1.0);
        double* z = jz.as<double>();
        if (!force) {
            launch_hv(*jp->j2, 0, y, z, s);
            double* r = rd.as<double>();
            dot_to(y, z, r + 40);
            dot_to(y, y, r + 41);
            double e[2];
            SM_HIP(hipMemcpyAsync(e, r + 40, 16, hipMemcpyDeviceToHost, s));
            SM_HIP(hipStreamSynchronize(s));
            if (std::fabs(e[0] / e[1] - cJ) <= jp->tol) return false;
        }
        std::vector<int> two;
        for (int t2 = jp->twoJ + 2; t2 <= jp->twoJmax; t2 += 2) two.push_back(t2);
        std::vector<int> order;
        for (size_t lo = 0, hi = two.size(); lo < hi;) {
            order.push_back(two[--hi]);
            if (lo < hi) order.push_back(two[lo++]);
        }
        for (int t2 : order) {
            const double hp = 0.5 * t2, c = hp * (hp + 1.0);
            launch_hv(*jp->j2, 0, y, z, s);
            hipLaunchKernelGGL(k_jfactor, dim3(gstore), dim3(TB), 0, s, y, (const double*)z, n, c, 1.0 / (cJ - c));
        }
        st.n_jproj++;
        return true;
    }

    // lowest min(nev, k) eigenvalues of T[0..k-1] into cur (entries past keep prev);
    // returns diagonalize_T!'s convergence flag all(prev - cur < tol)
    bool converged(int k, std::vector<double>& prev, std::vector<double>& cur) {
        const int m = std::min(nev, k);
        cur = prev;
        // band reduction while T is banded; the previous values bound these from above
        eig.lowest_values(T.data(), lmT, k, m, cur.data(), prev.data());
        bool ok = k >= nev;     // (stale values past k do not count as converged)
        for (int i = 0; i < nev; i++) ok = ok && (prev[i] - cur[i] < P.tol);
        prev = cur;
        return ok;
    }

    void store_block(int j0, const std::vector<int>& cols) {
        for (size_t i = 0; i < cols.size(); i++)
            hipLaunchKernelGGL((k_convert<double, TBv>), dim3(gstore), dim3(TB), 0, s, (const double*)Wb(cols[i]),
                               V(j0 + (int)i), n, 1.0);
    }

    void start_block(const double* v0) {
        for (int b = 0; b < q; b++) {
            if (v0) SM_HIP(hipMemcpyAsync(Wb(b), v0 + (int64_t)b * n, n * 8, hipMemcpyHostToDevice, s));
            else hipLaunchKernelGGL(k_randn, dim3(gstore), dim3(TB), 0, s, Wb(b), n,
                                    (uint64_t)P.seed + 0x9E3779B97F4A7C15ull * (uint64_t)b);
            if (jp) jproject(Wb(b), true);
        }
        std::vector<double> Rq;
        auto ok = qr_block(q, 1e-10, Rq);
        if ((int)ok.size() < q)
            throw std::invalid_argument(jp ? "start block has too few components with the requested J"
                                           : "start vectors are linearly dependent");
        store_block(0, ok);
    }

    void restart() {
        const int m = q * (nbl - 1);
        eig.reduce(T.data(), lmT, m);
        std::vector<double> vals(lsv), Y((size_t)lsv * m);
        eig.lowest_pairs(lsv, vals.data(), Y.data());
        std::vector<double> r((size_t)q * lsv, 0.0);     // r = B Y[m-q..m-1, :]
        for (int i = 0; i < q; i++)
            for (int k = 0; k < lsv; k++) {
                double a = 0.0;
                for (int c = 0; c < q; c++) a += Bq[(size_t)i * q + c] * Y[(size_t)k * m + (m - q + c)];
                r[(size_t)i * lsv + k] = a;
            }
        std::fill(T.begin(), T.end(), 0.0);
        for (int k = 0; k < lsv; k++) t(k, k) = vals[k];
        for (int i = 0; i < q; i++)
            for (int k = 0; k < lsv; k++) t(lsv + i, k) = t(k, lsv + i) = r[(size_t)i * lsv + k];
        // kept Ritz vectors U = V[0..m-1] Y, then the new block moves behind them
        SM_HIP(hipMemcpyAsync(Sd.p, Y.data(), Y.size() * 8, hipMemcpyHostToDevice, s));
        rotate_chunks<TBv, TBv>((const TBv*)V(0), ld, m, (const double*)Sd.as<double>(), lsv, U.as<TBv>(), ld, n, s);
        SM_HIP(hipMemcpyAsync(V(lsv), V(m), (size_t)q * ld * sizeof(TBv), hipMemcpyDeviceToDevice, s));
        SM_HIP(hipMemcpyAsync(V(0), U.p, (size_t)lsv * ld * sizeof(TBv), hipMemcpyDeviceToDevice, s));
        // blocks 1..ls/q re-orthonormalised against the earlier ones (bl_ThickRestart's final sweep)
        std::vector<double> Rq;
        for (int b = 1; b <= lsub; b++) {
            for (int i = 0; i < q; i++)
                cgs2_into<TBv>(b * q, (const TBv*)V(b * q + i), Wb(i), h1.as<double>(), h2.as<double>());
            auto ok = qr_block(q, 0.0, Rq);
            if ((int)ok.size() < q) throw std::runtime_error("TRBL restart: kept block lost rank");
            store_block(b * q, ok);
        }
        st.n_restart++;
    }

    void run(double* evals, double* evecs, const double* v0) {
        auto t0 = std::chrono::steady_clock::now();
        check_args();
        alloc();
        T.assign((size_t)lmT * lmT, 0.0);
        Bq.assign((size_t)q * q, 0.0);
        st.n_hv = 0; st.n_restart = 0;
        start_block(v0);
        std::vector<double> prev(nev, 1.0e4), cur(nev, 1.0e4);
        const double drop = prec_hv ? 1e-5 : 1e-10;   // rank-loss threshold relative to max|T|
        int itmin = 0, qe = q, kfin = 0;
        int64_t total = 0;
        bool done = false;
        while (!done) {
            for (int it = itmin; it < nbl - 1 || (qe < q && it < nbl); it++) {
                if (cancel && cancel->load(std::memory_order_relaxed)) throw Canceled();
                total += qe;
                if (total > P.max_iter) throw std::runtime_error("TRBL: iteration limit reached");
                const int j0 = it * q;
                launch_hv_multi(H, prec_hv, qe, V(j0), ld, HVb(0), ld, s);     // bl_operate_H!
                st.n_hv += qe;
                const int k = j0 + qe;
                for (int b = 0; b < qe; b++)
                    cgs2_into<TBv>(k, (const TBv*)HVb(b), Wb(b), h1.as<double>() + (int64_t)b * ldp,
                                   h2.as<double>() + (int64_t)b * ldp);
                std::vector<double> c1((size_t)qe * qe), c2((size_t)qe * qe);
                for (int b = 0; b < qe; b++) {
                    SM_HIP(hipMemcpyAsync(c1.data() + (size_t)b * qe, h1.as<double>() + (int64_t)b * ldp + j0, qe * 8,
                                          hipMemcpyDeviceToHost, s));
                    SM_HIP(hipMemcpyAsync(c2.data() + (size_t)b * qe, h2.as<double>() + (int64_t)b * ldp + j0, qe * 8,
                                          hipMemcpyDeviceToHost, s));
                }
                SM_HIP(hipStreamSynchronize(s));
                // Rr[i][j] = v_i . H v_j ; the diagonal block takes the lower triangle
                auto Rr = [&](int i, int j) { return c1[(size_t)j * qe + i] + c2[(size_t)j * qe + i]; };
                for (int i = 0; i < qe; i++)
                    for (int j = 0; j < qe; j++) t(j0 + i, j0 + j) = i >= j ? Rr(i, j) : Rr(j, i);
                const bool conv = converged(k, prev, cur);
                if (qe < q || k >= n) { kfin = k; done = true; break; }      // space exhausted
                double scale = 1.0;
                for (int i = 0; i < k; i++)
                    for (int j = std::max(0, i - 2 * q); j <= i; j++) scale = std::max(scale, std::fabs(t(i, j)));
                std::vector<double> Rq;
                auto ok = qr_block(qe, drop * scale, Rq);
                // src: columns of Wb to store (in order); rows: their rows of Rq
                std::vector<int> src = ok, rows = ok;
                if (jp && !ok.empty()) {
                    // doubleLanczos: stop when the new block vanishes, else filter it on J
                    double bn = 0.0;
                    for (int i = 0; i < q; i++) bn = std::max(bn, std::fabs(Rq[(size_t)i * q + i]));
                    if (bn < jp->tol) ok.clear();
                    bool any = false;
                    for (int i : ok) any = jproject(Wb(i), false) || any;
                    src = ok; rows = ok;
                    if (any) {
                        // filtered columns go back into the complement of the span with unit
                        // norm (compacted to Wb(0..)); one that keeps less than J_EXHAUSTED of
                        // its norm was noise outside the (exhausted) J space and is dropped
                        for (size_t a = 0; a < ok.size(); a++)
                            cgs2_into<double>(k, (const double*)Wb(ok[a]), Wb((int)a), h1.as<double>(),
                                              h2.as<double>());
                        std::vector<double> Rj;
                        src = qr_block((int)ok.size(), J_EXHAUSTED, Rj);
                        rows.clear();
                        for (int a : src) rows.push_back(ok[a]);
                    }
                }
                if ((int64_t)src.size() > n - k) { src.resize(n - k); rows.resize(n - k); }
                if (src.empty()) { kfin = k; done = true; break; }
                // coupling blocks (add_bl_T!): T[k+i][j0+c] = T[j0+c][k+i] = B[row_i][c]
                for (size_t i2 = 0; i2 < src.size(); i2++)
                    for (int c = 0; c < q; c++)
                        t(k + (int)i2, j0 + c) = t(j0 + c, k + (int)i2) = Rq[(size_t)rows[i2] * q + c];
                store_block(k, src);
                if (conv) { kfin = k; done = true; break; }
                if ((int)src.size() < q) qe = (int)src.size();   // last, partial block
                else Bq = Rq;
            }
            if (done) break;
            restart();
            itmin = lsub;
        }
        SM_HIP(hipStreamSynchronize(s));
        st.iterations = kfin;
        for (int i = 0; i < nev; i++) { evals[i] = cur[i]; st.ritz[i] = cur[i]; }
        if (kfin < nev) throw std::invalid_argument("the Krylov space holds fewer states than num_ev");
        finish(kfin, evals, evecs);
        st.wall_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }

    // Ritz vectors of T[0..k-1] (main_sm's Rvecs for is_block, shellmodel_main.jl:211-224),
    // FP64 Rayleigh quotients with the FP64 H·v when H·v ran in FP32
    void finish(int k, double* evals, double* evecs) {
        const bool refine = prec_hv == 1 && P.refine;
        if (!evecs && !refine) return;
        if (k < nev) throw std::invalid_argument("the Krylov space holds fewer states than num_ev");
        eig.reduce(T.data(), lmT, k);
        std::vector<double> vals(nev), Y((size_t)nev * k);
        eig.lowest_pairs(nev, vals.data(), Y.data());
        SM_HIP(hipMemcpyAsync(Sd.p, Y.data(), Y.size() * 8, hipMemcpyHostToDevice, s));
        rotate_chunks<TBv, double>((const TBv*)V(0), ld, k, (const double*)Sd.as<double>(), nev, R.as<double>(), ld, n,
                                   s);
        double* r = rd.as<double>();
        for (int i = 0; i < nev; i++) {
            double* x = R.as<double>() + (int64_t)i * ld;
            dot_to(x, x, r + 48);
            double nr2 = 0.0;
            SM_HIP(hipMemcpyAsync(&nr2, r + 48, 8, hipMemcpyDeviceToHost, s));
            SM_HIP(hipStreamSynchronize(s));
            hipLaunchKernelGGL(k_scale_h, dim3(gstore), dim3(TB), 0, s, x, n, 1.0 / std::sqrt(nr2));
            if (refine) {
                launch_hv(H, 0, x, hv64.p, s);
                st.n_hv64++;
                dot_to(x, hv64.as<double>(), r + 49);
                double rq = 0.0;
                SM_HIP(hipMemcpyAsync(&rq, r + 49, 8, hipMemcpyDeviceToHost, s));
                SM_HIP(hipStreamSynchronize(s));
                evals[i] = rq;
            }
        }
        if (refine) std::sort(evals, evals + nev);
        if (evecs)
            for (int i = 0; i < nev; i++)
                SM_HIP(hipMemcpyAsync(evecs + (int64_t)i * n, R.as<double>() + (int64_t)i * ld, n * 8,
                                      hipMemcpyDeviceToHost, s));
        SM_HIP(hipStreamSynchronize(s));
    }
};

}  // namespace

int64_t solve_workspace_bytes(int64_t mdim, const SolveParams& p) {
    int64_t total = 0;
    auto add = [&](DBuf Workspace::*, int64_t b) { total += b; };
    if (p.precision == 0) workspace_plan<double>(mdim, p, add);
    else workspace_plan<float>(mdim, p, add);
    return total;
}

int64_t workspace_capacity(const SolveCtx& ctx) {
    int64_t c = 0;
    for (int k = 0; k < 2; k++)
        if (ctx.ws[k]) c += static_cast<const Workspace*>(ctx.ws[k].get())->capacity();
    return c;
}

int64_t workspace_growth(const SolveCtx& ctx, int64_t mdim, const SolveParams& p) {
    int64_t g = 0;
    const int k = p.precision == 0 ? 0 : 1;
    const Workspace* W = static_cast<const Workspace*>(ctx.ws[k].get());
    auto f = [&](DBuf Workspace::*m, int64_t b) {
        int64_t cap = W ? (int64_t)(W->*m).cap : 0;
        if (b > cap) g += b - cap;
    };
    if (k == 0) workspace_plan<double>(mdim, p, f);
    else workspace_plan<float>(mdim, p, f);
    return g;
}

void release_workspace(SolveCtx& ctx) {
    for (int k = 0; k < 2; k++) ctx.ws[k].reset();
}


void trl_solve(DeviceHam& h, const SolveParams& p, double* evals, double* evecs, SolveStats& st,
               const double* v0, const SolveCtx* ctx, const JProj* jp) {
    SM_HIP(hipSetDevice(h.device));
    if (jp && (jp->j2 == nullptr || jp->j2->mdim != h.mdim))
        throw std::invalid_argument("J^2 handle does not match the Hamiltonian's basis");
    if (p.precision == 0) {
        Solver<double> S(h, p, st, ctx, jp);
        S.run(evals, evecs, v0);
        return;
    }
    try {
        Solver<float> S(h, p, st, ctx, jp);
        S.run(evals, evecs, v0);
    } catch (const Inconsistent&) {
        SolveParams pf = p;
        pf.reorth = SMGPU_REORTH_FULL;
        st = SolveStats{};
        Solver<float> S(h, pf, st, ctx, jp);
        S.run(evals, evecs, v0);
        st.n_fallback = 1;
    }
}

// device run of the restart small problem (tests compare it with the host build)
// Several TRL solves on one Hamiltonian in lockstep: each round issues the H·v of
// every active member in shared launches (launch_hv_set, up to four vectors per
// launch) on hv_stream, after which each member runs the rest of its step on its
// own stream.  With per-member Hamiltonians (io.hams, interaction variants on one
// index) the round's vectors are ordered by Hamiltonian and issued through
// launch_hv_variants: members of one Hamiltonian share launches of up to four,
// the others pair up with their own coefficients.  Members keep their own basis, workspace and convergence; each
// result is bit-identical to the member solved alone with the same parameters.
// Member slots (stream + workspace, ctx[b]) are reused: a member that finishes or
// fails frees its slot, and before every round join(b) may place a new solve in a
// free slot b (filling io at b), which then starts at its own first step.
template <typename TBv>
static void batch_run(DeviceHam& h, int cap, int nb, const BatchIO& io, const SolveCtx* const* ctx,
                      hipStream_t hv_stream, const std::function<bool(int)>& join,
                      const std::function<void(int)>& on_done) {
    std::vector<std::unique_ptr<Solver<TBv>>> S(cap);
    std::vector<char> busy(cap, 0);
    std::vector<int> active;
    std::vector<hipEvent_t> ready(cap, nullptr);
    std::vector<const DeviceHam*> mh(cap, &h);
    hipEvent_t evh = nullptr;
    struct Events {
        std::vector<hipEvent_t>& r;
        hipEvent_t& h;
        ~Events() {
            for (auto e : r) if (e) (void)hipEventDestroy(e);
            if (h) (void)hipEventDestroy(h);
        }
    } events{ready, evh};
    SM_HIP(hipEventCreateWithFlags(&evh, hipEventDisableTiming));
    for (int b = 0; b < cap; b++) SM_HIP(hipEventCreateWithFlags(&ready[b], hipEventDisableTiming));
    auto fail = [&](int b) {
        std::string m;
        io.status[b] = status_of_current_exception(m);
        if (io.msgs) io.msgs[b] = m;
        busy[b] = 0;
        if (on_done) on_done(b);
    };
    auto begin = [&](int b) {
        io.status[b] = SMGPU_OK;
        busy[b] = 1;
        try {
            DeviceHam& hb = io.hams ? *io.hams[b] : h;
            if (hb.idx != h.idx) throw std::invalid_argument("a batch member's Hamiltonian is on other index tables");
            S[b] = std::make_unique<Solver<TBv>>(hb, io.p[b], io.st[b], ctx[b]);
            mh[b] = &hb;
            S[b]->step_ready = ready[b];
            S[b]->start(io.v0 ? io.v0[b] : nullptr);
            SM_HIP(hipEventRecord(ready[b], S[b]->stream()));      // start vector ready
            active.push_back(b);
        } catch (...) {
            fail(b);
        }
    };
    for (int b = 0; b < nb; b++) begin(b);
    std::vector<const void*> vin;
    std::vector<void*> vout;
    std::vector<const DeviceHam*> hin;
    std::vector<int> order;
    const int prec = sizeof(TBv) == 4 ? 1 : 0;
    for (;;) {
        if (join)
            for (int b = 0; b < cap; b++)
                if (!busy[b] && join(b)) begin(b);
        if (active.empty()) break;
        order = active;
        if (io.hams)
            std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return mh[a] < mh[b]; });
        vin.clear();
        vout.clear();
        hin.clear();
        for (int b : order) {
            SM_HIP(hipStreamWaitEvent(hv_stream, ready[b], 0));
            vin.push_back(S[b]->hv_in());
            vout.push_back(S[b]->hv_out());
            hin.push_back(mh[b]);
        }
        // Members read the omega gate on the host in rounds of one Hamiltonian only.
        // Measured (served, two slots of two): Cr48 10.69 -> 10.87 solves/s with the
        // host-read gate; 64 interaction variants 13.69 -> 13.20 with it in every round
        // and 13.48 with it only in rounds of one Hamiltonian.
        bool same_h = true;
        for (const DeviceHam* x : hin) same_h = same_h && x == hin[0];
        if (io.hams) launch_hv_variants(hin.data(), prec, (int)order.size(), vin.data(), vout.data(), hv_stream);
        else launch_hv_set(h, prec, (int)order.size(), vin.data(), vout.data(), hv_stream);
        SM_HIP(hipEventRecord(evh, hv_stream));
        std::vector<int> next, live;
        for (int b : active) {
            try {
                SM_HIP(hipStreamWaitEvent(S[b]->stream(), evh, 0));
                S[b]->pre_advance(same_h);
                live.push_back(b);
            } catch (...) {
                fail(b);
            }
        }
        for (int b : live) {
            try {
                if (S[b]->advance()) {
                    try {
                        S[b]->end(io.evals[b], io.evecs ? io.evecs[b] : nullptr);
                    } catch (const Inconsistent&) {
                        // redo this member alone on its own stream and workspace with
                        // reorthogonalisation at every step
                        SolveParams pf = io.p[b];
                        pf.reorth = SMGPU_REORTH_FULL;
                        io.st[b] = SolveStats{};
                        Solver<TBv> R(io.hams ? *io.hams[b] : h, pf, io.st[b], ctx[b]);
                        R.run(io.evals[b], io.evecs ? io.evecs[b] : nullptr, io.v0 ? io.v0[b] : nullptr);
                        io.st[b].n_fallback = 1;
                    }
                    busy[b] = 0;
                    if (on_done) on_done(b);
                } else {
                    next.push_back(b);
                }
            } catch (...) {
                fail(b);
            }
        }
        active.swap(next);
    }
    SM_HIP(hipStreamSynchronize(hv_stream));
    for (int b = 0; b < cap; b++)
        if (S[b]) SM_HIP(hipStreamSynchronize(S[b]->stream()));
}

void trl_solve_batch(DeviceHam& h, int cap, int nb, const BatchIO& io, const SolveCtx* const* ctx,
                     hipStream_t hv_stream, const std::function<bool(int)>& join,
                     const std::function<void(int)>& on_done) {
    SM_HIP(hipSetDevice(h.device));
    if (cap < 1 || nb < 0 || nb > cap) throw std::invalid_argument("bad batch size");
    for (int b = 1; b < nb; b++)
        if ((io.p[b].precision == 0) != (io.p[0].precision == 0))
            throw std::invalid_argument("a batch shares one H·v precision");
    for (int b = 0; b < cap; b++)
        if (!ctx[b] || !ctx[b]->stream) throw std::invalid_argument("each batch member needs its own stream");
    const bool fp64 = nb > 0 ? io.p[0].precision == 0 : false;
    if (fp64) batch_run<double>(h, cap, nb, io, ctx, hv_stream, join, on_done);
    else batch_run<float>(h, cap, nb, io, ctx, hv_stream, join, on_done);
}

void trbl_solve(DeviceHam& h, int q, const SolveParams& p, double* evals, double* evecs, SolveStats& st,
                const double* v0, const JProj* jp) {
    SM_HIP(hipSetDevice(h.device));
    if (jp && (jp->j2 == nullptr || jp->j2->mdim != h.mdim))
        throw std::invalid_argument("J^2 handle does not match the Hamiltonian's basis");
    if (p.precision == 0) {
        BlockSolver<double> S(h, p, st, q, jp);
        S.run(evals, evecs, v0);
    } else {
        BlockSolver<float> S(h, p, st, q, jp);
        S.run(evals, evecs, v0);
    }
}

void restart_small_device(int device, const double* alpha, const double* beta, int m, int ls, double* S,
                          double* alpha_out, double* beta_out) {
    SM_HIP(hipSetDevice(device));
    DBuf a, b, Sd, vals, Y, Q, w;
    a.alloc((m + 1) * 8); b.alloc((m + 1) * 8); Sd.alloc((size_t)m * ls * 8);
    vals.alloc(ls * 8); Y.alloc((size_t)ls * m * 8); Q.alloc((size_t)ls * ls * 8);
    w.alloc((size_t)(ls * 5 * m + 5 * ls + ls * ls + 64) * 8);
    SM_HIP(hipMemcpy(a.p, alpha, m * 8, hipMemcpyHostToDevice));
    SM_HIP(hipMemcpy(b.p, beta, m * 8, hipMemcpyHostToDevice));
    restart_small_on_stream(a.as<double>(), b.as<double>(), m, ls, Sd.as<double>(), vals.as<double>(), Y.as<double>(),
                            Q.as<double>(), w.as<double>(), 0);
    SM_HIP(hipDeviceSynchronize());
    SM_HIP(hipMemcpy(S, Sd.p, (size_t)m * ls * 8, hipMemcpyDeviceToHost));
    SM_HIP(hipMemcpy(alpha_out, a.p, ls * 8, hipMemcpyDeviceToHost));
    SM_HIP(hipMemcpy(beta_out, b.p, ls * 8, hipMemcpyDeviceToHost));
}

void lanczos_steps(DeviceHam& h, const SolveParams& p, int nsteps, const double* v0, double* alpha, double* beta) {
    SM_HIP(hipSetDevice(h.device));
    SolveStats st{};
    if (p.precision == 0) {
        Solver<double> S(h, p, st, nullptr);
        S.run_steps(nsteps, v0, alpha, beta);
    } else {
        Solver<float> S(h, p, st, nullptr);
        S.run_steps(nsteps, v0, alpha, beta);
    }
}

// ------------------------------------------------ lanczos_steps without allocation
namespace {
constexpr int MAX_DEVICES = 64;
constexpr size_t SCRATCH_ALIGN = 256;

struct ScratchSlot {
    std::mutex mu;          // held for a whole call: calls on one device run one at a time
    void* ptr = nullptr;    // device address of `scratch` on this device
};
ScratchSlot g_scratch[MAX_DEVICES];

// the device address of this device's copy of `scratch`, looked up on the first
// call on the device (that device current); the caller holds the slot's mutex
void* scratch_on(int device) {
    ScratchSlot& sl = g_scratch[device];
    if (!sl.ptr) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            throw std::runtime_error(std::string("HIP error ") + hipGetErrorString(e) +
                                     " (hipGetSymbolAddress of the Lanczos scratch)");
        sl.ptr = p;
    }
    return sl.ptr;
}

// Points every buffer of the solve plan (and the 8-int flag ring k_ritz writes) at
// consecutive 256-byte aligned pieces of base; returns the bytes used.  base null:
// only counts.
template <typename TBv>
size_t carve(Workspace& W, int64_t n, const SolveParams& p, unsigned char* base) {
    size_t off = 0;
    auto take = [&](size_t b) {
        const size_t at = off;
        off += (b + SCRATCH_ALIGN - 1) / SCRATCH_ALIGN * SCRATCH_ALIGN;
        return at;
    };
    workspace_plan<TBv>(n, p, [&](DBuf Workspace::*m, int64_t b) {
        const size_t at = take((size_t)b);
        if (base) (W.*m).view(base + at, (size_t)b);
    });
    const size_t at = take(8 * sizeof(int));
    if (base) W.hflag = reinterpret_cast<int*>(base + at);
    return off;
}

template <typename TBv>
void steps_in_scratch(DeviceHam& h, const SolveParams& p, int nsteps, const double* v0, double* alpha,
                      double* beta) {
    Workspace W;
    W.view = true;
    const size_t need = carve<TBv>(W, h.mdim, p, nullptr);
    if (need > LANCZOS_SCRATCH_BYTES)
        throw std::invalid_argument("lanczos_steps_noalloc: mdim " + std::to_string(h.mdim) + " with " +
                                    std::to_string(nsteps) + " steps needs " + std::to_string(need) +
                                    " bytes of device scratch, above its maximum of " +
                                    std::to_string(LANCZOS_SCRATCH_BYTES));
    ScratchSlot& sl = g_scratch[h.device];
    std::lock_guard<std::mutex> lk(sl.mu);
    carve<TBv>(W, h.mdim, p, static_cast<unsigned char*>(scratch_on(h.device)));
    SolveStats st{};
    Solver<TBv> S(h, p, st, W, h.stream);
    S.run_steps(nsteps, v0, alpha, beta);
}
}  // namespace

int64_t lanczos_scratch_max() { return (int64_t)LANCZOS_SCRATCH_BYTES; }

void lanczos_steps_noalloc(DeviceHam& h, const SolveParams& p, int nsteps, const double* v0, double* alpha,
                           double* beta) {
    if (h.device < 0 || h.device >= MAX_DEVICES) throw std::invalid_argument("device index out of range");
    SM_HIP(hipSetDevice(h.device));
    if (p.precision == 0) steps_in_scratch<double>(h, p, nsteps, v0, alpha, beta);
    else steps_in_scratch<float>(h, p, nsteps, v0, alpha, beta);
}

}  // namespace smgpu
