// Thick-restart Lanczos on the device.
//
// Reference algorithm: TRL (src/ShellModel/lanczos_methods.jl:1-91) with
// diagonalize_T! (:250-258), ReORTH (src/ShellModel/shellmodel_main.jl:1005-1012),
// ThickRestart (:260-307) and the Ritz-vector assembly of main_sm (:208-224).
//
// Per iteration (it = index of the current basis vector v):
//   w  = H v                                   (FP32 or FP64 H·v kernel)
//   h  = V[0..it]^T w;   w -= V h              (classical Gram-Schmidt, FP64 sums)
//   h2 = V[0..it]^T w;   w -= V h2             (second pass: "twice is enough")
//   alpha[it] = h[it]  (= v.Hv, the reference's dot before ReORTH)
//   beta[it]  = |w|,  V[it+1] = w / beta
//   Ritz values of the tridiagonal T[0..it] by multisection bisection (FP64),
//   convergence flag: all(en_prev - en_new < tol), the reference's rule.
// Restart keeps T tridiagonal: the kept Ritz vectors are rotated within their
// span (tri_restart) so the projected matrix stays tridiagonal; its eigenvalues
// equal those of the reference's arrowhead Tmat.
// Precision 0: basis, H·v and all vectors FP64 (the plain FP64 path).
// Precision 1: FP32 H·v and FP32 basis storage; dot products, alpha/beta, T,
// the small eigen-problems and the restart rotation coefficients are FP64;
// the final Ritz vectors are assembled in FP64 and their eigenvalues refined by
// an FP64 Rayleigh quotient (one FP64 H·v per eigenpair).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <random>
#include <stdexcept>

#include "sm_device.h"
#include "sm_trl.h"
#include "tridiag.h"

namespace smgpu {

static constexpr int TB = 256;

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

// part[blk*nv + j] = sum over the block's chunk of B[j][x] * w[x]
template <typename TBv, typename TW>
__global__ __launch_bounds__(TB) void k_dots(const TBv* __restrict__ B, int64_t ldb, int nv,
                                             const TW* __restrict__ w, int64_t n, int64_t chunk,
                                             double* __restrict__ part) {
    __shared__ double red[TB / 64];
    const int64_t x0 = (int64_t)blockIdx.x * chunk;
    const int64_t x1 = min(n, x0 + chunk);
    for (int j = 0; j < nv; j++) {
        const TBv* b = B + (int64_t)j * ldb;
        double s = 0.0;
        for (int64_t x = x0 + threadIdx.x; x < x1; x += TB) s += (double)b[x] * (double)w[x];
        s = block_sum(s, red);
        if (threadIdx.x == 0) part[(int64_t)blockIdx.x * nv + j] = s;
    }
}

// out[j] = sum_g part[g*nv + j]  (+ optional sqrt into out2 for j == 0)
__global__ __launch_bounds__(TB) void k_reduce(const double* __restrict__ part, int G, int nv,
                                               double* __restrict__ out) {
    __shared__ double red[TB / 64];
    const int j = blockIdx.x;
    double s = 0.0;
    for (int g = threadIdx.x; g < G; g += TB) s += part[(int64_t)g * nv + j];
    s = block_sum(s, red);
    if (threadIdx.x == 0) out[j] = s;
}

// wout[x] = win[x] - sum_j h[j] B[j][x];  part[blk] = sum wout^2 over the chunk
template <typename TBv, typename TW>
__global__ __launch_bounds__(TB) void k_update(const TBv* __restrict__ B, int64_t ldb, int nv,
                                               const double* __restrict__ h, const TW* __restrict__ win,
                                               double* __restrict__ wout, int64_t n, int64_t chunk,
                                               double* __restrict__ part) {
    __shared__ double hs[1024];
    __shared__ double red[TB / 64];
    for (int j = threadIdx.x; j < nv; j += TB) hs[j] = h[j];
    __syncthreads();
    const int64_t x0 = (int64_t)blockIdx.x * chunk;
    const int64_t x1 = min(n, x0 + chunk);
    double nr = 0.0;
    for (int64_t x = x0 + threadIdx.x; x < x1; x += TB) {
        double a = (double)win[x];
        for (int j = 0; j < nv; j++) a -= hs[j] * (double)B[(int64_t)j * ldb + x];
        wout[x] = a;
        nr += a * a;
    }
    nr = block_sum(nr, red);
    if (threadIdx.x == 0) part[blockIdx.x] = nr;
}

// alpha/beta bookkeeping after the two Gram-Schmidt passes
__global__ void k_finish(const double* h1, int it, const double* part, int G, double* alpha, double* beta) {
    __shared__ double red[TB / 64];
    double s = 0.0;
    for (int g = threadIdx.x; g < G; g += TB) s += part[g];
    s = block_sum(s, red);
    if (threadIdx.x == 0) {
        alpha[it] = h1[it];
        beta[it] = sqrt(s);
    }
}

template <typename TBv>
__global__ __launch_bounds__(TB) void k_store(const double* __restrict__ w, const double* __restrict__ beta, int it,
                                              TBv* __restrict__ dst, int64_t n) {
    const double inv = 1.0 / beta[it];
    for (int64_t x = (int64_t)blockIdx.x * TB + threadIdx.x; x < n; x += (int64_t)gridDim.x * TB)
        dst[x] = (TBv)(w[x] * inv);
}

template <typename TA, typename TBv>
__global__ __launch_bounds__(TB) void k_convert(const TA* __restrict__ a, TBv* __restrict__ b, int64_t n, double scale) {
    for (int64_t x = (int64_t)blockIdx.x * TB + threadIdx.x; x < n; x += (int64_t)gridDim.x * TB)
        b[x] = (TBv)((double)a[x] * scale);
}

// Ritz values: block j computes eigenvalue j of T[0..n-1] by 256-way multisection.
__global__ __launch_bounds__(TB) void k_ritz(const double* __restrict__ alpha, const double* __restrict__ beta,
                                             int n, double* __restrict__ out) {
    __shared__ double sa[1024], sb[1024];
    __shared__ int firstbig;
    __shared__ double slo, shi;
    const int j = blockIdx.x;
    for (int i = threadIdx.x; i < n; i += TB) { sa[i] = alpha[i]; sb[i] = i < n - 1 ? beta[i] : 0.0; }
    __syncthreads();
    if (threadIdx.x == 0) {
        double lo, hi;
        tri_bounds(sa, sb, n, &lo, &hi);
        slo = lo; shi = hi;
    }
    __syncthreads();
    const double pm = tri_pivmin(sb, n);
    for (int round = 0; round < 12; round++) {
        const double lo = slo, hi = shi;
        if (threadIdx.x == 0) firstbig = TB;
        __syncthreads();
        const double x = lo + (hi - lo) * (double)(threadIdx.x + 1) / (double)(TB + 1);
        const int c = tri_count(sa, sb, n, x, pm);
        if (c > j) atomicMin(&firstbig, (int)threadIdx.x);
        __syncthreads();
        if (threadIdx.x == 0) {
            const int t = firstbig;
            const double nlo = t == 0 ? lo : lo + (hi - lo) * (double)t / (double)(TB + 1);
            const double nhi = t == TB ? hi : lo + (hi - lo) * (double)(t + 1) / (double)(TB + 1);
            slo = nlo; shi = nhi;
        }
        __syncthreads();
        if (shi - slo <= 8.0 * 2.220446049250313e-16 * fmax(1.0, fmax(fabs(slo), fabs(shi)))) break;
    }
    if (threadIdx.x == 0) out[j] = 0.5 * (slo + shi);
}

// history shift + convergence flag (diagonalize_T!): en[h][num_ev], h = 0 newest
__global__ void k_conv(double* en, int nh, int num_ev, const double* newv, int n, double tol, int* flag) {
    if (threadIdx.x != 0) return;
    for (int h = nh - 1; h > 0; h--)
        for (int k = 0; k < num_ev; k++) en[h * num_ev + k] = en[(h - 1) * num_ev + k];
    for (int k = 0; k < n; k++) en[k] = newv[k];
    int ok = 1;
    for (int k = 0; k < num_ev; k++) ok &= (en[num_ev + k] - en[k] < tol) ? 1 : 0;
    *flag = ok;
}

// restart / final small problems, one thread (rare, O(lm * ls) work)
__global__ void k_restart_small(double* alpha, double* beta, int m, int ls, double* S, double* work) {
    if (threadIdx.x == 0 && blockIdx.x == 0) tri_restart(alpha, beta, m, ls, S, work);
}
__global__ void k_lowest_small(const double* alpha, const double* beta, int n, int k, double* vals, double* Y,
                               double* work) {
    if (threadIdx.x == 0 && blockIdx.x == 0) tri_lowest(alpha, beta, n, k, vals, Y, work);
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

// ------------------------------------------------------------------ driver
namespace {

struct Workspace {
    DBuf basis, wH, w64, part, h1, h2, alpha, beta, en, newv, flag, S, swork, U, ritz;
};

int64_t grid_chunks(int64_t n, int64_t* chunk) {
    int64_t G = std::max<int64_t>(1, std::min<int64_t>(1024, (n + 4 * TB - 1) / (4 * TB)));
    *chunk = ((n + G - 1) / G + TB - 1) / TB * TB;
    G = (n + *chunk - 1) / *chunk;
    return std::max<int64_t>(G, 1);
}

template <typename TBv, typename TW>
void cgs_pass(const TBv* B, int64_t ldb, int nv, const TW* w_in, double* w_out, int64_t n, double* part,
              double* h, hipStream_t s, double* norm_part, int64_t* G_out) {
    int64_t chunk;
    int64_t G = grid_chunks(n, &chunk);
    hipLaunchKernelGGL((k_dots<TBv, TW>), dim3(G), dim3(TB), 0, s, B, ldb, nv, w_in, n, chunk, part);
    hipLaunchKernelGGL(k_reduce, dim3(nv), dim3(TB), 0, s, part, (int)G, nv, h);
    hipLaunchKernelGGL((k_update<TBv, TW>), dim3(G), dim3(TB), 0, s, B, ldb, nv, h, w_in, w_out, n, chunk, norm_part);
    *G_out = G;
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
    Workspace W;

    Solver(DeviceHam& h, const SolveParams& p, SolveStats& stats) : H(h), P(p), st(stats) {
        s = H.stream;
        n = H.mdim;
        ld = (n + 63) / 64 * 64;
        lm = P.lm;
        ls = P.ls;
        nev = (int)std::min<int64_t>(P.num_ev, n);
        prec_hv = sizeof(TBv) == 4 ? 1 : 0;
    }

    TBv* V(int j) { return W.basis.as<TBv>() + (int64_t)j * ld; }

    void hv_into(int j) {   // W.wH = H V(j)
        if (prec_hv == 0) launch_hv(H, 0, V(j), W.wH.p, s);
        else launch_hv(H, 1, V(j), W.wH.p, s);
    }

    void orth(int nv, bool from_hv) {
        int64_t G;
        double* part = W.part.as<double>();
        double* npart = part + (int64_t)1024 * (lm + 1);
        if (from_hv) {
            if (prec_hv == 0)
                cgs_pass<TBv, double>((const TBv*)V(0), ld, nv, W.wH.as<double>(), W.w64.as<double>(), n, part,
                                      W.h1.as<double>(), s, npart, &G);
            else
                cgs_pass<TBv, float>((const TBv*)V(0), ld, nv, W.wH.as<float>(), W.w64.as<double>(), n, part,
                                     W.h1.as<double>(), s, npart, &G);
        } else {
            cgs_pass<TBv, double>((const TBv*)V(0), ld, nv, W.w64.as<double>(), W.w64.as<double>(), n, part,
                                  W.h1.as<double>(), s, npart, &G);
        }
        cgs_pass<TBv, double>((const TBv*)V(0), ld, nv, W.w64.as<double>(), W.w64.as<double>(), n, part,
                              W.h2.as<double>(), s, npart, &G);
        last_G = G;
    }
    int64_t last_G = 1;

    void run(double* evals, double* evecs) {
        auto t0 = std::chrono::steady_clock::now();
        if (lm < 3 || ls < 1 || ls > 32 || ls >= lm - 1 || nev < 1 || nev > ls)
            throw std::invalid_argument("need 1 <= num_ev <= ls <= 32 and ls < lm - 1");
        if (lm > 1000) throw std::invalid_argument("lm above 1000 is not supported");
        int nh = std::max(2, P.num_history);
        W.basis.alloc((size_t)lm * ld * sizeof(TBv));
        W.wH.alloc((size_t)ld * (prec_hv ? 4 : 8));
        W.w64.alloc((size_t)ld * 8);
        W.part.alloc((size_t)1024 * (lm + 2) * 8);
        W.h1.alloc((lm + 1) * 8); W.h2.alloc((lm + 1) * 8);
        W.alpha.alloc((lm + 1) * 8); W.beta.alloc((lm + 1) * 8);
        W.en.alloc((size_t)nh * nev * 8); W.newv.alloc((lm + 1) * 8); W.flag.alloc(64);
        W.S.alloc((size_t)lm * 32 * 8);
        W.swork.alloc((size_t)(5 * lm + 32 * (lm + 3 * 32 + 8) + 64) * 8);
        W.U.alloc((size_t)32 * ld * sizeof(TBv));
        SM_HIP(hipMemsetAsync(W.basis.p, 0, W.basis.bytes, s));
        {
            std::vector<double> en0((size_t)nh * nev, 1.0e4);
            SM_HIP(hipMemcpyAsync(W.en.p, en0.data(), en0.size() * 8, hipMemcpyHostToDevice, s));
        }
        // start vector: seeded normal, normalised in FP64 (initialize_wf, :939-948)
        {
            std::mt19937_64 rng(P.seed);
            std::normal_distribution<double> nd(0.0, 1.0);
            std::vector<double> v0(n);
            double nr = 0.0;
            for (int64_t i = 0; i < n; i++) { v0[i] = nd(rng); nr += v0[i] * v0[i]; }
            SM_HIP(hipMemcpyAsync(W.w64.p, v0.data(), n * 8, hipMemcpyHostToDevice, s));
            int gb = (int)std::min<int64_t>(4096, (n + TB - 1) / TB);
            hipLaunchKernelGGL((k_convert<double, TBv>), dim3(gb), dim3(TB), 0, s, W.w64.as<double>(), V(0), n,
                               1.0 / std::sqrt(nr));
            SM_HIP(hipStreamSynchronize(s));
        }
        int* hflag = nullptr;
        SM_HIP(hipHostMalloc((void**)&hflag, 64, hipHostMallocDefault));
        int k = 0;
        bool conv = false;
        int elit = 1;
        int64_t total_steps = 0;
        st.n_hv = 0; st.n_restart = 0;
        int gstore = (int)std::min<int64_t>(4096, (n + TB - 1) / TB);
        while (!conv) {
            for (int it = k; it < lm - 1; it++) {
                hv_into(it);
                st.n_hv++;
                orth(it + 1, true);
                hipLaunchKernelGGL(k_finish, dim3(1), dim3(TB), 0, s, W.h1.as<double>(), it,
                                   W.part.as<double>() + (int64_t)1024 * (lm + 1), (int)last_G,
                                   W.alpha.as<double>(), W.beta.as<double>());
                int nn = std::min(nev, it + 1);
                hipLaunchKernelGGL(k_ritz, dim3(nn), dim3(TB), 0, s, W.alpha.as<double>(), W.beta.as<double>(),
                                   it + 1, W.newv.as<double>());
                hipLaunchKernelGGL(k_conv, dim3(1), dim3(64), 0, s, W.en.as<double>(), nh, nev,
                                   W.newv.as<double>(), nn, P.tol, W.flag.as<int>());
                SM_HIP(hipMemcpyAsync(hflag, W.flag.p, sizeof(int), hipMemcpyDeviceToHost, s));
                SM_HIP(hipStreamSynchronize(s));
                if (*hflag || it + 1 == n) { conv = true; elit = it + 1; break; }
                if (++total_steps >= P.max_iter) throw std::runtime_error("TRL: iteration limit reached");
                hipLaunchKernelGGL((k_store<TBv>), dim3(gstore), dim3(TB), 0, s, W.w64.as<double>(),
                                   W.beta.as<double>(), it, V(it + 1), n);
            }
            if (conv) break;
            restart();
            st.n_restart++;
            k = ls;
        }
        st.iterations = elit;
        // Ritz values and vectors (main_sm :208-224)
        finish(elit, evals, evecs);
        SM_HIP(hipHostFree(hflag));
        st.wall_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }

    void restart() {
        int m = lm - 1;
        hipLaunchKernelGGL(k_restart_small, dim3(1), dim3(64), 0, s, W.alpha.as<double>(), W.beta.as<double>(),
                           m, ls, W.S.as<double>(), W.swork.as<double>());
        int gb = (int)std::min<int64_t>(2048, (n + TB - 1) / TB);
        size_t sh = (size_t)ls * m * 8;
        if (sh > 64 * 1024) throw std::invalid_argument("ls*(lm-1) too large for the restart rotation");
        hipLaunchKernelGGL((k_rotate<TBv, TBv>), dim3(gb), dim3(TB), sh, s, (const TBv*)V(0), ld, m,
                           W.S.as<double>(), ls, W.U.as<TBv>(), ld, n);
        // v_ls = v_{lm-1}; kept vectors from U
        SM_HIP(hipMemcpyAsync(V(ls), V(lm - 1), n * sizeof(TBv), hipMemcpyDeviceToDevice, s));
        SM_HIP(hipMemcpyAsync(V(0), W.U.p, (size_t)ls * ld * sizeof(TBv), hipMemcpyDeviceToDevice, s));
        // re-orthonormalise the kept vectors (the reference's make_sure pass)
        int gstore = (int)std::min<int64_t>(4096, (n + TB - 1) / TB);
        for (int j = 0; j < ls; j++) {
            hipLaunchKernelGGL((k_convert<TBv, double>), dim3(gstore), dim3(TB), 0, s, (const TBv*)V(j),
                               W.w64.as<double>(), n, 1.0);
            if (j > 0) {
                int64_t G;
                double* part = W.part.as<double>();
                double* npart = part + (int64_t)1024 * (lm + 1);
                cgs_pass<TBv, double>((const TBv*)V(0), ld, j, W.w64.as<double>(), W.w64.as<double>(), n, part,
                                      W.h1.as<double>(), s, npart, &G);
            }
            norm_store(j);
        }
        // v_ls against the kept vectors (no renormalisation, as the reference)
        {
            hipLaunchKernelGGL((k_convert<TBv, double>), dim3(gstore), dim3(TB), 0, s, (const TBv*)V(ls),
                               W.w64.as<double>(), n, 1.0);
            int64_t G;
            double* part = W.part.as<double>();
            double* npart = part + (int64_t)1024 * (lm + 1);
            cgs_pass<TBv, double>((const TBv*)V(0), ld, ls, W.w64.as<double>(), W.w64.as<double>(), n, part,
                                  W.h1.as<double>(), s, npart, &G);
            hipLaunchKernelGGL((k_convert<double, TBv>), dim3(gstore), dim3(TB), 0, s, W.w64.as<double>(), V(ls),
                               n, 1.0);
        }
        SM_HIP(hipMemsetAsync(V(ls + 1), 0, (size_t)(lm - ls - 1) * ld * sizeof(TBv), s));
    }

    // V(j) = w64 / |w64|
    void norm_store(int j) {
        int64_t chunk;
        int64_t G = grid_chunks(n, &chunk);
        double* part = W.part.as<double>();
        hipLaunchKernelGGL((k_dots<double, double>), dim3(G), dim3(TB), 0, s, W.w64.as<double>(), ld, 1,
                           W.w64.as<double>(), n, chunk, part);
        hipLaunchKernelGGL(k_finish, dim3(1), dim3(TB), 0, s, W.h1.as<double>(), 0, part, (int)G,
                           W.newv.as<double>(), W.newv.as<double>() + 1);
        int gstore = (int)std::min<int64_t>(4096, (n + TB - 1) / TB);
        hipLaunchKernelGGL((k_store<TBv>), dim3(gstore), dim3(TB), 0, s, W.w64.as<double>(),
                           W.newv.as<double>() + 1, 0, V(j), n);
    }

    void finish(int elit, double* evals, double* evecs) {
        // lowest nev eigenpairs of T[0..elit-1]
        DBuf vals, Y, work;
        vals.alloc(nev * 8);
        Y.alloc((size_t)nev * elit * 8);
        work.alloc((size_t)(5 * elit + 8) * 8);
        hipLaunchKernelGGL(k_lowest_small, dim3(1), dim3(64), 0, s, W.alpha.as<double>(), W.beta.as<double>(),
                           elit, nev, vals.as<double>(), Y.as<double>(), work.as<double>());
        std::vector<double> ritz(nev);
        SM_HIP(hipMemcpyAsync(ritz.data(), vals.p, nev * 8, hipMemcpyDeviceToHost, s));
        // the converged Ritz values of the last check (en[1])
        std::vector<double> en_last(nev);
        SM_HIP(hipMemcpyAsync(en_last.data(), W.en.p, nev * 8, hipMemcpyDeviceToHost, s));
        SM_HIP(hipStreamSynchronize(s));
        for (int i = 0; i < nev; i++) { evals[i] = en_last[i]; st.ritz[i] = en_last[i]; }
        bool need_vecs = evecs != nullptr || (prec_hv == 1 && P.refine);
        if (!need_vecs) return;
        DBuf R;
        R.alloc((size_t)nev * ld * 8);
        int gb = (int)std::min<int64_t>(2048, (n + TB - 1) / TB);
        size_t sh = (size_t)nev * elit * 8;
        if (sh > 64 * 1024) throw std::invalid_argument("num_ev*iterations too large for the Ritz assembly");
        hipLaunchKernelGGL((k_rotate<TBv, double>), dim3(gb), dim3(TB), sh, s, (const TBv*)V(0), ld, elit,
                           Y.as<double>(), nev, R.as<double>(), ld, n);
        // normalise (Rvec .*= 1/sqrt(dot))
        int64_t chunk;
        int64_t G = grid_chunks(n, &chunk);
        double* part = W.part.as<double>();
        std::vector<double> nrm(nev), rq(nev);
        DBuf hv64;
        if (prec_hv == 1 && P.refine) hv64.alloc((size_t)ld * 8);
        for (int i = 0; i < nev; i++) {
            double* r = R.as<double>() + (int64_t)i * ld;
            hipLaunchKernelGGL((k_dots<double, double>), dim3(G), dim3(TB), 0, s, r, ld, 1, r, n, chunk, part);
            hipLaunchKernelGGL(k_reduce, dim3(1), dim3(TB), 0, s, part, (int)G, 1, W.newv.as<double>() + i);
        }
        SM_HIP(hipMemcpyAsync(nrm.data(), W.newv.p, nev * 8, hipMemcpyDeviceToHost, s));
        SM_HIP(hipStreamSynchronize(s));
        int gs = (int)std::min<int64_t>(4096, (n + TB - 1) / TB);
        for (int i = 0; i < nev; i++) {
            double* r = R.as<double>() + (int64_t)i * ld;
            hipLaunchKernelGGL((k_convert<double, double>), dim3(gs), dim3(TB), 0, s, r, r, n, 1.0 / std::sqrt(nrm[i]));
            if (prec_hv == 1 && P.refine) {
                // FP64 Rayleigh quotient with the FP64 H·v
                launch_hv(H, 0, r, hv64.p, s);
                st.n_hv64++;
                hipLaunchKernelGGL((k_dots<double, double>), dim3(G), dim3(TB), 0, s, r, ld, 1, hv64.as<double>(), n,
                                   chunk, part);
                hipLaunchKernelGGL(k_reduce, dim3(1), dim3(TB), 0, s, part, (int)G, 1, W.newv.as<double>() + i);
            }
        }
        if (prec_hv == 1 && P.refine) {
            SM_HIP(hipMemcpyAsync(rq.data(), W.newv.p, nev * 8, hipMemcpyDeviceToHost, s));
            SM_HIP(hipStreamSynchronize(s));
            for (int i = 0; i < nev; i++) evals[i] = rq[i];
            // keep ascending order
            std::sort(evals, evals + nev);
        }
        if (evecs) {
            for (int i = 0; i < nev; i++)
                SM_HIP(hipMemcpyAsync(evecs + (int64_t)i * n, R.as<double>() + (int64_t)i * ld, n * 8,
                                      hipMemcpyDeviceToHost, s));
            SM_HIP(hipStreamSynchronize(s));
        }
    }
};

}  // namespace

void trl_solve(DeviceHam& h, const SolveParams& p, double* evals, double* evecs, SolveStats& st) {
    SM_HIP(hipSetDevice(h.device));
    if (p.precision == 0) {
        Solver<double> S(h, p, st);
        S.run(evals, evecs);
    } else {
        Solver<float> S(h, p, st);
        S.run(evals, evecs);
    }
}

}  // namespace smgpu
