// Lowest eigenpairs of a small dense symmetric matrix in FP64 on the host: the
// projected matrix of the block Lanczos method, which is block tridiagonal with
// an arrowhead after each thick restart (lanczos_methods.jl:563-617), at most a
// few hundred rows.
//
// Householder reduction to tridiagonal form, A = Q T Q^T with
// Q = H_0 H_1 ... H_{n-3}, H_j = I - tau_j v_j v_j^T acting on rows j+1..n-1;
// eigenvalues of T by bisection, eigenvectors by inverse iteration
// (tridiag.h), back-transformed by Q.  The reference calls LAPACK's
// eigvals/eigen on the same matrix; the eigenvalues agree to rounding.
#pragma once
#include <cmath>
#include <vector>

#include "tridiag.h"

namespace smgpu {

struct DenseEig {
    int n = 0;
    std::vector<double> a;      // n x n; reflector j below the subdiagonal of column j
    std::vector<double> d, e, tau;

    // reduce the leading k x k block of T (row stride ldt)
    void reduce(const double* T, int ldt, int k) {
        n = k;
        a.assign((size_t)n * n, 0.0);
        for (int i = 0; i < n; i++)
            for (int j = 0; j < n; j++) a[(size_t)i * n + j] = T[(size_t)i * ldt + j];
        d.assign(n, 0.0);
        e.assign(n > 1 ? n - 1 : 1, 0.0);
        tau.assign(n, 0.0);
        std::vector<double> p(n), q(n);
        for (int j = 0; j + 2 < n; j++) {
            double* col = a.data() + j;               // column j: element i at col[i*n]
            double s2 = 0.0;
            for (int i = j + 2; i < n; i++) s2 += col[(size_t)i * n] * col[(size_t)i * n];
            const double x0 = col[(size_t)(j + 1) * n];
            if (s2 == 0.0) {                           // already tridiagonal in this column
                e[j] = x0;
                tau[j] = 0.0;
                for (int i = j + 1; i < n; i++) col[(size_t)i * n] = 0.0;
                continue;
            }
            const double xn = std::sqrt(x0 * x0 + s2);
            const double alpha = x0 >= 0.0 ? -xn : xn;
            const double v0 = x0 - alpha;
            col[(size_t)(j + 1) * n] = v0;            // v = x - alpha e_1 (in place)
            const double t = 2.0 / (v0 * v0 + s2);
            tau[j] = t;
            e[j] = alpha;
            // trailing block B (rows/cols j+1..n-1): B <- H B H
            //   p = t B v,  q = p - (t/2)(p.v) v,  B -= v q^T + q v^T
            double pv = 0.0;
            for (int i = j + 1; i < n; i++) {
                const double* row = a.data() + (size_t)i * n;
                double s = 0.0;
                for (int l = j + 1; l < n; l++) s += row[l] * col[(size_t)l * n];
                p[i] = t * s;
                pv += p[i] * col[(size_t)i * n];
            }
            const double K = 0.5 * t * pv;
            for (int i = j + 1; i < n; i++) q[i] = p[i] - K * col[(size_t)i * n];
            for (int i = j + 1; i < n; i++) {
                double* row = a.data() + (size_t)i * n;
                const double vi = col[(size_t)i * n], qi = q[i];
                for (int l = j + 1; l < n; l++) row[l] -= vi * q[l] + qi * col[(size_t)l * n];
            }
        }
        for (int j = 0; j < n; j++) d[j] = a[(size_t)j * n + j];
        if (n >= 2) e[n - 2] = a[(size_t)(n - 1) * n + (n - 2)];
    }

    // lowest k eigenvalues (ascending)
    void lowest(int k, double* vals) const {
        if (n == 1) { vals[0] = d[0]; return; }
        double lo, hi;
        tri_bounds(d.data(), e.data(), n, &lo, &hi);
        const double pm = tri_pivmin(e.data(), n);
        for (int j = 0; j < k; j++) vals[j] = tri_bisect(d.data(), e.data(), n, j, lo, hi, pm);
    }

    // lowest k eigenpairs; vector j at vecs + j*n (unit 2-norm)
    void lowest_pairs(int k, double* vals, double* vecs) const {
        if (n == 1) { vals[0] = d[0]; vecs[0] = 1.0; return; }
        std::vector<double> work(5 * (size_t)n + 8);
        tri_lowest(d.data(), e.data(), n, k, vals, vecs, work.data());
        for (int m = 0; m < k; m++) {
            double* x = vecs + (size_t)m * n;
            for (int j = n - 3; j >= 0; j--) {
                if (tau[j] == 0.0) continue;
                double s = 0.0;
                for (int l = j + 1; l < n; l++) s += a[(size_t)l * n + j] * x[l];
                s *= tau[j];
                for (int l = j + 1; l < n; l++) x[l] -= s * a[(size_t)l * n + j];
            }
        }
    }
};

}  // namespace smgpu
