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
#include <algorithm>
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

    // Lowest nv eigenvalues (ascending) of the leading k x k block of T (row stride ldt)
    // without eigenvectors.  A banded T (lower bandwidth bw < k / 4: the block
    // tridiagonal projection of the block Lanczos method, bw ~ q, or ~ ls + q with the
    // arrowhead of a thick restart) is reduced to tridiagonal form by Givens rotations
    // in O(k^2 bw) (band_reduce); otherwise by the O(k^3) Householder reduction.  Both
    // are orthogonal similarities, so the values agree to rounding.
    // hint (optional): the previous step's values, upper bounds of these by Cauchy
    // interlacing (rows were only added), used to bracket the bisections.
    void lowest_values(const double* T, int ldt, int k, int nv, double* vals, const double* hint = nullptr) {
        int bw = 0;
        for (int i = 0; i < k; i++)
            for (int j = 0; j < i - bw; j++)
                if (T[(size_t)i * ldt + j] != 0.0) { bw = i - j; break; }
        if (k >= 8 && 4 * bw < k) band_reduce(T, ldt, k, bw);
        else reduce(T, ldt, k);
        lowest(nv, vals, hint);
    }

    // Band to tridiagonal (Rutishauser 1963, Schwarz 1968): for m = bw .. 2, the
    // outermost diagonal is annihilated column by column with a rotation of rows and
    // columns (c + m - 1, c + m); each rotation fills one element outside the band at
    // distance m + 1 further down, which the next rotation (m rows lower) annihilates in
    // turn.  Columns are swept in increasing order, so the element left of each
    // rotation is already zero and no other fill appears.  Only the band window of the
    // two rows and columns is updated.  Sets d, e (tau = 0: no eigenvectors).
    void band_reduce(const double* T, int ldt, int k, int bw) {
        n = k;
        a.assign((size_t)n * n, 0.0);
        for (int i = 0; i < n; i++)
            for (int j = std::max(0, i - bw); j <= i; j++)
                a[(size_t)i * n + j] = a[(size_t)j * n + i] = T[(size_t)i * ldt + j];
        auto A = [&](int i, int j) -> double& { return a[(size_t)i * n + j]; };
        for (int m = bw; m > 1; m--) {
            for (int c0 = 0; c0 + m < n; c0++) {
                int col = c0, p = c0 + m - 1, q = c0 + m;
                for (;;) {
                    const double x = A(p, col), y = A(q, col);
                    if (y != 0.0) {
                        const double r = std::hypot(x, y), c = x / r, sn = y / r;
                        const int lo = std::max(0, p - m - 1), hi = std::min(n - 1, q + m + 1);
                        for (int t = lo; t <= hi; t++) {
                            const double ap = A(p, t), aq = A(q, t);
                            A(p, t) = c * ap + sn * aq;
                            A(q, t) = -sn * ap + c * aq;
                        }
                        for (int t = lo; t <= hi; t++) {
                            const double ap = A(t, p), aq = A(t, q);
                            A(t, p) = c * ap + sn * aq;
                            A(t, q) = -sn * ap + c * aq;
                        }
                        A(q, col) = 0.0;
                        A(col, q) = 0.0;
                    }
                    if (q + m >= n) break;
                    col = p;
                    p += m;
                    q += m;
                }
            }
        }
        d.assign(n, 0.0);
        e.assign(n > 1 ? n - 1 : 1, 0.0);
        tau.assign(n, 0.0);
        for (int i = 0; i < n; i++) d[i] = A(i, i);
        for (int i = 0; i + 1 < n; i++) e[i] = A(i + 1, i);
    }

    // lowest k eigenvalues (ascending).  With hint[j] an upper bound of eigenvalue j
    // (checked by a Sturm count), its bisection starts from [hint - delta, hint + pad]
    // with delta grown until the count confirms the lower end; otherwise from the
    // Gershgorin interval.
    void lowest(int k, double* vals, const double* hint = nullptr) const {
        if (n == 1) { vals[0] = d[0]; return; }
        double lo, hi;
        tri_bounds(d.data(), e.data(), n, &lo, &hi);
        const double pm = tri_pivmin(e.data(), n);
        const double scale = std::fmax(1.0, std::fmax(std::fabs(lo), std::fabs(hi)));
        for (int j = 0; j < k; j++) {
            double l = lo, h = hi;
            if (hint && hint[j] > lo && hint[j] < hi) {
                const double hh = std::fmin(hi, hint[j] + 1e-9 * scale);
                if (tri_count(d.data(), e.data(), n, hh, pm) > j) {
                    double del = 1e-8 * scale;
                    for (int t = 0; t < 24; t++) {
                        const double ll = std::fmax(lo, hh - del);
                        if (tri_count(d.data(), e.data(), n, ll, pm) <= j) { l = ll; h = hh; break; }
                        if (ll <= lo) break;
                        del *= 16.0;
                    }
                }
            }
            vals[j] = tri_bisect(d.data(), e.data(), n, j, l, h, pm);
        }
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
