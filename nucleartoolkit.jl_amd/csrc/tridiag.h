// Symmetric tridiagonal eigen-problems in FP64, written once for host and device.
//
// The Lanczos projection T is kept tridiagonal (the thick restart rotates the
// kept Ritz vectors so that T stays tridiagonal, see tri_restart), so every
// small eigen-problem here is tridiagonal:
//   * tri_count      — Sturm count: number of eigenvalues < x (LDL^T inertia)
//   * tri_bisect     — k-th eigenvalue by bisection
//   * tri_eigvec     — eigenvector by inverse iteration (LU with partial pivoting)
//   * tri_restart    — thick-restart bookkeeping: lowest ls Ritz pairs of T,
//                      rotation of the Ritz basis that makes the new leading
//                      block tridiagonal (Lanczos on diag(theta) from the
//                      coupling vector sigma, order reversed)
// The eigenvalues equal those of the reference's dense `eigvals`/`eigen` of the
// arrowhead Tmat (lanczos_methods.jl:250-307) up to rounding, since both
// matrices are similar.
#pragma once
#include <math.h>

#ifdef __HIPCC__
#define SMHD __host__ __device__
#else
#define SMHD
#endif

namespace smgpu {

// number of eigenvalues of T (alpha[0..n-1], beta[0..n-2]) strictly below x
SMHD inline int tri_count(const double* alpha, const double* beta, int n, double x, double pivmin) {
    int c = 0;
    double d = alpha[0] - x;
    if (fabs(d) < pivmin) d = -pivmin;
    if (d < 0.0) c++;
    for (int i = 1; i < n; i++) {
        d = alpha[i] - x - beta[i - 1] * beta[i - 1] / d;
        if (fabs(d) < pivmin) d = -pivmin;
        if (d < 0.0) c++;
    }
    return c;
}

SMHD inline void tri_bounds(const double* alpha, const double* beta, int n, double* lo, double* hi) {
    double l = 1e300, h = -1e300;
    for (int i = 0; i < n; i++) {
        double r = (i > 0 ? fabs(beta[i - 1]) : 0.0) + (i < n - 1 ? fabs(beta[i]) : 0.0);
        l = fmin(l, alpha[i] - r);
        h = fmax(h, alpha[i] + r);
    }
    double pad = 1e-12 * fmax(1.0, fmax(fabs(l), fabs(h)));
    *lo = l - pad;
    *hi = h + pad;
}

SMHD inline double tri_pivmin(const double* beta, int n) {
    double m = 1.0;
    for (int i = 0; i < n - 1; i++) m = fmax(m, beta[i] * beta[i]);
    return 2.2250738585072014e-308 / 2.220446049250313e-16 * m;
}

// k-th (0-based) eigenvalue, bisection to ~2 ulp of max(|lo|,|hi|)
SMHD inline double tri_bisect(const double* alpha, const double* beta, int n, int k, double lo, double hi,
                              double pivmin) {
    for (int it = 0; it < 200; it++) {
        double mid = 0.5 * (lo + hi);
        if (mid <= lo || mid >= hi) break;
        if (tri_count(alpha, beta, n, mid, pivmin) > k) hi = mid;
        else lo = mid;
        if (hi - lo <= 4.0 * 2.220446049250313e-16 * fmax(fabs(lo), fabs(hi))) break;
    }
    return 0.5 * (lo + hi);
}

// Inverse iteration for the eigenvector of eigenvalue lam.  work: 5n doubles.
// prev/nprev: already computed vectors (stride n) of nearby eigenvalues to
// orthogonalise against.  Output normalised to unit 2-norm.
SMHD inline void tri_eigvec(const double* alpha, const double* beta, int n, double lam, double* x,
                            double* work, const double* prev, int nprev, double tnorm) {
    if (n == 1) { x[0] = 1.0; return; }
    double* dl = work;          // L multipliers
    double* d = work + n;       // U diagonal
    double* u1 = work + 2 * n;  // U first superdiagonal
    double* u2 = work + 3 * n;  // U second superdiagonal (from pivoting)
    int* piv = (int*)(work + 4 * n);
    double eps = 2.220446049250313e-16;
    double small = eps * tnorm;
    if (small == 0.0) small = eps;
    // LU of (T - lam I) with partial pivoting (dgttrf-style)
    for (int i = 0; i < n; i++) { d[i] = alpha[i] - lam; u1[i] = i < n - 1 ? beta[i] : 0.0; u2[i] = 0.0; }
    for (int i = 0; i < n - 1; i++) {
        double sub = beta[i];
        if (fabs(d[i]) >= fabs(sub)) {
            piv[i] = 0;
            if (d[i] == 0.0) d[i] = small;
            double f = sub / d[i];
            dl[i] = f;
            d[i + 1] -= f * u1[i];
        } else {
            piv[i] = 1;
            double f = d[i] / sub;
            dl[i] = f;
            d[i] = sub;
            double t = u1[i];
            u1[i] = d[i + 1];
            d[i + 1] = t - f * d[i + 1];
            if (i < n - 2) { u2[i] = u1[i + 1]; u1[i + 1] = -f * u1[i + 1]; }
        }
    }
    if (d[n - 1] == 0.0) d[n - 1] = small;
    for (int i = 0; i < n; i++) if (fabs(d[i]) < small) d[i] = d[i] < 0 ? -small : small;
    // start vector: deterministic, not orthogonal to any eigenvector in practice
    for (int i = 0; i < n; i++) x[i] = 1.0 + 0.5 * sin(1.3 * (double)i + 0.7);
    for (int iter = 0; iter < 4; iter++) {
        // solve L U y = P x
        for (int i = 0; i < n - 1; i++) {
            if (piv[i] == 0) x[i + 1] -= dl[i] * x[i];
            else { double t = x[i]; x[i] = x[i + 1]; x[i + 1] = t - dl[i] * x[i + 1]; }
        }
        x[n - 1] /= d[n - 1];
        if (n >= 2) x[n - 2] = (x[n - 2] - u1[n - 2] * x[n - 1]) / d[n - 2];
        for (int i = n - 3; i >= 0; i--) x[i] = (x[i] - u1[i] * x[i + 1] - u2[i] * x[i + 2]) / d[i];
        // orthogonalise against nearby eigenvectors (twice)
        for (int rep = 0; rep < 2; rep++)
            for (int p = 0; p < nprev; p++) {
                const double* y = prev + (long)p * n;
                double s = 0.0;
                for (int i = 0; i < n; i++) s += x[i] * y[i];
                for (int i = 0; i < n; i++) x[i] -= s * y[i];
            }
        double nr = 0.0;
        for (int i = 0; i < n; i++) nr += x[i] * x[i];
        nr = sqrt(nr);
        if (nr == 0.0) { x[iter % n] = 1.0; continue; }
        for (int i = 0; i < n; i++) x[i] /= nr;
    }
    // sign convention: largest component positive
    int im = 0;
    for (int i = 1; i < n; i++) if (fabs(x[i]) > fabs(x[im])) im = i;
    if (x[im] < 0) for (int i = 0; i < n; i++) x[i] = -x[i];
}

SMHD inline double tri_norm(const double* alpha, const double* beta, int n) {
    double m = 0.0;
    for (int i = 0; i < n; i++) {
        double r = fabs(alpha[i]) + (i > 0 ? fabs(beta[i - 1]) : 0.0) + (i < n - 1 ? fabs(beta[i]) : 0.0);
        m = fmax(m, r);
    }
    return m;
}

// Lowest k eigenpairs of T (n x n).  vals[k], vecs[k*n] (vector j at vecs + j*n).
// work: 5n doubles.
SMHD inline void tri_lowest(const double* alpha, const double* beta, int n, int k, double* vals, double* vecs,
                            double* work) {
    double lo, hi;
    tri_bounds(alpha, beta, n, &lo, &hi);
    double pm = tri_pivmin(beta, n);
    double tn = tri_norm(alpha, beta, n);
    for (int j = 0; j < k; j++) vals[j] = tri_bisect(alpha, beta, n, j, lo, hi, pm);
    double ortol = 1e-3 * tn;
    for (int j = 0; j < k; j++) {
        int first = j;
        while (first > 0 && fabs(vals[j] - vals[first - 1]) < ortol) first--;
        tri_eigvec(alpha, beta, n, vals[j], vecs + (long)j * n, work, vecs + (long)first * n, j - first, tn);
    }
}

// First index of the eigenvalue cluster of j (eigenvalues closer than ortol
// are orthogonalised against each other during inverse iteration).
SMHD inline int tri_cluster_first(const double* vals, int j, double ortol) {
    int first = j;
    while (first > 0 && fabs(vals[j] - vals[first - 1]) < ortol) first--;
    return first;
}

// Eigenvectors of the cluster that starts at j0 (inclusive) and ends before j1,
// in order; work: 5n doubles.
SMHD inline void tri_cluster_vecs(const double* alpha, const double* beta, int n, const double* vals, int j0,
                                  int j1, double* vecs, double* work, double tnorm) {
    for (int j = j0; j < j1; j++)
        tri_eigvec(alpha, beta, n, vals[j], vecs + (long)j * n, work, vecs + (long)j0 * n, j - j0, tnorm);
}

// Lanczos on diag(theta) from sigma/|sigma| with full reorthogonalisation,
// order reversed so the coupling to the residual vector sits on the last kept
// vector.  Outputs: Qrev (ls x ls, column j at Qrev + j*ls = Lanczos vector
// ls-1-j), new alpha[0..ls-1], beta[0..ls-1] (beta[ls-1] = |sigma|).
// work: 4*ls + ls*ls doubles.
SMHD inline void tri_diag_lanczos(const double* theta, const double* sig, int ls, double* Qrev, double* alpha,
                                  double* beta, double* work) {
    double* Q = work;                      // ls * ls, column p = Lanczos vector p
    double* na = Q + (long)ls * ls;
    double* nb = na + ls;
    double* r = nb + ls;
    double snorm = 0.0;
    for (int j = 0; j < ls; j++) snorm += sig[j] * sig[j];
    snorm = sqrt(snorm);
    double bprev = 0.0;
    for (int i = 0; i < ls; i++) Q[i] = snorm > 0 ? sig[i] / snorm : (i == 0 ? 1.0 : 0.0);
    for (int p = 0; p < ls; p++) {
        double* q = Q + (long)p * ls;
        double a = 0.0;
        for (int i = 0; i < ls; i++) a += theta[i] * q[i] * q[i];
        na[p] = a;
        if (p == ls - 1) break;
        for (int i = 0; i < ls; i++) r[i] = theta[i] * q[i] - a * q[i] - (p > 0 ? bprev * Q[(long)(p - 1) * ls + i] : 0.0);
        for (int rep = 0; rep < 2; rep++)
            for (int t = 0; t <= p; t++) {
                double s = 0.0;
                for (int i = 0; i < ls; i++) s += r[i] * Q[(long)t * ls + i];
                for (int i = 0; i < ls; i++) r[i] -= s * Q[(long)t * ls + i];
            }
        double b = 0.0;
        for (int i = 0; i < ls; i++) b += r[i] * r[i];
        b = sqrt(b);
        double* qn = Q + (long)(p + 1) * ls;
        if (b <= 1e-13 * (fabs(theta[0]) + fabs(theta[ls - 1]) + 1.0)) {
            // invariant subspace: continue with the unit vector least represented so far
            b = 0.0;
            int best = 0; double bestv = 2.0;
            for (int e = 0; e < ls; e++) {
                double s = 0.0;
                for (int t = 0; t <= p; t++) s += Q[(long)t * ls + e] * Q[(long)t * ls + e];
                if (s < bestv) { bestv = s; best = e; }
            }
            for (int i = 0; i < ls; i++) qn[i] = i == best ? 1.0 : 0.0;
            for (int rep = 0; rep < 2; rep++)
                for (int t = 0; t <= p; t++) {
                    double s = 0.0;
                    for (int i = 0; i < ls; i++) s += qn[i] * Q[(long)t * ls + i];
                    for (int i = 0; i < ls; i++) qn[i] -= s * Q[(long)t * ls + i];
                }
            double nn = 0.0;
            for (int i = 0; i < ls; i++) nn += qn[i] * qn[i];
            nn = sqrt(nn);
            for (int i = 0; i < ls; i++) qn[i] /= nn;
        } else {
            for (int i = 0; i < ls; i++) qn[i] = r[i] / b;
        }
        nb[p] = b;
        bprev = b;
    }
    for (int j = 0; j < ls; j++) {
        const int pj = ls - 1 - j;
        alpha[j] = na[pj];
        if (j < ls - 1) beta[j] = nb[pj - 1];
        for (int i = 0; i < ls; i++) Qrev[(long)j * ls + i] = Q[(long)pj * ls + i];
    }
    beta[ls - 1] = snorm;
}

// Thick restart of a tridiagonal Lanczos matrix (host composition of the
// pieces above; the device runs the same pieces in parallel kernels).
//   in : alpha[0..m-1], beta[0..m-1]  (m = lm-1 basis vectors; beta[m-1] couples
//        the last kept vector to the residual vector v_m)
//   out: S[m x ls] (column j at S + j*m) — coefficients of the new basis vectors
//        u_j = sum_i S[i,j] v_i, j < ls; new alpha[0..ls-1], beta[0..ls-1]
//        (beta[ls-1] couples u_{ls-1} to v_m, which becomes basis vector ls).
// work: 5m + ls*(m + 3*ls + 8) doubles.
SMHD inline void tri_restart(double* alpha, double* beta, int m, int ls, double* S, double* work) {
    double* theta = work;                  // ls
    double* Y = theta + ls;                // ls * m : Ritz vectors of T
    double* Qrev = Y + (long)ls * m;       // ls * ls
    double* sig = Qrev + (long)ls * ls;    // ls
    double* w5 = sig + ls;                 // 5m, then reused for tri_diag_lanczos (4ls + ls^2)
    const double bcoup = beta[m - 1];
    tri_lowest(alpha, beta, m, ls, theta, Y, w5);
    for (int j = 0; j < ls; j++) sig[j] = bcoup * Y[(long)j * m + (m - 1)];
    tri_diag_lanczos(theta, sig, ls, Qrev, alpha, beta, w5);
    for (int j = 0; j < ls; j++) {
        const double* qv = Qrev + (long)j * ls;
        for (int i = 0; i < m; i++) {
            double s = 0.0;
            for (int t = 0; t < ls; t++) s += Y[(long)t * m + i] * qv[t];
            S[(long)j * m + i] = s;
        }
    }
}

}  // namespace smgpu
