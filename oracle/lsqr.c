/*
 * Oracle (test infrastructure only): CSR products and damped LSQR in FP64.
 * Restates Numerics/SparseMatrix.cs (CsrMatrix.MultiplyInto :56-73, Transpose :79-97),
 * Numerics/SimdVector.cs (:20-98) and Numerics/Lsqr.cs (Solve, :34-152).
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* SparseMatrix.cs:79-97: counting-sort transpose; row r of the transpose lists, in increasing row
 * order of A, every entry of column r. trs has cols+1 entries. */
void or_csr_transpose(int rows, int cols, const int64_t *rs, const int *ci, const double *va,
                      int64_t *trs, int *tci, double *tva)
{
    memset(trs, 0, sizeof(int64_t) * (cols + 1));
    int64_t nnz = rs[rows];
    for (int64_t k = 0; k < nnz; k++) trs[ci[k] + 1]++;
    for (int c = 0; c < cols; c++) trs[c + 1] += trs[c];
    int64_t *next = malloc(sizeof(int64_t) * (cols + 1));
    memcpy(next, trs, sizeof(int64_t) * (cols + 1));
    for (int i = 0; i < rows; i++)
        for (int64_t k = rs[i]; k < rs[i + 1]; k++) {
            int64_t p = next[ci[k]]++;
            tci[p] = i;
            tva[p] = va[k];
        }
    free(next);
}

/* SparseMatrix.cs:56-73: y_i = sum_k va[k] x[ci[k]] in row order, rows in blocks of 256 in parallel. */
void or_spmv(int rows, const int64_t *rs, const int *ci, const double *va, const double *x, double *y,
             int threads)
{
    int blocks = (rows + 255) / 256;
#ifdef _OPENMP
    if (threads <= 0) threads = omp_get_max_threads();
#pragma omp parallel for schedule(static) num_threads(threads)
#endif
    for (int b = 0; b < blocks; b++) {
        int end = (b + 1) * 256 < rows ? (b + 1) * 256 : rows;
        for (int i = b * 256; i < end; i++) {
            double sum = 0;
            for (int64_t k = rs[i]; k < rs[i + 1]; k++) sum += va[k] * x[ci[k]];
            y[i] = sum;
        }
    }
    (void)threads;
}

/* SimdVector.cs:20-29: lane-wise accumulation over 4 doubles (Vector<double> on AVX2), then the
 * lanes in order, then the scalar tail. */
double or_dot(int n, const double *a, const double *b)
{
    double acc[4] = {0, 0, 0, 0};
    int v = n / 4;
    for (int i = 0; i < v; i++)
        for (int l = 0; l < 4; l++) acc[l] += a[4 * i + l] * b[4 * i + l];
    double sum = acc[0] * 1.0 + acc[1] * 1.0 + acc[2] * 1.0 + acc[3] * 1.0;
    for (int i = 4 * v; i < n; i++) sum += a[i] * b[i];
    return sum;
}

/* LsqrOpenCl.cs:79-94 dot_partial with 64 groups of 256 items: item l of group g sums a[i] b[i] over
 * i = 256 g + l + k 16384, the group halves its 256 sums in a tree, the host adds the 64 group sums in
 * order (LsqrOpenCl.cs Dot). */
double or_dot_device(int n, const double *a, const double *b)
{
    double part[64];
    for (int g = 0; g < 64; g++) {
        double red[256];
        for (int l = 0; l < 256; l++) {
            double s = 0;
            for (int64_t i = (int64_t)g * 256 + l; i < n; i += 64 * 256) s += a[i] * b[i];
            red[l] = s;
        }
        for (int h = 128; h > 0; h >>= 1)
            for (int l = 0; l < h; l++) red[l] += red[l + h];
        part[g] = red[0];
    }
    double s = 0;
    for (int g = 0; g < 64; g++) s += part[g];
    return s;
}

/* which dot product or_lsqr uses: 0 SimdVector's (Lsqr.Solve), 1 the reference device's (LsqrOpenCl.Run,
 * whose norm is the square root of that dot product) */
static _Thread_local int lsqr_dot_mode = 0;
void or_lsqr_dot_mode(int mode) { lsqr_dot_mode = mode; }
static double dot_sel(int n, const double *a, const double *b)
{
    return lsqr_dot_mode ? or_dot_device(n, a, b) : or_dot(n, a, b);
}
static double norm_sel(int n, const double *a)
{
    if (!lsqr_dot_mode) return or_norm(n, a);
    double d = or_dot_device(n, a, a);
    return sqrt(d > 0 ? d : 0);
}

static double max_abs(int n, const double *a)
{
    double m = 0;
    for (int i = 0; i < n; i++) {
        double x = fabs(a[i]);
        if (x > m) m = x;
    }
    return m;
}

/* SimdVector.cs:32-45: the norm, rescaled by max |a| outside [1e-150, 1e150]. */
double or_norm(int n, const double *a)
{
    double m = max_abs(n, a);
    if (m == 0 || !isfinite(m)) return m;
    if (m > 1e150 || m < 1e-150) {
        double s = 0, inv = 1.0 / m;
        for (int i = 0; i < n; i++) { double y = a[i] * inv; s += y * y; }
        return m * sqrt(s);
    }
    return sqrt(or_dot(n, a, a));
}

static void scale(int n, double *a, double s) { for (int i = 0; i < n; i++) a[i] *= s; }
static void axpy(int n, double alpha, const double *x, double *y) { for (int i = 0; i < n; i++) y[i] += alpha * x[i]; }
static void xpby(int n, const double *x, double beta, double *y) { for (int i = 0; i < n; i++) y[i] = x[i] + beta * y[i]; }
static void xmby(int n, const double *x, double beta, double *y) { for (int i = 0; i < n; i++) y[i] = x[i] - beta * y[i]; }

/* Lsqr.cs:34-152 Solve. stats = {iterations, bnorm, rnorm, acond}; the return value is the stop
 * reason: 0 budget exhausted, 1 btol, 2 atol, 3 conlim, 4-6 machine precision, 7 bidiagonalisation
 * terminated, 8 empty system, 9 zero right-hand side, 10 A^T b = 0. rnorm_hist (may be NULL) gets
 * the residual norm of every iteration (the onIteration callback, Lsqr.cs:140). */
int or_lsqr(int m, int n, const int64_t *rs, const int *ci, const double *va,
            const int64_t *trs, const int *tci, const double *tva, const double *b, double damp,
            int max_it, double atol, double btol, double conlim, int threads,
            double *x, double stats[4], double *rnorm_hist)
{
    memset(x, 0, sizeof(double) * n);
    stats[0] = stats[1] = stats[2] = stats[3] = 0;
    if (m == 0 || n == 0) return 8;
    double *u = malloc(sizeof(double) * m), *v = calloc(n, sizeof(double)), *w = malloc(sizeof(double) * n);
    double *tm = malloc(sizeof(double) * m), *tn = malloc(sizeof(double) * n);
    memcpy(u, b, sizeof(double) * m);
    int stop = 0, iterations = 0;
    double beta = norm_sel(m, u);
    if (beta == 0) { stop = 9; goto done; }
    double bnorm = beta;
    stats[1] = bnorm;
    scale(m, u, 1.0 / beta);
    or_spmv(n, trs, tci, tva, u, v, threads);
    double alpha = norm_sel(n, v);
    if (alpha == 0) { stats[2] = bnorm; stop = 10; goto done; }
    scale(n, v, 1.0 / alpha);
    memcpy(w, v, sizeof(double) * n);
    double rhobar = alpha, phibar = beta;
    double ddnorm = 0, xxnorm = 0, z = 0, cs2 = -1, sn2 = 0, anorm2 = 0, damped_res2 = 0;
    double rnorm = beta, acond = 0;
    double ctol = conlim > 0 ? 1.0 / conlim : 0.0;
    for (int it = 0; it < max_it; it++) {
        iterations = it + 1;
        or_spmv(m, rs, ci, va, v, tm, threads);
        xmby(m, tm, alpha, u);
        beta = norm_sel(m, u);
        if (beta > 0) {
            scale(m, u, 1.0 / beta);
            anorm2 += alpha * alpha + beta * beta + damp * damp;
            or_spmv(n, trs, tci, tva, u, tn, threads);
            xmby(n, tn, beta, v);
            alpha = norm_sel(n, v);
            if (alpha > 0) scale(n, v, 1.0 / alpha);
        }
        double rhobar1 = sqrt(rhobar * rhobar + damp * damp);
        double cs1 = rhobar / rhobar1, sn1 = damp / rhobar1;
        double psi = sn1 * phibar;
        phibar = cs1 * phibar;
        double rho = sqrt(rhobar1 * rhobar1 + beta * beta);
        double cs = rhobar1 / rho, sn = beta / rho;
        double theta = sn * alpha;
        rhobar = -cs * alpha;
        double phi = cs * phibar;
        phibar = sn * phibar;
        double tau = sn * phi;
        ddnorm += dot_sel(n, w, w) / (rho * rho);
        axpy(n, phi / rho, w, x);
        xpby(n, v, -theta / rho, w);
        double delta = sn2 * rho, gambar = -cs2 * rho;
        double rhs = phi - delta * z;
        double gamma = sqrt(gambar * gambar + theta * theta);
        cs2 = gambar / gamma;
        sn2 = theta / gamma;
        z = rhs / gamma;
        xxnorm += z * z;
        damped_res2 += psi * psi;
        double anorm = sqrt(anorm2);
        acond = anorm * sqrt(ddnorm);
        double arnorm = alpha * fabs(tau);
        rnorm = sqrt(phibar * phibar + damped_res2);
        double xnorm = sqrt(xxnorm);
        double test1 = rnorm / bnorm;
        double test2 = arnorm / fmax(1e-300, anorm * rnorm);
        double test3 = 1.0 / fmax(1e-300, acond);
        double t1 = test1 / (1 + anorm * xnorm / bnorm);
        double rtol = btol + atol * anorm * xnorm / bnorm;
        if (rnorm_hist) rnorm_hist[it] = rnorm;
        if (test1 <= rtol) { stop = 1; break; }
        if (test2 <= atol) { stop = 2; break; }
        if (ctol > 0 && test3 <= ctol) { stop = 3; break; }
        if (1 + t1 <= 1) { stop = 4; break; }
        if (1 + test2 <= 1) { stop = 5; break; }
        if (1 + test3 <= 1) { stop = 6; break; }
        if (beta == 0 || alpha == 0) { stop = 7; break; }
    }
    stats[2] = rnorm;
    stats[3] = acond;
done:
    stats[0] = iterations;
    free(u); free(v); free(w); free(tm); free(tn);
    return stop;
}
