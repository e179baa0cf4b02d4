// The exact formal resolution on the device: FormalResolution.Compute (Tomography/FormalResolution.cs:
// 108-352). With N = A^T A of the weighted system, D its data part and Gamma = N - D its regularisation
// part, the kept columns' reduced normal matrix S (the hypocentre groups eliminated by their Schur
// complements) is factored, S = L L^T, and with Y = L^-T:
//   C_ii = |Y_i|^2 (the posterior variance), R_ii = 1 - sum_j Gamma_ij (Y_i . Y_j),
// and, when the columns have places, the resolution length of every placed column from its full row of
// R = I - S^-1 Gamma (:291-331).
//
// Layout: S, then L and L^-1 in place, row-major n_pad x n_pad FP64 in HBM, n_pad the kept count rounded
// up to 64; the padding is the identity, so the factor and the inverse are the true ones in the leading
// block. Y = (L^-1)^T is a second n_pad x n_pad array and S^-1 = Y Y^T reuses the first.
//
// Kernels: the normal matrix and the Schur update in the reference's summation order (one workgroup per
// kept row, the rows of A that cross the column in order; a row's entries in parallel, as they touch
// distinct columns), so S equals the reference's to the last bit. The factorisation is blocked with
// 64-wide panels: the diagonal block in LDS, the panel's triangular solve one wavefront per row, the
// trailing update, the triangular inverse's products and S^-1 on FP64 MFMA (v_mfma_f64_16x16x4f64,
// 64 x 64 tiles, 4 wavefronts of 32 x 32). The reference's dot products are SIMD-width dependent
// (Vector<double>, :354-365), so from the factor on the results agree to rounding, not to the bit.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <unordered_map>
#include <vector>

#include "common.hpp"

using namespace ts;

namespace ts {
namespace {

typedef double v4d __attribute__((ext_vector_type(4)));
constexpr int kT = 64;   // tile and panel edge
constexpr int kKC = 32;  // K chunk staged in LDS per step

// The K range of output tile (ti, tj) when operands are triangular: from max(kb_i ti, kb_j tj) tiles
// (A's row tile zero to the left of its diagonal: kb_i; op(B)'s column tile zero above its diagonal:
// kb_j) to (ti + 1) tiles when ke_i (A lower triangular), else to K.
struct KRange {
    int kb_i, kb_j, ke_i;
};

// C[M x N] = alpha A[M x K] op(B) + beta C, row-major with leading dimensions; op(B) is B[K x N] or,
// with kTransB, the transpose of B[N x K]. M, N, K are multiples of 64; one workgroup per 64 x 64 tile
// of C (grid x: column tiles, y: row tiles); kLower skips the tiles above the diagonal. beta = 0 does
// not read C.
template <bool kTransB, bool kLower>
__global__ __launch_bounds__(256) void fr_gemm(int K, KRange kr, double alpha, const double *__restrict__ A,
                                               int64_t lda, const double *__restrict__ B, int64_t ldb, double beta,
                                               double *__restrict__ C, int64_t ldc)
{
    const int ti = blockIdx.y, tj = blockIdx.x;
    if (kLower && tj > ti) return;
    const int kbeg = max(kr.kb_i * ti, kr.kb_j * tj) * kT;
    const int kend = kr.ke_i ? min(K, (ti + 1) * kT) : K;
    __shared__ double As[kT][kKC + 1];
    __shared__ double Bs[kKC][kT + 1];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wr = (wave >> 1) * 32, wc = (wave & 1) * 32;
    const int r16 = lane & 15, kq = lane >> 4;
    v4d acc[2][2];
    for (int a = 0; a < 2; a++)
        for (int b = 0; b < 2; b++) acc[a][b] = v4d{0, 0, 0, 0};
    const double *Ab = A + (int64_t)ti * kT * lda;
    for (int k0 = kbeg; k0 < kend; k0 += kKC) {
        for (int e = tid; e < kT * kKC; e += 256) {
            const int r = e / kKC, c = e % kKC;
            As[r][c] = Ab[(int64_t)r * lda + k0 + c];
        }
        if (kTransB) {
            for (int e = tid; e < kT * kKC; e += 256) {
                const int j = e / kKC, k = e % kKC;
                Bs[k][j] = B[(int64_t)(tj * kT + j) * ldb + k0 + k];
            }
        } else {
            for (int e = tid; e < kKC * kT; e += 256) {
                const int k = e / kT, j = e % kT;
                Bs[k][j] = B[(int64_t)(k0 + k) * ldb + tj * kT + j];
            }
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < kKC; kk += 4) {
            // A[i][k] in lane i + 16 k, B[k][j] in lane j + 16 k
            const double a0 = As[wr + r16][kk + kq], a1 = As[wr + 16 + r16][kk + kq];
            const double b0 = Bs[kk + kq][wc + r16], b1 = Bs[kk + kq][wc + 16 + r16];
            acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, acc[0][0], 0, 0, 0);
            acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b1, acc[0][1], 0, 0, 0);
            acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b0, acc[1][0], 0, 0, 0);
            acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, acc[1][1], 0, 0, 0);
        }
        __syncthreads();
    }
    // D[i][j] in register i / 4 of lane j + 16 (i % 4)
    for (int a = 0; a < 2; a++)
        for (int b = 0; b < 2; b++)
            for (int r = 0; r < 4; r++) {
                const int64_t row = (int64_t)ti * kT + wr + a * 16 + kq + 4 * r;
                const int64_t col = (int64_t)tj * kT + wc + b * 16 + r16;
                double *c = C + row * ldc + col;
                *c = (beta == 0 ? 0.0 : beta * *c) + alpha * acc[a][b][r];
            }
}

// S = N_kk, lower triangle (:150-166): workgroup i adds, for each row r of A crossing kept column
// kept[i] (in row order), v_r times row r's entries at kept columns j <= i.
__global__ __launch_bounds__(256) void fr_normal(int n, const int32_t *__restrict__ kept, const int32_t *__restrict__ index,
                                                 const int32_t *__restrict__ rs, const int32_t *__restrict__ ci,
                                                 const double *__restrict__ va, const int32_t *__restrict__ trs,
                                                 const int32_t *__restrict__ tci, const double *__restrict__ tva,
                                                 double *__restrict__ S, int64_t ld)
{
    const int i = blockIdx.x;
    if (i >= n) return;
    const int c = kept[i];
    double *row = S + (int64_t)i * ld;
    for (int p = trs[c]; p < trs[c + 1]; p++) {
        const int r = tci[p];
        const double v = tva[p];
        for (int q = rs[r] + threadIdx.x; q < rs[r + 1]; q += blockDim.x) {
            const int j = index[ci[q]];
            if (j >= 0 && j <= i) row[j] += v * va[q];
        }
        __syncthreads();
    }
}

// S -= U U^T for the eliminated groups (:231-246): workgroup i takes its entries of U (group g, column x,
// value v) in order and subtracts v u[z][x] at the group's kept rows j = rows[z] <= i.
__global__ __launch_bounds__(256) void fr_schur(int n, const int32_t *__restrict__ ucount, const int32_t *__restrict__ ublock,
                                                const int32_t *__restrict__ upos, const int32_t *__restrict__ brows_start,
                                                const int32_t *__restrict__ brows, const int64_t *__restrict__ bu_start,
                                                const double *__restrict__ bu, const int32_t *__restrict__ bm,
                                                double *__restrict__ S, int64_t ld)
{
    const int i = blockIdx.x;
    if (i >= n) return;
    double *row = S + (int64_t)i * ld;
    for (int p = ucount[i]; p < ucount[i + 1]; p++) {
        const int g = ublock[p], m = bm[g];
        const double *u = bu + bu_start[g];
        const int x = upos[p] % m;
        const double v = u[upos[p]];
        const int z0 = brows_start[g], nz = brows_start[g + 1] - z0;
        for (int z = threadIdx.x; z < nz; z += blockDim.x) {
            const int j = brows[z0 + z];
            if (j <= i) row[j] -= v * u[(int64_t)z * m + x];
        }
        __syncthreads();
    }
}

__global__ void fr_pad_identity(double *S, int64_t ld, int n, int npad)
{
    const int i = n + blockIdx.x * blockDim.x + threadIdx.x;
    if (i < npad) S[(int64_t)i * ld + i] = 1.0;
}

// Cholesky of the 64 x 64 diagonal block at (k0, k0) in LDS, lower triangle; zeros above the diagonal;
// bad = 1 when a pivot is not positive. Step j updates the trailing block with column j and its pivot
// d (L_il -= L_ij (L_lj / d)) and scales column j - 1, which step j no longer reads: one barrier per
// column.
__global__ __launch_bounds__(256) void fr_potrf_diag(double *S, int64_t ld, int k0, int32_t *bad)
{
    __shared__ double L[kT][kT + 1];
    const int tid = threadIdx.x;
    double *base = S + (int64_t)k0 * ld + k0;
    for (int e = tid; e < kT * kT; e += 256) {
        const int r = e / kT, c = e % kT;
        L[r][c] = c <= r ? base[(int64_t)r * ld + c] : 0.0;
    }
    __syncthreads();
    double dprev = 0;
    for (int j = 0; j < kT; j++) {
        const double d = L[j][j];
        if (tid == 0 && !(d > 0)) *bad = 1;
        if (j > 0) {
            const double sq = sqrt(dprev);
            for (int i = j - 1 + tid; i < kT; i += 256) L[i][j - 1] = i == j - 1 ? sq : L[i][j - 1] / sq;
        }
        const double rd = 1.0 / d;
        for (int e = tid; e < kT * kT; e += 256) {
            const int i = e / kT, l = e % kT;
            if (l > j && l <= i) L[i][l] -= L[i][j] * (L[l][j] * rd);
        }
        dprev = d;
        __syncthreads();
    }
    if (tid == 0) L[kT - 1][kT - 1] = sqrt(dprev);
    __syncthreads();
    for (int e = tid; e < kT * kT; e += 256) {
        const int r = e / kT, c = e % kT;
        base[(int64_t)r * ld + c] = L[r][c];
    }
}

// The panel below the diagonal block: rows i in [row0, row0 + rows) solve x L^T = a (:318-327), one
// wavefront per row, L's transpose in LDS.
__global__ __launch_bounds__(256) void fr_trsm_panel(double *S, int64_t ld, int k0, int row0, int rows)
{
    __shared__ double Lt[kT][kT + 1];
    const int tid = threadIdx.x, lane = tid & 63;
    const double *base = S + (int64_t)k0 * ld + k0;
    for (int e = tid; e < kT * kT; e += 256) {
        const int r = e / kT, c = e % kT;
        Lt[c][r] = base[(int64_t)r * ld + c];
    }
    __syncthreads();
    const int w = blockIdx.x * 4 + (tid >> 6);
    if (w >= rows) return;
    double *row = S + (int64_t)(row0 + w) * ld + k0;
    const double a = row[lane];
    double acc = 0, x = 0;
    for (int j = 0; j < kT; j++) {
        const double cand = (a - acc) / Lt[j][j];
        const double xj = __shfl(cand, j);
        if (lane == j) x = xj;
        if (lane > j) acc += xj * Lt[j][lane];
    }
    row[lane] = x;
}

// The inverse of every 64 x 64 diagonal block of the factor, in place (column c by forward substitution,
// the reference's recurrence, :263-272); zeros above the diagonal.
__global__ __launch_bounds__(64) void fr_trtri_leaf(double *S, int64_t ld)
{
    __shared__ double L[kT][kT];
    __shared__ double X[kT][kT];
    const int c = threadIdx.x;
    double *base = S + (int64_t)blockIdx.x * kT * ld + (int64_t)blockIdx.x * kT;
    for (int r = 0; r < kT; r++) L[r][c] = c <= r ? base[(int64_t)r * ld + c] : 0.0;
    __syncthreads();
    for (int k = 0; k < c; k++) X[k][c] = 0.0;
    X[c][c] = 1.0 / L[c][c];
    for (int k = c + 1; k < kT; k++) {
        double s = 0;
        for (int j = c; j < k; j++) s += L[k][j] * X[j][c];
        X[k][c] = -s / L[k][k];
    }
    __syncthreads();
    for (int r = 0; r < kT; r++) base[(int64_t)r * ld + c] = X[r][c];
}

// Y = X^T in 64 x 64 tiles through LDS.
__global__ __launch_bounds__(256) void fr_transpose(const double *__restrict__ X, double *__restrict__ Y, int64_t ld)
{
    __shared__ double t[kT][kT + 1];
    const int bi = blockIdx.y, bj = blockIdx.x, tid = threadIdx.x;
    for (int e = tid; e < kT * kT; e += 256) {
        const int r = e / kT, c = e % kT;
        t[r][c] = X[((int64_t)bi * kT + r) * ld + (int64_t)bj * kT + c];
    }
    __syncthreads();
    for (int e = tid; e < kT * kT; e += 256) {
        const int r = e / kT, c = e % kT;
        Y[((int64_t)bj * kT + r) * ld + (int64_t)bi * kT + c] = t[c][r];
    }
}

// The tiles above the diagonal from the transposed tiles below it (a symmetric matrix of which the
// lower tiles are computed).
__global__ __launch_bounds__(256) void fr_mirror(double *S, int64_t ld)
{
    __shared__ double t[kT][kT + 1];
    const int bi = blockIdx.y, bj = blockIdx.x, tid = threadIdx.x;
    if (bj <= bi) return;
    for (int e = tid; e < kT * kT; e += 256) {
        const int r = e / kT, c = e % kT;
        t[r][c] = S[((int64_t)bj * kT + r) * ld + (int64_t)bi * kT + c];
    }
    __syncthreads();
    for (int e = tid; e < kT * kT; e += 256) {
        const int r = e / kT, c = e % kT;
        S[((int64_t)bi * kT + r) * ld + (int64_t)bj * kT + c] = t[c][r];
    }
}

__device__ inline double block_sum(double v, double *red)
{
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) red[wave] = v;
    __syncthreads();
    double s = 0;
    for (int w = 0; w < (int)(blockDim.x >> 6); w++) s += red[w];
    return s;
}

// sd_i = |Y_i| and R_ii = 1 - sum over Gamma_i of g (Y_i . Y_j), the dots over k >= max(i, j) (:280-289)
__global__ __launch_bounds__(256) void fr_diag_sd(int n, const double *__restrict__ Y, int64_t ld,
                                                  const int32_t *__restrict__ grs, const int32_t *__restrict__ gci,
                                                  const double *__restrict__ gva, double *__restrict__ diag,
                                                  double *__restrict__ sd)
{
    __shared__ double red[4];
    const int i = blockIdx.x;
    if (i >= n) return;
    const double *yi = Y + (int64_t)i * ld;
    double s = 0;
    for (int64_t k = i + threadIdx.x; k < ld; k += blockDim.x) s += yi[k] * yi[k];
    s = block_sum(s, red);
    double sum = 0;
    for (int e = grs[i]; e < grs[i + 1]; e++) {
        const int j = gci[e];
        const int m = max(i, j);
        const double *yj = Y + (int64_t)j * ld;
        double d = 0;
        for (int64_t k = m + threadIdx.x; k < ld; k += blockDim.x) d += yi[k] * yj[k];
        d = block_sum(d, red);
        sum += gva[e] * d;
    }
    if (threadIdx.x == 0) {
        sd[i] = sqrt(s);
        diag[i] = 1 - sum;
    }
}

// The resolution length of placed column i (:291-331): row_j of R = [i == j] - sum over Gamma_j of
// g S^-1_ik, and sqrt(sum R_ij^2 d_ij^2 / sum R_ij^2) over the placed columns j of i's block.
__global__ __launch_bounds__(256) void fr_lengths(int n, const double *__restrict__ Sinv, int64_t ld,
                                                  const int32_t *__restrict__ grs, const int32_t *__restrict__ gci,
                                                  const double *__restrict__ gva, const double *__restrict__ pos,
                                                  const int32_t *__restrict__ block, double *__restrict__ length)
{
    __shared__ double red[4];
    const int i = blockIdx.x;
    if (i >= n) return;
    if (block[i] < 0) {
        if (threadIdx.x == 0) length[i] = NAN;
        return;
    }
    const double *si = Sinv + (int64_t)i * ld;
    const double xi = pos[3 * i], yi = pos[3 * i + 1], zi = pos[3 * i + 2];
    double num = 0, den = 0;
    for (int j = threadIdx.x; j < n; j += blockDim.x) {
        if (block[j] != block[i]) continue;
        double r = i == j ? 1.0 : 0.0;
        for (int e = grs[j]; e < grs[j + 1]; e++) r -= si[gci[e]] * gva[e];
        const double r2 = r * r;
        const double dx = pos[3 * j] - xi, dy = pos[3 * j + 1] - yi, dz = pos[3 * j + 2] - zi;
        num += r2 * (dx * dx + dy * dy + dz * dz);
        den += r2;
    }
    num = block_sum(num, red);
    den = block_sum(den, red);
    if (threadIdx.x == 0) length[i] = den > 0 ? sqrt(num / den) : NAN;
}

unsigned nblocks(int64_t n, int per) { return (unsigned)std::max<int64_t>(1, (n + per - 1) / per); }

template <bool kTransB, bool kLower>
void gemm(hipStream_t st, int M, int N, int K, double alpha, const double *A, int64_t lda, const double *B,
          int64_t ldb, double beta, double *C, int64_t ldc, KRange kr = {0, 0, 0})
{
    if (M <= 0 || N <= 0) return;
    fr_gemm<kTransB, kLower><<<dim3(N / kT, M / kT), 256, 0, st>>>(K, kr, alpha, A, lda, B, ldb, beta, C, ldc);
    TS_HIP(hipGetLastError());
}

// In-place inverse of the lower-triangular factor at rows/columns [off, off + size) whose 64 x 64 diagonal
// blocks are already inverted: X11, X22 of the halves, then X21 = -X22 (L21 X11) through T.
void trtri(hipStream_t st, double *S, int64_t ld, int off, int size, double *T)
{
    if (size <= kT) return;
    const int h = (size / kT / 2) * kT;
    trtri(st, S, ld, off, h, T);
    trtri(st, S, ld, off + h, size - h, T);
    double *L21 = S + (int64_t)(off + h) * ld + off;
    // X11 is lower triangular: column tile tj of L21 X11 sums from tile tj; X22 too: row tile ti of
    // X22 T sums up to tile ti
    gemm<false, false>(st, size - h, h, h, 1.0, L21, ld, S + (int64_t)off * ld + off, ld, 0.0, T, h, {0, 1, 0});
    gemm<false, false>(st, size - h, h, size - h, -1.0, S + (int64_t)(off + h) * ld + off + h, ld, T, h, 0.0, L21,
                       ld, {0, 0, 1});
}

// Counting-sort transpose with rows in increasing order within each column (SparseMatrix.cs:79-97).
void transpose(int rows, int cols, const int32_t *rs, const int32_t *ci, const double *va, std::vector<int32_t> &trs,
               std::vector<int32_t> &tci, std::vector<double> &tva)
{
    const int64_t nnz = rs[rows];
    trs.assign(cols + 1, 0);
    for (int64_t q = 0; q < nnz; q++) trs[ci[q] + 1]++;
    for (int c = 0; c < cols; c++) trs[c + 1] += trs[c];
    std::vector<int32_t> next(trs.begin(), trs.end() - 1);
    tci.resize(nnz);
    tva.resize(nnz);
    for (int r = 0; r < rows; r++)
        for (int q = rs[r]; q < rs[r + 1]; q++) {
            const int p = next[ci[q]]++;
            tci[p] = r;
            tva[p] = va[q];
        }
}

// SmallCholesky (:336-352)
bool small_cholesky(std::vector<double> &m, int n)
{
    for (int j = 0; j < n; j++) {
        double d = m[j * n + j];
        for (int k = 0; k < j; k++) d -= m[j * n + k] * m[j * n + k];
        if (d <= 0) return false;
        m[j * n + j] = std::sqrt(d);
        for (int i = j + 1; i < n; i++) {
            double s = m[i * n + j];
            for (int k = 0; k < j; k++) s -= m[i * n + k] * m[j * n + k];
            m[i * n + j] = s / m[j * n + j];
        }
    }
    return true;
}

// host arrays uploaded here stay alive until the stream is synchronised
template <typename T>
void put(dbuf<T> &d, const T *h, size_t n, hipStream_t st)
{
    d.alloc(std::max<size_t>(1, n));
    if (n) TS_HIP(hipMemcpyAsync(d.p, h, n * sizeof(T), hipMemcpyHostToDevice, st));
}

template <typename T>
void put(dbuf<T> &d, const std::vector<T> &h, hipStream_t st)
{
    put(d, h.data(), h.size(), st);
}

}  // namespace
}  // namespace ts

extern "C" {

int ts_formal_resolution(ts_context *ctx, int rows, int cols, const int32_t *row_start, const int32_t *column,
                         const double *values, int data_rows, int ngroups, const int32_t *group_start,
                         const int32_t *group_columns, int max_unknowns, const double *places,
                         const int32_t *place_block, int32_t *kept_out, int32_t *nkept, double *diagonal,
                         double *stddev, double *length_km, double *seconds, char *reason, int reason_len)
{
    if (!ctx || rows < 0 || cols < 0 || !row_start || (row_start[rows] > 0 && (!column || !values)) || ngroups < 0 ||
        (ngroups > 0 && (!group_start || !group_columns)) || !kept_out || !nkept || !diagonal || !stddev ||
        (places && !place_block) || (places && !length_km)) {
        set_reason(reason, reason_len, "invalid argument");
        return TS_EINVAL;
    }
    return guarded(reason, reason_len, [&]() -> int {
        TS_HIP(hipSetDevice(ctx->device));
        hipStream_t st = exec_stream(ctx);
        auto t0 = std::chrono::steady_clock::now();
        const int64_t nnz = row_start[rows];
        for (int r = 0; r < rows; r++)
            if (row_start[r + 1] < row_start[r]) fail(TS_EINVAL, "row %d: decreasing row start", r);
        std::vector<int32_t> seen(cols, -1);
        for (int r = 0; r < rows; r++)
            for (int q = row_start[r]; q < row_start[r + 1]; q++) {
                const int c = column[q];
                if (c < 0 || c >= cols) fail(TS_EINVAL, "row %d: column %d of %d", r, c, cols);
                if (seen[c] == r) fail(TS_EINVAL, "row %d repeats column %d", r, c);
                seen[c] = r;
            }
        // groups to eliminate and the kept columns (:111-126)
        std::vector<int32_t> group(cols, -1);
        for (int g = 0; g < ngroups; g++)
            for (int q = group_start[g]; q < group_start[g + 1]; q++) {
                const int c = group_columns[q];
                if (c < 0 || c >= cols) fail(TS_EINVAL, "group %d: column %d of %d", g, c, cols);
                group[c] = g;
            }
        std::vector<int32_t> kept, index(cols, -1);
        for (int c = 0; c < cols; c++)
            if (group[c] < 0) { index[c] = (int)kept.size(); kept.push_back(c); }
        const int n = (int)kept.size();
        if (n > max_unknowns)
            fail(TS_DECLINED, "Formal resolution: %d unknowns, above the limit of %d (%.1f GB of dense matrices); "
                              "use the adaptive grid, or the checkerboard and spike tests.",
                 n, max_unknowns, 2.0 * n * (double)n * 8 / 1e9);
        *nkept = n;
        std::copy(kept.begin(), kept.end(), kept_out);
        if (n == 0) {
            if (seconds) *seconds = 0;
            return TS_OK;
        }

        // Gamma on the kept columns, and the checks that the separation is exact (:128-158)
        std::vector<std::vector<std::pair<int, double>>> gamma(n);
        std::vector<std::unordered_map<int, int>> slot(n);
        for (int r = 0; r < rows; r++) {
            const int s0 = row_start[r], s1 = row_start[r + 1];
            int row_group = -1;
            bool touches_kept = false, touches_elim = false;
            for (int p = s0; p < s1; p++) {
                const int g = group[column[p]];
                if (g < 0) { touches_kept = true; continue; }
                touches_elim = true;
                if (row_group >= 0 && g != row_group)
                    fail(TS_EINVAL, "Formal resolution: a row couples the hypocentres of two events (double "
                                    "differences); not supported.");
                row_group = g;
            }
            if (r < data_rows) continue;
            if (touches_kept && touches_elim)
                fail(TS_EINVAL, "Formal resolution: a regularisation row couples model and hypocentre columns.");
            if (!touches_kept) continue;
            for (int p = s0; p < s1; p++) {
                const int i = index[column[p]];
                for (int q = s0; q < s1; q++) {
                    const int j = index[column[q]];
                    auto it = slot[i].find(j);
                    if (it == slot[i].end()) {
                        slot[i].emplace(j, (int)gamma[i].size());
                        gamma[i].push_back({j, values[p] * values[q]});
                    } else {
                        gamma[i][it->second].second += values[p] * values[q];
                    }
                }
            }
        }
        slot.clear();
        std::vector<int32_t> grs(n + 1, 0), gci;
        std::vector<double> gva;
        for (int i = 0; i < n; i++) {
            for (auto &e : gamma[i]) { gci.push_back(e.first); gva.push_back(e.second); }
            grs[i + 1] = (int)gci.size();
        }
        gamma.clear();

        std::vector<int32_t> trs, tci;
        std::vector<double> tva;
        transpose(rows, cols, row_start, column, values, trs, tci, tva);

        // the Schur blocks of the groups, on the host in the reference's order (:171-214)
        std::vector<int32_t> brows_start(ngroups + 1, 0), brows, bm(std::max(1, ngroups), 0);
        std::vector<int64_t> bu_start(ngroups + 1, 0);
        std::vector<double> bu;
        for (int g = 0; g < ngroups; g++) {
            const int m = group_start[g + 1] - group_start[g];
            const int32_t *gc = group_columns + group_start[g];
            std::vector<double> ngg((size_t)m * m, 0.0);
            std::unordered_map<int, std::vector<double>> nkg;
            for (int l = 0; l < m; l++)
                for (int p = trs[gc[l]]; p < trs[gc[l] + 1]; p++) {
                    const int r = tci[p];
                    const double v = tva[p];
                    for (int q = row_start[r]; q < row_start[r + 1]; q++) {
                        const int c = column[q];
                        if (group[c] == g) {
                            const int at = (int)(std::find(gc, gc + m, c) - gc);
                            ngg[(size_t)l * m + at] += v * values[q];
                            continue;
                        }
                        const int k = index[c];
                        if (k < 0) continue;
                        auto &e = nkg[k];
                        if (e.empty()) e.assign(m, 0.0);
                        e[l] += v * values[q];
                    }
                }
            std::vector<int> live;
            for (int l = 0; l < m; l++)
                if (ngg[(size_t)l * m + l] > 0) live.push_back(l);
            const int ml = (int)live.size();
            std::vector<double> lg((size_t)ml * ml);
            for (int x = 0; x < ml; x++)
                for (int y = 0; y < ml; y++) lg[(size_t)x * ml + y] = ngg[(size_t)live[x] * m + live[y]];
            if (ml > 0 && small_cholesky(lg, ml)) {
                std::vector<int> rws;
                for (auto &kv : nkg) rws.push_back(kv.first);
                std::sort(rws.begin(), rws.end());
                std::vector<double> u(rws.size() * (size_t)ml);
                for (size_t z = 0; z < rws.size(); z++) {
                    const auto &e = nkg[rws[z]];
                    for (int x = 0; x < ml; x++) {
                        double sum = e[live[x]];
                        for (int y = 0; y < x; y++) sum -= lg[(size_t)x * ml + y] * u[z * ml + y];
                        u[z * ml + x] = sum / lg[(size_t)x * ml + x];
                    }
                }
                brows.insert(brows.end(), rws.begin(), rws.end());
                bu.insert(bu.end(), u.begin(), u.end());
                bm[g] = ml;
            }
            brows_start[g + 1] = (int)brows.size();
            bu_start[g + 1] = (int64_t)bu.size();
        }
        // U by kept row (:215-230)
        std::vector<int32_t> ucount(n + 1, 0);
        for (int g = 0; g < ngroups; g++)
            for (int z = brows_start[g]; z < brows_start[g + 1]; z++) ucount[brows[z] + 1] += bm[g];
        for (int k = 0; k < n; k++) ucount[k + 1] += ucount[k];
        std::vector<int32_t> next(ucount.begin(), ucount.end() - 1), ublock(ucount[n]), upos(ucount[n]);
        for (int g = 0; g < ngroups; g++) {
            const int m = bm[g];
            for (int z = 0; z < brows_start[g + 1] - brows_start[g]; z++)
                for (int x = 0; x < m; x++) {
                    const int p = next[brows[brows_start[g] + z]]++;
                    ublock[p] = g;
                    upos[p] = z * m + x;
                }
        }

        // device: S, the factor, the inverse, Y, the diagonals and the lengths
        const int npad = (n + kT - 1) / kT * kT;
        const int64_t ld = npad;
        const int h_top = npad > kT ? (npad / kT / 2) * kT : kT;
        dbuf<double> S((size_t)npad * npad), Y((size_t)npad * npad), T((size_t)(npad - h_top + kT) * h_top);
        TS_HIP(hipMemsetAsync(S.p, 0, S.bytes(), st));
        dbuf<int32_t> d_kept, d_index, d_rs, d_ci, d_trs, d_tci, d_grs, d_gci, d_bad(1);
        dbuf<double> d_va, d_tva, d_gva;
        put(d_kept, kept, st);
        put(d_index, index, st);
        put(d_rs, row_start, (size_t)rows + 1, st);
        put(d_ci, column, (size_t)nnz, st);
        put(d_va, values, (size_t)nnz, st);
        put(d_trs, trs, st);
        put(d_tci, tci, st);
        put(d_tva, tva, st);
        put(d_grs, grs, st);
        put(d_gci, gci, st);
        put(d_gva, gva, st);
        TS_HIP(hipMemsetAsync(d_bad.p, 0, 4, st));
        fr_normal<<<n, 256, 0, st>>>(n, d_kept.p, d_index.p, d_rs.p, d_ci.p, d_va.p, d_trs.p, d_tci.p, d_tva.p, S.p, ld);
        TS_HIP(hipGetLastError());
        if (ucount[n] > 0) {
            dbuf<int32_t> d_uc, d_ub, d_up, d_brs, d_br, d_bm;
            dbuf<int64_t> d_bus;
            dbuf<double> d_bu;
            put(d_uc, ucount, st);
            put(d_ub, ublock, st);
            put(d_up, upos, st);
            put(d_brs, brows_start, st);
            put(d_br, brows, st);
            put(d_bm, bm, st);
            put(d_bus, bu_start, st);
            put(d_bu, bu, st);
            fr_schur<<<n, 256, 0, st>>>(n, d_uc.p, d_ub.p, d_up.p, d_brs.p, d_br.p, d_bus.p, d_bu.p, d_bm.p, S.p, ld);
            TS_HIP(hipGetLastError());
            TS_HIP(hipStreamSynchronize(st));
        }
        if (npad > n) fr_pad_identity<<<nblocks(npad - n, 64), 64, 0, st>>>(S.p, ld, n, npad);

        // blocked right-looking Cholesky
        for (int k0 = 0; k0 < npad; k0 += kT) {
            const int k1 = k0 + kT;
            fr_potrf_diag<<<1, 256, 0, st>>>(S.p, ld, k0, d_bad.p);
            if (k1 < npad) {
                fr_trsm_panel<<<nblocks(npad - k1, 4), 256, 0, st>>>(S.p, ld, k0, k1, npad - k1);
                gemm<true, true>(st, npad - k1, npad - k1, kT, -1.0, S.p + (int64_t)k1 * ld + k0, ld,
                                 S.p + (int64_t)k1 * ld + k0, ld, 1.0, S.p + (int64_t)k1 * ld + k1, ld);
            }
        }
        TS_HIP(hipGetLastError());
        int32_t bad = 0;
        TS_HIP(hipMemcpyAsync(&bad, d_bad.p, 4, hipMemcpyDeviceToHost, st));
        TS_HIP(hipStreamSynchronize(st));
        if (bad)
            fail(TS_EINVAL, "Formal resolution: the reduced normal matrix is not positive definite (an unknown "
                            "without data and without damping).");

        // X = L^-1 in place, Y = X^T
        fr_trtri_leaf<<<npad / kT, 64, 0, st>>>(S.p, ld);
        trtri(st, S.p, ld, 0, npad, T.p);
        fr_transpose<<<dim3(npad / kT, npad / kT), 256, 0, st>>>(S.p, Y.p, ld);
        TS_HIP(hipGetLastError());
        T.reset();

        dbuf<double> d_diag(n), d_sd(n), d_len(n);
        fr_diag_sd<<<n, 256, 0, st>>>(n, Y.p, ld, d_grs.p, d_gci.p, d_gva.p, d_diag.p, d_sd.p);
        TS_HIP(hipGetLastError());
        if (places) {
            // S^-1 = Y Y^T into the first array: Y is upper triangular, so tile (ti, tj <= ti) sums from
            // tile ti; the upper tiles are the mirror of the lower ones
            gemm<true, true>(st, npad, npad, npad, 1.0, Y.p, ld, Y.p, ld, 0.0, S.p, ld, {1, 1, 0});
            fr_mirror<<<dim3(npad / kT, npad / kT), 256, 0, st>>>(S.p, ld);
            TS_HIP(hipGetLastError());
            std::vector<double> pos(3 * (size_t)n);
            std::vector<int32_t> blk(n);
            for (int i = 0; i < n; i++) {
                for (int q = 0; q < 3; q++) pos[3 * (size_t)i + q] = places[3 * (size_t)kept[i] + q];
                blk[i] = place_block[kept[i]];
            }
            dbuf<double> d_pos;
            dbuf<int32_t> d_blk;
            put(d_pos, pos, st);
            put(d_blk, blk, st);
            fr_lengths<<<n, 256, 0, st>>>(n, S.p, ld, d_grs.p, d_gci.p, d_gva.p, d_pos.p, d_blk.p, d_len.p);
            TS_HIP(hipGetLastError());
            TS_HIP(hipMemcpyAsync(length_km, d_len.p, n * 8, hipMemcpyDeviceToHost, st));
            TS_HIP(hipStreamSynchronize(st));
        } else if (length_km) {
            std::fill(length_km, length_km + n, NAN);
        }
        TS_HIP(hipMemcpyAsync(diagonal, d_diag.p, n * 8, hipMemcpyDeviceToHost, st));
        TS_HIP(hipMemcpyAsync(stddev, d_sd.p, n * 8, hipMemcpyDeviceToHost, st));
        TS_HIP(hipStreamSynchronize(st));
        if (seconds) *seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        return TS_OK;
    });
}

}  // extern "C"
