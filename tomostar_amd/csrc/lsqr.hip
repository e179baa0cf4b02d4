// LSQR on gfx950: damped LSQR (Paige & Saunders; the SOL form of Numerics/Lsqr.cs:34-152) with A and
// its stored transpose resident in HBM as CSR (Numerics/SparseMatrix.cs:24-97).
//
// Device forms:
//   plain (TS_LSQR_PLAIN_FP64): the reference's device loop (Numerics/LsqrOpenCl.cs:272-533) on this GPU -
//     FP64 values and vectors, one right-hand side at a time, a row-per-thread CSR product, one launch
//     per vector operation, dot products as 64 fixed partial sums added on the host in order, and the
//     plane rotations on the host (three blocking reads per iteration).
//   fast (TS_LSQR_FAST_FP32, TS_LSQR_FAST_MIXED): k right-hand sides with their own damping share every
//     pass over A and A^T (SpMM over [row][k]-interleaved vectors: one gather of k contiguous values
//     per non-zero). Matrix values are FP32; the u, v, w vectors are FP32 (FAST_FP32) or FP64
//     (FAST_MIXED); row sums, the solution x and every scalar (the norms alpha and beta, the
//     renormalisation factors, the rotations and the stopping estimates) are FP64 on the device. u
//     and v are stored unnormalised with an FP64 scale each, so a renormalisation costs no pass. An
//     iteration is five kernels captured once in a graph; the host only looks at the stop flags every
//     few iterations.
//     Why two vector widths: u' = A v - alpha u cancels heavily on ill-conditioned systems, so the
//     rounding of the gathered operand v (not of A) is amplified; FP32 operands move the iterates by
//     ~1e-3 within 20 iterations on a regularised tomography system where FP64 operands agree with the
//     FP64 oracle to 1e-8 (FP32 matrix values alone perturb the problem consistently and cost little).
#include <algorithm>
#include <chrono>
#include <memory>
#include <random>
#include <vector>

#include <hipcub/hipcub.hpp>

#include "common.hpp"
#include "matrix.hpp"

using namespace ts;

namespace ts {

// Per right-hand-side state of the fast form, all FP64 (Lsqr.cs:55-66 and the loop's locals).
struct LsqrScal {
    double alpha, beta, rhobar, phibar, ddnorm, xxnorm, z, cs2, sn2, anorm2, dres2, rnorm, acond, bnorm, damp;
    double su, sv;      // stored U = u / su^-1 (u = su * U), v = sv * V
    double cx, cv, cw;  // x += cx w ; w = cv V + cw w
    double atol, btol, ctol;
    int it, stop, done, skip;
};

constexpr int kSpmmBlocks = 2048;
#ifndef TS_GATHER_BATCH
#define TS_GATHER_BATCH 16
#endif
constexpr int kGatherBatch = TS_GATHER_BATCH;  // FP64 gathers a row's lanes issue together (FP32: half)

// ------------------------------------------------------------------ matrix upload and transpose

__global__ void mx_row_of(const int32_t *__restrict__ rs, int rows, int32_t *__restrict__ row_of)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < rows; i += gridDim.x * blockDim.x)
        for (int32_t q = rs[i]; q < rs[i + 1]; q++) row_of[q] = i;
}

__global__ void mx_iota(int32_t *__restrict__ v, int64_t n)
{
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x)
        v[q] = (int32_t)q;
}

__global__ void mx_fill_i32(int32_t *__restrict__ v, int64_t n, int32_t x)
{
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x)
        v[q] = x;
}

__global__ void mx_count_cols(const int32_t *__restrict__ ci, int64_t nnz, int32_t *__restrict__ count)
{
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < nnz; q += (int64_t)gridDim.x * blockDim.x)
        atomicAdd(&count[ci[q]], 1);
}

__global__ void mx_to_f32(const double *__restrict__ a, float *__restrict__ b, int64_t n)
{
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x)
        b[q] = (float)a[q];
}

// transposed entry p is entry perm[p] of A: its row becomes the column index
__global__ void mx_gather(const int32_t *__restrict__ perm, const int32_t *__restrict__ row_of,
                          const double *__restrict__ va, int64_t nnz, int32_t *__restrict__ tci,
                          float *__restrict__ tva32, double *__restrict__ tva64)
{
    for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < nnz; p += (int64_t)gridDim.x * blockDim.x) {
        const int32_t q = perm[p];
        tci[p] = row_of[q];
        const double v = va[q];
        tva32[p] = (float)v;
        if (tva64) tva64[p] = v;
    }
}

static unsigned grid_for(int64_t n) { return (unsigned)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 16384)); }

// Row order of the stored matrix: rows sorted by (first column, last column), so rows that touch the
// same region of the model - rays ending at one station, the smoothing rows of neighbouring nodes -
// are products of neighbouring blocks and their gathers meet in the caches (7 % faster products on
// the 1e6-ray system than the order rows are given in). Empty rows go last.
__global__ void mx_row_keys(const int32_t *__restrict__ rs, const int32_t *__restrict__ ci, int rows,
                            uint64_t *__restrict__ key, int32_t *__restrict__ idx)
{
    for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < rows; r += gridDim.x * blockDim.x) {
        const int a = rs[r], b = rs[r + 1];
        key[r] = a < b ? ((uint64_t)(uint32_t)ci[a] << 32) | (uint32_t)ci[b - 1] : ~0ull;
        idx[r] = r;
    }
}

__global__ void mx_perm_lengths(const int32_t *__restrict__ rs, const int32_t *__restrict__ perm, int rows,
                                int32_t *__restrict__ len)
{
    for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < rows; r += gridDim.x * blockDim.x)
        len[r] = rs[perm[r] + 1] - rs[perm[r]];
}

// Row r of the new order is row perm[r] of the old one. A row shorter than kLongRow entries (the
// regularisation rows) is copied by one thread; a longer one (a ray's few hundred cells) is appended
// to `longs` and copied by a wavefront, its lanes on consecutive entries (one thread per ray row
// copied its entries serially: 40 ms of the 6.1e8-entry system's assembly).
constexpr int kLongRow = 64;

__global__ __launch_bounds__(256) void mx_perm_rows(const int32_t *__restrict__ rs, const int32_t *__restrict__ ci,
                                                    const double *__restrict__ va, const int32_t *__restrict__ perm,
                                                    const int32_t *__restrict__ nrs, int rows,
                                                    int32_t *__restrict__ nci, double *__restrict__ nva,
                                                    int32_t *__restrict__ longs, int32_t *__restrict__ nlong)
{
    for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < rows; r += gridDim.x * blockDim.x) {
        const int a = rs[perm[r]], b = rs[perm[r] + 1], o = nrs[r];
        if (b - a >= kLongRow) {
            longs[atomicAdd(nlong, 1)] = r;
            continue;
        }
        for (int q = a; q < b; q++) {
            nci[o + q - a] = ci[q];
            nva[o + q - a] = va[q];
        }
    }
}

// the long rows listed by mx_perm_rows, one wavefront each
__global__ __launch_bounds__(256) void mx_perm_long_rows(const int32_t *__restrict__ rs, const int32_t *__restrict__ ci,
                                                         const double *__restrict__ va, const int32_t *__restrict__ perm,
                                                         const int32_t *__restrict__ nrs, const int32_t *__restrict__ longs,
                                                         const int32_t *__restrict__ nlong, int32_t *__restrict__ nci,
                                                         double *__restrict__ nva)
{
    const int lane = threadIdx.x & 63;
    const int waves = (int)(gridDim.x * blockDim.x / 64);
    const int n = *nlong;
    for (int w = (int)((blockIdx.x * blockDim.x + threadIdx.x) / 64); w < n; w += waves) {
        const int r = longs[w];
        const int a = rs[perm[r]], b = rs[perm[r] + 1], o = nrs[r] - a;
        for (int q = a + lane; q < b; q += 64) {
            nci[o + q] = ci[q];
            nva[o + q] = va[q];
        }
    }
}

// ci[q] = map[ci[q]]: column indices into the stored column order
__global__ void mx_map_cols(int32_t *__restrict__ ci, int64_t nnz, const int32_t *__restrict__ map)
{
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < nnz; q += (int64_t)gridDim.x * blockDim.x)
        ci[q] = map[ci[q]];
}

// y[r] = sum_q va[q] x[ci[q]] in FP64 (CsrMatrix.MultiplyInto, SparseMatrix.cs:46-73), one row per thread
template <typename V>
__global__ void mx_spmv(int rows, const int32_t *__restrict__ rs, const int32_t *__restrict__ ci,
                        const V *__restrict__ va, const double *__restrict__ x, double *__restrict__ y)
{
    for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < rows; r += gridDim.x * blockDim.x) {
        double acc = 0;
        for (int q = rs[r]; q < rs[r + 1]; q++) acc += (double)va[q] * x[ci[q]];
        y[r] = acc;
    }
}

// dst[r] = src[perm[r]] over m rows
__global__ void mx_permute_vec(const double *__restrict__ src, const int32_t *__restrict__ perm, int m,
                               double *__restrict__ dst)
{
    for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < m; r += gridDim.x * blockDim.x) dst[r] = src[perm[r]];
}

static void reorder_rows(ts_matrix &A, dbuf<double> &va64, hipStream_t st)
{
    const int m = A.rows;
    const int64_t nnz = A.nnz;
    dbuf<uint64_t> key(m), key2(m);
    dbuf<int32_t> idx(m);
    A.row_perm.alloc(m);
    mx_row_keys<<<grid_for(m), 256, 0, st>>>(A.rs.p, A.ci.p, m, key.p, idx.p);
    size_t tb = 0;
    TS_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, key.p, key2.p, idx.p, A.row_perm.p, m, 0, 64, st));
    {
        dbuf<unsigned char> tmp(std::max<size_t>(1, tb));
        TS_HIP(hipcub::DeviceRadixSort::SortPairs(tmp.p, tb, key.p, key2.p, idx.p, A.row_perm.p, m, 0, 64, st));
    }
    dbuf<int32_t> len((size_t)m + 1), nrs((size_t)m + 1), nci(nnz);
    dbuf<double> nva(nnz);
    mx_fill_i32<<<1, 1, 0, st>>>(len.p + m, 1, 0);
    mx_perm_lengths<<<grid_for(m), 256, 0, st>>>(A.rs.p, A.row_perm.p, m, len.p);
    tb = 0;
    TS_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tb, len.p, nrs.p, m + 1, st));
    {
        dbuf<unsigned char> tmp(std::max<size_t>(1, tb));
        TS_HIP(hipcub::DeviceScan::ExclusiveSum(tmp.p, tb, len.p, nrs.p, m + 1, st));
    }
    dbuf<int32_t> longs(m), nlong(1);
    mx_fill_i32<<<1, 1, 0, st>>>(nlong.p, 1, 0);
    mx_perm_rows<<<grid_for(m), 256, 0, st>>>(A.rs.p, A.ci.p, va64.p, A.row_perm.p, nrs.p, m, nci.p, nva.p, longs.p,
                                             nlong.p);
    mx_perm_long_rows<<<4096, 256, 0, st>>>(A.rs.p, A.ci.p, va64.p, A.row_perm.p, nrs.p, longs.p, nlong.p, nci.p,
                                           nva.p);
    TS_HIP(hipGetLastError());
    TS_HIP(hipStreamSynchronize(st));
    A.rs = std::move(nrs);
    A.ci = std::move(nci);
    va64 = std::move(nva);
}

void matrix_finish(ts_matrix &A, dbuf<double> &&d_va64, bool keep_fp64, const int32_t *stored_col)
{
    hipStream_t st = A.ctx->stream;
    const int64_t nnz = A.nnz;
    if (stored_col && A.cols > 0) {
        std::vector<int32_t> given(A.cols);
        for (int c = 0; c < A.cols; c++) given[stored_col[c]] = c;
        dbuf<int32_t> map(A.cols);
        A.col_perm.alloc(A.cols);
        TS_HIP(hipMemcpyAsync(map.p, stored_col, (size_t)A.cols * 4, hipMemcpyHostToDevice, st));
        TS_HIP(hipMemcpyAsync(A.col_perm.p, given.data(), (size_t)A.cols * 4, hipMemcpyHostToDevice, st));
        if (nnz > 0) mx_map_cols<<<grid_for(nnz), 256, 0, st>>>(A.ci.p, nnz, map.p);
        TS_HIP(hipGetLastError());
        TS_HIP(hipStreamSynchronize(st));
    }
    if (nnz > 0 && A.rows > 1) reorder_rows(A, d_va64, st);
    A.trs.alloc((size_t)A.cols + 1);
    A.tci.alloc(std::max<int64_t>(1, nnz));
    A.va32.alloc(std::max<int64_t>(1, nnz));
    A.tva32.alloc(std::max<int64_t>(1, nnz));
    if (keep_fp64 && nnz > 0) A.tva64.alloc(nnz);
    // column counts -> transposed row starts
    {
        dbuf<int32_t> count((size_t)A.cols + 1);
        mx_fill_i32<<<grid_for(A.cols + 1), 256, 0, st>>>(count.p, A.cols + 1, 0);
        if (nnz > 0) mx_count_cols<<<grid_for(nnz), 256, 0, st>>>(A.ci.p, nnz, count.p);
        size_t tb = 0;
        TS_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tb, count.p, A.trs.p, A.cols + 1, st));
        dbuf<unsigned char> tmp(std::max<size_t>(1, tb));
        TS_HIP(hipcub::DeviceScan::ExclusiveSum(tmp.p, tb, count.p, A.trs.p, A.cols + 1, st));
        TS_HIP(hipGetLastError());
        TS_HIP(hipStreamSynchronize(st));
    }
    if (nnz > 0) {
        mx_to_f32<<<grid_for(nnz), 256, 0, st>>>(d_va64.p, A.va32.p, nnz);
        dbuf<int32_t> row_of(nnz), idx(nnz), perm(nnz), keys(nnz);
        mx_row_of<<<grid_for(A.rows), 256, 0, st>>>(A.rs.p, A.rows, row_of.p);
        mx_iota<<<grid_for(nnz), 256, 0, st>>>(idx.p, nnz);
        int end_bit = 1;
        while (end_bit < 31 && (1ll << end_bit) < (long long)A.cols) end_bit++;
        // LSD radix sort is stable: within a column the entries keep their row order
        size_t tb = 0;
        TS_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, A.ci.p, keys.p, idx.p, perm.p, (int)nnz, 0, end_bit, st));
        dbuf<unsigned char> tmp(std::max<size_t>(1, tb));
        TS_HIP(hipcub::DeviceRadixSort::SortPairs(tmp.p, tb, A.ci.p, keys.p, idx.p, perm.p, (int)nnz, 0, end_bit, st));
        mx_gather<<<grid_for(nnz), 256, 0, st>>>(perm.p, row_of.p, d_va64.p, nnz, A.tci.p, A.tva32.p,
                                                 keep_fp64 ? A.tva64.p : nullptr);
        TS_HIP(hipGetLastError());
        TS_HIP(hipStreamSynchronize(st));
    }
    if (keep_fp64 && nnz > 0) A.va64 = std::move(d_va64);
    else d_va64.reset();
}

// ------------------------------------------------------------------------------- fast kernels

// Y[r][j] = sc[j] * sum_q va[q] X[ci[q]][j] - sc[KP + j] * Y[r][j] for all rows, FP64 row sums; the
// squares of the stored results are summed per lane and reduced per block in a fixed order.
template <int KP, typename T, typename V>
__global__ __launch_bounds__(256) void lsqr_spmm(int rows, const int32_t *__restrict__ rs,
                                                 const int32_t *__restrict__ ci, const V *__restrict__ va,
                                                 const T *__restrict__ X, T *__restrict__ Y,
                                                 const double *__restrict__ sc, double *__restrict__ partial)
{
    constexpr int G = 256 / KP;
    const int j = threadIdx.x % KP, grp = threadIdx.x / KP;
    const double sx = sc[j], cy = sc[KP + j];
    // Block b runs on XCD b mod 8. A step of the grid covers gridDim.x * G consecutive rows, cut into
    // eight contiguous chunks, one per XCD: neighbouring rows of the stored order (rays between the
    // same regions, cells of one region) gather through one L2, and the eight chunks of a step are
    // neighbours in the shared last-level cache.
    const int chunk = (gridDim.x / 8) * G, lr = (blockIdx.x / 8) * G + grp;
    double sq = 0;
    for (int base = (blockIdx.x % 8) * chunk; base < rows; base += 8 * chunk) {
        const int r = base + lr;
        if (r >= rows) break;
        const int a = rs[r], b = rs[r + 1];
        double acc0 = 0, acc1 = 0;
        if constexpr (KP >= 8) {
            // The KP lanes of the row load KP column indices and values at once (one coalesced load
            // each) and pass them round with shuffles; the gathers of kGatherBatch entries are issued
            // together, so a wave keeps many gathers in flight instead of waiting on each index.
            constexpr int kBatch = sizeof(T) == 8 ? kGatherBatch : kGatherBatch / 2;
            constexpr int GB = KP < kBatch ? KP : kBatch;
            int cn = 0;
            V vn = 0;
            if (j < b - a) { cn = ci[a + j]; vn = va[a + j]; }
            for (int q0 = a; q0 < b; q0 += KP) {
                const int cnt = min(KP, b - q0);
                const int c = cn;
                const V v = vn;
                // the next chunk's indices and values load while this chunk gathers
                if (q0 + KP + j < b) { cn = ci[q0 + KP + j]; vn = va[q0 + KP + j]; }
                int t = 0;
                for (; t + GB <= cnt; t += GB) {
                    T x[GB];
                    V w[GB];
#pragma unroll
                    for (int u = 0; u < GB; u++) {
                        const int col = __shfl(c, t + u, KP);
                        w[u] = __shfl(v, t + u, KP);
                        x[u] = X[(size_t)col * KP + j];
                    }
#pragma unroll
                    for (int u = 0; u < GB; u += 2) {
                        acc0 += (double)w[u] * (double)x[u];
                        acc1 += (double)w[u + 1] * (double)x[u + 1];
                    }
                }
                for (; t < cnt; t++) {
                    const int col = __shfl(c, t, KP);
                    acc0 += (double)__shfl(v, t, KP) * (double)X[(size_t)col * KP + j];
                }
            }
        } else {
            int q = a;
            for (; q + 4 <= b; q += 4) {
                const int c0 = ci[q], c1 = ci[q + 1], c2 = ci[q + 2], c3 = ci[q + 3];
                const V v0 = va[q], v1 = va[q + 1], v2 = va[q + 2], v3 = va[q + 3];
                const T x0 = X[(size_t)c0 * KP + j], x1 = X[(size_t)c1 * KP + j];
                const T x2 = X[(size_t)c2 * KP + j], x3 = X[(size_t)c3 * KP + j];
                acc0 += (double)v0 * (double)x0;
                acc1 += (double)v1 * (double)x1;
                acc0 += (double)v2 * (double)x2;
                acc1 += (double)v3 * (double)x3;
            }
            for (; q < b; q++) acc0 += (double)va[q] * (double)X[(size_t)ci[q] * KP + j];
        }
        const size_t at = (size_t)r * KP + j;
        const double old = cy != 0.0 ? cy * (double)Y[at] : 0.0;
        const T y = (T)(sx * (acc0 + acc1) - old);
        Y[at] = y;
        sq += (double)y * (double)y;
    }
    __shared__ double red[256];
    red[threadIdx.x] = sq;
    __syncthreads();
    for (int s = G / 2; s > 0; s >>= 1) {
        if (grp < s) red[threadIdx.x] += red[threadIdx.x + s * KP];
        __syncthreads();
    }
    if (grp == 0) partial[(size_t)blockIdx.x * KP + j] = red[j];
}

// lsqr_spmm for KP < 8: 8 lanes per row, 8 / KP of them per right-hand side, each taking every
// (8 / KP)-th entry of the row; the lanes' sums meet in a fixed shuffle tree. A row of a few hundred
// entries (a ray among regularisation rows) then takes a few dozen dependent steps instead of holding
// its wave for the whole row. Same outputs and partials as lsqr_spmm, in another summation order.
template <int KP, typename T, typename V>
__global__ __launch_bounds__(256) void lsqr_spmm_split(int rows, const int32_t *__restrict__ rs,
                                                       const int32_t *__restrict__ ci, const V *__restrict__ va,
                                                       const T *__restrict__ X, T *__restrict__ Y,
                                                       const double *__restrict__ sc, double *__restrict__ partial)
{
    constexpr int L = 8 / KP;  // lanes per (row, right-hand side)
    constexpr int G = 256 / 8;
    const int slot = threadIdx.x >> 3, lane8 = threadIdx.x & 7, j = lane8 / L, sub = lane8 % L;
    const double sx = sc[j], cy = sc[KP + j];
    const int chunk = (gridDim.x / 8) * G, lr = (blockIdx.x / 8) * G + slot;
    double sq = 0;
    for (int base = (blockIdx.x % 8) * chunk; base < rows; base += 8 * chunk) {
        const int r = base + lr;
        if (r >= rows) break;
        const int a = rs[r], b = rs[r + 1];
        double acc0 = 0, acc1 = 0;
        int q = a + sub;
        for (; q + 3 * L < b; q += 4 * L) {
            const int c0 = ci[q], c1 = ci[q + L], c2 = ci[q + 2 * L], c3 = ci[q + 3 * L];
            const V v0 = va[q], v1 = va[q + L], v2 = va[q + 2 * L], v3 = va[q + 3 * L];
            const T x0 = X[(size_t)c0 * KP + j], x1 = X[(size_t)c1 * KP + j];
            const T x2 = X[(size_t)c2 * KP + j], x3 = X[(size_t)c3 * KP + j];
            acc0 += (double)v0 * (double)x0;
            acc1 += (double)v1 * (double)x1;
            acc0 += (double)v2 * (double)x2;
            acc1 += (double)v3 * (double)x3;
        }
        for (; q < b; q += L) acc0 += (double)va[q] * (double)X[(size_t)ci[q] * KP + j];
        double acc = acc0 + acc1;
#pragma unroll
        for (int m = 1; m < L; m <<= 1) acc += __shfl_xor(acc, m, 8);
        if (sub == 0) {
            const size_t at = (size_t)r * KP + j;
            const double old = cy != 0.0 ? cy * (double)Y[at] : 0.0;
            const T y = (T)(sx * acc - old);
            Y[at] = y;
            sq += (double)y * (double)y;
        }
    }
    // per right-hand side: the block's slots in a fixed tree, then its lanes in order
    __shared__ double red[256];
    red[threadIdx.x] = sq;
    __syncthreads();
    for (int s2 = G / 2; s2 > 0; s2 >>= 1) {
        if (slot < s2) red[threadIdx.x] += red[threadIdx.x + s2 * 8];
        __syncthreads();
    }
    if (threadIdx.x < KP) {
        double t = 0;
        for (int u = 0; u < L; u++) t += red[threadIdx.x * L + u];
        partial[(size_t)blockIdx.x * KP + threadIdx.x] = t;
    }
}

// One product: lsqr_spmm, or lsqr_spmm_split below 8 right-hand sides.
template <int KP, typename T, typename V>
void spmm(hipStream_t st, int rows, const int32_t *rs, const int32_t *ci, const V *va, const T *X, T *Y,
          const double *sc, double *partial)
{
    if constexpr (KP < 8) lsqr_spmm_split<KP, T, V><<<kSpmmBlocks, 256, 0, st>>>(rows, rs, ci, va, X, Y, sc, partial);
    else lsqr_spmm<KP, T, V><<<kSpmmBlocks, 256, 0, st>>>(rows, rs, ci, va, X, Y, sc, partial);
}

// The matrix's values in the type V: FP32 (always kept) or FP64 (kept when the matrix was created so).
template <typename V>
const V *values_of(const ts_matrix &A, bool transpose)
{
    if constexpr (sizeof(V) == 8) return transpose ? A.tva64.p : A.va64.p;
    else return transpose ? A.tva32.p : A.va32.p;
}

// x += cx w ; w = cv V + cw w over n rows; squares of the new w per block (for the next ddnorm).
template <int KP, typename T>
__global__ __launch_bounds__(256) void lsqr_xw(int n, const T *__restrict__ V, T *__restrict__ W,
                                               double *__restrict__ Xs, const LsqrScal *__restrict__ S,
                                               double *__restrict__ partial)
{
    constexpr int G = 256 / KP;
    const int j = threadIdx.x % KP, grp = threadIdx.x / KP;
    const double cx = S[j].cx, cv = S[j].cv, cw = S[j].cw;
    double sq = 0;
    for (int r = blockIdx.x * G + grp; r < n; r += gridDim.x * G) {
        const size_t at = (size_t)r * KP + j;
        const double w = W[at];
        Xs[at] += cx * w;
        const T wn = (T)(cv * (double)V[at] + cw * w);
        W[at] = wn;
        sq += (double)wn * (double)wn;
    }
    __shared__ double red[256];
    red[threadIdx.x] = sq;
    __syncthreads();
    for (int s = G / 2; s > 0; s >>= 1) {
        if (grp < s) red[threadIdx.x] += red[threadIdx.x + s * KP];
        __syncthreads();
    }
    if (grp == 0) partial[(size_t)blockIdx.x * KP + j] = red[j];
}

// The per-RHS sums of the product kernels' per-block partials, by the whole workgroup (kFinThreads
// threads, kp of them per stride): thread (part, j) sums blocks part, part + P, ... in order, then a
// fixed tree over the parts; every thread returns the sum of its j. The order is fixed: the same bits
// every run. Call from all threads (it synchronises).
constexpr int kFinThreads = 1024;

__device__ inline double sum_partials(const double *p, int blocks, int kp, double *red)
{
    const int t = threadIdx.x, j = t % kp, part = t / kp, parts = blockDim.x / kp;
    double s = 0;
    for (int b = part; b < blocks; b += parts) s += p[(size_t)b * kp + j];
    __syncthreads();  // red may hold an earlier call's values
    red[t] = s;
    __syncthreads();
    for (int h = parts / 2; h > 0; h >>= 1) {
        if (part < h) red[t] += red[t + h * kp];
        __syncthreads();
    }
    return red[j];
}

// After A v: beta, the u scale, ||A||^2 and the coefficients of the A^T product (Lsqr.cs:79-87).
__global__ __launch_bounds__(kFinThreads) void lsqr_fin_u(int k, int kp, LsqrScal *S, const double *partial, int blocks,
                                                         double *sc_t)
{
    __shared__ double red[kFinThreads];
    const double sum = sum_partials(partial, blocks, kp, red);
    const int j = threadIdx.x;
    if (j >= kp) return;
    if (j >= k) { sc_t[j] = 0; sc_t[kp + j] = -1; return; }
    LsqrScal s = S[j];
    if (s.done) { sc_t[j] = 0; sc_t[kp + j] = -1; return; }
    const double beta = sqrt(sum);
    s.beta = beta;
    if (beta > 0) {
        s.su = 1.0 / beta;
        s.anorm2 += s.alpha * s.alpha + beta * beta + s.damp * s.damp;
        s.skip = 0;
        sc_t[j] = s.su;               // V' = A^T u - beta v = su (A^T U) - beta sv V
        sc_t[kp + j] = beta * s.sv;
    } else {
        s.skip = 1;                   // u, v and alpha stay as they are (Lsqr.cs:81-90)
        sc_t[j] = 0;
        sc_t[kp + j] = -1;
    }
    S[j] = s;
}

// After A^T u: alpha, the v scale, the plane rotations, the norm estimates and the stopping rules of
// one iteration (Lsqr.cs:88-146), and the coefficients of the next A v product and of the x/w update.
__global__ __launch_bounds__(kFinThreads) void lsqr_fin_v(int k, int kp, LsqrScal *S, const double *partial_v,
                                                         const double *partial_w, int blocks_v, int blocks_w,
                                                         double *sc_a, double *hist, int max_it)
{
    __shared__ double red[kFinThreads];
    const double sum_w = sum_partials(partial_w, blocks_w, kp, red);
    const double sum_v = sum_partials(partial_v, blocks_v, kp, red);
    const int j = threadIdx.x;
    if (j >= kp) return;
    if (j >= k) { sc_a[j] = 0; sc_a[kp + j] = -1; return; }
    LsqrScal s = S[j];
    if (s.done) {
        // frozen: x and w keep their values, u and v are left as they are
        S[j].cx = 0;
        S[j].cv = 0;
        S[j].cw = 1;
        sc_a[j] = 0;
        sc_a[kp + j] = -1;
        return;
    }
    const double wnorm2 = sum_w;
    if (!s.skip) {
        s.alpha = sqrt(sum_v);
        if (s.alpha > 0) s.sv = 1.0 / s.alpha;
    }
    const double alpha = s.alpha, beta = s.beta, damp = s.damp;
    const double rhobar1 = sqrt(s.rhobar * s.rhobar + damp * damp);
    const double cs1 = s.rhobar / rhobar1, sn1 = damp / rhobar1;
    const double psi = sn1 * s.phibar;
    s.phibar = cs1 * s.phibar;
    const double rho = sqrt(rhobar1 * rhobar1 + beta * beta);
    const double cs = rhobar1 / rho, sn = beta / rho;
    const double theta = sn * alpha;
    s.rhobar = -cs * alpha;
    const double phi = cs * s.phibar;
    s.phibar = sn * s.phibar;
    const double tau = sn * phi;
    s.ddnorm += wnorm2 / (rho * rho);
    s.cx = phi / rho;
    s.cv = s.sv;
    s.cw = -theta / rho;
    const double delta = s.sn2 * rho, gambar = -s.cs2 * rho;
    const double rhs = phi - delta * s.z;
    const double gamma = sqrt(gambar * gambar + theta * theta);
    s.cs2 = gambar / gamma;
    s.sn2 = theta / gamma;
    s.z = rhs / gamma;
    s.xxnorm += s.z * s.z;
    s.dres2 += psi * psi;
    const double anorm = sqrt(s.anorm2);
    s.acond = anorm * sqrt(s.ddnorm);
    const double arnorm = alpha * fabs(tau);
    s.rnorm = sqrt(s.phibar * s.phibar + s.dres2);
    const double xnorm = sqrt(s.xxnorm);
    const double test1 = s.rnorm / s.bnorm;
    const double test2 = arnorm / fmax(1e-300, anorm * s.rnorm);
    const double test3 = 1.0 / fmax(1e-300, s.acond);
    const double t1 = test1 / (1 + anorm * xnorm / s.bnorm);
    const double rtol = s.btol + s.atol * anorm * xnorm / s.bnorm;
    if (hist) hist[(size_t)j * max_it + s.it] = s.rnorm;
    s.it += 1;
    int stop = 0;
    if (test1 <= rtol) stop = 1;
    else if (test2 <= s.atol) stop = 2;
    else if (s.ctol > 0 && test3 <= s.ctol) stop = 3;
    else if (1 + t1 <= 1) stop = 4;
    else if (1 + test2 <= 1) stop = 5;
    else if (1 + test3 <= 1) stop = 6;
    else if (beta == 0 || alpha == 0) stop = 7;
    else if (s.it >= max_it) stop = -1;  // budget exhausted: reported as 0
    if (stop != 0) {
        s.stop = stop < 0 ? 0 : stop;
        s.done = 1;
    }
    // next A v product: U' = A v - alpha u = sv (A V) - alpha su U
    if (s.done) { sc_a[j] = 0; sc_a[kp + j] = -1; }
    else { sc_a[j] = s.sv; sc_a[kp + j] = alpha * s.su; }
    S[j] = s;
}

// Start of the solve (Lsqr.cs:51-66) once U = b and beta = ||b|| are set and V = A^T U: alpha, the
// first w = v and the scalars.
__global__ __launch_bounds__(kFinThreads) void lsqr_fin_init(int k, int kp, LsqrScal *S, const double *partial_v,
                                                            int blocks, double *sc_a)
{
    __shared__ double red[kFinThreads];
    const double sum = sum_partials(partial_v, blocks, kp, red);
    const int j = threadIdx.x;
    if (j >= kp) return;
    if (j >= k) { sc_a[j] = 0; sc_a[kp + j] = -1; return; }
    LsqrScal s = S[j];
    if (!s.done) {
        s.alpha = sqrt(sum);
        if (s.alpha == 0) {
            s.done = 1;
            s.stop = 10;
            s.rnorm = s.bnorm;
        } else {
            s.sv = 1.0 / s.alpha;
            s.rhobar = s.alpha;
            s.phibar = s.beta;
            s.rnorm = s.beta;
        }
    }
    s.cx = 0;
    s.cv = s.done ? 0.0 : s.sv;  // w = v
    s.cw = 0;
    if (s.done) { sc_a[j] = 0; sc_a[kp + j] = -1; }
    else { sc_a[j] = s.sv; sc_a[kp + j] = s.alpha * s.su; }
    S[j] = s;
}

template <typename T>
__global__ void lsqr_load_b(int m, int k, int kp, const double *b, const int32_t *perm, T *U, const LsqrScal *S)
{
    size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= (size_t)m * kp) return;
    const int j = (int)(q % kp);
    const size_t r = q / kp;
    U[q] = j < k ? (T)b[(size_t)j * m + (perm ? perm[r] : (int)r)] : (T)0;
}

// x[j][given column] from the stored [column][kp] solution (col_perm null: the same order)
__global__ void lsqr_unpack_x(int n, int k, int kp, const double *Xs, const int32_t *col_perm, double *x)
{
    size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= (size_t)n * k) return;
    const int j = (int)(q / n);
    const size_t r = q - (size_t)j * n;
    x[(size_t)j * n + (col_perm ? col_perm[r] : (int)r)] = Xs[r * kp + j];
}

// dst[perm[c]] = src[c] over n columns
__global__ void mx_unpermute_vec(const double *__restrict__ src, const int32_t *__restrict__ perm, int n,
                                 double *__restrict__ dst)
{
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < n; c += gridDim.x * blockDim.x) dst[perm[c]] = src[c];
}

// ------------------------------------------------------------------------------- plain kernels

// The reference's device kernels in FP64 (LsqrOpenCl.cs:45-95): one work item per CSR row, BLAS-1
// updates, and a dot product of 64 groups x 256 items whose partials the host adds in order.
__global__ void plain_spmv(int rows, const int32_t *rs, const int32_t *ci, const double *va, const double *x,
                           double *y)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= rows) return;
    double sum = 0;
    for (int q = rs[i]; q < rs[i + 1]; q++) sum += va[q] * x[ci[q]];
    y[i] = sum;
}
__global__ void plain_xmby(int n, const double *x, double b, double *y)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = x[i] - b * y[i];
}
__global__ void plain_xpby(int n, const double *x, double b, double *y)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = x[i] + b * y[i];
}
__global__ void plain_axpy(int n, double a, const double *x, double *y)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] += a * x[i];
}
__global__ void plain_scal(int n, double *x, double s)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] *= s;
}
__global__ __launch_bounds__(256) void plain_dot_partial(int n, const double *a, const double *b, double *partial)
{
    __shared__ double red[256];
    double s = 0;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += 64 * 256) s += a[i] * b[i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

inline unsigned nblk(size_t n, unsigned bs) { return (unsigned)((n + bs - 1) / bs); }

// ------------------------------------------------------------------------------- host drivers

struct SolveOut {
    double *x;          // k * n
    int32_t *iterations, *stop;
    double *stats;      // k * 3: bnorm, rnorm, acond
    double *hist;       // k * max_it or null
};

void lsqr_plain(ts_matrix &A, int k, const double *b, const double *damp, int max_it, double atol, double btol,
                double conlim, const volatile int32_t *cancel, SolveOut o)
{
    hipStream_t st = exec_stream(A.ctx);
    const int m = A.rows, n = A.cols;
    if (!A.va64.p) fail(TS_EINVAL, "matrix was created without FP64 values");
    dbuf<double> u(m), v(n), w(n), x(n), tm(m), tn(n), part(64);
    double *hp = nullptr;
    TS_HIP(hipHostMalloc((void **)&hp, 64 * sizeof(double), hipHostMallocDefault));
    struct Free { double *p; ~Free() { if (p) (void)hipHostFree(p); } } fh{hp};
    auto dot = [&](const double *a, const double *bb, int len) {
        plain_dot_partial<<<64, 256, 0, st>>>(len, a, bb, part.p);
        TS_HIP(hipMemcpyAsync(hp, part.p, 64 * sizeof(double), hipMemcpyDeviceToHost, st));
        TS_HIP(hipStreamSynchronize(st));
        double s = 0;
        for (int g = 0; g < 64; g++) s += hp[g];
        return s;
    };
    auto norm = [&](const double *a, int len) { return sqrt(std::max(0.0, dot(a, a, len))); };
    auto spmv = [&](bool t, const double *xin, double *yout) {
        int rows = t ? n : m;
        plain_spmv<<<nblk(rows, 256), 256, 0, st>>>(rows, t ? A.trs.p : A.rs.p, t ? A.tci.p : A.ci.p,
                                                    t ? A.tva64.p : A.va64.p, xin, yout);
    };
    for (int j = 0; j < k; j++) {
        const double dmp = damp[j];
        if (A.row_perm.p) {
            // b in the stored row order
            TS_HIP(hipMemcpyAsync(tm.p, b + (size_t)j * m, (size_t)m * 8, hipMemcpyHostToDevice, st));
            mx_permute_vec<<<grid_for(m), 256, 0, st>>>(tm.p, A.row_perm.p, m, u.p);
        } else {
            TS_HIP(hipMemcpyAsync(u.p, b + (size_t)j * m, (size_t)m * 8, hipMemcpyHostToDevice, st));
        }
        TS_HIP(hipMemsetAsync(v.p, 0, (size_t)n * 8, st));
        TS_HIP(hipMemsetAsync(w.p, 0, (size_t)n * 8, st));
        TS_HIP(hipMemsetAsync(x.p, 0, (size_t)n * 8, st));
        TS_HIP(hipMemsetAsync(tn.p, 0, (size_t)n * 8, st));
        int it_done = 0, stop = 0;
        double bnorm = 0, rnorm = 0, acond = 0;
        double beta = norm(u.p, m);
        if (beta == 0) {
            stop = 9;
        } else {
            bnorm = beta;
            plain_scal<<<nblk(m, 256), 256, 0, st>>>(m, u.p, 1.0 / beta);
            spmv(true, u.p, v.p);
            double alpha = norm(v.p, n);
            if (alpha == 0) {
                stop = 10;
                rnorm = bnorm;
            } else {
                plain_scal<<<nblk(n, 256), 256, 0, st>>>(n, v.p, 1.0 / alpha);
                TS_HIP(hipMemcpyAsync(w.p, v.p, (size_t)n * 8, hipMemcpyDeviceToDevice, st));
                double rhobar = alpha, phibar = beta, ddnorm = 0, xxnorm = 0, z = 0, cs2 = -1, sn2 = 0;
                double anorm2 = 0, dres2 = 0;
                rnorm = beta;
                const double ctol = conlim > 0 ? 1.0 / conlim : 0.0;
                for (int it = 0; it < max_it; it++) {
                    if (cancelled(cancel)) fail(TS_CANCELLED, "cancelled");
                    it_done = it + 1;
                    spmv(false, v.p, tm.p);
                    plain_xmby<<<nblk(m, 256), 256, 0, st>>>(m, tm.p, alpha, u.p);
                    beta = norm(u.p, m);
                    if (beta > 0) {
                        plain_scal<<<nblk(m, 256), 256, 0, st>>>(m, u.p, 1.0 / beta);
                        anorm2 += alpha * alpha + beta * beta + dmp * dmp;
                        spmv(true, u.p, tn.p);
                        plain_xmby<<<nblk(n, 256), 256, 0, st>>>(n, tn.p, beta, v.p);
                        alpha = norm(v.p, n);
                        if (alpha > 0) plain_scal<<<nblk(n, 256), 256, 0, st>>>(n, v.p, 1.0 / alpha);
                    }
                    double rhobar1 = sqrt(rhobar * rhobar + dmp * dmp);
                    double cs1 = rhobar / rhobar1, sn1 = dmp / rhobar1;
                    double psi = sn1 * phibar;
                    phibar = cs1 * phibar;
                    double rho = sqrt(rhobar1 * rhobar1 + beta * beta);
                    double cs = rhobar1 / rho, sn = beta / rho;
                    double theta = sn * alpha;
                    rhobar = -cs * alpha;
                    double phi = cs * phibar;
                    phibar = sn * phibar;
                    double tau = sn * phi;
                    ddnorm += dot(w.p, w.p, n) / (rho * rho);
                    plain_axpy<<<nblk(n, 256), 256, 0, st>>>(n, phi / rho, w.p, x.p);
                    plain_xpby<<<nblk(n, 256), 256, 0, st>>>(n, v.p, -theta / rho, w.p);
                    double delta = sn2 * rho, gambar = -cs2 * rho;
                    double rhs = phi - delta * z;
                    double gamma = sqrt(gambar * gambar + theta * theta);
                    cs2 = gambar / gamma;
                    sn2 = theta / gamma;
                    z = rhs / gamma;
                    xxnorm += z * z;
                    dres2 += psi * psi;
                    double anorm = sqrt(anorm2);
                    acond = anorm * sqrt(ddnorm);
                    double arnorm = alpha * fabs(tau);
                    rnorm = sqrt(phibar * phibar + dres2);
                    double xnorm = sqrt(xxnorm);
                    double test1 = rnorm / bnorm, test2 = arnorm / std::max(1e-300, anorm * rnorm);
                    double test3 = 1.0 / std::max(1e-300, acond);
                    double t1 = test1 / (1 + anorm * xnorm / bnorm);
                    double rtol = btol + atol * anorm * xnorm / bnorm;
                    if (o.hist) o.hist[(size_t)j * max_it + it] = rnorm;
                    if (test1 <= rtol) { stop = 1; break; }
                    if (test2 <= atol) { stop = 2; break; }
                    if (ctol > 0 && test3 <= ctol) { stop = 3; break; }
                    if (1 + t1 <= 1) { stop = 4; break; }
                    if (1 + test2 <= 1) { stop = 5; break; }
                    if (1 + test3 <= 1) { stop = 6; break; }
                    if (beta == 0 || alpha == 0) { stop = 7; break; }
                }
            }
        }
        if (A.col_perm.p) {  // the solution in the given column order
            mx_unpermute_vec<<<grid_for(n), 256, 0, st>>>(x.p, A.col_perm.p, n, tn.p);
            TS_HIP(hipMemcpyAsync(o.x + (size_t)j * n, tn.p, (size_t)n * 8, hipMemcpyDeviceToHost, st));
        } else {
            TS_HIP(hipMemcpyAsync(o.x + (size_t)j * n, x.p, (size_t)n * 8, hipMemcpyDeviceToHost, st));
        }
        TS_HIP(hipStreamSynchronize(st));
        o.iterations[j] = it_done;
        o.stop[j] = stop;
        o.stats[3 * j] = bnorm;
        o.stats[3 * j + 1] = rnorm;
        o.stats[3 * j + 2] = acond;
    }
}

// The fast form's device state, kept on the matrix handle's side between the create and the solve.
template <typename T>
struct FastLsqr {
    int k = 0, kp = 0, max_it = 0;
    dbuf<T> U, V, W;
    dbuf<double> X, pu, pv, pw, sc_a, sc_t, hist, bd;
    dbuf<LsqrScal> S;
};

template <int KP, typename T, typename V>
void fast_iteration(ts_matrix &A, FastLsqr<T> &F, hipStream_t st, bool with_hist)
{
    spmm<KP, T, V>(st, A.rows, A.rs.p, A.ci.p, values_of<V>(A, false), F.V.p, F.U.p, F.sc_a.p, F.pu.p);
    lsqr_fin_u<<<1, kFinThreads, 0, st>>>(F.k, KP, F.S.p, F.pu.p, kSpmmBlocks, F.sc_t.p);
    spmm<KP, T, V>(st, A.cols, A.trs.p, A.tci.p, values_of<V>(A, true), F.U.p, F.V.p, F.sc_t.p, F.pv.p);
    lsqr_fin_v<<<1, kFinThreads, 0, st>>>(F.k, KP, F.S.p, F.pv.p, F.pw.p, kSpmmBlocks, kSpmmBlocks, F.sc_a.p,
                                 with_hist ? F.hist.p : nullptr, F.max_it);
    lsqr_xw<KP, T><<<kSpmmBlocks, 256, 0, st>>>(A.cols, F.V.p, F.W.p, F.X.p, F.S.p, F.pw.p);
}

template <int KP, typename T, typename V>
void lsqr_fast_kp(ts_matrix &A, int k, const double *b, const double *damp, int max_it, double atol, double btol,
                  double conlim, const volatile int32_t *cancel, SolveOut o, double *timing)
{
    hipStream_t st = exec_stream(A.ctx);
    const int m = A.rows, n = A.cols;
    FastLsqr<T> F;
    F.k = k;
    F.kp = KP;
    F.max_it = max_it;
    F.U.alloc((size_t)m * KP);
    F.V.alloc((size_t)n * KP);
    F.W.alloc((size_t)n * KP);
    F.X.alloc((size_t)n * KP);
    F.pu.alloc((size_t)kSpmmBlocks * KP);
    F.pv.alloc((size_t)kSpmmBlocks * KP);
    F.pw.alloc((size_t)kSpmmBlocks * KP);
    F.sc_a.alloc(2 * KP);
    F.sc_t.alloc(2 * KP);
    F.S.alloc(KP);
    if (o.hist && max_it > 0) F.hist.alloc((size_t)k * max_it);
    F.bd.alloc((size_t)k * m);
    // b and its FP64 norms (Lsqr.cs:51-55)
    TS_HIP(hipMemcpyAsync(F.bd.p, b, (size_t)k * m * 8, hipMemcpyHostToDevice, st));
    std::vector<LsqrScal> S(KP);
    for (int j = 0; j < KP; j++) {
        LsqrScal &s = S[j];
        std::memset(&s, 0, sizeof s);
        s.cs2 = -1;
        s.su = s.sv = 1;
        s.atol = atol;
        s.btol = btol;
        s.ctol = conlim > 0 ? 1.0 / conlim : 0.0;
        if (j >= k) { s.done = 1; continue; }
        s.damp = damp[j];
        double ss = 0, mx = 0;
        const double *bj = b + (size_t)j * m;
        for (int r = 0; r < m; r++) mx = std::max(mx, fabs(bj[r]));
        if (mx > 0 && (mx > 1e150 || mx < 1e-150)) {
            for (int r = 0; r < m; r++) { double y = bj[r] / mx; ss += y * y; }
            s.beta = mx * sqrt(ss);
        } else {
            for (int r = 0; r < m; r++) ss += bj[r] * bj[r];
            s.beta = sqrt(ss);
        }
        s.bnorm = s.beta;
        if (s.beta == 0) { s.done = 1; s.stop = 9; continue; }
        s.su = 1.0 / s.beta;
    }
    TS_HIP(hipMemcpyAsync(F.S.p, S.data(), KP * sizeof(LsqrScal), hipMemcpyHostToDevice, st));
    lsqr_load_b<T><<<nblk((size_t)m * KP, 256), 256, 0, st>>>(m, k, KP, F.bd.p, A.row_perm.p, F.U.p, F.S.p);
    TS_HIP(hipMemsetAsync(F.V.p, 0, (size_t)n * KP * sizeof(T), st));
    TS_HIP(hipMemsetAsync(F.W.p, 0, (size_t)n * KP * sizeof(T), st));
    TS_HIP(hipMemsetAsync(F.X.p, 0, (size_t)n * KP * 8, st));
    std::vector<double> sc(2 * KP);
    for (int j = 0; j < KP; j++) { sc[j] = S[j].done ? 0.0 : S[j].su; sc[KP + j] = 0; }
    TS_HIP(hipMemcpyAsync(F.sc_t.p, sc.data(), 2 * KP * 8, hipMemcpyHostToDevice, st));
    // v = A^T u, alpha, w = v
    spmm<KP, T, V>(st, n, A.trs.p, A.tci.p, values_of<V>(A, true), F.U.p, F.V.p, F.sc_t.p, F.pv.p);
    lsqr_fin_init<<<1, kFinThreads, 0, st>>>(k, KP, F.S.p, F.pv.p, kSpmmBlocks, F.sc_a.p);
    lsqr_xw<KP, T><<<kSpmmBlocks, 256, 0, st>>>(n, F.V.p, F.W.p, F.X.p, F.S.p, F.pw.p);
    TS_HIP(hipGetLastError());
    TS_HIP(hipStreamSynchronize(st));
    // one iteration as a graph, replayed; the stop flags are read every `check` iterations
    hipGraphExec_t ge = nullptr;
    // graphs only on the owner's own stream: a serving worker launches directly (captures on several
    // threads' streams at once are not reliable)
    if (A.ctx->use_graphs && !stream_override()) {
        hipGraph_t gr;
        TS_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        fast_iteration<KP, T, V>(A, F, st, o.hist != nullptr);
        TS_HIP(hipStreamEndCapture(st, &gr));
        hipError_t e = hipGraphInstantiate(&ge, gr, nullptr, nullptr, 0);
        (void)hipGraphDestroy(gr);
        TS_HIP(e);
    }
    struct FreeG { hipGraphExec_t g; ~FreeG() { if (g) (void)hipGraphExecDestroy(g); } } fg{ge};
    hipEvent_t ev0, ev1;
    TS_HIP(hipEventCreate(&ev0));
    TS_HIP(hipEventCreate(&ev1));
    struct FreeEv { hipEvent_t a, b; ~FreeEv() { (void)hipEventDestroy(a); (void)hipEventDestroy(b); } } fe{ev0, ev1};
    const int check = 8;
    std::vector<LsqrScal> cur(KP);
    TS_HIP(hipEventRecord(ev0, st));
    int it = 0;
    while (it < max_it) {
        if (cancelled(cancel)) fail(TS_CANCELLED, "cancelled");
        int todo = std::min(check, max_it - it);
        for (int q = 0; q < todo; q++) {
            if (ge) TS_HIP(hipGraphLaunch(ge, st));
            else fast_iteration<KP, T, V>(A, F, st, o.hist != nullptr);
        }
        it += todo;
        TS_HIP(hipMemcpyAsync(cur.data(), F.S.p, KP * sizeof(LsqrScal), hipMemcpyDeviceToHost, st));
        TS_HIP(hipStreamSynchronize(st));
        bool all = true;
        for (int j = 0; j < k; j++) all = all && cur[j].done;
        if (all) break;
    }
    TS_HIP(hipEventRecord(ev1, st));
    dbuf<double> xo((size_t)n * k);
    lsqr_unpack_x<<<nblk((size_t)n * k, 256), 256, 0, st>>>(n, k, KP, F.X.p, A.col_perm.p, xo.p);
    TS_HIP(hipMemcpyAsync(o.x, xo.p, (size_t)n * k * 8, hipMemcpyDeviceToHost, st));
    TS_HIP(hipMemcpyAsync(cur.data(), F.S.p, KP * sizeof(LsqrScal), hipMemcpyDeviceToHost, st));
    if (o.hist && F.hist.p) TS_HIP(hipMemcpyAsync(o.hist, F.hist.p, (size_t)k * max_it * 8, hipMemcpyDeviceToHost, st));
    TS_HIP(hipStreamSynchronize(st));
    float ms = 0;
    TS_HIP(hipEventElapsedTime(&ms, ev0, ev1));
    if (timing) *timing = ms * 1e-3;
    for (int j = 0; j < k; j++) {
        o.iterations[j] = cur[j].stop == 9 ? 0 : cur[j].it;
        o.stop[j] = cur[j].stop;
        o.stats[3 * j] = cur[j].bnorm;
        o.stats[3 * j + 1] = cur[j].stop == 9 ? 0 : cur[j].rnorm;
        o.stats[3 * j + 2] = cur[j].acond;
    }
}

template <typename T, typename V>
void lsqr_fast_t(ts_matrix &A, int k, const double *b, const double *damp, int max_it, double atol, double btol,
                 double conlim, const volatile int32_t *cancel, SolveOut o, double *timing)
{
    if (k <= 1) lsqr_fast_kp<1, T, V>(A, k, b, damp, max_it, atol, btol, conlim, cancel, o, timing);
    else if (k <= 2) lsqr_fast_kp<2, T, V>(A, k, b, damp, max_it, atol, btol, conlim, cancel, o, timing);
    else if (k <= 4) lsqr_fast_kp<4, T, V>(A, k, b, damp, max_it, atol, btol, conlim, cancel, o, timing);
    else if (k <= 8) lsqr_fast_kp<8, T, V>(A, k, b, damp, max_it, atol, btol, conlim, cancel, o, timing);
    else if (k <= 16) lsqr_fast_kp<16, T, V>(A, k, b, damp, max_it, atol, btol, conlim, cancel, o, timing);
    else if (k <= 32) lsqr_fast_kp<32, T, V>(A, k, b, damp, max_it, atol, btol, conlim, cancel, o, timing);
    else lsqr_fast_kp<64, T, V>(A, k, b, damp, max_it, atol, btol, conlim, cancel, o, timing);
}

// vectors in FP32 (TS_LSQR_FAST_FP32) or FP64 (TS_LSQR_FAST_MIXED, TS_LSQR_FAST_FP64); matrix values FP32,
// or FP64 in TS_LSQR_FAST_FP64 (the matrix must keep them)
void lsqr_fast(ts_matrix &A, int k, const double *b, const double *damp, int max_it, double atol, double btol,
               double conlim, const volatile int32_t *cancel, SolveOut o, double *timing, int mode)
{
    if (mode == TS_LSQR_FAST_FP64) {
        if (!A.va64.p || !A.tva64.p) fail(TS_EINVAL, "the FP64-valued form needs a matrix created with its FP64 values kept");
        lsqr_fast_t<double, double>(A, k, b, damp, max_it, atol, btol, conlim, cancel, o, timing);
    } else if (mode == TS_LSQR_FAST_MIXED) lsqr_fast_t<double, float>(A, k, b, damp, max_it, atol, btol, conlim, cancel, o, timing);
    else lsqr_fast_t<float, float>(A, k, b, damp, max_it, atol, btol, conlim, cancel, o, timing);
}

// ------------------------------------------------------------------------------- self-test

// Host LSQR in FP64 for the gate only (Lsqr.cs:34-152, with plain loops).
static std::vector<double> host_lsqr(int m, int n, const std::vector<int32_t> &rs, const std::vector<int32_t> &ci,
                                     const std::vector<double> &va, const std::vector<double> &b, double damp,
                                     int max_it, double atol, double btol, double conlim)
{
    std::vector<int32_t> trs(n + 1, 0), tci(ci.size());
    std::vector<double> tva(va.size());
    for (int32_t c : ci) trs[c + 1]++;
    for (int c = 0; c < n; c++) trs[c + 1] += trs[c];
    std::vector<int32_t> nx(trs.begin(), trs.end());
    for (int i = 0; i < m; i++)
        for (int q = rs[i]; q < rs[i + 1]; q++) { int p = nx[ci[q]]++; tci[p] = i; tva[p] = va[q]; }
    auto mv = [](int rows, const std::vector<int32_t> &r, const std::vector<int32_t> &c, const std::vector<double> &v,
                 const std::vector<double> &x, std::vector<double> &y) {
        for (int i = 0; i < rows; i++) { double s = 0; for (int q = r[i]; q < r[i + 1]; q++) s += v[q] * x[c[q]]; y[i] = s; }
    };
    auto nrm = [](const std::vector<double> &a) { double s = 0; for (double v : a) s += v * v; return sqrt(s); };
    std::vector<double> u(b), v(n), w(n), x(n, 0.0), tm(m), tn(n);
    double beta = nrm(u);
    if (beta == 0) return x;
    double bnorm = beta;
    for (auto &q : u) q /= beta;
    mv(n, trs, tci, tva, u, v);
    double alpha = nrm(v);
    if (alpha == 0) return x;
    for (auto &q : v) q /= alpha;
    w = v;
    double rhobar = alpha, phibar = beta, ddnorm = 0, xxnorm = 0, z = 0, cs2 = -1, sn2 = 0, anorm2 = 0, dres2 = 0;
    double ctol = conlim > 0 ? 1.0 / conlim : 0.0;
    for (int it = 0; it < max_it; it++) {
        mv(m, rs, ci, va, v, tm);
        for (int i = 0; i < m; i++) u[i] = tm[i] - alpha * u[i];
        beta = nrm(u);
        if (beta > 0) {
            for (auto &q : u) q /= beta;
            anorm2 += alpha * alpha + beta * beta + damp * damp;
            mv(n, trs, tci, tva, u, tn);
            for (int i = 0; i < n; i++) v[i] = tn[i] - beta * v[i];
            alpha = nrm(v);
            if (alpha > 0) for (auto &q : v) q /= alpha;
        }
        double rhobar1 = sqrt(rhobar * rhobar + damp * damp), cs1 = rhobar / rhobar1, sn1 = damp / rhobar1;
        double psi = sn1 * phibar;
        phibar = cs1 * phibar;
        double rho = sqrt(rhobar1 * rhobar1 + beta * beta), cs = rhobar1 / rho, sn = beta / rho;
        double theta = sn * alpha;
        rhobar = -cs * alpha;
        double phi = cs * phibar;
        phibar = sn * phibar;
        double tau = sn * phi;
        double ww = 0;
        for (double q : w) ww += q * q;
        ddnorm += ww / (rho * rho);
        for (int i = 0; i < n; i++) { x[i] += phi / rho * w[i]; w[i] = v[i] - theta / rho * w[i]; }
        double delta = sn2 * rho, gambar = -cs2 * rho, rhs = phi - delta * z;
        double gamma = sqrt(gambar * gambar + theta * theta);
        cs2 = gambar / gamma; sn2 = theta / gamma; z = rhs / gamma;
        xxnorm += z * z;
        dres2 += psi * psi;
        double anorm = sqrt(anorm2), acond = anorm * sqrt(ddnorm), arnorm = alpha * fabs(tau);
        double rnorm = sqrt(phibar * phibar + dres2), xnorm = sqrt(xxnorm);
        double test1 = rnorm / bnorm, test2 = arnorm / std::max(1e-300, anorm * rnorm), test3 = 1.0 / std::max(1e-300, acond);
        double t1 = test1 / (1 + anorm * xnorm / bnorm), rtol = btol + atol * anorm * xnorm / bnorm;
        if (test1 <= rtol || test2 <= atol || (ctol > 0 && test3 <= ctol) || 1 + t1 <= 1 || 1 + test2 <= 1 ||
            1 + test3 <= 1 || beta == 0 || alpha == 0)
            break;
    }
    return x;
}

// LsqrOpenCl.SelfTest (LsqrOpenCl.cs:541-573): a random 3000 x 1200 system with 12 non-zeros a row,
// values in [0.1, 1.1), right-hand side in [-0.5, 0.5), damp 0.05, 300 iterations. The plain FP64
// form must agree with the host to 1e-8 of max |x| (the reference's gate), the fast FP32 form to
// 1e-5 (its stated precision).
double lsqr_selftest(ts_context *ctx, double *fast_err, char *msg, int len)
{
    const int m = 3000, n = 1200, per = 12;
    std::mt19937_64 rng(17);
    std::uniform_real_distribution<double> u01(0.0, 1.0);
    std::vector<int32_t> rs{0}, ci;
    std::vector<double> va, b;
    std::vector<int> cols(n);
    for (int i = 0; i < m; i++) {
        std::vector<int> pick;
        while ((int)pick.size() < per) {
            int c = (int)(rng() % n);
            if (std::find(pick.begin(), pick.end(), c) == pick.end()) pick.push_back(c);
        }
        std::sort(pick.begin(), pick.end());
        for (int c : pick) { ci.push_back(c); va.push_back(u01(rng) + 0.1); }
        rs.push_back((int32_t)ci.size());
        b.push_back(u01(rng) - 0.5);
    }
    std::vector<double> host = host_lsqr(m, n, rs, ci, va, b, 0.05, 300, 1e-12, 1e-12, 1e10);
    ts_matrix *A = nullptr;
    char reason[256];
    int rc = ts_matrix_create(ctx, m, n, rs.data(), ci.data(), va.data(), 1, &A, reason, sizeof reason);
    if (rc != TS_OK) { snprintf(msg, len, "matrix: %s", reason); return INFINITY; }
    double worst[2] = {0, 0}, scale = 0;
    int iters[2] = {0, 0};
    for (int mode = 0; mode < 2; mode++) {
        std::vector<double> x(n), stats(3);
        int32_t it = 0, stop = 0;
        double damp = 0.05;
        rc = ts_lsqr_solve(A, 1, b.data(), &damp, 300, 1e-12, 1e-12, 1e10, mode, nullptr, x.data(), &it, &stop,
                           stats.data(), nullptr, reason, sizeof reason);
        if (rc != TS_OK) {
            ts_matrix_release(A);
            snprintf(msg, len, "mode %d: %s", mode, reason);
            return INFINITY;
        }
        iters[mode] = it;
        for (int i = 0; i < n; i++) {
            worst[mode] = std::max(worst[mode], fabs(x[i] - host[i]));
            scale = std::max(scale, fabs(host[i]));
        }
    }
    ts_matrix_release(A);
    double rel = worst[0] / std::max(1e-300, scale);
    if (fast_err) *fast_err = worst[1] / std::max(1e-300, scale);
    snprintf(msg, len, "max |dx| %.2e (plain FP64, %d it), %.2e (fast FP32, %d it) of max |x|", rel, iters[0],
             worst[1] / std::max(1e-300, scale), iters[1]);
    return rel;
}

}  // namespace ts

namespace ts {

// ts_lsqr_solve without the matrix's gate, on the calling thread's stream (exec_stream): concurrent
// solves on one resident matrix only read it. seconds: device time of the fast form's iterations.
int lsqr_solve_core(ts_matrix *a, int k, const double *b, const double *damp, int max_iterations, double atol,
                    double btol, double conlim, int mode, const volatile int32_t *cancel, double *x,
                    int32_t *iterations, int32_t *stop, double *stats, double *rnorm_history, double *seconds,
                    char *reason, int reason_len)
{
    if (!a || k < 1 || k > 64 || !b || !damp || max_iterations < 0 || !x || !iterations || !stop || !stats ||
        (mode != TS_LSQR_PLAIN_FP64 && mode != TS_LSQR_FAST_FP32 && mode != TS_LSQR_FAST_MIXED &&
         mode != TS_LSQR_FAST_FP64)) {
        set_reason(reason, reason_len, "invalid argument");
        return TS_EINVAL;
    }
    return guarded(reason, reason_len, [&]() -> int {
        TS_HIP(hipSetDevice(a->ctx->device));
        if (a->rows == 0 || a->cols == 0) {
            for (int j = 0; j < k; j++) {
                iterations[j] = 0;
                stop[j] = 8;  // empty system (Lsqr.cs:48)
                stats[3 * j] = stats[3 * j + 1] = stats[3 * j + 2] = 0;
            }
            std::fill(x, x + (size_t)k * a->cols, 0.0);
            return TS_OK;
        }
        SolveOut o{x, iterations, stop, stats, rnorm_history};
        if (mode == TS_LSQR_PLAIN_FP64) lsqr_plain(*a, k, b, damp, max_iterations, atol, btol, conlim, cancel, o);
        else lsqr_fast(*a, k, b, damp, max_iterations, atol, btol, conlim, cancel, o, seconds, mode);
        return TS_OK;
    });
}

}  // namespace ts

// ================================================================================== C ABI

extern "C" {

int ts_matrix_create(ts_context *ctx, int rows, int cols, const int32_t *row_start, const int32_t *column,
                     const double *values, int keep_fp64, ts_matrix **out, char *reason, int reason_len)
{
    if (out) *out = nullptr;
    if (!ctx || !out || rows < 0 || cols < 0 || !row_start || (row_start[rows] > 0 && (!column || !values))) {
        set_reason(reason, reason_len, "invalid argument");
        return TS_EINVAL;
    }
    return guarded(reason, reason_len, [&]() -> int {
        TS_HIP(hipSetDevice(ctx->device));
        const int64_t nnz = row_start[rows];
        if (row_start[0] != 0) fail(TS_EINVAL, "row_start[0] must be 0");
        for (int i = 0; i < rows; i++)
            if (row_start[i + 1] < row_start[i]) fail(TS_EINVAL, "row_start decreases at row %d", i);
        for (int64_t q = 0; q < nnz; q++)
            if ((uint32_t)column[q] >= (uint32_t)cols) fail(TS_EINVAL, "column index %d out of range at %lld", column[q], (long long)q);
        std::unique_ptr<ts_matrix> A(new ts_matrix);
        A->ctx = ctx;
        A->rows = rows;
        A->cols = cols;
        A->nnz = nnz;
        if (nnz >= INT32_MAX) fail(TS_DECLINED, "%lld non-zeros exceed the int32 index range", (long long)nnz);
        auto t0 = std::chrono::steady_clock::now();
        hipStream_t st = ctx->stream;
        A->rs.alloc(rows + 1);
        A->ci.alloc(std::max<int64_t>(1, nnz));
        dbuf<double> va(std::max<int64_t>(1, nnz));
        TS_HIP(hipMemcpyAsync(A->rs.p, row_start, (size_t)(rows + 1) * 4, hipMemcpyHostToDevice, st));
        if (nnz > 0) {
            TS_HIP(hipMemcpyAsync(A->ci.p, column, (size_t)nnz * 4, hipMemcpyHostToDevice, st));
            TS_HIP(hipMemcpyAsync(va.p, values, (size_t)nnz * 8, hipMemcpyHostToDevice, st));
        }
        matrix_finish(*A, std::move(va), keep_fp64 != 0);
        A->build_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        *out = A.release();
        return TS_OK;
    });
}

}  // extern "C"

namespace ts {

// y = A x or A^T x for k vectors through four device vectors the caller provides: hin[nin] (the input
// as given), din[nin] (in the stored order), dout[nout] (in the stored order), hout[nout] (as given).
// Runs on the matrix's stream and waits for each vector's copy back.
static void matrix_multiply_with(ts_matrix *a, int k, const double *x, double *y, int transpose, double *hin,
                                 double *din, double *dout, double *hout)
{
    hipStream_t st = a->ctx->stream;
    const int m = a->rows, n = a->cols;
    const int nin = transpose ? m : n, nout = transpose ? n : m;
    // permutation of the input into the stored order and of the output back to the given one
    const int32_t *pin = transpose ? a->row_perm.p : a->col_perm.p;
    const int32_t *pout = transpose ? a->col_perm.p : a->row_perm.p;
    for (int j = 0; j < k; j++) {
        TS_HIP(hipMemcpyAsync(hin, x + (size_t)j * nin, (size_t)nin * 8, hipMemcpyHostToDevice, st));
        if (pin) mx_permute_vec<<<grid_for(nin), 256, 0, st>>>(hin, pin, nin, din);
        const double *xin = pin ? din : hin;
        double *yout = pout ? dout : hout;
        if (transpose) {
            if (a->tva64.p) mx_spmv<double><<<grid_for(n), 256, 0, st>>>(n, a->trs.p, a->tci.p, a->tva64.p, xin, yout);
            else mx_spmv<float><<<grid_for(n), 256, 0, st>>>(n, a->trs.p, a->tci.p, a->tva32.p, xin, yout);
        } else {
            if (a->va64.p) mx_spmv<double><<<grid_for(m), 256, 0, st>>>(m, a->rs.p, a->ci.p, a->va64.p, xin, yout);
            else mx_spmv<float><<<grid_for(m), 256, 0, st>>>(m, a->rs.p, a->ci.p, a->va32.p, xin, yout);
        }
        if (pout) mx_unpermute_vec<<<grid_for(nout), 256, 0, st>>>(dout, pout, nout, hout);
        TS_HIP(hipGetLastError());
        TS_HIP(hipMemcpyAsync(y + (size_t)j * nout, hout, (size_t)nout * 8, hipMemcpyDeviceToHost, st));
        TS_HIP(hipStreamSynchronize(st));
    }
}

// The no-allocation product's staging memory: one array per GPU, allocated with the code object when
// the library loads and held until the process ends. Host code never passes `mul_scratch` itself to a
// kernel or a copy (in host code the name is only a stand-in); it uses the device address that
// hipGetSymbolAddress returns for the current GPU, kept per device in mul_scratch_ptr.
constexpr size_t kMulScratchDoubles = (size_t)8 << 20;  // 64 MiB
static __device__ __align__(256) double mul_scratch[kMulScratchDoubles];

constexpr int kMaxDevices = 64;
static std::mutex mul_scratch_setup;
static double *mul_scratch_ptr[kMaxDevices];
static std::mutex mul_scratch_use[kMaxDevices];  // calls on one GPU take turns at its array

// The device address of this GPU's array, looked up once per device with that device current.
static double *mul_scratch_on(int device)
{
    if (device < 0 || device >= kMaxDevices) fail(TS_EDEVICE, "device %d beyond the %d the library tracks", device, kMaxDevices);
    std::lock_guard<std::mutex> lk(mul_scratch_setup);
    if (!mul_scratch_ptr[device]) {
        void *ptr = nullptr;
        TS_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(mul_scratch)));
        if (!ptr) fail(TS_EDEVICE, "no device address for the product's scratch array");
        mul_scratch_ptr[device] = (double *)ptr;
    }
    return mul_scratch_ptr[device];
}

}  // namespace ts

extern "C" {

int ts_matrix_multiply(ts_matrix *a, int k, const double *x, double *y, int transpose, char *reason, int reason_len)
{
    if (!a || k < 0 || (k > 0 && (!x || !y))) {
        set_reason(reason, reason_len, "invalid argument");
        return TS_EINVAL;
    }
    return guarded(reason, reason_len, [&]() -> int {
        TS_HIP(hipSetDevice(a->ctx->device));
        std::lock_guard<std::mutex> lk(a->gate);
        const int nin = transpose ? a->rows : a->cols, nout = transpose ? a->cols : a->rows;
        dbuf<double> hin(std::max(1, nin)), din(std::max(1, nin)), dout(std::max(1, nout)), hout(std::max(1, nout));
        matrix_multiply_with(a, k, x, y, transpose, hin.p, din.p, dout.p, hout.p);
        return TS_OK;
    });
}

int ts_matrix_multiply_noalloc(ts_matrix *a, int k, const double *x, double *y, int transpose, char *reason,
                               int reason_len)
{
    if (!a || k < 0 || (k > 0 && (!x || !y))) {
        set_reason(reason, reason_len, "invalid argument");
        return TS_EINVAL;
    }
    const size_t need = 2 * ((size_t)a->rows + (size_t)a->cols);
    if (need > kMulScratchDoubles) {
        char msg[200];
        snprintf(msg, sizeof msg, "the product needs %zu doubles of scratch (2 x (rows + cols)); the limit is %zu (%zu MiB)",
                 need, kMulScratchDoubles, (kMulScratchDoubles * 8) >> 20);
        set_reason(reason, reason_len, msg);
        return TS_EINVAL;
    }
    return guarded(reason, reason_len, [&]() -> int {
        const int dev = a->ctx->device;
        TS_HIP(hipSetDevice(dev));
        double *scratch = mul_scratch_on(dev);
        std::lock_guard<std::mutex> use(mul_scratch_use[dev]);
        std::lock_guard<std::mutex> lk(a->gate);
        const size_t nin = transpose ? a->rows : a->cols, nout = transpose ? a->cols : a->rows;
        double *hin = scratch, *din = hin + nin, *dout = din + nin, *hout = dout + nout;
        matrix_multiply_with(a, k, x, y, transpose, hin, din, dout, hout);
        return TS_OK;
    });
}

int ts_matrix_stats(ts_matrix *a, int64_t *nnz, double *seconds)
{
    if (!a) return TS_EINVAL;
    if (nnz) *nnz = a->nnz;
    if (seconds) *seconds = a->last_seconds;
    return TS_OK;
}

void ts_matrix_release(ts_matrix *a)
{
    if (!a) return;
    (void)hipSetDevice(a->ctx->device);
    delete a;
}

int ts_lsqr_solve(ts_matrix *a, int k, const double *b, const double *damp, int max_iterations, double atol,
                  double btol, double conlim, int mode, const volatile int32_t *cancel, double *x,
                  int32_t *iterations, int32_t *stop, double *stats, double *rnorm_history, char *reason,
                  int reason_len)
{
    if (!a) {
        set_reason(reason, reason_len, "invalid argument");
        return TS_EINVAL;
    }
    std::lock_guard<std::mutex> lk(a->gate);
    double seconds = 0;
    const int rc = ts::lsqr_solve_core(a, k, b, damp, max_iterations, atol, btol, conlim, mode, cancel, x, iterations,
                                       stop, stats, rnorm_history, &seconds, reason, reason_len);
    if (rc == TS_OK) a->last_seconds = seconds;
    return rc;
}

}  // extern "C"
