// A sparse matrix resident in HBM for LSQR: CSR with int32 indices (CsrMatrix's own index type,
// Numerics/SparseMatrix.cs:24-46) plus its stored transpose (the cached CsrMatrix.Transpose,
// SparseMatrix.cs:79-97). Values are kept in FP32 for the fast form and, on request, in FP64 for the
// plain form.
#pragma once

#include <mutex>
#include <vector>

#include "common.hpp"

struct ts_matrix {
    ts_context *ctx = nullptr;
    int rows = 0, cols = 0;
    int64_t nnz = 0;
    ts::dbuf<int32_t> rs, ci, trs, tci;  // A and its transpose
    ts::dbuf<float> va32, tva32;
    ts::dbuf<double> va64, tva64;        // kept only when the plain FP64 form is enabled
    ts::dbuf<int32_t> row_perm;          // stored row r is row row_perm[r] of the matrix as given
    ts::dbuf<int32_t> col_perm;          // stored column c is column col_perm[c] as given (empty: same)
    std::vector<double> rhs;         // right-hand side of an assembled system (empty otherwise)
    double last_seconds = 0;         // device time of the last fast solve's iterations (HIP events)
    double build_seconds = 0;        // wall time of the assembly / upload and transpose
    std::mutex gate;
};

namespace ts {

// Completes a matrix whose rs, ci (device, int32) are set and whose FP64 values sit in d_va64 (device,
// nnz): the stored column order (stored_col[c], host, when given: the stored position of column c),
// the stored row order, the FP32 values, and the transpose built on the device - a stable radix sort
// of the entries by column, so each transposed row lists its rows in increasing order, as the counting
// sort of SparseMatrix.cs:79-97 does. d_va64 is adopted as va64 when keep_fp64, released otherwise.
void matrix_finish(ts_matrix &A, dbuf<double> &&d_va64, bool keep_fp64, const int32_t *stored_col = nullptr);

// ts_lsqr_solve without the matrix's gate, on the calling thread's stream.
int lsqr_solve_core(ts_matrix *a, int k, const double *b, const double *damp, int max_iterations, double atol,
                    double btol, double conlim, int mode, const volatile int32_t *cancel, double *x,
                    int32_t *iterations, int32_t *stop, double *stats, double *rnorm_history, double *seconds,
                    char *reason, int reason_len);

}  // namespace ts
