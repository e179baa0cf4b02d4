// The tomographic linear system assembled in HBM: TravelTimeTomography.BuildSystem
// (Tomography/TravelTimeTomography.cs:860-1034) for a plain node grid (no parameter lattice, no
// adaptive mesh, not layered), from ray rows already resident on the device, straight into the
// LSQR matrix (CSR, int32 indices) and its stored transpose.
//
// Column layout (:869-883): [P nodes | S nodes | 4 per event (T0, E, N, Z) | station P | station S].
// Rows, in the reference's order:
//   data rows (:885-973): one per arrival that has a ray, velocity terms (Terms), hypocentre terms,
//     the station term; entries below 1e-4 in magnitude dropped (Packed), empty rows skipped, the
//     row and its residual scaled by absoluteWeight / sigma (CsrBuilder.AddRow, SparseMatrix.cs:128-140);
//   regularisation (:976-1017): damping and smoothing of the P block, of the S block (times the S
//     factor), damping of the hypocentre columns, and per station block its damping plus the
//     zero-sum row; smoothing is the Laplacian (Observations.cs:410-443) or first differences
//     (Observations.cs:360-384), with AxisWeights (Observations.cs:390-397).
//
// Data rows need no sort: a row's nodes are ascending and the blocks are laid out in column order,
// so the entries land in the order Packed sorts them to. Entries are counted in one pass and written
// in a second (a wavefront per data row, a thread per regularisation row), between them one scan over
// the row counts gives the row starts. Every value is the reference's product in the reference's
// order, in FP64 (correctly rounded division; no contraction), so the assembled matrix is the
// reference's to the last bit.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <memory>
#include <vector>

#include "common.hpp"
#include "geometry.hpp"
#include "matrix.hpp"
#include "rays.hpp"

using namespace ts;

namespace ts {
namespace {

struct DataArgs {
    const int32_t *nodes;      // ray CSR (arrival order)
    const float *lengths;
    const int64_t *ray_start;  // [nrays + 1]
    const int32_t *ray_of_row, *phase, *event, *station;
    const double *residual, *sigma, *grad;  // grad: Ge, Gn, Gz per row
    const uint8_t *ev_fixed, *ev_outside;
    const double *sP, *sRefP, *sRefS, *kRef, *kCur;
    int nrows;
    int offP, offS, offH, offCP, offCS;
    int invert_p, vpvs;
    double absolute_weight;
};

__device__ inline bool kept(double v) { return fabs(v) >= 1e-4; }  // Packed (:921-926)

// the velocity entry of one node: block 0 is the P block, 1 the S block; false if the row has no
// entry of that block
__device__ inline bool node_value(const DataArgs &A, int ph, int block, int nd, float len, double &v)
{
    const double L = (double)len;
    if (ph == 0) {
        if (block != 0 || A.offP < 0 || !A.invert_p) return false;
        v = L * A.sRefP[nd];
        return true;
    }
    if (A.offS < 0) return false;
    if (A.vpvs) {
        if (block == 0) {
            if (A.offP < 0 || !A.invert_p) return false;
            v = L * A.kCur[nd] * A.sRefP[nd];
        } else {
            v = L * A.sP[nd] * A.kRef[nd];
        }
        return true;
    }
    if (block != 1) return false;
    v = L * A.sRefS[nd];
    return true;
}

// hypocentre and station entries of a row, in column order: (column, value) pairs; returns the count
__device__ inline int tail_entries(const DataArgs &A, int r, int cols[5], double vals[5])
{
    int c = 0;
    const int e = A.event[r];
    if (A.offH >= 0 && !A.ev_fixed[e]) {
        const int h = A.offH + 4 * e;
        cols[c] = h; vals[c] = 1.0; c++;
        if (!A.ev_outside[e]) {
            cols[c] = h + 1; vals[c] = A.grad[3 * r]; c++;
            cols[c] = h + 2; vals[c] = A.grad[3 * r + 1]; c++;
            cols[c] = h + 3; vals[c] = A.grad[3 * r + 2]; c++;
        }
    }
    const int coff = A.phase[r] == 0 ? A.offCP : A.offCS;
    if (coff >= 0) { cols[c] = coff + A.station[r]; vals[c] = 1.0; c++; }
    return c;
}

// pass 1: kept entries per data row (one wavefront per row)
__global__ __launch_bounds__(256) void sys_data_count(DataArgs A, int32_t *__restrict__ count)
{
    const int lane = threadIdx.x & 63;
    const int r = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (r >= A.nrows) return;
    const int ray = A.ray_of_row[r], ph = A.phase[r];
    const int64_t b = A.ray_start[ray], e = A.ray_start[ray + 1];
    int c = 0;
    for (int block = 0; block < 2; block++)
        for (int64_t q = b + lane; q < e; q += 64) {
            double v;
            if (node_value(A, ph, block, A.nodes[q], A.lengths[q], v) && kept(v)) c++;
        }
    for (int off = 32; off > 0; off >>= 1) c += __shfl_xor(c, off);
    if (lane == 0) {
        int cols[5];
        double vals[5];
        const int t = tail_entries(A, r, cols, vals);
        for (int q = 0; q < t; q++) c += kept(vals[q]);
        count[r] = c;
    }
}

// pass 2: the entries of each non-empty data row at its row start, times absoluteWeight / sigma
__global__ __launch_bounds__(256) void sys_data_fill(DataArgs A, const int32_t *__restrict__ count,
                                                     const int32_t *__restrict__ dest, const int64_t *__restrict__ rs,
                                                     int32_t *__restrict__ ci, double *__restrict__ va,
                                                     double *__restrict__ rhs)
{
    const int lane = threadIdx.x & 63;
    const int r = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (r >= A.nrows || count[r] == 0) return;
    const int row = dest[r];
    const double w = A.absolute_weight / A.sigma[r];
    const int ray = A.ray_of_row[r], ph = A.phase[r];
    const int64_t b = A.ray_start[ray], e = A.ray_start[ray + 1];
    int64_t at = rs[row];
    const int off[2] = {A.offP, A.offS};
    for (int block = 0; block < 2; block++)
        for (int64_t q0 = b; q0 < e; q0 += 64) {
            const int64_t q = q0 + lane;
            double v = 0;
            int nd = 0;
            bool k = false;
            if (q < e) {
                nd = A.nodes[q];
                k = node_value(A, ph, block, nd, A.lengths[q], v) && kept(v);
            }
            const uint64_t mask = __ballot(k);
            if (k) {
                const int pos = __popcll(mask & ((1ull << lane) - 1));
                ci[at + pos] = off[block] + nd;
                va[at + pos] = v * w;
            }
            at += __popcll(mask);
        }
    if (lane == 0) {
        int cols[5];
        double vals[5];
        const int t = tail_entries(A, r, cols, vals);
        for (int q = 0; q < t; q++)
            if (kept(vals[q])) { ci[at] = cols[q]; va[at] = vals[q] * w; at++; }
        rhs[row] = A.residual[r] * w;
    }
}

__global__ void sys_scatter_counts(const int32_t *__restrict__ count, const int32_t *__restrict__ dest, int nrows,
                                   int64_t *__restrict__ rowcnt)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < nrows && count[r] > 0) rowcnt[dest[r]] = count[r];
}

__global__ void sys_flags(const int32_t *__restrict__ count, int n, int32_t *__restrict__ flag)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < n) flag[r] = count[r] > 0;
}

// kappa = sS / sP and the deviations from the reference model the smoothing rows act on (:981-988)
__global__ void sys_models(const double *__restrict__ sP, const double *__restrict__ sS, const double *__restrict__ sRefP,
                           const double *__restrict__ sRefS, const double *__restrict__ kRef, int n, int vpvs,
                           double *__restrict__ kCur, double *__restrict__ devP, double *__restrict__ devS)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const double k = sS[q] / sP[q];
    kCur[q] = k;
    devP[q] = sP[q] / sRefP[q] - 1;
    devS[q] = vpvs ? k / kRef[q] - 1 : sS[q] / sRefS[q] - 1;
}

// ---- regularisation blocks
enum BlockKind { kDamp = 0, kLaplacian = 1, kGradient = 2, kZeroSum = 3 };

struct Block {
    int kind;
    int offset;        // first column
    int count;         // columns damped (kDamp, kZeroSum) or nodes (smoothing)
    int64_t first_row; // absolute row of the block's first row
    int64_t rows;
    double lambda;     // AddRow weight
    const double *current;  // deviation the smoothing rows' right-hand side is formed from
};

struct RegArgs {
    int nx, ny, nz;
    double w[3];   // AxisWeights
    double a[3];   // their square roots (first differences)
};

__device__ inline bool in_range(const RegArgs &G, int i, int j, int k)
{
    return (unsigned)i < (unsigned)G.nx && (unsigned)j < (unsigned)G.ny && (unsigned)k < (unsigned)G.nz;
}

// Laplacian row of node (i, j, k) in AddLaplacian's column order: i-1, i+1, j-1, j+1, k-1, k+1, self
__device__ inline int laplacian_row(const RegArgs &G, int i, int j, int k, int cols[7], double vals[7])
{
    const int d[6][3] = {{-1, 0, 0}, {1, 0, 0}, {0, -1, 0}, {0, 1, 0}, {0, 0, -1}, {0, 0, 1}};
    int c = 0;
    double diag = 0;
    for (int q = 0; q < 6; q++) {
        const int ii = i + d[q][0], jj = j + d[q][1], kk = k + d[q][2];
        if (!in_range(G, ii, jj, kk)) continue;
        const double w = G.w[q >> 1];
        cols[c] = (kk * G.ny + jj) * G.nx + ii;
        vals[c] = -w;
        c++;
        diag += w;
    }
    cols[c] = (k * G.ny + j) * G.nx + i;
    vals[c] = diag;
    return c + 1;
}

// first-difference rows before node (i, j, k) in node-major, axis-minor order (AddGradient's loop)
__device__ inline int64_t gradient_rows_before(const RegArgs &G, int i, int j, int k)
{
    const int64_t nx = G.nx, ny = G.ny, nz = G.nz;
    int64_t r = 0;
    if (G.w[0] > 0) r += k * ny * (nx - 1) + j * (nx - 1) + min((int64_t)i, nx - 1);
    if (G.w[1] > 0) r += k * (ny - 1) * nx + min((int64_t)j, ny - 1) * nx + (j < ny - 1 ? i : 0);
    if (G.w[2] > 0) r += min((int64_t)k, nz - 1) * ny * nx + (k < nz - 1 ? j * nx + i : 0);
    return r;
}

__global__ void sys_reg_count(Block B, RegArgs G, int64_t *__restrict__ rowcnt)
{
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    switch (B.kind) {
    case kDamp:
        if (t < B.rows) rowcnt[B.first_row + t] = 1;
        break;
    case kZeroSum:
        if (t == 0) rowcnt[B.first_row] = B.count;
        break;
    case kLaplacian: {
        if (t >= B.count) return;
        const int n = (int)t, i = n % G.nx, j = (n / G.nx) % G.ny, k = n / (G.nx * G.ny);
        int cols[7];
        double vals[7];
        const int c = laplacian_row(G, i, j, k, cols, vals);
        int kept = 0;
        for (int q = 0; q < c; q++) kept += vals[q] != 0;
        rowcnt[B.first_row + n] = kept;
        break;
    }
    case kGradient:
        if (t < B.rows) rowcnt[B.first_row + t] = 2;
        break;
    }
}

__global__ void sys_reg_fill(Block B, RegArgs G, const int64_t *__restrict__ rs, int32_t *__restrict__ ci,
                             double *__restrict__ va, double *__restrict__ rhs)
{
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    switch (B.kind) {
    case kDamp:  // AddDamping (Observations.cs:474-486)
        if (t < B.rows) {
            const int64_t row = B.first_row + t;
            ci[rs[row]] = B.offset + (int)t;
            va[rs[row]] = 1.0 * B.lambda;
            rhs[row] = 0.0 * B.lambda;
        }
        break;
    case kZeroSum:  // sum of the station terms = 0, weight 10 (:1010-1016)
        for (int64_t s = t; s < B.count; s += (int64_t)gridDim.x * blockDim.x) {
            ci[rs[B.first_row] + s] = B.offset + (int)s;
            va[rs[B.first_row] + s] = 1.0 * B.lambda;
        }
        if (t == 0) rhs[B.first_row] = 0.0 * B.lambda;
        break;
    case kLaplacian: {  // AddLaplacian (Observations.cs:410-443)
        if (t >= B.count) return;
        const int n = (int)t, i = n % G.nx, j = (n / G.nx) % G.ny, k = n / (G.nx * G.ny);
        int cols[7];
        double vals[7];
        const int c = laplacian_row(G, i, j, k, cols, vals);
        const int64_t row = B.first_row + n;
        int64_t at = rs[row];
        double r = 0;
        for (int q = 0; q < c; q++) {
            if (B.current) r -= vals[q] * B.current[cols[q]];
            if (vals[q] == 0) continue;
            ci[at] = B.offset + cols[q];
            va[at] = vals[q] * B.lambda;
            at++;
        }
        rhs[row] = r * B.lambda;
        break;
    }
    case kGradient: {  // AddGradient (Observations.cs:360-384)
        if (t >= B.count) return;
        const int n = (int)t, i = n % G.nx, j = (n / G.nx) % G.ny, k = n / (G.nx * G.ny);
        int64_t row = B.first_row + gradient_rows_before(G, i, j, k);
        const int d[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
        for (int q = 0; q < 3; q++) {
            const int ii = i + d[q][0], jj = j + d[q][1], kk = k + d[q][2];
            if (!in_range(G, ii, jj, kk) || G.w[q] <= 0) continue;
            const int m = (kk * G.ny + jj) * G.nx + ii;
            const double a = G.a[q];
            const int64_t at = rs[row];
            ci[at] = B.offset + n;
            va[at] = -a * B.lambda;
            ci[at + 1] = B.offset + m;
            va[at + 1] = a * B.lambda;
            rhs[row] = (B.current ? -a * (B.current[m] - B.current[n]) : 0.0) * B.lambda;
            row++;
        }
        break;
    }
    }
}

__global__ void sys_rs32(const int64_t *__restrict__ rs, int64_t n, int32_t *__restrict__ out)
{
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x)
        out[q] = (int32_t)rs[q];
}

// ---- Q tomography's t* rows (QTomography.SolveOnce, Attenuation/QTomography.cs:318-364)

struct QArgs {
    const int32_t *nodes;      // ray CSR (arrival order)
    const float *lengths;
    const int64_t *ray_start;  // [nrays + 1]
    const int32_t *ray_of_row;
    const double *slow;        // slowness of the phase inverted
    const double *q;           // q = 1/Q the rows are formed at; null: q0 at every node
    double q0;
    int nrows;
    const double *outside;     // per arrival: RayRow.OutsideTime; null: no part outside the grid
};

// dt*/dq of one node: (L s) q, the kernel at q (:337; Predict's term, :310)
__device__ inline double q_entry(const QArgs &A, int64_t e, int nd)
{
    return (double)A.lengths[e] * A.slow[nd] * (A.q ? A.q[nd] : A.q0);
}

// One thread per row, its entries in order: the predicted t* (Predict, :307-312: q0 times the time
// outside the grid, none here, plus sum L s q), the path time sum L s (:232-233), the entries AddRow
// keeps (not exactly zero), the row's sum of squared weighted entries (v / sigma)^2 for the model
// scale (:356-361), and the data right-hand side t* - Predict + sum (L s q) m (:344-349; without m,
// the start residual). Every output may be null.
__global__ __launch_bounds__(256) void q_rows(QArgs A, const double *__restrict__ tstar, const double *__restrict__ m,
                                              const double *__restrict__ sigma, double *__restrict__ pred,
                                              double *__restrict__ tt, int32_t *__restrict__ count,
                                              double *__restrict__ wsum, double *__restrict__ rhs)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= A.nrows) return;
    const int ray = A.ray_of_row[r];
    const int64_t b = A.ray_start[ray], e = A.ray_start[ray + 1];
    const double sg = sigma ? sigma[r] : 1.0;
    // Predict starts from q0 * OutsideTime, the Q0 estimate's path time from OutsideTime (:231, :309)
    const double out = A.outside ? A.outside[ray] : 0.0;
    double p = A.q0 * out, t = out, s2 = 0.0;
    int c = 0;
    for (int64_t k = b; k < e; k++) {
        const int nd = A.nodes[k];
        const double ls = (double)A.lengths[k] * A.slow[nd];
        const double v = ls * (A.q ? A.q[nd] : A.q0);
        p += v;
        t += ls;
        c += v != 0;
        const double w = v / sg;
        s2 += w * w;
    }
    if (pred) pred[r] = p;
    if (tt) tt[r] = t;
    if (count) count[r] = c;
    if (wsum) wsum[r] = s2;
    if (rhs) {
        double d = tstar[r] - p;
        if (m)
            for (int64_t k = b; k < e; k++) {
                const int nd = A.nodes[k];
                d += q_entry(A, k, nd) * m[nd];
            }
        rhs[r] = d;
    }
}

// entries per data row: the kept node entries, plus the station term's
__global__ void q_counts(const int32_t *__restrict__ count, int nrows, int station_terms, int64_t *__restrict__ rowcnt)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < nrows) rowcnt[r] = count[r] + station_terms;
}

// The entries of each data row at its row start (one wavefront per row): the node entries in node
// order without exact zeros, the station term 1 (:351), all times 1 / sigma, and the right-hand side
// times 1 / sigma (CsrBuilder.AddRow, SparseMatrix.cs:128-140).
__global__ __launch_bounds__(256) void q_fill(QArgs A, const int32_t *__restrict__ station,
                                              const double *__restrict__ sigma, const double *__restrict__ rhs_in,
                                              int off_st, const int64_t *__restrict__ rs, int32_t *__restrict__ ci,
                                              double *__restrict__ va, double *__restrict__ rhs)
{
    const int lane = threadIdx.x & 63;
    const int r = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (r >= A.nrows) return;
    const double w = 1 / sigma[r];
    const int ray = A.ray_of_row[r];
    const int64_t b = A.ray_start[ray], e = A.ray_start[ray + 1];
    int64_t at = rs[r];
    for (int64_t q0 = b; q0 < e; q0 += 64) {
        const int64_t q = q0 + lane;
        double v = 0;
        int nd = 0;
        if (q < e) {
            nd = A.nodes[q];
            v = q_entry(A, q, nd);
        }
        const bool k = q < e && v != 0;
        const uint64_t mask = __ballot(k);
        if (k) {
            const int pos = __popcll(mask & ((1ull << lane) - 1));
            ci[at + pos] = nd;
            va[at + pos] = v * w;
        }
        at += __popcll(mask);
    }
    if (lane == 0) {
        if (off_st >= 0) {
            ci[at] = off_st + station[r];
            va[at] = 1.0 * w;
        }
        rhs[r] = rhs_in[r] * w;
    }
}

unsigned blocks(int64_t n, int per) { return (unsigned)std::max<int64_t>(1, (n + per - 1) / per); }

// AxisWeights (Observations.cs:390-397): spacings at (Nx/2, Ny/2, 0) (SphericalGrid.Spacing, :152-156)
void axis_weights(const Grid &g, double vertical, bool in_nodes, double w[3])
{
    if (in_nodes) { w[0] = 1; w[1] = 1; w[2] = vertical; return; }
    const double r = kEarthRadiusKm - g.min_depth;
    const double lat = g.min_lat + (g.ny / 2) * g.dlat;
    const double hx = std::max(1e-9, r * cr::cos(lat * kDeg2Rad) * (g.dlon * kDeg2Rad));
    const double hy = r * (g.dlat * kDeg2Rad), hz = g.ddepth;
    const double h = std::sqrt(hx * hy);
    w[0] = std::pow(h / hx, 2);
    w[1] = std::pow(h / hy, 2);
    w[2] = vertical * std::pow(h / hz, 2);
}

// Stored column order of an assembled system (the stored position of each column): the nodes of each
// model block in tiles of 8 x 8 x 8 (TOMOSTAR_COL_TILE="XxYxZ" picks another shape, "0" keeps the
// given order), tiles in grid order, nodes in grid order within a tile; other columns stay where they
// are. Neighbouring cells then sit in neighbouring columns, so the rows of A^T that a stretch of the
// grid's rays cross are products of neighbouring blocks and gather the same rows of the rays' vectors.
std::vector<int32_t> stored_columns(const Grid &g, int cols, std::initializer_list<int> blocks_at)
{
    int tx = 8, ty = 8, tz = 8;
    if (const char *e = getenv("TOMOSTAR_COL_TILE")) {
        if (std::sscanf(e, "%dx%dx%d", &tx, &ty, &tz) != 3 || tx < 1 || ty < 1 || tz < 1) return {};
    }
    const int64_t n = (int64_t)g.nx * g.ny * g.nz;
    std::vector<int32_t> pos(cols);
    for (int c = 0; c < cols; c++) pos[c] = c;
    for (const int off : blocks_at) {
        if (off < 0 || off + n > cols) continue;
        int32_t at = off;
        for (int k0 = 0; k0 < g.nz; k0 += tz)
            for (int j0 = 0; j0 < g.ny; j0 += ty)
                for (int i0 = 0; i0 < g.nx; i0 += tx)
                    for (int k = k0; k < std::min(g.nz, k0 + tz); k++)
                        for (int j = j0; j < std::min(g.ny, j0 + ty); j++)
                            for (int i = i0; i < std::min(g.nx, i0 + tx); i++)
                                pos[off + ((int64_t)k * g.ny + j) * g.nx + i] = at++;
    }
    return pos;
}

template <typename T>
void upload(dbuf<T> &d, const T *h, size_t n, hipStream_t st)
{
    d.alloc(std::max<size_t>(1, n));
    if (n) TS_HIP(hipMemcpyAsync(d.p, h, n * sizeof(T), hipMemcpyHostToDevice, st));
}

// Completes a system whose data rows come first: their row counts (count_data writes rows
// [0, data rows) of the count array), the blocks' counts, one scan to row starts, the data rows'
// entries (fill_data) and the blocks', int32 row starts, the right-hand side to the host, and the
// stored orders and transpose (matrix_finish), the nodes of the blocks starting at node_blocks stored
// in tiles.
template <typename CountData, typename FillData>
std::unique_ptr<ts_matrix> assemble(ts_context *ctx, hipStream_t st, const Grid &g, const RegArgs &G,
                                    const std::vector<Block> &regs, int64_t m, int64_t col, bool keep_fp64,
                                    std::initializer_list<int> node_blocks, CountData count_data, FillData fill_data)
{
    if (m >= INT32_MAX) fail(TS_DECLINED, "%lld rows exceed the int32 index range", (long long)m);
    dbuf<int64_t> rowcnt(m + 1), rs64(m + 1);
    TS_HIP(hipMemsetAsync(rowcnt.p, 0, (m + 1) * 8, st));
    count_data(rowcnt.p);
    for (const Block &B : regs) {
        const int64_t threads = B.kind == kDamp || B.kind == kGradient ? B.rows : B.kind == kZeroSum ? 1 : B.count;
        sys_reg_count<<<blocks(threads, 256), 256, 0, st>>>(B, G, rowcnt.p);
    }
    TS_HIP(hipGetLastError());
    {
        size_t tb = 0;
        TS_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tb, rowcnt.p, rs64.p, m + 1, st));
        dbuf<unsigned char> tmp(std::max<size_t>(1, tb));
        TS_HIP(hipcub::DeviceScan::ExclusiveSum(tmp.p, tb, rowcnt.p, rs64.p, m + 1, st));
    }
    int64_t nnz = 0;
    TS_HIP(hipMemcpyAsync(&nnz, rs64.p + m, 8, hipMemcpyDeviceToHost, st));
    TS_HIP(hipStreamSynchronize(st));
    if (nnz >= INT32_MAX) fail(TS_DECLINED, "%lld non-zeros exceed the int32 index range", (long long)nnz);

    std::unique_ptr<ts_matrix> M(new ts_matrix);
    M->ctx = ctx;
    M->rows = (int)m;
    M->cols = (int)col;
    M->nnz = nnz;
    M->rs.alloc(m + 1);
    M->ci.alloc(std::max<int64_t>(1, nnz));
    dbuf<double> va(std::max<int64_t>(1, nnz)), rhs(std::max<int64_t>(1, m));
    fill_data(rs64.p, M->ci.p, va.p, rhs.p);
    for (const Block &B : regs) {
        const int64_t threads = B.kind == kDamp ? B.rows : B.kind == kZeroSum ? std::max(1, B.count) : B.count;
        sys_reg_fill<<<blocks(threads, 256), 256, 0, st>>>(B, G, rs64.p, M->ci.p, va.p, rhs.p);
    }
    sys_rs32<<<blocks(std::min<int64_t>(m + 1, 1 << 22), 256), 256, 0, st>>>(rs64.p, m + 1, M->rs.p);
    TS_HIP(hipGetLastError());
    M->rhs.resize(m);
    if (m > 0) TS_HIP(hipMemcpyAsync(M->rhs.data(), rhs.p, m * 8, hipMemcpyDeviceToHost, st));
    TS_HIP(hipStreamSynchronize(st));
    const std::vector<int32_t> order = stored_columns(g, (int)col, node_blocks);
    matrix_finish(*M, std::move(va), keep_fp64, order.empty() ? nullptr : order.data());
    return M;
}

}  // namespace
}  // namespace ts

// ================================================================================== C ABI

extern "C" {

int ts_system_build(ts_context *ctx, const ts_grid *grid, const ts_system_opts *o, ts_rays *rays, int nrows,
                    const int32_t *ray_of_row, const int32_t *phase, const int32_t *event, const int32_t *station,
                    const double *residual, const double *sigma, const uint8_t *event_fixed,
                    const uint8_t *event_outside, const double *sP, const double *sS, const double *sRefP,
                    const double *sRefS, const double *kRef, int keep_fp64, ts_matrix **out, int64_t layout[8],
                    char *reason, int reason_len)
{
    if (out) *out = nullptr;
    if (!ctx || !grid || !o || !rays || nrows < 0 || !out || !sP || !sS || !sRefP || !sRefS || !kRef ||
        (nrows > 0 && (!ray_of_row || !phase || !event || !station || !residual || !sigma)) ||
        o->nevents < 0 || o->nstations < 0 || (o->nevents > 0 && (!event_fixed || !event_outside)) || !grid_valid(*grid)) {
        set_reason(reason, reason_len, "invalid argument");
        return TS_EINVAL;
    }
    return guarded(reason, reason_len, [&]() -> int {
        TS_HIP(hipSetDevice(ctx->device));
        hipStream_t st = ctx->stream;
        auto t0 = std::chrono::steady_clock::now();
        const Grid g = make_grid(*grid);
        const int n = g.nx * g.ny * g.nz;
        if (rays->ctx != ctx) fail(TS_EINVAL, "the rays belong to another context");
        for (int r = 0; r < nrows; r++) {
            if (ray_of_row[r] < 0 || ray_of_row[r] >= rays->nrays || rays->counts[ray_of_row[r]] < 0)
                fail(TS_EINVAL, "row %d names arrival %d, which has no ray", r, ray_of_row[r]);
            if (phase[r] != 0 && phase[r] != 1) fail(TS_EINVAL, "row %d: phase %d", r, phase[r]);
            if (event[r] < 0 || event[r] >= o->nevents) fail(TS_EINVAL, "row %d: event %d of %d", r, event[r], o->nevents);
            if (o->station_corrections && (station[r] < 0 || station[r] >= o->nstations))
                fail(TS_EINVAL, "row %d: station %d of %d", r, station[r], o->nstations);
        }
        if (rays->inversion.nx != g.nx || rays->inversion.ny != g.ny || rays->inversion.nz != g.nz)
            fail(TS_EINVAL, "the rays were traced on another inversion grid");

        // column layout (:863-883)
        const bool has_p = o->invert_p && o->any_p, has_s = o->invert_s && o->any_s, vpvs = o->vpvs != 0;
        int64_t col = 0;
        int offP = -1, offS = -1, offH = -1, offCP = -1, offCS = -1;
        if (has_p || (has_s && vpvs)) { offP = (int)col; col += n; }
        if (has_s) { offS = (int)col; col += n; }
        if (o->joint_hypocentres) { offH = (int)col; col += 4 * (int64_t)o->nevents; }
        if (o->station_corrections) {
            if (o->any_p) { offCP = (int)col; col += o->nstations; }
            if (o->any_s) { offCS = (int)col; col += o->nstations; }
        }
        if (col >= INT32_MAX) fail(TS_DECLINED, "%lld columns exceed the int32 index range", (long long)col);

        // inputs to the device
        dbuf<int64_t> d_rstart;
        upload(d_rstart, rays->row_start.data(), rays->row_start.size(), st);
        std::vector<double> grad(3 * (size_t)nrows);
        for (int r = 0; r < nrows; r++)
            for (int q = 0; q < 3; q++) grad[3 * (size_t)r + q] = rays->tcalc_grad[4 * (size_t)ray_of_row[r] + 1 + q];
        dbuf<int32_t> d_ray, d_phase, d_event, d_station;
        dbuf<double> d_res, d_sig, d_grad;
        dbuf<uint8_t> d_fixed, d_outside;
        upload(d_ray, ray_of_row, nrows, st);
        upload(d_phase, phase, nrows, st);
        upload(d_event, event, nrows, st);
        std::vector<int32_t> zeros;
        if (!o->station_corrections) zeros.assign(nrows, 0);
        upload(d_station, o->station_corrections ? station : zeros.data(), nrows, st);
        upload(d_res, residual, nrows, st);
        upload(d_sig, sigma, nrows, st);
        upload(d_grad, grad.data(), grad.size(), st);
        upload(d_fixed, event_fixed, o->nevents, st);
        upload(d_outside, event_outside, o->nevents, st);
        dbuf<double> d_sP, d_sS, d_sRefP, d_sRefS, d_kRef, d_kCur(n), d_devP(n), d_devS(n);
        upload(d_sP, sP, n, st);
        upload(d_sS, sS, n, st);
        upload(d_sRefP, sRefP, n, st);
        upload(d_sRefS, sRefS, n, st);
        upload(d_kRef, kRef, n, st);
        sys_models<<<blocks(n, 256), 256, 0, st>>>(d_sP.p, d_sS.p, d_sRefP.p, d_sRefS.p, d_kRef.p, n, vpvs, d_kCur.p,
                                                   d_devP.p, d_devS.p);

        DataArgs A{rays->nodes.p, rays->lengths.p, d_rstart.p, d_ray.p, d_phase.p, d_event.p, d_station.p, d_res.p,
                   d_sig.p, d_grad.p, d_fixed.p, d_outside.p, d_sP.p, d_sRefP.p, d_sRefS.p, d_kRef.p, d_kCur.p, nrows,
                   offP, offS, offH, offCP, offCS, o->invert_p, vpvs ? 1 : 0, o->absolute_weight};
        const bool data = o->absolute_weight > 0 && nrows > 0;  // (:954)
        dbuf<int32_t> d_count(std::max(1, nrows)), d_flag(std::max(1, nrows) + 1), d_dest(std::max(1, nrows) + 1);
        int64_t data_rows = 0;
        if (data) {
            sys_data_count<<<blocks(nrows, 4), 256, 0, st>>>(A, d_count.p);
            sys_flags<<<blocks(nrows, 256), 256, 0, st>>>(d_count.p, nrows, d_flag.p);
            TS_HIP(hipGetLastError());
            size_t tb = 0;
            TS_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tb, d_flag.p, d_dest.p, nrows, st));
            dbuf<unsigned char> tmp(std::max<size_t>(1, tb));
            TS_HIP(hipcub::DeviceScan::ExclusiveSum(tmp.p, tb, d_flag.p, d_dest.p, nrows, st));
            int32_t last[2];
            TS_HIP(hipMemcpyAsync(&last[0], d_dest.p + nrows - 1, 4, hipMemcpyDeviceToHost, st));
            TS_HIP(hipMemcpyAsync(&last[1], d_flag.p + nrows - 1, 4, hipMemcpyDeviceToHost, st));
            TS_HIP(hipStreamSynchronize(st));
            data_rows = (int64_t)last[0] + last[1];
        }

        // regularisation blocks (:976-1017), plain node grid: regScale = 1
        double aw[3];
        axis_weights(g, o->vertical_weight, o->smoothing_in_nodes != 0, aw);
        RegArgs G{g.nx, g.ny, g.nz, {aw[0], aw[1], aw[2]}, {std::sqrt(aw[0]), std::sqrt(aw[1]), std::sqrt(aw[2])}};
        std::vector<Block> regs;
        int64_t m = data_rows;
        auto damp = [&](int off, int count, double lambda) {
            if (lambda <= 0 || count <= 0) return;
            regs.push_back(Block{kDamp, off, count, m, count, lambda, nullptr});
            m += count;
        };
        auto smooth = [&](int off, const double *current, double lambda) {
            if (lambda <= 0) return;
            if (o->gradient_smoothing) {
                int64_t rows = 0;
                if (aw[0] > 0) rows += (int64_t)(g.nx - 1) * g.ny * g.nz;
                if (aw[1] > 0) rows += (int64_t)g.nx * (g.ny - 1) * g.nz;
                if (aw[2] > 0) rows += (int64_t)g.nx * g.ny * (g.nz - 1);
                regs.push_back(Block{kGradient, off, n, m, rows, lambda, current});
                m += rows;
            } else {
                regs.push_back(Block{kLaplacian, off, n, m, n, lambda, current});
                m += n;
            }
        };
        if (offP >= 0) {
            damp(offP, n, o->damping * 1);
            smooth(offP, d_devP.p, o->smoothing * 1);
        }
        if (offS >= 0) {
            const double f = o->s_factor > 0 ? o->s_factor : 1;
            damp(offS, n, o->damping * f);
            smooth(offS, d_devS.p, o->smoothing * f);
        }
        if (offH >= 0) damp(offH, 4 * o->nevents, o->damping_hypocentre);
        for (int off : {offCP, offCS}) {
            if (off < 0) continue;
            damp(off, o->nstations, o->damping_station);
            regs.push_back(Block{kZeroSum, off, o->nstations, m, 1, 10.0, nullptr});
            m += 1;
        }
        std::unique_ptr<ts_matrix> M = assemble(
            ctx, st, g, G, regs, m, col, keep_fp64 != 0, {(int)offP, (int)offS},
            [&](int64_t *rowcnt) {
                if (data) sys_scatter_counts<<<blocks(nrows, 256), 256, 0, st>>>(d_count.p, d_dest.p, nrows, rowcnt);
            },
            [&](const int64_t *rs64, int32_t *ci, double *va, double *rhs) {
                if (data) sys_data_fill<<<blocks(nrows, 4), 256, 0, st>>>(A, d_count.p, d_dest.p, rs64, ci, va, rhs);
            });
        M->build_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        if (layout) {
            const int64_t L[8] = {offP, offS, offH, offCP, offCS, col, m, data_rows};
            std::copy(L, L + 8, layout);
        }
        *out = M.release();
        return TS_OK;
    });
}

int ts_q_predict(ts_rays *rays, int nrows, const int32_t *ray_of_row, const double *slowness, const double *q,
                 double q0, double *pred, double *travel_time, char *reason, int reason_len)
{
    if (!rays || nrows < 0 || !slowness || (nrows > 0 && !ray_of_row) || !(q0 > 0)) {
        set_reason(reason, reason_len, "invalid argument");
        return TS_EINVAL;
    }
    return guarded(reason, reason_len, [&]() -> int {
        ts_context *ctx = rays->ctx;
        TS_HIP(hipSetDevice(ctx->device));
        hipStream_t st = ctx->stream;
        const Grid g = make_grid(rays->inversion);
        const int n = g.nx * g.ny * g.nz;
        for (int r = 0; r < nrows; r++)
            if (ray_of_row[r] < 0 || ray_of_row[r] >= rays->nrays || rays->counts[ray_of_row[r]] < 0)
                fail(TS_EINVAL, "row %d names arrival %d, which has no ray", r, ray_of_row[r]);
        if (nrows == 0) return TS_OK;
        dbuf<int64_t> d_rstart;
        dbuf<int32_t> d_ray;
        dbuf<double> d_slow, d_q, d_pred(nrows), d_tt(nrows);
        upload(d_rstart, rays->row_start.data(), rays->row_start.size(), st);
        upload(d_ray, ray_of_row, nrows, st);
        upload(d_slow, slowness, n, st);
        if (q) upload(d_q, q, n, st);
        dbuf<double> d_out;
        if (rays->extended && rays->nrays > 0) upload(d_out, rays->outside.data(), rays->outside.size(), st);
        QArgs A{rays->nodes.p, rays->lengths.p, d_rstart.p, d_ray.p, d_slow.p, q ? d_q.p : nullptr, q0, nrows,
                rays->extended && rays->nrays > 0 ? d_out.p : nullptr};
        q_rows<<<blocks(nrows, 256), 256, 0, st>>>(A, nullptr, nullptr, nullptr, d_pred.p, d_tt.p, nullptr, nullptr,
                                                   nullptr);
        TS_HIP(hipGetLastError());
        if (pred) TS_HIP(hipMemcpyAsync(pred, d_pred.p, nrows * 8, hipMemcpyDeviceToHost, st));
        if (travel_time) TS_HIP(hipMemcpyAsync(travel_time, d_tt.p, nrows * 8, hipMemcpyDeviceToHost, st));
        TS_HIP(hipStreamSynchronize(st));
        return TS_OK;
    });
}

int ts_qsystem_build(ts_context *ctx, const ts_grid *grid, const ts_qsystem_opts *o, ts_rays *rays, int nrows,
                     const int32_t *ray_of_row, const int32_t *station, const double *tstar, const double *sigma,
                     const double *slowness, const double *q_old, const double *m_old, int keep_fp64,
                     double scales[2], ts_matrix **out, int64_t layout[4], char *reason, int reason_len)
{
    if (out) *out = nullptr;
    if (!ctx || !grid || !o || !rays || nrows < 0 || !out || !slowness || !scales || !(o->q0 > 0) ||
        (nrows > 0 && (!ray_of_row || !tstar || !sigma || (o->station_terms && !station))) ||
        (!q_old) != (!m_old) || o->nstations < 0 || !grid_valid(*grid)) {
        set_reason(reason, reason_len, "invalid argument");
        return TS_EINVAL;
    }
    return guarded(reason, reason_len, [&]() -> int {
        TS_HIP(hipSetDevice(ctx->device));
        hipStream_t st = ctx->stream;
        auto t0 = std::chrono::steady_clock::now();
        const Grid g = make_grid(*grid);
        const int n = g.nx * g.ny * g.nz;
        if (rays->ctx != ctx) fail(TS_EINVAL, "the rays belong to another context");
        if (rays->inversion.nx != g.nx || rays->inversion.ny != g.ny || rays->inversion.nz != g.nz)
            fail(TS_EINVAL, "the rays were traced on another inversion grid");
        for (int r = 0; r < nrows; r++) {
            if (ray_of_row[r] < 0 || ray_of_row[r] >= rays->nrays || rays->counts[ray_of_row[r]] < 0)
                fail(TS_EINVAL, "row %d names arrival %d, which has no ray", r, ray_of_row[r]);
            if (!(sigma[r] > 0)) fail(TS_EINVAL, "row %d: sigma %g", r, sigma[r]);
            if (o->station_terms && (station[r] < 0 || station[r] >= o->nstations))
                fail(TS_EINVAL, "row %d: station %d of %d", r, station[r], o->nstations);
        }
        // columns (:326-328): the nodes, then one term per station
        const int off_st = o->station_terms ? n : -1;
        const int64_t col = (int64_t)n + (o->station_terms ? o->nstations : 0);
        if (col >= INT32_MAX) fail(TS_DECLINED, "%lld columns exceed the int32 index range", (long long)col);

        const int nr = std::max(1, nrows);
        dbuf<int64_t> d_rstart;
        dbuf<int32_t> d_ray, d_station, d_count(nr);
        dbuf<double> d_slow, d_q, d_m, d_tstar, d_sigma, d_wsum(nr), d_rhs(nr);
        upload(d_rstart, rays->row_start.data(), rays->row_start.size(), st);
        upload(d_ray, ray_of_row, nrows, st);
        std::vector<int32_t> zeros;
        if (!o->station_terms) zeros.assign(nrows, 0);
        upload(d_station, o->station_terms ? station : zeros.data(), nrows, st);
        upload(d_tstar, tstar, nrows, st);
        upload(d_sigma, sigma, nrows, st);
        upload(d_slow, slowness, n, st);
        if (q_old) {
            upload(d_q, q_old, n, st);
            upload(d_m, m_old, n, st);
        }
        dbuf<double> d_out;
        if (rays->extended && rays->nrays > 0) upload(d_out, rays->outside.data(), rays->outside.size(), st);
        QArgs A{rays->nodes.p, rays->lengths.p, d_rstart.p, d_ray.p, d_slow.p, q_old ? d_q.p : nullptr, o->q0, nrows,
                rays->extended && rays->nrays > 0 ? d_out.p : nullptr};
        if (nrows > 0) {
            q_rows<<<blocks(nrows, 256), 256, 0, st>>>(A, d_tstar.p, q_old ? d_m.p : nullptr, d_sigma.p, nullptr,
                                                       nullptr, d_count.p, d_wsum.p, d_rhs.p);
            TS_HIP(hipGetLastError());
        }

        // the sensitivity scales (:356-372): the RMS of the weighted entries of each block. The station
        // block's sum runs over the rows in order, as the reference's; the model block's sums each row's
        // terms in order and then the rows in order.
        std::vector<double> wsum(nrows);
        if (nrows > 0) {
            TS_HIP(hipMemcpyAsync(wsum.data(), d_wsum.p, nrows * 8, hipMemcpyDeviceToHost, st));
            TS_HIP(hipStreamSynchronize(st));
        }
        double sum_model = 0, sum_station = 0;
        int64_t count_model = 0, count_station = 0;
        for (int r = 0; r < nrows; r++) {
            const int ray = ray_of_row[r];
            sum_model += wsum[r];
            count_model += rays->row_start[ray + 1] - rays->row_start[ray];
            if (o->station_terms) {
                const double w = 1.0 / sigma[r];
                sum_station += w * w;
                count_station++;
            }
        }
        double sm = count_model > 0 ? std::sqrt(sum_model / (double)count_model) : 1;
        double ss = count_station > 0 ? std::sqrt(sum_station / (double)count_station) : 1;
        if (o->fixed_scales) { sm = scales[0]; ss = scales[1]; }

        // regularisation (:373-392), plain node grid
        double aw[3];
        axis_weights(g, o->vertical_weight, o->smoothing_in_nodes != 0, aw);
        RegArgs G{g.nx, g.ny, g.nz, {aw[0], aw[1], aw[2]}, {std::sqrt(aw[0]), std::sqrt(aw[1]), std::sqrt(aw[2])}};
        std::vector<Block> regs;
        int64_t m = nrows;
        const double ld = o->damping * sm, ls = o->smoothing * sm;
        if (ld > 0) { regs.push_back(Block{kDamp, 0, n, m, n, ld, nullptr}); m += n; }
        if (ls > 0) {
            if (o->gradient_smoothing) {
                int64_t rows = 0;
                if (aw[0] > 0) rows += (int64_t)(g.nx - 1) * g.ny * g.nz;
                if (aw[1] > 0) rows += (int64_t)g.nx * (g.ny - 1) * g.nz;
                if (aw[2] > 0) rows += (int64_t)g.nx * g.ny * (g.nz - 1);
                regs.push_back(Block{kGradient, 0, n, m, rows, ls, nullptr});
                m += rows;
            } else {
                regs.push_back(Block{kLaplacian, 0, n, m, n, ls, nullptr});
                m += n;
            }
        }
        if (off_st >= 0) {
            const double lst = o->damping_station * ss;
            if (lst > 0 && o->nstations > 0) {
                regs.push_back(Block{kDamp, off_st, o->nstations, m, o->nstations, lst, nullptr});
                m += o->nstations;
            }
            regs.push_back(Block{kZeroSum, off_st, o->nstations, m, 1, 10 * ss, nullptr});
            m += 1;
        }
        std::unique_ptr<ts_matrix> M = assemble(
            ctx, st, g, G, regs, m, col, keep_fp64 != 0, {0},
            [&](int64_t *rowcnt) {
                if (nrows > 0)
                    q_counts<<<blocks(nrows, 256), 256, 0, st>>>(d_count.p, nrows, o->station_terms ? 1 : 0, rowcnt);
            },
            [&](const int64_t *rs64, int32_t *ci, double *va, double *rhs) {
                if (nrows > 0)
                    q_fill<<<blocks(nrows, 4), 256, 0, st>>>(A, d_station.p, d_sigma.p, d_rhs.p, off_st, rs64, ci, va,
                                                            rhs);
            });
        M->build_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        scales[0] = sm;
        scales[1] = ss;
        if (layout) {
            const int64_t L[4] = {off_st, col, m, nrows};
            std::copy(L, L + 4, layout);
        }
        *out = M.release();
        return TS_OK;
    });
}

int ts_matrix_shape(ts_matrix *a, int32_t *rows, int32_t *cols, int64_t *nnz, double *build_seconds)
{
    if (!a) return TS_EINVAL;
    if (rows) *rows = a->rows;
    if (cols) *cols = a->cols;
    if (nnz) *nnz = a->nnz;
    if (build_seconds) *build_seconds = a->build_seconds;
    return TS_OK;
}

int ts_matrix_read(ts_matrix *a, int32_t *row_start, int32_t *column, double *values, double *rhs)
{
    if (!a) return TS_EINVAL;
    return guarded(nullptr, 0, [&]() -> int {
        TS_HIP(hipSetDevice(a->ctx->device));
        std::lock_guard<std::mutex> lk(a->gate);
        const int m = a->rows;
        const int64_t nnz = a->nnz;
        if (rhs) {
            if (a->rhs.size() != (size_t)a->rows) fail(TS_EINVAL, "the matrix was not assembled here: no right-hand side");
            std::copy(a->rhs.begin(), a->rhs.end(), rhs);
        }
        if (!row_start && !column && !values) return TS_OK;
        // the stored arrays, then the rows back in the order they were given
        std::vector<int32_t> rs(m + 1), ci(std::max<int64_t>(1, nnz)), perm(std::max(1, m));
        std::vector<double> va(std::max<int64_t>(1, nnz));
        TS_HIP(hipMemcpy(rs.data(), a->rs.p, ((size_t)m + 1) * 4, hipMemcpyDeviceToHost));
        if (nnz > 0) {
            TS_HIP(hipMemcpy(ci.data(), a->ci.p, nnz * 4, hipMemcpyDeviceToHost));
            if (a->va64.p) TS_HIP(hipMemcpy(va.data(), a->va64.p, nnz * 8, hipMemcpyDeviceToHost));
            else {
                std::vector<float> f(nnz);
                TS_HIP(hipMemcpy(f.data(), a->va32.p, nnz * 4, hipMemcpyDeviceToHost));
                for (int64_t q = 0; q < nnz; q++) va[q] = f[q];
            }
        }
        if (a->row_perm.p) TS_HIP(hipMemcpy(perm.data(), a->row_perm.p, (size_t)m * 4, hipMemcpyDeviceToHost));
        else for (int r = 0; r < m; r++) perm[r] = r;
        if (a->col_perm.p && nnz > 0) {  // stored column indices back to the given ones
            std::vector<int32_t> cp(a->cols);
            TS_HIP(hipMemcpy(cp.data(), a->col_perm.p, (size_t)a->cols * 4, hipMemcpyDeviceToHost));
            for (int64_t q = 0; q < nnz; q++) ci[q] = cp[ci[q]];
        }
        std::vector<int32_t> where(std::max(1, m));  // stored position of each given row
        for (int r = 0; r < m; r++) where[perm[r]] = r;
        int64_t at = 0;
        if (row_start) row_start[0] = 0;
        for (int o = 0; o < m; o++) {
            const int r = where[o];
            for (int32_t q = rs[r]; q < rs[r + 1]; q++, at++) {
                if (column) column[at] = ci[q];
                if (values) values[at] = va[q];
            }
            if (row_start) row_start[o + 1] = (int32_t)at;
        }
        return TS_OK;
    });
}

int ts_matrix_stored_order(ts_matrix *a, int32_t *row_perm, int32_t *col_perm)
{
    if (!a) return TS_EINVAL;
    return guarded(nullptr, 0, [&]() -> int {
        TS_HIP(hipSetDevice(a->ctx->device));
        std::lock_guard<std::mutex> lk(a->gate);
        if (row_perm) {
            if (a->row_perm.p && a->rows > 0)
                TS_HIP(hipMemcpy(row_perm, a->row_perm.p, (size_t)a->rows * 4, hipMemcpyDeviceToHost));
            else
                for (int r = 0; r < a->rows; r++) row_perm[r] = r;
        }
        if (col_perm) {
            if (a->col_perm.p && a->cols > 0)
                TS_HIP(hipMemcpy(col_perm, a->col_perm.p, (size_t)a->cols * 4, hipMemcpyDeviceToHost));
            else
                for (int c = 0; c < a->cols; c++) col_perm[c] = c;
        }
        return TS_OK;
    });
}

int ts_matrix_read_transpose(ts_matrix *a, int32_t *row_start, int32_t *column, float *values)
{
    if (!a) return TS_EINVAL;
    return guarded(nullptr, 0, [&]() -> int {
        TS_HIP(hipSetDevice(a->ctx->device));
        std::lock_guard<std::mutex> lk(a->gate);
        const int n = a->cols, m = a->rows;
        const int64_t nnz = a->nnz;
        std::vector<int32_t> trs(n + 1), tci(std::max<int64_t>(1, nnz)), perm(std::max(1, m));
        std::vector<float> tva(std::max<int64_t>(1, nnz));
        TS_HIP(hipMemcpy(trs.data(), a->trs.p, ((size_t)n + 1) * 4, hipMemcpyDeviceToHost));
        if (nnz > 0) {
            TS_HIP(hipMemcpy(tci.data(), a->tci.p, nnz * 4, hipMemcpyDeviceToHost));
            TS_HIP(hipMemcpy(tva.data(), a->tva32.p, nnz * 4, hipMemcpyDeviceToHost));
        }
        if (a->row_perm.p && m > 0) {
            // stored row indices back to the given ones, each transposed row in increasing order again
            TS_HIP(hipMemcpy(perm.data(), a->row_perm.p, (size_t)m * 4, hipMemcpyDeviceToHost));
            std::vector<std::pair<int32_t, float>> e;
            for (int c = 0; c < n; c++) {
                e.clear();
                for (int32_t q = trs[c]; q < trs[c + 1]; q++) e.emplace_back(perm[tci[q]], tva[q]);
                std::sort(e.begin(), e.end(), [](const auto &x, const auto &y) { return x.first < y.first; });
                for (size_t i = 0; i < e.size(); i++) { tci[trs[c] + i] = e[i].first; tva[trs[c] + i] = e[i].second; }
            }
        }
        if (a->col_perm.p && n > 0) {
            // transposed rows back in the given column order
            std::vector<int32_t> cp(n), orow(n + 1, 0), oci(std::max<int64_t>(1, nnz));
            std::vector<float> ova(std::max<int64_t>(1, nnz));
            TS_HIP(hipMemcpy(cp.data(), a->col_perm.p, (size_t)n * 4, hipMemcpyDeviceToHost));
            std::vector<int32_t> stored(n);
            for (int c = 0; c < n; c++) stored[cp[c]] = c;
            for (int g = 0; g < n; g++) {
                const int c = stored[g];
                orow[g + 1] = orow[g] + (trs[c + 1] - trs[c]);
                std::copy(tci.begin() + trs[c], tci.begin() + trs[c + 1], oci.begin() + orow[g]);
                std::copy(tva.begin() + trs[c], tva.begin() + trs[c + 1], ova.begin() + orow[g]);
            }
            trs.swap(orow);
            tci.swap(oci);
            tva.swap(ova);
        }
        if (row_start) std::copy(trs.begin(), trs.end(), row_start);
        if (column && nnz > 0) std::copy(tci.begin(), tci.begin() + nnz, column);
        if (values && nnz > 0) std::copy(tva.begin(), tva.begin() + nnz, values);
        return TS_OK;
    });
}

}  // extern "C"
