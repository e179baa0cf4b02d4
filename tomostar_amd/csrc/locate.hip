// EventLocator.GridSearch (Location/EventLocator.cs:114-175) on the device: for every event, the node
// of the forward grid whose travel times best fit the event's picks, over the resident tables of a
// table set. At a node the origin time is the weighted median of (observed - predicted - correction)
// and the misfit the weighted L1 norm of what remains (a pick whose table is not finite there counts
// 10 x its weight). A global search scans a strided sub-grid of about CoarseCandidates nodes and then
// every node within two strides of the best one; a local search every node within 6 of the start.
//
// One workgroup per event; its threads take the candidates in turn and the workgroup keeps the
// lowest misfit, the first in grid order on ties (the reference's strict '<' over k, j, i loops).
// Every misfit is computed as the reference computes it - the residuals in pick order, a stable sort
// of the finite ones, the weights summed in sorted order, the misfit summed in pick order - so the
// chosen node, origin time and misfit are the reference's to the last bit.
#include <climits>
#include <cmath>

#include "common.hpp"
#include "geometry.hpp"

struct ts_tableset;

namespace ts {

const float *tableset_tables(const ts_tableset *t);
Grid tableset_grid(const ts_tableset *t);
int tableset_fields(const ts_tableset *t);
ts_context *tableset_ctx(const ts_tableset *t);

namespace {

constexpr int kLocThreads = 256;
constexpr int kLocMaxPicks = 512;

struct LocArgs {
    const float *tab;           // [field][N] travel times
    int nx, ny, nz, N;
    const int32_t *pick_start;  // [events + 1]
    const int32_t *pick_field;  // table of each pick
    const double *obs, *weight, *corr;
    const int32_t *center;      // [events][3] start node (i, j, k) of a local search; null: global
    const int32_t *k_fixed;     // [events] fixed depth index, -1 free; null: none fixed
    int stride;                 // coarse stride of a global search
    int32_t *best_node;
    double *best_t0, *best_misfit;
};

// Misfit(idx, out t0) of EventLocator.GridSearch (EventLocator.cs:121-133) with WeightedMedian (:311-323).
__device__ double misfit(const LocArgs &A, int e0, int q_count, int idx, double *t0_out, double *r, int16_t *ord)
{
    for (int q = 0; q < q_count; q++) {
        const float t = A.tab[(size_t)A.pick_field[e0 + q] * A.N + idx];
        r[q] = isfinite(t) ? A.obs[e0 + q] - (double)t - A.corr[e0 + q] : (double)NAN;
    }
    int n = 0;
    for (int q = 0; q < q_count; q++) {  // stable insertion sort of the finite residuals
        if (isnan(r[q])) continue;
        int p = n++;
        while (p > 0 && r[ord[p - 1]] > r[q]) {
            ord[p] = ord[p - 1];
            p--;
        }
        ord[p] = (int16_t)q;
    }
    double t0 = 0;
    if (n > 0) {
        double sum = 0;
        for (int i = 0; i < n; i++) sum += A.weight[e0 + ord[i]];
        const double half = sum / 2;
        double acc = 0;
        t0 = r[ord[n - 1]];
        for (int i = 0; i < n; i++) {
            acc += A.weight[e0 + ord[i]];
            if (acc >= half) { t0 = r[ord[i]]; break; }
        }
    }
    double m = 0;
    for (int q = 0; q < q_count; q++)
        m += isnan(r[q]) ? 10 * A.weight[e0 + q] : A.weight[e0 + q] * fabs(r[q] - t0);
    *t0_out = t0;
    return m;
}

// Search(i0, i1, j0, j1, k0, k1, stride) (EventLocator.cs:150-174): the lowest misfit over the
// workgroup's candidates, the first in grid order on ties.
__device__ int search(const LocArgs &A, int e, int e0, int q_count, int i0, int i1, int j0, int j1, int k0, int k1,
                      int stride, double *r, int16_t *ord, double *sm, int *si)
{
    i0 = max(0, i0); j0 = max(0, j0); k0 = max(0, k0);
    i1 = min(A.nx - 1, i1); j1 = min(A.ny - 1, j1); k1 = min(A.nz - 1, k1);
    if (A.k_fixed && A.k_fixed[e] >= 0) k0 = k1 = A.k_fixed[e];
    const int ni = i1 >= i0 ? (i1 - i0) / stride + 1 : 0;
    const int nj = j1 >= j0 ? (j1 - j0) / stride + 1 : 0;
    const int nk = k1 >= k0 ? (k1 - k0) / stride + 1 : 0;
    const int total = ni * nj * nk;
    double best_m = INFINITY;
    int best = INT_MAX;
    for (int c = threadIdx.x; c < total; c += blockDim.x) {
        const int kk = k0 + (c / (ni * nj)) * stride, jj = j0 + ((c / ni) % nj) * stride, ii = i0 + (c % ni) * stride;
        const int idx = (kk * A.ny + jj) * A.nx + ii;
        double t0;
        const double m = misfit(A, e0, q_count, idx, &t0, r, ord);
        if (m < best_m || (m == best_m && idx < best)) { best_m = m; best = idx; }
    }
    sm[threadIdx.x] = best_m;
    si[threadIdx.x] = best;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            const double m2 = sm[threadIdx.x + s];
            const int i2 = si[threadIdx.x + s];
            if (m2 < sm[threadIdx.x] || (m2 == sm[threadIdx.x] && i2 < si[threadIdx.x])) {
                sm[threadIdx.x] = m2;
                si[threadIdx.x] = i2;
            }
        }
        __syncthreads();
    }
    const int winner = si[0];
    __syncthreads();  // sm / si are reused by the next search
    return winner == INT_MAX ? (k0 * A.ny + j0) * A.nx + i0 : winner;
}

__global__ __launch_bounds__(kLocThreads) void loc_grid_search(LocArgs A, int nev)
{
    __shared__ double sm[kLocThreads];
    __shared__ int si[kLocThreads];
    const int e = blockIdx.x;
    if (e >= nev) return;
    const int e0 = A.pick_start[e], q_count = A.pick_start[e + 1] - e0;
    double r[kLocMaxPicks];
    int16_t ord[kLocMaxPicks];
    int best;
    if (A.center) {
        const int ci = A.center[3 * e], cj = A.center[3 * e + 1], ck = A.center[3 * e + 2];
        best = search(A, e, e0, q_count, ci - 6, ci + 6, cj - 6, cj + 6, ck - 6, ck + 6, 1, r, ord, sm, si);
    } else {
        const int s = A.stride;
        best = search(A, e, e0, q_count, 0, A.nx - 1, 0, A.ny - 1, 0, A.nz - 1, s, r, ord, sm, si);
        const int bi = best % A.nx, bj = (best / A.nx) % A.ny, bk = best / (A.nx * A.ny);
        best = search(A, e, e0, q_count, bi - 2 * s, bi + 2 * s, bj - 2 * s, bj + 2 * s, bk - 2 * s, bk + 2 * s, 1, r, ord,
                      sm, si);
    }
    if (threadIdx.x == 0) {
        double t0;
        const double m = misfit(A, e0, q_count, best, &t0, r, ord);
        A.best_node[e] = best;
        A.best_t0[e] = t0;
        A.best_misfit[e] = m;
    }
}

}  // namespace
}  // namespace ts

using namespace ts;

extern "C" {

int ts_locate_grid_search(ts_tableset *t, int nevents, const int32_t *pick_start, const int32_t *pick_field,
                          const double *obs, const double *weight, const double *correction, int coarse_candidates,
                          const int32_t *center, const int32_t *k_fixed, int32_t *best_node, double *best_t0,
                          double *best_misfit, char *reason, int reason_len)
{
    if (!t || nevents < 0 || (nevents > 0 && (!pick_start || !best_node || !best_t0 || !best_misfit))) {
        set_reason(reason, reason_len, "invalid argument");
        return TS_EINVAL;
    }
    return guarded(reason, reason_len, [&]() -> int {
        ts_context *ctx = tableset_ctx(t);
        TS_HIP(hipSetDevice(ctx->device));
        const float *tab = tableset_tables(t);
        if (!tab) fail(TS_EINVAL, "the table set has not been solved");
        if (nevents == 0) return TS_OK;
        const Grid g = tableset_grid(t);
        const int nf = tableset_fields(t);
        const int npicks = pick_start[nevents];
        if (pick_start[0] != 0) fail(TS_EINVAL, "pick_start[0] must be 0");
        for (int e = 0; e < nevents; e++) {
            const int q = pick_start[e + 1] - pick_start[e];
            if (q < 0) fail(TS_EINVAL, "pick_start decreases at event %d", e);
            if (q > kLocMaxPicks) fail(TS_DECLINED, "event %d has %d picks (at most %d on the device)", e, q, kLocMaxPicks);
        }
        if (npicks > 0 && (!pick_field || !obs || !weight || !correction)) fail(TS_EINVAL, "pick arrays are null");
        for (int q = 0; q < npicks; q++)
            if ((uint32_t)pick_field[q] >= (uint32_t)nf) fail(TS_EINVAL, "pick %d names table %d of %d", q, pick_field[q], nf);
        const int64_t N = (int64_t)g.nx * g.ny * g.nz;
        if (center)
            for (int e = 0; e < nevents; e++)
                if (center[3 * e] < -1000000 || center[3 * e] > 1000000 || center[3 * e + 1] < -1000000 ||
                    center[3 * e + 1] > 1000000 || center[3 * e + 2] < -1000000 || center[3 * e + 2] > 1000000)
                    fail(TS_EINVAL, "start node of event %d out of range", e);
        if (k_fixed)
            for (int e = 0; e < nevents; e++)
                if (k_fixed[e] >= g.nz) fail(TS_EINVAL, "fixed depth index of event %d beyond the grid", e);
        // the coarse stride: max(1, ceil(cbrt(N / CoarseCandidates))) (EventLocator.cs:139)
        const int stride = std::max(1, (int)std::ceil(std::cbrt((double)N / std::max(1, coarse_candidates))));
        hipStream_t st = ctx->stream;
        dbuf<int32_t> d_start(nevents + 1), d_field(std::max(1, npicks)), d_center(center ? 3 * nevents : 1),
            d_kf(k_fixed ? nevents : 1), d_node(nevents);
        dbuf<double> d_obs(std::max(1, npicks)), d_w(std::max(1, npicks)), d_c(std::max(1, npicks)), d_t0(nevents),
            d_m(nevents);
        TS_HIP(hipMemcpyAsync(d_start.p, pick_start, (size_t)(nevents + 1) * 4, hipMemcpyHostToDevice, st));
        if (npicks > 0) {
            TS_HIP(hipMemcpyAsync(d_field.p, pick_field, (size_t)npicks * 4, hipMemcpyHostToDevice, st));
            TS_HIP(hipMemcpyAsync(d_obs.p, obs, (size_t)npicks * 8, hipMemcpyHostToDevice, st));
            TS_HIP(hipMemcpyAsync(d_w.p, weight, (size_t)npicks * 8, hipMemcpyHostToDevice, st));
            TS_HIP(hipMemcpyAsync(d_c.p, correction, (size_t)npicks * 8, hipMemcpyHostToDevice, st));
        }
        if (center) TS_HIP(hipMemcpyAsync(d_center.p, center, (size_t)nevents * 12, hipMemcpyHostToDevice, st));
        if (k_fixed) TS_HIP(hipMemcpyAsync(d_kf.p, k_fixed, (size_t)nevents * 4, hipMemcpyHostToDevice, st));
        LocArgs A;
        A.tab = tab;
        A.nx = g.nx; A.ny = g.ny; A.nz = g.nz; A.N = (int)N;
        A.pick_start = d_start.p; A.pick_field = d_field.p;
        A.obs = d_obs.p; A.weight = d_w.p; A.corr = d_c.p;
        A.center = center ? d_center.p : nullptr;
        A.k_fixed = k_fixed ? d_kf.p : nullptr;
        A.stride = stride;
        A.best_node = d_node.p; A.best_t0 = d_t0.p; A.best_misfit = d_m.p;
        loc_grid_search<<<nevents, kLocThreads, 0, st>>>(A, nevents);
        TS_HIP(hipGetLastError());
        TS_HIP(hipMemcpyAsync(best_node, d_node.p, (size_t)nevents * 4, hipMemcpyDeviceToHost, st));
        TS_HIP(hipMemcpyAsync(best_t0, d_t0.p, (size_t)nevents * 8, hipMemcpyDeviceToHost, st));
        TS_HIP(hipMemcpyAsync(best_misfit, d_m.p, (size_t)nevents * 8, hipMemcpyDeviceToHost, st));
        TS_HIP(hipStreamSynchronize(st));
        return TS_OK;
    });
}

}  // extern "C"
