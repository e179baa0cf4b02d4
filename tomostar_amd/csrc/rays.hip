// Curved rays on gfx950: the back-trace of every arrival on its station's travel-time table and the
// path integral of the inversion grid's trilinear basis along it - the rows of the sensitivity matrix G.
//
// Reference: Forward/RayTracing.cs:73-159 (Backtrack, Kernel), Forward/TravelTimeTable.cs:116-157
// (Time, Gradient), Tomography/RayKernels.cs:62-105 (step rule, nodes ascending, float32 lengths).
// Positions are FP64 on FP32 tables, as the reference computes them.
//
// Two kernels per batch of arrivals:
//   ray_trace: one work item per arrival walks the steepest descent (7 trilinear table reads and the
//     spherical frame per step) and writes its path; arrivals are ordered by table and position so a
//     wavefront's rays read the same table in the same region.
//   ray_rows: one wavefront per arrival integrates the path. Lanes compute the pieces (midpoint,
//     trilinear weights) in parallel; the contributions then enter an LDS hash table piece by piece,
//     eight lanes (the eight corners of one piece, distinct nodes) at a time, so every node's length
//     is summed in the reference's order (piece order, RayTracing.cs:150-155); the nodes are finally
//     sorted ascending with a bitonic network and cast to float32.
#include <algorithm>
#include <chrono>
#include <memory>
#include <numeric>
#include <vector>

#include "common.hpp"
#include "rays.hpp"
#include "geometry.hpp"

struct ts_tableset;

namespace ts {
// accessors of the table set defined in eikonal.hip
const float *tableset_tables(const ts_tableset *t);
Grid tableset_grid(const ts_tableset *t);
int tableset_fields(const ts_tableset *t);
ts_context *tableset_ctx(const ts_tableset *t);
}  // namespace ts

using namespace ts;


namespace ts {

// TravelTimeTable.TimeFractional (TravelTimeTable.cs:122-136): trilinear, +inf if a corner with a
// non-zero weight is unreached.
__device__ inline double table_time_frac(const Grid &g, const float *__restrict__ T, double fx, double fy, double fz)
{
    int i0, j0, k0;
    double tx, ty, tz;
    trilinear_base(g, fx, fy, fz, i0, j0, k0, tx, ty, tz);
    double t = 0;
#pragma unroll
    for (int dk = 0; dk < 2; dk++)
#pragma unroll
        for (int dj = 0; dj < 2; dj++)
#pragma unroll
            for (int di = 0; di < 2; di++) {
                const double w = (di == 0 ? 1 - tx : tx) * (dj == 0 ? 1 - ty : ty) * (dk == 0 ? 1 - tz : tz);
                if (w == 0) continue;
                const float v = T[g.index(i0 + di, j0 + dj, k0 + dk)];
                if (!__builtin_isfinite(v)) return INFINITY;
                t += v * w;
            }
    return t;
}

__device__ inline double table_time(const Grid &g, const float *T, double lon, double lat, double depth)
{
    double fx, fy, fz;
    fractional(g, lon, lat, depth, fx, fy, fz);
    return table_time_frac(g, T, fx, fy, fz);
}

// TravelTimeTable.Gradient (TravelTimeTable.cs:143-157), s/km in (east, north, down)
__device__ inline void table_gradient(const Grid &g, const float *T, double lon, double lat, double depth, double &ge,
                                      double &gn, double &gz)
{
    double fx, fy, fz;
    fractional(g, lon, lat, depth, fx, fy, fz);
    const double d = 0.5;
    const double tx = (table_time_frac(g, T, fx + d, fy, fz) - table_time_frac(g, T, fx - d, fy, fz)) / (2 * d);
    const double ty = (table_time_frac(g, T, fx, fy + d, fz) - table_time_frac(g, T, fx, fy - d, fz)) / (2 * d);
    const double tz = (table_time_frac(g, T, fx, fy, fz + d) - table_time_frac(g, T, fx, fy, fz - d)) / (2 * d);
    const double r = kEarthRadiusKm - depth;
    ge = tx / (r * cr::cos(lat * kDeg2Rad) * g.dlon_rad);
    gn = ty / (r * g.dlat_rad);
    gz = tz / g.ddepth;
}

struct RayJob {
    int ray;          // arrival index
    int field;        // table of the arrival
    int64_t path_at;  // first path point slot
    int path_cap;     // points that fit
    int phase;        // 0 P, 1 S: the background velocity of the pieces outside the inversion grid
};

// VelocityModel1D.Evaluate (Model/VelocityModel1D.cs:31-45): piecewise linear in depth; at or above the
// first node its value; a repeated depth is a step (a point on it takes the upper value); below the last
// node its value.
__device__ inline double background_velocity(const double *depth, const double *v, int count, double d)
{
    if (d <= depth[0]) return v[0];
    for (int i = 1; i < count; i++) {
        const double a = depth[i - 1], b = depth[i];
        if (d > b) continue;
        if (b - a <= 1e-12) continue;
        const double t = (d - a) / (b - a);
        return v[i - 1] + t * (v[i] - v[i - 1]);
    }
    return v[count - 1];
}

// RayTracing.Backtrack (RayTracing.cs:73-111). path: 3 doubles (lon, lat, depth) per point.
// npts[ray] = points, -1 no ray (stalled / non-finite), -2 path longer than its slot.
__global__ __launch_bounds__(64) void ray_trace(Grid g, const float *__restrict__ tables, const double *__restrict__ stations,
                                                const double *__restrict__ events, const RayJob *__restrict__ jobs,
                                                int njobs, double step, double *__restrict__ path,
                                                int32_t *__restrict__ npts, double *__restrict__ tcg)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= njobs) return;
    const RayJob job = jobs[q];
    const float *T = tables + (size_t)job.field * g.count();
    const double slon = stations[3 * job.field], slat = stations[3 * job.field + 1];
    const double sdep = dclamp(stations[3 * job.field + 2], g.depth(0), g.depth_last);
    const double lon = events[3 * job.ray], lat = events[3 * job.ray + 1], depth = events[3 * job.ray + 2];
    double *P = path + 3 * job.path_at;
    int n = 0;
    const Vec3 station = to_cartesian(slon, slat, sdep);
    Vec3 p = to_cartesian(lon, lat, depth);
    // TCalc and the hypocentre gradient at the event (RayKernels.cs:81-82)
    const double tcalc = table_time(g, T, lon, lat, depth);
    double ge, gn, gz;
    table_gradient(g, T, lon, lat, depth, ge, gn, gz);
    tcg[4 * job.ray] = tcalc;
    tcg[4 * job.ray + 1] = ge;
    tcg[4 * job.ray + 2] = gn;
    tcg[4 * job.ray + 3] = gz;
    int status = 0;
    P[0] = lon; P[1] = lat; P[2] = depth;
    n = 1;
    if (!__builtin_isfinite(tcalc)) status = -1;
    const int max_steps = 20 * (int)ceil(vlen(vsub(p, station)) / step + 10);
    double last_t = tcalc;
    int stalls = 0;
    for (int s = 0; s < max_steps && status == 0; s++) {
        if (vlen(vsub(p, station)) <= 1.5 * step) break;
        const GeoPoint gp = from_cartesian(p);
        double e1, e2, e3;
        table_gradient(g, T, gp.lon, gp.lat, gp.depth, e1, e2, e3);
        if (!__builtin_isfinite(e1) || !__builtin_isfinite(e2) || !__builtin_isfinite(e3)) { status = -1; break; }
        // GeoMath.LocalFrame (GeoMath.cs:71-80)
        const double la = gp.lat * kDeg2Rad, lo = gp.lon * kDeg2Rad;
        double sl, cl, so, co;
        cr::sincos(la, sl, cl);
        cr::sincos(lo, so, co);
        const Vec3 east{-so, co, 0}, north{-sl * co, -sl * so, cl}, up{cl * co, cl * so, sl};
        const Vec3 grad = vadd(vadd(vscale(east, e1), vscale(north, e2)), vscale(up, -e3));
        const double norm = vlen(grad);
        if (norm < 1e-12) { status = -1; break; }
        p = vsub(p, vscale(grad, step / norm));
        GeoPoint qp = from_cartesian(p);
        qp.lon = dclamp(unwrap_lon(qp.lon, g.centre_lon), g.min_lon, g.lon_last);
        qp.lat = dclamp(qp.lat, g.min_lat, g.lat_last);
        qp.depth = dclamp(qp.depth, g.min_depth, g.depth_last);
        p = to_cartesian(qp.lon, qp.lat, qp.depth);
        if (n + 1 >= job.path_cap) { status = -2; break; }
        P[3 * n] = qp.lon; P[3 * n + 1] = qp.lat; P[3 * n + 2] = qp.depth;
        n++;
        const double t = table_time(g, T, qp.lon, qp.lat, qp.depth);
        if (!(t < last_t)) {
            if (++stalls > 20) { status = -1; break; }
        } else stalls = 0;
        last_t = fmin(last_t, t);
    }
    if (status == 0) {
        P[3 * n] = slon; P[3 * n + 1] = slat; P[3 * n + 2] = sdep;
        n++;
    }
    npts[job.ray] = status == 0 ? n : status;
}

constexpr int kChunkSeg = 32;         // path segments integrated per pass
constexpr int kMaxPiecesSeg = 4;      // pieces a segment may split into within one pass buffer
// Rows are formed in an LDS hash of kSlots entries (load kept under 1/2: a row holds at most
// kSlots / 2 nodes). Most rays fit the small table; the few longer ones are formed again with the
// large one.
constexpr int kSlotsSmall = 2048;
constexpr int kSlotsLarge = 8192;

struct RowsArgs {
    Grid inv;
    double h;                          // piece length: half the inversion grid's smallest spacing
    const double *path;
    const RayJob *jobs;
    const int32_t *npts;
    const int32_t *job_of_block;       // the jobs this launch forms (null: block q forms job q)
    int nblocks;
    int32_t *row_nodes;                // [job][row_cap]
    float *row_len;
    int64_t row_cap;
    int32_t *row_count;                // per job: nodes, -1 no row, -3 too many nodes
    // the extended forward domain's background (bg_count 0: none, every piece goes to the grid's nodes)
    int bg_count;
    const double *bg_depth, *bg_vp, *bg_vs;
    double *row_outside;               // per job: the time of the pieces outside the inversion grid
};

template <int kHashSlots>
__device__ inline uint32_t slot_hash(int32_t node) { return ((uint32_t)node * 2654435761u) & (kHashSlots - 1); }

// RayTracing.Kernel (RayTracing.cs:126-159) for one arrival per wavefront.
template <int kHashSlots>
__global__ __launch_bounds__(64) void ray_rows(RowsArgs A)
{
    __shared__ int32_t hkey[kHashSlots];
    __shared__ double hval[kHashSlots];
    __shared__ int32_t cnode[kChunkSeg * kMaxPiecesSeg * 8];
    __shared__ double cval[kChunkSeg * kMaxPiecesSeg * 8];
    __shared__ int32_t pstart[kChunkSeg + 1];
    __shared__ double cout[kChunkSeg * kMaxPiecesSeg];  // a piece's time outside the grid (0: inside)
    __shared__ int overflow;
    const int lane = threadIdx.x;
    if ((int)blockIdx.x >= A.nblocks) return;
    const int q = A.job_of_block ? A.job_of_block[blockIdx.x] : (int)blockIdx.x;
    const int slot = blockIdx.x;  // the row's place in this launch's output
    const RayJob job = A.jobs[q];
    const int np = A.npts[job.ray];
    if (np < 0) {
        if (lane == 0) {
            A.row_count[slot] = -1;
            if (A.row_outside) A.row_outside[slot] = 0.0;
        }
        return;
    }
    double outside = 0.0;  // lane 0's running sum, in piece order (RayTracing.cs:143)
    for (int s = lane; s < kHashSlots; s += 64) hkey[s] = -1;
    if (lane == 0) overflow = 0;
    __syncthreads();
    const double *P = A.path + 3 * job.path_at;
    const Grid &g = A.inv;
    for (int s0 = 1; s0 < np; s0 += kChunkSeg) {
        // pass 1: each lane one segment; its pieces' eight (node, dl*w) into the chunk buffer
        const int s = s0 + lane;
        int pieces = 0;
        Vec3 a{0, 0, 0}, d{0, 0, 0};
        double dl = 0;
        if (lane < kChunkSeg && s < np) {
            a = to_cartesian(P[3 * (s - 1)], P[3 * (s - 1) + 1], P[3 * (s - 1) + 2]);
            const Vec3 b = to_cartesian(P[3 * s], P[3 * s + 1], P[3 * s + 2]);
            d = vsub(b, a);
            const double len = vlen(d);
            if (len > 0) {
                pieces = (int)ceil(len / A.h);
                if (pieces < 1) pieces = 1;
                dl = len / pieces;
            }
        }
        // pieces of this chunk beyond the buffer are integrated in a further pass of the same chunk
        if (lane == 0) pstart[0] = 0;
        // exclusive scan of the piece counts over the chunk's lanes
        int incl = pieces;
        for (int off = 1; off < 64; off <<= 1) {
            int y = __shfl_up(incl, off);
            if (lane >= off) incl += y;
        }
        if (lane < kChunkSeg) pstart[lane + 1] = incl;
        __syncthreads();
        const int total = pstart[kChunkSeg];
        const int first = incl - pieces;
        for (int base = 0; base < total; base += kChunkSeg * kMaxPiecesSeg) {
            const int limit = min(total - base, kChunkSeg * kMaxPiecesSeg);
            for (int k = 0; k < pieces; k++) {
                const int pi = first + k - base;
                if (pi < 0 || pi >= limit) continue;
                const double f = (k + 0.5) / pieces;
                const GeoPoint mid = from_cartesian(vadd(a, vscale(d, f)));
                double fx, fy, fz;
                fractional(g, mid.lon, mid.lat, mid.depth, fx, fy, fz);
                // SphericalGrid.Contains(mid, 1e-6) (SphericalGrid.cs:186-192) when a background is given
                const double tol = 1e-6;
                const bool in = A.bg_count == 0 || (fx >= -tol && fx <= g.nx - 1 + tol && fy >= -tol &&
                                                    fy <= g.ny - 1 + tol && fz >= -tol && fz <= g.nz - 1 + tol);
                if (!in) {
                    const double v = background_velocity(A.bg_depth, job.phase ? A.bg_vs : A.bg_vp, A.bg_count, mid.depth);
                    cout[pi] = dl * (1.0 / v);
#pragma unroll
                    for (int c = 0; c < 8; c++) cnode[pi * 8 + c] = -1;
                    continue;
                }
                cout[pi] = 0.0;
                int nodes[8];
                double w[8];
                trilinear_weights(g, fx, fy, fz, nodes, w);
#pragma unroll
                for (int c = 0; c < 8; c++) {
                    cnode[pi * 8 + c] = w[c] == 0 ? -1 : nodes[c];
                    cval[pi * 8 + c] = dl * w[c] * 1.0;
                }
            }
            __syncthreads();
            // pass 2: pieces in order, the eight corner lanes of one piece at a time
            const int c = lane & 7;
            const int sub = lane >> 3;
            for (int pi = 0; pi < limit; pi++) {
                if (lane == 0) outside += cout[pi];
                if (sub == 0) {
                    const int node = cnode[pi * 8 + c];
                    if (node >= 0) {
                        uint32_t slot = slot_hash<kHashSlots>(node);
                        for (int probe = 0; probe < kHashSlots; probe++) {
                            int k0 = hkey[slot];
                            if (k0 == node) { hval[slot] += cval[pi * 8 + c]; break; }
                            if (k0 == -1) {
                                int prev = atomicCAS(&hkey[slot], -1, node);
                                if (prev == -1) { hval[slot] = cval[pi * 8 + c]; break; }
                                if (prev == node) { hval[slot] += cval[pi * 8 + c]; break; }
                            }
                            slot = (slot + 1) & (kHashSlots - 1);
                            if (probe == kHashSlots / 2) { overflow = 1; }
                        }
                    }
                }
                __syncthreads();
            }
        }
    }
    __syncthreads();
    // the row: occupied slots, ascending by node. The occupied slots are moved to the front of the
    // table in slot order, 64 slots at a time (all lanes read their slot, then write it at or below
    // the slots read so far), and only the next power of two of their count is sorted, with a
    // bitonic network.
    int count = 0;
    for (int base = 0; base < kHashSlots; base += 64) {
        const int32_t k = hkey[base + lane];
        const double v = hval[base + lane];
        const uint64_t occ = __ballot(k != -1);
        __syncthreads();
        if (k != -1) {
            const int at = count + __popcll(occ & ((1ull << lane) - 1));
            hkey[at] = k;
            hval[at] = v;
        }
        count += __popcll(occ);
        __syncthreads();
    }
    if (lane == 0 && A.row_outside) A.row_outside[slot] = outside;
    if (count > A.row_cap || overflow) {
        if (lane == 0) A.row_count[slot] = -3;
        return;
    }
    int span = 64;
    while (span < count) span <<= 1;
    for (int i = count + lane; i < span; i += 64) hkey[i] = -1;  // pads sort last
    __syncthreads();
    int32_t *key = hkey;
    double *val = hval;
    for (int size = 2; size <= span; size <<= 1)
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            for (int i = lane; i < span; i += 64) {
                const int jx = i ^ stride;
                if (jx > i) {
                    const bool asc = (i & size) == 0;
                    // empty slots and pads (-1) sort last
                    const uint32_t ki = (uint32_t)key[i], kj = (uint32_t)key[jx];
                    if ((ki > kj) == asc) {
                        key[i] = (int32_t)kj; key[jx] = (int32_t)ki;
                        const double t = val[i]; val[i] = val[jx]; val[jx] = t;
                    }
                }
            }
            __syncthreads();
        }
    for (int i = lane; i < count; i += 64) {
        A.row_nodes[(size_t)slot * A.row_cap + i] = key[i];
        A.row_len[(size_t)slot * A.row_cap + i] = (float)val[i];
    }
    if (lane == 0) A.row_count[slot] = count;
}

// rows of a batch (each in its row_cap slot) to their places in the arrival-ordered CSR; dst_at < 0:
// the job has no row to keep
__global__ void ray_place(const int32_t *row_nodes, const float *row_len, int64_t row_cap, const int32_t *row_count,
                          const int64_t *dst_at, int njobs, int32_t *nodes, float *lengths)
{
    const int q = blockIdx.x;
    if (q >= njobs) return;
    const int c = row_count[q];
    const int64_t at = dst_at[q];
    if (at < 0 || c <= 0) return;
    for (int i = threadIdx.x; i < c; i += blockDim.x) {
        nodes[at + i] = row_nodes[(size_t)q * row_cap + i];
        lengths[at + i] = row_len[(size_t)q * row_cap + i];
    }
}

// tight rows of a batch to their places in the arrival-ordered CSR
__global__ void ray_move(const int32_t *src_nodes, const float *src_len, const int64_t *src_at, const int64_t *dst_at,
                         const int32_t *count, int n, int32_t *nodes, float *lengths)
{
    const int q = blockIdx.x;
    if (q >= n) return;
    const int c = count[q];
    const int64_t a = src_at[q], b = dst_at[q];
    for (int i = threadIdx.x; i < c; i += blockDim.x) {
        nodes[b + i] = src_nodes[a + i];
        lengths[b + i] = src_len[a + i];
    }
}

}  // namespace ts

// ================================================================================== C ABI

extern "C" {

int ts_rays_trace_extended(ts_tableset *tables, const ts_grid *inversion, const double *field_station, int nrays,
                           const int32_t *field_of_ray, const int32_t *phase_of_ray, const double *events,
                           double step_km, const ts_background *background, ts_rays **out, char *reason,
                           int reason_len)
{
    if (out) *out = nullptr;
    if (!tables || !inversion || !field_station || nrays < 0 || (nrays > 0 && (!field_of_ray || !events)) || !out ||
        !(step_km > 0) || !grid_valid(*inversion) ||
        (background && (background->count < 1 || !background->depth_km || !background->vp || !background->vs ||
                        (nrays > 0 && !phase_of_ray)))) {
        set_reason(reason, reason_len, "invalid argument");
        return TS_EINVAL;
    }
    return guarded(reason, reason_len, [&]() -> int {
        // VelocityModel1D.Validate (Model/VelocityModel1D.cs:47-56)
        const int nbg = background ? background->count : 0;
        for (int i = 0; i < nbg; i++) {
            const double vp = background->vp[i], vs = background->vs[i], z = background->depth_km[i];
            if (!(vp > 0) || !(vs > 0) || !std::isfinite(vp) || !std::isfinite(vs))
                fail(TS_EINVAL, "background node %d: velocities must be positive", i);
            if (vs >= vp) fail(TS_EINVAL, "background node %d: Vs must be smaller than Vp", i);
            if (!std::isfinite(z) || (i > 0 && z < background->depth_km[i - 1]))
                fail(TS_EINVAL, "background node %d: depths must not decrease", i);
        }
        if (nbg > 0)
            for (int r = 0; r < nrays; r++)
                if (phase_of_ray[r] != 0 && phase_of_ray[r] != 1) fail(TS_EINVAL, "arrival %d: phase %d", r, phase_of_ray[r]);
        ts_context *ctx = tableset_ctx(tables);
        const float *tab = tableset_tables(tables);
        if (!tab) fail(TS_EINVAL, "the table set has not been solved");
        TS_HIP(hipSetDevice(ctx->device));
        hipStream_t st = ctx->stream;
        const Grid fwd = tableset_grid(tables);
        const Grid inv = make_grid(*inversion);
        const int nf = tableset_fields(tables);
        for (int r = 0; r < nrays; r++)
            if (field_of_ray[r] < 0 || field_of_ray[r] >= nf) fail(TS_EINVAL, "arrival %d names table %d of %d", r, field_of_ray[r], nf);
        std::unique_ptr<ts_rays> R(new ts_rays);
        R->ctx = ctx;
        R->nrays = nrays;
        R->inversion = *inversion;
        R->counts.assign(nrays, -1);
        R->tcalc_grad.assign(4 * (size_t)nrays, 0.0);
        R->path_points.assign(nrays, 0);
        R->outside.assign(nrays, 0.0);
        R->extended = nbg > 0;
        dbuf<double> d_bg(3 * (size_t)std::max(1, nbg));
        if (nbg > 0) {
            TS_HIP(hipMemcpyAsync(d_bg.p, background->depth_km, nbg * 8, hipMemcpyHostToDevice, st));
            TS_HIP(hipMemcpyAsync(d_bg.p + nbg, background->vp, nbg * 8, hipMemcpyHostToDevice, st));
            TS_HIP(hipMemcpyAsync(d_bg.p + 2 * nbg, background->vs, nbg * 8, hipMemcpyHostToDevice, st));
        }
        const double h = 0.5 * min_spacing_km(inv);
        // arrivals ordered by table, then by event position, so a wavefront shares a table and a region
        std::vector<int> order(nrays);
        std::iota(order.begin(), order.end(), 0);
        std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
            if (field_of_ray[a] != field_of_ray[b]) return field_of_ray[a] < field_of_ray[b];
            return std::make_pair(events[3 * a + 1], events[3 * a]) < std::make_pair(events[3 * b + 1], events[3 * b]);
        });
        dbuf<double> d_st(3 * (size_t)nf), d_ev(3 * (size_t)std::max(1, nrays)), d_tcg(4 * (size_t)std::max(1, nrays));
        dbuf<int32_t> d_npts(std::max(1, nrays));
        TS_HIP(hipMemcpyAsync(d_st.p, field_station, 3 * (size_t)nf * 8, hipMemcpyHostToDevice, st));
        if (nrays > 0) TS_HIP(hipMemcpyAsync(d_ev.p, events, 3 * (size_t)nrays * 8, hipMemcpyHostToDevice, st));
        // A batch keeps its formed rows on the device until every batch is done; the rows are then
        // placed in arrival order (RayKernels.cs:114) without passing through the host.
        struct Formed {
            dbuf<int32_t> nodes;             // the rows of a launch, back to back
            dbuf<float> len;
            std::vector<int32_t> ray, count; // per kept row: arrival and length
            std::vector<int64_t> at;         // per kept row: offset in nodes / len
        };
        std::vector<Formed> formed;
        const int batch = 65536;
        auto t0 = std::chrono::steady_clock::now();
        // rows of one launch (row_cap slots) packed back to back; keep[q] < 0 drops job q
        auto pack = [&](const dbuf<int32_t> &rn, const dbuf<float> &rl, int64_t cap, const dbuf<int32_t> &rc,
                        const std::vector<int32_t> &cnt, const std::vector<int32_t> &ray_of) {
            const int n = (int)cnt.size();
            Formed F;
            std::vector<int64_t> dst(n, -1);
            int64_t tot = 0;
            for (int q = 0; q < n; q++)
                if (ray_of[q] >= 0 && cnt[q] > 0) {
                    dst[q] = tot;
                    F.ray.push_back(ray_of[q]);
                    F.count.push_back(cnt[q]);
                    F.at.push_back(tot);
                    tot += cnt[q];
                }
            if (tot == 0) return;
            F.nodes.alloc(tot);
            F.len.alloc(tot);
            dbuf<int64_t> d_dst(n);
            TS_HIP(hipMemcpyAsync(d_dst.p, dst.data(), n * 8, hipMemcpyHostToDevice, st));
            ray_place<<<n, 256, 0, st>>>(rn.p, rl.p, cap, rc.p, d_dst.p, n, F.nodes.p, F.len.p);
            TS_HIP(hipGetLastError());
            TS_HIP(hipStreamSynchronize(st));
            formed.push_back(std::move(F));
        };
        auto form_rows = [&](const dbuf<double> &d_path, const dbuf<RayJob> &d_jobs, const std::vector<RayJob> &jobs,
                             const std::vector<int32_t> &npts_all) {
            const int nb = (int)jobs.size();
            const int64_t cap = kSlotsSmall / 2;
            dbuf<int32_t> rn((size_t)nb * cap), rc(nb);
            dbuf<float> rl((size_t)nb * cap);
            dbuf<double> ro(nb);
            RowsArgs A{inv, h, d_path.p, d_jobs.p, d_npts.p, nullptr, nb, rn.p, rl.p, cap, rc.p,
                       nbg, d_bg.p, d_bg.p + nbg, d_bg.p + 2 * nbg, ro.p};
            ray_rows<kSlotsSmall><<<nb, 64, 0, st>>>(A);
            TS_HIP(hipGetLastError());
            std::vector<int32_t> cnt(nb);
            std::vector<double> outs(nb);
            TS_HIP(hipMemcpyAsync(cnt.data(), rc.p, nb * 4, hipMemcpyDeviceToHost, st));
            TS_HIP(hipMemcpyAsync(outs.data(), ro.p, nb * 8, hipMemcpyDeviceToHost, st));
            TS_HIP(hipStreamSynchronize(st));
            std::vector<int32_t> ray_of(nb), longer;
            for (int q = 0; q < nb; q++) {
                const int r = jobs[q].ray;
                R->path_points[r] = npts_all[r];
                ray_of[q] = r;
                if (cnt[q] == -3) longer.push_back(q);
                else {
                    R->counts[r] = cnt[q] >= 0 ? cnt[q] : -1;
                    R->outside[r] = cnt[q] >= 0 ? outs[q] : 0.0;
                }
            }
            pack(rn, rl, cap, rc, cnt, ray_of);
            // rows too long for the small table: formed again with the large one
            if (!longer.empty()) {
                const int nl = (int)longer.size();
                const int64_t capl = kSlotsLarge / 2;
                dbuf<int32_t> ln((size_t)nl * capl), lc(nl), which(nl);
                dbuf<float> ll((size_t)nl * capl);
                dbuf<double> lo(nl);
                TS_HIP(hipMemcpyAsync(which.p, longer.data(), nl * 4, hipMemcpyHostToDevice, st));
                RowsArgs B{inv, h, d_path.p, d_jobs.p, d_npts.p, which.p, nl, ln.p, ll.p, capl, lc.p,
                           nbg, d_bg.p, d_bg.p + nbg, d_bg.p + 2 * nbg, lo.p};
                ray_rows<kSlotsLarge><<<(unsigned)nl, 64, 0, st>>>(B);
                TS_HIP(hipGetLastError());
                std::vector<int32_t> cnt2(nl), ray2(nl);
                std::vector<double> outs2(nl);
                TS_HIP(hipMemcpyAsync(cnt2.data(), lc.p, nl * 4, hipMemcpyDeviceToHost, st));
                TS_HIP(hipMemcpyAsync(outs2.data(), lo.p, nl * 8, hipMemcpyDeviceToHost, st));
                TS_HIP(hipStreamSynchronize(st));
                for (int i = 0; i < nl; i++) {
                    const int r = jobs[longer[i]].ray;
                    if (cnt2[i] == -3) fail(TS_DECLINED, "ray %d touches more than %lld nodes", r, (long long)capl);
                    ray2[i] = r;
                    R->counts[r] = cnt2[i] >= 0 ? cnt2[i] : -1;
                    R->outside[r] = cnt2[i] >= 0 ? outs2[i] : 0.0;
                }
                pack(ln, ll, capl, lc, cnt2, ray2);
            }
        };
        for (int b0 = 0; b0 < nrays; b0 += batch) {
            const int nb = std::min(batch, nrays - b0);
            // path slots sized from the straight distance; a path that outgrows its slot is traced again
            std::vector<RayJob> jobs(nb);
            int64_t at = 0;
            for (int q = 0; q < nb; q++) {
                const int r = order[b0 + q];
                const int f = field_of_ray[r];
                const Vec3 a = to_cartesian(events[3 * r], events[3 * r + 1], events[3 * r + 2]);
                const Vec3 s = to_cartesian(field_station[3 * f], field_station[3 * f + 1], field_station[3 * f + 2]);
                const int cap = 3 * (int)ceil(vlen(vsub(a, s)) / step_km) + 64;
                jobs[q] = RayJob{r, f, at, cap, nbg > 0 ? phase_of_ray[r] : 0};
                at += cap;
            }
            std::vector<int32_t> npts_all(nrays);
            std::vector<RayJob> redo;
            {
                dbuf<RayJob> d_jobs(nb);
                dbuf<double> d_path(3 * (size_t)at);
                TS_HIP(hipMemcpyAsync(d_jobs.p, jobs.data(), nb * sizeof(RayJob), hipMemcpyHostToDevice, st));
                ray_trace<<<(nb + 63) / 64, 64, 0, st>>>(fwd, tab, d_st.p, d_ev.p, d_jobs.p, nb, step_km, d_path.p,
                                                          d_npts.p, d_tcg.p);
                TS_HIP(hipGetLastError());
                TS_HIP(hipMemcpyAsync(npts_all.data(), d_npts.p, (size_t)nrays * 4, hipMemcpyDeviceToHost, st));
                TS_HIP(hipStreamSynchronize(st));
                // paths that outgrew their slots are traced again with the bound of Backtrack's loop
                int64_t at2 = 0;
                for (int q = 0; q < nb; q++) {
                    if (npts_all[jobs[q].ray] == -2) {
                        RayJob j = jobs[q];
                        const int r = j.ray, f = j.field;
                        const Vec3 a = to_cartesian(events[3 * r], events[3 * r + 1], events[3 * r + 2]);
                        const Vec3 s = to_cartesian(field_station[3 * f], field_station[3 * f + 1], field_station[3 * f + 2]);
                        j.path_cap = 20 * (int)ceil(vlen(vsub(a, s)) / step_km + 10) + 4;
                        j.path_at = at2;
                        at2 += j.path_cap;
                        redo.push_back(j);
                    }
                }
                form_rows(d_path, d_jobs, jobs, npts_all);
                if (!redo.empty()) {
                    // their first paths were cut short (no row was formed): the retrace below forms them
                    dbuf<RayJob> d_jobs2(redo.size());
                    dbuf<double> d_path2(3 * (size_t)at2);
                    TS_HIP(hipMemcpyAsync(d_jobs2.p, redo.data(), redo.size() * sizeof(RayJob), hipMemcpyHostToDevice, st));
                    ray_trace<<<((unsigned)redo.size() + 63) / 64, 64, 0, st>>>(fwd, tab, d_st.p, d_ev.p, d_jobs2.p,
                                                                                (int)redo.size(), step_km, d_path2.p,
                                                                                d_npts.p, d_tcg.p);
                    TS_HIP(hipGetLastError());
                    TS_HIP(hipMemcpyAsync(npts_all.data(), d_npts.p, (size_t)nrays * 4, hipMemcpyDeviceToHost, st));
                    TS_HIP(hipStreamSynchronize(st));
                    for (const RayJob &j : redo)
                        if (npts_all[j.ray] == -2) fail(TS_EDEVICE, "ray path exceeded the back-trace bound");
                    form_rows(d_path2, d_jobs2, redo, npts_all);
                }
            }
        }
        std::vector<double> tcg(4 * (size_t)nrays);
        if (nrays > 0) TS_HIP(hipMemcpy(tcg.data(), d_tcg.p, tcg.size() * 8, hipMemcpyDeviceToHost));
        R->tcalc_grad = tcg;
        for (int r = 0; r < nrays; r++)
            if (R->counts[r] >= 0 && !std::isfinite(tcg[4 * r])) {
                R->counts[r] = -1;  // RayKernels.cs:94
                R->outside[r] = 0.0;
            }
        // rows in arrival order (RayKernels.cs:114)
        R->row_start.assign(nrays + 1, 0);
        for (int r = 0; r < nrays; r++) R->row_start[r + 1] = R->row_start[r] + std::max(0, R->counts[r]);
        const int64_t nnz = R->row_start[nrays];
        R->nodes.alloc(std::max<int64_t>(1, nnz));
        R->lengths.alloc(std::max<int64_t>(1, nnz));
        for (Formed &F : formed) {
            const int n = (int)F.ray.size();
            std::vector<int64_t> dst(n);
            for (int q = 0; q < n; q++) dst[q] = R->row_start[F.ray[q]];
            std::vector<int32_t> cnt(n);
            for (int q = 0; q < n; q++) cnt[q] = R->counts[F.ray[q]] > 0 ? F.count[q] : 0;  // rows voided since
            dbuf<int64_t> d_src(n), d_dst(n);
            dbuf<int32_t> d_cnt(n);
            TS_HIP(hipMemcpyAsync(d_src.p, F.at.data(), n * 8, hipMemcpyHostToDevice, st));
            TS_HIP(hipMemcpyAsync(d_dst.p, dst.data(), n * 8, hipMemcpyHostToDevice, st));
            TS_HIP(hipMemcpyAsync(d_cnt.p, cnt.data(), n * 4, hipMemcpyHostToDevice, st));
            ray_move<<<n, 256, 0, st>>>(F.nodes.p, F.len.p, d_src.p, d_dst.p, d_cnt.p, n, R->nodes.p, R->lengths.p);
            TS_HIP(hipGetLastError());
            TS_HIP(hipStreamSynchronize(st));
        }
        R->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        *out = R.release();
        return TS_OK;
    });
}

int ts_rays_trace(ts_tableset *tables, const ts_grid *inversion, const double *field_station, int nrays,
                  const int32_t *field_of_ray, const double *events, double step_km, ts_rays **out, char *reason,
                  int reason_len)
{
    return ts_rays_trace_extended(tables, inversion, field_station, nrays, field_of_ray, nullptr, events, step_km,
                                  nullptr, out, reason, reason_len);
}

int ts_rays_outside_time(ts_rays *r, double *outside)
{
    if (!r || !outside) return TS_EINVAL;
    std::copy(r->outside.begin(), r->outside.end(), outside);
    return TS_OK;
}

int ts_rays_rows(ts_rays *r, int32_t *counts, int64_t *row_start, int32_t *nodes, float *lengths, double *tcalc_grad,
                 int32_t *path_points)
{
    if (!r) return TS_EINVAL;
    return guarded(nullptr, 0, [&]() -> int {
        if (counts) std::copy(r->counts.begin(), r->counts.end(), counts);
        if (row_start) std::copy(r->row_start.begin(), r->row_start.end(), row_start);
        const int64_t nnz = r->row_start.empty() ? 0 : r->row_start.back();
        if (nnz > 0 && (nodes || lengths)) TS_HIP(hipSetDevice(r->ctx->device));
        if (nnz > 0 && nodes) copy_to_host(r->stage, exec_stream(r->ctx), nodes, r->nodes.p, nnz * 4);
        if (nnz > 0 && lengths) copy_to_host(r->stage, exec_stream(r->ctx), lengths, r->lengths.p, nnz * 4);
        if (tcalc_grad) std::copy(r->tcalc_grad.begin(), r->tcalc_grad.end(), tcalc_grad);
        if (path_points) std::copy(r->path_points.begin(), r->path_points.end(), path_points);
        return TS_OK;
    });
}

int ts_rays_stats(ts_rays *r, int64_t *nnz, double *seconds)
{
    if (!r) return TS_EINVAL;
    if (nnz) *nnz = r->row_start.empty() ? 0 : r->row_start.back();
    if (seconds) *seconds = r->seconds;
    return TS_OK;
}

void ts_rays_release(ts_rays *r)
{
    if (!r) return;
    (void)hipSetDevice(r->ctx->device);
    delete r;
}

}  // extern "C"
