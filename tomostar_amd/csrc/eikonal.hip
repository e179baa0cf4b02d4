// The batched 3-D spherical eikonal on gfx950.
//
// Scheme (the reference's, Forward/SphericalEikonal.cs:284-396 Relax and Forward/EikonalOpenCl.cs:387-623):
// Godunov upwind update, Gauss-Seidel in the eight fast-sweeping orderings, each walked along the
// diagonal planes i'+j'+k' = p; a first-order phase to its fixed point, then a phase with the
// second-order promotion; a round of eight sweeps that lowers no node by more than 1e-7 s ends a
// phase; a phase that has not settled within max(32, (Nx+Ny+Nz)/4) rounds is declined.
//
// Two device forms:
//   plain (TS_EIK_PLAIN_FP64): the reference form on this GPU - FP64 times, one field per work item,
//     one launch per diagonal plane, a per-(node,field) wake byte, batch-level convergence read back
//     by the host after every round. It is the baseline the fast form is measured against.
//   fast (TS_EIK_FAST_FP32 mixed / TS_EIK_FAST_FP64): all fields of a table set in one pass. A wavefront
//     updates one node for 64 fields, one field per lane, over times stored node-major in groups of
//     64 fields ([group][node][64]), so every neighbour read is one contiguous 256-byte (FP32) line;
//     the diagonal planes of all groups go out as one graph per round; each field keeps its own phase
//     and round count (so every field follows Relax exactly), and a wave skips a node when no node it
//     reads moved since its last update (one wake byte per node and group, raised with a ballot).
//     The mixed form runs the first-order phase on FP32 times in the offset form of the update (every
//     term of the quadratic is a cell crossing time measured from the smallest upwind neighbour; the
//     first-order fixed point is unique and FP32 reaches it to rounding) and the second-order phase
//     on FP64 times, where a node's readers wake only on drops above 1e-10 s; the FP64 form runs both
//     phases in FP64 with Relax's arithmetic term for term, wakes on every drop, and agrees with Relax
//     to the last bit.
#include <algorithm>
#include <chrono>
#include <memory>
#include <type_traits>
#include <vector>

#include "common.hpp"
#include "eikonal_update.hpp"
#include "geometry.hpp"

namespace ts {

constexpr double kSettleTol = 1e-7;  // SphericalEikonal.cs:332 SettleTolerance
constexpr int kLanes = 64;
// The mixed form's first-order phase on FP32 times counts a node as moved when it dropped by more than
// the settle tolerance and by more than this many FP32 ulps of its time (4096 ulps: 4.9e-4 of the
// time). The phase only seeds the FP64 second-order phase, which lowers every time from there until its
// own rounds settle at 1e-7 s: times only ever drop, so a first-order phase stopped early leaves times
// above its fixed point, and the second-order phase takes them down the rest of the way. Measured on
// 400 fields of 201 x 201 x 101 (first-order rounds, solve seconds, largest relative distance to the
// FP64 form over every tenth field): 4 ulps (5, 1.089 s, 4.571e-6), 256 (4, 1.068 s, 4.571e-6), 4096
// (3, 1.029 s, 4.571e-6), 16384 (3, 1.031 s, 4.571e-6), 65536 (3, 1.014 s, 2.2e-5) - the second-order
// phase takes 6 rounds in each case; from 65536 ulps on its result moves.
constexpr float kFp32SettleUlps = 4096;  // TOMOSTAR_FP32_SETTLE_ULPS overrides
// The mixed form's first-order phase wakes a node's readers only when it dropped by more than this many
// FP32 ulps of its time (1.9e-6 of it): smaller drops of a phase that only seeds the second-order phase
// are left to that phase, which starts with every node dirty. Measured on 400 fields of
// 201 x 201 x 101 (solve seconds, first-order phase seconds, largest and RMS relative distance to the
// FP64 form over 35 fields): 0 ulps (0.799, 0.226, 9.5e-6, 6.3e-9), 16 (0.755, 0.183, 1.6e-5, 1.2e-8),
// 32 (0.749, 0.176, 1.8e-5, 2.2e-8), 64 (0.741, 0.168, 2.4e-5, 4.4e-8), 128 (0.735, 0.163, 6.5e-5, 9.7e-8).
constexpr float kFp32WakeUlps = 16;  // TOMOSTAR_FP32_WAKE_ULPS overrides

inline int sweep_round_cap(const Grid &g)  // SphericalEikonal.cs:335
{
    int c = (g.nx + g.ny + g.nz) / 4;
    return c > 32 ? c : 32;
}

// ------------------------------------------------------------------------------- host geometry

// GridMetric (SphericalEikonal.cs:17-29)
void grid_metric(const Grid &g, double *hlon, double *hlat, double *hr)
{
    for (int k = 0; k < g.nz; k++) {
        double r = kEarthRadiusKm - g.depth(k);
        hlat[k] = r * g.dlat_rad;
        for (int j = 0; j < g.ny; j++) {
            double v = r * cr::cos(g.lat(j) * kDeg2Rad) * g.dlon_rad;
            hlon[k * g.ny + j] = v > 1e-9 ? v : 1e-9;
        }
    }
    *hr = g.ddepth;
}

static double interpolate(const Grid &g, const double *field, double lon, double lat, double depth)
{
    double fx, fy, fz, w[8];
    int nodes[8];
    fractional(g, lon, lat, depth, fx, fy, fz);
    trilinear_weights(g, fx, fy, fz, nodes, w);
    double sum = 0;
    for (int n = 0; n < 8; n++) sum += field[nodes[n]] * w[n];
    return sum;
}

// SphericalEikonal.PointSource (SphericalEikonal.cs:68-100)
int point_source(const Grid &g, const double *slowness, double lon, double lat, double depth, double origin,
                 double radius_cells, int32_t *idx_out, double *t_out, int cap)
{
    double dep = dclamp(depth, g.depth(0), g.depth_last);
    double fx, fy, fz;
    fractional(g, lon, lat, dep, fx, fy, fz);
    fx = dclamp(fx, 0, g.nx - 1);
    fy = dclamp(fy, 0, g.ny - 1);
    fz = dclamp(fz, 0, g.nz - 1);
    Vec3 src = to_cartesian(lon, lat, dep);
    double s_src = interpolate(g, slowness, lon, lat, dep);
    double r = kEarthRadiusKm - dep;
    double h_lon = r * cr::cos(lat * kDeg2Rad) * g.dlon_rad;
    if (!(h_lon > 1e-9)) h_lon = 1e-9;
    double h_lat = r * g.dlat_rad, h_r = g.ddepth;
    double radius = radius_cells * std::max(h_lon, std::max(h_lat, h_r));
    int ri = (int)ceil(radius / h_lon), rj = (int)ceil(radius / h_lat), rk = (int)ceil(radius / h_r);
    int i0 = (int)nearbyint(fx), j0 = (int)nearbyint(fy), k0 = (int)nearbyint(fz);  // Math.Round: half to even
    int count = 0;
    for (int k = std::max(0, k0 - rk); k <= std::min(g.nz - 1, k0 + rk); k++)
        for (int j = std::max(0, j0 - rj); j <= std::min(g.ny - 1, j0 + rj); j++)
            for (int i = std::max(0, i0 - ri); i <= std::min(g.nx - 1, i0 + ri); i++) {
                int idx = g.index(i, j, k);
                double d = vlen(vsub(to_cartesian(g.lon(i), g.lat(j), g.depth(k)), src));
                bool inside = fabs(i - fx) <= 1 && fabs(j - fy) <= 1 && fabs(k - fz) <= 1;
                if (d > radius && !inside) continue;
                if (count < cap) {
                    idx_out[count] = idx;
                    t_out[count] = origin + d * 0.5 * (s_src + slowness[idx]);
                }
                count++;
            }
    return count;
}

// ------------------------------------------------------------------------------- the update

// ------------------------------------------------------------------------------- plain form

// One diagonal plane of one sweep ordering for a batch of fields of one slowness (the reference's
// launch shape: item (a*batch + e, c) is i' = i0 + a, k' = k0 + c of field e, node-major fields).
__global__ __launch_bounds__(64) void eik_plain_plane(double *__restrict__ f, const double *__restrict__ slow,
                                                      const double *__restrict__ hlon,
                                                      const double *__restrict__ hlat, double hr,
                                                      const uint8_t *__restrict__ seeded, int nx, int ny, int nz,
                                                      int second, int *__restrict__ changed, int dir, int p,
                                                      int batch, uint8_t *__restrict__ dirty)
{
    const int i0 = max(0, p - (ny - 1) - (nz - 1));
    const int iw = min(nx - 1, p) - i0 + 1;
    const int k0 = max(0, p - (nx - 1) - (ny - 1));
    const int t = blockIdx.x * 64 + threadIdx.x;
    const int a = t / batch;
    if (a >= iw) return;
    const int e = t - a * batch;
    const int ip = i0 + a, kp = k0 + blockIdx.y, jp = p - ip - kp;
    if (jp < 0 || jp >= ny) return;
    const int i = (dir & 1) ? nx - 1 - ip : ip;
    const int j = (dir & 2) ? ny - 1 - jp : jp;
    const int k = (dir & 4) ? nz - 1 - kp : kp;
    const int nxy = nx * ny;
    const int node = (k * ny + j) * nx + i;
    const size_t b = (size_t)batch;
    const size_t q = (size_t)node * b + e;
    if (seeded[q] || !dirty[q]) return;
    dirty[q] = 0;
    const double s = slow[node];
    if (!(s > 0.0)) return;
    const double *fe = f + e;
    const double cur = fe[(size_t)node * b];
    const int pos[3] = {i, j, k}, n[3] = {nx, ny, nz}, st[3] = {1, nx, nxy};
    const double h[3] = {hlon[k * ny + j], hlat[k], hr};
    double first[3], t2[3];
#pragma unroll
    for (int ax = 0; ax < 3; ax++) {
        double mn = pos[ax] > 0 ? fe[(size_t)(node - st[ax]) * b] : INFINITY;
        double pl = pos[ax] + 1 < n[ax] ? fe[(size_t)(node + st[ax]) * b] : INFINITY;
        int sec;
        if (mn <= pl) { first[ax] = mn; sec = pos[ax] > 1 ? node - 2 * st[ax] : -1; }
        else { first[ax] = pl; sec = pos[ax] + 2 < n[ax] ? node + 2 * st[ax] : -1; }
        t2[ax] = (second && sec >= 0 && isfinite(first[ax])) ? fe[(size_t)sec * b] : INFINITY;
    }
    const double cand = godunov_abs<double>(s, first, t2, h);
    if (!(cand < cur)) return;
    f[q] = cand;
    if (!isfinite(cur) || cur - cand > kSettleTol) atomicOr(changed, 1);
    uint8_t *d = dirty + e;
#pragma unroll
    for (int ax = 0; ax < 3; ax++) {
        if (pos[ax] > 0) d[(size_t)(node - st[ax]) * b] = 1;
        if (pos[ax] > 1) d[(size_t)(node - 2 * st[ax]) * b] = 1;
        if (pos[ax] + 1 < n[ax]) d[(size_t)(node + st[ax]) * b] = 1;
        if (pos[ax] + 2 < n[ax]) d[(size_t)(node + 2 * st[ax]) * b] = 1;
    }
}

__global__ void eik_plain_clear(double *f, uint8_t *seeded, uint8_t *dirty, size_t total)
{
    size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= total) return;
    f[q] = INFINITY;
    seeded[q] = 0;
    dirty[q] = 1;
}

__global__ void eik_fill_u8(uint8_t *p, size_t total, uint8_t v)
{
    size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < total) p[q] = v;
}

__global__ void eik_plain_seed(double *f, uint8_t *seeded, const int64_t *at, const double *t, int count)
{
    int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= count) return;
    f[at[q]] = t[q];
    seeded[at[q]] = 1;
}

// node-major batch (x*batch + e) -> field-major FP64 and FP32 tables
__global__ void eik_plain_unpack(const double *f, int batch, int N, double *out64, float *out32)
{
    size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= (size_t)batch * N) return;
    int e = (int)(q / N);
    size_t node = q - (size_t)e * N;
    double v = f[node * batch + e];
    if (out64) out64[q] = v;
    if (out32) out32[q] = isfinite(v) ? (float)v : INFINITY;
}

// ------------------------------------------------------------------------------- fast form

template <typename R>
struct FastArgs {
    R *T;                         // [groups][N][64]
    uint8_t *dirty;               // [groups][N]: some node this one reads moved since its last update
    const uint64_t *seedmask;     // [groups][N]: lanes whose field holds this node fixed
    const R *slow;                // [models][N]
    const int32_t *model;         // [groups*64]
    const R *wlon, *wlat;         // spacing weights (first, second order): [ny*nz][2], [nz][2]
    R wr0, wr1;                   // and of the depth axis
    const uint64_t *active;       // [groups]: lanes still relaxing
    const uint64_t *phase1;       // [groups]: lanes in the second-order phase
    uint64_t *changed;            // [groups]: lanes that moved a node by more than the settle tolerance
    unsigned long long *evals;    // [1024] spread counters of executed node updates
    const uint32_t *planes;       // nodes of the plane family, packed i' | j' << 11 | k' << 22
    double wake_tol;              // a drop wakes the node's readers when larger than this (0: any drop)
    float settle_ulps;            // FP32 times: a node moved when it dropped by more than this many ulps
    float wake_ulps;              // FP32 times: a drop wakes the node's readers when larger than this many ulps
    int nx, ny, nz, N;
    int narrow;                   // the group of at most 16 fields whose waves update four nodes (-1: none)
};

// A load at a uniform base plus a 32-bit byte offset (one address register per lane).
template <typename T>
__device__ inline T at_off(const T *base, uint32_t byte_off)
{
    return *(const T *)((const char *)base + byte_off);
}

// How an update treats the node's dirty flag: read it with every other load of the update (the seed
// mask, the slowness, the node and its first and second neighbours on each axis), so a wave waits for
// memory once; read it first, so a clean node costs one load; or know it is set (checked beforehand).
enum FlagRead { kFlagWithLoads = 0, kFlagFirst = 1, kFlagKnown = 2 };

// Waves per SIMD the sparse sweep kernel is held to (1: no hold).
#ifndef TS_SPARSE_WAVES32
#define TS_SPARSE_WAVES32 7
#endif
#ifndef TS_SPARSE_WAVES64
#define TS_SPARSE_WAVES64 1
#endif

// Grid node of packed plane entry pk (i' | j' << 11 | k' << 22 in the ordering's mirrored coordinates),
// and its unmirrored coordinates.
template <typename R>
__device__ inline int plane_node(const FastArgs<R> &A, int dir, uint32_t pk, int &i, int &j, int &k)
{
    const int ip = pk & 2047, jp = (pk >> 11) & 2047, kp = pk >> 22;
    i = (dir & 1) ? A.nx - 1 - ip : ip;
    j = (dir & 2) ? A.ny - 1 - jp : jp;
    k = (dir & 4) ? A.nz - 1 - kp : kp;
    return (k * A.ny + j) * A.nx + i;
}
template <typename R>
__device__ inline int plane_node(const FastArgs<R> &A, int dir, uint32_t pk)
{
    int i, j, k;
    return plane_node(A, dir, pk, i, j, k);
}

// The node a lane wakes when the update moves (lanes 0-11): lane 4a + w reads axis a (0 x, 1 y, 2 z)
// at offset -1, -2, +1, +2 nodes for w = 0, 1, 2, 3. The offset in nodes, fixed per lane for a kernel.
__device__ inline int wake_offset(int lane, int nx, int nxy)
{
    const int ax = (lane >> 2) & 3, w = lane & 3;
    const int d = w == 0 ? -1 : w == 1 ? -2 : w == 2 ? 1 : 2;
    return d * (ax == 0 ? 1 : ax == 1 ? nx : nxy);
}

// What the update of one node reads: its coordinates (wave-uniform), this lane's slowness and time,
// and the times of its first and second neighbours on each axis (+inf off the grid, and for the second
// neighbours when no field of the wave is in its second-order phase).
template <typename R>
struct NodeLoad {
    int node, i, j, k;
    uint8_t fl;
    uint64_t fixed;
    R s, cur, mn[3], pl[3], m2[3], p2[3];
    R wlon0, wlon1, wlat0, wlat1;  // spacing weights of the x and y axes (first, second order)
};

// Plane entry q, for a wave-uniform q: a scalar load (the plane lists are never written by a sweep, so
// the constant address space, read through the scalar cache, holds them).
template <typename R>
__device__ __forceinline__ uint32_t plane_entry(const FastArgs<R> &A, int q)
{
    return ((const __attribute__((address_space(4))) uint32_t *)A.planes)[q];
}

// The lane whose field's times a lane loads: its own while its field relaxes this round, else the
// first such lane of the wave. A lane whose field is settled, done or unused never stores what it
// loads (node_update returns its current time with on false), so its loads only cost HBM rows: sent
// to a row the wave fetches anyway, they add no cache lines (a group of 16 fields fetches 16 fields'
// lines of each row rather than 64). act != 0 (the kernels return on an idle group).
__device__ __forceinline__ int load_lane(uint64_t act, int lane)
{
    return ((act >> lane) & 1) ? lane : __builtin_ctzll(act);
}

// The loads of the update of plane node pk for the 64 fields of group g (one wavefront, one lane per
// field; lane: the field whose times this lane loads, load_lane). No node of a plane reads another
// node of the same plane (neighbours lie one or two planes away), so the loads of several nodes of a
// plane may be issued before any of them is updated.
template <typename R, bool kOff32, int kFlag, bool kSecond, bool kBranchLoads = (sizeof(R) == 8)>
__device__ __forceinline__ void load_node(const FastArgs<R> &A, int dir, uint32_t pk, int g, int lane, int mdl,
                                 NodeLoad<R> &L)
{
    const int nx = A.nx, ny = A.ny, nz = A.nz, nxy = nx * ny;
    // every lane of the wave updates the same node (one field per lane): the node's coordinates and
    // addresses are wave-uniform, held in scalar registers
    pk = __builtin_amdgcn_readfirstlane(pk);
    const int node = plane_node(A, dir, pk, L.i, L.j, L.k);
    L.node = node;
    const size_t gn = (size_t)g * A.N;
    L.fl = 1;
    if constexpr (kFlag != kFlagKnown) L.fl = A.dirty[gn + node];
    if (kFlag == kFlagFirst && L.fl == 0) return;  // nothing this node reads has moved: the update would return its value
    L.fixed = A.seedmask[gn + node];
    const R *Tg = A.T + gn * kLanes;  // this group's times (uniform base)
    auto T = [&](int q) -> R {        // time of node q for this lane's field
        if constexpr (kOff32) return at_off(Tg, ((uint32_t)q * kLanes + lane) * (uint32_t)sizeof(R));
        else return Tg[(size_t)q * kLanes + lane];
    };
    if constexpr (kOff32) L.s = at_off(A.slow, ((uint32_t)mdl * (uint32_t)A.N + node) * (uint32_t)sizeof(R));
    else L.s = A.slow[(size_t)mdl * A.N + node];
    L.cur = T(node);
    // the weights with the node's other loads: a wave waits for its loads in the order they were issued
    L.wlon0 = A.wlon[2 * (L.k * ny + L.j)];
    L.wlon1 = A.wlon[2 * (L.k * ny + L.j) + 1];
    L.wlat0 = A.wlat[2 * L.k];
    L.wlat1 = A.wlat[2 * L.k + 1];
    const int pos[3] = {L.i, L.j, L.k}, n[3] = {nx, ny, nz}, st[3] = {1, nx, nxy};
    if constexpr (kBranchLoads) {
        // one node per wave at a time (FP64): loads off the grid are skipped
#pragma unroll
        for (int ax = 0; ax < 3; ax++) {
            L.mn[ax] = pos[ax] > 0 ? T(node - st[ax]) : (R)INFINITY;
            L.pl[ax] = pos[ax] + 1 < n[ax] ? T(node + st[ax]) : (R)INFINITY;
            L.m2[ax] = kSecond && pos[ax] > 1 ? T(node - 2 * st[ax]) : (R)INFINITY;
            L.p2[ax] = kSecond && pos[ax] + 2 < n[ax] ? T(node + 2 * st[ax]) : (R)INFINITY;
        }
        return;
    }
    // Two nodes per wave at a time (FP32): every load is issued, off the grid from the node itself, and
    // replaced by +inf afterwards: with no load behind a branch the compiler counts the loads in flight
    // exactly, so the update of one node waits for its own loads only, not for those of the next node
    // already issued.
#pragma unroll
    for (int ax = 0; ax < 3; ax++) {
        const bool hm = pos[ax] > 0, hp = pos[ax] + 1 < n[ax];
        const R vm = T(hm ? node - st[ax] : node), vp = T(hp ? node + st[ax] : node);
        L.mn[ax] = hm ? vm : (R)INFINITY;
        L.pl[ax] = hp ? vp : (R)INFINITY;
        if constexpr (kSecond) {
            const bool hm2 = pos[ax] > 1, hp2 = pos[ax] + 2 < n[ax];
            const R vm2 = T(hm2 ? node - 2 * st[ax] : node), vp2 = T(hp2 ? node + 2 * st[ax] : node);
            L.m2[ax] = hm2 ? vm2 : (R)INFINITY;
            L.p2[ax] = hp2 ? vp2 : (R)INFINITY;
        } else {
            L.m2[ax] = L.p2[ax] = (R)INFINITY;
        }
    }
}

// The new time of the node L describes for this lane's field, from what load_node read: Relax's
// arithmetic (FP64: bit for bit; FP32: the correctly rounded offset form) or, with kLean, the mixed
// form's (godunov_lean). on: the lane's field relaxes and does not hold the node fixed (else the
// result is the current time).
template <typename R, bool kLean, bool kSecond>
__device__ __forceinline__ R node_update(const FastArgs<R> &A, const NodeLoad<R> &L, uint64_t act, uint64_t ph,
                                         int lane, bool &on)
{
    constexpr bool any_second = kSecond;  // some field of the wave in its second-order phase
    const R wa[3][2] = {{L.wlon0, L.wlon1}, {L.wlat0, L.wlat1}, {A.wr0, A.wr1}};
    on = ((act & ~L.fixed) >> lane) & 1;
    const R s = L.s;
    if (!on) return L.cur;
    const bool second = (ph >> lane) & 1;
    R first[3];
    bool minus[3];
#pragma unroll
    for (int ax = 0; ax < 3; ax++) {
        // the smaller neighbour of the axis
        minus[ax] = L.mn[ax] <= L.pl[ax];
        first[ax] = minus[ax] ? L.mn[ax] : L.pl[ax];
    }
    // second order: the next node beyond the smaller neighbour
    auto beyond = [&](R t2[3]) {
#pragma unroll
        for (int ax = 0; ax < 3; ax++)
            t2[ax] = (second && isfinite(first[ax])) ? (minus[ax] ? L.m2[ax] : L.p2[ax]) : (R)INFINITY;
    };
    R next, t2[3];
    if constexpr (kLean && sizeof(R) == 4) {
        beyond(t2);
        next = (s > 0.0f) ? godunov_lean32(s, first, t2, wa) : (R)INFINITY;
    } else if constexpr (kLean) {
        beyond(t2);
        next = (s > 0.0) ? godunov_lean64(s, first, t2, wa) : (R)INFINITY;
    } else if constexpr (sizeof(R) == 4) {
        // no field of the wave in its second-order phase (wave-uniform): the first-order arithmetic
        if (!any_second) next = (s > 0.0f) ? godunov_offset_first(s, first, wa) : (R)INFINITY;
        else {
            beyond(t2);
            next = (s > 0.0f) ? godunov_offset(s, first, t2, wa) : (R)INFINITY;
        }
    } else {
        beyond(t2);
        next = (s > 0.0) ? godunov_abs_w<double>(s, first, t2, wa) : (R)INFINITY;
    }
    return next;
}

// A lane's drop of the node's time: `dropped` wakes its readers (more than the wake tolerance),
// `moved` counts for the settle test (more than the settle tolerance; FP32 times: and more than the
// settle threshold in ulps of the time).
template <typename R>
__device__ __forceinline__ void drop_tests(const FastArgs<R> &A, R cur, R next, bool &dropped, bool &moved)
{
    if constexpr (sizeof(R) == 4)
        dropped = !isfinite(cur) || (cur - next) > fmaxf((float)A.wake_tol, A.wake_ulps * 1.1920929e-7f * cur);
    else
        dropped = !isfinite(cur) || (double)(cur - next) > A.wake_tol;
    if constexpr (sizeof(R) == 4)
        moved = !isfinite(cur) || (cur - next) > fmaxf((float)kSettleTol, A.settle_ulps * 1.1920929e-7f * cur);
    else
        moved = !isfinite(cur) || (double)(cur - next) > kSettleTol;
}

// Raise the dirty flags of the nodes that read node (i, j, k) of group g (the wave's fields dropped
// it): one and two steps on each axis (lanes 0-11, wake_offset) when a field of the group is in its
// second-order phase, else one step - a first-order update reads no node two steps away, and a field
// entering its second-order phase wakes every node. Bit 4a + w of the wave-uniform mask: the lane's
// node lies in the grid.
template <typename R, bool kSecond>
__device__ __forceinline__ void wake_readers(const FastArgs<R> &A, size_t gn, int node, int i, int j, int k,
                                             int lane, int wake_off)
{
    const int pos[3] = {i, j, k}, n[3] = {A.nx, A.ny, A.nz};
    uint32_t in = 0;
#pragma unroll
    for (int ax = 0; ax < 3; ax++)
        in |= (uint32_t)((pos[ax] >= 1) | (pos[ax] >= 2) << 1 | (pos[ax] + 1 < n[ax]) << 2 |
                         (pos[ax] + 2 < n[ax]) << 3) << (4 * ax);
    if (!kSecond) in &= 0x555u;
    if (((uint64_t)in >> lane) & 1) A.dirty[gn + node + wake_off] = 1;
}

// The update itself, from what load_node read: the new time, and the node's readers woken when it
// dropped by more than the wake tolerance. FP32 times: returns the lanes whose field moved the node by
// more than the settle tolerance (the caller gathers them for the group's changed mask).
// kLean: the mixed form's arithmetic (godunov_lean) instead of the reference's (FP64: Relax's, bit for
// bit; FP32: the correctly rounded offset form).
template <typename R, bool kCount, bool kLean, bool kSecond>
__device__ __forceinline__ uint64_t finish_node(const FastArgs<R> &A, const NodeLoad<R> &L, int g, uint64_t act,
                                       uint64_t ph, int lane, unsigned counter, int wake_off)
{
    if (L.fl == 0) return 0;
    const int node = L.node;
    const size_t gn = (size_t)g * A.N;
    if (lane == 0) A.dirty[gn + node] = 0;
    if (kCount && lane == 0) atomicAdd(&A.evals[counter & 1023], (unsigned long long)__popcll(act & ~L.fixed));
    bool on, dropped = false, moved = false;
    const R next = node_update<R, kLean, kSecond>(A, L, act, ph, lane, on);
    if (on && next < L.cur) {
        A.T[(gn + node) * kLanes + lane] = next;
        drop_tests(A, L.cur, next, dropped, moved);
    }
    if (__ballot(dropped)) wake_readers<R, kSecond>(A, gn, node, L.i, L.j, L.k, lane, wake_off);
    const uint64_t mv = __ballot(moved);
    if constexpr (sizeof(R) == 8) {
        // FP64 times: into the group's changed mask at once, when the mask (read at the device's scope)
        // lacks them - measured faster here than gathering them per wave
        if (mv && lane == 0) {
            const uint64_t seen = __hip_atomic_load(&A.changed[g], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (mv & ~seen) atomicOr((unsigned long long *)&A.changed[g], (unsigned long long)mv);
        }
        return 0;
    }
    return mv;
}

// The update of a streaming sweep's node: as finish_node, with the new time returned (the current
// one when no lane lowers it) for the tile's ring, the moved lanes left to the caller, and the lanes
// that woke the node's readers returned. dev (wave-uniform): the node lies within two columns of the
// tile's edge, so its time, flag and wakes are stored device-coherent (tiles beside it read or
// write them within the launch).
template <typename R, bool kLean, bool kSecond>
__device__ __forceinline__ uint64_t stream_update(const FastArgs<R> &A, const NodeLoad<R> &L, int g, uint64_t act,
                                                  uint64_t ph, int lane, int wake_off, bool dev, R &next, bool &moved)
{
    const size_t gn = (size_t)g * A.N;
    if (lane == 0) {
        if (dev) st_dev(&A.dirty[gn + L.node], (uint8_t)0);
        else A.dirty[gn + L.node] = 0;
    }
    bool on, dropped = false;
    moved = false;
    next = node_update<R, kLean, kSecond>(A, L, act, ph, lane, on);
    if (on && next < L.cur) {
        R *at = &A.T[(gn + L.node) * kLanes + lane];
        if (dev) st_dev(at, next);
        else *at = next;
        drop_tests(A, L.cur, next, dropped, moved);
    } else {
        next = L.cur;
    }
    const uint64_t woke = __ballot(dropped);
    if (woke) {
        const int pos[3] = {L.i, L.j, L.k}, n[3] = {A.nx, A.ny, A.nz};
        uint32_t in = 0;
#pragma unroll
        for (int ax = 0; ax < 3; ax++)
            in |= (uint32_t)((pos[ax] >= 1) | (pos[ax] >= 2) << 1 | (pos[ax] + 1 < n[ax]) << 2 |
                             (pos[ax] + 2 < n[ax]) << 3) << (4 * ax);
        if (!kSecond) in &= 0x555u;
        if (((uint64_t)in >> lane) & 1) {
            uint8_t *f = &A.dirty[gn + L.node + wake_off];
            if (dev) st_dev(f, (uint8_t)1);
            else *f = 1;
        }
    }
    return woke;
}

// The lanes a wave saw move, into the group's changed mask: at most one atomic per wave, at its end, and
// none when the mask as this XCD's cache holds it already has them (the mask only gains bits during a
// round, so a stale copy can only ask for a redundant atomic, never skip a needed one).
__device__ inline void mark_changed(uint64_t *changed, uint64_t mv, int lane)
{
    if (lane == 0 && mv) {
        const uint64_t seen = __hip_atomic_load(changed, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        if (mv & ~seen) atomicOr((unsigned long long *)changed, (unsigned long long)mv);
    }
}

// Kernels that only ever see one kind of wave compile that body alone (one body runs faster than two
// in one kernel): the mixed form's FP64 kernels run its second-order phase only (and the second-order
// body is right for any wave, only slower for a first-order one); rounds on FP32 times hold every field
// at the end of its first-order phase (hold in launch_round), so FP32 kernels see first-order waves only.
template <typename R, bool kLean>
constexpr bool kSecondOnly = kLean && sizeof(R) == 8;
template <typename R>
constexpr bool kFirstOnly = sizeof(R) == 4;

}  // namespace ts

#include "eikonal_stream.hpp"

namespace ts {

// A group of at most 16 fields (a table set's last group, when its field count is 1-16 over a multiple
// of 64): a wave updates four plane nodes at once, 16 lanes each (lane = 16 x quarter + field), so the
// group takes a quarter of the waves of a full one. Per lane, the loads, the arithmetic, the tests and
// the flags are those of the one-node-per-wave kernels (finish_node), and no node of a plane reads
// another node of the same plane, so the times are the same to the bit. Plane nodes a0 + quarter,
// a0 + stride + quarter, ... (a0 and stride wave-uniform, multiples of 4). Times and slowness by 32-bit
// byte offsets from uniform bases (the caller checks they fit), one address register per load.
template <typename R, bool kCount, bool kLean, bool kSecond>
__device__ __forceinline__ void narrow_nodes(const FastArgs<R> &A, int dir, int off, int cnt, int g, uint64_t act,
                                             uint64_t ph, int a0, int stride)
{
    const int lane = threadIdx.x & 63, qd = lane >> 4, fl = lane & 15;
    const int nx = A.nx, ny = A.ny, nz = A.nz, nxy = nx * ny;
    const size_t gn = (size_t)g * A.N;
    const R *Tg = A.T + gn * kLanes;
    const uint8_t *dg = A.dirty + gn;
    const int ll = load_lane(act, fl);  // the group's fields are lanes 0-15
    const int mdl = A.model[g * kLanes + ll];
    const int wake_off = wake_offset(fl, nx, nxy);
    uint64_t mv = 0;
    for (; a0 < cnt; a0 += stride) {
        const int a = a0 + qd;
        NodeLoad<R> L;
        L.fl = 0;
        L.node = L.i = L.j = L.k = 0;
        if (a < cnt) {
            L.node = plane_node(A, dir, A.planes[off + a], L.i, L.j, L.k);
            L.fl = dg[L.node];
        }
        bool dropped = false, moved = false;
        if (L.fl) {
            const int node = L.node;
            auto T = [&](int q) -> R { return at_off(Tg, ((uint32_t)q * kLanes + ll) * (uint32_t)sizeof(R)); };
            L.fixed = A.seedmask[gn + node];
            L.s = at_off(A.slow, ((uint32_t)mdl * (uint32_t)A.N + node) * (uint32_t)sizeof(R));
            L.cur = T(node);
            L.wlon0 = A.wlon[2 * (L.k * ny + L.j)];
            L.wlon1 = A.wlon[2 * (L.k * ny + L.j) + 1];
            L.wlat0 = A.wlat[2 * L.k];
            L.wlat1 = A.wlat[2 * L.k + 1];
            const int pos[3] = {L.i, L.j, L.k}, n[3] = {nx, ny, nz}, st[3] = {1, nx, nxy};
#pragma unroll
            for (int ax = 0; ax < 3; ax++) {
                L.mn[ax] = pos[ax] > 0 ? T(node - st[ax]) : (R)INFINITY;
                L.pl[ax] = pos[ax] + 1 < n[ax] ? T(node + st[ax]) : (R)INFINITY;
            }
            // the second neighbour beyond the smaller first one only (node_update reads no other): three
            // values live rather than six
#pragma unroll
            for (int ax = 0; ax < 3; ax++) {
                R t2 = (R)INFINITY;
                if constexpr (kSecond) {
                    if (L.mn[ax] <= L.pl[ax]) { if (pos[ax] > 1) t2 = T(node - 2 * st[ax]); }
                    else if (pos[ax] + 2 < n[ax]) t2 = T(node + 2 * st[ax]);
                }
                L.m2[ax] = L.p2[ax] = t2;
            }
            if (fl == 0) A.dirty[gn + node] = 0;
            if (kCount && fl == 0) atomicAdd(&A.evals[blockIdx.x & 1023], (unsigned long long)__popcll(act & ~L.fixed));
            bool on;
            const R next = node_update<R, kLean, kSecond>(A, L, act, ph, fl, on);
            if (on && next < L.cur) {
                A.T[(gn + node) * kLanes + fl] = next;
                drop_tests(A, L.cur, next, dropped, moved);
            }
        }
        // a quarter whose node dropped for some field wakes the node's readers (its lanes 0-11)
        if ((__ballot(dropped) >> (16 * qd)) & 0xFFFFull) {
            const int pos[3] = {L.i, L.j, L.k}, n[3] = {nx, ny, nz};
            uint32_t in = 0;
#pragma unroll
            for (int ax = 0; ax < 3; ax++)
                in |= (uint32_t)((pos[ax] >= 1) | (pos[ax] >= 2) << 1 | (pos[ax] + 1 < n[ax]) << 2 |
                                 (pos[ax] + 2 < n[ax]) << 3) << (4 * ax);
            if (!kSecond) in &= 0x555u;
            if ((in >> fl) & 1) A.dirty[gn + L.node + wake_off] = 1;
        }
        const uint64_t m = __ballot(moved);
        mv |= (m | m >> 16 | m >> 32 | m >> 48) & 0xFFFFull;
    }
    mark_changed(&A.changed[g], mv, lane);
}

// The narrow group's part of a plane launch (g == A.narrow): its waves take four nodes each, the
// kernel's XCD slices of the plane in steps of four.
template <typename R, bool kCount, bool kLean, int WPB>
__device__ __forceinline__ void narrow_plane(const FastArgs<R> &A, int dir, int off, int cnt, int g, uint64_t act)
{
    const int per_xcd = (gridDim.x + 7) / 8;
    const int lb = (blockIdx.x % 8) * per_xcd + blockIdx.x / 8;
    const int a0 = (lb * WPB + __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6))) * 4;
    const int stride = 8 * per_xcd * WPB * 4;
    const uint64_t ph = A.phase1[g];
    if constexpr (kFirstOnly<R>) narrow_nodes<R, kCount, kLean, false>(A, dir, off, cnt, g, act, ph, a0, stride);
    else if (kSecondOnly<R, kLean> || (ph & act) != 0)
        narrow_nodes<R, kCount, kLean, true>(A, dir, off, cnt, g, act, ph, a0, stride);
    else narrow_nodes<R, kCount, kLean, false>(A, dir, off, cnt, g, act, ph, a0, stride);
}

// The nodes a wave of eik_fast_plane updates: plane nodes a, a + stride, ... (kSecond: some field of
// the group in its second-order phase, so the loads two nodes away are issued).
template <typename R, bool kCount, bool kOff32, int kFlag, bool kLean, bool kSecond>
__device__ __forceinline__ void plane_nodes(const FastArgs<R> &A, int dir, int off, int cnt, int g, uint64_t act,
                                            uint64_t ph, int lane, int a, int stride)
{
    const int wake_off = wake_offset(lane, A.nx, A.nx * A.ny);
    const int ll = load_lane(act, lane);
    const int mdl = A.model[g * kLanes + ll];
    uint64_t mv = 0;
    for (; a < cnt; a += stride) {
        NodeLoad<R> L;
        load_node<R, kOff32, kFlag, kSecond, false>(A, dir, plane_entry(A, off + a), g, ll, mdl, L);
        mv |= finish_node<R, kCount, kLean, kSecond>(A, L, g, act, ph, lane, blockIdx.x, wake_off);
    }
    mark_changed(&A.changed[g], mv, lane);
}

// One diagonal plane of one sweep ordering for the groups g0 .. g0 + gridDim.y - 1: each wavefront
// updates plane nodes a, a + W, a + 2W, ... (W the waves of a group's row of blocks), so a launch
// is a few resident waves per SIMD rather than one short-lived workgroup per node. kFlag: the dirty
// flag read first (rounds where most nodes are clean) or with every other load (where most are not).
// The all-loads variant is held to 8 waves per SIMD (its FP64 forms otherwise take 101 scalar registers
// and 7 waves; held, 78 and no spills): the round it runs, where every node is dirty, is latency-bound.
template <typename R, int WPB, bool kCount, bool kOff32, int kFlag, bool kLean, bool kNarrow = false>
__global__ __launch_bounds__(64 * WPB) __attribute__((amdgpu_waves_per_eu(kFlag == kFlagWithLoads ? 8 : 1))) void eik_fast_plane(FastArgs<R> A, int dir, int off, int cnt, int g0)
{
    const int lane = threadIdx.x & 63;
    const int g = g0 + blockIdx.y;
    const uint64_t act = A.active[g];
    if (act == 0) return;
    if constexpr (kNarrow && kOff32) {
        if (g == A.narrow) {
            narrow_plane<R, kCount, kLean, WPB>(A, dir, off, cnt, g, act);
            return;
        }
    }
    // Workgroups go to the eight XCDs in turn (block b to XCD b mod 8); block b takes the b/8-th slice
    // of XCD (b mod 8)'s contiguous share of the plane, so the nodes one XCD updates are neighbours and
    // the rows their stencils share meet in that XCD's L2.
    const int per_xcd = (gridDim.x + 7) / 8;
    const int lb = (blockIdx.x % 8) * per_xcd + blockIdx.x / 8;
    const int stride = 8 * per_xcd * WPB;
    const int a = lb * WPB + __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));  // wave-uniform: scalar loads of the plane entry
    const uint64_t ph = A.phase1[g];
    // The mixed form's FP64 kernels (kLean) run its second-order phase only: one body, the kernel's code
    // half the size (the second-order body is right for any wave, only slower for a first-order one).
    if constexpr (kFirstOnly<R>) plane_nodes<R, kCount, kOff32, kFlag, kLean, false>(A, dir, off, cnt, g, act, ph, lane, a, stride);
    else if (kSecondOnly<R, kLean> || (ph & act) != 0)
        plane_nodes<R, kCount, kOff32, kFlag, kLean, true>(A, dir, off, cnt, g, act, ph, lane, a, stride);
    else plane_nodes<R, kCount, kOff32, kFlag, kLean, false>(A, dir, off, cnt, g, act, ph, lane, a, stride);
}

// The dirty nodes of a wave of eik_fast_plane_sparse (todo: the lanes holding them in pk), in plane order,
// K at a time: the loads of K nodes are issued before the first of them is updated, so the later nodes'
// loads are in flight while the earlier ones are updated.
template <typename R, bool kCount, bool kOff32, bool kLean, int K, bool kSecond>
__device__ __forceinline__ void sparse_nodes(const FastArgs<R> &A, int dir, int g, uint64_t act, uint64_t ph,
                                             int lane, uint32_t pk, uint64_t todo)
{
    const int wake_off = wake_offset(lane, A.nx, A.nx * A.ny);
    const int ll = load_lane(act, lane);
    const int mdl = A.model[g * kLanes + ll];
    uint64_t mv = 0;
    while (todo) {
        uint32_t p[K];
        bool valid[K];
#pragma unroll
        for (int q = 0; q < K; q++) {
            valid[q] = todo != 0;
            p[q] = __shfl(pk, valid[q] ? __builtin_ctzll(todo) : 0);
            // a missing node loads the first one again rather than branch around its loads
            if (!valid[q]) p[q] = p[0];
            if (valid[q]) todo &= todo - 1;
        }
        NodeLoad<R> L[K];
#pragma unroll
        for (int q = 0; q < K; q++) load_node<R, kOff32, kFlagKnown, kSecond, (K == 1)>(A, dir, p[q], g, ll, mdl, L[q]);
#pragma unroll
        for (int q = 0; q < K; q++)
            if (valid[q]) mv |= finish_node<R, kCount, kLean, kSecond>(A, L[q], g, act, ph, lane, blockIdx.x, wake_off);
    }
    mark_changed(&A.changed[g], mv, lane);
}

// Rounds in which most nodes are clean: each wavefront reads the dirty flags of M consecutive plane
// nodes at once (lanes 0 .. M-1) and then updates its dirty ones in plane order, so a clean stretch of
// the plane costs one load per M nodes and M times fewer waves are launched. FP32: held to 7 waves per
// SIMD (72 vector registers instead of 76 and 6 waves, no spills; 8 waves would spill).
template <typename R, int WPB, int M, bool kCount, bool kOff32, bool kLean, int K, bool kNarrow = false>
__global__ __launch_bounds__(64 * WPB) __attribute__((amdgpu_waves_per_eu(sizeof(R) == 4 ? TS_SPARSE_WAVES32 : kNarrow ? 8 : TS_SPARSE_WAVES64))) void eik_fast_plane_sparse(FastArgs<R> A, int dir, int off, int cnt, int g0)
{
    const int lane = threadIdx.x & 63;
    const int g = g0 + blockIdx.y;
    const uint64_t act = A.active[g];
    if (act == 0) return;
    if constexpr (kNarrow && kOff32) {
        if (g == A.narrow) {
            narrow_plane<R, kCount, kLean, WPB>(A, dir, off, cnt, g, act);
            return;
        }
    }
    const int per_xcd = (gridDim.x + 7) / 8;
    const int lb = (blockIdx.x % 8) * per_xcd + blockIdx.x / 8;
    const int a0 = (lb * WPB + (threadIdx.x >> 6)) * M;
    if (a0 >= cnt) return;
    // lanes 0 .. M-1 read the plane entries and their flags side by side (vector loads: as scalar loads
    // the M nodes' coordinates are worked out one after another, and the sparse rounds took longer)
    uint32_t pk = 0;
    bool dirty = false;
    if (lane < M && a0 + lane < cnt) {
        pk = A.planes[off + a0 + lane];
        dirty = A.dirty[(size_t)g * A.N + plane_node(A, dir, pk)] != 0;
    }
    const uint64_t ph = A.phase1[g];
    const uint64_t todo = __ballot(dirty);
    if constexpr (kFirstOnly<R>) sparse_nodes<R, kCount, kOff32, kLean, K, false>(A, dir, g, act, ph, lane, pk, todo);
    else if (kSecondOnly<R, kLean> || (ph & act) != 0)
        sparse_nodes<R, kCount, kOff32, kLean, K, true>(A, dir, g, act, ph, lane, pk, todo);
    else sparse_nodes<R, kCount, kOff32, kLean, K, false>(A, dir, g, act, ph, lane, pk, todo);
}

// Per-field bookkeeping after a round (one wave per group): a field whose round moved no node by
// more than the tolerance ends its phase (first order -> second order -> done); a field that
// reaches the round cap in a phase is unsettled. state: 0/1 phase, 2 done, 3 unsettled, 4 unused,
// 5 first order settled and held (hold != 0: the second-order phase starts for all fields at once,
// in FP64, once every field has settled its first-order phase in FP32).
__global__ __launch_bounds__(64) void eik_fast_round_end(uint64_t *active, uint64_t *phase1, uint64_t *changed,
                                                         int32_t *state, int32_t *rounds, int cap,
                                                         int32_t *status, uint8_t *reset, int hold, int first_cap)
{
    const int g = blockIdx.x, lane = threadIdx.x, f = g * 64 + lane;
    int st = state[f];
    const bool ch = (changed[g] >> lane) & 1;
    bool need_reset = false;
    if (st < 2) {
        int r = ++rounds[2 * f + st];
        // first_cap (held first-order phases only): the phase ends after that many rounds, settled or not
        if (!ch || (hold && st == 0 && first_cap > 0 && r >= first_cap)) {
            if (st == 0) {
                st = hold ? 5 : 1;
                need_reset = !hold;
            } else st = 2;
        } else if (r >= cap) st = 3;
        state[f] = st;
    }
    const uint64_t act = __ballot(st < 2), ph = __ballot(st == 1), rs = __ballot(need_reset);
    const uint64_t bad = __ballot(st == 3);
    if (lane == 0) {
        active[g] = act;
        phase1[g] = ph;
        changed[g] = 0;
        reset[g] = rs != 0;
        if (act) atomicAdd(&status[0], __popcll(act));
        if (bad) atomicAdd(&status[1], __popcll(bad));
    }
}

// A field entering its second-order phase needs every node re-evaluated (the update changed).
__global__ void eik_fast_reset_dirty(uint8_t *dirty, const uint8_t *reset, int N)
{
    const int g = blockIdx.y;
    if (!reset[g]) return;
    size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < (size_t)N) dirty[(size_t)g * N + q] = 1;
}

// Second-order phase of the mixed form: the FP32 first-order times become the FP64 starting point.
__global__ void eik_mixed_widen(const float *T32, double *T64, size_t n)
{
    size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) T64[q] = (double)T32[q];
}

__global__ __launch_bounds__(64) void eik_mixed_promote(uint64_t *active, uint64_t *phase1, uint64_t *changed,
                                                        int32_t *state)
{
    const int g = blockIdx.x, lane = threadIdx.x, f = g * 64 + lane;
    int st = state[f];
    if (st == 5) st = 1;
    state[f] = st;
    const uint64_t act = __ballot(st == 1);
    if (lane == 0) {
        active[g] = act;
        phase1[g] = act;
        changed[g] = 0;
    }
}

// Every time unreached and every node clean: an update of a node whose neighbours are all unreached
// returns +inf and changes nothing, so only the readers of the seeds need a first visit.
template <typename R>
__global__ void eik_fast_clear(R *T, uint8_t *dirty, uint64_t *seedmask, size_t nodes_total)
{
    size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < nodes_total * kLanes) T[q] = (R)INFINITY;
    if (q < nodes_total) { dirty[q] = 0; seedmask[q] = 0; }
}

// Seed times, seed masks and, when wake is set, the dirty flags of the twelve nodes that read each
// seed (one and two steps on each axis).
template <typename R>
__global__ void eik_fast_seed(R *T, uint64_t *seedmask, uint8_t *dirty, const int32_t *field, const int32_t *node,
                              const double *time, int count, int nx, int ny, int nz, bool wake)
{
    int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= count) return;
    const int N = nx * ny * nz;
    const int f = field[q], g = f / kLanes, lane = f % kLanes, n = node[q];
    const size_t at = (size_t)g * N + n;
    T[at * kLanes + lane] = (R)time[q];
    atomicOr((unsigned long long *)&seedmask[at], 1ull << lane);
    if (!wake) return;
    const int pos[3] = {n % nx, (n / nx) % ny, n / (nx * ny)}, ext[3] = {nx, ny, nz}, st[3] = {1, nx, nx * ny};
    for (int ax = 0; ax < 3; ax++)
        for (int d = -2; d <= 2; d++)
            if (d != 0 && pos[ax] + d >= 0 && pos[ax] + d < ext[ax]) dirty[at + d * st[ax]] = 1;
}

// [group][node][64] -> field-major tables (FP32, and FP64 when requested), through an LDS tile so
// both the read and the write are coalesced.
template <typename R>
__global__ __launch_bounds__(256) void eik_fast_unpack(const R *T, int N, const int32_t *field_of_slot, float *out32,
                                                       double *out64)
{
    __shared__ R tile[64][65];
    const int g = blockIdx.y, n0 = blockIdx.x * 64, t = threadIdx.x;
    const R *Tg = T + (size_t)g * N * kLanes;
#pragma unroll
    for (int r = 0; r < 16; r++) {
        int nn = (t >> 6) + r * 4, node = n0 + nn;
        tile[nn][t & 63] = node < N ? Tg[(size_t)node * kLanes + (t & 63)] : (R)INFINITY;
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < 16; r++) {
        int fl = (t >> 6) + r * 4, f = field_of_slot[g * kLanes + fl], node = n0 + (t & 63);
        if (f < 0 || node >= N) continue;
        R v = tile[t & 63][fl];
        size_t at = (size_t)f * N + node;
        if (out32) out32[at] = isfinite(v) ? (float)v : INFINITY;
        if (out64) out64[at] = (double)v;
    }
}

__global__ void eik_f32_to_f64(const float *in, double *out, size_t n)
{
    size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) out[q] = (double)in[q];
}

}  // namespace ts

// ================================================================================== host side

using namespace ts;

struct ts_tableset {
    ts_context *ctx = nullptr;
    Grid grid{};
    int nmodels = 0, nfields = 0, N = 0;
    std::vector<int32_t> model_of_field;
    // fields go to 64-lane groups in the Morton order of their sources, so the fields a wavefront
    // relaxes together have nearby fronts and wake the same nodes
    std::vector<int32_t> slot_of_field, field_of_slot;
    dbuf<int32_t> field_of_slot_dev;
    std::vector<std::vector<int32_t>> seed_node;  // merged seeds of each field (smallest time wins)
    std::vector<std::vector<double>> seed_time;
    std::vector<double> hlon, hlat;
    double hr = 0;
    dbuf<double> slow64, hlon64, hlat64;  // [models][N]
    dbuf<float> slow32, hlon32, hlat32;
    dbuf<double> wlon64, wlat64;          // spacing weights of the fast form: [ny*nz][2], [nz][2]
    dbuf<float> wlon32, wlat32;
    double wr64[2] = {0, 0};
    float wr32[2] = {0, 0};
    dbuf<uint32_t> planes;
    std::vector<int> plane_off;           // nx+ny+nz-1 entries
    // results
    dbuf<float> tab32;                    // [fields][N] stored tables
    dbuf<double> tab64;                   // [fields][N] when the last solve ran in FP64
    bool have64 = false, solved = false;
    bool pending64 = false;               // the fast form left its FP64 times in T64: tab64 on the first FP64 read
    std::unique_ptr<HostStage> stage;     // pinned staging of large reads (copy_to_host)
    std::vector<int32_t> rounds;          // 2 per field
    double evaluations = 0, evaluations32 = 0, evaluations64 = 0, seconds = 0, sweep_seconds = 0;
    double sweep_by_width[2] = {0, 0};  // sweep-graph device seconds on FP32 / FP64 times
    int32_t rounds_by_width[2] = {0, 0};  // rounds run on FP32 / FP64 times
    int32_t plane_launches_per_round = 0;
    std::vector<int32_t> round_kind;      // per round of the last solve: 16 for FP64 times | round order
    std::vector<double> round_seconds;    // and its device seconds
    // fast-form working storage, kept across solves
    dbuf<float> T32;
    dbuf<double> T64;
    dbuf<uint8_t> dirty, reset;
    dbuf<uint64_t> seedmask, active, phase1, changed;
    dbuf<int32_t> model_dev, state, rounds_dev, status;
    dbuf<unsigned long long> evals;
    double wake_tol[2] = {0, 0};          // FastArgs::wake_tol of the running solve's FP32 / FP64 sweeps
    bool lean[2] = {false, false};        // the running solve's FP32 / FP64 sweeps use the mixed form's arithmetic
    float settle_ulps = 4096;             // the running solve's FP32 settle threshold (ulps of the time)
    float wake_ulps = 0;                  // and its FP32 wake threshold (ulps of the time)
    int first_rounds = 0;                 // mixed form: first-order rounds after which a field is held (0: no
                                          // cap; TOMOSTAR_FP32_MAX_ROUNDS)
    // launch-shape knobs, read once when the table set is created (graphs captured with them stay valid)
    int knob_group_chunk = 0, knob_plane_blocks = 0;
    bool narrow = true;   // a last group of 1-16 fields updates four nodes per wave (TOMOSTAR_NARROW=0: not)
    // what a captured graph's kernel arguments hold of the running solve: its tolerances, thresholds
    // and arithmetic
    struct SolveKey {
        double wake_tol = 0;
        float settle_ulps = 0, wake_ulps = 0;
        int first_rounds = 0;
        bool lean = false;
        bool operator==(const SolveKey &o) const
        {
            return wake_tol == o.wake_tol && settle_ulps == o.settle_ulps &&
                   wake_ulps == o.wake_ulps && first_rounds == o.first_rounds && lean == o.lean;
        }
    };
    SolveKey solve_key(int w) const
    {
        return {wake_tol[w], settle_ulps, wake_ulps, first_rounds, lean[w]};
    }
    static constexpr int kRoundOrders = 5;  // how a round reads the dirty flags (round_order)
    hipGraphExec_t graphs[2][2][kRoundOrders][2] = {};  // [FP64][counting][round order][wake_tol > 0]
    SolveKey graph_key[2][2][kRoundOrders][2] = {};     // what each graph was captured with
    // chunks of groups sweeping side by side, each on its own stream with its own graph of plane launches
    static constexpr int kMaxStreams = 8;
    hipStream_t gstream[kMaxStreams] = {};
    hipEvent_t gfork = nullptr, gjoin[kMaxStreams] = {};
    hipGraphExec_t cgraphs[2][2][kRoundOrders][2][kMaxStreams] = {};
    SolveKey cgraph_key[2][2][kRoundOrders][2][kMaxStreams] = {};
    int cgraph_groups[2][2][kRoundOrders][2][kMaxStreams][2] = {};  // the groups [first, first + count) each swept
    // streaming rounds, per width (FP32 / FP64 times): tile shape, the task list of every group's tiles
    // in an order where each tile follows its upwind tiles, and the queues and progress of a round
    struct StreamPlan {
        int tu = 0, tv = 0, ntu = 0, ntv = 0, ntasks = 0, slots = 0, grid = 0;
        dbuf<uint32_t> tasks;
    } splan[2];
    dbuf<int> stream_queue, stream_prog;
    int stream_spin_limit = 1 << 18;
    int stream_shape[2] = {0, 0};         // tile shape per width (with_shape)
    int stream_publish = 1;               // planes a streaming tile completes between two publications
    std::mutex gate;
    ~ts_tableset()
    {
        for (auto &a : graphs)
            for (auto &b : a)
                for (auto &c : b)
                    for (auto &gx : c)
                        if (gx) (void)hipGraphExecDestroy(gx);
        for (auto &a : cgraphs)
            for (auto &b : a)
                for (auto &c : b)
                    for (auto &d : c)
                        for (auto &gx : d)
                            if (gx) (void)hipGraphExecDestroy(gx);
        for (auto &sx : gstream)
            if (sx) (void)hipStreamDestroy(sx);
        for (auto &ex : gjoin)
            if (ex) (void)hipEventDestroy(ex);
        if (gfork) (void)hipEventDestroy(gfork);
    }
};

namespace {

// Nodes of the diagonal planes i'+j'+k' = p of the unmirrored ordering, by plane then k' then i'.
// The same list serves every ordering: the kernel mirrors each coordinate whose bit is set in dir.
void build_planes(ts_tableset &t)
{
    const Grid &g = t.grid;
    const int P = g.nx + g.ny + g.nz - 2;
    std::vector<int> count(P + 1, 0);
    for (int p = 0; p < P; p++) {
        int i0 = std::max(0, p - (g.ny - 1) - (g.nz - 1)), i1 = std::min(g.nx - 1, p);
        int k0 = std::max(0, p - (g.nx - 1) - (g.ny - 1)), k1 = std::min(g.nz - 1, p);
        for (int kp = k0; kp <= k1; kp++) {
            int lo = std::max(i0, p - kp - (g.ny - 1)), hi = std::min(i1, p - kp);
            if (hi >= lo) count[p + 1] += hi - lo + 1;
        }
    }
    for (int p = 0; p < P; p++) count[p + 1] += count[p];
    std::vector<uint32_t> list(count[P]);
    for (int p = 0; p < P; p++) {
        int at = count[p];
        int i0 = std::max(0, p - (g.ny - 1) - (g.nz - 1)), i1 = std::min(g.nx - 1, p);
        int k0 = std::max(0, p - (g.nx - 1) - (g.ny - 1)), k1 = std::min(g.nz - 1, p);
        for (int kp = k0; kp <= k1; kp++) {
            int lo = std::max(i0, p - kp - (g.ny - 1)), hi = std::min(i1, p - kp);
            for (int ip = lo; ip <= hi; ip++) {
                int jp = p - ip - kp;
                list[at++] = (uint32_t)ip | ((uint32_t)jp << 11) | ((uint32_t)kp << 22);
            }
        }
    }
    if ((int64_t)list.size() != (int64_t)g.count()) fail(TS_EDEVICE, "plane list covers %zu of %d nodes", list.size(), g.count());
    t.plane_off = count;
    t.planes.alloc(list.size());
    TS_HIP(hipMemcpy(t.planes.p, list.data(), list.size() * 4, hipMemcpyHostToDevice));
}

inline unsigned blocks_for(size_t n, unsigned bs) { return (unsigned)((n + bs - 1) / bs); }

// EikonalOpenCl.SuggestBatch (EikonalOpenCl.cs:121-132) with this device's memory.
int suggest_batch(const ts_context *ctx, int nodes)
{
    const int64_t real = 8;
    int64_t n = std::max(1, nodes);
    int64_t per_field = n * real + 2 * n;
    int64_t budget = (int64_t)(ctx->global_mem * 0.5) - 2 * n * real;
    int64_t by_alloc = (int64_t)ctx->global_mem / std::max<int64_t>(1, n * real);
    int64_t by_host = (768LL << 20) / std::max<int64_t>(1, n * 8);
    int64_t enough = std::max<int64_t>(4, (128LL << 20) / n);
    int64_t b = std::min(std::min(budget / std::max<int64_t>(1, per_field), by_alloc), std::min(by_host, enough));
    return (int)std::clamp<int64_t>(b, 0, 32);
}

// ---- plain form: the reference's batches, planes and per-round host flag read
void solve_plain(ts_tableset &t, const volatile int32_t *cancel)
{
    ts_context *ctx = t.ctx;
    hipStream_t st = ctx->stream;
    const Grid &g = t.grid;
    const int N = t.N;
    const int cap = sweep_round_cap(g);
    const int planes = g.nx + g.ny + g.nz - 2;
    int B = suggest_batch(ctx, N);
    if (B < 1) fail(TS_DECLINED, "no room for one field of %d nodes", N);
    t.tab64.ensure((size_t)t.nfields * N);
    t.tab32.ensure((size_t)t.nfields * N);
    dbuf<double> f((size_t)B * N);
    dbuf<uint8_t> seeded((size_t)B * N), dirty((size_t)B * N);
    dbuf<int> changed(1);
    int *h_changed = nullptr;
    TS_HIP(hipHostMalloc((void **)&h_changed, sizeof(int), hipHostMallocDefault));
    struct Free { int *p; ~Free() { if (p) (void)hipHostFree(p); } } free_h{h_changed};
    for (int m = 0; m < t.nmodels; m++) {
        std::vector<int> fields;
        for (int q = 0; q < t.nfields; q++)
            if (t.model_of_field[q] == m) fields.push_back(q);
        for (size_t b0 = 0; b0 < fields.size(); b0 += B) {
            const int batch = (int)std::min<size_t>(B, fields.size() - b0);
            const size_t total = (size_t)batch * N;
            std::vector<int64_t> at;
            std::vector<double> tm;
            for (int e = 0; e < batch; e++) {
                int q = fields[b0 + e];
                for (size_t s = 0; s < t.seed_node[q].size(); s++) {
                    at.push_back((int64_t)t.seed_node[q][s] * batch + e);
                    tm.push_back(t.seed_time[q][s]);
                }
            }
            dbuf<int64_t> d_at(at.size());
            dbuf<double> d_tm(tm.size());
            TS_HIP(hipMemcpyAsync(d_at.p, at.data(), at.size() * 8, hipMemcpyHostToDevice, st));
            TS_HIP(hipMemcpyAsync(d_tm.p, tm.data(), tm.size() * 8, hipMemcpyHostToDevice, st));
            eik_plain_clear<<<blocks_for(total, 256), 256, 0, st>>>(f.p, seeded.p, dirty.p, total);
            eik_plain_seed<<<blocks_for(at.size(), 256), 256, 0, st>>>(f.p, seeded.p, d_at.p, d_tm.p, (int)at.size());
            TS_HIP(hipGetLastError());
            for (int phase = 0; phase < 2; phase++) {
                if (phase > 0) eik_fill_u8<<<blocks_for(total, 256), 256, 0, st>>>(dirty.p, total, 1);
                bool settled = false;
                int round;
                for (round = 0; round < cap && !settled; round++) {
                    if (cancelled(cancel)) fail(TS_CANCELLED, "cancelled");
                    TS_HIP(hipMemsetAsync(changed.p, 0, sizeof(int), st));
                    for (int dir = 0; dir < 8; dir++)
                        for (int p = 0; p < planes; p++) {
                            int i0 = std::max(0, p - (g.ny - 1) - (g.nz - 1)), iw = std::min(g.nx - 1, p) - i0 + 1;
                            int k0 = std::max(0, p - (g.nx - 1) - (g.ny - 1)), kw = std::min(g.nz - 1, p) - k0 + 1;
                            dim3 grid(blocks_for((size_t)iw * batch, 64), kw);
                            eik_plain_plane<<<grid, 64, 0, st>>>(f.p, t.slow64.p + (size_t)m * N, t.hlon64.p,
                                                                 t.hlat64.p, t.hr, seeded.p, g.nx, g.ny, g.nz,
                                                                 phase, changed.p, dir, p, batch, dirty.p);
                        }
                    TS_HIP(hipGetLastError());
                    TS_HIP(hipMemcpyAsync(h_changed, changed.p, sizeof(int), hipMemcpyDeviceToHost, st));
                    TS_HIP(hipStreamSynchronize(st));
                    settled = *h_changed == 0;
                }
                for (int e = 0; e < batch; e++) t.rounds[2 * fields[b0 + e] + phase] = round;
                if (!settled) fail(TS_DECLINED, "sweeping not settled after %d rounds", cap);
            }
            // field-major results: the batch's fields are not contiguous in the table set in general
            dbuf<double> o64(total);
            dbuf<float> o32(total);
            eik_plain_unpack<<<blocks_for(total, 256), 256, 0, st>>>(f.p, batch, N, o64.p, o32.p);
            for (int e = 0; e < batch; e++) {
                int q = fields[b0 + e];
                TS_HIP(hipMemcpyAsync(t.tab64.p + (size_t)q * N, o64.p + (size_t)e * N, (size_t)N * 8,
                                      hipMemcpyDeviceToDevice, st));
                TS_HIP(hipMemcpyAsync(t.tab32.p + (size_t)q * N, o32.p + (size_t)e * N, (size_t)N * 4,
                                      hipMemcpyDeviceToDevice, st));
            }
            TS_HIP(hipStreamSynchronize(st));
        }
    }
    t.have64 = true;
    t.pending64 = false;
}

// ---- fast form
template <typename R>
FastArgs<R> fast_args(ts_tableset &t, R *T)
{
    FastArgs<R> A;
    A.T = T;
    A.dirty = t.dirty.p;
    A.seedmask = t.seedmask.p;
    if constexpr (sizeof(R) == 4) {
        A.slow = t.slow32.p; A.wlon = t.wlon32.p; A.wlat = t.wlat32.p; A.wr0 = t.wr32[0]; A.wr1 = t.wr32[1];
    } else {
        A.slow = t.slow64.p; A.wlon = t.wlon64.p; A.wlat = t.wlat64.p; A.wr0 = t.wr64[0]; A.wr1 = t.wr64[1];
    }
    A.model = t.model_dev.p;
    A.active = t.active.p;
    A.phase1 = t.phase1.p;
    A.changed = t.changed.p;
    A.evals = t.evals.p + (sizeof(R) == 8 ? 1024 : 0);  // FP32 and FP64 updates counted apart
    A.planes = t.planes.p;
    A.wake_tol = t.wake_tol[sizeof(R) == 8];
    A.settle_ulps = t.settle_ulps;
    A.wake_ulps = t.wake_ulps;
    A.nx = t.grid.nx; A.ny = t.grid.ny; A.nz = t.grid.nz; A.N = t.N;
    const int rest = t.nfields % kLanes;
    A.narrow = t.narrow && rest >= 1 && rest <= 16 ? t.nfields / kLanes : -1;
    return A;
}

#ifndef TS_WPB
#define TS_WPB 2
#endif
constexpr int kWPB = TS_WPB;  // waves per block of the plane kernels
constexpr int kSparse = 2;  // round order after kFlagWithLoads, kFlagFirst: kSparseNodes flags per wave, then the dirty nodes
constexpr int kStream = 3;  // the streaming sweep: one persistent launch per ordering over column tiles
constexpr int kSparseWide = 4;  // as kSparse with kWideNodes flags per wave: the late rounds, few nodes dirty
static_assert(kSparseWide < ts_tableset::kRoundOrders, "a graph slot per round order");
constexpr int kStreamShapes = 2;  // tile shapes per width (with_shape)
#ifndef TS_SPARSE_NODES32
#define TS_SPARSE_NODES32 4
#endif
#ifndef TS_INFLIGHT32
#define TS_INFLIGHT32 4
#endif
#ifndef TS_SPARSE_NODES64
#define TS_SPARSE_NODES64 2
#endif
#ifndef TS_INFLIGHT64
#define TS_INFLIGHT64 1
#endif
// Flags read per wave of a sparse launch (M) and dirty nodes whose loads are in flight together (K): FP32
// waves carry the loads of K first-order nodes within the registers of full occupancy, FP64 waves (twelve
// FP64 neighbours) one.
template <typename R>
constexpr int kSparseNodes = sizeof(R) == 4 ? TS_SPARSE_NODES32 : TS_SPARSE_NODES64;
template <typename R>
constexpr int kInFlight = sizeof(R) == 4 ? TS_INFLIGHT32 : TS_INFLIGHT64;
#ifndef TS_WIDE_NODES32
#define TS_WIDE_NODES32 8
#endif
#ifndef TS_WIDE_NODES64
#define TS_WIDE_NODES64 4
#endif
template <typename R>
constexpr int kWideNodes = sizeof(R) == 4 ? TS_WIDE_NODES32 : TS_WIDE_NODES64;

// Cap on the blocks of a plane launch over all its groups (TOMOSTAR_PLANE_BLOCKS): waves stride over
// the plane beyond it. Measured on the 400-field case: fewer, longer-lived waves are slower (the
// sweep is bound by the memory requests in flight), so by default every plane node gets its own wave.
int plane_blocks(const ts_tableset &t) { return t.knob_plane_blocks > 0 ? t.knob_plane_blocks : 1 << 30; }

// Groups swept together: TOMOSTAR_GROUP_CHUNK when set, else all of them.
int group_chunk(const ts_tableset &t, size_t width)
{
    (void)width;
    return t.knob_group_chunk > 0 ? t.knob_group_chunk : (t.nfields + kLanes - 1) / kLanes;
}

// The launch-shape knobs of a new table set (TOMOSTAR_GROUP_CHUNK, TOMOSTAR_PLANE_BLOCKS), read once.
void read_launch_knobs(ts_tableset &t)
{
    // polls of a streaming tile's wait for an upwind tile before it raises the error flag (a wait
    // normally ends within a few planes' time; each poll takes about a microsecond)
    if (const char *e = getenv("TOMOSTAR_STREAM_SPIN")) t.stream_spin_limit = std::max(0, atoi(e));
    if (const char *e = getenv("TOMOSTAR_STREAM_SHAPE32")) t.stream_shape[0] = std::clamp(atoi(e), 0, kStreamShapes - 1);
    if (const char *e = getenv("TOMOSTAR_STREAM_SHAPE64")) t.stream_shape[1] = std::clamp(atoi(e), 0, kStreamShapes - 1);
    if (const char *e = getenv("TOMOSTAR_STREAM_PUBLISH")) t.stream_publish = std::clamp(atoi(e), 1, 64);
    if (const char *e = getenv("TOMOSTAR_GROUP_CHUNK")) t.knob_group_chunk = std::max(1, atoi(e));
    if (const char *e = getenv("TOMOSTAR_PLANE_BLOCKS")) t.knob_plane_blocks = std::max(1, atoi(e));
    if (const char *e = getenv("TOMOSTAR_NARROW")) t.narrow = atoi(e) != 0;
}

// The plane launches of groups [g0, g0 + gc): all eight orderings, every plane. A chunk holding the
// narrow group (FastArgs::narrow) launches the kernels that carry its four-nodes-per-wave body, on FP32
// times and in the FP64 rounds that read two flags per wave; the others launch kernels without it (the
// body takes vector registers the full groups' waves would then lack: the FP64 sparse kernel carrying it
// is held to 8 waves per SIMD, 64 registers). Measured on 400 fields (six groups and one of 16) of
// 201 x 201 x 101, the times the same to the bit: first-order rounds 79, 58, 30 ms against 81, 61, 30 ms,
// second-order rounds 1-2 151, 92 ms against 158, 96 ms. In the FP64 all-loads round the body costs the
// chunk's full groups more than it saves (193 ms against 186 ms), but that chunk's launches reading two
// flags per wave with it take the round to 180 ms (the full groups run that round as fast either way):
// the solve 0.689 s against 0.710 s. The late rounds (four flags per wave) keep one node per wave (with
// the body: 17 ms against 14 ms).
template <typename R, bool kCount, int kOrder, bool kLean, bool kNarrow>
void launch_planes_n(ts_tableset &t, const FastArgs<R> &A, R *T, hipStream_t st, int g0, int gc, bool off32)
{
    const int P = t.grid.nx + t.grid.ny + t.grid.nz - 2;
    for (int dir = 0; dir < 8; dir++)
        for (int p = 0; p < P; p++) {
            int off = t.plane_off[p], cnt = t.plane_off[p + 1] - off;
            if constexpr (kOrder == kSparse || kOrder == kSparseWide) {
                constexpr int M = kOrder == kSparse ? kSparseNodes<R> : kWideNodes<R>, K = kInFlight<R>;
                const unsigned nb = blocks_for(cnt, kWPB * M);
                dim3 grid((nb + 7) / 8 * 8, gc);
                if (off32)
                    eik_fast_plane_sparse<R, kWPB, M, kCount, true, kLean, K, kNarrow><<<grid, 64 * kWPB, 0, st>>>(A, dir, off, cnt, g0);
                else
                    eik_fast_plane_sparse<R, kWPB, M, kCount, false, kLean, K><<<grid, 64 * kWPB, 0, st>>>(A, dir, off, cnt, g0);
            } else {
                // a multiple of 8 blocks: each XCD gets the same number of slices
                const unsigned nb = std::min<unsigned>(blocks_for(cnt, kWPB), std::max(1, plane_blocks(t) / gc));
                dim3 grid((nb + 7) / 8 * 8, gc);
                if (off32) eik_fast_plane<R, kWPB, kCount, true, kOrder, kLean, kNarrow><<<grid, 64 * kWPB, 0, st>>>(A, dir, off, cnt, g0);
                else eik_fast_plane<R, kWPB, kCount, false, kOrder, kLean><<<grid, 64 * kWPB, 0, st>>>(A, dir, off, cnt, g0);
            }
        }
}

template <typename R, bool kCount, int kOrder, bool kLean>
void launch_planes(ts_tableset &t, R *T, hipStream_t st, int g0, int gc)
{
    FastArgs<R> A = fast_args<R>(t, T);
    // 32-bit byte offsets when a group's times and all slowness models fit in 4 GB
    const bool off32 = (uint64_t)t.N * kLanes * sizeof(R) < (1ull << 32) &&
                       (uint64_t)t.nmodels * t.N * sizeof(R) < (1ull << 32);
    if constexpr (sizeof(R) == 4 || kOrder == kSparse) {  // where it pays (above)
        if (off32 && A.narrow >= g0 && A.narrow < g0 + gc) {
            launch_planes_n<R, kCount, kOrder, kLean, true>(t, A, T, st, g0, gc, off32);
            return;
        }
    } else if constexpr (kOrder == kFlagWithLoads) {
        // the all-loads round: that chunk reads two flags per wave instead, with the body
        if (off32 && A.narrow >= g0 && A.narrow < g0 + gc) {
            launch_planes_n<R, kCount, kSparse, kLean, true>(t, A, T, st, g0, gc, off32);
            return;
        }
    }
    A.narrow = -1;
    launch_planes_n<R, kCount, kOrder, kLean, false>(t, A, T, st, g0, gc, off32);
}

// The per-field bookkeeping that ends a round.
void round_end(ts_tableset &t, hipStream_t st, int hold)
{
    const int groups = (t.nfields + kLanes - 1) / kLanes;
    eik_fast_round_end<<<groups, 64, 0, st>>>(t.active.p, t.phase1.p, t.changed.p, t.state.p, t.rounds_dev.p,
                                              sweep_round_cap(t.grid), t.status.p, t.reset.p, hold,
                                              hold ? t.first_rounds : 0);
    dim3 rg(blocks_for(t.N, 256), groups);
    eik_fast_reset_dirty<<<rg, 256, 0, st>>>(t.dirty.p, t.reset.p, t.N);
}

// Tile shapes of the streaming sweep: TU x TV z-columns over NW waves, each column a ring of
// kStreamRing entries of 64 fields in LDS (4 x 4 columns of FP64 times or 8 x 4 of FP32: 48 KB). The
// table set picks one per width (TOMOSTAR_STREAM_SHAPE32 / TOMOSTAR_STREAM_SHAPE64: index below).
template <int TU_, int TV_, int NW_>
struct Shp {
    static constexpr int TU = TU_, TV = TV_, NW = NW_;
};
template <typename R, typename F>
void with_shape(int idx, F &&f)
{
    if constexpr (sizeof(R) == 4) {
        if (idx == 1) f(Shp<8, 4, 16>{});
        else f(Shp<8, 4, 8>{});
    } else {
        if (idx == 1) f(Shp<4, 4, 8>{});
        else f(Shp<4, 4, 4>{});
    }
}
template <typename R, int TU, int TV>
constexpr size_t stream_lds_bytes()
{
    return (size_t)TU * TV * kStreamRing * (kLanes * sizeof(R) + 1);
}

// The task list of the streaming rounds on FP32 (w = 0) or FP64 (w = 1) times: every group's tiles, by
// tile diagonal, then group, then tile, so each tile follows its upwind tiles (one column to the left
// and one below in the ordering's mirrored coordinates), and the progress slots of a round.
void stream_plan(ts_tableset &t, int w, int TU, int TV)
{
    auto &pl = t.splan[w];
    const int groups = (t.nfields + kLanes - 1) / kLanes;
    const int ntu = (t.grid.nx + TU - 1) / TU, ntv = (t.grid.ny + TV - 1) / TV;
    if (ntu >= 4096 || ntv >= 4096 || groups > 255) fail(TS_DECLINED, "streaming tiles: grid or groups too large");
    std::vector<uint32_t> list;
    list.reserve((size_t)groups * ntu * ntv);
    for (int d = 0; d <= ntu - 1 + ntv - 1; d++)
        for (int g = 0; g < groups; g++)
            for (int tu = std::max(0, d - (ntv - 1)); tu <= std::min(ntu - 1, d); tu++)
                list.push_back((uint32_t)tu | (uint32_t)(d - tu) << 12 | (uint32_t)g << 24);
    pl.tu = TU;
    pl.tv = TV;
    pl.ntu = ntu;
    pl.ntv = ntv;
    pl.slots = groups * ntu * ntv;
    pl.ntasks = (int)list.size();
    pl.tasks.alloc(list.size());
    TS_HIP(hipMemcpy(pl.tasks.p, list.data(), list.size() * 4, hipMemcpyHostToDevice));
}

// Workgroups of one streaming launch per CU: as many as stay resident (a persistent grid). Queried once
// per kernel (host calls only, none on a stream, so the first launch may be under capture).
template <typename K>
int stream_blocks_per_cu(K kernel, int threads, size_t lds)
{
    TS_HIP(hipFuncSetAttribute((const void *)kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    int per_cu = 0;
    TS_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, threads, lds));
    return std::max(1, per_cu);
}

template <typename R, typename Sh, int DIR, bool kCount, bool kLean>
int stream_per_cu_dir()
{
    static const int per_cu = stream_blocks_per_cu(eik_stream<R, Sh::TU, Sh::TV, Sh::NW, DIR, kCount, kLean>,
                                                   64 * Sh::NW, stream_lds_bytes<R, Sh::TU, Sh::TV>());
    return per_cu;
}

// The resident workgroups per CU of the eight launches of a streaming round, queried before any round
// is captured (the queries are host calls, kept out of stream capture).
template <typename R, bool kCount, bool kLean>
void stream_prepare(int shape)
{
    if constexpr (sizeof(R) == 4 && kLean) return;  // not built: such rounds sweep planes
    else
        with_shape<R>(shape, [](auto sh) {
            using Sh = decltype(sh);
            (void)stream_per_cu_dir<R, Sh, 0, kCount, kLean>(); (void)stream_per_cu_dir<R, Sh, 1, kCount, kLean>();
            (void)stream_per_cu_dir<R, Sh, 2, kCount, kLean>(); (void)stream_per_cu_dir<R, Sh, 3, kCount, kLean>();
            (void)stream_per_cu_dir<R, Sh, 4, kCount, kLean>(); (void)stream_per_cu_dir<R, Sh, 5, kCount, kLean>();
            (void)stream_per_cu_dir<R, Sh, 6, kCount, kLean>(); (void)stream_per_cu_dir<R, Sh, 7, kCount, kLean>();
        });
}

template <typename R, typename Sh, int DIR, bool kCount, bool kLean>
void launch_stream_dir(ts_tableset &t, const StreamArgs<R> &S, hipStream_t st)
{
    constexpr size_t lds = stream_lds_bytes<R, Sh::TU, Sh::TV>();
    const int grid = std::max(1, std::min(S.ntasks, stream_per_cu_dir<R, Sh, DIR, kCount, kLean>() * t.ctx->cu_count));
    eik_stream<R, Sh::TU, Sh::TV, Sh::NW, DIR, kCount, kLean><<<grid, 64 * Sh::NW, lds, st>>>(S);
}

// One round of the streaming sweep: queues and progress cleared by a kernel (no memset node in a
// captured graph), the eight orderings as one persistent launch each, then the per-field bookkeeping.
template <typename R, bool kCount, bool kLean>
void launch_stream_round(ts_tableset &t, R *T, hipStream_t st, int hold)
{
    auto &pl = t.splan[sizeof(R) == 8];
    StreamArgs<R> S;
    S.A = fast_args<R>(t, T);
    S.tasks = pl.tasks.p;
    S.ntasks = pl.ntasks;
    S.g0 = 0;
    S.ntu = pl.ntu;
    S.ntv = pl.ntv;
    S.slots = pl.slots;
    S.spin_limit = t.stream_spin_limit;
    S.publish = t.stream_publish;
    S.sync.queue = t.stream_queue.p;
    S.sync.prog = t.stream_prog.p;
    S.sync.err = t.status.p + 2;
    const int n = 8 * pl.slots;
    eik_stream_reset<<<blocks_for(std::max(n, 8), 256), 256, 0, st>>>(t.stream_queue.p, t.stream_prog.p, n);
    with_shape<R>(t.stream_shape[sizeof(R) == 8], [&](auto sh) {
        using Sh = decltype(sh);
        launch_stream_dir<R, Sh, 0, kCount, kLean>(t, S, st);
        launch_stream_dir<R, Sh, 1, kCount, kLean>(t, S, st);
        launch_stream_dir<R, Sh, 2, kCount, kLean>(t, S, st);
        launch_stream_dir<R, Sh, 3, kCount, kLean>(t, S, st);
        launch_stream_dir<R, Sh, 4, kCount, kLean>(t, S, st);
        launch_stream_dir<R, Sh, 5, kCount, kLean>(t, S, st);
        launch_stream_dir<R, Sh, 6, kCount, kLean>(t, S, st);
        launch_stream_dir<R, Sh, 7, kCount, kLean>(t, S, st);
    });
    TS_HIP(hipGetLastError());
    t.plane_launches_per_round = 8;
    round_end(t, st, hold);
}

template <typename R, bool kCount, int kOrder, bool kLean>
void launch_round(ts_tableset &t, R *T, hipStream_t st, int hold)
{
    const int groups = (t.nfields + kLanes - 1) / kLanes;
    const int P = t.grid.nx + t.grid.ny + t.grid.nz - 2;
    // groups are independent fields: a chunk of them sweeps all eight orderings before the next chunk
    // starts, so the planes a sweep keeps re-reading belong to fewer fields and stay in cache
    const int chunk = std::max(1, std::min(groups, group_chunk(t, sizeof(R))));
    t.plane_launches_per_round = ((groups + chunk - 1) / chunk) * 8 * P;
    for (int g0 = 0; g0 < groups; g0 += chunk)
        launch_planes<R, kCount, kOrder, kLean>(t, T, st, g0, std::min(chunk, groups - g0));
    round_end(t, st, hold);
}

template <typename R, bool kLean>
void launch_planes_lean(ts_tableset &t, R *T, hipStream_t st, bool count, int order, int g0, int gc)
{
    if (count) {
        if (order == kFlagWithLoads) launch_planes<R, true, kFlagWithLoads, kLean>(t, T, st, g0, gc);
        else if (order == kFlagFirst) launch_planes<R, true, kFlagFirst, kLean>(t, T, st, g0, gc);
        else if (order == kSparseWide) launch_planes<R, true, kSparseWide, kLean>(t, T, st, g0, gc);
        else launch_planes<R, true, kSparse, kLean>(t, T, st, g0, gc);
    } else {
        if (order == kFlagWithLoads) launch_planes<R, false, kFlagWithLoads, kLean>(t, T, st, g0, gc);
        else if (order == kFlagFirst) launch_planes<R, false, kFlagFirst, kLean>(t, T, st, g0, gc);
        else if (order == kSparseWide) launch_planes<R, false, kSparseWide, kLean>(t, T, st, g0, gc);
        else launch_planes<R, false, kSparse, kLean>(t, T, st, g0, gc);
    }
}

template <typename R>
void launch_planes_as(ts_tableset &t, R *T, hipStream_t st, bool count, int order, int g0, int gc)
{
    if (t.lean[sizeof(R) == 8]) launch_planes_lean<R, true>(t, T, st, count, order, g0, gc);
    else launch_planes_lean<R, false>(t, T, st, count, order, g0, gc);
}

// Streams the groups sweep on side by side when rounds run as graphs (TOMOSTAR_GROUP_STREAMS overrides;
// 0 or 1: one stream, every group in every launch). The groups are independent fields, so each chunk of
// them runs its own graph of plane launches on its own stream between a fork and a join, the results
// unchanged to the bit. Measured on 400 fields (7 groups) of 201 x 201 x 101: one stream 0.802 s per
// mixed solve, 2 streams 0.723 s, 3 streams 0.716 s, 4 streams 0.928 s (with the solve's stream, more
// streams than the 4 hardware queues hold apart). The plane launches of two chunks overlap: one
// chunk's launch gaps are filled by the other's waves.
int group_streams(int groups, bool graphs)
{
    int n = graphs ? 3 : 0;  // direct launches (TOMOSTAR_NO_GRAPHS, the serving workers) keep one stream
    if (const char *e = getenv("TOMOSTAR_GROUP_STREAMS")) n = atoi(e);
    return std::max(0, std::min({n, groups, (int)ts_tableset::kMaxStreams}));
}

template <typename R, bool kCount, bool kLean>
void launch_round_order(ts_tableset &t, R *T, hipStream_t st, int hold, int order)
{
    if (order == kStream) {
        // the streaming kernels are built for the arithmetic each width runs by default and for
        // the FP64 form's (FP32 times: the correctly rounded offset form)
        if constexpr (sizeof(R) == 4 && kLean) launch_round<R, kCount, kFlagFirst, kLean>(t, T, st, hold);
        else launch_stream_round<R, kCount, kLean>(t, T, st, hold);
    } else if (order == kFlagWithLoads) launch_round<R, kCount, kFlagWithLoads, kLean>(t, T, st, hold);
    else if (order == kFlagFirst) launch_round<R, kCount, kFlagFirst, kLean>(t, T, st, hold);
    else if (order == kSparseWide) launch_round<R, kCount, kSparseWide, kLean>(t, T, st, hold);
    else launch_round<R, kCount, kSparse, kLean>(t, T, st, hold);
}

template <typename R>
void launch_round_as(ts_tableset &t, R *T, hipStream_t st, int hold, bool count, int order)
{
    const bool lean = t.lean[sizeof(R) == 8];
    if (count) {
        if (lean) launch_round_order<R, true, true>(t, T, st, hold, order);
        else launch_round_order<R, true, false>(t, T, st, hold, order);
    } else {
        if (lean) launch_round_order<R, false, true>(t, T, st, hold, order);
        else launch_round_order<R, false, false>(t, T, st, hold, order);
    }
}

// How round `round` of a run reads the dirty flags:
// - the mixed form's first-order phase reads kSparseNodes flags per wave from round 0
//   (TOMOSTAR_SPARSE_FROM overrides; rounds before it read the flag first). Measured on 400 fields of
//   201 x 201 x 101, the phase takes 0.306 s; 0.313 s with round 0 issuing all loads at once, 0.321 s
//   with rounds 0-1 so;
// - the mixed form's second-order phase starts with every node dirty and most of them clean after its
//   first round: all loads at once in round 0 (TOMOSTAR_FP64_FLAG_FIRST moves the end of the all-loads
//   rounds), kSparseNodes flags per wave from round 1 (TOMOSTAR_SECOND_SPARSE_FROM overrides; rounds
//   between the two read the flag first). Sparse from round 1, 2 or 3 takes the solve 0.938, 0.946 and
//   0.957 s at two flags per wave (0.965, 0.965 and 0.974 s at four);
// - the FP64 form reads the flag first in every round: a field's second-order phase starts in mid-run
//   with every node of its group dirty again.
// Sparse rounds first took the solve from 1.26 s to 1.20 s; with the rounds above, 1, 2, 4 and 8 flags
// per wave (kSparseNodes) take it 0.986, 0.946, 0.959 and 1.024 s (16 or 32: slower still; fewer waves
// are left to hide the latency of the dirty nodes). With the loads of several dirty nodes in flight
// (kInFlight), the FP32 phase takes 0.237 s at 2 flags and 2 nodes, 0.227 s at 4 and 4, 0.239 s at 8 and
// 4; 64 flags per wave from a late round on is slower in every round tried (FP64 rounds 2, 3, 4: 1.103,
// 0.761, 0.584 s against 0.569 s; FP32 rounds 1, 2: 0.617, 0.407 s against 0.226 s) - the late rounds
// still hold dirty nodes enough that a wave working through 64 flags' worth becomes the launch's tail.
// - the second-order phase reads kWideNodes flags per wave from its round 5 (TOMOSTAR_SECOND_WIDE_FROM;
//   TOMOSTAR_WIDE_FROM for the first-order phase, off by default): a round that finds few nodes dirty is
//   bound by the waves that read the flags. Measured on the same case (results identical to the bit),
//   four flags per wave take FP64 round 5 from 18 to 14 ms but rounds 2-4 from 96, 44, 24 ms to 99, 50,
//   25 ms, and eight take the first-order phase's rounds 1-2 from 62, 30 ms to 69, 38 ms: wide from
//   round 5, 0.711 s per solve; from 2, 3, 4: 0.721, 0.719, 0.712 s; never: 0.716 s.
// - rounds before TOMOSTAR_STREAM_ROUNDS (a run from the seeds) or TOMOSTAR_STREAM_ROUNDS2 (the mixed
//   form's second-order phase) run the streaming sweep.
int round_order(bool all_dirty, bool one_phase, int round)
{
    int stream = 0;
    if (const char *e = getenv(all_dirty ? "TOMOSTAR_STREAM_ROUNDS2" : "TOMOSTAR_STREAM_ROUNDS")) stream = std::max(0, atoi(e));
    if (round < stream) return kStream;
    int first = 0;
    if (all_dirty) {
        first = 1;
        if (const char *e = getenv("TOMOSTAR_FP64_FLAG_FIRST")) first = std::max(0, atoi(e));
    }
    int sparse = all_dirty ? first : 0;
    if (const char *e = getenv(all_dirty ? "TOMOSTAR_SECOND_SPARSE_FROM" : "TOMOSTAR_SPARSE_FROM"))
        sparse = std::max(0, atoi(e));
    int wide = all_dirty ? 5 : 1 << 20;
    if (const char *e = getenv(all_dirty ? "TOMOSTAR_SECOND_WIDE_FROM" : "TOMOSTAR_WIDE_FROM")) wide = std::max(0, atoi(e));
    return round < first ? kFlagWithLoads : round < sparse || !one_phase ? kFlagFirst : round < wide ? kSparse : kSparseWide;
}

// One round as a graph: 8 orderings x all planes x all groups, then the per-field bookkeeping. The
// FP32 graph holds fields at the end of their first-order phase (mixed form), the FP64 one does not.
template <typename R>
hipGraphExec_t round_graph(ts_tableset &t, R *T, bool count, int order)
{
    const int w = sizeof(R) == 8;
    const int woken = t.wake_tol[w] > 0;
    hipGraphExec_t &ge = t.graphs[w][count][order][woken];
    ts_tableset::SolveKey &captured = t.graph_key[w][count][order][woken];
    // captured with another tolerance, threshold or arithmetic (kernel arguments are captured)
    if (ge && !(captured == t.solve_key(w))) {
        TS_HIP(hipGraphExecDestroy(ge));
        ge = nullptr;
    }
    if (ge) return ge;
    captured = t.solve_key(w);
    hipStream_t st = t.ctx->stream;
    hipGraph_t gr;
    TS_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    const int hold = sizeof(R) == 4;
    launch_round_as<R>(t, T, st, hold, count, order);
    TS_HIP(hipStreamEndCapture(st, &gr));
    hipError_t e = hipGraphInstantiate(&ge, gr, nullptr, nullptr, 0);
    (void)hipGraphDestroy(gr);
    TS_HIP(e);
    return ge;
}

// The plane launches of chunk c (groups [g0, g0 + gc)) as a graph captured on stream st.
template <typename R>
hipGraphExec_t chunk_graph(ts_tableset &t, R *T, hipStream_t st, bool count, int order, int c, int g0, int gc)
{
    const int w = sizeof(R) == 8;
    const int woken = t.wake_tol[w] > 0;
    hipGraphExec_t &ge = t.cgraphs[w][count][order][woken][c];
    ts_tableset::SolveKey &captured = t.cgraph_key[w][count][order][woken][c];
    int *range = t.cgraph_groups[w][count][order][woken][c];
    // captured with another tolerance, threshold or arithmetic, or for other groups (the stream count
    // changed between solves)
    if (ge && (!(captured == t.solve_key(w)) || range[0] != g0 || range[1] != gc)) {
        TS_HIP(hipGraphExecDestroy(ge));
        ge = nullptr;
    }
    if (ge) return ge;
    captured = t.solve_key(w);
    range[0] = g0;
    range[1] = gc;
    hipGraph_t gr;
    TS_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    launch_planes_as<R>(t, T, st, count, order, g0, gc);
    TS_HIP(hipStreamEndCapture(st, &gr));
    hipError_t e = hipGraphInstantiate(&ge, gr, nullptr, nullptr, 0);
    (void)hipGraphDestroy(gr);
    TS_HIP(e);
    return ge;
}

// Rounds of one graph until no field is relaxing; the sweep time is bracketed by HIP events on the
// solve's stream. Declines when a field exhausts its round cap.
template <typename R>
void run_rounds(ts_tableset &t, R *T, bool count, const volatile int32_t *cancel, int32_t *h_status, bool all_dirty,
                bool one_phase)
{
    const int hold = sizeof(R) == 4;
    hipStream_t st = t.ctx->stream;
    hipEvent_t ev0, ev1;
    TS_HIP(hipEventCreate(&ev0));
    TS_HIP(hipEventCreate(&ev1));
    struct FreeEv { hipEvent_t a, b; ~FreeEv() { (void)hipEventDestroy(a); (void)hipEventDestroy(b); } } free_ev{ev0, ev1};
    const int max_rounds = 2 * sweep_round_cap(t.grid) + 2;
    const int groups = (t.nfields + kLanes - 1) / kLanes;
    const int S = group_streams(groups, t.ctx->use_graphs);
    // chunk c holds groups [cut[c], cut[c + 1]): the groups spread evenly over the streams
    std::vector<int> cut(S + 1, 0);
    for (int c = 0; c <= S; c++) cut[c] = S > 1 ? (int)((int64_t)c * groups / S) : (c ? groups : 0);
    if (S > 1) {
        for (int c = 1; c < S; c++) {
            if (!t.gstream[c]) TS_HIP(hipStreamCreateWithFlags(&t.gstream[c], hipStreamNonBlocking));
            if (!t.gjoin[c]) TS_HIP(hipEventCreateWithFlags(&t.gjoin[c], hipEventDisableTiming));
        }
        if (!t.gfork) TS_HIP(hipEventCreateWithFlags(&t.gfork, hipEventDisableTiming));
        // a plane step (one ordering's plane) is S kernels side by side: counted once
        t.plane_launches_per_round = 8 * (t.grid.nx + t.grid.ny + t.grid.nz - 2);
    }
    for (int round = 0; round < max_rounds; round++) {
        if (cancelled(cancel)) fail(TS_CANCELLED, "cancelled");
        TS_HIP(hipMemsetAsync(t.status.p, 0, 3 * sizeof(int32_t), st));
        const int order = round_order(all_dirty, one_phase, round);
        if (S > 1 && order != kStream) {
            // chunk 0 sweeps on the solve's stream, every other chunk on its own stream between a fork
            // and a join on the solve's stream
            std::vector<hipGraphExec_t> cg(S, nullptr);
            for (int c = 0; c < S; c++)
                if (t.ctx->use_graphs)
                    cg[c] = chunk_graph<R>(t, T, c == 0 ? st : t.gstream[c], count, order, c, cut[c],
                                           cut[c + 1] - cut[c]);
            TS_HIP(hipEventRecord(ev0, st));
            TS_HIP(hipEventRecord(t.gfork, st));
            for (int c = 1; c < S; c++) {
                TS_HIP(hipStreamWaitEvent(t.gstream[c], t.gfork, 0));
                if (cg[c]) TS_HIP(hipGraphLaunch(cg[c], t.gstream[c]));
                else launch_planes_as<R>(t, T, t.gstream[c], count, order, cut[c], cut[c + 1] - cut[c]);
                TS_HIP(hipEventRecord(t.gjoin[c], t.gstream[c]));
            }
            if (cg[0]) TS_HIP(hipGraphLaunch(cg[0], st));
            else launch_planes_as<R>(t, T, st, count, order, 0, cut[1]);
            for (int c = 1; c < S; c++) TS_HIP(hipStreamWaitEvent(st, t.gjoin[c], 0));
            round_end(t, st, hold);
        } else {
            hipGraphExec_t ge = t.ctx->use_graphs ? round_graph<R>(t, T, count, order) : nullptr;
            TS_HIP(hipEventRecord(ev0, st));
            if (ge) TS_HIP(hipGraphLaunch(ge, st));
            else launch_round_as<R>(t, T, st, hold, count, order);
        }
        TS_HIP(hipEventRecord(ev1, st));
        TS_HIP(hipMemcpyAsync(h_status, t.status.p, 3 * sizeof(int32_t), hipMemcpyDeviceToHost, st));
        TS_HIP(hipStreamSynchronize(st));
        float ms = 0;
        TS_HIP(hipEventElapsedTime(&ms, ev0, ev1));
        t.sweep_seconds += ms * 1e-3;
        t.round_kind.push_back((sizeof(R) == 8 ? 16 : 0) | order);
        t.round_seconds.push_back(ms * 1e-3);
        t.sweep_by_width[sizeof(R) == 8] += ms * 1e-3;
        t.rounds_by_width[sizeof(R) == 8]++;
        if (h_status[2] > 0) fail(TS_EDEVICE, "streaming sweep: a tile waited for its upwind tiles beyond the spin limit");
        if (h_status[1] > 0) fail(TS_DECLINED, "sweeping not settled after %d rounds", sweep_round_cap(t.grid));
        if (h_status[0] == 0) return;
    }
    fail(TS_DECLINED, "sweeping not settled");
}

// The fast forms. FP64 (mode TS_EIK_FAST_FP64): Relax per field, both phases in FP64. Mixed (mode
// TS_EIK_FAST_FP32): the first-order phase in FP32 offset form (its fixed point is unique, and FP32
// reaches it to rounding), then the second-order phase in FP64 from the widened FP32 times - the
// second-order promotion amplifies storage rounding (an FP32 second-order phase drifts by up to
// 4e-4 of the time on a 0.5 km grid), so it keeps FP64 times.
// The mixed form's second-order phase wakes a node's readers only on drops above 1e-8 s (1/10 of the
// settle tolerance; TOMOSTAR_MIXED_WAKE_TOL overrides): the phase starts near its fixed point, and
// without the threshold rounding-level drops keep almost every node dirty in every round. Measured on
// 400 fields of 201 x 201 x 101 (solve seconds; mean and largest change of the times against the
// 1e-10 s threshold over every tenth field): 1e-10 s 1.031 s; 1e-9 s 1.005 s, 6.2e-12 / 8.8e-9 s;
// 1e-8 s 0.975 s, 5.9e-11 / 1.1e-7 s; 3e-8 s 0.962 s, 2.1e-10 / 8.1e-7 s. At 1e-8 s no time moves by
// more than the settle tolerance, and the largest distance to the FP64 form (4.571e-6 relative) is
// unchanged. Without any threshold the phase took 1.25 s against 0.76 s at 1e-10 s.
double mixed_wake_tol()
{
    if (const char *e = getenv("TOMOSTAR_MIXED_WAKE_TOL")) return std::max(0.0, atof(e));
    return 1e-8;
}

void solve_fast(ts_tableset &t, int mode, const volatile int32_t *cancel, bool count)
{
    ts_context *ctx = t.ctx;
    hipStream_t st = ctx->stream;
    const bool mixed = mode == TS_EIK_FAST_FP32;
    const int N = t.N, groups = (t.nfields + kLanes - 1) / kLanes, slots = groups * kLanes;
    const size_t nodes_total = (size_t)groups * N;
    if (mixed) t.T32.ensure(nodes_total * kLanes);
    t.T64.ensure(nodes_total * kLanes);
    t.dirty.ensure(nodes_total);
    t.seedmask.ensure(nodes_total);
    t.active.ensure(groups); t.phase1.ensure(groups); t.changed.ensure(groups); t.reset.ensure(groups);
    t.model_dev.ensure(slots); t.state.ensure(slots); t.rounds_dev.ensure(2 * slots); t.status.ensure(3);
    t.evals.ensure(2048);
    // the streaming kernels' launch shapes, outside any stream capture
    {
        const int s32 = t.stream_shape[0], s64 = t.stream_shape[1];
        if (count) {
            stream_prepare<float, true, false>(s32); stream_prepare<double, true, false>(s64);
            stream_prepare<double, true, true>(s64);
        } else {
            stream_prepare<float, false, false>(s32); stream_prepare<double, false, false>(s64);
            stream_prepare<double, false, true>(s64);
        }
    }
    // the streaming rounds' task lists (once per table set: captured graphs keep these buffers)
    if (!t.splan[0].ntasks) {
        with_shape<float>(t.stream_shape[0], [&](auto sh) { stream_plan(t, 0, decltype(sh)::TU, decltype(sh)::TV); });
        with_shape<double>(t.stream_shape[1], [&](auto sh) { stream_plan(t, 1, decltype(sh)::TU, decltype(sh)::TV); });
        t.stream_queue.alloc(8);
        t.stream_prog.alloc((size_t)8 * std::max(t.splan[0].slots, t.splan[1].slots));
    }
    t.tab32.ensure((size_t)t.nfields * N);
    t.tab64.ensure((size_t)t.nfields * N);

    // per-field state: all fields start in the first-order phase
    std::vector<int32_t> model(slots, 0), state(slots, 4);
    std::vector<uint64_t> active(groups, 0);
    for (int f = 0; f < t.nfields; f++) {
        const int sl = t.slot_of_field[f];
        model[sl] = t.model_of_field[f];
        state[sl] = 0;
        active[sl / kLanes] |= 1ull << (sl % kLanes);
    }
    TS_HIP(hipMemcpyAsync(t.model_dev.p, model.data(), slots * 4, hipMemcpyHostToDevice, st));
    TS_HIP(hipMemcpyAsync(t.state.p, state.data(), slots * 4, hipMemcpyHostToDevice, st));
    TS_HIP(hipMemcpyAsync(t.active.p, active.data(), groups * 8, hipMemcpyHostToDevice, st));
    TS_HIP(hipMemsetAsync(t.phase1.p, 0, groups * 8, st));
    TS_HIP(hipMemsetAsync(t.changed.p, 0, groups * 8, st));
    TS_HIP(hipMemsetAsync(t.rounds_dev.p, 0, 2 * slots * 4, st));
    TS_HIP(hipMemsetAsync(t.evals.p, 0, 2048 * 8, st));
    std::vector<int32_t> sf, sn;
    std::vector<double> sv;
    for (int f = 0; f < t.nfields; f++)
        for (size_t q = 0; q < t.seed_node[f].size(); q++) {
            sf.push_back(t.slot_of_field[f]);
            sn.push_back(t.seed_node[f][q]);
            sv.push_back(t.seed_time[f][q]);
        }
    dbuf<int32_t> df(sf.size()), dn(sn.size());
    dbuf<double> dv(sv.size());
    TS_HIP(hipMemcpyAsync(df.p, sf.data(), sf.size() * 4, hipMemcpyHostToDevice, st));
    TS_HIP(hipMemcpyAsync(dn.p, sn.data(), sn.size() * 4, hipMemcpyHostToDevice, st));
    TS_HIP(hipMemcpyAsync(dv.p, sv.data(), sv.size() * 8, hipMemcpyHostToDevice, st));
    const unsigned cb = blocks_for(nodes_total * kLanes, 256), sb = blocks_for(sf.size(), 256);
    if (mixed) {
        eik_fast_clear<float><<<cb, 256, 0, st>>>(t.T32.p, t.dirty.p, t.seedmask.p, nodes_total);
        eik_fast_seed<float><<<sb, 256, 0, st>>>(t.T32.p, t.seedmask.p, t.dirty.p, df.p, dn.p, dv.p, (int)sf.size(),
                                                 t.grid.nx, t.grid.ny, t.grid.nz, true);
    } else {
        eik_fast_clear<double><<<cb, 256, 0, st>>>(t.T64.p, t.dirty.p, t.seedmask.p, nodes_total);
        eik_fast_seed<double><<<sb, 256, 0, st>>>(t.T64.p, t.seedmask.p, t.dirty.p, df.p, dn.p, dv.p, (int)sf.size(),
                                                  t.grid.nx, t.grid.ny, t.grid.nz, true);
    }
    TS_HIP(hipGetLastError());
    TS_HIP(hipStreamSynchronize(st));
    int32_t *h_status = nullptr;
    TS_HIP(hipHostMalloc((void **)&h_status, 3 * sizeof(int32_t), hipHostMallocDefault));
    struct Free { int32_t *p; ~Free() { if (p) (void)hipHostFree(p); } } free_h{h_status};
    t.sweep_seconds = 0;
    t.round_kind.clear();
    t.round_seconds.clear();
    t.sweep_by_width[0] = t.sweep_by_width[1] = 0;
    t.rounds_by_width[0] = t.rounds_by_width[1] = 0;
    // The FP64 form wakes a node's readers on every drop (Relax's result, bit for bit); the mixed form's
    // first-order phase on drops above kFp32WakeUlps of the time, its second-order phase on drops above
    // mixed_wake_tol().
    t.wake_tol[0] = 0.0;
    t.wake_tol[1] = mixed ? mixed_wake_tol() : 0.0;
    // The mixed form's second-order phase uses godunov_lean64 (TOMOSTAR_MIXED_EXACT=1: Relax's arithmetic);
    // its first-order phase the correctly rounded FP32 offset form (TOMOSTAR_LEAN_FIRST=1: godunov_lean32).
    // The FP64 form keeps Relax's arithmetic.
    const bool exact = getenv("TOMOSTAR_MIXED_EXACT") && atoi(getenv("TOMOSTAR_MIXED_EXACT")) != 0;
    const bool lean_first = getenv("TOMOSTAR_LEAN_FIRST") && atoi(getenv("TOMOSTAR_LEAN_FIRST")) != 0;
    t.lean[0] = mixed && lean_first;
    t.lean[1] = mixed && !exact;
    t.settle_ulps = getenv("TOMOSTAR_FP32_SETTLE_ULPS") ? (float)atof(getenv("TOMOSTAR_FP32_SETTLE_ULPS")) : kFp32SettleUlps;
    t.wake_ulps = !mixed ? 0.0f : getenv("TOMOSTAR_FP32_WAKE_ULPS") ? (float)atof(getenv("TOMOSTAR_FP32_WAKE_ULPS")) : kFp32WakeUlps;
    t.first_rounds = mixed && getenv("TOMOSTAR_FP32_MAX_ROUNDS") ? std::max(0, atoi(getenv("TOMOSTAR_FP32_MAX_ROUNDS"))) : 0;
    if (mixed) {
        run_rounds<float>(t, t.T32.p, count, cancel, h_status, false, true);
        eik_mixed_widen<<<blocks_for(nodes_total * kLanes, 256), 256, 0, st>>>(t.T32.p, t.T64.p, nodes_total * kLanes);
        // the fixed seed times again, unrounded
        eik_fast_seed<double><<<sb, 256, 0, st>>>(t.T64.p, t.seedmask.p, t.dirty.p, df.p, dn.p, dv.p, (int)sf.size(),
                                                  t.grid.nx, t.grid.ny, t.grid.nz, false);
        eik_mixed_promote<<<groups, 64, 0, st>>>(t.active.p, t.phase1.p, t.changed.p, t.state.p);
        eik_fill_u8<<<blocks_for(nodes_total, 256), 256, 0, st>>>(t.dirty.p, nodes_total, 1);
        TS_HIP(hipGetLastError());
    }
    run_rounds<double>(t, t.T64.p, count, cancel, h_status, mixed, mixed);
    dim3 ug(blocks_for(N, 64), groups);
    // the stored tables are FP32 (TravelTimeTable's float32 tables); the FP64 times stay in T64 until a
    // caller reads them in FP64 (ts_tables_read), which unpacks them then
    eik_fast_unpack<double><<<ug, 256, 0, st>>>(t.T64.p, N, t.field_of_slot_dev.p, t.tab32.p, nullptr);
    TS_HIP(hipGetLastError());
    std::vector<int32_t> rounds(2 * slots);
    TS_HIP(hipMemcpyAsync(rounds.data(), t.rounds_dev.p, rounds.size() * 4, hipMemcpyDeviceToHost, st));
    std::vector<unsigned long long> ev(2048);
    TS_HIP(hipMemcpyAsync(ev.data(), t.evals.p, 2048 * 8, hipMemcpyDeviceToHost, st));
    TS_HIP(hipStreamSynchronize(st));
    for (int f = 0; f < t.nfields; f++) {
        const int sl = t.slot_of_field[f];
        t.rounds[2 * f] = rounds[2 * sl];
        t.rounds[2 * f + 1] = rounds[2 * sl + 1];
    }
    if (count) {
        double e32 = 0, e64 = 0;
        for (int q = 0; q < 1024; q++) { e32 += (double)ev[q]; e64 += (double)ev[1024 + q]; }
        t.evaluations = e32 + e64;
        t.evaluations32 = e32;
        t.evaluations64 = e64;
    }
    t.have64 = false;
    t.pending64 = true;
}

}  // namespace

namespace ts {
const float *tableset_tables(const ts_tableset *t) { return t->solved ? t->tab32.p : nullptr; }
Grid tableset_grid(const ts_tableset *t) { return t->grid; }
int tableset_fields(const ts_tableset *t) { return t->nfields; }
ts_context *tableset_ctx(const ts_tableset *t) { return t->ctx; }
}  // namespace ts

namespace ts {

// ts_tables_create, with the slowness models either on the host (slowness) or resident on the device
// (dev64 / dev32: FP64 and FP32 copies, the serving path).
int tables_create_core(ts_context *ctx, const ts_grid *grid, int nmodels, const double *const *slowness,
                       const double *const *dev64, const float *const *dev32, int nfields,
                       const int32_t *model_of_field, const int64_t *seed_offsets, const int32_t *seed_index,
                       const double *seed_time, ts_tableset **out, char *reason, int reason_len)
{
    if (out) *out = nullptr;
    if (!ctx || !grid || !out || nmodels < 1 || (!slowness && !(dev64 && dev32)) || nfields < 1 || !seed_offsets ||
        !grid_valid(*grid)) {
        set_reason(reason, reason_len, "invalid argument");
        return TS_EINVAL;
    }
    return guarded(reason, reason_len, [&]() -> int {
        TS_HIP(hipSetDevice(ctx->device));
        std::unique_ptr<ts_tableset> t(new ts_tableset);
        t->ctx = ctx;
        t->grid = make_grid(*grid);
        t->nmodels = nmodels;
        t->nfields = nfields;
        t->N = t->grid.count();
        const int N = t->N;
        t->model_of_field.assign(nfields, 0);
        for (int f = 0; f < nfields; f++) {
            int m = model_of_field ? model_of_field[f] : 0;
            if (m < 0 || m >= nmodels) fail(TS_EINVAL, "field %d names slowness model %d of %d", f, m, nmodels);
            t->model_of_field[f] = m;
        }
        // merge seeds: per field, smallest time per node (EikonalOpenCl.cs:150-163)
        t->seed_node.resize(nfields);
        t->seed_time.resize(nfields);
        std::vector<std::pair<int32_t, double>> seeds;
        for (int f = 0; f < nfields; f++) {
            int64_t a = seed_offsets[f], b = seed_offsets[f + 1];
            if (b < a) fail(TS_EINVAL, "seed offsets decrease at field %d", f);
            if (b > a && (!seed_index || !seed_time)) fail(TS_EINVAL, "seed arrays are null");
            seeds.clear();
            for (int64_t q = a; q < b; q++)
                if ((uint32_t)seed_index[q] < (uint32_t)N) seeds.emplace_back(seed_index[q], seed_time[q]);
            if (seeds.empty()) fail(TS_DECLINED, "no seed nodes for field %d", f);
            std::stable_sort(seeds.begin(), seeds.end(),
                             [](const auto &x, const auto &y) { return x.first < y.first; });
            for (size_t q = 0; q < seeds.size(); q++) {
                if (q > 0 && seeds[q].first == t->seed_node[f].back()) {
                    if (seeds[q].second < t->seed_time[f].back()) t->seed_time[f].back() = seeds[q].second;
                    continue;
                }
                t->seed_node[f].push_back(seeds[q].first);
                t->seed_time[f].push_back(seeds[q].second);
            }
        }
        {
            // Morton code of each field's first seed node (its source), ties kept in field order
            auto spread = [](uint32_t v) {
                uint64_t x = v & 0x3ff;
                x = (x | (x << 16)) & 0x30000ff;
                x = (x | (x << 8)) & 0x300f00f;
                x = (x | (x << 4)) & 0x30c30c3;
                x = (x | (x << 2)) & 0x9249249;
                return x;
            };
            const Grid &gg = t->grid;
            std::vector<uint64_t> code(nfields);
            for (int f = 0; f < nfields; f++) {
                const int32_t n0 = t->seed_node[f][std::min_element(t->seed_time[f].begin(), t->seed_time[f].end()) -
                                                    t->seed_time[f].begin()];
                const int k = n0 / (gg.nx * gg.ny), rest = n0 - k * gg.nx * gg.ny, j = rest / gg.nx, i = rest - j * gg.nx;
                const uint32_t si = (uint32_t)((int64_t)i * 1024 / gg.nx), sj = (uint32_t)((int64_t)j * 1024 / gg.ny),
                               sk = (uint32_t)((int64_t)k * 1024 / gg.nz);
                code[f] = spread(si) | (spread(sj) << 1) | (spread(sk) << 2);
            }
            std::vector<int32_t> order(nfields);
            for (int f = 0; f < nfields; f++) order[f] = f;
            std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return code[a] < code[b]; });
            const int slots = (nfields + kLanes - 1) / kLanes * kLanes;
            t->slot_of_field.assign(nfields, 0);
            t->field_of_slot.assign(slots, -1);
            for (int q = 0; q < nfields; q++) {
                t->slot_of_field[order[q]] = q;
                t->field_of_slot[q] = order[q];
            }
            t->field_of_slot_dev.alloc(slots);
            TS_HIP(hipMemcpy(t->field_of_slot_dev.p, t->field_of_slot.data(), slots * 4, hipMemcpyHostToDevice));
        }
        t->hlon.resize((size_t)t->grid.ny * t->grid.nz);
        t->hlat.resize(t->grid.nz);
        grid_metric(t->grid, t->hlon.data(), t->hlat.data(), &t->hr);
        hipStream_t st = exec_stream(ctx);
        TS_HIP(hipSetDevice(ctx->device));
        t->slow64.alloc((size_t)nmodels * N);
        t->slow32.alloc((size_t)nmodels * N);
        std::vector<float> s32(N);
        for (int m = 0; m < nmodels; m++) {
            if (dev64 && dev32) {
                // resident models: device-to-device copies on the solve's stream
                if (!dev64[m] || !dev32[m]) fail(TS_EINVAL, "resident model %d is null", m);
                TS_HIP(hipMemcpyAsync(t->slow64.p + (size_t)m * N, dev64[m], (size_t)N * 8, hipMemcpyDeviceToDevice, st));
                TS_HIP(hipMemcpyAsync(t->slow32.p + (size_t)m * N, dev32[m], (size_t)N * 4, hipMemcpyDeviceToDevice, st));
                continue;
            }
            if (!slowness[m]) fail(TS_EINVAL, "slowness model %d is null", m);
            TS_HIP(hipMemcpyAsync(t->slow64.p + (size_t)m * N, slowness[m], (size_t)N * 8, hipMemcpyHostToDevice, st));
            for (int q = 0; q < N; q++) s32[q] = (float)slowness[m][q];
            TS_HIP(hipMemcpyAsync(t->slow32.p + (size_t)m * N, s32.data(), (size_t)N * 4, hipMemcpyHostToDevice, st));
            TS_HIP(hipStreamSynchronize(st));
        }
        TS_HIP(hipStreamSynchronize(st));
        t->hlon64.alloc(t->hlon.size()); t->hlat64.alloc(t->hlat.size());
        t->hlon32.alloc(t->hlon.size()); t->hlat32.alloc(t->hlat.size());
        std::vector<float> hl32(t->hlon.begin(), t->hlon.end()), ht32(t->hlat.begin(), t->hlat.end());
        TS_HIP(hipMemcpy(t->hlon64.p, t->hlon.data(), t->hlon.size() * 8, hipMemcpyHostToDevice));
        TS_HIP(hipMemcpy(t->hlat64.p, t->hlat.data(), t->hlat.size() * 8, hipMemcpyHostToDevice));
        TS_HIP(hipMemcpy(t->hlon32.p, hl32.data(), hl32.size() * 4, hipMemcpyHostToDevice));
        TS_HIP(hipMemcpy(t->hlat32.p, ht32.data(), ht32.size() * 4, hipMemcpyHostToDevice));
        {
            // the fast form's spacing weights, FP64 from the FP64 metric and FP32 from the FP32 one
            std::vector<double> wl(2 * t->hlon.size()), wt(2 * t->hlat.size());
            std::vector<float> wl32(wl.size()), wt32(wt.size());
            for (size_t q = 0; q < t->hlon.size(); q++) {
                spacing_weights<double>(t->hlon[q], &wl[2 * q]);
                spacing_weights<float>(hl32[q], &wl32[2 * q]);
            }
            for (size_t q = 0; q < t->hlat.size(); q++) {
                spacing_weights<double>(t->hlat[q], &wt[2 * q]);
                spacing_weights<float>(ht32[q], &wt32[2 * q]);
            }
            spacing_weights<double>(t->hr, t->wr64);
            spacing_weights<float>((float)t->hr, t->wr32);
            t->wlon64.alloc(wl.size()); t->wlat64.alloc(wt.size());
            t->wlon32.alloc(wl.size()); t->wlat32.alloc(wt.size());
            TS_HIP(hipMemcpy(t->wlon64.p, wl.data(), wl.size() * 8, hipMemcpyHostToDevice));
            TS_HIP(hipMemcpy(t->wlat64.p, wt.data(), wt.size() * 8, hipMemcpyHostToDevice));
            TS_HIP(hipMemcpy(t->wlon32.p, wl32.data(), wl32.size() * 4, hipMemcpyHostToDevice));
            TS_HIP(hipMemcpy(t->wlat32.p, wt32.data(), wt32.size() * 4, hipMemcpyHostToDevice));
        }
        build_planes(*t);
        read_launch_knobs(*t);
        t->rounds.assign(2 * (size_t)nfields, 0);
        *out = t.release();
        return TS_OK;
    });
}


}  // namespace ts

// ================================================================================== C ABI

extern "C" {

int ts_grid_metric(const ts_grid *grid, double *hlon, double *hlat, double *hr)
{
    if (!grid || !hlon || !hlat || !hr || !grid_valid(*grid)) return TS_EINVAL;
    grid_metric(make_grid(*grid), hlon, hlat, hr);
    return TS_OK;
}

int ts_point_source(const ts_grid *grid, const double *slowness, double lon, double lat, double depth_km,
                    double origin_time, double radius_cells, int32_t *index_out, double *time_out, int cap)
{
    if (!grid || !slowness || !grid_valid(*grid) || cap < 0 || (cap > 0 && (!index_out || !time_out)))
        return TS_EINVAL;
    if (!std::isfinite(lon) || !std::isfinite(lat) || !std::isfinite(depth_km)) return TS_EINVAL;
    return point_source(make_grid(*grid), slowness, lon, lat, depth_km, origin_time, radius_cells, index_out,
                        time_out, cap);
}

int ts_eik_suggest_batch(ts_context *ctx, int nodes)
{
    if (!ctx || nodes <= 0) return TS_EINVAL;
    return suggest_batch(ctx, nodes);
}

int ts_tables_create(ts_context *ctx, const ts_grid *grid, int nmodels, const double *const *slowness,
                     int nfields, const int32_t *model_of_field, const int64_t *seed_offsets,
                     const int32_t *seed_index, const double *seed_time, ts_tableset **out, char *reason,
                     int reason_len)
{
    if (!slowness) {
        if (out) *out = nullptr;
        set_reason(reason, reason_len, "invalid argument");
        return TS_EINVAL;
    }
    return ts::tables_create_core(ctx, grid, nmodels, slowness, nullptr, nullptr, nfields, model_of_field,
                                  seed_offsets, seed_index, seed_time, out, reason, reason_len);
}

int ts_tables_solve(ts_tableset *t, int mode, const volatile int32_t *cancel, int32_t *rounds_out, char *reason,
                    int reason_len)
{
    if (!t || mode < TS_EIK_PLAIN_FP64 || mode > TS_EIK_FAST_FP64) {
        set_reason(reason, reason_len, "invalid argument");
        return TS_EINVAL;
    }
    std::lock_guard<std::mutex> lk(t->gate);
    return guarded(reason, reason_len, [&]() -> int {
        TS_HIP(hipSetDevice(t->ctx->device));
        t->solved = false;
        auto t0 = std::chrono::steady_clock::now();
        bool count = t->evaluations == 0;
        if (mode == TS_EIK_PLAIN_FP64) solve_plain(*t, cancel);
        else solve_fast(*t, mode, cancel, count);
        t->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        t->solved = true;
        if (rounds_out) {
            rounds_out[0] = rounds_out[1] = 0;
            for (int f = 0; f < t->nfields; f++) {
                rounds_out[0] = std::max(rounds_out[0], t->rounds[2 * f]);
                rounds_out[1] = std::max(rounds_out[1], t->rounds[2 * f + 1]);
            }
        }
        return TS_OK;
    });
}

int ts_tables_read(ts_tableset *t, int first, int count, double *out)
{
    if (!t || !out || first < 0 || count < 0 || first + count > t->nfields) return TS_EINVAL;
    if (!t->solved) return TS_EINVAL;
    std::lock_guard<std::mutex> lk(t->gate);
    return guarded(nullptr, 0, [&]() -> int {
        TS_HIP(hipSetDevice(t->ctx->device));
        const size_t n = (size_t)count * t->N, at = (size_t)first * t->N;
        if (t->pending64) {
            // the last fast solve's FP64 times, field-major, once
            const int groups = (t->nfields + kLanes - 1) / kLanes;
            dim3 ug(blocks_for(t->N, 64), groups);
            eik_fast_unpack<double><<<ug, 256, 0, t->ctx->stream>>>(t->T64.p, t->N, t->field_of_slot_dev.p, nullptr,
                                                                    t->tab64.p);
            TS_HIP(hipGetLastError());
            TS_HIP(hipStreamSynchronize(t->ctx->stream));
            t->pending64 = false;
            t->have64 = true;
        }
        if (t->have64) {
            copy_to_host(t->stage, t->ctx->stream, out, t->tab64.p + at, n * 8);
        } else {
            // widen on the device in slabs, so the host never holds a second copy
            const size_t slab = std::min<size_t>(n, 32u << 20);
            dbuf<double> w(slab);
            for (size_t q = 0; q < n; q += slab) {
                size_t len = std::min(slab, n - q);
                eik_f32_to_f64<<<blocks_for(len, 256), 256, 0, t->ctx->stream>>>(t->tab32.p + at + q, w.p, len);
                TS_HIP(hipMemcpyAsync(out + q, w.p, len * 8, hipMemcpyDeviceToHost, t->ctx->stream));
                TS_HIP(hipStreamSynchronize(t->ctx->stream));
            }
        }
        return TS_OK;
    });
}

int ts_tables_read_f32(ts_tableset *t, int first, int count, float *out)
{
    if (!t || !out || first < 0 || count < 0 || first + count > t->nfields || !t->solved) return TS_EINVAL;
    std::lock_guard<std::mutex> lk(t->gate);
    return guarded(nullptr, 0, [&]() -> int {
        TS_HIP(hipSetDevice(t->ctx->device));
        copy_to_host(t->stage, t->ctx->stream, out, t->tab32.p + (size_t)first * t->N, (size_t)count * t->N * 4);
        return TS_OK;
    });
}

int ts_tables_rounds(ts_tableset *t, int32_t *rounds)
{
    if (!t || !rounds) return TS_EINVAL;
    std::copy(t->rounds.begin(), t->rounds.end(), rounds);
    return TS_OK;
}

int ts_tables_stats(ts_tableset *t, double *evaluations_fp32, double *evaluations_fp64, double *seconds,
                    double *sweep_seconds)
{
    if (!t) return TS_EINVAL;
    if (evaluations_fp32) *evaluations_fp32 = t->evaluations32;
    if (evaluations_fp64) *evaluations_fp64 = t->evaluations64;
    if (seconds) *seconds = t->seconds;
    if (sweep_seconds) *sweep_seconds = t->sweep_seconds;
    return TS_OK;
}

int ts_tables_round_log(ts_tableset *t, int max, int32_t *kind, double *seconds)
{
    if (!t || max < 0 || (max > 0 && (!kind || !seconds))) return TS_EINVAL;
    const int n = (int)std::min<size_t>((size_t)max, t->round_kind.size());
    for (int q = 0; q < n; q++) {
        kind[q] = t->round_kind[q];
        seconds[q] = t->round_seconds[q];
    }
    return (int)t->round_kind.size();
}

int ts_tables_sweep_stats(ts_tableset *t, int32_t rounds[2], int32_t *plane_launches_per_round, double seconds[2])
{
    if (!t) return TS_EINVAL;
    if (rounds) { rounds[0] = t->rounds_by_width[0]; rounds[1] = t->rounds_by_width[1]; }
    if (plane_launches_per_round) *plane_launches_per_round = t->plane_launches_per_round;
    if (seconds) { seconds[0] = t->sweep_by_width[0]; seconds[1] = t->sweep_by_width[1]; }
    return TS_OK;
}

void ts_tables_release(ts_tableset *t)
{
    if (!t) return;
    (void)hipSetDevice(t->ctx->device);
    delete t;
}

int ts_eik_solve(ts_context *ctx, const ts_grid *grid, const double *slowness, int batch,
                 const int64_t *seed_offsets, const int32_t *seed_index, const double *seed_time, double *output,
                 int mode, const volatile int32_t *cancel, int32_t *rounds_out, char *reason, int reason_len)
{
    if (!ctx || !grid || !slowness || batch < 1 || !seed_offsets || !output) {
        set_reason(reason, reason_len, "invalid argument");
        return TS_EINVAL;
    }
    std::lock_guard<std::mutex> lk(ctx->gate);
    ts_tableset *t = nullptr;
    const double *models[1] = {slowness};
    int rc = ts_tables_create(ctx, grid, 1, models, batch, nullptr, seed_offsets, seed_index, seed_time, &t, reason,
                              reason_len);
    if (rc != TS_OK) return rc;
    rc = ts_tables_solve(t, mode, cancel, rounds_out, reason, reason_len);
    if (rc == TS_OK) rc = ts_tables_read(t, 0, batch, output);
    ts_tables_release(t);
    return rc;
}

}  // extern "C"
