// Many eikonal problems, each on its own grid, solved together: the uneven batch of an adaptive or
// resolution-test run, where the reference calls EikonalOpenCl.TrySolve once per grid
// (Forward/EikonalOpenCl.cs:138-343) and each call sweeps its own planes.
//
// Here every sweep step is one launch over all problems: plane p of ordering dir of every problem
// that has a plane p. A per-step segment table is the launch's work queue - segment (problem,
// first node in the problem's plane list, node count, first block) - and a block finds its segment
// by binary search. Lanes are nodes (one field per problem, so there is nothing to batch across
// lanes), times are stored per problem in node order. The update, dirty flags, settle rule, round
// cap and phases are those of Relax (SphericalEikonal.cs:284-396): the FP64 form reproduces Relax to
// the last bit per problem; the mixed form runs the first-order phase on FP32 times and the
// second-order phase on FP64 times, as the batched fast form does.
//
// A tuner picks, per size class of problem (by node count), how many nodes a thread updates per
// step (1, 2 or 4), by timing one trial round of the class alone with each choice.
//
// The streaming sweep (mg_stream) instead runs each ordering as ONE persistent launch over a work
// queue of column tiles of every problem: a wave takes a tile of 64 z-columns of one problem (one
// lane per column) and walks the ordering's diagonal planes x' + y' + z' = P through it, lane (a, b)
// updating node z' = P - x' - y' of its column. A lane keeps its column's nodes z'-2 .. z'+2 in
// registers and takes its x / y neighbours from the lanes beside it (shuffles; the lanes are in
// lockstep, so a neighbour upwind already holds this plane's new time and one downwind still its
// old one), and the neighbours beyond the tile's edge from memory, a plane ahead: old times of the
// tiles downwind, new times of the tiles upwind once those have published the plane. Tiles are
// queued so that each follows its upwind tiles; times, flags and progress that cross a tile edge
// within a launch are stored and loaded device-coherent, and a wait beyond a spin limit raises an
// error flag instead of hanging. A node is updated when the plane kernels would update it (its dirty
// flag, or a drop of an upwind neighbour in the last two planes, which the flag cannot show yet), so
// the results equal the plane sweep's to the bit. The tuner picks the tile width per size class and
// the sweep (planes or streaming) for the batch.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <memory>
#include <numeric>
#include <vector>

#include "common.hpp"
#include "eikonal_update.hpp"
#include "geometry.hpp"

namespace ts {
void grid_metric(const Grid &g, double *hlon, double *hlat, double *hr);
}

using namespace ts;

struct MgProblem {
    int nx, ny, nz, nodes_per_thread;
    int64_t node_base;   // first node in the concatenated arrays
    int wlon_base, wlat_base;
    double wr0, wr1;     // depth-axis spacing weights (FP64; the FP32 kernel narrows them)
    float wr0f, wr1f;
    // streaming sweep: tile width along x' (4, 8 or 16 columns; 64 / width along y'), tiles per
    // axis and the problem's first progress slot
    int tile_w, ntu, ntv, slot_base;
};

struct MgSeg {
    int problem, list_off, count, block_start;
};

// A chunk of the problems sweeping on its own stream: its segment tables and, per width, its graph of
// plane launches.
struct MgChunk {
    std::vector<int> seg_off, seg_blocks;
    dbuf<MgSeg> segs;
    hipGraphExec_t graph[2] = {nullptr, nullptr};
    double graph_wake[2] = {0, 0};        // the wake tolerance each graph was captured with
};

struct ts_gridbatch {
    ts_context *ctx = nullptr;
    // streaming sweep: method (0 plane launches, 1 streaming), tile width per size class, the task
    // list (problem | tu << 32 | tv << 48, each tile after its upwind tiles), queues and progress
    int method = 0;
    int tile_w[3] = {8, 8, 8};
    dbuf<uint64_t> stasks;
    int sntasks = 0, sslots = 0;
    dbuf<int32_t> squeue, sprog;
    int spin_limit = 1 << 18;
    int publish = 4;  // planes between two publications of a tile's progress (measured: 1, 2, 4, 8 planes
                      // take the 64-problem mixed solve 0.373, 0.342, 0.335, 0.351 s)
    int nprob = 0;
    std::vector<Grid> grids;
    std::vector<int64_t> base;            // node base per problem (+ total)
    std::vector<std::vector<int>> plane_off;
    std::vector<int64_t> list_base;       // plane-list base per problem
    std::vector<MgProblem> prob;
    std::vector<std::vector<int32_t>> seed_node;
    std::vector<std::vector<double>> seed_time;
    int maxP = 0;
    dbuf<uint32_t> planes;
    dbuf<double> slow64, wlon64, wlat64, T64;
    dbuf<float> slow32, wlon32, wlat32, T32;
    dbuf<uint8_t> dirty, seeded;
    dbuf<MgProblem> prob_dev;
    dbuf<int32_t> phase, changed, rounds, status, reset, cap;
    dbuf<unsigned long long> evals;       // executed updates of the last solve: FP32, FP64
    double updates[2] = {0, 0};
    dbuf<MgSeg> segs;
    std::vector<int> seg_off;             // per plane index: first segment
    std::vector<int> seg_blocks;          // per plane index: blocks of its launch
    std::vector<int32_t> rounds_host;
    std::vector<int> tuned;               // nodes per thread chosen per size class
    double tune_ms[3] = {0, 0, 0};        // the tuner's trial mixed solve of the batch per method (ms)
    double seconds = 0, sweep_seconds = 0;
    std::vector<int32_t> round_kind;      // per round of the last solve: 16 for FP64 times | method
    std::vector<double> round_seconds;    // and its device seconds
    int last_mode = -1;
    double wake_tol[2] = {0, 0};          // MgArgs::wake_tol of the running FP32 / FP64 sweeps
    static constexpr int kMaxChunks = 4;
    MgChunk chunk[kMaxChunks];
    int nchunks = 0;                      // > 1: the problems sweep in chunks side by side
    hipStream_t cstream[kMaxChunks] = {};
    hipEvent_t cfork = nullptr, cjoin[kMaxChunks] = {};
    std::mutex gate;
    ~ts_gridbatch()
    {
        for (auto &c : chunk)
            for (auto &g : c.graph)
                if (g) (void)hipGraphExecDestroy(g);
        for (auto &sx : cstream)
            if (sx) (void)hipStreamDestroy(sx);
        for (auto &e : cjoin)
            if (e) (void)hipEventDestroy(e);
        if (cfork) (void)hipEventDestroy(cfork);
    }
};

namespace ts {
namespace {

constexpr double kMgSettleTol = 1e-7;  // SphericalEikonal.cs:332
// The mixed form's first-order phase on FP32 times: a node has moved when it dropped by more than the
// settle tolerance and by more than this many FP32 ulps of its time, the batched form's rule (its
// second-order phase takes the times the rest of the way down).
constexpr float kMgFp32SettleUlps = 4096;
// The mixed form's second-order phase wakes a node's readers only on drops above this (s), as the
// batched form's does; the FP64 form wakes them on every drop, as Relax does.
constexpr double kMgMixedWakeTol = 1e-8;
constexpr int kMgBlock = 256;
constexpr int kWave = 64;

inline int mg_round_cap(const Grid &g)  // SphericalEikonal.cs:335
{
    const int c = (g.nx + g.ny + g.nz) / 4;
    return c > 32 ? c : 32;
}

template <typename R>
struct MgArgs {
    R *T;
    uint8_t *dirty;
    const uint8_t *seeded;
    const R *slow, *wlon, *wlat;
    const MgProblem *prob;
    const uint32_t *planes;
    const int32_t *phase;
    int32_t *changed;
    double wake_tol;  // a drop wakes the node's readers when larger than this (0: any drop)
    unsigned long long *evals;  // executed updates (one counter per width)
};

// Returns whether the node was evaluated (dirty, not a seed, positive slowness).
template <typename R>
__device__ inline bool mg_update(const MgArgs<R> &A, const MgProblem &P, int q, int second, int dir, uint32_t pk)
{
    const int nx = P.nx, ny = P.ny, nz = P.nz, nxy = nx * ny;
    const int ip = pk & 2047, jp = (pk >> 11) & 2047, kp = pk >> 22;
    const int i = (dir & 1) ? nx - 1 - ip : ip;
    const int j = (dir & 2) ? ny - 1 - jp : jp;
    const int k = (dir & 4) ? nz - 1 - kp : kp;
    const int node = (k * ny + j) * nx + i;
    const int64_t at = P.node_base + node;
    if (A.seeded[at] || !A.dirty[at]) return false;
    A.dirty[at] = 0;
    const R s = A.slow[at];
    if (!(s > (R)0)) return false;
    R *Tp = A.T + P.node_base;
    const R cur = Tp[node];
    const int pos[3] = {i, j, k}, n[3] = {nx, ny, nz}, st[3] = {1, nx, nxy};
    R first[3], t2[3];
#pragma unroll
    for (int ax = 0; ax < 3; ax++) {
        const R mn = pos[ax] > 0 ? Tp[node - st[ax]] : (R)INFINITY;
        const R pl = pos[ax] + 1 < n[ax] ? Tp[node + st[ax]] : (R)INFINITY;
        int sec;
        if (mn <= pl) { first[ax] = mn; sec = pos[ax] > 1 ? node - 2 * st[ax] : -1; }
        else { first[ax] = pl; sec = pos[ax] + 2 < n[ax] ? node + 2 * st[ax] : -1; }
        t2[ax] = (second && sec >= 0 && isfinite(first[ax])) ? Tp[sec] : (R)INFINITY;
    }
    R next;
    if constexpr (sizeof(R) == 4) {
        const float wa[3][2] = {{A.wlon[P.wlon_base + 2 * (k * ny + j)], A.wlon[P.wlon_base + 2 * (k * ny + j) + 1]},
                                {A.wlat[P.wlat_base + 2 * k], A.wlat[P.wlat_base + 2 * k + 1]}, {P.wr0f, P.wr1f}};
        next = godunov_offset(s, first, t2, wa);
    } else {
        const double wa[3][2] = {{A.wlon[P.wlon_base + 2 * (k * ny + j)], A.wlon[P.wlon_base + 2 * (k * ny + j) + 1]},
                                 {A.wlat[P.wlat_base + 2 * k], A.wlat[P.wlat_base + 2 * k + 1]}, {P.wr0, P.wr1}};
        next = godunov_abs_w<double>(s, first, t2, wa);
    }
    if (!(next < cur)) return true;
    Tp[node] = next;
    bool moved;
    if constexpr (sizeof(R) == 4)
        moved = !isfinite(cur) || (cur - next) > fmaxf((float)kMgSettleTol, kMgFp32SettleUlps * 1.1920929e-7f * cur);
    else
        moved = !isfinite(cur) || (double)(cur - next) > kMgSettleTol;
    if (moved) atomicOr(&A.changed[q], 1);
    if (isfinite(cur) && !((double)(cur - next) > A.wake_tol)) return true;
    uint8_t *d = A.dirty + P.node_base;
#pragma unroll
    for (int ax = 0; ax < 3; ax++) {
        if (pos[ax] > 0) d[node - st[ax]] = 1;
        if (pos[ax] > 1) d[node - 2 * st[ax]] = 1;
        if (pos[ax] + 1 < n[ax]) d[node + st[ax]] = 1;
        if (pos[ax] + 2 < n[ax]) d[node + 2 * st[ax]] = 1;
    }
    return true;
}

// One sweep step over every problem: block b serves the segment whose block range holds b.
template <typename R>
__global__ __launch_bounds__(kMgBlock) void mg_plane(MgArgs<R> A, const MgSeg *__restrict__ segs, int nsegs, int dir)
{
    __shared__ MgSeg sg;
    __shared__ unsigned block_evals;
    if (threadIdx.x < kWave) {
        // the last segment whose first block is <= blockIdx.x: the first wave reads 64 segment starts at
        // a time and ballots (one wait for memory per 64 segments, where a binary search by one thread
        // waits once per halving)
        const int lane = threadIdx.x;
        int found = 0;
        for (int c0 = 0; c0 < nsegs; c0 += kWave) {
            const int q = c0 + lane;
            const uint64_t le = __ballot(q < nsegs && segs[q].block_start <= (int)blockIdx.x);
            if (le) found = c0 + 63 - __builtin_clzll(le);
            if (le != ~0ull) break;  // starts ascend: a segment beyond this block ends the search
        }
        if (lane == 0) {
            block_evals = 0;
            sg = segs[found];
        }
    }
    __syncthreads();
    const int q = sg.problem;
    const int ph = A.phase[q];
    if (ph >= 2) return;  // done, held or unsettled: not swept
    const MgProblem P = A.prob[q];
    const int per = P.nodes_per_thread;
    const int t = (blockIdx.x - sg.block_start) * kMgBlock + threadIdx.x;
    const int width = ((sg.count + per - 1) / per + kMgBlock - 1) / kMgBlock * kMgBlock;  // threads of the segment
    unsigned done = 0;
    for (int r = 0; r < per; r++) {
        const int a = t + r * width;
        if (a < sg.count) done += mg_update<R>(A, P, q, ph, dir, A.planes[sg.list_off + a]);
    }
    // executed updates: one count per block
    if (done) atomicAdd(&block_evals, done);
    __syncthreads();
    if (threadIdx.x == 0 && block_evals) atomicAdd(&A.evals[sizeof(R) == 8], (unsigned long long)block_evals);
}

// After a round (one thread per problem): a problem whose round moved no node by more than the
// tolerance ends its phase; one that reaches the round cap is unsettled. phase: 0 first order,
// 1 second order, 2 done, 3 first order settled and held (hold != 0), 4 unsettled.
__global__ void mg_round_end(int nprob, int32_t *phase, int32_t *changed, int32_t *rounds, const int32_t *cap,
                             int32_t *status, int32_t *reset, int hold)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nprob) return;
    int ph = phase[q];
    reset[q] = 0;
    if (ph < 2) {
        const int r = ++rounds[2 * q + ph];
        if (!changed[q]) {
            if (ph == 0) {
                ph = hold ? 3 : 1;
                reset[q] = !hold;
            } else ph = 2;
        } else if (r >= cap[q]) ph = 4;
        phase[q] = ph;
    }
    changed[q] = 0;
    if (ph < 2) atomicAdd(&status[0], 1);
    if (ph == 4) atomicAdd(&status[1], 1);
}

// every node of the problems flagged in reset[] (or all problems: reset null) dirty again
__global__ void mg_reset_dirty(uint8_t *dirty, const int64_t *base, int nprob, const int32_t *reset, int64_t total)
{
    const int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (q >= total) return;
    int lo = 0, hi = nprob - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (base[mid] <= q) lo = mid;
        else hi = mid - 1;
    }
    if (!reset || reset[lo]) dirty[q] = 1;
}

__global__ void mg_promote(int nprob, int32_t *phase, int32_t *changed)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nprob) return;
    if (phase[q] == 3) phase[q] = 1;
    changed[q] = 0;
}

template <typename R>
__global__ void mg_clear(R *T, uint8_t *dirty, uint8_t *seeded, int64_t total)
{
    const int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (q >= total) return;
    T[q] = (R)INFINITY;
    dirty[q] = 1;
    seeded[q] = 0;
}

template <typename R>
__global__ void mg_seed(R *T, uint8_t *seeded, const int64_t *at, const double *time, int count)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= count) return;
    T[at[q]] = (R)time[q];
    seeded[at[q]] = 1;
}

__global__ void mg_widen(const float *a, double *b, int64_t n)
{
    const int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (q < n) b[q] = (double)a[q];
}

// ------------------------------------------------------------------------------ the streaming sweep

template <typename R>
struct MgStreamArgs {
    MgArgs<R> A;
    const uint64_t *tasks;  // problem | tu << 32 | tv << 48, each tile after its upwind tiles
    int ntasks, slots, spin_limit, publish;
    int32_t *queue;         // [8]: next task of ordering d
    int32_t *prog;          // [8][slots]: planes completed + 1 per tile (INT_MAX when done)
    int32_t *err;           // set when a wait exceeded the spin limit
};

// One tile of problem q (tile (tu, tv) of the ordering's mirrored coordinates), one wave, one lane per
// column: the planes from the first that any column's node z' = 0 needs loaded to the last.
template <typename R, int DIR, bool kSecond>
__device__ __forceinline__ void mg_stream_tile(const MgStreamArgs<R> &S, const MgProblem &P, int q, int tu, int tv,
                                               int slot, int32_t *prog, int lane, unsigned long long &evals,
                                               bool &moved_any)
{
    const MgArgs<R> &A = S.A;
    const int nx = P.nx, ny = P.ny, nz = P.nz, nxy = nx * ny;
    const int TU = P.tile_w, TV = kWave / TU, lg = TU == 4 ? 2 : TU == 8 ? 3 : 4;
    const int a = lane & (TU - 1), b = lane >> lg;
    const int u0 = tu * TU, v0 = tv * TV, xp = u0 + a, yp = v0 + b;
    const bool real = xp < nx && yp < ny;
    const int x = (DIR & 1) ? nx - 1 - xp : xp, y = (DIR & 2) ? ny - 1 - yp : yp;
    const int colb = real ? y * nx + x : 0;
    auto az = [&](int zp) { return (DIR & 4) ? nz - 1 - zp : zp; };
    const int sx = (DIR & 1) ? -1 : 1, sy = (DIR & 2) ? -nx : nx;  // node offset of mirrored +1 along x', y'
    R *Tp = A.T + P.node_base;
    uint8_t *Dp = A.dirty + P.node_base;
    const uint8_t *Sd = A.seeded + P.node_base;
    const R *Sl = A.slow + P.node_base;
    const R inf = (R)INFINITY;
    // lanes whose stores tiles downwind read within the launch, and whose wakes may cross an edge
    const bool dev = a >= TU - 2 || b >= TV - 2;
    const bool edge = dev || a <= 1 || b <= 1;
    const int Pb = u0 + v0 - 3, Pe = u0 + TU - 1 + v0 + TV - 1 + nz - 1;
    int seen_u = tu > 0 ? 0 : INT_MAX, seen_v = tv > 0 ? 0 : INT_MAX;
    // w[i]: this column's time at z' = kc - 2 + i (+inf off the grid or not loaded); pre: z' = kc + 3
    R w[5] = {inf, inf, inf, inf, inf};
    R pre = inf;
    {
        const int k3 = Pb - xp - yp + 3;
        if (real && k3 >= 0 && k3 < nz) pre = Tp[az(k3) * nxy + colb];
    }
    // what plane kc and the next read beyond the window ([0]: this plane, [1]: the next, loaded two
    // planes ahead): slowness, seed mark, dirty flag, and the neighbours beyond the tile (nearest first)
    R sl[2] = {(R)0, (R)0}, hx[2][2] = {{inf, inf}, {inf, inf}}, hy[2][2] = {{inf, inf}, {inf, inf}};
    uint32_t seed[2] = {1, 1}, flag[2] = {0, 0};
    uint32_t d1 = 0, d2 = 0;  // this column dropped its node at z' = kc - 1, kc - 2 (this sweep)
    for (int Pl = Pb; Pl <= Pe; Pl++) {
        const int kc = Pl - xp - yp;
        const bool in = real && kc >= 0 && kc < nz;
        // (1) the update of node kc: dirty when its flag (loaded two planes ago) is set or an upwind
        // neighbour dropped in the last two planes (lanes beside it, or this column)
        uint32_t dl = flag[0] | d1 | d2;
        {
            const uint32_t m1 = __shfl(d1, lane - 1), m2 = __shfl(d2, lane - 2);
            const uint32_t n1 = __shfl(d1, lane - TU), n2 = __shfl(d2, lane - 2 * TU);
            if (a >= 1) dl |= m1;
            if (a >= 2) dl |= m2;
            if (b >= 1) dl |= n1;
            if (b >= 2) dl |= n2;
        }
        const bool ev = in && dl != 0 && seed[0] == 0;  // evaluated: its flag is cleared
        uint32_t dnew = 0;
        if (__ballot(ev)) {
            // neighbours at z' = kc: a lane upwind already holds this plane's new time at its w[1] / w[0]
            // (one / two planes ago), a lane downwind still the old time at its w[3] / w[4]
            R xm1 = __shfl(w[1], lane - 1), xm2 = __shfl(w[0], lane - 2);
            R xq1 = __shfl(w[3], lane + 1), xq2 = __shfl(w[4], lane + 2);
            R ym1 = __shfl(w[1], lane - TU), ym2 = __shfl(w[0], lane - 2 * TU);
            R yq1 = __shfl(w[3], lane + TU), yq2 = __shfl(w[4], lane + 2 * TU);
            xm1 = a >= 1 ? xm1 : (xp >= 1 ? hx[0][0] : inf);
            xm2 = a >= 2 ? xm2 : (xp >= 2 ? hx[0][1] : inf);
            xq1 = a + 1 < TU ? xq1 : (xp + 1 < nx ? hx[0][0] : inf);
            xq2 = a + 2 < TU ? xq2 : (xp + 2 < nx ? hx[0][1] : inf);
            ym1 = b >= 1 ? ym1 : (yp >= 1 ? hy[0][0] : inf);
            ym2 = b >= 2 ? ym2 : (yp >= 2 ? hy[0][1] : inf);
            yq1 = b + 1 < TV ? yq1 : (yp + 1 < ny ? hy[0][0] : inf);
            yq2 = b + 2 < TV ? yq2 : (yp + 2 < ny ? hy[0][1] : inf);
            // the grid's own orientation: minus side first on every axis
            const R mn[3] = {(DIR & 1) ? xq1 : xm1, (DIR & 2) ? yq1 : ym1, (DIR & 4) ? w[3] : w[1]};
            const R pl[3] = {(DIR & 1) ? xm1 : xq1, (DIR & 2) ? ym1 : yq1, (DIR & 4) ? w[1] : w[3]};
            const R m2[3] = {(DIR & 1) ? xq2 : xm2, (DIR & 2) ? yq2 : ym2, (DIR & 4) ? w[4] : w[0]};
            const R p2[3] = {(DIR & 1) ? xm2 : xq2, (DIR & 2) ? ym2 : yq2, (DIR & 4) ? w[0] : w[4]};
            const int k = az(kc), node = k * nxy + colb;
            const R cur = w[2];
            if (ev) {
                if (dev) st_dev(&Dp[node], (uint8_t)0);
                else Dp[node] = 0;
            }
            const bool on = ev && sl[0] > (R)0;
            R next = cur;
            if (on) {
                R first[3], t2[3];
#pragma unroll
                for (int ax = 0; ax < 3; ax++) {
                    const bool minus = mn[ax] <= pl[ax];
                    first[ax] = minus ? mn[ax] : pl[ax];
                    t2[ax] = (kSecond && isfinite(first[ax])) ? (minus ? m2[ax] : p2[ax]) : inf;
                }
                const int jl = P.wlon_base + 2 * (k * ny + y), kl = P.wlat_base + 2 * k;
                if constexpr (sizeof(R) == 4) {
                    const float wa[3][2] = {{A.wlon[jl], A.wlon[jl + 1]}, {A.wlat[kl], A.wlat[kl + 1]}, {P.wr0f, P.wr1f}};
                    next = godunov_offset(sl[0], first, t2, wa);
                } else {
                    const double wa[3][2] = {{A.wlon[jl], A.wlon[jl + 1]}, {A.wlat[kl], A.wlat[kl + 1]}, {P.wr0, P.wr1}};
                    next = godunov_abs_w<double>(sl[0], first, t2, wa);
                }
            }
            bool moved = false, dropped = false;
            if (on && next < cur) {
                if (dev) st_dev(&Tp[node], next);
                else Tp[node] = next;
                w[2] = next;
                if constexpr (sizeof(R) == 4)
                    moved = !isfinite(cur) || (cur - next) > fmaxf((float)kMgSettleTol, kMgFp32SettleUlps * 1.1920929e-7f * cur);
                else
                    moved = !isfinite(cur) || (double)(cur - next) > kMgSettleTol;
                dropped = !isfinite(cur) || (double)(cur - next) > A.wake_tol;
            }
            if (dropped) {
                // wake the twelve readers (one and two steps on each axis) inside the grid
                const int pos[3] = {x, y, k}, n[3] = {nx, ny, nz}, st[3] = {1, nx, nxy};
#pragma unroll
                for (int ax = 0; ax < 3; ax++)
#pragma unroll
                    for (int dd = -2; dd <= 2; dd++) {
                        if (dd == 0 || pos[ax] + dd < 0 || pos[ax] + dd >= n[ax]) continue;
                        uint8_t *f = &Dp[node + dd * st[ax]];
                        if (edge) st_dev(f, (uint8_t)1);
                        else *f = 1;
                    }
            }
            dnew = dropped;
            moved_any |= __ballot(moved) != 0;
            evals += (unsigned long long)__popcll(__ballot(on));
        }
        // (2) the window moves one node down the column, and the next plane's reads become this one's
        w[0] = w[1];
        w[1] = w[2];
        w[2] = w[3];
        w[3] = w[4];
        w[4] = pre;
        d2 = d1;
        d1 = dnew;
        sl[0] = sl[1];
        seed[0] = seed[1];
        flag[0] = flag[1];
        hx[0][0] = hx[1][0];
        hx[0][1] = hx[1][1];
        hy[0][0] = hy[1][0];
        hy[0][1] = hy[1][1];
        if ((Pl - Pb + 1) % S.publish == 0 || Pl == Pe) {
            // the stores other tiles read are complete before the plane is published
            wait_stores();
            if (lane == 0) __hip_atomic_store(&prog[slot], Pl + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        // (3) loads for later planes: this column's node kc + 4 (the window's new bottom a plane later),
        // and for node kc + 2 (two planes later) its slowness, seed mark, neighbours beyond the tile
        // and dirty flag
        const int kn = kc + 2, k4 = kc + 4;
        pre = (real && k4 >= 0 && k4 < nz) ? Tp[az(k4) * nxy + colb] : inf;
        const bool inn = real && kn >= 0 && kn < nz;
        const int nn = az(inn ? kn : 0) * nxy + colb;
        sl[1] = inn ? Sl[nn] : (R)0;
        seed[1] = inn ? Sd[nn] : 1;
        if (inn && a >= TU - 2) {  // downwind along x': old times
            hx[1][0] = (a == TU - 1 && xp + 1 < nx) ? Tp[nn + sx] : inf;
            if (kSecond) hx[1][1] = xp + 2 < nx ? Tp[nn + 2 * sx] : inf;
        }
        if (inn && b >= TV - 2) {
            hy[1][0] = (b == TV - 1 && yp + 1 < ny) ? Tp[nn + sy] : inf;
            if (kSecond) hy[1][1] = yp + 2 < ny ? Tp[nn + 2 * sy] : inf;
        }
        // the upwind tiles must have completed plane Pl + 1 (their nodes beside plane Pl + 2's, and
        // the flags they raised)
        auto wait_for = [&](int &seen, int at) {
            if (seen > Pl + 1) return;
            int spins = 0;
            for (;;) {
                int v = __hip_atomic_load(&prog[at], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                v = __builtin_amdgcn_readfirstlane(v);
                if (v > Pl + 1) { seen = v; break; }
                if (++spins > S.spin_limit) {
                    if (lane == 0) atomicExch(S.err, 1);
                    seen = INT_MAX;
                    break;
                }
                __builtin_amdgcn_s_sleep(2);
            }
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        };
        if (tu > 0) wait_for(seen_u, slot - 1);
        if (tv > 0) wait_for(seen_v, slot - P.ntu);
        if (inn && a <= 1) {  // upwind along x': new times
            hx[1][0] = (a == 0 && xp >= 1) ? ld_dev(&Tp[nn - sx]) : inf;
            if (kSecond) hx[1][1] = xp >= 2 ? ld_dev(&Tp[nn - 2 * sx]) : inf;
        }
        if (inn && b <= 1) {
            hy[1][0] = (b == 0 && yp >= 1) ? ld_dev(&Tp[nn - sy]) : inf;
            if (kSecond) hy[1][1] = yp >= 2 ? ld_dev(&Tp[nn - 2 * sy]) : inf;
        }
        flag[1] = !inn ? 0u : (a <= 1 || b <= 1) ? (uint32_t)ld_dev(&Dp[nn]) : (uint32_t)Dp[nn];
    }
}

// One sweep ordering (DIR) of every problem's tiles as a persistent launch: each wave takes tiles from
// the queue until it is empty (the waves of a workgroup are independent).
template <typename R, int DIR>
__global__ __launch_bounds__(256) void mg_stream(MgStreamArgs<R> S)
{
    const int lane = threadIdx.x & 63;
    int32_t *prog = S.prog + (size_t)DIR * S.slots;
    unsigned long long evals = 0;
    for (;;) {
        int task = 0;
        if (lane == 0) task = atomicAdd(&S.queue[DIR], 1);
        task = __builtin_amdgcn_readfirstlane(task);
        if (task >= S.ntasks) break;
        const uint64_t pk = S.tasks[task];
        const int q = (int)(pk & 0xffffffffu), tu = (int)((pk >> 32) & 0xffff), tv = (int)(pk >> 48);
        const MgProblem P = S.A.prob[q];
        const int slot = P.slot_base + tv * P.ntu + tu;
        const int ph = S.A.phase[q];
        bool moved = false;
        if (ph == 1) mg_stream_tile<R, DIR, true>(S, P, q, tu, tv, slot, prog, lane, evals, moved);
        else if (ph == 0) mg_stream_tile<R, DIR, false>(S, P, q, tu, tv, slot, prog, lane, evals, moved);
        if (moved && lane == 0) atomicOr(&S.A.changed[q], 1);
        // done: the tiles downwind wait no more
        wait_stores();
        if (lane == 0) __hip_atomic_store(&prog[slot], INT_MAX, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    if (lane == 0 && evals) atomicAdd(&S.A.evals[sizeof(R) == 8], evals);
}

__global__ void mg_stream_reset(int32_t *queue, int32_t *prog, int n)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q < 8) queue[q] = 0;
    if (q < n) prog[q] = 0;
}

unsigned mg_blocks(int64_t n, int bs) { return (unsigned)std::max<int64_t>(1, (n + bs - 1) / bs); }

int size_class(int64_t nodes) { return nodes < 200000 ? 0 : nodes < 800000 ? 1 : 2; }

// The segment tables of every plane index for the problems in `which` (all when empty).
void build_segments_into(ts_gridbatch &b, const std::vector<int> &which, std::vector<int> &seg_off,
                         std::vector<int> &seg_blocks, dbuf<MgSeg> &segs)
{
    std::vector<MgSeg> all;
    seg_off.assign(b.maxP + 1, 0);
    seg_blocks.assign(b.maxP, 0);
    std::vector<char> on(b.nprob, which.empty() ? 1 : 0);
    for (int q : which) on[q] = 1;
    for (int p = 0; p < b.maxP; p++) {
        seg_off[p] = (int)all.size();
        int blocks = 0;
        for (int q = 0; q < b.nprob; q++) {
            const auto &po = b.plane_off[q];
            if (!on[q] || p + 1 >= (int)po.size()) continue;
            const int cnt = po[p + 1] - po[p];
            if (cnt <= 0) continue;
            const int per = b.prob[q].nodes_per_thread;
            const int nb = (int)mg_blocks((cnt + per - 1) / per, kMgBlock);
            all.push_back(MgSeg{q, (int)(b.list_base[q] + po[p]), cnt, blocks});
            blocks += nb;
        }
        seg_blocks[p] = blocks;
    }
    seg_off[b.maxP] = (int)all.size();
    segs.alloc(std::max<size_t>(1, all.size()));
    if (!all.empty()) TS_HIP(hipMemcpy(segs.p, all.data(), all.size() * sizeof(MgSeg), hipMemcpyHostToDevice));
}

// Chunks sweeping side by side when rounds can run as graphs (TOMOSTAR_GRIDBATCH_STREAMS overrides;
// 0 or 1: every problem in every launch).
int gridbatch_streams(const ts_gridbatch &b)
{
    int n = b.ctx->use_graphs ? 3 : 0;
    if (const char *e = getenv("TOMOSTAR_GRIDBATCH_STREAMS")) n = atoi(e);
    return std::max(0, std::min({n, b.nprob, (int)ts_gridbatch::kMaxChunks}));
}

void build_segments(ts_gridbatch &b, const std::vector<int> &which)
{
    build_segments_into(b, which, b.seg_off, b.seg_blocks, b.segs);
    // the chunks: every problem, largest first to the chunk with the fewest nodes so far
    for (auto &c : b.chunk)
        for (auto &g : c.graph)
            if (g) { (void)hipGraphExecDestroy(g); g = nullptr; }
    b.nchunks = 0;
    const int S = which.empty() ? gridbatch_streams(b) : 0;
    if (S < 2) return;
    std::vector<int> order(b.nprob);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(),
                     [&](int x, int y) { return b.base[x + 1] - b.base[x] > b.base[y + 1] - b.base[y]; });
    std::vector<std::vector<int>> part(S);
    std::vector<int64_t> load(S, 0);
    for (int q : order) {
        const int c = (int)(std::min_element(load.begin(), load.end()) - load.begin());
        part[c].push_back(q);
        load[c] += b.base[q + 1] - b.base[q];
    }
    for (int c = 0; c < S; c++) {
        std::sort(part[c].begin(), part[c].end());
        build_segments_into(b, part[c], b.chunk[c].seg_off, b.chunk[c].seg_blocks, b.chunk[c].segs);
        if (c > 0) {
            if (!b.cstream[c]) TS_HIP(hipStreamCreateWithFlags(&b.cstream[c], hipStreamNonBlocking));
            if (!b.cjoin[c]) TS_HIP(hipEventCreateWithFlags(&b.cjoin[c], hipEventDisableTiming));
        }
    }
    if (!b.cfork) TS_HIP(hipEventCreateWithFlags(&b.cfork, hipEventDisableTiming));
    b.nchunks = S;
}

template <typename R>
MgArgs<R> mg_args(ts_gridbatch &b)
{
    MgArgs<R> A;
    if constexpr (sizeof(R) == 4) { A.T = b.T32.p; A.slow = b.slow32.p; A.wlon = b.wlon32.p; A.wlat = b.wlat32.p; }
    else { A.T = b.T64.p; A.slow = b.slow64.p; A.wlon = b.wlon64.p; A.wlat = b.wlat64.p; }
    A.dirty = b.dirty.p;
    A.seeded = b.seeded.p;
    A.prob = b.prob_dev.p;
    A.planes = b.planes.p;
    A.phase = b.phase.p;
    A.changed = b.changed.p;
    A.wake_tol = b.wake_tol[sizeof(R) == 8];
    A.evals = b.evals.p;
    return A;
}

template <typename R>
void mg_planes(ts_gridbatch &b, hipStream_t st, const std::vector<int> &seg_off, const std::vector<int> &seg_blocks,
               const MgSeg *segs)
{
    const MgArgs<R> A = mg_args<R>(b);
    for (int dir = 0; dir < 8; dir++)
        for (int p = 0; p < b.maxP; p++) {
            const int ns = seg_off[p + 1] - seg_off[p];
            if (ns == 0 || seg_blocks[p] == 0) continue;
            mg_plane<R><<<seg_blocks[p], kMgBlock, 0, st>>>(A, segs + seg_off[p], ns, dir);
        }
}

// The streaming sweep's tiles: per problem the tile width of its size class, its tiles per axis and
// progress slots; the task list by tile diagonal, then problem, then tile, so that every tile follows
// its upwind tiles (tu - 1 and tv - 1 of the same problem).
void build_stream_plan(ts_gridbatch &b)
{
    int slots = 0, maxd = 0;
    for (int q = 0; q < b.nprob; q++) {
        MgProblem &P = b.prob[q];
        const int tw = b.tile_w[size_class(b.base[q + 1] - b.base[q])];
        P.tile_w = tw;
        P.ntu = (P.nx + tw - 1) / tw;
        P.ntv = (P.ny + kWave / tw - 1) / (kWave / tw);
        P.slot_base = slots;
        slots += P.ntu * P.ntv;
        maxd = std::max(maxd, P.ntu - 1 + P.ntv - 1);
    }
    std::vector<uint64_t> list;
    list.reserve(slots);
    for (int d = 0; d <= maxd; d++)
        for (int q = 0; q < b.nprob; q++) {
            const MgProblem &P = b.prob[q];
            for (int tu = std::max(0, d - (P.ntv - 1)); tu <= std::min(P.ntu - 1, d); tu++)
                list.push_back((uint64_t)q | (uint64_t)tu << 32 | (uint64_t)(d - tu) << 48);
        }
    if ((int64_t)list.size() != slots) fail(TS_EDEVICE, "streaming plan covers %zu of %d tiles", list.size(), slots);
    b.stasks.alloc(list.size());
    TS_HIP(hipMemcpy(b.stasks.p, list.data(), list.size() * 8, hipMemcpyHostToDevice));
    b.sntasks = (int)list.size();
    b.sslots = slots;
    b.sprog.alloc((size_t)8 * slots);
    b.squeue.alloc(8);
    TS_HIP(hipMemcpy(b.prob_dev.p, b.prob.data(), b.nprob * sizeof(MgProblem), hipMemcpyHostToDevice));
}

// Workgroups (of four independent waves) of a streaming launch: as many as stay resident.
template <typename R, int DIR>
int mg_stream_grid(const ts_gridbatch &b)
{
    static const int per_cu = [] {
        int n = 0;
        TS_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, mg_stream<R, DIR>, 256, 0));
        return std::max(1, n);
    }();
    const int needed = (b.sntasks + 3) / 4;
    return std::max(1, std::min(needed, per_cu * b.ctx->cu_count));
}

template <typename R, int DIR>
void mg_stream_dir(ts_gridbatch &b, const MgStreamArgs<R> &S, hipStream_t st)
{
    mg_stream<R, DIR><<<mg_stream_grid<R, DIR>(b), 256, 0, st>>>(S);
}

template <typename R>
void mg_stream_planes(ts_gridbatch &b, hipStream_t st)
{
    MgStreamArgs<R> S;
    S.A = mg_args<R>(b);
    S.tasks = b.stasks.p;
    S.ntasks = b.sntasks;
    S.slots = b.sslots;
    S.spin_limit = b.spin_limit;
    S.publish = b.publish;
    S.queue = b.squeue.p;
    S.prog = b.sprog.p;
    S.err = b.status.p + 2;
    const int n = 8 * b.sslots;
    mg_stream_reset<<<mg_blocks(std::max(n, 8), 256), 256, 0, st>>>(b.squeue.p, b.sprog.p, n);
    mg_stream_dir<R, 0>(b, S, st);
    mg_stream_dir<R, 1>(b, S, st);
    mg_stream_dir<R, 2>(b, S, st);
    mg_stream_dir<R, 3>(b, S, st);
    mg_stream_dir<R, 4>(b, S, st);
    mg_stream_dir<R, 5>(b, S, st);
    mg_stream_dir<R, 6>(b, S, st);
    mg_stream_dir<R, 7>(b, S, st);
    TS_HIP(hipGetLastError());
}

void mg_end(ts_gridbatch &b, hipStream_t st, int hold)
{
    mg_round_end<<<mg_blocks(b.nprob, 64), 64, 0, st>>>(b.nprob, b.phase.p, b.changed.p, b.rounds.p, b.cap.p,
                                                       b.status.p, b.reset.p, hold);
}

// The sweep round `round` of a run uses: method 0 plane launches, 1 streaming, 2 (hybrid) plane
// launches for a run's first round (every node dirty: the plane launches keep more updates in flight)
// and streaming after it (few dirty nodes: one launch per ordering instead of one per plane step).
inline int round_method(const ts_gridbatch &b, int round) { return b.method == 2 ? (round == 0 ? 0 : 1) : b.method; }

template <typename R>
void mg_round(ts_gridbatch &b, hipStream_t st, int hold, int method)
{
    if (method == 1) mg_stream_planes<R>(b, st);
    else mg_planes<R>(b, st, b.seg_off, b.seg_blocks, b.segs.p);
    mg_end(b, st, hold);
}

// Chunk c's plane launches as a graph captured on stream cs (the arrays they read stay where they are
// for the batch's life; the segment tables are rebuilt only with the graphs).
template <typename R>
hipGraphExec_t mg_chunk_graph(ts_gridbatch &b, int c, hipStream_t cs)
{
    const int w = sizeof(R) == 8;
    hipGraphExec_t &ge = b.chunk[c].graph[w];
    if (ge && b.chunk[c].graph_wake[w] != b.wake_tol[w]) {  // captured with another tolerance
        TS_HIP(hipGraphExecDestroy(ge));
        ge = nullptr;
    }
    if (ge) return ge;
    b.chunk[c].graph_wake[w] = b.wake_tol[w];
    hipGraph_t gr;
    TS_HIP(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
    mg_planes<R>(b, cs, b.chunk[c].seg_off, b.chunk[c].seg_blocks, b.chunk[c].segs.p);
    TS_HIP(hipStreamEndCapture(cs, &gr));
    hipError_t e = hipGraphInstantiate(&ge, gr, nullptr, nullptr, 0);
    (void)hipGraphDestroy(gr);
    TS_HIP(e);
    return ge;
}

// A round of the chunks side by side: chunk 0 on the batch's stream, the others on their own streams
// between a fork and a join.
template <typename R>
void mg_round_chunks(ts_gridbatch &b, hipStream_t st, int hold)
{
    const int S = b.nchunks;
    std::vector<hipGraphExec_t> g(S);
    for (int c = 0; c < S; c++) g[c] = mg_chunk_graph<R>(b, c, c == 0 ? st : b.cstream[c]);
    TS_HIP(hipEventRecord(b.cfork, st));
    for (int c = 1; c < S; c++) {
        TS_HIP(hipStreamWaitEvent(b.cstream[c], b.cfork, 0));
        TS_HIP(hipGraphLaunch(g[c], b.cstream[c]));
        TS_HIP(hipEventRecord(b.cjoin[c], b.cstream[c]));
    }
    TS_HIP(hipGraphLaunch(g[0], st));
    for (int c = 1; c < S; c++) TS_HIP(hipStreamWaitEvent(st, b.cjoin[c], 0));
    mg_end(b, st, hold);
}

void mg_start(ts_gridbatch &b, bool fp32, hipStream_t st)
{
    const int64_t total = b.base[b.nprob];
    if (fp32) mg_clear<float><<<mg_blocks(total, 256), 256, 0, st>>>(b.T32.p, b.dirty.p, b.seeded.p, total);
    else mg_clear<double><<<mg_blocks(total, 256), 256, 0, st>>>(b.T64.p, b.dirty.p, b.seeded.p, total);
    std::vector<int64_t> at;
    std::vector<double> tm;
    for (int q = 0; q < b.nprob; q++)
        for (size_t s = 0; s < b.seed_node[q].size(); s++) {
            at.push_back(b.base[q] + b.seed_node[q][s]);
            tm.push_back(b.seed_time[q][s]);
        }
    dbuf<int64_t> d_at(std::max<size_t>(1, at.size()));
    dbuf<double> d_tm(std::max<size_t>(1, tm.size()));
    if (!at.empty()) {
        TS_HIP(hipMemcpyAsync(d_at.p, at.data(), at.size() * 8, hipMemcpyHostToDevice, st));
        TS_HIP(hipMemcpyAsync(d_tm.p, tm.data(), tm.size() * 8, hipMemcpyHostToDevice, st));
        if (fp32) mg_seed<float><<<mg_blocks(at.size(), 256), 256, 0, st>>>(b.T32.p, b.seeded.p, d_at.p, d_tm.p, (int)at.size());
        else mg_seed<double><<<mg_blocks(at.size(), 256), 256, 0, st>>>(b.T64.p, b.seeded.p, d_at.p, d_tm.p, (int)at.size());
    }
    TS_HIP(hipMemsetAsync(b.phase.p, 0, b.nprob * 4, st));
    TS_HIP(hipMemsetAsync(b.changed.p, 0, b.nprob * 4, st));
    TS_HIP(hipMemsetAsync(b.rounds.p, 0, 2 * b.nprob * 4, st));
    TS_HIP(hipGetLastError());
    TS_HIP(hipStreamSynchronize(st));
}

// Rounds until no problem relaxes; declines when a problem reaches its round cap.
template <typename R>
void mg_rounds(ts_gridbatch &b, const volatile int32_t *cancel, int hold, int max_rounds, int32_t *h_status,
               const int64_t *d_base)
{
    hipStream_t st = b.ctx->stream;
    hipEvent_t e0, e1;
    TS_HIP(hipEventCreate(&e0));
    TS_HIP(hipEventCreate(&e1));
    struct FreeEv { hipEvent_t a, c; ~FreeEv() { (void)hipEventDestroy(a); (void)hipEventDestroy(c); } } fe{e0, e1};
    const int64_t total = b.base[b.nprob];
    for (int round = 0; round < max_rounds; round++) {
        if (cancel && *cancel) fail(TS_CANCELLED, "cancelled");
        TS_HIP(hipMemsetAsync(b.status.p, 0, 12, st));
        TS_HIP(hipEventRecord(e0, st));
        const int method = round_method(b, round);
        if (b.nchunks > 1 && method == 0) mg_round_chunks<R>(b, st, hold);
        else mg_round<R>(b, st, hold, method);
        TS_HIP(hipEventRecord(e1, st));
        if (!hold) mg_reset_dirty<<<mg_blocks(total, 256), 256, 0, st>>>(b.dirty.p, d_base, b.nprob, b.reset.p, total);
        TS_HIP(hipMemcpyAsync(h_status, b.status.p, 12, hipMemcpyDeviceToHost, st));
        TS_HIP(hipStreamSynchronize(st));
        float ms = 0;
        TS_HIP(hipEventElapsedTime(&ms, e0, e1));
        b.sweep_seconds += ms * 1e-3;
        b.round_kind.push_back((sizeof(R) == 8 ? 16 : 0) | method);
        b.round_seconds.push_back(ms * 1e-3);
        if (h_status[2] > 0) fail(TS_EDEVICE, "streaming sweep: a tile waited for its upwind tiles beyond the spin limit");
        if (h_status[1] > 0) fail(TS_DECLINED, "%d problem(s) not settled within their round cap", h_status[1]);
        if (h_status[0] == 0) return;
    }
    fail(TS_DECLINED, "sweeping not settled");
}

void mg_solve(ts_gridbatch &b, int mode, const volatile int32_t *cancel)
{
    hipStream_t st = b.ctx->stream;
    const bool mixed = mode == TS_EIK_FAST_FP32;
    const int64_t total = b.base[b.nprob];
    dbuf<int64_t> d_base(b.nprob);
    TS_HIP(hipMemcpy(d_base.p, b.base.data(), b.nprob * 8, hipMemcpyHostToDevice));
    int32_t *h_status = nullptr;
    TS_HIP(hipHostMalloc((void **)&h_status, 12, hipHostMallocDefault));
    struct FreeH { int32_t *p; ~FreeH() { if (p) (void)hipHostFree(p); } } fh{h_status};
    int max_rounds = 0;
    for (const Grid &g : b.grids) max_rounds = std::max(max_rounds, 2 * mg_round_cap(g) + 2);
    b.sweep_seconds = 0;
    b.round_kind.clear();
    b.round_seconds.clear();
    TS_HIP(hipMemsetAsync(b.evals.p, 0, 2 * sizeof(unsigned long long), st));
    b.wake_tol[0] = 0.0;
    b.wake_tol[1] = mixed ? kMgMixedWakeTol : 0.0;
    mg_start(b, mixed, st);
    if (mixed) {
        mg_rounds<float>(b, cancel, 1, max_rounds, h_status, d_base.p);
        // second-order phase on FP64 times from the first-order FP32 times; the seeds as given
        std::vector<int64_t> at;
        std::vector<double> tm;
        for (int q = 0; q < b.nprob; q++)
            for (size_t s = 0; s < b.seed_node[q].size(); s++) {
                at.push_back(b.base[q] + b.seed_node[q][s]);
                tm.push_back(b.seed_time[q][s]);
            }
        mg_widen<<<mg_blocks(total, 256), 256, 0, st>>>(b.T32.p, b.T64.p, total);
        dbuf<int64_t> d_at(std::max<size_t>(1, at.size()));
        dbuf<double> d_tm(std::max<size_t>(1, tm.size()));
        if (!at.empty()) {
            TS_HIP(hipMemcpyAsync(d_at.p, at.data(), at.size() * 8, hipMemcpyHostToDevice, st));
            TS_HIP(hipMemcpyAsync(d_tm.p, tm.data(), tm.size() * 8, hipMemcpyHostToDevice, st));
            mg_seed<double><<<mg_blocks(at.size(), 256), 256, 0, st>>>(b.T64.p, b.seeded.p, d_at.p, d_tm.p, (int)at.size());
        }
        mg_promote<<<mg_blocks(b.nprob, 64), 64, 0, st>>>(b.nprob, b.phase.p, b.changed.p);
        mg_reset_dirty<<<mg_blocks(total, 256), 256, 0, st>>>(b.dirty.p, d_base.p, b.nprob, nullptr, total);
        TS_HIP(hipGetLastError());
        TS_HIP(hipStreamSynchronize(st));
    }
    mg_rounds<double>(b, cancel, 0, max_rounds, h_status, d_base.p);
    TS_HIP(hipMemcpy(b.rounds_host.data(), b.rounds.p, 2 * b.nprob * 4, hipMemcpyDeviceToHost));
}

// Per size class, the nodes each thread updates per step: one trial first-order round of the class
// alone with each choice, timed with HIP events; the fastest is kept.
void mg_tune(ts_gridbatch &b)
{
    hipStream_t st = b.ctx->stream;
    b.tuned.assign(3, 1);
    const int choices[3] = {1, 2, 4};
    for (int c = 0; c < 3; c++) {
        std::vector<int> which;
        for (int q = 0; q < b.nprob; q++)
            if (size_class(b.base[q + 1] - b.base[q]) == c) which.push_back(q);
        if (which.empty()) continue;
        double best = 1e300;
        for (int per : choices) {
            for (int q : which) b.prob[q].nodes_per_thread = per;
            TS_HIP(hipMemcpy(b.prob_dev.p, b.prob.data(), b.nprob * sizeof(MgProblem), hipMemcpyHostToDevice));
            build_segments(b, which);
            mg_start(b, true, st);
            hipEvent_t e0, e1;
            TS_HIP(hipEventCreate(&e0));
            TS_HIP(hipEventCreate(&e1));
            TS_HIP(hipEventRecord(e0, st));
            mg_round<float>(b, st, 1, 0);
            TS_HIP(hipEventRecord(e1, st));
            TS_HIP(hipEventSynchronize(e1));
            float ms = 0;
            TS_HIP(hipEventElapsedTime(&ms, e0, e1));
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
            if (ms < best) { best = ms; b.tuned[c] = per; }
        }
        for (int q : which) b.prob[q].nodes_per_thread = b.tuned[c];
    }
    TS_HIP(hipMemcpy(b.prob_dev.p, b.prob.data(), b.nprob * sizeof(MgProblem), hipMemcpyHostToDevice));
    build_segments(b, {});
    // the streaming sweep: per class the tile width whose trial round of the class alone (the other
    // problems held done) is fastest, then the batch's sweep by one trial round of each
    auto trial = [&](int method, const std::vector<int> &which) {
        mg_start(b, true, st);
        if (!which.empty()) {
            std::vector<int32_t> ph(b.nprob, 2);
            for (int q : which) ph[q] = 0;
            TS_HIP(hipMemcpy(b.phase.p, ph.data(), b.nprob * 4, hipMemcpyHostToDevice));
        }
        hipEvent_t e0, e1;
        TS_HIP(hipEventCreate(&e0));
        TS_HIP(hipEventCreate(&e1));
        TS_HIP(hipMemsetAsync(b.status.p, 0, 12, st));
        TS_HIP(hipEventRecord(e0, st));
        mg_round<float>(b, st, 1, method);
        TS_HIP(hipEventRecord(e1, st));
        TS_HIP(hipEventSynchronize(e1));
        float ms = 0;
        TS_HIP(hipEventElapsedTime(&ms, e0, e1));
        int32_t err = 0;
        TS_HIP(hipMemcpy(&err, b.status.p + 2, 4, hipMemcpyDeviceToHost));
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        if (err) fail(TS_EDEVICE, "streaming sweep: a tile waited for its upwind tiles beyond the spin limit");
        return (double)ms;
    };
    const int widths[3] = {4, 8, 16};
    for (int c = 0; c < 3; c++) {
        std::vector<int> which;
        for (int q = 0; q < b.nprob; q++)
            if (size_class(b.base[q + 1] - b.base[q]) == c) which.push_back(q);
        if (which.empty()) continue;
        double best = 1e300;
        int pick = b.tile_w[c];
        for (int tw : widths) {
            b.tile_w[c] = tw;
            build_stream_plan(b);
            const double ms = trial(1, which);
            if (ms < best) { best = ms; pick = tw; }
        }
        b.tile_w[c] = pick;
    }
    build_stream_plan(b);
    int pick_method = b.method;
    // the method: a whole mixed solve with each (a first round alone misleads: streaming is slower
    // while every node is dirty and faster once few are)
    double best = 1e300;
    for (int m = 0; m < 3; m++) {
        b.method = m;
        const auto t0 = std::chrono::steady_clock::now();
        mg_solve(b, TS_EIK_FAST_FP32, nullptr);
        b.tune_ms[m] = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        if (b.tune_ms[m] < best) { best = b.tune_ms[m]; pick_method = m; }
    }
    b.method = pick_method;
}

}  // namespace
}  // namespace ts

// ================================================================================== C ABI

extern "C" {

int ts_gridbatch_create(ts_context *ctx, int nproblems, const ts_grid *grids, const double *const *slowness,
                        const int64_t *seed_offsets, const int32_t *seed_index, const double *seed_time,
                        ts_gridbatch **out, char *reason, int reason_len)
{
    if (out) *out = nullptr;
    if (!ctx || nproblems < 1 || !grids || !slowness || !seed_offsets || !out) {
        set_reason(reason, reason_len, "invalid argument");
        return TS_EINVAL;
    }
    return guarded(reason, reason_len, [&]() -> int {
        TS_HIP(hipSetDevice(ctx->device));
        std::unique_ptr<ts_gridbatch> B(new ts_gridbatch);
        ts_gridbatch &b = *B;
        b.ctx = ctx;
        b.nprob = nproblems;
        b.base.assign(nproblems + 1, 0);
        b.list_base.assign(nproblems, 0);
        b.seed_node.resize(nproblems);
        b.seed_time.resize(nproblems);
        b.plane_off.resize(nproblems);
        std::vector<uint32_t> list;
        std::vector<double> wl, wt;
        std::vector<float> wl32, wt32;
        for (int q = 0; q < nproblems; q++) {
            if (!grid_valid(grids[q])) fail(TS_EINVAL, "grid %d is invalid", q);
            if (!slowness[q]) fail(TS_EINVAL, "slowness of problem %d is null", q);
            const Grid g = make_grid(grids[q]);
            b.grids.push_back(g);
            const int N = g.count();
            b.base[q + 1] = b.base[q] + N;
            // seeds: in-grid, the smallest time per node
            std::vector<std::pair<int32_t, double>> sd;
            for (int64_t s = seed_offsets[q]; s < seed_offsets[q + 1]; s++)
                if (seed_index[s] >= 0 && seed_index[s] < N) sd.emplace_back(seed_index[s], seed_time[s]);
            std::sort(sd.begin(), sd.end());
            for (size_t s = 0; s < sd.size(); s++)
                if (s == 0 || sd[s].first != sd[s - 1].first) {
                    b.seed_node[q].push_back(sd[s].first);
                    b.seed_time[q].push_back(sd[s].second);
                }
            if (b.seed_node[q].empty()) fail(TS_DECLINED, "problem %d has no seed node on its grid", q);
            // plane lists (as the batched fast form's), concatenated
            const int P = g.nx + g.ny + g.nz - 2;
            b.maxP = std::max(b.maxP, P);
            std::vector<int> &po = b.plane_off[q];
            po.assign(P + 1, 0);
            b.list_base[q] = (int64_t)list.size();
            for (int p = 0; p < P; p++) {
                po[p] = (int)(list.size() - b.list_base[q]);
                const int i0 = std::max(0, p - (g.ny - 1) - (g.nz - 1)), i1 = std::min(g.nx - 1, p);
                const int k0 = std::max(0, p - (g.nx - 1) - (g.ny - 1)), k1 = std::min(g.nz - 1, p);
                for (int kp = k0; kp <= k1; kp++) {
                    const int lo = std::max(i0, p - kp - (g.ny - 1)), hi = std::min(i1, p - kp);
                    for (int ip = lo; ip <= hi; ip++)
                        list.push_back((uint32_t)ip | ((uint32_t)(p - ip - kp) << 11) | ((uint32_t)kp << 22));
                }
            }
            po[P] = (int)(list.size() - b.list_base[q]);
            // spacing weights of the problem's metric
            std::vector<double> hl((size_t)g.ny * g.nz), ht(g.nz);
            double hr = 0;
            grid_metric(g, hl.data(), ht.data(), &hr);
            MgProblem P_{};
            P_.nx = g.nx; P_.ny = g.ny; P_.nz = g.nz; P_.nodes_per_thread = 1;
            P_.node_base = b.base[q];
            P_.wlon_base = (int)wl.size();
            P_.wlat_base = (int)wt.size();
            for (double h : hl) {
                double w[2]; float wf[2];
                spacing_weights<double>(h, w);
                spacing_weights<float>((float)h, wf);
                wl.push_back(w[0]); wl.push_back(w[1]);
                wl32.push_back(wf[0]); wl32.push_back(wf[1]);
            }
            for (double h : ht) {
                double w[2]; float wf[2];
                spacing_weights<double>(h, w);
                spacing_weights<float>((float)h, wf);
                wt.push_back(w[0]); wt.push_back(w[1]);
                wt32.push_back(wf[0]); wt32.push_back(wf[1]);
            }
            double w[2]; float wf[2];
            spacing_weights<double>(hr, w);
            spacing_weights<float>((float)hr, wf);
            P_.wr0 = w[0]; P_.wr1 = w[1]; P_.wr0f = wf[0]; P_.wr1f = wf[1];
            b.prob.push_back(P_);
        }
        const int64_t total = b.base[nproblems];
        b.planes.alloc(list.size());
        TS_HIP(hipMemcpy(b.planes.p, list.data(), list.size() * 4, hipMemcpyHostToDevice));
        b.slow64.alloc(total);
        b.slow32.alloc(total);
        std::vector<float> s32;
        for (int q = 0; q < nproblems; q++) {
            const int64_t N = b.base[q + 1] - b.base[q];
            TS_HIP(hipMemcpy(b.slow64.p + b.base[q], slowness[q], N * 8, hipMemcpyHostToDevice));
            s32.assign(slowness[q], slowness[q] + N);
            TS_HIP(hipMemcpy(b.slow32.p + b.base[q], s32.data(), N * 4, hipMemcpyHostToDevice));
        }
        b.wlon64.alloc(wl.size()); b.wlat64.alloc(wt.size());
        b.wlon32.alloc(wl.size()); b.wlat32.alloc(wt.size());
        TS_HIP(hipMemcpy(b.wlon64.p, wl.data(), wl.size() * 8, hipMemcpyHostToDevice));
        TS_HIP(hipMemcpy(b.wlat64.p, wt.data(), wt.size() * 8, hipMemcpyHostToDevice));
        TS_HIP(hipMemcpy(b.wlon32.p, wl32.data(), wl32.size() * 4, hipMemcpyHostToDevice));
        TS_HIP(hipMemcpy(b.wlat32.p, wt32.data(), wt32.size() * 4, hipMemcpyHostToDevice));
        b.T64.alloc(total);
        b.T32.alloc(total);
        b.dirty.alloc(total);
        b.seeded.alloc(total);
        b.prob_dev.alloc(nproblems);
        b.phase.alloc(nproblems); b.changed.alloc(nproblems); b.rounds.alloc(2 * nproblems);
        b.reset.alloc(nproblems); b.cap.alloc(nproblems); b.status.alloc(3);
        b.evals.alloc(2);
        std::vector<int32_t> caps(nproblems);
        for (int q = 0; q < nproblems; q++) caps[q] = mg_round_cap(b.grids[q]);
        TS_HIP(hipMemcpy(b.cap.p, caps.data(), nproblems * 4, hipMemcpyHostToDevice));
        TS_HIP(hipMemcpy(b.prob_dev.p, b.prob.data(), nproblems * sizeof(MgProblem), hipMemcpyHostToDevice));
        build_segments(b, {});
        if (const char *e = getenv("TOMOSTAR_STREAM_SPIN")) b.spin_limit = std::max(0, atoi(e));
        if (const char *e = getenv("TOMOSTAR_STREAM_PUBLISH")) b.publish = std::clamp(atoi(e), 1, 64);
        build_stream_plan(b);
        b.rounds_host.assign(2 * (size_t)nproblems, 0);
        *out = B.release();
        return TS_OK;
    });
}

int ts_gridbatch_tune(ts_gridbatch *b, int32_t choice[3])
{
    if (!b) return TS_EINVAL;
    std::lock_guard<std::mutex> lk(b->gate);
    return guarded(nullptr, 0, [&]() -> int {
        TS_HIP(hipSetDevice(b->ctx->device));
        mg_tune(*b);
        if (choice) for (int c = 0; c < 3; c++) choice[c] = b->tuned[c];
        return TS_OK;
    });
}

int ts_gridbatch_schedule(ts_gridbatch *b, int32_t *method, int32_t tile_w[3], int32_t nodes_per_thread[3],
                          double trial_ms[3])
{
    if (!b) return TS_EINVAL;
    std::lock_guard<std::mutex> lk(b->gate);
    if (method) *method = b->method;
    for (int c = 0; c < 3; c++) {
        if (tile_w) tile_w[c] = b->tile_w[c];
        if (nodes_per_thread) nodes_per_thread[c] = b->tuned.empty() ? 1 : b->tuned[c];
    }
    if (trial_ms)
        for (int m = 0; m < 3; m++) trial_ms[m] = b->tune_ms[m];
    return TS_OK;
}

int ts_gridbatch_round_log(ts_gridbatch *b, int max, int32_t *kind, double *seconds)
{
    if (!b || max < 0 || (max > 0 && (!kind || !seconds))) return TS_EINVAL;
    std::lock_guard<std::mutex> lk(b->gate);
    const int n = (int)std::min<size_t>((size_t)max, b->round_kind.size());
    for (int q = 0; q < n; q++) {
        kind[q] = b->round_kind[q];
        seconds[q] = b->round_seconds[q];
    }
    return (int)b->round_kind.size();
}

int ts_gridbatch_set_schedule(ts_gridbatch *b, int method, const int32_t tile_w[3])
{
    if (!b || method < 0 || method > 2) return TS_EINVAL;
    if (tile_w)
        for (int c = 0; c < 3; c++)
            if (tile_w[c] != 4 && tile_w[c] != 8 && tile_w[c] != 16) return TS_EINVAL;
    std::lock_guard<std::mutex> lk(b->gate);
    return guarded(nullptr, 0, [&]() -> int {
        TS_HIP(hipSetDevice(b->ctx->device));
        b->method = method;
        if (tile_w) {
            for (int c = 0; c < 3; c++) b->tile_w[c] = tile_w[c];
            build_stream_plan(*b);
        }
        return TS_OK;
    });
}

int ts_gridbatch_solve(ts_gridbatch *b, int mode, const volatile int32_t *cancel, int32_t *rounds_out, char *reason,
                       int reason_len)
{
    if (!b || (mode != TS_EIK_FAST_FP32 && mode != TS_EIK_FAST_FP64)) {
        set_reason(reason, reason_len, "invalid argument");
        return TS_EINVAL;
    }
    std::lock_guard<std::mutex> lk(b->gate);
    return guarded(reason, reason_len, [&]() -> int {
        TS_HIP(hipSetDevice(b->ctx->device));
        auto t0 = std::chrono::steady_clock::now();
        b->last_mode = -1;
        mg_solve(*b, mode, cancel);
        b->last_mode = mode;
        b->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        unsigned long long ev[2] = {0, 0};
        TS_HIP(hipMemcpy(ev, b->evals.p, sizeof ev, hipMemcpyDeviceToHost));
        b->updates[0] = (double)ev[0];
        b->updates[1] = (double)ev[1];
        if (rounds_out) std::copy(b->rounds_host.begin(), b->rounds_host.end(), rounds_out);
        return TS_OK;
    });
}

int ts_gridbatch_read(ts_gridbatch *b, int problem, double *out)
{
    if (!b || problem < 0 || problem >= b->nprob || !out) return TS_EINVAL;
    std::lock_guard<std::mutex> lk(b->gate);
    return guarded(nullptr, 0, [&]() -> int {
        if (b->last_mode < 0) fail(TS_EINVAL, "the batch has not been solved");
        TS_HIP(hipSetDevice(b->ctx->device));
        const int64_t N = b->base[problem + 1] - b->base[problem];
        TS_HIP(hipMemcpy(out, b->T64.p + b->base[problem], N * 8, hipMemcpyDeviceToHost));
        return TS_OK;
    });
}

int ts_gridbatch_updates(ts_gridbatch *b, double *fp32, double *fp64)
{
    if (!b) return TS_EINVAL;
    if (fp32) *fp32 = b->updates[0];
    if (fp64) *fp64 = b->updates[1];
    return TS_OK;
}

int ts_gridbatch_stats(ts_gridbatch *b, int64_t *nodes, double *seconds, double *sweep_seconds)
{
    if (!b) return TS_EINVAL;
    if (nodes) *nodes = b->base[b->nprob];
    if (seconds) *seconds = b->seconds;
    if (sweep_seconds) *sweep_seconds = b->sweep_seconds;
    return TS_OK;
}

void ts_gridbatch_release(ts_gridbatch *b)
{
    if (!b) return;
    (void)hipSetDevice(b->ctx->device);
    delete b;
}

}  // extern "C"
