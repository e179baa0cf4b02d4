// The streaming sweep: one persistent launch per sweep ordering, over a work queue of column tiles.
//
// One ordering of Relax's Gauss-Seidel sweep (SphericalEikonal.cs:284-396) updates node X after every
// node X - e_a and before every node X + e_a. In the ordering's mirrored coordinates (x', y', z') a
// task is a tile of TU x TV z-columns of one 64-field group; its workgroup walks the global diagonal
// planes P = x' + y' + z' from the tile's first to its last, column (a, b) of the tile updating node
// z' = P - (u0 + a) - (v0 + b) at plane P. That is a topological order of the ordering's dependences,
// so the results are those of the plane sweep to the bit, and each node is read from HBM once and
// written once per ordering:
//   - every column keeps a ring of its nodes z'-2 .. z'+3 in LDS (the node's own time, its z
//     neighbours, and the x / y neighbours the tile's other columns read), filled one plane ahead by
//     a load issued a plane earlier;
//   - neighbours outside the tile (two columns deep on each side) are loaded into registers a plane
//     ahead: old times of the downwind tiles (they do not change until this tile has passed them),
//     new times of the upwind tiles once those have published that they completed the plane;
//   - the workgroups take tasks from a queue in an order where every tile follows its upwind tiles
//     (tile diagonal, group, tile), so a tile only ever waits for a tile held by a running workgroup.
//     What tiles exchange within a launch (the times, flags and wakes of the two columns next to a
//     tile edge, and the per-tile count of completed planes) is stored and loaded device-coherent
//     (write-through / miss-forcing), and a wave waits for its stores to complete before its tile
//     publishes a plane: no L2 write-back or invalidation per plane (agent-scope fences cost ~150 us a
//     plane here, 20x the plane's work). A wait that exceeds a spin limit raises an error flag
//     instead of hanging.
// A node is updated only when something it reads changed since its last update (the dirty flag of
// the plane kernels: the byte loaded a plane ahead, or, for changes within the last planes of this
// tile, a per-entry drop bit in LDS), and it raises the dirty flags of its twelve readers as the plane
// kernels do, so the rounds that follow can run either kernel.
//
// Included by eikonal.hip after the fast form's node update (FastArgs, NodeLoad, stream_update).
#pragma once

#include <hip/hip_runtime.h>

#include <climits>
#include <cstdint>

namespace ts {

constexpr int kStreamRing = 6;  // ring entries per column: z'-2 .. z'+2 read at a plane, z'+3 filled

// Queue and progress of the launches of one round (one ordering each).
struct StreamSync {
    int *queue;       // [8]: next task of ordering d
    int *prog;        // [8][slots]: planes completed + 1 per tile (INT_MAX when done)
    int *err;         // set when a wait exceeded the spin limit
};

template <typename R>
struct StreamArgs {
    FastArgs<R> A;
    const uint32_t *tasks;  // tu | tv << 12 | (group - g0) << 24, every tile after its upwind tiles
    int ntasks, g0, ntu, ntv, slots, spin_limit;
    int publish;            // planes between two publications of a tile's progress (and store waits)
    StreamSync sync;
};

// Per-column state carried from one plane to the next (registers).
template <typename R>
struct StreamCol {
    R pre;          // old time of node z'+4: the ring entry the next plane fills
    R s;            // slowness of the next plane's node (this lane's field)
    R hu[2], hv[2]; // the next plane's neighbours outside the tile along x' and y' (nearest first)
    uint64_t fixed; // seed lanes of the next plane's node
    R wx0, wx1, wy0, wy1;  // spacing weights of the next plane's node (x: longitude, y: latitude)
    uint32_t flag;  // the next plane's node was dirty when loaded
};

template <typename R, int TU, int TV, int NW, int DIR, bool kCount, bool kLean, bool kSecond>
__device__ __forceinline__ void stream_tile(const StreamArgs<R> &S, R *ring, uint8_t *drop, int g, int gl, int tu,
                                            int tv, int slot, uint64_t act, uint64_t ph, int d)
{
    constexpr int NC = TU * TV, C = NC / NW, RG = kStreamRing;
    const FastArgs<R> &A = S.A;
    // the wave index through readfirstlane: the compiler then knows each wave's columns are uniform
    const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int nx = A.nx, ny = A.ny, nz = A.nz, nxy = nx * ny;
    const size_t gn = (size_t)g * A.N;
    R *Tg = A.T + gn * kLanes;
    const int ll = load_lane(act, lane);  // loads of an idle lane go to a row the wave fetches anyway
    const int mdl = A.model[g * kLanes + ll];
    const R inf = (R)INFINITY;
    const int u0 = tu * TU, v0 = tv * TV;
    int *prog = S.sync.prog + (size_t)d * S.slots;
    // the upwind tiles this wave's columns read from, and the progress seen of them
    int seen_u = tu > 0 ? 0 : INT_MAX, seen_v = tv > 0 ? 0 : INT_MAX;
    bool need_u = false, need_v = false;
    StreamCol<R> cs[C];
    int ca[C], cb[C];
    bool real[C];
#pragma unroll
    for (int c = 0; c < C; c++) {
        const int q = wave * C + c;
        ca[c] = q % TU;
        cb[c] = q / TU;
        real[c] = u0 + ca[c] < nx && v0 + cb[c] < ny;
        need_u |= real[c] && ca[c] <= 1 && tu > 0;
        need_v |= real[c] && cb[c] <= 1 && tv > 0;
    }
    // actual grid coordinates of mirrored (x', y', z')
    auto ax = [&](int xp) { return (DIR & 1) ? nx - 1 - xp : xp; };
    auto ay = [&](int yp) { return (DIR & 2) ? ny - 1 - yp : yp; };
    auto az = [&](int zp) { return (DIR & 4) ? nz - 1 - zp : zp; };
    auto node_at = [&](int xp, int yp, int zp) { return (az(zp) * ny + ay(yp)) * nx + ax(xp); };
    auto tval = [&](int node) -> R { return Tg[(size_t)node * kLanes + ll]; };

    const int Pb = u0 + v0 - 3, Pe = u0 + TU - 1 + v0 + TV - 1 + nz - 1;
    // prologue: the entry the first plane fills (z' = kc(Pb) + 3)
#pragma unroll
    for (int c = 0; c < C; c++) {
        const int k3 = Pb - (u0 + ca[c]) - (v0 + cb[c]) + 3;
        cs[c].pre = (real[c] && k3 >= 0 && k3 < nz) ? tval(node_at(u0 + ca[c], v0 + cb[c], k3)) : inf;
        cs[c].flag = 0;
    }
    const int wake_off = wake_offset(lane, nx, nxy);
    uint64_t mv = 0;
    unsigned long long evals = 0;
    for (int P = Pb; P <= Pe; P++) {
        // (A) the updates of plane P
#pragma unroll
        for (int c = 0; c < C; c++) {
            const int a = ca[c], b = cb[c], col = b * TU + a;
            const int kc = P - (u0 + a) - (v0 + b);
            if (!real[c] || kc < 0 || kc >= nz) continue;
            const int sk = kc % RG;
            // (a drop two nodes away wakes a node only while a field of the group is in its
            // second-order phase, as in the plane kernels)
            bool dirty = cs[c].flag != 0;
            if (kc >= 1) dirty |= drop[col * RG + (kc - 1) % RG] != 0;
            if (a >= 1) dirty |= drop[(col - 1) * RG + sk] != 0;
            if (b >= 1) dirty |= drop[(col - TU) * RG + sk] != 0;
            if constexpr (kSecond) {
                if (kc >= 2) dirty |= drop[col * RG + (kc - 2) % RG] != 0;
                if (a >= 2) dirty |= drop[(col - 2) * RG + sk] != 0;
                if (b >= 2) dirty |= drop[(col - 2 * TU) * RG + sk] != 0;
            }
            if (!dirty) continue;
            auto ent = [&](int cl, int k) -> R { return ring[((cl * RG) + k % RG) * kLanes + lane]; };
            const int xp = u0 + a, yp = v0 + b;
            const R cur = ent(col, kc);
            // mirrored neighbours: m = toward the upwind side of the ordering, p = downwind
            R zm1 = kc >= 1 ? ent(col, kc - 1) : inf, zp1 = kc + 1 < nz ? ent(col, kc + 1) : inf;
            R xm1 = xp >= 1 ? (a >= 1 ? ent(col - 1, kc) : cs[c].hu[0]) : inf;
            R xp1 = xp + 1 < nx ? (a + 1 < TU ? ent(col + 1, kc) : cs[c].hu[0]) : inf;
            R ym1 = yp >= 1 ? (b >= 1 ? ent(col - TU, kc) : cs[c].hv[0]) : inf;
            R yp1 = yp + 1 < ny ? (b + 1 < TV ? ent(col + TU, kc) : cs[c].hv[0]) : inf;
            R zm2 = inf, zp2 = inf, xm2 = inf, xp2 = inf, ym2 = inf, yp2 = inf;
            if constexpr (kSecond) {
                zm2 = kc >= 2 ? ent(col, kc - 2) : inf;
                zp2 = kc + 2 < nz ? ent(col, kc + 2) : inf;
                xm2 = xp >= 2 ? (a >= 2 ? ent(col - 2, kc) : cs[c].hu[1]) : inf;
                xp2 = xp + 2 < nx ? (a + 2 < TU ? ent(col + 2, kc) : cs[c].hu[1]) : inf;
                ym2 = yp >= 2 ? (b >= 2 ? ent(col - 2 * TU, kc) : cs[c].hv[1]) : inf;
                yp2 = yp + 2 < ny ? (b + 2 < TV ? ent(col + 2 * TU, kc) : cs[c].hv[1]) : inf;
            }
            // the update in the grid's own orientation (axis order and the minus side first, as the
            // plane kernels present it, so ties break the same way)
            NodeLoad<R> L;
            L.i = ax(xp);
            L.j = ay(yp);
            L.k = az(kc);
            L.node = (L.k * ny + L.j) * nx + L.i;
            L.fl = 1;
            L.fixed = cs[c].fixed;
            L.s = cs[c].s;
            L.cur = cur;
            L.mn[0] = (DIR & 1) ? xp1 : xm1; L.pl[0] = (DIR & 1) ? xm1 : xp1;
            L.m2[0] = (DIR & 1) ? xp2 : xm2; L.p2[0] = (DIR & 1) ? xm2 : xp2;
            L.mn[1] = (DIR & 2) ? yp1 : ym1; L.pl[1] = (DIR & 2) ? ym1 : yp1;
            L.m2[1] = (DIR & 2) ? yp2 : ym2; L.p2[1] = (DIR & 2) ? ym2 : yp2;
            L.mn[2] = (DIR & 4) ? zp1 : zm1; L.pl[2] = (DIR & 4) ? zm1 : zp1;
            L.m2[2] = (DIR & 4) ? zp2 : zm2; L.p2[2] = (DIR & 4) ? zm2 : zp2;
            L.wlon0 = cs[c].wx0; L.wlon1 = cs[c].wx1; L.wlat0 = cs[c].wy0; L.wlat1 = cs[c].wy1;
            R next;
            bool moved;
            const bool dev = a <= 1 || a >= TU - 2 || b <= 1 || b >= TV - 2;
            const uint64_t woke = stream_update<R, kLean, kSecond>(A, L, g, act, ph, lane, wake_off, dev, next, moved);
            if (next < cur) ring[((col * RG) + sk) * kLanes + lane] = next;
            if (lane == 0) drop[col * RG + sk] = woke != 0;
            mv |= __ballot(moved);
            if (kCount) evals += (unsigned long long)__popcll(act & ~L.fixed);
        }
        // (B) the ring entry z'+3 of each column, for the planes that follow
#pragma unroll
        for (int c = 0; c < C; c++) {
            const int a = ca[c], b = cb[c], col = b * TU + a;
            const int k3 = P - (u0 + a) - (v0 + b) + 3;
            if (!real[c] || k3 < 0 || k3 >= nz) continue;
            ring[((col * RG) + k3 % RG) * kLanes + lane] = cs[c].pre;
            if (lane == 0) drop[col * RG + k3 % RG] = 0;
        }
        // (C) the device-coherent stores this wave made up to plane P have completed before the plane is
        // published (every S.publish planes and at the last; the fences only keep the compiler from
        // moving memory operations across the wait)
        const bool pub = (P - Pb + 1) % S.publish == 0 || P == Pe;
        if (pub) {
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
            __builtin_amdgcn_s_waitcnt(0);
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        }
        // (D) what plane P + 1 reads beyond the ring: old times first (no wait needed)
#pragma unroll
        for (int c = 0; c < C; c++) {
            const int a = ca[c], b = cb[c];
            const int kc = P - (u0 + a) - (v0 + b), kn = kc + 1, k4 = kc + 4;
            const int xp = u0 + a, yp = v0 + b;
            if (!real[c]) continue;
            if (k4 >= 0 && k4 < nz) cs[c].pre = tval(node_at(xp, yp, k4));
            if (kn < 0 || kn >= nz) continue;
            const int nd = node_at(xp, yp, kn);
            cs[c].s = A.slow[(size_t)mdl * A.N + nd];
            cs[c].fixed = A.seedmask[gn + nd];
            const int zj = az(kn) * ny + ay(yp);
            cs[c].wx0 = A.wlon[2 * zj];
            cs[c].wx1 = A.wlon[2 * zj + 1];
            cs[c].wy0 = A.wlat[2 * az(kn)];
            cs[c].wy1 = A.wlat[2 * az(kn) + 1];
            // downwind neighbours outside the tile: old times
            if (a >= TU - 2) {
                if (a == TU - 1) cs[c].hu[0] = xp + 1 < nx ? tval(node_at(xp + 1, yp, kn)) : inf;
                if (kSecond) cs[c].hu[1] = xp + 2 < nx ? tval(node_at(xp + 2, yp, kn)) : inf;
            }
            if (b >= TV - 2) {
                if (b == TV - 1) cs[c].hv[0] = yp + 1 < ny ? tval(node_at(xp, yp + 1, kn)) : inf;
                if (kSecond) cs[c].hv[1] = yp + 2 < ny ? tval(node_at(xp, yp + 2, kn)) : inf;
            }
        }
        // the upwind tiles must have completed plane P before their times of plane P + 1's
        // neighbours (and the dirty flags they raised) are read
        auto wait_for = [&](int &seen, int at) {
            if (seen > P) return;
            int spins = 0;
            for (;;) {
                int v = __hip_atomic_load(&prog[at], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                v = __builtin_amdgcn_readfirstlane(v);
                if (v > P) { seen = v; break; }
                if (++spins > S.spin_limit) {
                    if (lane == 0) atomicExch(S.sync.err, 1);
                    seen = INT_MAX;
                    break;
                }
                __builtin_amdgcn_s_sleep(2);
            }
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        };
        if (need_u) wait_for(seen_u, slot - 1);
        if (need_v) wait_for(seen_v, slot - S.ntu);
#pragma unroll
        for (int c = 0; c < C; c++) {
            const int a = ca[c], b = cb[c];
            const int kn = P - (u0 + a) - (v0 + b) + 1;
            const int xp = u0 + a, yp = v0 + b;
            if (!real[c] || kn < 0 || kn >= nz) continue;
            // times the upwind tiles wrote within this launch, and flags they may have raised: read
            // device-coherent
            auto tdev = [&](int node) -> R { return ld_dev(&Tg[(size_t)node * kLanes + ll]); };
            if (a <= 1) {
                if (a == 0) cs[c].hu[0] = xp >= 1 ? tdev(node_at(xp - 1, yp, kn)) : inf;
                if (kSecond) cs[c].hu[1] = xp >= 2 ? tdev(node_at(xp - 2, yp, kn)) : inf;
            }
            if (b <= 1) {
                if (b == 0) cs[c].hv[0] = yp >= 1 ? tdev(node_at(xp, yp - 1, kn)) : inf;
                if (kSecond) cs[c].hv[1] = yp >= 2 ? tdev(node_at(xp, yp - 2, kn)) : inf;
            }
            uint8_t *fp = &A.dirty[gn + node_at(xp, yp, kn)];
            cs[c].flag = (a <= 1 || b <= 1) ? ld_dev(fp) : *fp;
        }
        __syncthreads();
        if (pub && threadIdx.x == 0) __hip_atomic_store(&prog[slot], P + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    if (kCount && lane == 0 && evals) atomicAdd(&A.evals[blockIdx.x & 1023], evals);
    mark_changed(&A.changed[g], mv, lane);
}

// One sweep ordering (DIR) of every tile of groups g0 .. g0 + groups - 1, as a persistent launch.
template <typename R, int TU, int TV, int NW, int DIR, bool kCount, bool kLean>
__global__ __launch_bounds__(64 * NW) void eik_stream(StreamArgs<R> S)
{
    constexpr int NC = TU * TV;
    static_assert(NC % NW == 0 && TU >= 4 && TV >= 4, "tile shape");
    extern __shared__ __align__(16) unsigned char stream_lds[];
    R *ring = (R *)stream_lds;
    uint8_t *drop = stream_lds + (size_t)NC * kStreamRing * kLanes * sizeof(R);
    __shared__ int task_sh;
    int *prog = S.sync.prog + (size_t)DIR * S.slots;
    // Every branch around a barrier is on a value the compiler can see is uniform (kernel arguments,
    // readfirstlane results), and the thread-0 stores are each fenced by barriers: a loop exit on a
    // value read from LDS counts as divergent, and the compiler then lets thread 0 leave the loop body
    // for its own stores while the rest of its wave runs the barriers ahead - the waves' barrier counts
    // part and the launch never ends.
    for (;;) {
        if (threadIdx.x == 0) task_sh = atomicAdd(&S.sync.queue[DIR], 1);
        __syncthreads();
        const int task = __builtin_amdgcn_readfirstlane(task_sh);
        __syncthreads();
        if (task >= S.ntasks) break;
        const uint32_t pk = S.tasks[task];
        const int tu = pk & 4095, tv = (pk >> 12) & 4095, gl = (int)(pk >> 24), g = S.g0 + gl;
        const int slot = (gl * S.ntv + tv) * S.ntu + tu;
        const uint64_t act = S.A.active[g];
        if (act != 0) {
            const uint64_t ph = S.A.phase1[g];
            if constexpr (kFirstOnly<R>)
                stream_tile<R, TU, TV, NW, DIR, kCount, kLean, false>(S, ring, drop, g, gl, tu, tv, slot, act, ph, DIR);
            else if (kSecondOnly<R, kLean> || (ph & act) != 0)
                stream_tile<R, TU, TV, NW, DIR, kCount, kLean, true>(S, ring, drop, g, gl, tu, tv, slot, act, ph, DIR);
            else
                stream_tile<R, TU, TV, NW, DIR, kCount, kLean, false>(S, ring, drop, g, gl, tu, tv, slot, act, ph, DIR);
        }
        // the tile is done: its downwind tiles wait no more
        __syncthreads();
        if (threadIdx.x == 0) __hip_atomic_store(&prog[slot], INT_MAX, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __syncthreads();
    }
}

// Queues and progress of the eight launches of a round start from zero.
__global__ void eik_stream_reset(int *queue, int *prog, int n)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q < 8) queue[q] = 0;
    if (q < n) prog[q] = 0;
}

}  // namespace ts
