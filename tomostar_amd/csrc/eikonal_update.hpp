// The Godunov upwind update of the spherical eikonal, shared by the device kernels and the host
// Relax that gates the device (SphericalEikonal.cs:349-396; EikonalOpenCl.cs:429-606).
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

namespace ts {

// Godunov update in absolute times with Relax's arithmetic (SphericalEikonal.cs:349-396). first[a]
// is the smaller neighbour of axis a (+inf: no term), t2[a] its second-order partner (+inf: none).
// Terms are held in three fixed slots (an absent axis is +inf) and ordered with the stable
// three-element network of SphericalEikonal.cs:246-254, which orders the present terms exactly as
// the insertion sort of the reference orders its compacted list.
template <typename R>
__host__ __device__ inline R godunov_abs(R s, const R first[3], const R t2[3], const R h[3])
{
    const R inf = (R)INFINITY;
    R te[3], he[3], tn[3];
    int used = 0;
#pragma unroll
    for (int a = 0; a < 3; a++) {
        bool on = __builtin_isfinite(first[a]);
        bool so = on && __builtin_isfinite(t2[a]) && t2[a] <= first[a];
        te[a] = !on ? inf : so ? (4 * first[a] - t2[a]) / 3 : first[a];
        he[a] = so ? 2 * h[a] / 3 : h[a];
        tn[a] = first[a];
        used += on;
    }
    if (used == 0) return inf;
#define TS_SWAP(x, y) { R t_ = x; x = y; y = t_; }
    if (te[0] > te[1]) { TS_SWAP(te[0], te[1]) TS_SWAP(he[0], he[1]) TS_SWAP(tn[0], tn[1]) }
    if (te[1] > te[2]) {
        TS_SWAP(te[1], te[2]) TS_SWAP(he[1], he[2]) TS_SWAP(tn[1], tn[2])
        if (te[0] > te[1]) { TS_SWAP(te[0], te[1]) TS_SWAP(he[0], he[1]) TS_SWAP(tn[0], tn[1]) }
    }
#undef TS_SWAP
    R cand = inf;
    bool found = false;
#pragma unroll
    for (int terms = 3; terms >= 1; terms--) {
        if (found || terms > used) continue;
        R sw = 0, swt = 0, swt2 = 0;
#pragma unroll
        for (int q = 0; q < 3; q++) {
            if (q >= terms) continue;
            R w = (R)1 / (he[q] * he[q]);
            sw += w;
            swt += w * te[q];
            swt2 += w * te[q] * te[q];
        }
        R disc = swt * swt - sw * (swt2 - s * s);
        if (disc < 0) continue;
        R root = (swt + sqrt(disc)) / sw;
        bool ok = true;
#pragma unroll
        for (int q = 0; q < 3; q++)
            if (q < terms && root < tn[q]) ok = false;
        if (ok) { cand = root; found = true; }
    }
    return cand;
}

// The spacing weights of an axis: w[0] = 1 / (h h) for a first-order term, w[1] = 1 / (h' h') with
// h' = 2 h / 3 for a second-order one - the values the update above forms from h, computed once per
// grid line (they depend on the node's (j, k) only) with the same operations, so the same bits.
template <typename R>
__host__ __device__ inline void spacing_weights(R h, R w[2])
{
    const R h2 = 2 * h / 3;
    w[0] = (R)1 / (h * h);
    w[1] = (R)1 / (h2 * h2);
}

// godunov_abs with the spacing weights of the three axes given (wa[a][0] first order, wa[a][1]
// second order): the same arithmetic in the same order, without the six divisions that form them.
template <typename R>
__device__ inline R godunov_abs_w(R s, const R first[3], const R t2[3], const R wa[3][2])
{
    const R inf = (R)INFINITY;
    R te[3], we[3], tn[3];
    int used = 0;
#pragma unroll
    for (int a = 0; a < 3; a++) {
        bool on = __builtin_isfinite(first[a]);
        bool so = on && __builtin_isfinite(t2[a]) && t2[a] <= first[a];
        te[a] = !on ? inf : so ? (4 * first[a] - t2[a]) / 3 : first[a];
        we[a] = so ? wa[a][1] : wa[a][0];
        tn[a] = first[a];
        used += on;
    }
    if (used == 0) return inf;
#define TS_SWAP(x, y) { R t_ = x; x = y; y = t_; }
    if (te[0] > te[1]) { TS_SWAP(te[0], te[1]) TS_SWAP(we[0], we[1]) TS_SWAP(tn[0], tn[1]) }
    if (te[1] > te[2]) {
        TS_SWAP(te[1], te[2]) TS_SWAP(we[1], we[2]) TS_SWAP(tn[1], tn[2])
        if (te[0] > te[1]) { TS_SWAP(te[0], te[1]) TS_SWAP(we[0], we[1]) TS_SWAP(tn[0], tn[1]) }
    }
#undef TS_SWAP
    R cand = inf;
    bool found = false;
#pragma unroll
    for (int terms = 3; terms >= 1; terms--) {
        if (found || terms > used) continue;
        R sw = 0, swt = 0, swt2 = 0;
#pragma unroll
        for (int q = 0; q < 3; q++) {
            if (q >= terms) continue;
            const R w = we[q];
            sw += w;
            swt += w * te[q];
            swt2 += w * te[q] * te[q];
        }
        R disc = swt * swt - sw * (swt2 - s * s);
        if (disc < 0) continue;
        R root = (swt + sqrt(disc)) / sw;
        bool ok = true;
#pragma unroll
        for (int q = 0; q < 3; q++)
            if (q < terms && root < tn[q]) ok = false;
        if (ok) { cand = root; found = true; }
    }
    return cand;
}

// FP32 update in the offset form (the float path of EikonalOpenCl.cs:454-496, :568-606 restated for
// plain float times): every term is taken relative to r, the smallest upwind neighbour, so the
// quadratic's terms are cell crossing times and only the final r + x rounds at the size of t.
__device__ inline float godunov_offset(float s, const float first[3], const float t2[3], const float wa[3][2])
{
    const float inf = INFINITY;
    float r = inf;
#pragma unroll
    for (int a = 0; a < 3; a++)
        if (first[a] < r) r = first[a];
    if (!(r < inf)) return inf;
    float te[3], we[3], tn[3];
    int used = 0;
#pragma unroll
    for (int a = 0; a < 3; a++) {
        bool on = __builtin_isfinite(first[a]);
        float of = first[a] - r;
        bool so = on && __builtin_isfinite(t2[a]) && t2[a] <= first[a];
        te[a] = !on ? inf : so ? of + (first[a] - t2[a]) / 3.0f : of;
        we[a] = so ? wa[a][1] : wa[a][0];  // 1 / (h h), or 1 / (h' h') with h' = 2 h / 3
        tn[a] = of;
        used += on;
    }
#define TS_SWAP(x, y) { float t_ = x; x = y; y = t_; }
    if (te[0] > te[1]) { TS_SWAP(te[0], te[1]) TS_SWAP(we[0], we[1]) TS_SWAP(tn[0], tn[1]) }
    if (te[1] > te[2]) {
        TS_SWAP(te[1], te[2]) TS_SWAP(we[1], we[2]) TS_SWAP(tn[1], tn[2])
        if (te[0] > te[1]) { TS_SWAP(te[0], te[1]) TS_SWAP(we[0], we[1]) TS_SWAP(tn[0], tn[1]) }
    }
#undef TS_SWAP
    const float base = te[0];
    float cand = inf;
    bool found = false;
#pragma unroll
    for (int terms = 3; terms >= 1; terms--) {
        if (found || terms > used) continue;
        float sw = 0.0f, swd = 0.0f, swd2 = 0.0f;
#pragma unroll
        for (int q = 0; q < 3; q++) {
            if (q >= terms) continue;
            float d = te[q] - base;
            sw += we[q];
            swd += we[q] * d;
            swd2 += we[q] * d * d;
        }
        float disc = swd * swd - sw * (swd2 - s * s);
        if (disc < 0.0f) continue;
        float root = base + (swd + sqrtf(disc)) / sw;
        bool ok = true;
#pragma unroll
        for (int q = 0; q < 3; q++)
            if (q < terms && root < tn[q]) ok = false;
        if (ok) { cand = root; found = true; }
    }
    return found ? r + cand : inf;
}

// godunov_offset on a node none of whose axes has a second-order term (every t2 +inf, as in a
// first-order phase): the same operations in the same order on what is left of them, so the same bits.
// Without second-order terms te[a] = tn[a] = first[a] - r (+inf for an absent axis, as r is finite) and
// we[a] = wa[a][0], so the divisions by 3, the tn payload of the ordering network and its selections drop
// out.
__device__ inline float godunov_offset_first(float s, const float first[3], const float wa[3][2])
{
    const float inf = INFINITY;
    float r = inf;
#pragma unroll
    for (int a = 0; a < 3; a++)
        if (first[a] < r) r = first[a];
    if (!(r < inf)) return inf;
    float te[3], we[3];
    int used = 0;
#pragma unroll
    for (int a = 0; a < 3; a++) {
        te[a] = first[a] - r;
        we[a] = wa[a][0];
        used += __builtin_isfinite(first[a]);
    }
#define TS_SWAP(x, y) { float t_ = x; x = y; y = t_; }
    if (te[0] > te[1]) { TS_SWAP(te[0], te[1]) TS_SWAP(we[0], we[1]) }
    if (te[1] > te[2]) {
        TS_SWAP(te[1], te[2]) TS_SWAP(we[1], we[2])
        if (te[0] > te[1]) { TS_SWAP(te[0], te[1]) TS_SWAP(we[0], we[1]) }
    }
#undef TS_SWAP
    const float base = te[0];
    float cand = inf;
    bool found = false;
#pragma unroll
    for (int terms = 3; terms >= 1; terms--) {
        if (found || terms > used) continue;
        float sw = 0.0f, swd = 0.0f, swd2 = 0.0f;
#pragma unroll
        for (int q = 0; q < 3; q++) {
            if (q >= terms) continue;
            float d = te[q] - base;
            sw += we[q];
            swd += we[q] * d;
            swd2 += we[q] * d * d;
        }
        float disc = swd * swd - sw * (swd2 - s * s);
        if (disc < 0.0f) continue;
        float root = base + (swd + sqrtf(disc)) / sw;
        bool ok = true;
#pragma unroll
        for (int q = 0; q < 3; q++)
            if (q < terms && root < te[q]) ok = false;
        if (ok) { cand = root; found = true; }
    }
    return found ? r + cand : inf;
}

// The mixed form's updates (no bit-for-bit obligation: the mixed form is held to 1e-4 of Relax). The
// terms of the reference's quadratic (the same selection as godunov_abs: the largest number of terms
// whose root is causal, SphericalEikonal.cs:365-395) are formed as offsets from r, the smallest upwind
// neighbour, and all three candidates are formed without branches (the 64 lanes of a wave are 64 fields
// that choose differently); the result is r + the chosen offset.
//
// godunov_lean64 (the second-order phase on FP64 times, the default): FP64 throughout, with the
// hardware's reciprocal and reciprocal square root refined by two Newton steps each (relative error
// ~1e-16) instead of the correctly rounded division and square root sequences. FP32 arithmetic here
// would add ~1e-8 s of noise to near-degenerate quadratics, enough to keep the phase's 1e-7 s settle
// test from passing (13 rounds instead of 6 on the 400-field case).
// godunov_lean32 (the first-order phase on FP32 times; off by default, TOMOSTAR_LEAN_FIRST=1): the
// quadratic in FP32 with the hardware's square root, reciprocal square root and reciprocal (1 ulp).
// It equals godunov_offset to 1 ulp (2e7 random first-order inputs), and the phase runs 17 % faster,
// but Relax's second-order phase has more than one fixed point: started from first-order times that
// differ by an ulp here and there, it settles elsewhere on large regions (400 fields of 201 x 201 x 101:
// largest distance to Relax 5.7e-5 against 4.6e-6, 9e4 nodes of one field moved by more than 1e-6 s
// against 208), so the first-order phase keeps the correctly rounded offset form.
__device__ inline float godunov_lean32(float s, const float first[3], const float t2[3], const float wa[3][2])
{
    const float r = fminf(first[0], fminf(first[1], first[2]));
    const float inf = INFINITY;
    if (!(r < inf)) return inf;
    float te[3], we[3], tn[3];
#pragma unroll
    for (int a = 0; a < 3; a++) {
        const bool on = first[a] < inf;
        const float of = on ? first[a] - r : inf;
        const bool so = on && t2[a] <= first[a];
        te[a] = so ? of + (first[a] - t2[a]) * (1.0f / 3.0f) : of;
        we[a] = so ? wa[a][1] : wa[a][0];
        tn[a] = of;
    }
#define TS_SWAP(x, y) { auto t_ = x; x = y; y = t_; }
    if (te[0] > te[1]) { TS_SWAP(te[0], te[1]) TS_SWAP(we[0], we[1]) TS_SWAP(tn[0], tn[1]) }
    if (te[1] > te[2]) {
        TS_SWAP(te[1], te[2]) TS_SWAP(we[1], we[2]) TS_SWAP(tn[1], tn[2])
        if (te[0] > te[1]) { TS_SWAP(te[0], te[1]) TS_SWAP(we[0], we[1]) TS_SWAP(tn[0], tn[1]) }
    }
    const float s2 = s * s;
    const float e1 = te[1] - te[0], e2 = te[2] - te[0];
    const float root1 = te[0] + s * __builtin_amdgcn_rsqf(we[0]);
    const float a2 = we[0] + we[1], b2 = we[1] * e1;
    const float disc2 = a2 * s2 - we[0] * we[1] * e1 * e1;
    const float root2 = te[0] + (b2 + __builtin_amdgcn_sqrtf(disc2)) * __builtin_amdgcn_rcpf(a2);
    const float a3 = a2 + we[2], b3 = b2 + we[2] * e2;
    const float c3 = we[1] * e1 * e1 + we[2] * e2 * e2 - s2;
    const float disc3 = b3 * b3 - a3 * c3;
    const float root3 = te[0] + (b3 + __builtin_amdgcn_sqrtf(disc3)) * __builtin_amdgcn_rcpf(a3);
    const bool ok3 = tn[2] < inf && disc3 >= 0.0f && root3 >= tn[0] && root3 >= tn[1] && root3 >= tn[2];
    const bool ok2 = tn[1] < inf && disc2 >= 0.0f && root2 >= tn[0] && root2 >= tn[1];
    const bool ok1 = root1 >= tn[0];
    const float cand = ok3 ? root3 : ok2 ? root2 : ok1 ? root1 : inf;
    return cand < inf ? r + cand : inf;
}

// 1 / x and 1 / sqrt(x) for x > 0: the hardware's estimates and two Newton steps (one step: the
// 400-field solve 0.5 % faster, which does not pay for the weaker bound)
__device__ inline double rcp_nr(double x)
{
    double y = __builtin_amdgcn_rcp(x);
    y = fma(y, fma(-x, y, 1.0), y);
    return fma(y, fma(-x, y, 1.0), y);
}
__device__ inline double rsq_nr(double x)
{
    double y = __builtin_amdgcn_rsq(x);
    y = y * fma(-0.5 * x * y, y, 1.5);
    return y * fma(-0.5 * x * y, y, 1.5);
}
// sqrt(x) for x >= 0 (0 for x <= 0)
__device__ inline double sqrt_nr(double x) { return x > 0 ? x * rsq_nr(x) : 0.0; }

__device__ inline double godunov_lean64(double s, const double first[3], const double t2[3], const double wa[3][2])
{
    // the smallest neighbour (compare and select: the times are never NaN, so no quieting is needed)
    const double r01 = first[1] < first[0] ? first[1] : first[0];
    const double r = first[2] < r01 ? first[2] : r01;
    const double inf = INFINITY;
    if (!(r < inf)) return inf;
    double te[3], we[3], tn[3];
#pragma unroll
    for (int a = 0; a < 3; a++) {
        const bool on = first[a] < inf;
        const double of = on ? first[a] - r : inf;
        const bool so = on && t2[a] <= first[a];
        te[a] = so ? of + (first[a] - t2[a]) * (1.0 / 3.0) : of;
        we[a] = so ? wa[a][1] : wa[a][0];
        tn[a] = of;
    }
    if (te[0] > te[1]) { TS_SWAP(te[0], te[1]) TS_SWAP(we[0], we[1]) TS_SWAP(tn[0], tn[1]) }
    if (te[1] > te[2]) {
        TS_SWAP(te[1], te[2]) TS_SWAP(we[1], we[2]) TS_SWAP(tn[1], tn[2])
        if (te[0] > te[1]) { TS_SWAP(te[0], te[1]) TS_SWAP(we[0], we[1]) TS_SWAP(tn[0], tn[1]) }
    }
#undef TS_SWAP
    const double s2 = s * s;
    const double e1 = te[1] - te[0], e2 = te[2] - te[0];
    const double root1 = te[0] + s * rsq_nr(we[0]);
    const double a2 = we[0] + we[1], b2 = we[1] * e1;
    const double disc2 = a2 * s2 - we[0] * we[1] * e1 * e1;
    const double root2 = te[0] + (b2 + sqrt_nr(disc2)) * rcp_nr(a2);
    const double a3 = a2 + we[2], b3 = b2 + we[2] * e2;
    const double c3 = we[1] * e1 * e1 + we[2] * e2 * e2 - s2;
    const double disc3 = b3 * b3 - a3 * c3;
    const double root3 = te[0] + (b3 + sqrt_nr(disc3)) * rcp_nr(a3);
    const bool ok3 = tn[2] < inf && disc3 >= 0.0 && root3 >= tn[0] && root3 >= tn[1] && root3 >= tn[2];
    const bool ok2 = tn[1] < inf && disc2 >= 0.0 && root2 >= tn[0] && root2 >= tn[1];
    const bool ok1 = root1 >= tn[0];
    const double cand = ok3 ? root3 : ok2 ? root2 : ok1 ? root1 : inf;
    return cand < inf ? r + cand : inf;
}

// Stores and loads coherent across the device's L2 caches (write-through / miss-forcing, "sc1"): the
// data a streaming tile exchanges with the tiles beside it within one launch. Everything else a tile
// touches is its own, or was written by an earlier launch.
template <typename T>
__device__ inline void st_dev(T *p, T v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T>
__device__ inline T ld_dev(T *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// This wave's stores have completed (the fences keep the compiler from moving memory operations
// across the wait).
__device__ inline void wait_stores()
{
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_s_waitcnt(0);
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
}

}  // namespace ts
