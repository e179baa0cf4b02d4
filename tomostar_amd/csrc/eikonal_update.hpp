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

}  // namespace ts
