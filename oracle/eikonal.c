/*
 * Oracle (test infrastructure only): the spherical eikonal of the reference in plain C, FP64.
 * Restates Forward/SphericalEikonal.cs: GridMetric (:15-35), PointSource (:68-100), the fast
 * marching Solve with its 4-ary decrease-key heap (:102-260, :404-475) and the Gauss-Seidel
 * diagonal-plane sweeping Relax (:284-396), which is the scheme the device kernels must reproduce.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define OR_R 6371.0
#define OR_D2R (M_PI / 180.0)
#define SETTLE_TOL 1e-7 /* SphericalEikonal.cs:332 SettleTolerance */

static double clampd(double v, double lo, double hi) { return v < lo ? lo : (v > hi ? hi : v); }
static double gd_dlon(const or_griddef *g) { return (g->max_lon - g->min_lon) / (g->nx - 1 > 1 ? g->nx - 1 : 1); }
static double gd_dlat(const or_griddef *g) { return (g->max_lat - g->min_lat) / (g->ny - 1 > 1 ? g->ny - 1 : 1); }
static double gd_ddep(const or_griddef *g) { return (g->max_depth - g->min_depth) / (g->nz - 1 > 1 ? g->nz - 1 : 1); }

/* SphericalEikonal.cs:17-29 GridMetric */
void or_metric(const or_griddef *g, double *hlon, double *hlat, double *hr)
{
    double dlon_rad = gd_dlon(g) * OR_D2R, dlat_rad = gd_dlat(g) * OR_D2R, dd = gd_ddep(g);
    for (int k = 0; k < g->nz; k++) {
        double depth = g->min_depth + k * dd;
        double r = OR_R - depth;
        hlat[k] = r * dlat_rad;
        for (int j = 0; j < g->ny; j++) {
            double c = or_cos((g->min_lat + j * gd_dlat(g)) * OR_D2R);
            double v = r * c * dlon_rad;
            hlon[k * g->ny + j] = v > 1e-9 ? v : 1e-9;
        }
    }
    *hr = dd;
}

/* SphericalEikonal.cs:68-100 PointSource. Returns the seed count (may exceed cap: only cap written). */
int or_point_source(const or_griddef *g, const double *slowness, double lon, double lat, double depth,
                    double origin_time, double radius_cells, int *idx_out, double *time_out, int cap)
{
    double dd = gd_ddep(g);
    double dep = clampd(depth, g->min_depth, g->min_depth + (g->nz - 1) * dd);
    double f[3];
    or_fractional(g, lon, lat, dep, f);
    double fx = clampd(f[0], 0, g->nx - 1), fy = clampd(f[1], 0, g->ny - 1), fz = clampd(f[2], 0, g->nz - 1);
    double src[3];
    or_to_cartesian(lon, lat, dep, src);
    double s_src = or_interpolate(g, slowness, lon, lat, dep);
    double r = OR_R - dep;
    double h_lon = r * or_cos(lat * OR_D2R) * (gd_dlon(g) * OR_D2R);
    if (!(h_lon > 1e-9)) h_lon = 1e-9;
    double h_lat = r * (gd_dlat(g) * OR_D2R);
    double h_r = dd;
    double m = h_lat > h_r ? h_lat : h_r;
    if (h_lon > m) m = h_lon;
    double radius = radius_cells * m;
    int ri = (int)ceil(radius / h_lon), rj = (int)ceil(radius / h_lat), rk = (int)ceil(radius / h_r);
    int i0 = (int)rint(fx), j0 = (int)rint(fy), k0 = (int)rint(fz);
    double *lonv = malloc(sizeof(double) * g->nx), *latv = malloc(sizeof(double) * g->ny),
           *depv = malloc(sizeof(double) * g->nz);
    or_node_coords(g, lonv, latv, depv);
    int count = 0;
    int ka = k0 - rk > 0 ? k0 - rk : 0, kb = k0 + rk < g->nz - 1 ? k0 + rk : g->nz - 1;
    int ja = j0 - rj > 0 ? j0 - rj : 0, jb = j0 + rj < g->ny - 1 ? j0 + rj : g->ny - 1;
    int ia = i0 - ri > 0 ? i0 - ri : 0, ib = i0 + ri < g->nx - 1 ? i0 + ri : g->nx - 1;
    for (int k = ka; k <= kb; k++)
        for (int j = ja; j <= jb; j++)
            for (int i = ia; i <= ib; i++) {
                int idx = (k * g->ny + j) * g->nx + i;
                double c[3];
                or_to_cartesian(lonv[i], latv[j], depv[k], c);
                double dx = c[0] - src[0], dy = c[1] - src[1], dz = c[2] - src[2];
                double d = sqrt(dx * dx + dy * dy + dz * dz);
                int inside = fabs(i - fx) <= 1 && fabs(j - fy) <= 1 && fabs(k - fz) <= 1;
                if (d > radius && !inside) continue;
                if (count < cap) {
                    idx_out[count] = idx;
                    time_out[count] = origin_time + d * 0.5 * (s_src + slowness[idx]);
                }
                count++;
            }
    free(lonv); free(latv); free(depv);
    return count;
}

/* Up to three upwind terms sorted by effective time: SphericalEikonal.cs:262-275 (insertion sort). */
static void sort3(double *te, double *he, double *tn, int used)
{
    for (int a = 1; a < used; a++) {
        double tt = te[a], hh = he[a], nn = tn[a];
        int b = a - 1;
        while (b >= 0 && te[b] > tt) {
            te[b + 1] = te[b]; he[b + 1] = he[b]; tn[b + 1] = tn[b];
            b--;
        }
        te[b + 1] = tt; he[b + 1] = hh; tn[b + 1] = nn;
    }
}

/* The causal root of sum w (t - te)^2 = s^2, dropping terms until it exceeds every neighbour:
 * SphericalEikonal.cs:179-199 and :364-379 (identical arithmetic). */
static double solve_terms(const double *te, const double *he, const double *tn, int used, double s)
{
    for (int terms = used; terms >= 1; terms--) {
        double sw = 0, swt = 0, swt2 = 0;
        for (int q = 0; q < terms; q++) {
            double w = 1.0 / (he[q] * he[q]);
            sw += w;
            swt += w * te[q];
            swt2 += w * te[q] * te[q];
        }
        double disc = swt * swt - sw * (swt2 - s * s);
        if (disc < 0) continue;
        double root = (swt + sqrt(disc)) / sw;
        int ok = 1;
        for (int q = 0; q < terms; q++)
            if (root < tn[q]) { ok = 0; break; }
        if (ok) return root;
    }
    return INFINITY;
}

/* SphericalEikonal.cs:349-396 RelaxNode */
static double relax_node(const or_griddef *g, const double *hlon, const double *hlat, double hr,
                         const double *s_, const double *f, int idx, int i, int j, int k, int second)
{
    double s = s_[idx];
    if (!(s > 0)) return INFINITY;
    double te[3], he[3], tn[3];
    int used = 0;
    int nx = g->nx, ny = g->ny, nz = g->nz, nxy = nx * ny;
    const int pos[3] = {i, j, k}, n[3] = {nx, ny, nz}, stride[3] = {1, nx, nxy};
    const double h[3] = {hlon[k * ny + j], hlat[k], hr};
    for (int a = 0; a < 3; a++) {
        double minus = pos[a] > 0 ? f[idx - stride[a]] : INFINITY;
        double plus = pos[a] + 1 < n[a] ? f[idx + stride[a]] : INFINITY;
        double first;
        int sec;
        if (minus <= plus) { first = minus; sec = pos[a] > 1 ? idx - 2 * stride[a] : -1; }
        else { first = plus; sec = pos[a] + 2 < n[a] ? idx + 2 * stride[a] : -1; }
        if (!isfinite(first)) continue;
        double t2 = second && sec >= 0 ? f[sec] : INFINITY;
        if (isfinite(t2) && t2 <= first) { te[used] = (4 * first - t2) / 3; he[used] = 2 * h[a] / 3; }
        else { te[used] = first; he[used] = h[a]; }
        tn[used] = first;
        used++;
    }
    if (used == 0) return INFINITY;
    sort3(te, he, tn, used);
    return solve_terms(te, he, tn, used, s);
}

/* SphericalEikonal.cs:284-325 Relax (+ :335 SweepRoundCap, :341-347 PlaneBounds).
 * Returns 1 if both phases settled within the cap, 0 otherwise (the device declines such a field,
 * EikonalOpenCl.cs:269); rounds_out gets the rounds each phase took. */
int or_relax(const or_griddef *g, const double *slowness, int nseeds, const int *seed_idx,
             const double *seed_time, int max_rounds, double *a, int rounds_out[2])
{
    int nx = g->nx, ny = g->ny, nz = g->nz, count = nx * ny * nz;
    double *hlon = malloc(sizeof(double) * ny * nz), *hlat = malloc(sizeof(double) * nz), hr;
    or_metric(g, hlon, hlat, &hr);
    unsigned char *fixed = calloc(count, 1);
    for (int q = 0; q < count; q++) a[q] = INFINITY;
    for (int q = 0; q < nseeds; q++) {
        int idx = seed_idx[q];
        if (seed_time[q] < a[idx]) a[idx] = seed_time[q];
        fixed[idx] = 1;
    }
    int cap = max_rounds > 0 ? max_rounds : ((nx + ny + nz) / 4 > 32 ? (nx + ny + nz) / 4 : 32);
    int settled_all = 1;
    for (int phase = 0; phase < 2; phase++) {
        int settled = 0, round;
        for (round = 0; round < cap; round++) {
            int changed = 0;
            for (int dir = 0; dir < 8; dir++)
                for (int p = 0; p < nx + ny + nz - 2; p++) {
                    int i0 = p - (ny - 1) - (nz - 1) > 0 ? p - (ny - 1) - (nz - 1) : 0;
                    int i1 = nx - 1 < p ? nx - 1 : p;
                    int k0 = p - (nx - 1) - (ny - 1) > 0 ? p - (nx - 1) - (ny - 1) : 0;
                    int k1 = nz - 1 < p ? nz - 1 : p;
                    for (int kp = k0; kp <= k1; kp++)
                        for (int ip = i0; ip <= i1; ip++) {
                            int jp = p - ip - kp;
                            if (jp < 0 || jp >= ny) continue;
                            int i = (dir & 1) ? nx - 1 - ip : ip;
                            int j = (dir & 2) ? ny - 1 - jp : jp;
                            int k = (dir & 4) ? nz - 1 - kp : kp;
                            int idx = (k * ny + j) * nx + i;
                            if (fixed[idx]) continue;
                            double cur = a[idx];
                            double cand = relax_node(g, hlon, hlat, hr, slowness, a, idx, i, j, k, phase == 1);
                            if (!(cand < cur)) continue;
                            a[idx] = cand;
                            if (!isfinite(cur) || cur - cand > SETTLE_TOL) changed = 1;
                        }
                }
            if (!changed) { settled = 1; round++; break; }
        }
        if (rounds_out) rounds_out[phase] = round;
        if (!settled) settled_all = 0;
    }
    free(hlon); free(hlat); free(fixed);
    return settled_all;
}

/* ---- fast marching: SphericalEikonal.cs:102-223 and the heap of :404-475 ---------------------- */

typedef struct { double key; int node; } heap_entry;
typedef struct { heap_entry *e; int *at; int count, cap; } front_heap;

static void heap_set(front_heap *h, int node, double key)
{
    int i = h->at[node] - 1;
    if (i < 0) {
        if (h->count == h->cap) { h->cap *= 2; h->e = realloc(h->e, sizeof(heap_entry) * h->cap); }
        i = h->count++;
    }
    while (i > 0) {
        int p = (i - 1) >> 2;
        if (h->e[p].key <= key) break;
        h->e[i] = h->e[p];
        h->at[h->e[i].node] = i + 1;
        i = p;
    }
    h->e[i].key = key;
    h->e[i].node = node;
    h->at[node] = i + 1;
}

static int heap_pop(front_heap *h, int *node)
{
    if (h->count == 0) { *node = -1; return 0; }
    heap_entry *e = h->e;
    *node = e[0].node;
    h->at[*node] = 0;
    heap_entry last = e[--h->count];
    int count = h->count;
    if (count == 0) return 1;
    int i = 0;
    for (;;) {
        int c = 4 * i + 1;
        if (c >= count) break;
        int best = c;
        double best_key = e[c].key;
        int end = c + 4 < count ? c + 4 : count;
        for (int q = c + 1; q < end; q++)
            if (e[q].key < best_key) { best = q; best_key = e[q].key; }
        if (best_key >= last.key) break;
        e[i] = e[best];
        h->at[e[i].node] = i + 1;
        i = best;
    }
    e[i] = last;
    h->at[last.node] = i + 1;
    return 1;
}

enum { FAR = 0, TRIAL = 1, ALIVE = 2, SEED = 3 };

typedef struct {
    const or_griddef *g;
    const double *hlon, *hlat, *s;
    double hr;
    double *time;
    unsigned char *status;
} march_t;

/* SphericalEikonal.cs:167-222 March.SolveNode / Axis */
static double march_solve(const march_t *m, int idx, int i, int j, int k)
{
    double s = m->s[idx];
    if (!(s > 0) || !isfinite(s)) return INFINITY;
    int nx = m->g->nx, ny = m->g->ny, nz = m->g->nz, nxy = nx * ny;
    double te[3], he[3], tn[3];
    int used = 0;
    const int pos[3] = {i, j, k}, n[3] = {nx, ny, nz}, stride[3] = {1, nx, nxy};
    const double h[3] = {m->hlon[k * ny + j], m->hlat[k], m->hr};
    for (int a = 0; a < 3; a++) {
        int st = stride[a];
        double minus = pos[a] > 0 && m->status[idx - st] == ALIVE ? m->time[idx - st] : INFINITY;
        double plus = pos[a] + 1 < n[a] && m->status[idx + st] == ALIVE ? m->time[idx + st] : INFINITY;
        double first;
        int sec;
        if (minus <= plus) { first = minus; sec = pos[a] > 1 ? idx - 2 * st : -1; }
        else { first = plus; sec = pos[a] + 2 < n[a] ? idx + 2 * st : -1; }
        if (!isfinite(first)) continue;
        double t2 = sec >= 0 && m->status[sec] == ALIVE ? m->time[sec] : INFINITY;
        if (isfinite(t2) && t2 <= first) { te[used] = (4 * first - t2) / 3; he[used] = 2 * h[a] / 3; }
        else { te[used] = first; he[used] = h[a]; }
        tn[used] = first;
        used++;
    }
    if (used == 0) return INFINITY;
    sort3(te, he, tn, used);
    return solve_terms(te, he, tn, used, s);
}

static int fmm_with_metric(const or_griddef *g, const double *hlon, const double *hlat, double hr,
                           const double *slowness, int nseeds, const int *seed_idx,
                           const double *seed_time, double *time)
{
    int nx = g->nx, ny = g->ny, nz = g->nz, count = nx * ny * nz, nxy = nx * ny;
    if (nseeds == 0) return -1;
    for (int q = 0; q < count; q++) time[q] = INFINITY;
    unsigned char *status = calloc(count, 1);
    front_heap h;
    h.cap = count / 64 > 1024 ? count / 64 : 1024;
    h.e = malloc(sizeof(heap_entry) * h.cap);
    h.at = calloc(count, sizeof(int));
    h.count = 0;
    for (int q = 0; q < nseeds; q++) {
        int idx = seed_idx[q];
        if ((unsigned)idx >= (unsigned)count || !(seed_time[q] < time[idx])) continue;
        time[idx] = seed_time[q];
        status[idx] = SEED;
        heap_set(&h, idx, seed_time[q]);
    }
    march_t m = {g, hlon, hlat, slowness, hr, time, status};
    int idx;
    while (heap_pop(&h, &idx)) {
        status[idx] = ALIVE;
        int k = idx / nxy, rest = idx - k * nxy, j = rest / nx, i = rest - j * nx;
        const int nb[6] = {idx - 1, idx + 1, idx - nx, idx + nx, idx - nxy, idx + nxy};
        const int ok[6] = {i > 0, i + 1 < nx, j > 0, j + 1 < ny, k > 0, k + 1 < nz};
        const int ni[6] = {i - 1, i + 1, i, i, i, i}, nj[6] = {j, j, j - 1, j + 1, j, j},
                  nk[6] = {k, k, k, k, k - 1, k + 1};
        for (int q = 0; q < 6; q++) {
            if (!ok[q]) continue;
            int n = nb[q];
            if (status[n] >= ALIVE) continue;
            double t = march_solve(&m, n, ni[q], nj[q], nk[q]);
            if (!(t < time[n])) continue;
            time[n] = t;
            status[n] = TRIAL;
            heap_set(&h, n, t);
        }
    }
    free(status); free(h.e); free(h.at);
    return 0;
}

/* SphericalEikonal.cs:102-152 Solve */
int or_fmm(const or_griddef *g, const double *slowness, int nseeds, const int *seed_idx,
           const double *seed_time, double *out)
{
    double *hlon = malloc(sizeof(double) * g->ny * g->nz), *hlat = malloc(sizeof(double) * g->nz), hr;
    or_metric(g, hlon, hlat, &hr);
    int rc = fmm_with_metric(g, hlon, hlat, hr, slowness, nseeds, seed_idx, seed_time, out);
    free(hlon); free(hlat);
    return rc;
}

/* One fast-marching table per core, the CPU path of TravelTimeTable.cs:332-354. Field f uses
 * slowness[f] and seeds [seed_offsets[f], seed_offsets[f+1]); out is field-major nfields*N. */
int or_fmm_batch(const or_griddef *g, int nfields, const double *const *slowness,
                 const int *seed_offsets, const int *seed_idx, const double *seed_time,
                 double *out, int threads)
{
    long count = (long)g->nx * g->ny * g->nz;
    double *hlon = malloc(sizeof(double) * g->ny * g->nz), *hlat = malloc(sizeof(double) * g->nz), hr;
    or_metric(g, hlon, hlat, &hr);
    int rc = 0;
#ifdef _OPENMP
    if (threads <= 0) threads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads) reduction(| : rc)
#endif
    for (int f = 0; f < nfields; f++) {
        int a = seed_offsets[f], b = seed_offsets[f + 1];
        rc |= fmm_with_metric(g, hlon, hlat, hr, slowness[f], b - a, seed_idx + a, seed_time + a,
                              out + (long)f * count);
    }
    free(hlon); free(hlat);
    (void)threads;
    return rc;
}
