/*
 * Oracle (test infrastructure only): travel-time tables, the curved-ray back-trace and the ray kernel.
 * Restates Forward/TravelTimeTable.cs (Time/TimeFractional :116-136, Gradient :143-157),
 * Forward/RayTracing.cs (Backtrack :73-111, Kernel :126-159) and the per-arrival row of
 * Tomography/RayKernels.cs:62-105 (step rule, ascending node order, float32 lengths).
 * Tables are float32, as the reference stores and memory-maps them (TravelTimeTable.cs:73).
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

static double clampd(double v, double lo, double hi) { return v < lo ? lo : (v > hi ? hi : v); }
static double gd_dlon(const or_griddef *g) { return (g->max_lon - g->min_lon) / (g->nx - 1 > 1 ? g->nx - 1 : 1); }
static double gd_dlat(const or_griddef *g) { return (g->max_lat - g->min_lat) / (g->ny - 1 > 1 ? g->ny - 1 : 1); }
static double gd_ddep(const or_griddef *g) { return (g->max_depth - g->min_depth) / (g->nz - 1 > 1 ? g->nz - 1 : 1); }

/* TravelTimeTable.cs:122-136 */
double or_table_time_fractional(const or_griddef *g, const float *table, double fx, double fy, double fz)
{
    int nodes[8];
    double w[8];
    or_trilinear_weights_fractional(g, fx, fy, fz, nodes, w);
    double t = 0;
    for (int n = 0; n < 8; n++) {
        if (w[n] == 0) continue;
        float v = table[nodes[n]];
        if (!isfinite(v)) return INFINITY;
        t += v * w[n];
    }
    return t;
}

/* TravelTimeTable.cs:116-120 */
double or_table_time(const or_griddef *g, const float *table, double lon, double lat, double depth)
{
    double f[3];
    or_fractional(g, lon, lat, depth, f);
    return or_table_time_fractional(g, table, f[0], f[1], f[2]);
}

/* TravelTimeTable.cs:143-157: central differences of the interpolant at +-0.5 node, per km. */
void or_table_gradient(const or_griddef *g, const float *table, double lon, double lat, double depth,
                       double out[3])
{
    double f[3];
    or_fractional(g, lon, lat, depth, f);
    const double d = 0.5;
    double tx = (or_table_time_fractional(g, table, f[0] + d, f[1], f[2]) -
                 or_table_time_fractional(g, table, f[0] - d, f[1], f[2])) / (2 * d);
    double ty = (or_table_time_fractional(g, table, f[0], f[1] + d, f[2]) -
                 or_table_time_fractional(g, table, f[0], f[1] - d, f[2])) / (2 * d);
    double tz = (or_table_time_fractional(g, table, f[0], f[1], f[2] + d) -
                 or_table_time_fractional(g, table, f[0], f[1], f[2] - d)) / (2 * d);
    double r = OR_R - depth;
    double hx = r * or_cos(lat * OR_D2R) * (gd_dlon(g) * OR_D2R);
    double hy = r * (gd_dlat(g) * OR_D2R);
    double hz = gd_ddep(g);
    out[0] = tx / hx;
    out[1] = ty / hy;
    out[2] = tz / hz;
}

static double vdist(const double a[3], const double b[3])
{
    double x = a[0] - b[0], y = a[1] - b[1], z = a[2] - b[2];
    return sqrt(x * x + y * y + z * z);
}

/* RayTracing.cs:73-111 Backtrack. Writes the path (lon, lat, depth), returns its point count, -1
 * when the descent stalls (null in the reference), or -2 when the path exceeds cap. */
int or_backtrack(const or_griddef *g, const float *table, double st_lon, double st_lat, double st_depth,
                 double lon, double lat, double depth, double step_km, int cap,
                 double *plon, double *plat, double *pdep)
{
    double d0 = g->min_depth, d1 = g->min_depth + (g->nz - 1) * gd_ddep(g);
    double lon0 = g->min_lon, lon1 = g->min_lon + (g->nx - 1) * gd_dlon(g);
    double lat0 = g->min_lat, lat1 = g->min_lat + (g->ny - 1) * gd_dlat(g);
    double centre = 0.5 * (g->min_lon + g->max_lon);
    double station[3], p[3];
    or_to_cartesian(st_lon, st_lat, clampd(st_depth, d0, d1), station);
    or_to_cartesian(lon, lat, depth, p);
    int n = 0;
#define PUSH(a, b, c)                                   \
    do {                                                \
        if (n >= cap) return -2;                        \
        plon[n] = (a); plat[n] = (b); pdep[n] = (c); n++; \
    } while (0)
    PUSH(lon, lat, depth);
    double t0 = or_table_time(g, table, lon, lat, depth);
    if (!isfinite(t0)) return -1;
    int max_steps = 20 * (int)ceil(vdist(p, station) / step_km + 10);
    double last_t = t0;
    int stalls = 0;
    for (int step = 0; step < max_steps; step++) {
        if (vdist(p, station) <= 1.5 * step_km) break;
        double gp[3], gr[3];
        or_from_cartesian(p, gp);
        or_table_gradient(g, table, gp[0], gp[1], gp[2], gr);
        if (!isfinite(gr[0]) || !isfinite(gr[1]) || !isfinite(gr[2])) return -1;
        double e[3], nn[3], up[3];
        or_local_frame(gp[0], gp[1], e, nn, up);
        double grad[3];
        for (int c = 0; c < 3; c++) grad[c] = e[c] * gr[0] + nn[c] * gr[1] + up[c] * (-gr[2]);
        double norm = sqrt(grad[0] * grad[0] + grad[1] * grad[1] + grad[2] * grad[2]);
        if (norm < 1e-12) return -1;
        double s = step_km / norm;
        for (int c = 0; c < 3; c++) p[c] = p[c] - grad[c] * s;
        double q[3];
        or_from_cartesian(p, q);
        q[0] = clampd(or_unwrap_lon(q[0], centre), lon0, lon1);
        q[1] = clampd(q[1], lat0, lat1);
        q[2] = clampd(q[2], d0, d1);
        or_to_cartesian(q[0], q[1], q[2], p);
        PUSH(q[0], q[1], q[2]);
        double t = or_table_time(g, table, q[0], q[1], q[2]);
        if (!(t < last_t)) {
            if (++stalls > 20) return -1;
        } else stalls = 0;
        last_t = fmin(last_t, t);
    }
    PUSH(st_lon, st_lat, clampd(st_depth, d0, d1));
#undef PUSH
    return n;
}

typedef struct { int node; int order; double v; } contrib_t;

static int contrib_cmp(const void *a, const void *b)
{
    const contrib_t *x = a, *y = b;
    if (x->node != y->node) return x->node < y->node ? -1 : 1;
    return x->order < y->order ? -1 : (x->order > y->order);
}

/* VelocityModel1D.Evaluate (Model/VelocityModel1D.cs:31-45): piecewise linear in depth; at or above the
 * first node its value; a repeated depth is a step (a point on it takes the upper value); below the
 * last node its value. phase 0 Vp, 1 Vs. */
double or_background_velocity(const or_background *bg, int phase, double depth)
{
    const double *v = phase ? bg->vs : bg->vp;
    if (depth <= bg->depth_km[0]) return v[0];
    for (int i = 1; i < bg->count; i++) {
        const double a = bg->depth_km[i - 1], b = bg->depth_km[i];
        if (depth > b) continue;
        if (b - a <= 1e-12) continue;
        const double t = (depth - a) / (b - a);
        return v[i - 1] + t * (v[i] - v[i - 1]);
    }
    return v[bg->count - 1];
}

/* RayTracing.cs:126-159 Kernel: kernel[node] = sum of dl * w over the pieces, each node accumulated in
 * the order the pieces reach it (the Dictionary's old + v), zero weights skipped. With a background
 * (bg != NULL, the extended forward domain of RayKernels.cs:62-64), a piece whose midpoint lies outside
 * the inversion grid (SphericalGrid.Contains, tolerance 1e-6 in node units, SphericalGrid.cs:186-192)
 * adds dl / v_bg(phase, depth) to *outside instead (:141-145). Output sorted by node (RayKernels.cs:96).
 * Returns the node count, or -2 beyond cap. */
int or_ray_kernel_bg(const or_griddef *inv, int npts, const double *plon, const double *plat, const double *pdep,
                     const or_background *bg, int phase, int cap, int *nodes_out, double *len_out, double *outside)
{
    double outside_time = 0;
    double h = 0.5 * or_min_spacing_km(inv);
    int ncap = 1024, nc = 0;
    contrib_t *c = malloc(sizeof(contrib_t) * ncap);
    for (int i = 1; i < npts; i++) {
        double a[3], b[3];
        or_to_cartesian(plon[i - 1], plat[i - 1], pdep[i - 1], a);
        or_to_cartesian(plon[i], plat[i], pdep[i], b);
        double d[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
        double len = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
        if (len <= 0) continue;
        int pieces = (int)ceil(len / h);
        if (pieces < 1) pieces = 1;
        double dl = len / pieces;
        for (int q = 0; q < pieces; q++) {
            double f = (q + 0.5) / pieces;
            double m[3] = {a[0] + d[0] * f, a[1] + d[1] * f, a[2] + d[2] * f};
            double mid[3], fr[3];
            or_from_cartesian(m, mid);
            or_fractional(inv, mid[0], mid[1], mid[2], fr);
            if (bg) {
                const double tol = 1e-6;
                const int in = fr[0] >= -tol && fr[0] <= inv->nx - 1 + tol && fr[1] >= -tol &&
                               fr[1] <= inv->ny - 1 + tol && fr[2] >= -tol && fr[2] <= inv->nz - 1 + tol;
                if (!in) {
                    outside_time += dl * (1.0 / or_background_velocity(bg, phase, mid[2]));
                    continue;
                }
            }
            int nodes[8];
            double w[8];
            or_trilinear_weights_fractional(inv, fr[0], fr[1], fr[2], nodes, w);
            for (int k = 0; k < 8; k++) {
                if (w[k] == 0) continue;
                if (nc == ncap) { ncap *= 2; c = realloc(c, sizeof(contrib_t) * ncap); }
                c[nc].node = nodes[k];
                c[nc].order = nc;
                c[nc].v = dl * w[k] * 1.0;
                nc++;
            }
        }
    }
    qsort(c, nc, sizeof(contrib_t), contrib_cmp);
    int count = 0;
    for (int q = 0; q < nc;) {
        int node = c[q].node;
        double s = c[q].v;
        int r = q + 1;
        while (r < nc && c[r].node == node) { s = s + c[r].v; r++; }
        if (count >= cap) { free(c); return -2; }
        nodes_out[count] = node;
        len_out[count] = s;
        count++;
        q = r;
    }
    free(c);
    if (outside) *outside = outside_time;
    return count;
}

int or_ray_kernel(const or_griddef *inv, int npts, const double *plon, const double *plat,
                  const double *pdep, int cap, int *nodes_out, double *len_out)
{
    return or_ray_kernel_bg(inv, npts, plon, plat, pdep, NULL, 0, cap, nodes_out, len_out, NULL);
}

/* One arrival of RayKernels.Compute (RayKernels.cs:77-105): back-trace on the station's table, the
 * time and gradient at the event, the kernel on the inversion grid with float32 lengths.
 * tcalc_grad = {TCalc, Ge, Gn, Gz}. Returns the node count, -1 when the arrival yields no row. */
static int ray_row_bg(const or_griddef *fwd, const float *table, double st_lon, double st_lat, double st_depth,
                      const or_griddef *inv, double ev_lon, double ev_lat, double ev_depth, double step_km,
                      const or_background *bg, int phase, int cap, int *nodes_out, float *len_out,
                      double tcalc_grad[4], int *npath_out, double *outside)
{
    int pcap = 20 * ((int)ceil(2 * 6371.0 / step_km) + 40);
    double *pl = malloc(sizeof(double) * pcap * 3);
    int np = or_backtrack(fwd, table, st_lon, st_lat, st_depth, ev_lon, ev_lat, ev_depth, step_km, pcap,
                          pl, pl + pcap, pl + 2 * pcap);
    if (npath_out) *npath_out = np;
    if (np < 0) { free(pl); return -1; }
    double tcalc = or_table_time(fwd, table, ev_lon, ev_lat, ev_depth);
    double gr[3];
    or_table_gradient(fwd, table, ev_lon, ev_lat, ev_depth, gr);
    if (!isfinite(tcalc)) { free(pl); return -1; }
    double *len = malloc(sizeof(double) * (cap > 0 ? cap : 1));
    int n = or_ray_kernel_bg(inv, np, pl, pl + pcap, pl + 2 * pcap, bg, phase, cap, nodes_out, len, outside);
    for (int q = 0; q < n; q++) len_out[q] = (float)len[q];
    tcalc_grad[0] = tcalc;
    tcalc_grad[1] = gr[0];
    tcalc_grad[2] = gr[1];
    tcalc_grad[3] = gr[2];
    free(len);
    free(pl);
    return n;
}

int or_ray_row(const or_griddef *fwd, const float *table, double st_lon, double st_lat, double st_depth,
               const or_griddef *inv, double ev_lon, double ev_lat, double ev_depth, double step_km,
               int cap, int *nodes_out, float *len_out, double tcalc_grad[4], int *npath_out)
{
    return ray_row_bg(fwd, table, st_lon, st_lat, st_depth, inv, ev_lon, ev_lat, ev_depth, step_km, NULL, 0, cap,
                      nodes_out, len_out, tcalc_grad, npath_out, NULL);
}

/* Many arrivals in parallel over the host cores (Parallel.ForEach, RayKernels.cs:67): ray r uses
 * table ray_table[r] (stations[3*t..] its station) and event events[3*r..]; its row goes to
 * nodes/lens[r*cap_per_ray ...], counts[r] = node count or -1. */
int or_ray_rows_batch_bg(const or_griddef *fwd, const float *const *tables, const double *stations,
                         const or_griddef *inv, int nrays, const int *ray_table, const double *events,
                         double step_km, const or_background *bg, const int *phase_of_ray, int cap_per_ray,
                         int threads, int *counts, int *nodes, float *lens, double *tcalc_grad, double *outside)
{
#ifdef _OPENMP
    if (threads <= 0) threads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 16) num_threads(threads)
#endif
    for (int r = 0; r < nrays; r++) {
        int t = ray_table[r];
        double out_t = 0;
        counts[r] = ray_row_bg(fwd, tables[t], stations[3 * t], stations[3 * t + 1], stations[3 * t + 2], inv,
                               events[3 * r], events[3 * r + 1], events[3 * r + 2], step_km, bg,
                               phase_of_ray ? phase_of_ray[r] : 0, cap_per_ray, nodes + (size_t)r * cap_per_ray,
                               lens + (size_t)r * cap_per_ray, tcalc_grad + 4 * (size_t)r, NULL, &out_t);
        if (outside) outside[r] = out_t;
    }
    (void)threads;
    return 0;
}

int or_ray_rows_batch(const or_griddef *fwd, const float *const *tables, const double *stations,
                      const or_griddef *inv, int nrays, const int *ray_table, const double *events,
                      double step_km, int cap_per_ray, int threads,
                      int *counts, int *nodes, float *lens, double *tcalc_grad)
{
    return or_ray_rows_batch_bg(fwd, tables, stations, inv, nrays, ray_table, events, step_km, NULL, NULL, cap_per_ray,
                                threads, counts, nodes, lens, tcalc_grad, NULL);
}
