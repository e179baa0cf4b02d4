/*
 * Oracle (test infrastructure only): spherical geometry and the regular grid.
 * Restates Geo/GeoMath.cs and Geo/SphericalGrid.cs of the reference in plain C.
 */
#include "oracle.h"

#include <math.h>
#include <quadmath.h>

/* Correctly rounded trigonometry: the 113-bit quadmath value (error below 2 ulp of quad) rounded to
 * the nearest double. It misses the correct rounding only when the exact value lies within 2^-110 of a
 * midpoint between doubles. */
double or_sin(double x) { return (double)sinq((__float128)x); }
double or_cos(double x) { return (double)cosq((__float128)x); }
double or_asin(double x) { return (double)asinq((__float128)x); }
double or_atan2(double y, double x) { return (double)atan2q((__float128)y, (__float128)x); }

/* Arrays of either: out[q] = f(a[q]) (func 0 sin, 1 cos, 2 asin) or atan2(a[q], b[q]) (func 3), the
 * correctly rounded values when correct != 0, else the platform's. */
void or_trig_array(int func, int correct, int n, const double *a, const double *b, double *out)
{
#pragma omp parallel for schedule(static)
    for (int q = 0; q < n; q++) {
        const double x = a[q], y = b ? b[q] : 0.0;
        if (!correct) out[q] = or_libm(func, x, y);
        else if (func == 0) out[q] = or_sin(x);
        else if (func == 1) out[q] = or_cos(x);
        else if (func == 2) out[q] = or_asin(x);
        else out[q] = or_atan2(x, y);
    }
}

/* func 0 sin(a), 1 cos(a), 2 asin(a), 3 atan2(a, b) */
double or_libm(int func, double a, double b)
{
    switch (func) {
    case 0: return sin(a);
    case 1: return cos(a);
    case 2: return asin(a);
    default: return atan2(a, b);
    }
}

#define OR_R 6371.0                 /* GeoMath.cs:43 EarthRadiusKm */
#define OR_D2R (M_PI / 180.0)       /* GeoMath.cs:45 */
#define OR_R2D (180.0 / M_PI)       /* GeoMath.cs:46 */

static double clampd(double v, double lo, double hi) { return v < lo ? lo : (v > hi ? hi : v); }

static double dlon(const or_griddef *g) { return (g->max_lon - g->min_lon) / (g->nx - 1 > 1 ? g->nx - 1 : 1); }
static double dlat(const or_griddef *g) { return (g->max_lat - g->min_lat) / (g->ny - 1 > 1 ? g->ny - 1 : 1); }
static double ddep(const or_griddef *g) { return (g->max_depth - g->min_depth) / (g->nz - 1 > 1 ? g->nz - 1 : 1); }

/* GeoMath.cs:49-56 */
void or_to_cartesian(double lon, double lat, double depth, double out[3])
{
    double r = OR_R - depth;
    double la = lat * OR_D2R, lo = lon * OR_D2R;
    double cl = or_cos(la);
    out[0] = r * cl * or_cos(lo);
    out[1] = r * cl * or_sin(lo);
    out[2] = r * or_sin(la);
}

/* GeoMath.cs:61-68 (out = lon, lat, depth) */
void or_from_cartesian(const double v[3], double out[3])
{
    double r = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    if (r <= 0) { out[0] = 0; out[1] = 0; out[2] = OR_R; return; }
    out[1] = or_asin(clampd(v[2] / r, -1, 1)) * OR_R2D;
    out[0] = or_atan2(v[1], v[0]) * OR_R2D;
    out[2] = OR_R - r;
}

/* GeoMath.cs:71-80 */
void or_local_frame(double lon, double lat, double east[3], double north[3], double up[3])
{
    double la = lat * OR_D2R, lo = lon * OR_D2R;
    double sl = or_sin(la), cl = or_cos(la), so = or_sin(lo), co = or_cos(lo);
    east[0] = -so; east[1] = co; east[2] = 0;
    north[0] = -sl * co; north[1] = -sl * so; north[2] = cl;
    up[0] = cl * co; up[1] = cl * so; up[2] = sl;
}

/* GeoMath.cs:146-151 */
double or_unwrap_lon(double lon, double reference)
{
    if (lon < reference - 180) lon += 360 * ceil((reference - 180 - lon) / 360);
    else if (lon > reference + 180) lon -= 360 * ceil((lon - reference - 180) / 360);
    return lon;
}

/* GeoMath.cs:83-95 (haversine) */
double or_surface_distance_km(double lon1, double lat1, double lon2, double lat2)
{
    double p1 = lat1 * OR_D2R, p2 = lat2 * OR_D2R;
    double dp = p2 - p1, dl = (lon2 - lon1) * OR_D2R;
    double a = or_sin(dp / 2) * or_sin(dp / 2) + or_cos(p1) * or_cos(p2) * or_sin(dl / 2) * or_sin(dl / 2);
    double s = sqrt(a);
    return 2 * or_asin(s < 1 ? s : 1) * OR_R;
}

/* SphericalGrid.cs:59-74 GridDefinition.FromSpacing; Math.Round is round-half-even (rint). */
void or_grid_from_spacing(double min_lon, double max_lon, double min_lat, double max_lat,
                          double min_depth, double max_depth, double h_km, double v_km, or_griddef *out)
{
    double mid = 0.5 * (min_lat + max_lat) * OR_D2R;
    double width = (max_lon - min_lon) * OR_D2R * OR_R * or_cos(mid);
    double height = (max_lat - min_lat) * OR_D2R * OR_R;
    int nx = (int)rint(width / h_km) + 1, ny = (int)rint(height / h_km) + 1;
    int nz = (int)rint((max_depth - min_depth) / v_km) + 1;
    out->min_lon = min_lon; out->max_lon = max_lon; out->min_lat = min_lat; out->max_lat = max_lat;
    out->min_depth = min_depth; out->max_depth = max_depth;
    out->nx = nx > 2 ? nx : 2; out->ny = ny > 2 ? ny : 2; out->nz = nz > 2 ? nz : 2;
}

/* SphericalGrid.cs:96-111: node coordinates (lon[nx], lat[ny], depth[nz]) */
void or_node_coords(const or_griddef *g, double *lon, double *lat, double *depth)
{
    double a = dlon(g), b = dlat(g), c = ddep(g);
    if (lon) for (int i = 0; i < g->nx; i++) lon[i] = g->min_lon + i * a;
    if (lat) for (int j = 0; j < g->ny; j++) lat[j] = g->min_lat + j * b;
    if (depth) for (int k = 0; k < g->nz; k++) depth[k] = g->min_depth + k * c;
}

/* SphericalGrid.cs:179-184 */
void or_fractional(const or_griddef *g, double lon, double lat, double depth, double out[3])
{
    double centre = 0.5 * (g->min_lon + g->max_lon);
    if (lon < centre - 180 || lon > centre + 180) lon = or_unwrap_lon(lon, centre);
    out[0] = (lon - g->min_lon) / dlon(g);
    out[1] = (lat - g->min_lat) / dlat(g);
    out[2] = (depth - g->min_depth) / ddep(g);
}

/* SphericalGrid.cs:205-224 */
void or_trilinear_weights_fractional(const or_griddef *g, double fx, double fy, double fz,
                                     int nodes[8], double weights[8])
{
    fx = clampd(fx, 0, g->nx - 1);
    fy = clampd(fy, 0, g->ny - 1);
    fz = clampd(fz, 0, g->nz - 1);
    int i0 = (int)fx, j0 = (int)fy, k0 = (int)fz;
    if (i0 > g->nx - 2) i0 = g->nx - 2;
    if (j0 > g->ny - 2) j0 = g->ny - 2;
    if (k0 > g->nz - 2) k0 = g->nz - 2;
    double tx = fx - i0, ty = fy - j0, tz = fz - k0;
    int n = 0;
    for (int dk = 0; dk < 2; dk++)
        for (int dj = 0; dj < 2; dj++)
            for (int di = 0; di < 2; di++) {
                nodes[n] = ((k0 + dk) * g->ny + (j0 + dj)) * g->nx + (i0 + di);
                weights[n] = (di == 0 ? 1 - tx : tx) * (dj == 0 ? 1 - ty : ty) * (dk == 0 ? 1 - tz : tz);
                n++;
            }
}

/* SphericalGrid.cs:227-235 */
double or_interpolate(const or_griddef *g, const double *field, double lon, double lat, double depth)
{
    double f[3];
    int nodes[8];
    double w[8];
    or_fractional(g, lon, lat, depth, f);
    or_trilinear_weights_fractional(g, f[0], f[1], f[2], nodes, w);
    double sum = 0;
    for (int n = 0; n < 8; n++) sum += field[nodes[n]] * w[n];
    return sum;
}

/* SphericalGrid.cs:159-164 */
double or_min_spacing_km(const or_griddef *g)
{
    double rmin = OR_R - (g->min_depth + (g->nz - 1) * ddep(g));
    double cosmin = INFINITY;
    for (int j = 0; j < g->ny; j++) {
        double c = or_cos((g->min_lat + j * dlat(g)) * OR_D2R);
        if (c < cosmin) cosmin = c;
    }
    double dlon_rad = dlon(g) * OR_D2R, dlat_rad = dlat(g) * OR_D2R;
    double a = rmin * dlat_rad, b = rmin * cosmin * dlon_rad;
    double m = a < b ? a : b;
    double c = ddep(g);
    return c < m ? c : m;
}
