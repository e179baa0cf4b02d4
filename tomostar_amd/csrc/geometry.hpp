// Spherical geometry of the reference grid, usable on host and device.
//
// GeoMath (Geo/GeoMath.cs:40-151) and SphericalGrid (Geo/SphericalGrid.cs:84-235): a sphere of radius
// 6371 km, geographic latitude used as geocentric, nodes ON the bounds, index (k*Ny + j)*Nx + i.
// Every formula keeps the reference's operation order so positions round the same way, and every
// sine, cosine, arcsine and arctangent is the correctly rounded one (crmath.hpp), so positions are the
// same doubles on the device as on any host that rounds its trigonometry correctly.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "crmath.hpp"
#include "tomostar_hip.h"

namespace ts {

constexpr double kEarthRadiusKm = 6371.0;         // GeoMath.cs:43
constexpr double kDeg2Rad = 3.14159265358979323846 / 180.0;
constexpr double kRad2Deg = 180.0 / 3.14159265358979323846;

struct Vec3 {
    double x, y, z;
};

__host__ __device__ inline Vec3 vsub(Vec3 a, Vec3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__host__ __device__ inline Vec3 vadd(Vec3 a, Vec3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__host__ __device__ inline Vec3 vscale(Vec3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
__host__ __device__ inline double vlen(Vec3 a) { return sqrt(a.x * a.x + a.y * a.y + a.z * a.z); }
__host__ __device__ inline double dclamp(double v, double lo, double hi) { return v < lo ? lo : (v > hi ? hi : v); }

// GeoMath.ToCartesian (GeoMath.cs:49-56)
__host__ __device__ inline Vec3 to_cartesian(double lon_deg, double lat_deg, double depth_km)
{
    double r = kEarthRadiusKm - depth_km;
    double lat = lat_deg * kDeg2Rad;
    double lon = lon_deg * kDeg2Rad;
    double sl, cl, so, co;
    cr::sincos(lat, sl, cl);
    cr::sincos(lon, so, co);
    return {r * cl * co, r * cl * so, r * sl};
}

struct GeoPoint {
    double lon, lat, depth;
};

// GeoMath.FromCartesian (GeoMath.cs:61-68)
__host__ __device__ inline GeoPoint from_cartesian(Vec3 v)
{
    double r = vlen(v);
    if (r <= 0) return {0, 0, kEarthRadiusKm};
    double lat = cr::asin(dclamp(v.z / r, -1, 1)) * kRad2Deg;
    double lon = cr::atan2(v.y, v.x) * kRad2Deg;
    return {lon, lat, kEarthRadiusKm - r};
}

// GeoMath.UnwrapLon (GeoMath.cs:146-151)
__host__ __device__ inline double unwrap_lon(double lon, double reference)
{
    if (lon < reference - 180) lon += 360 * ceil((reference - 180 - lon) / 360);
    else if (lon > reference + 180) lon -= 360 * ceil((lon - reference - 180) / 360);
    return lon;
}

// The derived quantities of a GridDefinition / SphericalGrid (SphericalGrid.cs:28-31, :96-116).
struct Grid {
    int nx, ny, nz;
    double min_lon, min_lat, min_depth;
    double dlon, dlat, ddepth;          // degrees, degrees, km per step
    double dlon_rad, dlat_rad;
    double centre_lon;
    double lon_last, lat_last, depth_last; // LonDeg[^1], LatDeg[^1], DepthKm[^1]

    __host__ __device__ int count() const { return nx * ny * nz; }
    __host__ __device__ double lon(int i) const { return min_lon + i * dlon; }
    __host__ __device__ double lat(int j) const { return min_lat + j * dlat; }
    __host__ __device__ double depth(int k) const { return min_depth + k * ddepth; }
    __host__ __device__ int index(int i, int j, int k) const { return (k * ny + j) * nx + i; }
};

inline Grid make_grid(const ts_grid &g)
{
    Grid r;
    r.nx = g.nx; r.ny = g.ny; r.nz = g.nz;
    r.min_lon = g.min_lon; r.min_lat = g.min_lat; r.min_depth = g.min_depth;
    r.dlon = (g.max_lon - g.min_lon) / (g.nx - 1 > 1 ? g.nx - 1 : 1);
    r.dlat = (g.max_lat - g.min_lat) / (g.ny - 1 > 1 ? g.ny - 1 : 1);
    r.ddepth = (g.max_depth - g.min_depth) / (g.nz - 1 > 1 ? g.nz - 1 : 1);
    r.dlon_rad = r.dlon * kDeg2Rad;
    r.dlat_rad = r.dlat * kDeg2Rad;
    r.centre_lon = 0.5 * (g.min_lon + g.max_lon);
    r.lon_last = r.lon(g.nx - 1);
    r.lat_last = r.lat(g.ny - 1);
    r.depth_last = r.depth(g.nz - 1);
    return r;
}

inline bool grid_valid(const ts_grid &g)
{
    if (g.nx < 2 || g.ny < 2 || g.nz < 2) return false;
    if ((int64_t)g.nx * g.ny * g.nz > INT32_MAX) return false;
    if (g.nx > 2047 || g.ny > 2047 || g.nz > 1023) return false;
    if (!std::isfinite(g.min_lon) || !std::isfinite(g.max_lon) || !std::isfinite(g.min_lat) ||
        !std::isfinite(g.max_lat) || !std::isfinite(g.min_depth) || !std::isfinite(g.max_depth))
        return false;
    if (!(g.max_lon > g.min_lon) || !(g.max_lat > g.min_lat) || !(g.max_depth > g.min_depth)) return false;
    if (g.max_lon - g.min_lon >= 360 || g.min_lat < -89.9 || g.max_lat > 89.9) return false;
    return g.max_depth < kEarthRadiusKm;
}

// SphericalGrid.Fractional (SphericalGrid.cs:179-184)
__host__ __device__ inline void fractional(const Grid &g, double lon, double lat, double depth, double &fx,
                                           double &fy, double &fz)
{
    if (lon < g.centre_lon - 180 || lon > g.centre_lon + 180) lon = unwrap_lon(lon, g.centre_lon);
    fx = (lon - g.min_lon) / g.dlon;
    fy = (lat - g.min_lat) / g.dlat;
    fz = (depth - g.min_depth) / g.ddepth;
}

// SphericalGrid.TrilinearWeightsFractional (SphericalGrid.cs:205-224): corner (i0,j0,k0) and the
// three fractions; weight of corner (di,dj,dk) is (di?tx:1-tx)*(dj?ty:1-ty)*(dk?tz:1-tz).
__host__ __device__ inline void trilinear_base(const Grid &g, double fx, double fy, double fz, int &i0, int &j0,
                                               int &k0, double &tx, double &ty, double &tz)
{
    fx = dclamp(fx, 0, g.nx - 1);
    fy = dclamp(fy, 0, g.ny - 1);
    fz = dclamp(fz, 0, g.nz - 1);
    i0 = (int)fx; if (i0 > g.nx - 2) i0 = g.nx - 2;
    j0 = (int)fy; if (j0 > g.ny - 2) j0 = g.ny - 2;
    k0 = (int)fz; if (k0 > g.nz - 2) k0 = g.nz - 2;
    tx = fx - i0; ty = fy - j0; tz = fz - k0;
}

__host__ __device__ inline void trilinear_weights(const Grid &g, double fx, double fy, double fz, int nodes[8],
                                                  double w[8])
{
    int i0, j0, k0;
    double tx, ty, tz;
    trilinear_base(g, fx, fy, fz, i0, j0, k0, tx, ty, tz);
    int n = 0;
    for (int dk = 0; dk < 2; dk++)
        for (int dj = 0; dj < 2; dj++)
            for (int di = 0; di < 2; di++) {
                nodes[n] = g.index(i0 + di, j0 + dj, k0 + dk);
                w[n] = (di == 0 ? 1 - tx : tx) * (dj == 0 ? 1 - ty : ty) * (dk == 0 ? 1 - tz : tz);
                n++;
            }
}

// SphericalGrid.MinSpacingKm (SphericalGrid.cs:159-164)
inline double min_spacing_km(const Grid &g)
{
    double rmin = kEarthRadiusKm - g.depth_last;
    double cosmin = INFINITY;
    for (int j = 0; j < g.ny; j++) {
        double c = cr::cos(g.lat(j) * kDeg2Rad);
        if (c < cosmin) cosmin = c;
    }
    double a = rmin * g.dlat_rad, b = rmin * cosmin * g.dlon_rad;
    double m = a < b ? a : b;
    return g.ddepth < m ? g.ddepth : m;
}

}  // namespace ts
