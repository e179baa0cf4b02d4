/*
 * CPU oracle for the TomoSTAR eikonal / ray / LSQR hot path: a plain-C FP64 restatement of the
 * reference's CPU solvers, written from the published algorithm and the reference's behaviour.
 *
 * TEST INFRASTRUCTURE ONLY. Nothing in the shipped library links, loads or calls this code; it is
 * built into oracle/liboracle.so and used by tests/, by __graft_entry__.smoke() as the checker and by
 * bench.py's cpu_baseline leg. Every function names the reference file:line it restates.
 *
 * Conventions (as the reference): grid index (k*Ny + j)*Nx + i, longitude fastest; times in s,
 * distances in km, angles in degrees at the interface; +inf marks an unreached node.
 * Compiled with -ffp-contract=off so products and sums round as the .NET JIT rounds them.
 *
 * Trigonometry: the reference calls .NET's Math.Sin / Cos / Asin / Atan2, which return the platform C
 * library's results, and those differ between platforms in the last bit (glibc 2.35's are not
 * correctly rounded on 0.07-0.6 % of a regional grid's angles). The oracle uses the correctly rounded
 * values, the one choice every platform can share: libquadmath's 113-bit sinq / cosq / asinq / atan2q
 * rounded to double (or_sin, or_cos, or_asin, or_atan2).
 */
#ifndef TOMOSTAR_ORACLE_H
#define TOMOSTAR_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* A regular lon/lat/depth grid: Geo/SphericalGrid.cs:13-31 (GridDefinition) and :84-133. */
typedef struct {
    double min_lon, max_lon, min_lat, max_lat, min_depth, max_depth;
    int nx, ny, nz;
} or_griddef;

/* --- correctly rounded trigonometry (the double nearest to the exact value) -------------------- */
double or_sin(double x);
double or_cos(double x);
double or_asin(double x);
double or_atan2(double y, double x);
/* the platform's (glibc) results, for comparisons only: func 0 sin(a), 1 cos(a), 2 asin(a), 3 atan2(a, b) */
double or_libm(int func, double a, double b);
void or_trig_array(int func, int correct, int n, const double *a, const double *b, double *out);

/* --- geometry (Geo/GeoMath.cs, Geo/SphericalGrid.cs) --------------------------------------- */
void or_to_cartesian(double lon, double lat, double depth, double out[3]);
void or_from_cartesian(const double v[3], double out[3]);
void or_local_frame(double lon, double lat, double east[3], double north[3], double up[3]);
double or_unwrap_lon(double lon, double reference);
double or_surface_distance_km(double lon1, double lat1, double lon2, double lat2);
void or_grid_from_spacing(double min_lon, double max_lon, double min_lat, double max_lat,
                          double min_depth, double max_depth, double h_km, double v_km, or_griddef *out);
void or_fractional(const or_griddef *g, double lon, double lat, double depth, double out[3]);
void or_trilinear_weights_fractional(const or_griddef *g, double fx, double fy, double fz,
                                     int nodes[8], double weights[8]);
double or_interpolate(const or_griddef *g, const double *field, double lon, double lat, double depth);
double or_min_spacing_km(const or_griddef *g);
void or_node_coords(const or_griddef *g, double *lon, double *lat, double *depth);

/* --- eikonal (Forward/SphericalEikonal.cs) -------------------------------------------------- */
void or_metric(const or_griddef *g, double *hlon /* ny*nz */, double *hlat /* nz */, double *hr);
int or_point_source(const or_griddef *g, const double *slowness, double lon, double lat, double depth,
                    double origin_time, double radius_cells, int *idx_out, double *time_out, int cap);
int or_relax(const or_griddef *g, const double *slowness, int nseeds, const int *seed_idx,
             const double *seed_time, int max_rounds, double *out, int rounds_out[2]);
int or_fmm(const or_griddef *g, const double *slowness, int nseeds, const int *seed_idx,
           const double *seed_time, double *out);
int or_fmm_batch(const or_griddef *g, int nfields, const double *const *slowness,
                 const int *seed_offsets, const int *seed_idx, const double *seed_time,
                 double *out, int threads);

/* --- travel-time tables (Forward/TravelTimeTable.cs) --------------------------------------- */
double or_table_time_fractional(const or_griddef *g, const float *table, double fx, double fy, double fz);
double or_table_time(const or_griddef *g, const float *table, double lon, double lat, double depth);
void or_table_gradient(const or_griddef *g, const float *table, double lon, double lat, double depth,
                       double out[3]);

/* --- rays (Forward/RayTracing.cs, Tomography/RayKernels.cs) -------------------------------- */
/* A 1-D background model (Model/VelocityModel1D.cs): node depths (km, non-decreasing), Vp, Vs (km/s). */
typedef struct {
    int count;
    const double *depth_km, *vp, *vs;
} or_background;
double or_background_velocity(const or_background *bg, int phase, double depth);
int or_ray_kernel_bg(const or_griddef *inv, int npts, const double *plon, const double *plat, const double *pdep,
                     const or_background *bg, int phase, int cap, int *nodes_out, double *len_out, double *outside);
int or_ray_rows_batch_bg(const or_griddef *fwd, const float *const *tables, const double *stations,
                         const or_griddef *inv, int nrays, const int *ray_table, const double *events,
                         double step_km, const or_background *bg, const int *phase_of_ray, int cap_per_ray,
                         int threads, int *counts, int *nodes, float *lens, double *tcalc_grad, double *outside);
int or_backtrack(const or_griddef *g, const float *table, double st_lon, double st_lat, double st_depth,
                 double lon, double lat, double depth, double step_km, int cap,
                 double *plon, double *plat, double *pdep);
int or_ray_kernel(const or_griddef *inv, int npts, const double *plon, const double *plat,
                  const double *pdep, int cap, int *nodes_out, double *len_out);
int or_ray_row(const or_griddef *fwd, const float *table, double st_lon, double st_lat, double st_depth,
               const or_griddef *inv, double ev_lon, double ev_lat, double ev_depth, double step_km,
               int cap, int *nodes_out, float *len_out, double tcalc_grad[4], int *npath_out);
int or_ray_rows_batch(const or_griddef *fwd, const float *const *tables, const double *stations /* 3/table */,
                      const or_griddef *inv, int nrays, const int *ray_table, const double *events /* 3/ray */,
                      double step_km, int cap_per_ray, int threads,
                      int *counts, int *nodes, float *lens, double *tcalc_grad /* 4/ray */);

/* --- sparse algebra and LSQR (Numerics/SparseMatrix.cs, SimdVector.cs, Lsqr.cs) ------------ */
void or_csr_transpose(int rows, int cols, const int64_t *rs, const int *ci, const double *va,
                      int64_t *trs, int *tci, double *tva);
void or_spmv(int rows, const int64_t *rs, const int *ci, const double *va, const double *x, double *y,
             int threads);
double or_dot(int n, const double *a, const double *b);
double or_norm(int n, const double *a);
/* LsqrOpenCl's device dot product (64 partials of 256 tree-reduced item sums), and the switch that
 * makes or_lsqr (on the calling thread) use it and its square root for the norms, as LsqrOpenCl.Run does */
double or_dot_device(int n, const double *a, const double *b);
void or_lsqr_dot_mode(int mode);
int or_lsqr(int m, int n, const int64_t *rs, const int *ci, const double *va,
            const int64_t *trs, const int *tci, const double *tva, const double *b, double damp,
            int max_it, double atol, double btol, double conlim, int threads,
            double *x, double stats[4], double *rnorm_hist);

/* --- the linear system (Tomography/TravelTimeTomography.cs:860-1034 BuildSystem, plain node grid) --- */
typedef struct {
    int vpvs, joint_hypocentres, station_corrections, invert_p, invert_s, gradient_smoothing;
    int nevents, nstations, any_p, any_s, smoothing_in_nodes;
    double damping, smoothing, vertical_weight, damping_hypocentre, damping_station, s_factor, absolute_weight;
} or_system_opts;

int or_build_system(const or_griddef *g, const or_system_opts *o, int nrows, const int64_t *row_start,
                    const int *row_nodes, const float *row_len, const double *row_tg, const int *phase,
                    const int *event, const int *station, const double *residual, const double *sigma,
                    const unsigned char *ev_fixed, const unsigned char *ev_outside, const double *sP,
                    const double *sS, const double *sRefP, const double *sRefS, const double *kRef,
                    int64_t cap_rows, int64_t cap_nnz, int64_t *rs, int *ci, double *va, double *rhs,
                    int64_t dims[4]);

/* --- Q tomography (Attenuation/QTomography.cs): t* rows of a plain node grid --- */
typedef struct {
    int station_terms, gradient_smoothing, smoothing_in_nodes, nstations, fixed_scales;
    double damping, smoothing, vertical_weight, damping_station, q0;
} or_qsystem_opts;

void or_q_predict(int nrows, const int64_t *row_start, const int *row_nodes, const float *row_len,
                  const double *slow, const double *q, double q0, const double *outside, double *pred, double *tt);
int or_build_qsystem(const or_griddef *g, const or_qsystem_opts *o, int nrows, const int64_t *row_start,
                     const int *row_nodes, const float *row_len, const double *outside, const int *station,
                     const double *tstar, const double *sigma, const double *slow, const double *q_old,
                     const double *m_old, int64_t cap_rows, int64_t cap_nnz, int64_t *rs, int *ci, double *va,
                     double *rhs, int64_t dims[4], double scales[2]);

#ifdef __cplusplus
}
#endif
#endif
