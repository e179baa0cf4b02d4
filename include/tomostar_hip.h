/*
 * tomostar_hip.h - C ABI of libtomostar_hip.so, the MI355X (gfx950) implementation of TomoSTAR's GPU
 * hot path: the batched 3-D spherical eikonal, the curved-ray back-trace that fills the sensitivity
 * matrix G, and LSQR's A*x / A^T*y products with the Golub-Kahan updates.
 *
 * The ABI is shaped for P/Invoke from TomoSTAR's C# host ([DllImport("tomostar_hip")], blittable
 * arrays, int returns; the pattern of Compute/MpiSession.cs:158-168) and for ctypes.
 *
 * Conventions of every entry point
 *   - status: TS_OK (0); > 0 = the device declines (reason written to `reason`, the caller falls back
 *     to its CPU solver exactly as after EikonalOpenCl.TrySolve or LsqrOpenCl.TrySolve return
 *     false/null); < 0 = invalid argument or cancellation. Nothing throws or aborts across the ABI.
 *   - host pointers are caller-owned and only read or written during the call; the library owns
 *     every device allocation. Opaque handles (ts_context, ts_tableset, ts_matrix) keep models,
 *     tables and G resident in HBM between calls and are released explicitly.
 *   - `cancel` (may be NULL) is polled between device rounds / iterations; a non-zero value ends the
 *     call with TS_CANCELLED (the adapter maps it to OperationCanceledException).
 *   - grids use the reference's node order (k*Ny + j)*Nx + i, longitude fastest; times in s, lengths
 *     in km, +inf for an unreached node. FP64 in and out at the API; the fast modes compute in FP32.
 *   - calls on one handle are serialised by the library (the reference's lock(_gate)); different
 *     handles may be driven from different host threads.
 */
#ifndef TOMOSTAR_HIP_H
#define TOMOSTAR_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TS_OK 0
#define TS_DECLINED 1
#define TS_EINVAL (-1)
#define TS_CANCELLED (-2)
#define TS_EDEVICE (-3)

/* ts_hip_create flags */
#define TS_CREATE_SKIP_SELFTEST 1

/* eikonal modes */
#define TS_EIK_PLAIN_FP64 0 /* the reference scheme: FP64, one launch per diagonal plane, batch-level
                               convergence, host flag read per round (EikonalOpenCl.cs:168-343) */
#define TS_EIK_FAST_FP32 1  /* batched, mixed precision: the first-order phase on FP32 times (offset form),
                               the second-order phase on FP64 times; 64 fields per wavefront, per-field
                               convergence, wavefront-ballot wake flags, one graph launch per round */
#define TS_EIK_FAST_FP64 2  /* the batched schedule with FP64 times: equal to SphericalEikonal.Relax of
                               each field to the last bit */

/* LSQR modes */
#define TS_LSQR_PLAIN_FP64 0 /* the reference's device loop (LsqrOpenCl.cs:272-533): FP64, one right-hand
                                side at a time, host-side rotations, three blocking reads an iteration */
#define TS_LSQR_FAST_FP32 1  /* k right-hand sides share each pass over A and A^T (SpMM); FP32 values and
                                u/v/w, FP64 row sums, x and scalars, all on the device, graph per iteration */
#define TS_LSQR_FAST_MIXED 2 /* as TS_LSQR_FAST_FP32 with FP64 u/v/w: only the matrix values are FP32.
                                FP64 operands keep the iterates on the FP64 solver's path; FP32 operands
                                of an ill-conditioned system drift by ~1e-3 within tens of iterations */
#define TS_LSQR_FAST_FP64 3  /* the multi-RHS SpMM form with FP64 matrix values and FP64 u/v/w: the FP64
                                solver's arithmetic in another summation order (the matrix must be created
                                with keep_fp64); for weakly damped systems, where FP32 values move the
                                model by more than the FP64 solver's own rounding spread */

typedef struct ts_context ts_context;
typedef struct ts_tableset ts_tableset;
typedef struct ts_matrix ts_matrix;
typedef struct ts_rays ts_rays;
typedef struct ts_gridbatch ts_gridbatch;
typedef struct ts_server ts_server;

/* The spherical grid and its metric (SphericalEikonal.cs:15-35 GridMetric): node counts, HLon[Ny*Nz],
 * HLat[Nz], HR, plus the grid bounds the ray tracer needs (GridDefinition, SphericalGrid.cs:13-31). */
typedef struct {
    int32_t nx, ny, nz;
    double min_lon, max_lon, min_lat, max_lat, min_depth, max_depth;
} ts_grid;

/* ---------------------------------------------------------------- device
 * Replaces EikonalOpenCl.TryCreate (EikonalOpenCl.cs:88-118) and LsqrOpenCl.TryCreate
 * (LsqrOpenCl.cs:208-229): opens a HIP device and runs the self-tests (eikonal <= 1e-6 of the
 * field range against a host Relax; LSQR <= 1e-8 relative against a host LSQR). Declines (> 0)
 * when there is no device or a self-test fails. */
int ts_hip_create(int device, int flags, ts_context **out, char *reason, int reason_len);
void ts_hip_destroy(ts_context *ctx);
/* Device name and the self-test results (EikonalOpenCl.Device, LsqrOpenCl.Device, SelfTest). */
int ts_hip_device_info(ts_context *ctx, char *buf, int len);
/* Self-test errors measured at create time: eikonal (FP32 and FP64 forms, relative to the field
 * range), LSQR (plain FP64 form, relative to max |x|) and the fast FP32 LSQR form. */
int ts_hip_selftest(ts_context *ctx, double *eikonal_err, double *lsqr_err);
int ts_hip_selftest_lsqr_fast(ts_context *ctx, double *err);

/* ---------------------------------------------------------------- grid helpers (host side)
 * GridMetric (SphericalEikonal.cs:17-29): hlon[ny*nz], hlat[nz], *hr. */
int ts_grid_metric(const ts_grid *grid, double *hlon, double *hlat, double *hr);
/* SphericalEikonal.PointSource (SphericalEikonal.cs:68-100): seeds of a point source. Returns the
 * seed count (only `cap` are written; call again with a larger buffer when it exceeds cap), < 0 on
 * invalid input. */
int ts_point_source(const ts_grid *grid, const double *slowness, double lon, double lat, double depth_km,
                    double origin_time, double radius_cells, int32_t *index_out, double *time_out, int cap);

/* The correctly rounded sine, cosine, arcsine and arctangent every position of the library is computed
 * with, in place of the platform's Math.Sin / Cos / Asin / Atan2 that GeoMath calls (GeoMath.cs:49-80):
 * out[q] = f(a[q]) for func 0 sin, 1 cos, 2 asin, or atan2(a[q], b[q]) for func 3; on the host when ctx
 * is NULL, else on the context's device. */
int ts_crmath_eval(ts_context *ctx, int func, int n, const double *a, const double *b, double *out);

/* ---------------------------------------------------------------- eikonal
 * EikonalOpenCl.SuggestBatch (EikonalOpenCl.cs:121-132): fields per plain batch for a grid. */
int ts_eik_suggest_batch(ts_context *ctx, int nodes);

/* EikonalOpenCl.TrySolve (EikonalOpenCl.cs:138-166): one field per seed set, one slowness model.
 * seed_offsets[batch+1] delimit each field's seeds in seed_index/seed_time (duplicates keep the
 * smallest time). output is field-major batch*N doubles. rounds_out (may be NULL) receives the
 * largest rounds any field took in phase 1 and 2 (EikonalOpenCl.LastStats). */
int ts_eik_solve(ts_context *ctx, const ts_grid *grid, const double *slowness, int batch,
                 const int64_t *seed_offsets, const int32_t *seed_index, const double *seed_time,
                 double *output, int mode, const volatile int32_t *cancel, int32_t *rounds_out,
                 char *reason, int reason_len);

/* Resident travel-time tables (the device half of TravelTimeTableSet, TravelTimeTable.cs:251-430):
 * `nmodels` slowness fields (e.g. P and S) on one grid, `nfields` point-source problems each using
 * slowness model_of_field[f]. Create uploads everything; solve runs on the device and leaves the
 * times in HBM; read copies fields out. */
int ts_tables_create(ts_context *ctx, const ts_grid *grid, int nmodels, const double *const *slowness,
                     int nfields, const int32_t *model_of_field, const int64_t *seed_offsets,
                     const int32_t *seed_index, const double *seed_time, ts_tableset **out,
                     char *reason, int reason_len);
int ts_tables_solve(ts_tableset *t, int mode, const volatile int32_t *cancel, int32_t *rounds_out,
                    char *reason, int reason_len);
/* Fields [first, first+count) as FP64, field-major. */
int ts_tables_read(ts_tableset *t, int first, int count, double *out);
/* The same fields as stored tables: float32, +inf unreached (TravelTimeTable.Write, :57-80). */
int ts_tables_read_f32(ts_tableset *t, int first, int count, float *out);
/* Per-field rounds: rounds[2*f] (first order), rounds[2*f+1] (second order) of the last solve. */
int ts_tables_rounds(ts_tableset *t, int32_t *rounds);
/* Node updates (stencil evaluations) a fast solve executes on FP32 and on FP64 times (counted on the
 * first fast solve of the set), the last solve's wall time, and the device time of its sweep graphs
 * (HIP events on its stream). */
int ts_tables_stats(ts_tableset *t, double *evaluations_fp32, double *evaluations_fp64, double *seconds,
                    double *sweep_seconds);
/* The last fast solve's sweep rounds and sweep-graph device seconds split by the width of the times
 * swept ([0] FP32, [1] FP64), and the plane kernels one round launches (8 orderings x planes). No
 * counterpart in the reference: it is the figure a roofline of the sweep kernel is computed from. */
int ts_tables_sweep_stats(ts_tableset *t, int32_t rounds[2], int32_t *plane_launches_per_round, double seconds[2]);
/* The rounds of the last fast solve, in order: kind[q] = 16 for a round on FP64 times (0: FP32) | how
 * the round swept (0 dirty flag read with every load, 1 flag first, 2 a few flags per wave, 3 the
 * streaming sweep over column tiles, 4 more flags per wave), seconds[q] its device seconds. Fills min(max, rounds) entries and
 * returns the number of rounds. No counterpart in the reference (diagnostics of the device schedule). */
int ts_tables_round_log(ts_tableset *t, int max, int32_t *kind, double *seconds);
void ts_tables_release(ts_tableset *t);

/* ---------------------------------------------------------------- uneven batches
 * Many single-field problems, each on its own grid (the adaptive and resolution-test runs, where
 * the reference calls EikonalOpenCl.TrySolve once per grid, EikonalOpenCl.cs:138-343): grids[q],
 * slowness[q] (nodes of grid q), the seeds of problem q at seed_offsets[q] .. seed_offsets[q+1].
 * Every sweep step is one launch over all problems, or (the streaming method, below) every sweep
 * ordering one persistent launch over all problems' tiles. mode: TS_EIK_FAST_FP64 (Relax to the last bit
 * per problem) or TS_EIK_FAST_FP32 (first order on FP32 times, second order on FP64 times). */
int ts_gridbatch_create(ts_context *ctx, int nproblems, const ts_grid *grids, const double *const *slowness,
                        const int64_t *seed_offsets, const int32_t *seed_index, const double *seed_time,
                        ts_gridbatch **out, char *reason, int reason_len);
/* Per size class of problem (< 2e5, < 8e5, >= 8e5 nodes), how many nodes a thread updates per step
 * (1, 2 or 4), chosen by timing one trial round of each class with each choice; choice[3] receives
 * the picks. Then the streaming tile width per class and the method (ts_gridbatch_schedule). No
 * counterpart in the reference: the batch's own tuner. */
int ts_gridbatch_tune(ts_gridbatch *b, int32_t choice[3]);
/* The batch's sweep: method 0 (one launch per plane step over every problem), 1 (streaming: one
 * persistent launch per ordering over a work queue of every problem's column tiles) or 2 (a run's
 * first round in plane launches, the rounds after it streaming), the streaming tile width (4, 8 or 16
 * columns along x; 64 / width along y) per size class, the plane method's nodes per thread per class,
 * and the tuner's trial mixed solve of the batch with each method (ms; 0 before ts_gridbatch_tune).
 * ts_gridbatch_tune picks all of them by timing; ts_gridbatch_set_schedule forces a method and
 * (tile_w non-null) the tile widths. No counterpart in the reference: the batch's own schedule. */
int ts_gridbatch_schedule(ts_gridbatch *b, int32_t *method, int32_t tile_w[3], int32_t nodes_per_thread[3],
                          double trial_ms[3]);
int ts_gridbatch_set_schedule(ts_gridbatch *b, int method, const int32_t tile_w[3]);
/* The rounds of the last solve, in order: kind[q] = 16 for a round on FP64 times (0: FP32) | the method
 * that swept it, seconds[q] its device seconds. Fills min(max, rounds) entries and returns the number of
 * rounds. No counterpart in the reference (diagnostics of the device schedule). */
int ts_gridbatch_round_log(ts_gridbatch *b, int max, int32_t *kind, double *seconds);
/* rounds_out[2 * nproblems]: rounds of each problem's first- and second-order phases. */
int ts_gridbatch_solve(ts_gridbatch *b, int mode, const volatile int32_t *cancel, int32_t *rounds_out,
                       char *reason, int reason_len);
/* Travel times of problem q (its grid's nodes, +inf unreached), FP64. */
int ts_gridbatch_read(ts_gridbatch *b, int problem, double *out);
/* Total nodes, the last solve's wall time and the device time of its sweep rounds (HIP events). */
int ts_gridbatch_stats(ts_gridbatch *b, int64_t *nodes, double *seconds, double *sweep_seconds);
/* Node updates the last solve executed (a dirty, unseeded node evaluated), on FP32 and on FP64 times:
 * the work behind the sweep time (EikonalOpenCl.LastStats, EikonalOpenCl.cs:77). */
int ts_gridbatch_updates(ts_gridbatch *b, double *fp32, double *fp64);
void ts_gridbatch_release(ts_gridbatch *b);

/* ---------------------------------------------------------------- rays
 * The device half of RayKernels.Compute (Tomography/RayKernels.cs:30-115): for every arrival r, the
 * back-trace RayTracing.Backtrack (Forward/RayTracing.cs:73-111) from events[3r..3r+2] (lon, lat,
 * depth) to the station of table field_of_ray[r] (field_station[3f..3f+2]) on the solved table set,
 * TCalc and the hypocentre gradient at the event (TravelTimeTable.Time/Gradient, :116-157), and
 * the kernel on the inversion grid (RayTracing.Kernel, :126-159): nodes ascending, lengths float32.
 * step_km is RayKernels' step, 0.5 * min(forward, inversion MinSpacingKm) (RayKernels.cs:62). */
int ts_rays_trace(ts_tableset *tables, const ts_grid *inversion, const double *field_station, int nrays,
                  const int32_t *field_of_ray, const double *events, double step_km, ts_rays **out,
                  char *reason, int reason_len);
/* The 1-D background of an extended forward domain (ForwardDomain.Background; VelocityModel1D,
 * Model/VelocityModel1D.cs:8-45): count nodes of depth_km (not decreasing; a repeated depth is a step,
 * a point on it takes the upper value), vp and vs (km/s, 0 < vs < vp), piecewise linear in depth,
 * constant above the first node and below the last. */
typedef struct ts_background {
    int count;
    const double *depth_km;
    const double *vp;
    const double *vs;
} ts_background;
/* ts_rays_trace on an extended forward domain (RayKernels.cs:62-64, 95): with background non-NULL,
 * every kernel piece whose midpoint lies outside the inversion grid (SphericalGrid.Contains, tolerance
 * 1e-6 in node units) is not attributed to the grid's edge nodes; its length times 1/v_bg(depth) of the
 * arrival's phase (phase_of_ray[r]: 0 P, 1 S) is summed, in piece order, into the row's OutsideTime
 * (RayTracing.Kernel, RayTracing.cs:126-147). background NULL: ts_rays_trace (phase_of_ray unused,
 * may be NULL). The background is validated as VelocityModel1D.Validate (:47-56) does. */
int ts_rays_trace_extended(ts_tableset *tables, const ts_grid *inversion, const double *field_station, int nrays,
                           const int32_t *field_of_ray, const int32_t *phase_of_ray, const double *events,
                           double step_km, const ts_background *background, ts_rays **out, char *reason,
                           int reason_len);
/* RayRow.OutsideTime (RayKernels.cs:18) per arrival, outside[nrays]: 0 for arrivals without a row and
 * for rays traced without a background. */
int ts_rays_outside_time(ts_rays *r, double *outside);
/* Rows in arrival order: counts[nrays] (nodes per row, -1 = no row: stalled descent or unreached
 * event, as RayKernels skips it), row_start[nrays+1] (CSR over all arrivals, empty rows for the
 * skipped ones), nodes/lengths[row_start[nrays]], tcalc_grad[4*nrays] (TCalc, Ge, Gn, Gz) and
 * path_points[nrays]. Any pointer may be NULL. */
int ts_rays_rows(ts_rays *r, int32_t *counts, int64_t *row_start, int32_t *nodes, float *lengths,
                 double *tcalc_grad, int32_t *path_points);
/* Non-zeros of all rows and the wall time of the trace. */
int ts_rays_stats(ts_rays *r, int64_t *nnz, double *seconds);
void ts_rays_release(ts_rays *r);

/* ---------------------------------------------------------------- event location
 * EventLocator.GridSearch (Location/EventLocator.cs:114-175, WeightedMedian :311-323) over the tables
 * of a solved table set, for many events at once: per event the picks
 * [pick_start[e], pick_start[e+1]) with their table (field index), observed time relative to the
 * event's reference (s), weight and station correction. center[3e..3e+2] (i, j, k), when given, makes
 * the search local (every node within 6 of it; LocatorSettings.GlobalSearch = false); otherwise a
 * strided sub-grid of about coarse_candidates nodes, then every node within two strides of the best.
 * k_fixed[e] >= 0 fixes the depth index (FixDepth). Out: the best node index, its origin time (the
 * weighted median) and its misfit, equal to the reference's. At most 512 picks per event (more:
 * declined). */
int ts_locate_grid_search(ts_tableset *t, int nevents, const int32_t *pick_start, const int32_t *pick_field,
                          const double *obs, const double *weight, const double *correction, int coarse_candidates,
                          const int32_t *center, const int32_t *k_fixed, int32_t *best_node, double *best_t0,
                          double *best_misfit, char *reason, int reason_len);

/* ---------------------------------------------------------------- LSQR
 * A CSR matrix resident in HBM with its stored transpose (CsrMatrix and CsrMatrix.Transpose,
 * Numerics/SparseMatrix.cs:24-97; int32 indices as CsrMatrix's). row_start[rows+1], column[nnz],
 * values[nnz]. keep_fp64 != 0 keeps FP64 values for the plain form (FP32 values are always kept). */
int ts_matrix_create(ts_context *ctx, int rows, int cols, const int32_t *row_start, const int32_t *column,
                     const double *values, int keep_fp64, ts_matrix **out, char *reason, int reason_len);
/* Non-zeros, and the device time of the last fast solve's iterations (HIP events on its stream). */
int ts_matrix_stats(ts_matrix *a, int64_t *nnz, double *seconds);
/* Rows, columns, non-zeros and the wall time of the upload or assembly (transpose included). */
int ts_matrix_shape(ts_matrix *a, int32_t *rows, int32_t *cols, int64_t *nnz, double *build_seconds);
/* The matrix back in host memory (CsrMatrix's arrays): row_start[rows+1], column[nnz], values[nnz]
 * (FP64 when kept, else the FP32 values widened) and, for a matrix from ts_system_build, the
 * right-hand side rhs[rows]. Any pointer may be NULL. */
int ts_matrix_read(ts_matrix *a, int32_t *row_start, int32_t *column, double *values, double *rhs);
/* The stored transpose (CsrMatrix.Transpose, SparseMatrix.cs:79-97): row_start[cols+1], column[nnz]
 * (each transposed row's rows ascending) and its FP32 values. */
int ts_matrix_read_transpose(ts_matrix *a, int32_t *row_start, int32_t *column, float *values);
/* The order the matrix is stored in on the device (the products and solves visit rows and columns in
 * it; CsrMatrix keeps the given order, SparseMatrix.cs:24-44): stored row r is given row row_perm[r],
 * stored column c is given column col_perm[c]; each stored row keeps its entries in the given order.
 * Either pointer may be NULL. */
int ts_matrix_stored_order(ts_matrix *a, int32_t *row_perm, int32_t *col_perm);
/* y = A x (transpose = 0) or y = A^T x (transpose != 0) for k vectors on the device, FP64 sums over the
 * FP64 values when kept, else the FP32 ones (CsrMatrix.MultiplyInto, Numerics/SparseMatrix.cs:46-73;
 * ResolutionDiagonal.Estimate's A z and TradeOff's G x, Tomography/ResolutionDiagonal.cs:48,
 * TravelTimeTomography.cs:599). x[k][cols or rows], y[k][rows or cols], in the given orders. */
int ts_matrix_multiply(ts_matrix *a, int k, const double *x, double *y, int transpose, char *reason, int reason_len);
/* ts_matrix_multiply without any device allocation: no counterpart in the reference. It is for callers
 * that cannot allocate device memory during the call (code under stream capture, tight serving loops,
 * harnesses that forbid hipMalloc). The input and output are staged in one 64 MiB array per GPU,
 * allocated with the library's code object and held until the process ends, so a call needs
 * 2 x (rows + cols) <= 8,388,608 doubles; larger matrices are rejected with TS_EINVAL and a reason.
 * Results equal ts_matrix_multiply's bit for bit (the same kernels). Calls on one GPU share its array:
 * the library makes them take turns, and the array keeps the last call's contents. */
int ts_matrix_multiply_noalloc(ts_matrix *a, int k, const double *x, double *y, int transpose, char *reason,
                               int reason_len);
void ts_matrix_release(ts_matrix *a);

/* ---------------------------------------------------------------- the linear system
 * TravelTimeTomography.BuildSystem (Tomography/TravelTimeTomography.cs:860-1034) on a plain node grid
 * (no parameter lattice, adaptive mesh or layered model), assembled on the device from ray rows
 * resident in HBM into an LSQR matrix. Settings follow TomographySettings (TravelTimeTomography.cs:
 * 30-110); any_p / any_s say whether the data hold P / S arrivals at all (rejected ones included,
 * :867-868 and :879-882), which fixes the column layout. */
typedef struct {
    int32_t vpvs;                /* VelocityParameterization.VpVpVs: S rows on (P, Vp/Vs) columns */
    int32_t invert_p, invert_s;  /* InvertP, InvertS */
    int32_t any_p, any_s;
    int32_t joint_hypocentres;   /* JointHypocentres: 4 columns per event */
    int32_t station_corrections; /* StationCorrections: one column per station and phase */
    int32_t gradient_smoothing;  /* 0: SmoothingMethod.Laplacian, 1: first differences (TotalVariation
                                    without reweighting) */
    int32_t smoothing_in_nodes;  /* SmoothingScale.Nodes (axis weights 1, 1, vertical), else Kilometres */
    int32_t nevents, nstations;
    double damping, smoothing;   /* the iteration's damping and smoothing weights */
    double vertical_weight;      /* VerticalSmoothingWeight */
    double damping_hypocentre;   /* DampingHypocentre */
    double damping_station;      /* DampingStation */
    double s_factor;             /* SRegularisationFactor (<= 0: 1) */
    double absolute_weight;      /* weight of the absolute-time rows (> 0 adds them) */
} ts_system_opts;

/* rays: the traced rows of the arrivals (ts_rays_trace). Data row r of the system is arrival
 * ray_of_row[r] (in the order given: the reference adds the used rows in arrival order, :953-972),
 * with its phase (0 P, 1 S), event, station, residual (observed - calculated, s) and sigma (s).
 * event_fixed / event_outside [nevents]: Event.Fixed / Event.Outside. sP, sS: the current slowness
 * models; sRefP, sRefS: the reference models; kRef: the reference Vp/Vs (sRefS / sRefP); all [nodes
 * of grid]. Output: the matrix (FP32 values, and FP64 when keep_fp64), its right-hand side
 * (ts_matrix_read) and layout[8] = {offP, offS, offH, offCP, offCS (-1: absent), columns, rows, data
 * rows} (MatrixLayout, :1031). */
int ts_system_build(ts_context *ctx, const ts_grid *grid, const ts_system_opts *opts, ts_rays *rays, int nrows,
                    const int32_t *ray_of_row, const int32_t *phase, const int32_t *event, const int32_t *station,
                    const double *residual, const double *sigma, const uint8_t *event_fixed,
                    const uint8_t *event_outside, const double *sP, const double *sS, const double *sRefP,
                    const double *sRefS, const double *kRef, int keep_fp64, ts_matrix **out, int64_t layout[8],
                    char *reason, int reason_len);

/* ---------------------------------------------------------------- Q (attenuation) tomography
 * QTomography.Predict (Attenuation/QTomography.cs:307-312) over rows of resident rays: per row r (arrival
 * ray_of_row[r]) the predicted t* = q0 * OutsideTime + sum L s q and the path time OutsideTime +
 * sum L s (Prepare's Q0 estimate, :231-233), each summed from the outside part on in the row's node
 * order (OutsideTime is 0 for rays traced without a background). slowness, q: [nodes of the rays'
 * inversion grid]; q NULL = q0 at every node. */
int ts_q_predict(ts_rays *rays, int nrows, const int32_t *ray_of_row, const double *slowness, const double *q,
                 double q0, double *pred, double *travel_time, char *reason, int reason_len);

/* QTomography.SolveOnce's system (QTomography.cs:318-392) on a plain node grid (no parameter lattice or
 * adaptive mesh; Laplacian smoothing or unweighted first differences), assembled on the device. */
typedef struct {
    int32_t station_terms;       /* StationTerms: one column per station after the nodes */
    int32_t gradient_smoothing;  /* 0: SmoothingMethod.Laplacian, 1: first differences (no reweighting) */
    int32_t smoothing_in_nodes;  /* SmoothingScale.Nodes */
    int32_t nstations;           /* data.Stations.Count */
    int32_t fixed_scales;        /* 1: scales[] as given (the Gauss-Newton iterations after the first) */
    double damping, smoothing;   /* relative to the sensitivity scales */
    double vertical_weight;      /* VerticalSmoothingWeight */
    double damping_station;      /* DampingStation */
    double q0;                   /* the reference q = 1 / Q0 */
} ts_qsystem_opts;

/* Data row r is arrival ray_of_row[r] (the used rows, in order) with station[r], t*[r] and sigma[r]:
 * entries L s q at the nodes (q = q_old, or q0 everywhere when q_old is NULL), 1 at the station's
 * column, right-hand side t* - Predict(q0) (q_old NULL) or t* - Predict(q_old) + sum (L s q_old) m_old
 * (:344-349), all times 1 / sigma; then damping and smoothing of the nodes times the model scale, the
 * station damping and the zero-sum row (weight 10) times the station scale. q_old and m_old are both
 * given or both NULL. scales[2] (model, station): the RMS of each block's weighted entries (:356-372),
 * written on return, or read when fixed_scales. layout[4] = {offset of the station columns (-1: none),
 * columns, rows, data rows}. */
int ts_qsystem_build(ts_context *ctx, const ts_grid *grid, const ts_qsystem_opts *opts, ts_rays *rays, int nrows,
                     const int32_t *ray_of_row, const int32_t *station, const double *tstar, const double *sigma,
                     const double *slowness, const double *q_old, const double *m_old, int keep_fp64,
                     double scales[2], ts_matrix **out, int64_t layout[4], char *reason, int reason_len);

/* ---------------------------------------------------------------- formal resolution
 * FormalResolution.Compute (Tomography/FormalResolution.cs:108-333) on the device: for a weighted system
 * (CSR in the given order, FP64 values; rows below data_rows are data, the rest regularisation) and
 * ngroups groups of columns to separate out (group g is group_columns[group_start[g] ..
 * group_start[g+1]): the 4 hypocentre columns of an event), the kept columns (kept_out[*nkept], in
 * order), their resolution diagonal R_ii, posterior standard deviation sqrt(C_ii) and, when places
 * (xyz[3 cols], geocentric km) and place_block[cols] (-1: no place) are given, the resolution length
 * in km (NaN without a place). Output arrays hold cols entries. More than max_unknowns kept columns
 * declines (> 0); a row coupling two groups, a regularisation row coupling kept and separated columns,
 * or a reduced normal matrix that is not positive definite are errors (< 0), with the reference's
 * reasons. seconds: wall time of the call. */
int ts_formal_resolution(ts_context *ctx, int rows, int cols, const int32_t *row_start, const int32_t *column,
                         const double *values, int data_rows, int ngroups, const int32_t *group_start,
                         const int32_t *group_columns, int max_unknowns, const double *places,
                         const int32_t *place_block, int32_t *kept_out, int32_t *nkept, double *diagonal,
                         double *stddev, double *length_km, double *seconds, char *reason, int reason_len);

/* Lsqr.Solve (Numerics/Lsqr.cs:34-152) / LsqrOpenCl.TrySolve (LsqrOpenCl.cs:244-270) for k (1..64)
 * right-hand sides b[k*m] (RHS-major), each with its own damping damp[k]: minimises
 * ||A x_j - b_j||^2 + damp_j^2 ||x_j||^2 with the reference's stopping rules (atol, btol, conlim).
 * Outputs per RHS: x[k*n], iterations[k], stop[k] (0 budget exhausted, 1 btol, 2 atol, 3 conlim,
 * 4-6 machine precision, 7 bidiagonalisation terminated, 8 empty system, 9 zero right-hand side,
 * 10 A^T b = 0: the order of Lsqr.cs:141-147), stats[3k] (bnorm, rnorm, acond) and, if not NULL,
 * the residual norm of every iteration rnorm_history[k*max_iterations] (onIteration, Lsqr.cs:140). */
int ts_lsqr_solve(ts_matrix *a, int k, const double *b, const double *damp, int max_iterations, double atol,
                  double btol, double conlim, int mode, const volatile int32_t *cancel, double *x,
                  int32_t *iterations, int32_t *stop, double *stats, double *rnorm_history, char *reason,
                  int reason_len);

/* ---------------------------------------------------------------- serving
 * A request loop on one GPU (no counterpart in the reference, whose device types serialise their
 * solves with lock(_gate), EikonalOpenCl.cs:64,165 and LsqrOpenCl.cs:184,257): `workers` threads, each
 * with its own stream, run table and LSQR requests concurrently; a request waits in a FIFO until its
 * estimated device memory fits what memory_budget (bytes) leaves; the head of the queue is never
 * overtaken. Velocity models are resident (ts_server_add_model); LSQR requests name a resident
 * matrix. A request gives the same result alone and in the mix. */
int ts_server_create(ts_context *ctx, int workers, uint64_t memory_budget, ts_server **out, char *reason,
                     int reason_len);
/* A slowness model resident on the device (FP64 and FP32), kept until the server is destroyed. */
int ts_server_add_model(ts_server *s, const ts_grid *grid, const double *slowness, int32_t *model_id, char *reason,
                        int reason_len);
/* Tables of sources[3*nsources] (lon, lat, depth; point sources, SphericalEikonal.PointSource) in the
 * P model and, when model_s >= 0, the S model: fields 2s (P) and 2s+1 (S), or s (P only). mode: a
 * TS_EIK_* mode. The tables stay on the device (and counted against the budget) until the request is
 * released. */
int ts_server_submit_tables(ts_server *s, int32_t model_p, int32_t model_s, int nsources, const double *sources,
                            int mode, int64_t *request_id, char *reason, int reason_len);
/* ts_lsqr_solve on a resident matrix (b is copied at submission). */
int ts_server_submit_lsqr(ts_server *s, ts_matrix *a, int k, const double *b, const double *damp, int max_iterations,
                          double atol, double btol, double conlim, int mode, int64_t *request_id, char *reason,
                          int reason_len);
/* Waits up to timeout_s for a request (TS_DECLINED on time-out); status is the request's own status
 * and times[3] its submission, start and end (seconds since the server started). */
int ts_server_wait(ts_server *s, int64_t request_id, double timeout_s, int32_t *status, double times[3], char *reason,
                   int reason_len);
int ts_server_result_tables(ts_server *s, int64_t request_id, int first, int count, float *out);
int ts_server_result_lsqr(ts_server *s, int64_t request_id, double *x, int32_t *iterations, int32_t *stop,
                          double *stats);
/* Frees a finished request and its device memory. */
int ts_server_release_request(ts_server *s, int64_t request_id);
int ts_server_stats(ts_server *s, uint64_t *peak_bytes, int32_t *peak_inflight, uint64_t *budget);
void ts_server_destroy(ts_server *s);

#ifdef __cplusplus
}
#endif
#endif
