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
#define TS_EIK_FAST_FP32 1  /* batched: FP32 times in the offset form, 64 fields per wavefront, per-field
                               convergence, wavefront-ballot wake flags, one graph launch per round */
#define TS_EIK_FAST_FP64 2  /* the batched schedule with FP64 times: equal to SphericalEikonal.Relax of
                               each field to the last bit */

typedef struct ts_context ts_context;
typedef struct ts_tableset ts_tableset;
typedef struct ts_matrix ts_matrix;

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
/* Self-test errors measured at create time (eikonal relative to range, LSQR relative). */
int ts_hip_selftest(ts_context *ctx, double *eikonal_err, double *lsqr_err);

/* ---------------------------------------------------------------- grid helpers (host side)
 * GridMetric (SphericalEikonal.cs:17-29): hlon[ny*nz], hlat[nz], *hr. */
int ts_grid_metric(const ts_grid *grid, double *hlon, double *hlat, double *hr);
/* SphericalEikonal.PointSource (SphericalEikonal.cs:68-100): seeds of a point source. Returns the
 * seed count (only `cap` are written; call again with a larger buffer when it exceeds cap), < 0 on
 * invalid input. */
int ts_point_source(const ts_grid *grid, const double *slowness, double lon, double lat, double depth_km,
                    double origin_time, double radius_cells, int32_t *index_out, double *time_out, int cap);

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
/* Node updates (stencil evaluations) the solve executed (counted on the first fast solve of the set),
 * the last solve's wall time, and the device time of its sweep graphs (HIP events on its stream). */
int ts_tables_stats(ts_tableset *t, double *evaluations, double *seconds, double *sweep_seconds);
void ts_tables_release(ts_tableset *t);

#ifdef __cplusplus
}
#endif
#endif
