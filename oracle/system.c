/*
 * Oracle (test infrastructure only): the tomographic linear system of the reference on a plain node
 * grid. Restates TravelTimeTomography.BuildSystem (Tomography/TravelTimeTomography.cs:860-1034: the
 * column layout, Terms, Packed, AddRows, the row weight and the regularisation order) and the row
 * generators of Tomography/Observations.cs (AddDamping :474-486, AddLaplacian :410-443, AddGradient
 * :360-384, AxisWeights :390-397), with CsrBuilder.AddRow's weighting (Numerics/SparseMatrix.cs:128-140).
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define OR_R 6371.0
#define OR_D2R (M_PI / 180.0)

typedef struct {
    int64_t rows, nnz, cap_rows, cap_nnz;
    int64_t *rs;
    int *ci;
    double *va, *rhs;
    int overflow;
} builder;

/* CsrBuilder.AddRow (SparseMatrix.cs:128-140): exact zeros dropped, row and rhs times the weight */
static void add_row(builder *b, int n, const int *cols, const double *vals, double rhs, double weight)
{
    if (b->rows >= b->cap_rows) { b->overflow = 1; return; }
    for (int q = 0; q < n; q++) {
        if (vals[q] == 0) continue;
        if (b->nnz >= b->cap_nnz) { b->overflow = 1; return; }
        b->ci[b->nnz] = cols[q];
        b->va[b->nnz] = vals[q] * weight;
        b->nnz++;
    }
    b->rhs[b->rows] = rhs * weight;
    b->rows++;
    b->rs[b->rows] = b->nnz;
}

static void add_damping(builder *b, int offset, int count, double lambda)
{
    if (lambda <= 0) return;
    for (int n = 0; n < count; n++) {
        int c = offset + n;
        double v = 1;
        add_row(b, 1, &c, &v, 0, lambda);
    }
}

/* Observations.cs:390-397 AxisWeights, kilometre scale */
static void axis_weights(const or_griddef *g, double vw, int in_nodes, double w[3])
{
    if (in_nodes) { w[0] = 1; w[1] = 1; w[2] = vw; return; } /* SmoothingScale.Nodes */
    /* GridDefinition.DLon / DLat / DDepth (Geo/SphericalGrid.cs:28-30) */
    double dlon = (g->max_lon - g->min_lon) / (g->nx - 1 > 1 ? g->nx - 1 : 1);
    double dlat = (g->max_lat - g->min_lat) / (g->ny - 1 > 1 ? g->ny - 1 : 1);
    double ddep = (g->max_depth - g->min_depth) / (g->nz - 1 > 1 ? g->nz - 1 : 1);
    double r = OR_R - g->min_depth;
    double lat = g->min_lat + (g->ny / 2) * dlat;
    double hx = r * or_cos(lat * OR_D2R) * (dlon * OR_D2R);
    if (!(hx > 1e-9)) hx = 1e-9;
    double hy = r * (dlat * OR_D2R), hz = ddep;
    double h = sqrt(hx * hy);
    w[0] = pow(h / hx, 2);
    w[1] = pow(h / hy, 2);
    w[2] = vw * pow(h / hz, 2);
}

static void add_laplacian(builder *b, const or_griddef *g, int offset, double lambda, double vw, int in_nodes,
                          const double *cur)
{
    if (lambda <= 0) return;
    double w[3];
    axis_weights(g, vw, in_nodes, w);
    int nx = g->nx, ny = g->ny, nz = g->nz;
    for (int k = 0; k < nz; k++)
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++) {
                int n = (k * ny + j) * nx + i, c = 0, cols[7];
                double vals[7], diag = 0;
                const int d[6][3] = {{-1, 0, 0}, {1, 0, 0}, {0, -1, 0}, {0, 1, 0}, {0, 0, -1}, {0, 0, 1}};
                for (int q = 0; q < 6; q++) {
                    int ii = i + d[q][0], jj = j + d[q][1], kk = k + d[q][2];
                    if (ii < 0 || ii >= nx || jj < 0 || jj >= ny || kk < 0 || kk >= nz) continue;
                    double wq = w[q / 2];
                    cols[c] = offset + (kk * ny + jj) * nx + ii;
                    vals[c] = -wq;
                    c++;
                    diag += wq;
                }
                cols[c] = offset + n;
                vals[c] = diag;
                c++;
                double rhs = 0;
                if (cur)
                    for (int q = 0; q < c; q++) rhs -= vals[q] * cur[cols[q] - offset];
                add_row(b, c, cols, vals, rhs, lambda);
            }
}

static void add_gradient(builder *b, const or_griddef *g, int offset, double lambda, double vw, int in_nodes,
                         const double *cur)
{
    if (lambda <= 0) return;
    double w[3];
    axis_weights(g, vw, in_nodes, w);
    int nx = g->nx, ny = g->ny, nz = g->nz;
    for (int k = 0; k < nz; k++)
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++) {
                int n = (k * ny + j) * nx + i;
                const int d[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
                for (int q = 0; q < 3; q++) {
                    int ii = i + d[q][0], jj = j + d[q][1], kk = k + d[q][2];
                    if (ii >= nx || jj >= ny || kk >= nz || w[q] <= 0) continue;
                    int m = (kk * ny + jj) * nx + ii;
                    double a = sqrt(w[q]);
                    int cols[2] = {offset + n, offset + m};
                    double vals[2] = {-a, a};
                    double rhs = cur ? -a * (cur[m] - cur[n]) : 0;
                    add_row(b, 2, cols, vals, rhs, lambda);
                }
            }
}

typedef struct { int col; double v; } entry;
static int entry_cmp(const void *a, const void *b)
{
    int x = ((const entry *)a)->col, y = ((const entry *)b)->col;
    return x < y ? -1 : x > y;
}

/* Returns 0, or -1 when the output arrays are too small. dims = {rows, cols, nnz, data rows}. */
int or_build_system(const or_griddef *g, const or_system_opts *o, int nrows, const int64_t *row_start,
                    const int *row_nodes, const float *row_len, const double *row_tg, const int *phase,
                    const int *event, const int *station, const double *residual, const double *sigma,
                    const unsigned char *ev_fixed, const unsigned char *ev_outside, const double *sP,
                    const double *sS, const double *sRefP, const double *sRefS, const double *kRef,
                    int64_t cap_rows, int64_t cap_nnz, int64_t *rs, int *ci, double *va, double *rhs,
                    int64_t dims[4])
{
    const int n = g->nx * g->ny * g->nz;
    const int has_p = o->invert_p && o->any_p, has_s = o->invert_s && o->any_s;
    int col = 0, offP = -1, offS = -1, offH = -1, offCP = -1, offCS = -1;
    if (has_p || (has_s && o->vpvs)) { offP = col; col += n; }
    if (has_s) { offS = col; col += n; }
    if (o->joint_hypocentres) { offH = col; col += 4 * o->nevents; }
    if (o->station_corrections) {
        if (o->any_p) { offCP = col; col += o->nstations; }
        if (o->any_s) { offCS = col; col += o->nstations; }
    }
    builder b = {0, 0, cap_rows, cap_nnz, rs, ci, va, rhs, 0};
    rs[0] = 0;
    double *kcur = malloc(sizeof(double) * n);
    for (int q = 0; q < n; q++) kcur[q] = sS[q] / sP[q];
    /* data rows: Terms (:890-918) with the station term, Packed (:921-926), AddRow with abs weight / sigma */
    int ecap = 64;
    entry *acc = malloc(sizeof(entry) * ecap);
    for (int r = 0; r < nrows; r++) {
        int cnt = (int)(row_start[r + 1] - row_start[r]);
        int need = 2 * cnt + 8;
        if (need > ecap) { ecap = need; acc = realloc(acc, sizeof(entry) * ecap); }
        int na = 0;
        const int *nodes = row_nodes + row_start[r];
        const float *len = row_len + row_start[r];
        if (phase[r] == 0) {
            if (offP >= 0 && o->invert_p)
                for (int q = 0; q < cnt; q++) { acc[na].col = offP + nodes[q]; acc[na].v = len[q] * sRefP[nodes[q]]; na++; }
        } else if (offS >= 0) {
            if (o->vpvs) {
                for (int q = 0; q < cnt; q++) {
                    int nd = nodes[q];
                    if (offP >= 0 && o->invert_p) { acc[na].col = offP + nd; acc[na].v = len[q] * kcur[nd] * sRefP[nd]; na++; }
                    acc[na].col = offS + nd; acc[na].v = len[q] * sP[nd] * kRef[nd]; na++;
                }
            } else {
                for (int q = 0; q < cnt; q++) { acc[na].col = offS + nodes[q]; acc[na].v = len[q] * sRefS[nodes[q]]; na++; }
            }
        }
        if (offH >= 0 && !ev_fixed[event[r]]) {
            int h = offH + 4 * event[r];
            acc[na].col = h; acc[na].v = 1; na++;
            if (!ev_outside[event[r]]) {
                acc[na].col = h + 1; acc[na].v = row_tg[4 * r + 1]; na++;
                acc[na].col = h + 2; acc[na].v = row_tg[4 * r + 2]; na++;
                acc[na].col = h + 3; acc[na].v = row_tg[4 * r + 3]; na++;
            }
        }
        int coff = phase[r] == 0 ? offCP : offCS;
        if (coff >= 0) { acc[na].col = coff + station[r]; acc[na].v = 1; na++; }
        qsort(acc, na, sizeof(entry), entry_cmp);
        int kept = 0;
        int cols[na > 0 ? na : 1];
        double vals[na > 0 ? na : 1];
        for (int q = 0; q < na; q++)
            if (fabs(acc[q].v) >= 1e-4) { cols[kept] = acc[q].col; vals[kept] = acc[q].v; kept++; }
        if (kept > 0) {
            double w = o->absolute_weight / sigma[r];
            add_row(&b, kept, cols, vals, residual[r], w);
        }
    }
    free(acc);
    int64_t data_rows = b.rows;
    /* regularisation (:976-1017), plain node grid, not layered */
    double *dev = malloc(sizeof(double) * n);
    if (offP >= 0) {
        add_damping(&b, offP, n, o->damping * 1);
        for (int q = 0; q < n; q++) dev[q] = sP[q] / sRefP[q] - 1;
        if (o->gradient_smoothing) add_gradient(&b, g, offP, o->smoothing * 1, o->vertical_weight, o->smoothing_in_nodes, dev);
        else add_laplacian(&b, g, offP, o->smoothing * 1, o->vertical_weight, o->smoothing_in_nodes, dev);
    }
    if (offS >= 0) {
        double f = o->s_factor > 0 ? o->s_factor : 1;
        add_damping(&b, offS, n, o->damping * f);
        for (int q = 0; q < n; q++) dev[q] = o->vpvs ? kcur[q] / kRef[q] - 1 : sS[q] / sRefS[q] - 1;
        if (o->gradient_smoothing) add_gradient(&b, g, offS, o->smoothing * f, o->vertical_weight, o->smoothing_in_nodes, dev);
        else add_laplacian(&b, g, offS, o->smoothing * f, o->vertical_weight, o->smoothing_in_nodes, dev);
    }
    if (offH >= 0) add_damping(&b, offH, 4 * o->nevents, o->damping_hypocentre);
    int offs[2] = {offCP, offCS};
    for (int q = 0; q < 2; q++) {
        if (offs[q] < 0) continue;
        add_damping(&b, offs[q], o->nstations, o->damping_station);
        int *zc = malloc(sizeof(int) * o->nstations);
        double *zv = malloc(sizeof(double) * o->nstations);
        for (int s = 0; s < o->nstations; s++) { zc[s] = offs[q] + s; zv[s] = 1.0; }
        add_row(&b, o->nstations, zc, zv, 0, 10);
        free(zc); free(zv);
    }
    free(dev);
    free(kcur);
    dims[0] = b.rows;
    dims[1] = col;
    dims[2] = b.nnz;
    dims[3] = data_rows;
    return b.overflow ? -1 : 0;
}

/* ---------------------------------------------------------------- Q tomography
 * QTomography.Predict (Attenuation/QTomography.cs:307-312) and the path time of Prepare's Q0 estimate
 * (:231-233): per row t* = q0 * OutsideTime + sum L s q and T = OutsideTime + sum L s, in the row's
 * order. outside[r]: the row's OutsideTime (NULL: 0, rows inside the grid). q NULL: q0 at every node. */
void or_q_predict(int nrows, const int64_t *row_start, const int *row_nodes, const float *row_len,
                  const double *slow, const double *q, double q0, const double *outside, double *pred, double *tt)
{
    for (int r = 0; r < nrows; r++) {
        /* Predict starts from q0 * OutsideTime (:309), the Q0 estimate's path time from OutsideTime (:231) */
        const double out = outside ? outside[r] : 0.0;
        double t = q0 * out, T = out;
        for (int64_t k = row_start[r]; k < row_start[r + 1]; k++) {
            int nd = row_nodes[k];
            t += (double)row_len[k] * slow[nd] * (q ? q[nd] : q0);
            T += (double)row_len[k] * slow[nd];
        }
        if (pred) pred[r] = t;
        if (tt) tt[r] = T;
    }
}

/* QTomography.SolveOnce's system (Attenuation/QTomography.cs:318-392) on a plain node grid: columns
 * [nodes | stations (StationTerms)]; per used row (in the order given) the node entries L s q (q_old,
 * or q0 everywhere when q_old is NULL), the station entry 1 and the data right-hand side - the start
 * residual t* - Predict(q0) without m_old, else t* - Predict(q_old) + sum (L s q_old) m_old (:344-349)
 * - all times 1 / sigma; the sensitivity scales (:365-372: RMS of the weighted entries of each block,
 * or the given ones when fixed_scales); damping and smoothing of the nodes times the model scale, the
 * station damping and the zero-sum row (weight 10) times the station scale (:379-392).
 * scales[2] in/out (model, station). Returns 0, or -1 when the output arrays are too small;
 * dims = {rows, cols, nnz, data rows}. */
int or_build_qsystem(const or_griddef *g, const or_qsystem_opts *o, int nrows, const int64_t *row_start,
                     const int *row_nodes, const float *row_len, const double *outside, const int *station,
                     const double *tstar, const double *sigma, const double *slow, const double *q_old,
                     const double *m_old, int64_t cap_rows, int64_t cap_nnz, int64_t *rs, int *ci, double *va,
                     double *rhs, int64_t dims[4], double scales[2])
{
    const int n = g->nx * g->ny * g->nz;
    const int off_st = o->station_terms ? n : -1;
    const int cols = n + (o->station_terms ? o->nstations : 0);
    builder b = {0, 0, cap_rows, cap_nnz, rs, ci, va, rhs, 0};
    rs[0] = 0;
    int maxlen = 1;
    for (int r = 0; r < nrows; r++)
        if (row_start[r + 1] - row_start[r] + 1 > maxlen) maxlen = (int)(row_start[r + 1] - row_start[r] + 1);
    int *c = malloc(sizeof(int) * maxlen);
    double *v = malloc(sizeof(double) * maxlen);
    double sum_model = 0, sum_station = 0;
    int64_t count_model = 0, count_station = 0;
    for (int r = 0; r < nrows; r++) {
        int len = 0;
        for (int64_t k = row_start[r]; k < row_start[r + 1]; k++) {
            int nd = row_nodes[k];
            c[len] = nd;
            v[len] = (double)row_len[k] * slow[nd] * (q_old ? q_old[nd] : o->q0);
            len++;
        }
        double d;
        if (!q_old) {
            double p = 0;
            or_q_predict(1, row_start + r, row_nodes, row_len, slow, NULL, o->q0, outside ? outside + r : NULL, &p, NULL);
            d = tstar[r] - p;
        } else {
            double p = 0;
            or_q_predict(1, row_start + r, row_nodes, row_len, slow, q_old, o->q0, outside ? outside + r : NULL, &p, NULL);
            d = tstar[r] - p;
            for (int k = 0; k < len; k++) d += v[k] * m_old[c[k]];
        }
        if (off_st >= 0) { c[len] = off_st + station[r]; v[len] = 1; len++; }
        for (int k = 0; k < len; k++) {
            double w = v[k] / sigma[r];
            if (c[k] < n) { sum_model += w * w; count_model++; }
            else { sum_station += w * w; count_station++; }
        }
        add_row(&b, len, c, v, d, 1 / sigma[r]);
    }
    free(c);
    free(v);
    const int64_t data_rows = b.rows;
    double sm = count_model > 0 ? sqrt(sum_model / count_model) : 1;
    double ss = count_station > 0 ? sqrt(sum_station / count_station) : 1;
    if (o->fixed_scales) { sm = scales[0]; ss = scales[1]; }
    scales[0] = sm;
    scales[1] = ss;
    add_damping(&b, 0, n, o->damping * sm);
    if (o->gradient_smoothing)
        add_gradient(&b, g, 0, o->smoothing * sm, o->vertical_weight, o->smoothing_in_nodes, NULL);
    else
        add_laplacian(&b, g, 0, o->smoothing * sm, o->vertical_weight, o->smoothing_in_nodes, NULL);
    if (off_st >= 0) {
        add_damping(&b, off_st, o->nstations, o->damping_station * ss);
        int *zc = malloc(sizeof(int) * (o->nstations + 1));
        double *zv = malloc(sizeof(double) * (o->nstations + 1));
        for (int s = 0; s < o->nstations; s++) { zc[s] = off_st + s; zv[s] = 1; }
        add_row(&b, o->nstations, zc, zv, 0, 10 * ss);
        free(zc);
        free(zv);
    }
    dims[0] = b.rows;
    dims[1] = cols;
    dims[2] = b.nnz;
    dims[3] = data_rows;
    return b.overflow ? -1 : 0;
}
