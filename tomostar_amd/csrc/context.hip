// Device handle of libtomostar_hip: opens a HIP device and gates it with the reference's self-tests
// before any solve is accepted (EikonalOpenCl.TryCreate/SelfTest, EikonalOpenCl.cs:88-118, :346-374;
// LsqrOpenCl.TryCreate/SelfTest, LsqrOpenCl.cs:208-229, :541-573).
//
// The gates compare the device with a host solve on a small problem, as the reference does: the
// eikonal against a host Relax (the same update arithmetic, walked plane by plane on the CPU), LSQR
// against a host LSQR in FP64. These host solves exist only to judge the device; the library never
// returns a host-computed result to a caller - a failed gate declines the device, and the caller's
// own CPU solver takes over, as after EikonalOpenCl.TryCreate returns null.
#include <algorithm>
#include <memory>
#include <vector>

#include "common.hpp"
#include "eikonal_update.hpp"
#include "geometry.hpp"

namespace ts {
double lsqr_selftest(ts_context *ctx, double *fast_err, char *msg, int len);
void grid_metric(const Grid &g, double *hlon, double *hlat, double *hr);
int point_source(const Grid &g, const double *slowness, double lon, double lat, double depth, double origin,
                 double radius_cells, int32_t *idx_out, double *t_out, int cap);

// Host Relax for the eikonal gate (SphericalEikonal.cs:284-325), one field, FP64.
static std::vector<double> host_relax(const Grid &g, const std::vector<double> &s, const std::vector<int32_t> &si,
                                      const std::vector<double> &st)
{
    const int nx = g.nx, ny = g.ny, nz = g.nz, nxy = nx * ny, count = g.count();
    std::vector<double> hlon((size_t)ny * nz), hlat(nz);
    double hr;
    grid_metric(g, hlon.data(), hlat.data(), &hr);
    std::vector<double> a(count, INFINITY);
    std::vector<uint8_t> fixed(count, 0);
    for (size_t q = 0; q < si.size(); q++) {
        a[si[q]] = std::min(a[si[q]], st[q]);
        fixed[si[q]] = 1;
    }
    const int cap = std::max(32, (nx + ny + nz) / 4);
    for (int phase = 0; phase < 2; phase++)
        for (int round = 0; round < cap; round++) {
            bool changed = false;
            for (int dir = 0; dir < 8; dir++)
                for (int p = 0; p < nx + ny + nz - 2; p++) {
                    int i0 = std::max(0, p - (ny - 1) - (nz - 1)), i1 = std::min(nx - 1, p);
                    int k0 = std::max(0, p - (nx - 1) - (ny - 1)), k1 = std::min(nz - 1, p);
                    for (int kp = k0; kp <= k1; kp++)
                        for (int ip = i0; ip <= i1; ip++) {
                            int jp = p - ip - kp;
                            if (jp < 0 || jp >= ny) continue;
                            int i = (dir & 1) ? nx - 1 - ip : ip, j = (dir & 2) ? ny - 1 - jp : jp;
                            int k = (dir & 4) ? nz - 1 - kp : kp;
                            int idx = (k * ny + j) * nx + i;
                            if (fixed[idx] || !(s[idx] > 0)) continue;
                            const int pos[3] = {i, j, k}, n[3] = {nx, ny, nz}, stv[3] = {1, nx, nxy};
                            const double h[3] = {hlon[k * ny + j], hlat[k], hr};
                            double first[3], t2[3];
                            for (int ax = 0; ax < 3; ax++) {
                                double mn = pos[ax] > 0 ? a[idx - stv[ax]] : INFINITY;
                                double pl = pos[ax] + 1 < n[ax] ? a[idx + stv[ax]] : INFINITY;
                                int sec;
                                if (mn <= pl) { first[ax] = mn; sec = pos[ax] > 1 ? idx - 2 * stv[ax] : -1; }
                                else { first[ax] = pl; sec = pos[ax] + 2 < n[ax] ? idx + 2 * stv[ax] : -1; }
                                t2[ax] = (phase == 1 && sec >= 0) ? a[sec] : INFINITY;
                            }
                            double cur = a[idx], cand = godunov_abs<double>(s[idx], first, t2, h);
                            if (!(cand < cur)) continue;
                            a[idx] = cand;
                            if (!std::isfinite(cur) || cur - cand > 1e-7) changed = true;
                        }
                }
            if (!changed) break;
        }
    return a;
}

// EikonalOpenCl.SelfTest (EikonalOpenCl.cs:346-374): 22 x 20 x 16 heterogeneous grid, one point
// source, max |device - host| over the field's range. The fast FP32 form is the one judged (it is
// the one the library runs by default); the plain FP64 form must agree too.
static double eikonal_selftest(ts_context *ctx, char *msg, int len)
{
    ts_grid gd{22, 20, 16, 10, 12, 42, 44, -2, 120};
    Grid g = make_grid(gd);
    std::vector<double> s(g.count());
    for (int k = 0; k < g.nz; k++)
        for (int j = 0; j < g.ny; j++)
            for (int i = 0; i < g.nx; i++)
                s[g.index(i, j, k)] = 1 / (5.8 + 0.02 * g.depth(k) + 0.2 * sin(0.7 * i) + 0.15 * cos(0.5 * j));
    std::vector<int32_t> si(4096);
    std::vector<double> st(4096);
    int ns = point_source(g, s.data(), 10.37, 42.61, 8.3, 0, 3, si.data(), st.data(), 4096);
    si.resize(ns);
    st.resize(ns);
    std::vector<double> host = host_relax(g, s, si, st);
    double worst = 0, lo = INFINITY, hi = -INFINITY;
    for (int mode : {TS_EIK_FAST_FP32, TS_EIK_PLAIN_FP64}) {
        std::vector<double> dev(g.count());
        int64_t offs[2] = {0, ns};
        char reason[256];
        int rc = ts_eik_solve(ctx, &gd, s.data(), 1, offs, si.data(), st.data(), dev.data(), mode, nullptr, nullptr,
                              reason, sizeof reason);
        if (rc != TS_OK) {
            snprintf(msg, len, "eikonal mode %d: %s", mode, reason);
            return INFINITY;
        }
        for (int q = 0; q < g.count(); q++) {
            if (!std::isfinite(host[q])) continue;
            if (!std::isfinite(dev[q])) {
                snprintf(msg, len, "device left a node unreached");
                return INFINITY;
            }
            lo = std::min(lo, host[q]);
            hi = std::max(hi, host[q]);
            worst = std::max(worst, fabs(dev[q] - host[q]));
        }
    }
    double rel = worst / std::max(1e-12, hi - lo);
    snprintf(msg, len, "max |dt| %.2e s over a %.2f s field (%.2e relative)", worst, hi - lo, rel);
    return rel;
}

}  // namespace ts

using namespace ts;

extern "C" {

int ts_hip_create(int device, int flags, ts_context **out, char *reason, int reason_len)
{
    if (!out) return TS_EINVAL;
    *out = nullptr;
    return guarded(reason, reason_len, [&]() -> int {
        int n = 0;
        if (hipGetDeviceCount(&n) != hipSuccess || n <= device || device < 0) {
            (void)hipGetLastError();
            fail(TS_DECLINED, "no HIP device %d (%d present)", device, n);
        }
        std::unique_ptr<ts_context> ctx(new ts_context);
        ctx->device = device;
        TS_HIP(hipSetDevice(device));
        hipDeviceProp_t prop;
        TS_HIP(hipGetDeviceProperties(&prop, device));
        ctx->name = std::string(prop.name) + " (" + prop.gcnArchName + ")";
        ctx->global_mem = prop.totalGlobalMem;
        ctx->cu_count = prop.multiProcessorCount;
        TS_HIP(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
        const char *ng = getenv("TOMOSTAR_NO_GRAPHS");
        ctx->use_graphs = !(ng && ng[0] == '1');
        if (!(flags & TS_CREATE_SKIP_SELFTEST)) {
            char msg[256];
            ctx->eik_selftest = eikonal_selftest(ctx.get(), msg, sizeof msg);
            if (!(ctx->eik_selftest <= 1e-6)) {
                (void)hipStreamDestroy(ctx->stream);
                fail(TS_DECLINED, "eikonal self-test failed (%s)", msg);
            }
            ctx->lsqr_selftest = lsqr_selftest(ctx.get(), &ctx->lsqr_fast_selftest, msg, sizeof msg);
            if (!(ctx->lsqr_selftest <= 1e-8) || !(ctx->lsqr_fast_selftest <= 1e-5)) {
                (void)hipStreamDestroy(ctx->stream);
                fail(TS_DECLINED, "LSQR self-test failed (%s)", msg);
            }
        }
        *out = ctx.release();
        return TS_OK;
    });
}

void ts_hip_destroy(ts_context *ctx)
{
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

int ts_hip_device_info(ts_context *ctx, char *buf, int len)
{
    if (!ctx || !buf || len <= 0) return TS_EINVAL;
    snprintf(buf, len, "%s, %d CUs, %.1f GB; self-test eikonal %.2e, LSQR %.2e (FP32 form %.2e)", ctx->name.c_str(),
             ctx->cu_count, ctx->global_mem / 1e9, ctx->eik_selftest, ctx->lsqr_selftest, ctx->lsqr_fast_selftest);
    return TS_OK;
}

int ts_hip_selftest_lsqr_fast(ts_context *ctx, double *err)
{
    if (!ctx || !err) return TS_EINVAL;
    *err = ctx->lsqr_fast_selftest;
    return TS_OK;
}

int ts_hip_selftest(ts_context *ctx, double *eikonal_err, double *lsqr_err)
{
    if (!ctx) return TS_EINVAL;
    if (eikonal_err) *eikonal_err = ctx->eik_selftest;
    if (lsqr_err) *lsqr_err = ctx->lsqr_selftest;
    return TS_OK;
}

}  // extern "C"
