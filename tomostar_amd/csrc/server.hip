// A serving loop on one GPU: travel-time-table and LSQR requests against resident velocity models and
// matrices, admitted by a device-memory budget and run concurrently.
//
// The reference answers one request at a time per device (EikonalOpenCl and LsqrOpenCl serialise
// their solves with lock(_gate), Forward/EikonalOpenCl.cs:64,165; Numerics/LsqrOpenCl.cs:184,257).
// Here W worker threads each own a stream; a request waits in a FIFO until its estimated device
// memory fits the budget left by the requests in flight (the head of the queue is never overtaken, so
// a large request is not starved), then runs on a free worker's stream while the others run theirs.
// A table request's tables stay resident (and counted) until the client releases the request.
//
// Results do not depend on what else runs: every kernel of the path reduces in a fixed order (block
// partials summed in index order, per-field flags only OR-ed), so a request gives the same bits alone
// and in the mix.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <memory>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "common.hpp"
#include "geometry.hpp"
#include "matrix.hpp"

namespace ts {
int point_source(const Grid &g, const double *slowness, double lon, double lat, double depth, double origin,
                 double radius_cells, int32_t *idx_out, double *t_out, int cap);
int tables_create_core(ts_context *ctx, const ts_grid *grid, int nmodels, const double *const *slowness,
                       const double *const *dev64, const float *const *dev32, int nfields,
                       const int32_t *model_of_field, const int64_t *seed_offsets, const int32_t *seed_index,
                       const double *seed_time, ts_tableset **out, char *reason, int reason_len);
}

using namespace ts;

struct SrvModel {
    ts_grid grid{};
    int N = 0;
    std::vector<double> host;
    dbuf<double> d64;
    dbuf<float> d32;
};

struct SrvRequest {
    int64_t id = 0;
    int kind = 0;  // 0 tables, 1 LSQR
    // tables
    ts_grid grid{};
    int models[2] = {-1, -1};
    const SrvModel *model_ptr[2] = {nullptr, nullptr};  // taken under the server's lock at submission
    std::vector<int32_t> model_of_field, seed_idx;
    std::vector<int64_t> seed_off;
    std::vector<double> seed_time;
    int nfields = 0;
    ts_tableset *tables = nullptr;
    // LSQR
    ts_matrix *matrix = nullptr;
    int k = 0, max_it = 0;
    double atol = 0, btol = 0, conlim = 0;
    std::vector<double> b, damp, x, stats;
    std::vector<int32_t> its, stop;
    int mode = 0;
    // bookkeeping
    uint64_t mem = 0;
    bool holds_mem = false;
    double t_submit = 0, t_start = 0, t_end = 0;
    int status = 0;
    bool done = false;
    std::string reason;
};

struct ts_server {
    ts_context *ctx = nullptr;
    std::vector<ts_context *> workers;
    std::vector<std::thread> threads;
    uint64_t budget = 0, in_use = 0, peak = 0;
    int inflight = 0, peak_inflight = 0;
    std::deque<SrvRequest *> queue;
    std::unordered_map<int64_t, std::unique_ptr<SrvRequest>> reqs;
    std::vector<std::unique_ptr<SrvModel>> models;
    std::mutex m;
    std::condition_variable cv_work, cv_done;
    bool stopping = false;
    int64_t next_id = 1;
    std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
    double now() const { return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(); }
};

namespace {

// Device memory a table request holds: FP32 and FP64 times of its 64-field groups, the dirty and
// seed-mask words, the float32 and FP64 tables, the slowness copies, the plane list and the metric.
uint64_t tables_bytes(int64_t N, int nfields, int nmodels)
{
    const int64_t groups = (nfields + 63) / 64;
    return (uint64_t)(groups * 64 * N * 12 + groups * N * 9 + (int64_t)nfields * N * 12 + nmodels * N * 12 +
                      N * 4 + (64ll << 20));
}

// Device memory an LSQR request adds to its resident matrix: the right-hand sides and the vectors of
// the fast form (FP64 for the mixed form) or of the plain form.
uint64_t lsqr_bytes(const ts_matrix *a, int k, int mode)
{
    int kp = 1;
    while (kp < k) kp <<= 1;
    const int64_t m = a->rows, n = a->cols;
    const int64_t width = mode == TS_LSQR_FAST_FP32 ? 4 : 8;
    if (mode == TS_LSQR_PLAIN_FP64) return (uint64_t)((2 * m + 5 * n) * 8 + (32ll << 20));
    return (uint64_t)((int64_t)k * m * 8 + kp * (m + 2 * n) * width + kp * n * 8 + (int64_t)n * k * 8 + (32ll << 20));
}

void run_request(ts_server &S, ts_context *w, SrvRequest &r)
{
    char why[256] = {0};
    if (r.kind == 0) {
        const SrvModel *const *mp = r.model_ptr;
        const int nm = mp[1] ? 2 : 1;
        const double *d64[2] = {mp[0]->d64.p, mp[1] ? mp[1]->d64.p : nullptr};
        const float *d32[2] = {mp[0]->d32.p, mp[1] ? mp[1]->d32.p : nullptr};
        int rc = tables_create_core(w, &r.grid, nm, nullptr, d64, d32, r.nfields, r.model_of_field.data(),
                                    r.seed_off.data(), r.seed_idx.data(), r.seed_time.data(), &r.tables, why,
                                    sizeof why);
        if (rc == TS_OK) rc = ts_tables_solve(r.tables, r.mode, nullptr, nullptr, why, sizeof why);
        r.status = rc;
    } else {
        stream_override() = w->stream;
        double secs = 0;
        r.status = lsqr_solve_core(r.matrix, r.k, r.b.data(), r.damp.data(), r.max_it, r.atol, r.btol, r.conlim,
                                   r.mode, nullptr, r.x.data(), r.its.data(), r.stop.data(), r.stats.data(), nullptr,
                                   &secs, why, sizeof why);
        stream_override() = nullptr;
        std::vector<double>().swap(r.b);
    }
    r.reason = why;
}

void worker_loop(ts_server *S, ts_context *w)
{
    (void)hipSetDevice(w->device);
    for (;;) {
        SrvRequest *r = nullptr;
        {
            std::unique_lock<std::mutex> lk(S->m);
            S->cv_work.wait(lk, [&] {
                return S->stopping || (!S->queue.empty() && S->queue.front()->mem <= S->budget - S->in_use);
            });
            if (S->stopping) return;
            r = S->queue.front();
            S->queue.pop_front();
            S->in_use += r->mem;
            r->holds_mem = true;
            S->peak = std::max(S->peak, S->in_use);
            S->inflight++;
            S->peak_inflight = std::max(S->peak_inflight, S->inflight);
            r->t_start = S->now();
        }
        run_request(*S, w, *r);
        {
            std::lock_guard<std::mutex> lk(S->m);
            r->t_end = S->now();
            S->inflight--;
            // LSQR buffers are gone with the solve; tables stay until the request is released
            if (r->kind == 1 || r->status != TS_OK) {
                S->in_use -= r->mem;
                r->holds_mem = false;
            }
            r->done = true;
        }
        S->cv_done.notify_all();
        S->cv_work.notify_all();
    }
}

int64_t enqueue(ts_server *S, std::unique_ptr<SrvRequest> r)
{
    std::lock_guard<std::mutex> lk(S->m);
    r->id = S->next_id++;
    r->t_submit = S->now();
    SrvRequest *p = r.get();
    S->reqs[p->id] = std::move(r);
    S->queue.push_back(p);
    S->cv_work.notify_all();
    return p->id;
}

}  // namespace

// ================================================================================== C ABI

extern "C" {

int ts_server_create(ts_context *ctx, int workers, uint64_t memory_budget, ts_server **out, char *reason,
                     int reason_len)
{
    if (out) *out = nullptr;
    if (!ctx || workers < 1 || workers > 16 || memory_budget == 0 || !out) {
        set_reason(reason, reason_len, "invalid argument");
        return TS_EINVAL;
    }
    return guarded(reason, reason_len, [&]() -> int {
        std::unique_ptr<ts_server> S(new ts_server);
        S->ctx = ctx;
        S->budget = std::min<uint64_t>(memory_budget, (uint64_t)(0.9 * ctx->global_mem));
        for (int i = 0; i < workers; i++) {
            ts_context *w = nullptr;
            char why[256];
            const int rc = ts_hip_create(ctx->device, TS_CREATE_SKIP_SELFTEST, &w, why, sizeof why);
            if (rc != TS_OK) {
                for (ts_context *x : S->workers) ts_hip_destroy(x);
                fail(rc, "worker context: %s", why);
            }
            w->use_graphs = false;  // direct launches: graph captures on several threads at once are not reliable
            S->workers.push_back(w);
        }
        for (ts_context *w : S->workers) S->threads.emplace_back(worker_loop, S.get(), w);
        *out = S.release();
        return TS_OK;
    });
}

int ts_server_add_model(ts_server *S, const ts_grid *grid, const double *slowness, int32_t *model_id, char *reason,
                        int reason_len)
{
    if (!S || !grid || !slowness || !model_id || !grid_valid(*grid)) {
        set_reason(reason, reason_len, "invalid argument");
        return TS_EINVAL;
    }
    return guarded(reason, reason_len, [&]() -> int {
        TS_HIP(hipSetDevice(S->ctx->device));
        std::unique_ptr<SrvModel> M(new SrvModel);
        M->grid = *grid;
        M->N = grid->nx * grid->ny * grid->nz;
        M->host.assign(slowness, slowness + M->N);
        M->d64.alloc(M->N);
        M->d32.alloc(M->N);
        std::vector<float> f(M->host.begin(), M->host.end());
        TS_HIP(hipMemcpy(M->d64.p, slowness, (size_t)M->N * 8, hipMemcpyHostToDevice));
        TS_HIP(hipMemcpy(M->d32.p, f.data(), (size_t)M->N * 4, hipMemcpyHostToDevice));
        std::lock_guard<std::mutex> lk(S->m);
        *model_id = (int32_t)S->models.size();
        S->models.push_back(std::move(M));
        return TS_OK;
    });
}

int ts_server_submit_tables(ts_server *S, int32_t model_p, int32_t model_s, int nsources, const double *sources,
                            int mode, int64_t *request_id, char *reason, int reason_len)
{
    if (!S || nsources < 1 || !sources || !request_id || (mode != TS_EIK_FAST_FP32 && mode != TS_EIK_FAST_FP64 &&
                                                           mode != TS_EIK_PLAIN_FP64)) {
        set_reason(reason, reason_len, "invalid argument");
        return TS_EINVAL;
    }
    return guarded(reason, reason_len, [&]() -> int {
        std::unique_ptr<SrvRequest> r(new SrvRequest);
        const int nm = model_s >= 0 ? 2 : 1;
        // the model pointers are taken under the lock (a concurrent add_model may reallocate the vector);
        // the models themselves stay where they are until the server is destroyed
        const SrvModel *mp = nullptr, *ms = nullptr;
        {
            std::lock_guard<std::mutex> lk(S->m);
            if (model_p < 0 || model_p >= (int)S->models.size() || model_s >= (int)S->models.size())
                fail(TS_EINVAL, "no such model");
            mp = S->models[model_p].get();
            ms = nm == 2 ? S->models[model_s].get() : nullptr;
        }
        if (ms && (ms->grid.nx != mp->grid.nx || ms->grid.ny != mp->grid.ny || ms->grid.nz != mp->grid.nz))
            fail(TS_EINVAL, "the P and S models lie on different grids");
        r->kind = 0;
        r->grid = mp->grid;
        r->models[0] = model_p;
        r->models[1] = nm == 2 ? model_s : -1;
        r->model_ptr[0] = mp;
        r->model_ptr[1] = ms;
        r->mode = mode;
        // one field per source and phase (SphericalEikonal.PointSource seeds, :68-100)
        const Grid g = make_grid(mp->grid);
        std::vector<int32_t> idx(4096);
        std::vector<double> tm(4096);
        r->seed_off.push_back(0);
        for (int s = 0; s < nsources; s++)
            for (int ph = 0; ph < nm; ph++) {
                const SrvModel *M = ph == 0 ? mp : ms;
                int c = point_source(g, M->host.data(), sources[3 * s], sources[3 * s + 1], sources[3 * s + 2], 0.0, 3.0,
                                     idx.data(), tm.data(), (int)idx.size());
                if (c > (int)idx.size()) {
                    idx.resize(c);
                    tm.resize(c);
                    c = point_source(g, M->host.data(), sources[3 * s], sources[3 * s + 1], sources[3 * s + 2], 0.0,
                                     3.0, idx.data(), tm.data(), (int)idx.size());
                }
                r->seed_idx.insert(r->seed_idx.end(), idx.begin(), idx.begin() + c);
                r->seed_time.insert(r->seed_time.end(), tm.begin(), tm.begin() + c);
                r->seed_off.push_back((int64_t)r->seed_idx.size());
                r->model_of_field.push_back(ph);
            }
        r->nfields = nsources * nm;
        r->mem = tables_bytes(mp->N, r->nfields, nm);
        if (r->mem > S->budget) fail(TS_DECLINED, "request needs %.2f GB, the budget is %.2f GB", r->mem / 1e9, S->budget / 1e9);
        *request_id = enqueue(S, std::move(r));
        return TS_OK;
    });
}

int ts_server_submit_lsqr(ts_server *S, ts_matrix *a, int k, const double *b, const double *damp, int max_iterations,
                          double atol, double btol, double conlim, int mode, int64_t *request_id, char *reason,
                          int reason_len)
{
    if (!S || !a || k < 1 || k > 64 || !b || !damp || max_iterations < 0 || !request_id ||
        (mode != TS_LSQR_PLAIN_FP64 && mode != TS_LSQR_FAST_FP32 && mode != TS_LSQR_FAST_MIXED &&
         mode != TS_LSQR_FAST_FP64)) {
        set_reason(reason, reason_len, "invalid argument");
        return TS_EINVAL;
    }
    return guarded(reason, reason_len, [&]() -> int {
        std::unique_ptr<SrvRequest> r(new SrvRequest);
        r->kind = 1;
        r->matrix = a;
        r->k = k;
        r->b.assign(b, b + (size_t)k * a->rows);
        r->damp.assign(damp, damp + k);
        r->max_it = max_iterations;
        r->atol = atol;
        r->btol = btol;
        r->conlim = conlim;
        r->mode = mode;
        r->x.assign((size_t)k * a->cols, 0.0);
        r->its.assign(k, 0);
        r->stop.assign(k, 0);
        r->stats.assign(3 * (size_t)k, 0.0);
        r->mem = lsqr_bytes(a, k, mode);
        if (r->mem > S->budget) fail(TS_DECLINED, "request needs %.2f GB, the budget is %.2f GB", r->mem / 1e9, S->budget / 1e9);
        *request_id = enqueue(S, std::move(r));
        return TS_OK;
    });
}

int ts_server_wait(ts_server *S, int64_t id, double timeout_s, int32_t *status, double times[3], char *reason,
                   int reason_len)
{
    if (!S) return TS_EINVAL;
    std::unique_lock<std::mutex> lk(S->m);
    auto it = S->reqs.find(id);
    if (it == S->reqs.end()) return TS_EINVAL;
    SrvRequest *r = it->second.get();
    if (!S->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), [&] { return r->done; })) return TS_DECLINED;
    if (status) *status = r->status;
    if (times) { times[0] = r->t_submit; times[1] = r->t_start; times[2] = r->t_end; }
    set_reason(reason, reason_len, r->reason.c_str());
    return TS_OK;
}

int ts_server_result_tables(ts_server *S, int64_t id, int first, int count, float *out)
{
    if (!S || !out) return TS_EINVAL;
    ts_tableset *t = nullptr;
    {
        std::lock_guard<std::mutex> lk(S->m);
        auto it = S->reqs.find(id);
        if (it == S->reqs.end() || !it->second->done || it->second->kind != 0 || it->second->status != TS_OK)
            return TS_EINVAL;
        t = it->second->tables;
    }
    return ts_tables_read_f32(t, first, count, out);
}

int ts_server_result_lsqr(ts_server *S, int64_t id, double *x, int32_t *iterations, int32_t *stop, double *stats)
{
    if (!S) return TS_EINVAL;
    std::lock_guard<std::mutex> lk(S->m);
    auto it = S->reqs.find(id);
    if (it == S->reqs.end() || !it->second->done || it->second->kind != 1 || it->second->status != TS_OK)
        return TS_EINVAL;
    const SrvRequest &r = *it->second;
    if (x) std::copy(r.x.begin(), r.x.end(), x);
    if (iterations) std::copy(r.its.begin(), r.its.end(), iterations);
    if (stop) std::copy(r.stop.begin(), r.stop.end(), stop);
    if (stats) std::copy(r.stats.begin(), r.stats.end(), stats);
    return TS_OK;
}

int ts_server_release_request(ts_server *S, int64_t id)
{
    if (!S) return TS_EINVAL;
    std::unique_ptr<SrvRequest> r;
    {
        std::lock_guard<std::mutex> lk(S->m);
        auto it = S->reqs.find(id);
        if (it == S->reqs.end() || !it->second->done) return TS_EINVAL;
        r = std::move(it->second);
        S->reqs.erase(it);
        if (r->holds_mem) S->in_use -= r->mem;
    }
    if (r->tables) ts_tables_release(r->tables);
    S->cv_work.notify_all();
    return TS_OK;
}

int ts_server_stats(ts_server *S, uint64_t *peak_bytes, int32_t *peak_inflight, uint64_t *budget)
{
    if (!S) return TS_EINVAL;
    std::lock_guard<std::mutex> lk(S->m);
    if (peak_bytes) *peak_bytes = S->peak;
    if (peak_inflight) *peak_inflight = S->peak_inflight;
    if (budget) *budget = S->budget;
    return TS_OK;
}

void ts_server_destroy(ts_server *S)
{
    if (!S) return;
    {
        std::lock_guard<std::mutex> lk(S->m);
        S->stopping = true;
    }
    S->cv_work.notify_all();
    for (std::thread &t : S->threads) t.join();
    for (auto &kv : S->reqs)
        if (kv.second->tables) ts_tables_release(kv.second->tables);
    S->reqs.clear();
    (void)hipSetDevice(S->ctx->device);
    S->models.clear();
    for (ts_context *w : S->workers) ts_hip_destroy(w);
    delete S;
}

}  // extern "C"
