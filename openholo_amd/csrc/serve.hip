// Serving engine: continuous batching of point-cloud hologram jobs.
//
// Jobs arrive at any time (ohx_server_submit) and are admitted in arrival
// order while a slot is free and their device memory fits the budget.  An
// admitted job uploads its points once, runs the per-point prep kernel and
// keeps its point constants and FP64 field resident until it finishes.
//
// The scheduler thread runs iterations.  One iteration advances every active
// job by one slice of its points:
//   * all direct-sum jobs (RS, Fresnel FP32/FP64) together in ONE persistent
//     kernel, k_pc_direct_jobs, whose workgroups stride over the (job, tile)
//     items of a job table -- the table is copied from pinned host memory and
//     the copy + kernel are captured once as a hipGraph and replayed;
//   * matrix-core Fresnel jobs launch their GEMM slice on the same stream.
// New arrivals join at the next iteration; finished jobs are phase-encoded,
// normalised to 8 bits and copied out, and their memory is released.
//
// Slice boundaries are a function of the job alone (its point count,
// resolution and mode), and each (job, tile) item is owned by one workgroup,
// so a job's field is bit-identical whether it ran alone or in any mix.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <thread>
#include <vector>

#include "ohx_internal.h"

namespace ohx {

constexpr int MAX_TABLE = 64;          // direct jobs per iteration
constexpr int PERSISTENT_GRID = 512;   // 2 workgroups per CU

struct SJob {
    uint64_t id = 0;
    std::vector<unsigned char> vtx;    // host copy of the vertices until upload
    int64_t n = 0;
    ohx_pc_params params{};
    uint32_t diff = 0, mode = 0;
    Geometry g{};
    bool mfma = false, fp64 = false;
    double* out_field = nullptr;
    uint8_t* out_u8 = nullptr;
    // device
    void* block = nullptr;             // SoA + field + enc + u8 + GEMM scratch
    PointSoA soa{};
    double* field = nullptr;
    double* enc = nullptr;
    uint8_t* u8 = nullptr;
    void* gemm = nullptr;
    size_t gemm_bytes = 0;
    size_t bytes = 0;
    int64_t next = 0, slice = 0;
    // state
    bool done = false, cancel = false;
    ohx_job_stats st{};
};

}  // namespace ohx

struct ohx_server {
    ohx_ctx* ctx = nullptr;
    ohx_server_opts opts{};
    std::thread worker;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<ohx::SJob*> pending;
    std::vector<ohx::SJob*> active;
    std::map<uint64_t, std::unique_ptr<ohx::SJob>> jobs;
    uint64_t next_id = 1;
    bool stop = false;
    size_t reserved = 0, peak = 0, budget = 0;
    uint64_t jobs_done = 0, iterations = 0, graph_replays = 0;
    std::chrono::steady_clock::time_point t0;
    // job table
    ohx::DirectJob* h_table = nullptr;
    int* h_njobs = nullptr;
    ohx::DirectJob* d_table = nullptr;
    int* d_njobs = nullptr;
    hipGraphExec_t graph = nullptr;
    // staging buffer for vertex uploads
    void* d_vtx = nullptr;
    size_t d_vtx_bytes = 0;

    double now_ms() const
    {
        return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
};

namespace ohx {

static size_t job_bytes(const SJob& j)
{
    const int64_t npx = (int64_t)j.g.pn_x * j.g.pn_y;
    size_t b = points_bytes(j.n) + 256;
    b += (size_t)npx * (16 + 8 + 1) + 3 * 256;
    if (j.mfma) b += fresnel_gemm_scratch_bytes(j.g, j.slice) + 256;
    return b;
}

static int64_t slice_points(const SJob& j, const ohx_server_opts& o)
{
    const double npx = (double)j.g.pn_x * j.g.pn_y;
    const double pairs = j.mfma ? o.slice_pairs_mfma : o.slice_pairs;
    int64_t s = (int64_t)(pairs / npx);
    s = std::max<int64_t>(s, 256);
    s = (s + DIRECT_CHUNK - 1) / DIRECT_CHUNK * DIRECT_CHUNK;   // whole LDS chunks of the direct kernel
    return std::min<int64_t>(s, std::max<int64_t>(j.n, 1));
}

// allocate, upload and prepare one admitted job; runs on the worker thread
static int setup_job(ohx_server* sv, SJob& j)
{
    ohx_ctx* c = sv->ctx;
    const int64_t npx = (int64_t)j.g.pn_x * j.g.pn_y;
    int st = check_hip(c, hipMalloc(&j.block, j.bytes), "hipMalloc(job)");
    if (st) return st;
    char* b = (char*)j.block;
    auto take = [&](size_t bytes) {
        char* r = b;
        b += (bytes + 255) & ~size_t(255);
        return (void*)r;
    };
    void* soa_base = take(points_bytes(j.n));
    carve_points(soa_base, j.n, &j.soa);
    j.field = (double*)take((size_t)npx * 16);
    j.enc = (double*)take((size_t)npx * 8);
    j.u8 = (uint8_t*)take((size_t)npx);
    if (j.mfma) {
        j.gemm_bytes = fresnel_gemm_scratch_bytes(j.g, j.slice);
        j.gemm = take(j.gemm_bytes);
    }
    const size_t vb = (size_t)j.n * 56;
    if (vb > sv->d_vtx_bytes) {
        if (sv->d_vtx) {
            (void)hipStreamSynchronize(c->stream);
            (void)hipFree(sv->d_vtx);
        }
        sv->d_vtx = nullptr;
        sv->d_vtx_bytes = 0;
        if ((st = check_hip(c, hipMalloc(&sv->d_vtx, vb), "hipMalloc(staging)"))) return st;
        sv->d_vtx_bytes = vb;
    }
    if (j.n > 0) {
        if ((st = check_hip(c, hipMemcpyAsync(sv->d_vtx, j.vtx.data(), vb, hipMemcpyHostToDevice, c->stream),
                            "hipMemcpyAsync(vertices)")))
            return st;
        if ((st = launch_prep(c, sv->d_vtx, j.g, j.soa))) return st;
    }
    if ((st = ohx_field_zero(c, j.field, npx))) return st;
    // the staging buffer is reused by the next admission: finish the upload
    st = check_hip(c, hipStreamSynchronize(c->stream), "hipStreamSynchronize(admission)");
    std::vector<unsigned char>().swap(j.vtx);
    return st;
}

static void release_job(ohx_server* sv, SJob& j)
{
    if (j.block) (void)hipFree(j.block);
    j.block = nullptr;
    sv->reserved -= j.bytes;
}

static int finish_job(ohx_server* sv, SJob& j)
{
    ohx_ctx* c = sv->ctx;
    const int64_t npx = (int64_t)j.g.pn_x * j.g.pn_y;
    int st = OHX_OK;
    if (j.out_u8) {
        const double* f = j.field;
        double* e = j.enc;
        uint8_t* o = j.u8;
        st = ohx_encode_normalize_u8(c, &f, 1, npx, OHX_ENCODE_PHASE, &e, &o);
        if (!st) st = check_hip(c, hipMemcpyAsync(j.out_u8, j.u8, npx, hipMemcpyDeviceToHost, c->stream), "d2h u8");
    }
    if (!st && j.out_field)
        st = check_hip(c, hipMemcpyAsync(j.out_field, j.field, npx * 16, hipMemcpyDeviceToHost, c->stream),
                       "d2h field");
    if (!st) st = check_hip(c, hipStreamSynchronize(c->stream), "hipStreamSynchronize(finish)");
    return st;
}

static int build_graph(ohx_server* sv)
{
    ohx_ctx* c = sv->ctx;
    hipGraph_t g = nullptr;
    int st = check_hip(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal), "begin capture");
    if (st) return st;
    (void)hipMemcpyAsync(sv->d_njobs, sv->h_njobs, sizeof(int), hipMemcpyHostToDevice, c->stream);
    (void)hipMemcpyAsync(sv->d_table, sv->h_table, sizeof(DirectJob) * MAX_TABLE, hipMemcpyHostToDevice, c->stream);
    launch_direct_jobs(c->stream, sv->d_table, sv->d_njobs, PERSISTENT_GRID);
    st = check_hip(c, hipStreamEndCapture(c->stream, &g), "end capture");
    if (st) return st;
    st = check_hip(c, hipGraphInstantiate(&sv->graph, g, nullptr, nullptr, 0), "graph instantiate");
    (void)hipGraphDestroy(g);
    return st;
}

static void worker_loop(ohx_server* sv)
{
    ohx_ctx* c = sv->ctx;
    (void)hipSetDevice(c->device);
    for (;;) {
        std::vector<SJob*> admit;
        {
            std::unique_lock<std::mutex> lk(sv->mu);
            sv->cv_work.wait(lk, [&] { return sv->stop || !sv->pending.empty() || !sv->active.empty(); });
            if (sv->stop) {
                for (SJob* j : sv->pending) {
                    j->st.status = OHX_ERR_CANCELLED;
                    j->done = true;
                }
                sv->pending.clear();
                for (SJob* j : sv->active) {
                    release_job(sv, *j);
                    j->st.status = OHX_ERR_CANCELLED;
                    j->done = true;
                }
                sv->active.clear();
                sv->cv_done.notify_all();
                return;
            }
            // admission in arrival order: slots and memory budget
            while (!sv->pending.empty() &&
                   (int)(sv->active.size() + admit.size()) < sv->opts.max_active) {
                SJob* j = sv->pending.front();
                if (sv->reserved + j->bytes > sv->budget) break;
                sv->pending.pop_front();
                sv->reserved += j->bytes;
                sv->peak = std::max(sv->peak, sv->reserved);
                admit.push_back(j);
            }
        }
        for (SJob* j : admit) {
            const int st = setup_job(sv, *j);
            std::lock_guard<std::mutex> lk(sv->mu);
            j->st.admit_ms = sv->now_ms();
            if (st) {
                release_job(sv, *j);
                j->st.status = st;
                j->done = true;
                sv->cv_done.notify_all();
            } else {
                sv->active.push_back(j);
            }
        }
        std::vector<SJob*> act;
        {
            std::lock_guard<std::mutex> lk(sv->mu);
            // drop cancelled jobs
            for (auto it = sv->active.begin(); it != sv->active.end();) {
                if ((*it)->cancel) {
                    release_job(sv, **it);
                    (*it)->st.status = OHX_ERR_CANCELLED;
                    (*it)->done = true;
                    it = sv->active.erase(it);
                    sv->cv_done.notify_all();
                } else {
                    ++it;
                }
            }
            act = sv->active;
        }
        if (act.empty()) continue;

        // ---- one iteration --------------------------------------------------
        const double t_iter = sv->now_ms();
        int nd = 0, items = 0;
        std::vector<SJob*> gemm_jobs, in_iter;
        for (SJob* j : act) {
            const int64_t p0 = j->next, p1 = std::min(j->n, p0 + j->slice);
            if (j->mfma) {
                gemm_jobs.push_back(j);
            } else {
                if (nd == MAX_TABLE) continue;   // waits for the next iteration
                DirectJob& d = sv->h_table[nd++];
                d.g = j->g;
                d.s = j->soa;
                d.field = (double2*)j->field;
                d.p0 = p0;
                d.p1 = p1;
                d.tiles = direct_tiles(j->g, j->fp64, &d.tiles_x);
                d.item_base = items;
                d.fp64 = j->fp64 ? 1 : 0;
                items += d.tiles;
            }
            in_iter.push_back(j);
            if (j->st.iterations == 0) j->st.start_ms = t_iter;
        }
        *sv->h_njobs = nd;
        int st = OHX_OK;
        if (nd > 0) {
            if (sv->graph) {
                st = check_hip(c, hipGraphLaunch(sv->graph, c->stream), "hipGraphLaunch");
                ++sv->graph_replays;
            } else {
                st = check_hip(c, hipMemcpyAsync(sv->d_njobs, sv->h_njobs, sizeof(int), hipMemcpyHostToDevice,
                                                 c->stream), "h2d table");
                if (!st) st = check_hip(c, hipMemcpyAsync(sv->d_table, sv->h_table, sizeof(DirectJob) * nd,
                                                          hipMemcpyHostToDevice, c->stream), "h2d table");
                if (!st) st = launch_direct_jobs(c->stream, sv->d_table, sv->d_njobs, PERSISTENT_GRID);
            }
        }
        for (SJob* j : gemm_jobs) {
            if (st) break;
            const int64_t p0 = j->next, p1 = std::min(j->n, p0 + j->slice);
            st = launch_fresnel_gemm_range(c->stream, j->gemm, j->gemm_bytes, j->g, j->soa, p0, p1, j->field, j->mode);
        }
        if (!st) st = check_hip(c, hipStreamSynchronize(c->stream), "hipStreamSynchronize(iteration)");

        std::vector<SJob*> finished;
        for (SJob* j : in_iter) {
            j->next = std::min(j->n, j->next + j->slice);
            ++j->st.iterations;
            if (st || j->next >= j->n) finished.push_back(j);
        }
        for (SJob* j : finished) {
            const int fs = st ? st : finish_job(sv, *j);
            std::lock_guard<std::mutex> lk(sv->mu);
            release_job(sv, *j);
            j->st.status = fs;
            j->st.end_ms = sv->now_ms();
            j->done = true;
            ++sv->jobs_done;
            sv->active.erase(std::find(sv->active.begin(), sv->active.end(), j));
            sv->cv_done.notify_all();
        }
        std::lock_guard<std::mutex> lk(sv->mu);
        ++sv->iterations;
    }
}

}  // namespace ohx

using namespace ohx;

extern "C" {

int ohx_server_create(const ohx_server_opts* opts, ohx_server** out)
{
    if (!out) return OHX_ERR_INVALID_ARG;
    *out = nullptr;
    ohx_server_opts o = opts ? *opts : ohx_server_opts{};
    if (o.max_active <= 0) o.max_active = 32;
    if (o.slice_pairs <= 0) o.slice_pairs = 1.5e10;
    if (o.slice_pairs_mfma <= 0) o.slice_pairs_mfma = 1.5e11;
    auto sv = std::make_unique<ohx_server>();
    int st = ohx_ctx_create(o.device, &sv->ctx);
    if (st) return st;
    ohx_ctx* c = sv->ctx;
    size_t fr = 0, tot = 0;
    if (hipMemGetInfo(&fr, &tot) != hipSuccess) {
        ohx_ctx_destroy(c);
        return OHX_ERR_DEVICE;
    }
    sv->budget = o.mem_budget ? (size_t)o.mem_budget : (fr > (4ull << 30) ? fr - (4ull << 30) : fr / 2);
    sv->opts = o;
    if (hipHostMalloc((void**)&sv->h_table, sizeof(DirectJob) * MAX_TABLE, hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc((void**)&sv->h_njobs, sizeof(int), hipHostMallocDefault) != hipSuccess ||
        hipMalloc((void**)&sv->d_table, sizeof(DirectJob) * MAX_TABLE) != hipSuccess ||
        hipMalloc((void**)&sv->d_njobs, sizeof(int)) != hipSuccess) {
        ohx_server_destroy(sv.release());
        return OHX_ERR_OUT_OF_MEMORY;
    }
    std::memset(sv->h_table, 0, sizeof(DirectJob) * MAX_TABLE);
    *sv->h_njobs = 0;
    if (o.use_graph && (st = build_graph(sv.get()))) {
        ohx_server_destroy(sv.release());
        return st;
    }
    sv->t0 = std::chrono::steady_clock::now();
    ohx_server* raw = sv.release();
    raw->worker = std::thread(worker_loop, raw);
    *out = raw;
    return OHX_OK;
}

int ohx_server_submit(ohx_server* sv, const ohx_job* job, uint64_t* id)
{
    if (!sv || !job || !id || job->n_points < 0 || (job->n_points > 0 && !job->vertices)) return OHX_ERR_INVALID_ARG;
    auto j = std::make_unique<SJob>();
    int st = geometry_from_params(sv->ctx, &job->params, job->diff_flag, &j->g);
    if (st) return st;
    j->n = job->n_points;
    j->params = job->params;
    j->diff = job->diff_flag;
    j->mode = job->mode;
    const uint32_t mf = job->mode & (OHX_MODE_MFMA_BF16X1 | OHX_MODE_MFMA_BF16X3 | OHX_MODE_MFMA_FP16X3);
    j->mfma = mf && job->diff_flag == OHX_DIFF_FRESNEL;
    j->fp64 = !j->mfma && (job->mode & (OHX_MODE_FLOAT | OHX_MODE_FASTMATH)) == 0 && mf == 0;
    j->out_field = job->out_field;
    j->out_u8 = job->out_u8;
    j->slice = slice_points(*j, sv->opts);
    j->bytes = job_bytes(*j);
    j->st.device_bytes = j->bytes;
    if (j->bytes > sv->budget) return OHX_ERR_ADMISSION;
    j->vtx.assign((const unsigned char*)job->vertices, (const unsigned char*)job->vertices + 56 * (size_t)j->n);
    std::lock_guard<std::mutex> lk(sv->mu);
    j->id = sv->next_id++;
    j->st.submit_ms = sv->now_ms();
    *id = j->id;
    sv->pending.push_back(j.get());
    sv->jobs.emplace(j->id, std::move(j));
    sv->cv_work.notify_one();
    return OHX_OK;
}

int ohx_server_wait(ohx_server* sv, uint64_t id, double timeout_s, ohx_job_stats* stats)
{
    if (!sv) return OHX_ERR_INVALID_ARG;
    std::unique_lock<std::mutex> lk(sv->mu);
    auto it = sv->jobs.find(id);
    if (it == sv->jobs.end()) return OHX_ERR_NOT_FOUND;
    SJob* j = it->second.get();
    auto ready = [&] { return j->done; };
    if (timeout_s < 0) sv->cv_done.wait(lk, ready);
    else if (!sv->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), ready)) return OHX_ERR_TIMEOUT;
    if (stats) *stats = j->st;
    const int st = j->st.status;
    sv->jobs.erase(it);
    return st;
}

int ohx_server_cancel(ohx_server* sv, uint64_t id)
{
    if (!sv) return OHX_ERR_INVALID_ARG;
    std::lock_guard<std::mutex> lk(sv->mu);
    auto it = sv->jobs.find(id);
    if (it == sv->jobs.end()) return OHX_ERR_NOT_FOUND;
    SJob* j = it->second.get();
    auto p = std::find(sv->pending.begin(), sv->pending.end(), j);
    if (p != sv->pending.end()) {
        sv->pending.erase(p);
        j->st.status = OHX_ERR_CANCELLED;
        j->done = true;
        sv->cv_done.notify_all();
    } else if (!j->done) {
        j->cancel = true;
        sv->cv_work.notify_one();
    }
    return OHX_OK;
}

int ohx_server_counters_get(ohx_server* sv, ohx_server_counters* out)
{
    if (!sv || !out) return OHX_ERR_INVALID_ARG;
    std::lock_guard<std::mutex> lk(sv->mu);
    out->jobs_done = sv->jobs_done;
    out->iterations = sv->iterations;
    out->graph_replays = sv->graph_replays;
    out->reserved_bytes = sv->reserved;
    out->peak_reserved_bytes = sv->peak;
    out->budget_bytes = sv->budget;
    return OHX_OK;
}

int ohx_server_destroy(ohx_server* sv)
{
    if (!sv) return OHX_ERR_INVALID_ARG;
    {
        std::lock_guard<std::mutex> lk(sv->mu);
        sv->stop = true;
        sv->cv_work.notify_all();
    }
    if (sv->worker.joinable()) sv->worker.join();
    if (sv->ctx) {
        (void)hipSetDevice(sv->ctx->device);
        (void)hipStreamSynchronize(sv->ctx->stream);
        if (sv->graph) (void)hipGraphExecDestroy(sv->graph);
        if (sv->d_vtx) (void)hipFree(sv->d_vtx);
        if (sv->d_table) (void)hipFree(sv->d_table);
        if (sv->d_njobs) (void)hipFree(sv->d_njobs);
        if (sv->h_table) (void)hipHostFree(sv->h_table);
        if (sv->h_njobs) (void)hipHostFree(sv->h_njobs);
        ohx_ctx_destroy(sv->ctx);
    }
    delete sv;
    return OHX_OK;
}

}  // extern "C"
