// Serving engine: continuous batching of point-cloud hologram jobs.
//
// Jobs arrive at any time (ohx_server_submit) and are admitted in arrival
// order while a slot is free and their device memory fits the budget.  Device
// memory (and the page-locked host staging of the outputs) comes from pools
// of arenas that grow on demand and are kept until the server is destroyed,
// so a warm server allocates nothing per job.
//
// Three streams, all at the default priority:
//   * setup  : an admitted job's vertex upload, per-point prep and field zero
//              (RS with a BF16 matrix-core mode: also the partition of its
//              points into matrix-core and direct ones, whose count is copied
//              to page-locked memory), then an event its first slice waits for
//              (no host sync; an RS matrix-core job is scheduled once that
//              event has completed);
//   * compute: iterations.  One iteration advances a set of jobs by one slice
//              each, chosen shortest remaining work first up to a GPU-time
//              budget: the GEMM slices of the matrix-core jobs (Fresnel, and
//              RS over its matrix-core points before its direct ones), then
//              ONE persistent kernel, k_pc_direct_jobs, striding over the
//              (job, tile) items of a table of the direct-sum jobs (RS,
//              Fresnel FP32/FP64), copied from pinned host memory (copy +
//              kernel captured once per table as a hipGraph).  The direct
//              kernel and the GEMMs share one stream: run side by side, the
//              persistent kernel held the LDS of every CU and the GEMM slices
//              starved between its launches;
//   * finish : when a job's last slice has been enqueued, its phase encoding,
//              8-bit normalisation and copies to page-locked staging are
//              enqueued here behind an event of that slice.
// Up to two iterations are in flight (two job tables, two graphs), so the
// host builds iteration k+1 while iteration k runs; the worker polls events
// instead of synchronising streams.  The staged outputs are copied into the
// caller's buffers by ohx_server_wait.
//
// Slice boundaries are a function of the job alone (its point count,
// resolution and mode), a job's slices run in order on the compute stream,
// and each (job, tile) item is owned by one workgroup, so a job's field is
// bit-identical whether it ran alone or in any mix.
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
constexpr int DEPTH = 2;               // iterations in flight
constexpr size_t MIN_ARENA = size_t(1) << 30;
constexpr size_t MIN_HOST_ARENA = size_t(256) << 20;

enum JobState { J_PENDING, J_ACTIVE, J_FINISHING, J_CANCELLING, J_DONE };

// Grow-only pool of arenas (device or page-locked host memory); first fit
// with coalescing, 256-byte units.
struct Pool {
    bool host = false;
    struct Arena {
        char* base = nullptr;
        size_t size = 0;
        std::map<size_t, size_t> free;   // offset -> length
    };
    std::vector<Arena> arenas;
    size_t total = 0;

    // nullptr with *status == OHX_OK: no room under cap now (wait for releases)
    void* take(size_t bytes, size_t cap, size_t min_arena, int* status)
    {
        bytes = (bytes + 255) & ~size_t(255);
        *status = OHX_OK;
        for (Arena& a : arenas) {
            for (auto it = a.free.begin(); it != a.free.end(); ++it) {
                if (it->second < bytes) continue;
                const size_t off = it->first, len = it->second;
                a.free.erase(it);
                if (len > bytes) a.free.emplace(off + bytes, len - bytes);
                return a.base + off;
            }
        }
        if (total + bytes > cap) return nullptr;
        const size_t size = std::max(bytes, std::min(min_arena, cap - total));
        Arena a;
        const hipError_t e = host ? hipHostMalloc((void**)&a.base, size, hipHostMallocDefault)
                                  : hipMalloc((void**)&a.base, size);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            *status = OHX_ERR_OUT_OF_MEMORY;
            return nullptr;
        }
        a.size = size;
        if (size > bytes) a.free.emplace(bytes, size - bytes);
        total += size;
        arenas.push_back(std::move(a));
        return arenas.back().base;
    }

    void give(void* p, size_t bytes)
    {
        bytes = (bytes + 255) & ~size_t(255);
        for (Arena& a : arenas) {
            if ((char*)p < a.base || (char*)p >= a.base + a.size) continue;
            size_t off = (size_t)((char*)p - a.base), len = bytes;
            auto nx = a.free.lower_bound(off);
            if (nx != a.free.end() && off + len == nx->first) {
                len += nx->second;
                nx = a.free.erase(nx);
            }
            if (nx != a.free.begin()) {
                auto pv = std::prev(nx);
                if (pv->first + pv->second == off) {
                    off = pv->first;
                    len += pv->second;
                    a.free.erase(pv);
                }
            }
            a.free.emplace(off, len);
            return;
        }
    }

    // free the arenas that hold nothing (called when no job is admitted and a
    // request does not fit the existing arenas)
    void trim()
    {
        for (auto it = arenas.begin(); it != arenas.end();) {
            if (it->free.size() == 1 && it->free.begin()->first == 0 && it->free.begin()->second == it->size) {
                (void)(host ? hipHostFree(it->base) : hipFree(it->base));
                total -= it->size;
                it = arenas.erase(it);
            } else {
                ++it;
            }
        }
    }

    void destroy()
    {
        for (Arena& a : arenas) (void)(host ? hipHostFree(a.base) : hipFree(a.base));
        arenas.clear();
        total = 0;
    }
};

struct SJob {
    uint64_t id = 0;
    std::vector<unsigned char> vtx;    // host copy of the vertices (kept until the job ends)
    int64_t n = 0;
    ohx_pc_params params{};
    uint32_t diff = 0, mode = 0;
    Geometry g{};
    bool mfma = false, fp64 = false;
    // RS on the matrix cores: the setup partitions the points (eligible ones
    // first, in order); GEMM slices cover part[0, ne), direct slices the rest
    bool rsm = false;
    PointSoA part{};
    void* rs_tmp = nullptr;
    size_t rs_tmp_bytes = 0;
    int64_t ne = -1;                   // matrix-core points; -1 until the setup has completed
    int64_t n_full = 0;                // of which the first n_full have the whole hologram as box
    int64_t slice_d = 0;               // direct slice of an RS matrix-core job
    double* out_field = nullptr;
    uint8_t* out_u8 = nullptr;
    // device: one block from the pool
    void* block = nullptr;             // vertices + SoA + field + enc + u8 + GEMM scratch
    PointSoA soa{};
    void* d_vtx = nullptr;
    double* field = nullptr;
    double* enc = nullptr;
    uint8_t* u8 = nullptr;
    void* gemm = nullptr;
    size_t gemm_bytes = 0;
    size_t bytes = 0;
    // page-locked staging of the outputs, copied out by ohx_server_wait
    char* stage = nullptr;
    size_t stage_bytes = 0;
    int64_t next = 0, slice = 0;
    // state
    JobState state = J_PENDING;
    bool cancel = false, done = false;
    hipEvent_t ev_ready = nullptr;     // setup finished (consumed by the first slice)
    hipEvent_t ev_last = nullptr;      // the latest work enqueued for the job
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
    std::vector<ohx::SJob*> active;    // admitted and not yet retired
    std::map<uint64_t, std::unique_ptr<ohx::SJob>> jobs;
    uint64_t next_id = 1;
    bool stop = false;
    size_t reserved = 0, peak = 0, budget = 0;
    uint64_t jobs_done = 0, iterations = 0, graph_replays = 0;
    std::chrono::steady_clock::time_point t0;
    hipStream_t s_setup = nullptr, s_finish = nullptr;   // compute: ctx->stream
    ohx::Pool pool, host_pool;
    unsigned long long* d_mm = nullptr;   // finish stream's normaliser: min/max words, then per-block keys
    // direct-sum job tables, one per iteration in flight
    ohx::DirectJob* h_table[ohx::DEPTH] = {};
    int* h_njobs[ohx::DEPTH] = {};
    ohx::DirectJob* d_table[ohx::DEPTH] = {};
    int* d_njobs[ohx::DEPTH] = {};
    hipGraphExec_t graph[ohx::DEPTH] = {};
    std::deque<hipEvent_t> flights;    // end events of the iterations in flight
    int next_set = 0;
    std::vector<hipEvent_t> free_events;

    double now_ms() const
    {
        return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
};

namespace ohx {

static hipEvent_t get_event(ohx_server* sv)
{
    if (!sv->free_events.empty()) {
        hipEvent_t e = sv->free_events.back();
        sv->free_events.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) {
        (void)hipGetLastError();
        return nullptr;
    }
    return e;
}
static void put_event(ohx_server* sv, hipEvent_t& e)
{
    if (e) sv->free_events.push_back(e);
    e = nullptr;
}

static size_t job_bytes(const SJob& j)
{
    const int64_t npx = (int64_t)j.g.pn_x * j.g.pn_y;
    size_t b = (size_t)j.n * 56 + 256;
    b += points_bytes(j.n) + 256;
    b += (size_t)npx * (16 + 8 + 1) + 3 * 256;
    if (j.mfma) b += fresnel_gemm_scratch_bytes(j.g, j.slice) + 256;
    if (j.rsm)
        b += rs_partition_scratch_bytes(j.n) + points_bytes(j.n) + rs_gemm_scratch_bytes(j.g, j.slice) + 3 * 256;
    return b;
}

static size_t stage_bytes(const SJob& j)
{
    const size_t npx = (size_t)j.g.pn_x * j.g.pn_y;
    // RS matrix-core jobs: one more word, the eligible-point count read back by the setup
    return (j.out_field ? npx * 16 : 0) + (j.out_u8 ? (npx + 255) / 256 * 256 : 0) + (j.rsm ? 256 : 0);
}
static unsigned long long* rs_count(const SJob& j)
{
    return (unsigned long long*)(j.stage + j.stage_bytes - 256);
}

// points per slice for `pairs` point-pixel pairs; RS on the matrix cores
// takes half the matrix-core pairs per slice (two K entries per point)
static int64_t slice_points(const SJob& j, double pairs)
{
    const double npx = (double)j.g.pn_x * j.g.pn_y;
    int64_t s = (int64_t)(pairs / npx);
    s = std::max<int64_t>(s, 256);
    s = (s + DIRECT_CHUNK - 1) / DIRECT_CHUNK * DIRECT_CHUNK;   // whole LDS chunks of the direct kernel
    return std::min<int64_t>(s, std::max<int64_t>(j.n, 1));
}

// carve the job's block and enqueue its setup on the setup stream (no sync)
static int setup_job(ohx_server* sv, SJob& j)
{
    const int64_t npx = (int64_t)j.g.pn_x * j.g.pn_y;
    char* b = (char*)j.block;
    auto take = [&](size_t bytes) {
        char* r = b;
        b += (bytes + 255) & ~size_t(255);
        return (void*)r;
    };
    j.d_vtx = take((size_t)j.n * 56);
    carve_points(take(points_bytes(j.n)), j.n, &j.soa);
    j.field = (double*)take((size_t)npx * 16);
    j.enc = (double*)take((size_t)npx * 8);
    j.u8 = (uint8_t*)take((size_t)npx);
    if (j.mfma) {
        j.gemm_bytes = fresnel_gemm_scratch_bytes(j.g, j.slice);
        j.gemm = take(j.gemm_bytes);
    }
    if (j.rsm) {
        j.gemm_bytes = rs_gemm_scratch_bytes(j.g, j.slice);
        j.gemm = take(j.gemm_bytes);
        j.rs_tmp_bytes = rs_partition_scratch_bytes(j.n);
        j.rs_tmp = take(j.rs_tmp_bytes);
        carve_points(take(points_bytes(j.n)), j.n, &j.part);
    }
    int st = OHX_OK;
    if (j.n > 0) {
        if (hipMemcpyAsync(j.d_vtx, j.vtx.data(), (size_t)j.n * 56, hipMemcpyHostToDevice, sv->s_setup) !=
            hipSuccess)
            st = OHX_ERR_DEVICE;
        if (!st) st = launch_prep_on(sv->s_setup, j.d_vtx, j.g, j.soa);
        if (!st && j.rsm) {
            unsigned long long* dc = nullptr;
            st = launch_rs_partition(sv->s_setup, j.rs_tmp, j.rs_tmp_bytes, j.g, j.soa, j.part, j.mode, false, &dc);
            if (!st && hipMemcpyAsync(rs_count(j), dc, 2 * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                                      sv->s_setup) != hipSuccess)
                st = OHX_ERR_DEVICE;
        }
    }
    if (!st) st = field_zero_on(sv->s_setup, j.field, npx);
    if (!st && !(j.ev_ready = get_event(sv))) st = OHX_ERR_DEVICE;
    if (!st && hipEventRecord(j.ev_ready, sv->s_setup) != hipSuccess) st = OHX_ERR_DEVICE;
    if (st) (void)hipGetLastError();
    return st;
}

// encode, normalise and stage the outputs behind ev (the event after the
// job's last slice, or its setup for a job without points); records ev_last
static int enqueue_finish(ohx_server* sv, SJob& j, hipEvent_t ev)
{
    const int64_t npx = (int64_t)j.g.pn_x * j.g.pn_y;
    hipStream_t s = sv->s_finish;
    if (hipStreamWaitEvent(s, ev, 0) != hipSuccess) return OHX_ERR_DEVICE;
    int st = OHX_OK;
    char* stage = j.stage;
    if (j.out_field) {
        if (hipMemcpyAsync(stage, j.field, npx * 16, hipMemcpyDeviceToHost, s) != hipSuccess) st = OHX_ERR_DEVICE;
        stage += npx * 16;
    }
    if (!st && j.out_u8) {
        const double* f = j.field;
        double* e = j.enc;
        uint8_t* o = j.u8;
        st = launch_encode_normalize_u8(s, sv->d_mm, sv->d_mm + 8, &f, 1, npx, OHX_ENCODE_PHASE, &e, &o);
        if (!st && hipMemcpyAsync(stage, j.u8, npx, hipMemcpyDeviceToHost, s) != hipSuccess) st = OHX_ERR_DEVICE;
    }
    if (!st && hipEventRecord(j.ev_last, s) != hipSuccess) st = OHX_ERR_DEVICE;
    if (st) (void)hipGetLastError();
    return st;
}

static int build_graph(ohx_server* sv, int set)
{
    ohx_ctx* c = sv->ctx;
    hipGraph_t g = nullptr;
    int st = check_hip(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal), "begin capture");
    if (st) return st;
    (void)hipMemcpyAsync(sv->d_njobs[set], sv->h_njobs[set], sizeof(int), hipMemcpyHostToDevice, c->stream);
    (void)hipMemcpyAsync(sv->d_table[set], sv->h_table[set], sizeof(DirectJob) * MAX_TABLE, hipMemcpyHostToDevice,
                         c->stream);
    launch_direct_jobs(c->stream, sv->d_table[set], sv->d_njobs[set], PERSISTENT_GRID);
    st = check_hip(c, hipStreamEndCapture(c->stream, &g), "end capture");
    if (st) return st;
    st = check_hip(c, hipGraphInstantiate(&sv->graph[set], g, nullptr, nullptr, 0), "graph instantiate");
    (void)hipGraphDestroy(g);
    return st;
}

// mark a job done (status st) and give its device memory back (the staging
// stays until ohx_server_wait has copied it out); caller holds sv->mu
static void retire(ohx_server* sv, SJob* j, int st)
{
    if (j->block) sv->pool.give(j->block, j->bytes);
    j->block = nullptr;
    sv->reserved -= j->bytes;
    put_event(sv, j->ev_ready);
    put_event(sv, j->ev_last);
    std::vector<unsigned char>().swap(j->vtx);
    if (st != OHX_OK && j->stage) {
        sv->host_pool.give(j->stage, j->stage_bytes);
        j->stage = nullptr;
    }
    j->st.status = st;
    j->st.end_ms = sv->now_ms();
    j->state = J_DONE;
    j->done = true;
    if (st == OHX_OK) ++sv->jobs_done;
    sv->active.erase(std::find(sv->active.begin(), sv->active.end(), j));
    sv->cv_done.notify_all();
}

// The job's next slice is a matrix-core slice
static bool in_gemm(const SJob& j) { return j.mfma || (j.rsm && j.next < j.ne); }
// End (exclusive) of the job's next slice
static int64_t slice_end(const SJob& j)
{
    if (j.rsm) return j.next < j.ne ? std::min(j.ne, j.next + j.slice) : std::min(j.n, j.next + j.slice_d);
    return std::min(j.n, j.next + j.slice);
}
static int64_t slices_left(const SJob& j)
{
    if (!j.rsm) return (j.n - j.next + j.slice - 1) / j.slice;
    const int64_t g = j.next < j.ne ? (j.ne - j.next + j.slice - 1) / j.slice : 0;
    return g + (j.n - std::max(j.next, j.ne) + j.slice_d - 1) / j.slice_d;
}
// Estimated GPU time of one slice (ms) from nominal rates of the kernels:
// FP32 direct ~5e9 pairs/ms, FP64 direct ~1e9, Fresnel matrix cores ~5e10,
// RS matrix cores ~2.5e10.
static double slice_ms(const SJob& j)
{
    const double pairs = (double)(slice_end(j) - j.next) * j.g.pn_x * j.g.pn_y;
    return pairs / (j.mfma ? 5e10 : in_gemm(j) ? 2.5e10 : j.fp64 ? 1e9 : 5e9);
}

// Enqueue one iteration over the jobs in act; true when something was
// enqueued.  Jobs are taken shortest remaining work first (in slices, then
// arrival order) until the iteration's estimated GPU time reaches
// ITER_BUDGET_MS, so small jobs are not held behind large ones.  The GEMM
// slices go first and get an event of their own; the direct jobs' table and
// persistent kernel follow.  Jobs whose last slice went into the iteration
// get their finish work enqueued behind the event that covers it.
constexpr double ITER_BUDGET_MS = 3.0;
static bool enqueue_iteration(ohx_server* sv, const std::vector<SJob*>& act, int* err)
{
    ohx_ctx* c = sv->ctx;
    hipStream_t s = c->stream;
    std::vector<SJob*> cand;
    for (SJob* j : act) {
        if (j->state != J_ACTIVE || j->cancel || j->next >= j->n) continue;
        if (j->rsm && j->ne < 0) {   // its split into GEMM and direct points is known once the setup is done
            if (!j->ev_ready || hipEventQuery(j->ev_ready) != hipSuccess) continue;
            j->ne = (int64_t)rs_count(*j)[0];
            j->n_full = (int64_t)rs_count(*j)[1];
        }
        cand.push_back(j);
    }
    if (cand.empty()) return false;
    auto left = [](const SJob* j) { return slices_left(*j); };
    std::stable_sort(cand.begin(), cand.end(), [&](const SJob* a, const SJob* b) {
        return left(a) != left(b) ? left(a) < left(b) : a->id < b->id;
    });
    std::vector<SJob*> gemm_jobs, direct_jobs;
    double ms = 0;
    for (SJob* j : cand) {
        const double t = slice_ms(*j);
        if (!gemm_jobs.empty() || !direct_jobs.empty())
            if (ms + t > ITER_BUDGET_MS) continue;
        if (!in_gemm(*j) && (int)direct_jobs.size() == MAX_TABLE) continue;
        ms += t;
        (in_gemm(*j) ? gemm_jobs : direct_jobs).push_back(j);
    }
    auto wait_ready = [&](SJob* j) {
        if (!j->ev_ready) return true;
        if (hipStreamWaitEvent(s, j->ev_ready, 0) != hipSuccess) return false;
        put_event(sv, j->ev_ready);
        return true;
    };
    // advance the jobs of one part behind its event
    auto advance = [&](const std::vector<SJob*>& part, hipEvent_t ev, double t) {
        for (SJob* j : part) {
            if (j->st.iterations == 0) j->st.start_ms = t;
            ++j->st.iterations;
            j->next = slice_end(*j);
            if (j->next >= j->n) {
                j->state = J_FINISHING;
                if (int st = enqueue_finish(sv, *j, ev)) return st;
            } else if (hipEventRecord(j->ev_last, s) != hipSuccess) {
                return (int)OHX_ERR_DEVICE;
            }
        }
        return (int)OHX_OK;
    };
    const double t0 = sv->now_ms();
    int st = OHX_OK;
    hipEvent_t ev_g = nullptr, ev_d = nullptr;
    // A job whose last slice this is gets an event right behind that slice,
    // so its encoding and read-back start before the iteration's other
    // slices have run (the small jobs come first).
    std::vector<SJob*> gemm_rest;
    for (SJob* j : gemm_jobs) {
        if (!wait_ready(j)) st = OHX_ERR_DEVICE;
        const int64_t p0 = j->next, p1 = slice_end(*j);
        if (!st)
            st = j->mfma ? launch_fresnel_gemm_range(s, j->gemm, j->gemm_bytes, j->g, j->soa, p0, p1, j->field, j->mode)
                         : launch_rs_gemm_range(s, j->gemm, j->gemm_bytes, j->g, rs_partition_records(j->rs_tmp, j->n),
                                                p0, p1, j->n_full, j->field, j->mode);
        if (!st && p1 >= j->n) {
            hipEvent_t e = get_event(sv);
            if (!e || hipEventRecord(e, s) != hipSuccess) st = OHX_ERR_DEVICE;
            if (!st) st = advance(std::vector<SJob*>{j}, e, t0);
            put_event(sv, e);   // the finish stream's wait has captured it
        } else if (!st) {
            gemm_rest.push_back(j);
        }
        if (st) break;
    }
    if (!st && !gemm_jobs.empty()) {
        if (!(ev_g = get_event(sv)) || hipEventRecord(ev_g, s) != hipSuccess) st = OHX_ERR_DEVICE;
        if (!st) st = advance(gemm_rest, ev_g, t0);
    }
    if (!st && !direct_jobs.empty()) {
        const int set = sv->next_set;
        int nd = 0, items = 0;
        for (SJob* j : direct_jobs) {
            if (!wait_ready(j)) {
                st = OHX_ERR_DEVICE;
                break;
            }
            DirectJob& d = sv->h_table[set][nd++];
            d.g = j->g;
            d.s = j->rsm ? j->part : j->soa;
            d.field = (double2*)j->field;
            d.p0 = j->next;
            d.p1 = slice_end(*j);
            d.tiles = direct_tiles(j->g, j->fp64, &d.tiles_x);
            d.item_base = items;
            d.fp64 = j->fp64 ? 1 : 0;
            items += d.tiles;
        }
        if (!st) {
            *sv->h_njobs[set] = nd;
            if (sv->graph[set]) {
                st = check_hip(c, hipGraphLaunch(sv->graph[set], s), "hipGraphLaunch");
                ++sv->graph_replays;
            } else {
                st = check_hip(c, hipMemcpyAsync(sv->d_njobs[set], sv->h_njobs[set], sizeof(int),
                                                 hipMemcpyHostToDevice, s), "h2d table");
                if (!st)
                    st = check_hip(c, hipMemcpyAsync(sv->d_table[set], sv->h_table[set], sizeof(DirectJob) * nd,
                                                     hipMemcpyHostToDevice, s), "h2d table");
                if (!st) st = launch_direct_jobs(s, sv->d_table[set], sv->d_njobs[set], PERSISTENT_GRID);
            }
            sv->next_set = (set + 1) % DEPTH;
        }
        if (!st && (!(ev_d = get_event(sv)) || hipEventRecord(ev_d, s) != hipSuccess)) st = OHX_ERR_DEVICE;
        if (!st) st = advance(direct_jobs, ev_d, t0);
    }
    // the iteration's last event marks its end; the other one only fed
    // enqueue_finish (the finish stream's wait has captured it)
    hipEvent_t last = ev_d ? ev_d : ev_g;
    if (ev_d && ev_g) put_event(sv, ev_g);
    if (last) sv->flights.push_back(last);
    ++sv->iterations;
    if (st) *err = st;
    return true;
}

static void sync_streams(ohx_server* sv)
{
    (void)hipStreamSynchronize(sv->s_setup);
    (void)hipStreamSynchronize(sv->ctx->stream);
    (void)hipStreamSynchronize(sv->s_finish);
    (void)hipGetLastError();
}

static void drop_flights(ohx_server* sv)
{
    for (hipEvent_t& e : sv->flights) put_event(sv, e);
    sv->flights.clear();
}

static void worker_loop(ohx_server* sv)
{
    ohx_ctx* c = sv->ctx;
    (void)hipSetDevice(c->device);
    for (;;) {
        bool progressed = false;
        int err = OHX_OK;
        std::vector<SJob*> act;
        {
            std::unique_lock<std::mutex> lk(sv->mu);
            if (sv->flights.empty() && sv->active.empty())
                sv->cv_work.wait(lk, [&] { return sv->stop || !sv->pending.empty() || !sv->active.empty(); });
            if (sv->stop) break;
            // admission in arrival order: slots, memory budget, pool space
            while (!sv->pending.empty() && (int)sv->active.size() < sv->opts.max_active) {
                SJob* j = sv->pending.front();
                if (sv->reserved + j->bytes > sv->budget) break;
                int hs = OHX_OK;
                void* blk = sv->pool.take(j->bytes, sv->budget, MIN_ARENA, &hs);
                if (!blk && hs == OHX_OK && sv->active.empty()) {   // idle but fragmented
                    sv->pool.trim();
                    blk = sv->pool.take(j->bytes, sv->budget, MIN_ARENA, &hs);
                }
                char* stg = nullptr;
                if (blk && j->stage_bytes) {
                    stg = (char*)sv->host_pool.take(j->stage_bytes, ~size_t(0) >> 1, MIN_HOST_ARENA, &hs);
                    if (!stg) {
                        sv->pool.give(blk, j->bytes);
                        blk = nullptr;
                    }
                }
                if (!blk && hs == OHX_OK) break;   // fragmented: wait for a release
                sv->pending.pop_front();
                if (!blk) {
                    j->st.status = hs;
                    j->state = J_DONE;
                    j->done = true;
                    sv->cv_done.notify_all();
                    continue;
                }
                j->block = blk;
                j->stage = stg;
                sv->reserved += j->bytes;
                sv->peak = std::max(sv->peak, sv->reserved);
                sv->active.push_back(j);
                j->st.admit_ms = sv->now_ms();
                j->ev_last = get_event(sv);
                int st = j->ev_last ? setup_job(sv, *j) : OHX_ERR_DEVICE;
                if (!st && j->n == 0) {   // nothing to diffract: finish behind the setup
                    j->state = J_FINISHING;
                    st = enqueue_finish(sv, *j, j->ev_ready);
                } else if (!st) {
                    j->state = J_ACTIVE;
                }
                if (st) retire(sv, j, st);
                progressed = true;
            }
            // cancelled admitted jobs take no further slices; they are retired
            // once the work already enqueued for them has completed
            for (SJob* j : sv->active)
                if (j->cancel && j->state == J_ACTIVE) j->state = J_CANCELLING;
            act = sv->active;
        }
        // completed iterations
        while (!sv->flights.empty() && hipEventQuery(sv->flights.front()) == hipSuccess) {
            put_event(sv, sv->flights.front());
            sv->flights.pop_front();
            progressed = true;
        }
        // new iterations while fewer than DEPTH are in flight
        while ((int)sv->flights.size() < DEPTH && !err && enqueue_iteration(sv, act, &err)) progressed = true;
        if (err) sync_streams(sv);   // the streams may hold failed work: drain them
        // finished and cancelled jobs whose work has completed
        {
            std::lock_guard<std::mutex> lk(sv->mu);
            const std::vector<SJob*> snapshot = sv->active;
            for (SJob* j : snapshot) {
                if (err) {
                    retire(sv, j, err);
                    progressed = true;
                    continue;
                }
                if (j->state != J_FINISHING && j->state != J_CANCELLING) continue;
                if (j->ev_ready && hipEventQuery(j->ev_ready) != hipSuccess) continue;
                if (hipEventQuery(j->ev_last) != hipSuccess) continue;
                retire(sv, j, j->state == J_FINISHING ? OHX_OK : OHX_ERR_CANCELLED);
                progressed = true;
            }
        }
        if (err) drop_flights(sv);
        if (!progressed) {
            std::unique_lock<std::mutex> lk(sv->mu);
            sv->cv_work.wait_for(lk, std::chrono::microseconds(50), [&] { return sv->stop; });
        }
    }
    // stop: drain the streams, then cancel everything still queued or admitted
    sync_streams(sv);
    std::lock_guard<std::mutex> lk(sv->mu);
    for (SJob* j : sv->pending) {
        j->st.status = OHX_ERR_CANCELLED;
        j->state = J_DONE;
        j->done = true;
    }
    sv->pending.clear();
    while (!sv->active.empty()) retire(sv, sv->active.back(), OHX_ERR_CANCELLED);
    drop_flights(sv);
    sv->cv_done.notify_all();
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
    // ~1.5 ms slices: a job that arrives waits for at most two of them (the
    // iterations in flight); twice that size gave 3 % more throughput but a
    // p99 slowdown of 15-18 instead of 10 on the mixed request stream
    if (o.slice_pairs <= 0) o.slice_pairs = 7.5e9;
    if (o.slice_pairs_mfma <= 0) o.slice_pairs_mfma = 7.5e10;
    auto sv = std::make_unique<ohx_server>();
    sv->host_pool.host = true;
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
    bool ok = hipStreamCreateWithFlags(&sv->s_setup, hipStreamNonBlocking) == hipSuccess &&
              hipStreamCreateWithFlags(&sv->s_finish, hipStreamNonBlocking) == hipSuccess &&
              hipMalloc((void**)&sv->d_mm, (8 + ENCODE_MM_PART_WORDS) * sizeof(unsigned long long)) == hipSuccess;
    for (int s = 0; ok && s < DEPTH; ++s) {
        ok = hipHostMalloc((void**)&sv->h_table[s], sizeof(DirectJob) * MAX_TABLE, hipHostMallocDefault) ==
                 hipSuccess &&
             hipHostMalloc((void**)&sv->h_njobs[s], sizeof(int), hipHostMallocDefault) == hipSuccess &&
             hipMalloc((void**)&sv->d_table[s], sizeof(DirectJob) * MAX_TABLE) == hipSuccess &&
             hipMalloc((void**)&sv->d_njobs[s], sizeof(int)) == hipSuccess;
        if (ok) {
            std::memset(sv->h_table[s], 0, sizeof(DirectJob) * MAX_TABLE);
            *sv->h_njobs[s] = 0;
        }
    }
    if (!ok) {
        (void)hipGetLastError();
        ohx_server_destroy(sv.release());
        return OHX_ERR_OUT_OF_MEMORY;
    }
    for (int s = 0; o.use_graph && s < DEPTH; ++s)
        if ((st = build_graph(sv.get(), s))) {
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
    // RS with a BF16 matrix-core mode, as ohx_pc_diffract runs it
    j->rsm = mf && job->diff_flag == OHX_DIFF_RS && !(job->mode & OHX_MODE_MFMA_FP16X3) && j->n > 0;
    j->fp64 = !j->mfma && (job->mode & (OHX_MODE_FLOAT | OHX_MODE_FASTMATH)) == 0 && mf == 0;
    j->out_field = job->out_field;
    j->out_u8 = job->out_u8;
    j->slice = slice_points(*j, j->mfma ? sv->opts.slice_pairs_mfma
                                        : j->rsm ? 0.5 * sv->opts.slice_pairs_mfma : sv->opts.slice_pairs);
    j->slice_d = slice_points(*j, sv->opts.slice_pairs);
    j->bytes = job_bytes(*j);
    j->stage_bytes = stage_bytes(*j);
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
    std::unique_ptr<SJob> owned = std::move(it->second);
    sv->jobs.erase(it);
    lk.unlock();
    // the outputs, staged in page-locked memory by the finish stream
    if (st == OHX_OK && owned->stage) {
        const size_t npx = (size_t)owned->g.pn_x * owned->g.pn_y;
        const char* src = owned->stage;
        if (owned->out_field) {
            std::memcpy(owned->out_field, src, npx * 16);
            src += npx * 16;
        }
        if (owned->out_u8) std::memcpy(owned->out_u8, src, npx);
    }
    if (owned->stage) {
        std::lock_guard<std::mutex> lk2(sv->mu);
        sv->host_pool.give(owned->stage, owned->stage_bytes);
    }
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
        j->state = J_DONE;
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
        if (sv->s_setup && sv->s_finish) sync_streams(sv);
        for (int s = 0; s < DEPTH; ++s) {
            if (sv->graph[s]) (void)hipGraphExecDestroy(sv->graph[s]);
            if (sv->d_table[s]) (void)hipFree(sv->d_table[s]);
            if (sv->d_njobs[s]) (void)hipFree(sv->d_njobs[s]);
            if (sv->h_table[s]) (void)hipHostFree(sv->h_table[s]);
            if (sv->h_njobs[s]) (void)hipHostFree(sv->h_njobs[s]);
        }
        if (sv->d_mm) (void)hipFree(sv->d_mm);
        sv->pool.destroy();
        // staging of jobs that were never waited for
        sv->jobs.clear();
        sv->host_pool.destroy();
        for (hipEvent_t e : sv->free_events) (void)hipEventDestroy(e);
        if (sv->s_setup) (void)hipStreamDestroy(sv->s_setup);
        if (sv->s_finish) (void)hipStreamDestroy(sv->s_finish);
        ohx_ctx_destroy(sv->ctx);
    }
    delete sv;
    return OHX_OK;
}

}  // extern "C"
