// Context, device memory and the C-ABI entry points of libopenholo_amd.
//
// Replaces the host orchestration of the reference's GPU path
// (ophgen/src/ophPointCloud_GPU.cpp:179-412) and its cudaWrapper singleton
// (ophgen/src/cudaWrapper.cpp:39-179): one context per device owns a stream
// and grow-only scratch; nothing is allocated per launch once warm, and
// errors are returned as status codes instead of exit().
#include <cmath>
#include <cstdio>
#include <cstring>

#include "ohx_internal.h"

namespace ohx {

int set_error(ohx_ctx* ctx, int status, const std::string& msg)
{
    if (ctx) ctx->last_error = msg;
    return status;
}

int check_hip(ohx_ctx* ctx, hipError_t e, const char* what)
{
    if (e == hipSuccess) return OHX_OK;
    const int st = (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation)
                       ? OHX_ERR_OUT_OF_MEMORY : OHX_ERR_DEVICE;
    return set_error(ctx, st, std::string(what) + ": " + hipGetErrorString(e));
}

int geometry_from_params(ohx_ctx* ctx, const ohx_pc_params* p, uint32_t diff, Geometry* g)
{
#pragma clang fp contract(off)
    if (!p) return set_error(ctx, OHX_ERR_INVALID_ARG, "params is null");
    if (diff != OHX_DIFF_RS && diff != OHX_DIFF_FRESNEL)
        return set_error(ctx, OHX_ERR_INVALID_ARG, "diff_flag must be 0 (RS) or 1 (Fresnel)");
    if (p->pn_x <= 0 || p->pn_y <= 0 || p->pn_x > (1 << 20) || p->pn_y > (1 << 20))
        return set_error(ctx, OHX_ERR_INVALID_ARG, "pixel_number out of range");
    if (!(p->pp_x > 0) || !(p->pp_y > 0) || !(p->lambda > 0))
        return set_error(ctx, OHX_ERR_INVALID_ARG, "pixel_pitch and lambda must be > 0");
    if (!std::isfinite(p->pp_x) || !std::isfinite(p->pp_y) || !std::isfinite(p->lambda) ||
        !std::isfinite(p->distance) || !std::isfinite(p->scale[0]) || !std::isfinite(p->scale[1]) ||
        !std::isfinite(p->scale[2]))
        return set_error(ctx, OHX_ERR_INVALID_ARG, "pitch, lambda, distance and scale must be finite");
    if (p->channel < 0 || p->channel > 2)
        return set_error(ctx, OHX_ERR_INVALID_ARG, "channel must index Vertex::color (0..2)");
    if (diff == OHX_DIFF_RS && !(p->lambda < 2 * p->pp_x && p->lambda < 2 * p->pp_y))
        return set_error(ctx, OHX_ERR_INVALID_ARG, "RS needs lambda < 2*pixel_pitch");
    g->pn_x = p->pn_x;
    g->pn_y = p->pn_y;
    g->off_x = p->offset_x;
    g->off_y = p->offset_y;
    g->pp_x = p->pp_x;
    g->pp_y = p->pp_y;
    // ss = pn * pp and x = -ss/2, as ophGen.cpp:367-368,377-378
    const double ssX = p->pn_x * p->pp_x, ssY = p->pn_y * p->pp_y;
    g->x0 = -ssX / 2;
    g->y0 = -ssY / 2;
    g->lambda = p->lambda;
    g->inv_lambda = 1.0 / p->lambda;
    // RS half-width coefficient tx/sqrt(1-tx^2), ophGen.cpp:373-376
    const double tx = p->lambda / (2 * p->pp_x), ty = p->lambda / (2 * p->pp_y);
    const double sqrtX = std::sqrt(1 - (tx * tx)), sqrtY = std::sqrt(1 - (ty * ty));
    g->hx_coef = tx / sqrtX;
    g->hy_coef = ty / sqrtY;
    for (int i = 0; i < 3; ++i) g->scale[i] = p->scale[i];
    g->distance = p->distance;
    g->channel = p->channel;
    g->use_phase = p->use_phase ? 1 : 0;
    g->diff = (int32_t)diff;
    g->rs_skip_mask = nullptr;
    g->rs_mask_words = 0;
    return OHX_OK;
}

size_t points_bytes(int64_t n) { return (size_t)n * (6 * sizeof(double) + sizeof(int4)) + 8 * 256; }

void carve_points(void* base, int64_t n, PointSoA* soa)
{
    char* b = (char*)base;
    auto take = [&](size_t bytes) {
        char* r = b;
        b += (bytes + 255) & ~size_t(255);
        return (void*)r;
    };
    soa->px = (double*)take(n * sizeof(double));
    soa->py = (double*)take(n * sizeof(double));
    soa->pz = (double*)take(n * sizeof(double));
    soa->cyc0 = (double*)take(n * sizeof(double));
    soa->kamp = (double*)take(n * sizeof(double));
    soa->q = (double*)take(n * sizeof(double));
    soa->box = (int4*)take(n * sizeof(int4));
    soa->n = n;
}

int ensure_points(ohx_ctx* ctx, int64_t n, PointSoA* soa)
{
    const size_t need = points_bytes(n);
    if (need > ctx->pts_bytes) {
        if (ctx->pts) {
            int st = check_hip(ctx, hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
            if (st) return st;
            (void)hipFree(ctx->pts);
            ctx->pts = nullptr;
            ctx->pts_bytes = 0;
        }
        size_t cap = need + need / 4;
        int st = check_hip(ctx, hipMalloc(&ctx->pts, cap), "hipMalloc(points)");
        if (st) return st;
        ctx->pts_bytes = cap;
    }
    carve_points(ctx->pts, n, soa);
    return OHX_OK;
}

int ensure_gemm_scratch(ohx_ctx* ctx, size_t bytes)
{
    if (bytes <= ctx->gemm_scratch_bytes) return OHX_OK;
    if (ctx->gemm_scratch) {
        int st = check_hip(ctx, hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
        if (st) return st;
        (void)hipFree(ctx->gemm_scratch);
        ctx->gemm_scratch = nullptr;
        ctx->gemm_scratch_bytes = 0;
    }
    int st = check_hip(ctx, hipMalloc(&ctx->gemm_scratch, bytes), "hipMalloc(gemm scratch)");
    if (st) return st;
    ctx->gemm_scratch_bytes = bytes;
    return OHX_OK;
}

__global__ void k_field_zero(double2* __restrict__ f, int64_t n)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) f[i] = make_double2(0.0, 0.0);
}

__global__ void k_field_acc(double2* __restrict__ d, const double2* __restrict__ s, int64_t n)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) {
        double2 a = d[i], b = s[i];
        d[i] = make_double2(a.x + b.x, a.y + b.y);
    }
}

__global__ void k_sincos_rev(const float* __restrict__ x, int64_t n, float* __restrict__ s,
                             float* __restrict__ c)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) {
        const float v = x[i];
        s[i] = sin_rev(v);
        c[i] = cos_rev(v);
    }
}

static unsigned grid_for(int64_t n, int block)
{
    int64_t g = (n + block - 1) / block;
    if (g > 65536) g = 65536;
    if (g < 1) g = 1;
    return (unsigned)g;
}

int field_zero_on(hipStream_t stream, double* f, int64_t n)
{
    if (n <= 0) return OHX_OK;
    hipLaunchKernelGGL(k_field_zero, dim3(grid_for(n, 256)), dim3(256), 0, stream, (double2*)f, n);
    return hipGetLastError() == hipSuccess ? OHX_OK : OHX_ERR_DEVICE;
}

}  // namespace ohx

using namespace ohx;

extern "C" {

const char* ohx_status_string(int s)
{
    switch (s) {
    case OHX_OK: return "ok";
    case OHX_ERR_INVALID_ARG: return "invalid argument";
    case OHX_ERR_OUT_OF_MEMORY: return "out of device memory";
    case OHX_ERR_DEVICE: return "device error";
    case OHX_ERR_ADMISSION: return "admission denied";
    case OHX_ERR_NOT_FOUND: return "not found";
    case OHX_ERR_TIMEOUT: return "timeout";
    case OHX_ERR_UNSUPPORTED: return "unsupported";
    case OHX_ERR_CANCELLED: return "cancelled";
    default: return "unknown status";
    }
}

int ohx_ctx_create(int device, ohx_ctx** out)
{
    if (!out) return OHX_ERR_INVALID_ARG;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return OHX_ERR_DEVICE;
    if (device < 0 || device >= ndev) return OHX_ERR_INVALID_ARG;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) return OHX_ERR_DEVICE;
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
        std::fprintf(stderr, "openholo_amd: device %d is %s; this library is built for gfx950\n",
                     device, prop.gcnArchName);
        return OHX_ERR_UNSUPPORTED;
    }
    ohx_ctx* c = new ohx_ctx();
    c->device = device;
    if (hipSetDevice(device) != hipSuccess ||
        hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&c->copy_stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&c->copy_ready, hipEventDisableTiming) != hipSuccess ||
        hipMalloc(&c->minmax, (8 + ENCODE_MM_PART_WORDS) * sizeof(unsigned long long)) != hipSuccess) {
        if (c->copy_ready) (void)hipEventDestroy(c->copy_ready);
        if (c->copy_stream) (void)hipStreamDestroy(c->copy_stream);
        if (c->stream) (void)hipStreamDestroy(c->stream);
        delete c;
        return OHX_ERR_DEVICE;
    }
    noalloc_init(device);   // address of the allocation-free entry point's scratch
    *out = c;
    return OHX_OK;
}

int ohx_ctx_destroy(ohx_ctx* c)
{
    if (!c) return OHX_ERR_INVALID_ARG;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    if (c->pts) (void)hipFree(c->pts);
    if (c->minmax) (void)hipFree(c->minmax);
    if (c->gemm_scratch) (void)hipFree(c->gemm_scratch);
    if (c->rs_scratch) (void)hipFree(c->rs_scratch);
    if (c->rs_mask) (void)hipFree(c->rs_mask);
    wrp_release(c);
    (void)hipStreamSynchronize(c->copy_stream);
    (void)hipEventDestroy(c->copy_ready);
    (void)hipStreamDestroy(c->copy_stream);
    (void)hipStreamDestroy(c->stream);
    delete c;
    return OHX_OK;
}

void* ohx_ctx_stream(ohx_ctx* c) { return c ? (void*)c->stream : nullptr; }
int ohx_ctx_device(ohx_ctx* c) { return c ? c->device : -1; }
const char* ohx_last_error(ohx_ctx* c) { return c ? c->last_error.c_str() : "null context"; }

int ohx_mem_info(ohx_ctx* c, size_t* fr, size_t* tot)
{
    if (!c || !fr || !tot) return OHX_ERR_INVALID_ARG;
    (void)hipSetDevice(c->device);
    return check_hip(c, hipMemGetInfo(fr, tot), "hipMemGetInfo");
}

int ohx_malloc(ohx_ctx* c, size_t bytes, void** p)
{
    if (!c || !p) return OHX_ERR_INVALID_ARG;
    *p = nullptr;
    if (bytes == 0) return OHX_OK;
    (void)hipSetDevice(c->device);
    return check_hip(c, hipMalloc(p, bytes), "hipMalloc");
}

int ohx_free(ohx_ctx* c, void* p)
{
    if (!c) return OHX_ERR_INVALID_ARG;
    if (!p) return OHX_OK;
    (void)hipSetDevice(c->device);
    int st = check_hip(c, hipStreamSynchronize(c->stream), "hipStreamSynchronize");
    if (st) return st;
    return check_hip(c, hipFree(p), "hipFree");
}

int ohx_host_alloc(size_t bytes, void** p)
{
    if (!p) return OHX_ERR_INVALID_ARG;
    *p = nullptr;
    if (bytes == 0) return OHX_OK;
    if (hipHostMalloc(p, bytes, hipHostMallocDefault) != hipSuccess) {
        (void)hipGetLastError();
        *p = nullptr;
        return OHX_ERR_DEVICE;
    }
    std::memset(*p, 0, bytes);
    return OHX_OK;
}

int ohx_host_free(void* p)
{
    if (!p) return OHX_OK;
    if (hipHostFree(p) != hipSuccess) {
        (void)hipGetLastError();
        return OHX_ERR_DEVICE;
    }
    return OHX_OK;
}

static int copy(ohx_ctx* c, void* d, const void* s, size_t n, hipMemcpyKind k)
{
    if (!c || (n && (!d || !s))) return OHX_ERR_INVALID_ARG;
    if (!n) return OHX_OK;
    (void)hipSetDevice(c->device);
    int st = check_hip(c, hipMemcpyAsync(d, s, n, k, c->stream), "hipMemcpyAsync");
    if (st) return st;
    return check_hip(c, hipStreamSynchronize(c->stream), "hipStreamSynchronize");
}

int ohx_memcpy_h2d(ohx_ctx* c, void* d, const void* s, size_t n) { return copy(c, d, s, n, hipMemcpyHostToDevice); }
int ohx_memcpy_d2h(ohx_ctx* c, void* d, const void* s, size_t n) { return copy(c, d, s, n, hipMemcpyDeviceToHost); }
int ohx_memcpy_d2d(ohx_ctx* c, void* d, const void* s, size_t n) { return copy(c, d, s, n, hipMemcpyDeviceToDevice); }

int ohx_memcpy_d2h_async_after(ohx_ctx* c, void* d, const void* s, size_t n)
{
    if (!c || (n && (!d || !s))) return OHX_ERR_INVALID_ARG;
    if (!n) return OHX_OK;
    // The event marks the end of the work enqueued on c->stream so far; the
    // copy stream's wait binds to this record, so re-recording the event for
    // the next copy does not move this copy's dependency.
    std::lock_guard<std::mutex> lk(c->mu);
    (void)hipSetDevice(c->device);
    int st = check_hip(c, hipEventRecord(c->copy_ready, c->stream), "hipEventRecord");
    if (!st) st = check_hip(c, hipStreamWaitEvent(c->copy_stream, c->copy_ready, 0), "hipStreamWaitEvent");
    if (!st) st = check_hip(c, hipMemcpyAsync(d, s, n, hipMemcpyDeviceToHost, c->copy_stream), "hipMemcpyAsync");
    return st;
}

int ohx_copy_synchronize(ohx_ctx* c)
{
    if (!c) return OHX_ERR_INVALID_ARG;
    (void)hipSetDevice(c->device);
    return check_hip(c, hipStreamSynchronize(c->copy_stream), "hipStreamSynchronize(copy)");
}

int ohx_memcpy_d2h_after(ohx_ctx* c, void* d, const void* s, size_t n)
{
    int st = ohx_memcpy_d2h_async_after(c, d, s, n);
    const int st2 = c ? ohx_copy_synchronize(c) : OHX_ERR_INVALID_ARG;
    return st ? st : st2;
}

int ohx_synchronize(ohx_ctx* c)
{
    if (!c) return OHX_ERR_INVALID_ARG;
    (void)hipSetDevice(c->device);
    return check_hip(c, hipStreamSynchronize(c->stream), "hipStreamSynchronize");
}

int ohx_field_zero(ohx_ctx* c, double* f, int64_t n)
{
    if (!c || n < 0 || (n && !f)) return OHX_ERR_INVALID_ARG;
    if (!n) return OHX_OK;
    (void)hipSetDevice(c->device);
    hipLaunchKernelGGL(k_field_zero, dim3(grid_for(n, 256)), dim3(256), 0, c->stream, (double2*)f, n);
    return check_hip(c, hipGetLastError(), "k_field_zero");
}

int ohx_field_accumulate(ohx_ctx* c, double* d, const double* s, int64_t n)
{
    if (!c || n < 0 || (n && (!d || !s))) return OHX_ERR_INVALID_ARG;
    if (!n) return OHX_OK;
    (void)hipSetDevice(c->device);
    hipLaunchKernelGGL(k_field_acc, dim3(grid_for(n, 256)), dim3(256), 0, c->stream,
                       (double2*)d, (const double2*)s, n);
    return check_hip(c, hipGetLastError(), "k_field_acc");
}

int ohx_pc_diffract(ohx_ctx* c, const void* d_vertices, int64_t n, double* d_field,
                    const ohx_pc_params* p, uint32_t diff, uint32_t mode)
{
    if (!c) return OHX_ERR_INVALID_ARG;
    if (n < 0 || (n > 0 && !d_vertices) || !d_field)
        return set_error(c, OHX_ERR_INVALID_ARG, "null buffer or negative point count");
    Geometry g;
    int st = geometry_from_params(c, p, diff, &g);
    if (st) return st;
    std::lock_guard<std::mutex> lk(c->mu);
    (void)hipSetDevice(c->device);
    c->last = ohx_launch_stats{};
    c->last.pairs_nominal = n * (int64_t)g.pn_x * g.pn_y;
    c->boxsum_pending = false;
    if (n == 0) return OHX_OK;  // setPointCloudModel with n < 1 is a no-op upstream
    PointSoA soa;
    if ((st = ensure_points(c, n, &soa))) return st;
    if ((st = launch_prep(c, d_vertices, g, soa, c->minmax + 4))) return st;
    c->boxsum_pending = true;
    c->boxsum_src = c->minmax + 4;
    const uint32_t mfma = mode & (OHX_MODE_MFMA_BF16X1 | OHX_MODE_MFMA_BF16X3 | OHX_MODE_MFMA_FP16X3);
    if (mfma && diff == OHX_DIFF_FRESNEL) return launch_fresnel_gemm(c, g, soa, d_field, mode);
    if (mfma && !(mode & OHX_MODE_MFMA_FP16X3)) return launch_rs_gemm(c, g, soa, d_field, mode);
    const bool fp64 = (mode & (OHX_MODE_FLOAT | OHX_MODE_FASTMATH)) == 0 && mfma == 0;
    return launch_direct(c, g, soa, d_field, fp64);
}

int ohx_sincos_rev(ohx_ctx* c, const float* d_x, int64_t n, float* d_sin, float* d_cos)
{
    if (!c || n < 0 || (n && (!d_x || !d_sin || !d_cos))) return OHX_ERR_INVALID_ARG;
    if (!n) return OHX_OK;
    (void)hipSetDevice(c->device);
    hipLaunchKernelGGL(k_sincos_rev, dim3(grid_for(n, 256)), dim3(256), 0, c->stream, d_x, n, d_sin, d_cos);
    return check_hip(c, hipGetLastError(), "k_sincos_rev");
}

int ohx_pc_probe_f64(ohx_ctx* c, const void* d_vertices, int64_t n, const ohx_pc_params* p,
                     uint32_t diff, const int64_t* d_pix, int64_t n_pix, double* d_out)
{
    if (!c) return OHX_ERR_INVALID_ARG;
    if (n < 0 || n_pix < 0 || (n > 0 && !d_vertices) || (n_pix > 0 && (!d_pix || !d_out)))
        return set_error(c, OHX_ERR_INVALID_ARG, "null buffer or negative count");
    Geometry g;
    int st = geometry_from_params(c, p, diff, &g);
    if (st) return st;
    std::lock_guard<std::mutex> lk(c->mu);
    (void)hipSetDevice(c->device);
    if (n_pix == 0) return OHX_OK;
    PointSoA soa;
    if ((st = ensure_points(c, n > 0 ? n : 1, &soa))) return st;
    soa.n = n;
    if (n > 0 && (st = launch_prep(c, d_vertices, g, soa))) return st;
    return launch_probe(c, g, soa, d_pix, n_pix, d_out);
}

int ohx_wrp_generate(ohx_ctx* c, const void* d_vertices, int64_t n, const ohx_wrp_params* p, double* d_field,
                     double* d_wrp)
{
    if (!c || !p) return OHX_ERR_INVALID_ARG;
    if (n < 0 || (n > 0 && !d_vertices) || !d_field)
        return set_error(c, OHX_ERR_INVALID_ARG, "null buffer or negative point count");
    if (p->pn_x < 1 || p->pn_y < 1 || (int64_t)p->pn_x * p->pn_y > (int64_t)1 << 28)
        return set_error(c, OHX_ERR_INVALID_ARG, "pixel number out of range");
    if (n > INT32_MAX) return set_error(c, OHX_ERR_INVALID_ARG, "WRP point count above 2^31 - 1");
    if (!(p->pp_x > 0) || !(p->pp_y > 0) || !(p->lambda > 0) || !std::isfinite(p->pp_x) ||
        !std::isfinite(p->pp_y) || !std::isfinite(p->lambda) || !std::isfinite(p->wrp_location) ||
        !std::isfinite(p->propagation_distance) || !std::isfinite(p->scale[0]) || !std::isfinite(p->scale[1]) ||
        !std::isfinite(p->scale[2]))
        return set_error(c, OHX_ERR_INVALID_ARG, "pitch, wave length, scale or distances not positive/finite");
    if (p->channel < 0 || p->channel > 2) return set_error(c, OHX_ERR_INVALID_ARG, "channel must be 0..2");
    if (p->random_phase) return set_error(c, OHX_ERR_INVALID_ARG, "random phase is not supported");
    std::lock_guard<std::mutex> lk(c->mu);
    (void)hipSetDevice(c->device);
    c->boxsum_pending = false;
    return wrp_generate(c, d_vertices, n, p, d_field, d_wrp);
}

int ohx_last_launch_stats(ohx_ctx* c, ohx_launch_stats* out)
{
    if (!c || !out) return OHX_ERR_INVALID_ARG;
    std::lock_guard<std::mutex> lk(c->mu);
    if (c->boxsum_pending) {   // the box-pixel count is read back on request only
        (void)hipSetDevice(c->device);
        unsigned long long v = 0;
        int st = check_hip(c, hipMemcpyAsync(&v, c->boxsum_src, sizeof(v), hipMemcpyDeviceToHost, c->stream),
                           "hipMemcpyAsync");
        if (!st) st = check_hip(c, hipStreamSynchronize(c->stream), "hipStreamSynchronize");
        if (st) return st;
        c->last.pairs_in_box = (int64_t)v;
        c->boxsum_pending = false;
    }
    *out = c->last;
    return OHX_OK;
}

}  // extern "C"
