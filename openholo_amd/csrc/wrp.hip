// WRP (wavefront recording plane) generator for gfx950, FP64 throughout.
//
// Replaces ophWRP::calculateWRPCPU / calculateWRPGPU (ophgen/src/ophWRP.cpp:
// 321-435, ophgen/src/ophWRP_GPU.cpp:48-200, ophWRPKernel.cu:170-275) and
// the propagation ophGen::Fresnel_FFT (ophgen/src/ophGen.cpp:516-581).
//
// Reference algorithm, per channel:
//  * every point adds, on a (2w) x (2w) window of the WRP centred on pixel
//    (tx, ty) = ((int)(x/pp_x) + pn_x/2, (int)(y/pp_y) + pn_y/2), the terms
//    (a cos(kr) cos(2pi) / r, -a sin(kr) sin(2pi) / r),
//    r = +-sqrt(dx^2 + dy^2 + dz^2), dx = wx pp_x, dy = wy pp_y, with ONE
//    depth dz = wrp_location - zmax for all points and
//    w = (int)(|lambda dz / (2 pp_x^2)| * 2);
//  * the WRP is zero-padded to 2 pn_x x 2 pn_y, FFT'd with centring shifts,
//    multiplied by exp(j 2pi d sqrt(1/lambda^2 - fx^2 - fy^2)), inverse FFT'd
//    (unnormalised) with centring shifts, and the centre is cropped.
//
// Design:
//  * Since dz is common, the window terms depend on (wx, wy) only: they are
//    tabulated once per launch (k_wrp_table), and a point's contribution to a
//    pixel is amplitude x table entry.
//  * Deterministic gather instead of the reference's atomics: points are
//    sorted by the cell of their window centre (hipCUB radix sort, stable:
//    point order within a cell), and each padded-plane pixel sums, in cell
//    then point order, the points whose window covers it (k_wrp_gather).  The
//    gather writes the zero-padded, centre-shifted FFT input directly.
//  * hipFFT Z2Z forward / inverse in place; k_wrp_asm multiplies by the
//    transfer function at the shifted frequency index, k_wrp_crop undoes the
//    output shift while cropping.
// Shifts are out of place (the reference's Openholo::fft2 calls fftShift
// with src == dst there, Openholo/src/Openholo.cpp:716, which overwrites half
// of every quadrant swap; the shifts here do what that function intends).
#include "ohx_internal.h"

#include <hipcub/hipcub.hpp>
#include <hipfft/hipfft.h>

#include <cmath>

namespace ohx {

namespace {

constexpr double MIN_DOUBLE = 2.2250738585072014e-308;   // Openholo/src/define.h:148
constexpr double PI = 3.14159265358979323846;

struct WrpGeom {
    int pn_x, pn_y;
    double pp_x, pp_y;
    double scale[3];
    int channel;
    int w;                 // window half width (pixels)
    int cell;              // cell edge of the point grid (pixels)
    int cells_x, cells_y;  // cell grid over window centres tx in [-w, pn_x + w)
};

// ---------------------------------------------------------------------------
// zmax = max(MIN_DOUBLE, max_i z_i * scale_z) (ophWRP.cpp:84-93)
__global__ __launch_bounds__(256) void k_wrp_zmax_part(const double* __restrict__ vtx, int64_t n, double sz,
                                                      double* __restrict__ part)
{
    __shared__ double sh[256];
    double m = MIN_DOUBLE;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
        const double z = vtx[7 * i + 2] * sz;
        if (m < z) m = z;
    }
    sh[threadIdx.x] = m;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s && sh[threadIdx.x] < sh[threadIdx.x + s]) sh[threadIdx.x] = sh[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

__global__ __launch_bounds__(256) void k_wrp_zmax_final(const double* __restrict__ part, int nparts,
                                                       double* __restrict__ out)
{
    __shared__ double sh[256];
    double m = MIN_DOUBLE;
    for (int i = threadIdx.x; i < nparts; i += 256)
        if (m < part[i]) m = part[i];
    sh[threadIdx.x] = m;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s && sh[threadIdx.x] < sh[threadIdx.x + s]) sh[threadIdx.x] = sh[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = sh[0];
}

// ---------------------------------------------------------------------------
// per point: window centre, amplitude and cell key (cells_x * cells_y for a
// window that cannot touch the plane: sorted to the end and never read)
__global__ __launch_bounds__(256) void k_wrp_prep(const double* __restrict__ vtx, int64_t n, WrpGeom wg,
                                                 uint32_t* __restrict__ key, uint32_t* __restrict__ idx,
                                                 int2* __restrict__ txy, double* __restrict__ amp)
{
#pragma clang fp contract(off)
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double* v = vtx + 7 * i;
    const double qx = (v[0] * wg.scale[0]) / wg.pp_x;   // (int)(x / ppX), ophWRP.cpp:371
    const double qy = (v[1] * wg.scale[1]) / wg.pp_y;
    uint32_t k = (uint32_t)(wg.cells_x * wg.cells_y);
    int tx = 0, ty = 0;
    if (fabs(qx) < 1.0e9 && fabs(qy) < 1.0e9) {
        tx = (int)qx + (wg.pn_x >> 1);
        ty = (int)qy + (wg.pn_y >> 1);
        if (tx >= -wg.w && tx < wg.pn_x + wg.w && ty >= -wg.w && ty < wg.pn_y + wg.w)
            k = (uint32_t)(((ty + wg.w) / wg.cell) * wg.cells_x + (tx + wg.w) / wg.cell);
    }
    key[i] = k;
    idx[i] = (uint32_t)i;
    txy[i] = make_int2(tx, ty);
    amp[i] = v[3 + wg.channel];
}

__global__ __launch_bounds__(256) void k_wrp_count(const uint32_t* __restrict__ key, int64_t n, uint32_t ncells,
                                                  uint32_t* __restrict__ count)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && key[i] < ncells) atomicAdd(&count[key[i]], 1u);
}

__global__ __launch_bounds__(256) void k_wrp_zero_u32(uint32_t* __restrict__ p, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

// point data in cell order
__global__ __launch_bounds__(256) void k_wrp_permute(const uint32_t* __restrict__ order, int64_t n,
                                                    const int2* __restrict__ txy, const double* __restrict__ amp,
                                                    int2* __restrict__ stxy, double* __restrict__ samp)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t p = order[i];
    stxy[i] = txy[p];
    samp[i] = amp[p];
}

// window terms without the amplitude: (cos(kr) cos(2pi) / r, -sin(kr) sin(2pi) / r)
// at (wx, wy) in [-w, w)^2, table index (wy + w) * 2w + (wx + w)
__global__ __launch_bounds__(256) void k_wrp_table(int w, double pp_x, double pp_y, double dz, double k,
                                                  double2* __restrict__ tab)
{
#pragma clang fp contract(off)
    const int n = 2 * w;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n * n) return;
    const int wy = i / n - w, wx = i % n - w;
    const double dy = wy * pp_y, dx = wx * pp_x;
    const double dyy = dy * dy, dzz = dz * dz;
    const double s = sqrt(dx * dx + dyy + dzz);
    const double r = (dz > 0.0) ? s : -s;
    const double kr = k * r;
    const double rp = 2 * PI * 1.0;   // random phase off: randVal = 1 (ophWRP.cpp:399-400)
    tab[i] = make_double2((cos(kr) * cos(rp)) / r, (-sin(kr) * sin(rp)) / r);
}

// One thread per sample of the shifted, zero-padded FFT input (L = 2 pn_y
// rows, M = 2 pn_x columns).  Sample (a, b) holds padded pixel
// (Y, X) = ((a + pn_y) mod L, (b + pn_x) mod M), i.e. WRP pixel
// (Y - pn_y/2, X - pn_x/2) when that lies on the plane, else zero.
__global__ __launch_bounds__(256) void k_wrp_gather(WrpGeom wg, const uint32_t* __restrict__ cell_start,
                                                   const int2* __restrict__ stxy, const double* __restrict__ samp,
                                                   const double2* __restrict__ tab, double2* __restrict__ fft_in,
                                                   double2* __restrict__ wrp_out)
{
    const int L = 2 * wg.pn_y, M = 2 * wg.pn_x;
    const int b = blockIdx.x * 64 + (threadIdx.x & 63);
    const int a = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (a >= L || b >= M) return;
    const int Y = a + wg.pn_y < L ? a + wg.pn_y : a + wg.pn_y - L;
    const int X = b + wg.pn_x < M ? b + wg.pn_x : b + wg.pn_x - M;
    const int row = Y - (wg.pn_y >> 1), col = X - (wg.pn_x >> 1);
    double2 acc = make_double2(0.0, 0.0);
    if (row >= 0 && row < wg.pn_y && col >= 0 && col < wg.pn_x) {
        const int w = wg.w, n = 2 * w;
        // window centres tx in [col - w + 1, col + w], i.e. tx + w in [col + 1, col + 2w]
        const int cx0 = (col + 1) / wg.cell, cx1 = min((col + n) / wg.cell, wg.cells_x - 1);
        const int cy0 = (row + 1) / wg.cell, cy1 = min((row + n) / wg.cell, wg.cells_y - 1);
        for (int cy = cy0; cy <= cy1; ++cy) {
            for (int cx = cx0; cx <= cx1; ++cx) {
                const int c = cy * wg.cells_x + cx;
                const uint32_t p0 = cell_start[c], p1 = cell_start[c + 1];
                for (uint32_t p = p0; p < p1; ++p) {
                    const int2 t = stxy[p];
                    const int wx = col - t.x, wy = row - t.y;
                    if (wx >= -w && wx < w && wy >= -w && wy < w) {
                        const double2 k = tab[(wy + w) * n + (wx + w)];
                        const double am = samp[p];
                        acc.x = fma(am, k.x, acc.x);
                        acc.y = fma(am, k.y, acc.y);
                    }
                }
            }
        }
        if (wrp_out) wrp_out[(int64_t)row * wg.pn_x + col] = acc;
    }
    fft_in[(int64_t)a * M + b] = acc;
}

// F[u][v] *= exp(j 2pi d sqrt(1/lambda^2 - fx_j^2 - fy_i^2)) with the
// centred frequency index (i, j) = ((u + pn_y) mod L, (v + pn_x) mod M) and
// fx_j = (j - pn_x) / (2 ss_x) (ophGen.cpp:544-566, same operation order)
__global__ __launch_bounds__(256) void k_wrp_asm(double2* __restrict__ f, int pn_x, int pn_y, double pp_x, double pp_y,
                                                double lambda, double distance)
{
#pragma clang fp contract(off)
    const int L = 2 * pn_y, M = 2 * pn_x;
    const int v = blockIdx.x * 64 + (threadIdx.x & 63);
    const int u = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (u >= L || v >= M) return;
    const int i = u + pn_y < L ? u + pn_y : u + pn_y - L;
    const int j = v + pn_x < M ? v + pn_x : v + pn_x - M;
    const double ss_x2 = (pn_x * pp_x) * 2, ss_y2 = (pn_y * pp_y) * 2;
    double fx = (j - pn_x) / ss_x2;
    fx *= fx;
    double fy = (i - pn_y) / ss_y2;
    fy *= fy;
    const double sqrt_part = sqrt(1 / (lambda * lambda) - fx - fy);
    const double th = (2 * PI * distance) * sqrt_part;
    double sn, cs;
    sincos(th, &sn, &cs);
    const int64_t id = (int64_t)u * M + v;
    const double2 z = f[id];
    f[id] = make_double2(z.x * cs - z.y * sn, z.x * sn + z.y * cs);
}

// out[y][x] = G[(y + pn_y/2 + pn_y) mod L][(x + pn_x/2 + pn_x) mod M]
__global__ __launch_bounds__(256) void k_wrp_crop(const double2* __restrict__ g, int pn_x, int pn_y,
                                                 double2* __restrict__ out)
{
    const int L = 2 * pn_y, M = 2 * pn_x;
    const int x = blockIdx.x * 64 + (threadIdx.x & 63);
    const int y = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (y >= pn_y || x >= pn_x) return;
    int Y = y + (pn_y >> 1) + pn_y, X = x + (pn_x >> 1) + pn_x;
    if (Y >= L) Y -= L;
    if (X >= M) X -= M;
    out[(int64_t)y * pn_x + x] = g[(int64_t)Y * M + X];
}

__global__ __launch_bounds__(256) void k_wrp_zero2(double2* __restrict__ p, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = make_double2(0.0, 0.0);
}

size_t align256(size_t b) { return (b + 255) / 256 * 256; }

}  // namespace

// ---------------------------------------------------------------------------
int wrp_generate(ohx_ctx* c, const void* d_vertices, int64_t n, const ohx_wrp_params* p, double* d_field,
                 double* d_wrp)
{
    const int pn_x = p->pn_x, pn_y = p->pn_y;
    const int64_t npx = (int64_t)pn_x * pn_y;
    const int L = 2 * pn_y, M = 2 * pn_x;
    hipStream_t st = c->stream;
    const unsigned gx = (unsigned)((pn_x + 63) / 64), gy = (unsigned)((pn_y + 3) / 4);
    const dim3 pgrid((unsigned)((M + 63) / 64), (unsigned)((L + 3) / 4));

    // zmax and the window
    WrpGeom wg{};
    wg.pn_x = pn_x;
    wg.pn_y = pn_y;
    wg.pp_x = p->pp_x;
    wg.pp_y = p->pp_y;
    for (int i = 0; i < 3; ++i) wg.scale[i] = p->scale[i];
    wg.channel = p->channel;
    double zmax = MIN_DOUBLE;
    const int zparts = (int)std::min<int64_t>(1024, std::max<int64_t>(1, (n + 255) / 256));
    int st_ = ensure_gemm_scratch(c, 256 + align256((size_t)zparts * 8));
    if (st_) return st_;
    if (n > 0) {
        double* zpart = (double*)((char*)c->gemm_scratch + 256);
        double* zout = (double*)c->gemm_scratch;
        hipLaunchKernelGGL(k_wrp_zmax_part, dim3(zparts), dim3(256), 0, st, (const double*)d_vertices, n,
                           p->scale[2], zpart);
        hipLaunchKernelGGL(k_wrp_zmax_final, dim3(1), dim3(256), 0, st, (const double*)zpart, zparts, zout);
        if ((st_ = check_hip(c, hipMemcpyAsync(&zmax, zout, 8, hipMemcpyDeviceToHost, st), "zmax copy"))) return st_;
        if ((st_ = check_hip(c, hipStreamSynchronize(st), "zmax sync"))) return st_;
    }
    const double dz = p->wrp_location - zmax;
    const double k = (2 * PI) / p->lambda;
    const double ppxx = p->pp_x * p->pp_x * 2;
    const double tw = fabs(p->lambda * dz / ppxx) * 2;
    if (!(tw < 4096.0)) return set_error(c, OHX_ERR_INVALID_ARG, "WRP window wider than 8192 pixels");
    wg.w = (int)tw;
    c->last = ohx_launch_stats{};
    c->last.kernel_kind = 3;
    c->last.pairs_nominal = n * (int64_t)(2 * wg.w) * (2 * wg.w);
    c->last.workgroups = (int64_t)pgrid.x * pgrid.y;

    // scratch: FFT buffer, window table, point arrays, cell offsets, sort storage
    const int w = wg.w;
    wg.cell = std::max(2 * w, 16);
    wg.cells_x = (pn_x + 2 * w + wg.cell - 1) / wg.cell;
    wg.cells_y = (pn_y + 2 * w + wg.cell - 1) / wg.cell;
    const uint32_t ncells = (uint32_t)wg.cells_x * (uint32_t)wg.cells_y;
    const int64_t np = std::max<int64_t>(n, 1);
    size_t sort_bytes = 0, scan_bytes = 0;
    hipcub::DeviceRadixSort::SortPairs(nullptr, sort_bytes, (uint32_t*)nullptr, (uint32_t*)nullptr,
                                       (uint32_t*)nullptr, (uint32_t*)nullptr, (int)np, 0, 32, st);
    hipcub::DeviceScan::ExclusiveSum(nullptr, scan_bytes, (uint32_t*)nullptr, (uint32_t*)nullptr,
                                     (int)(ncells + 1), st);
    const size_t b_fft = align256((size_t)L * M * sizeof(double2));
    const size_t b_tab = align256((size_t)std::max(1, 4 * w * w) * sizeof(double2));
    const size_t b_u32 = align256((size_t)np * 4);
    const size_t b_i2 = align256((size_t)np * 8);
    const size_t b_cells = align256((size_t)(ncells + 1) * 4);
    const size_t b_tmp = align256(std::max(sort_bytes, scan_bytes));
    const size_t total = b_fft + b_tab + 4 * b_u32 + 2 * b_i2 + 2 * b_i2 + 2 * b_cells + b_tmp;
    if ((st_ = ensure_gemm_scratch(c, total))) return st_;
    char* q = (char*)c->gemm_scratch;
    double2* fft = (double2*)q;              q += b_fft;
    double2* tab = (double2*)q;              q += b_tab;
    uint32_t* key = (uint32_t*)q;            q += b_u32;
    uint32_t* key_s = (uint32_t*)q;          q += b_u32;
    uint32_t* idx = (uint32_t*)q;            q += b_u32;
    uint32_t* idx_s = (uint32_t*)q;          q += b_u32;
    int2* txy = (int2*)q;                    q += b_i2;
    double* amp = (double*)q;                q += b_i2;
    int2* stxy = (int2*)q;                   q += b_i2;
    double* samp = (double*)q;               q += b_i2;
    uint32_t* count = (uint32_t*)q;          q += b_cells;
    uint32_t* start = (uint32_t*)q;          q += b_cells;
    void* tmp = (void*)q;

    if (w > 0 && n > 0) {
        const unsigned pb = (unsigned)((n + 255) / 256);
        hipLaunchKernelGGL(k_wrp_prep, dim3(pb), dim3(256), 0, st, (const double*)d_vertices, n, wg, key, idx, txy,
                           amp);
        if (hipcub::DeviceRadixSort::SortPairs(tmp, sort_bytes, key, key_s, idx, idx_s, (int)n, 0, 32, st) !=
            hipSuccess)
            return set_error(c, OHX_ERR_DEVICE, "WRP point sort failed");
        hipLaunchKernelGGL(k_wrp_zero_u32, dim3((ncells + 1 + 255) / 256), dim3(256), 0, st, count,
                           (int64_t)ncells + 1);
        hipLaunchKernelGGL(k_wrp_count, dim3(pb), dim3(256), 0, st, (const uint32_t*)key, n, ncells, count);
        if (hipcub::DeviceScan::ExclusiveSum(tmp, scan_bytes, count, start, (int)(ncells + 1), st) != hipSuccess)
            return set_error(c, OHX_ERR_DEVICE, "WRP cell scan failed");
        hipLaunchKernelGGL(k_wrp_permute, dim3(pb), dim3(256), 0, st, (const uint32_t*)idx_s, n,
                           (const int2*)txy, (const double*)amp, stxy, samp);
        const int nt = 4 * w * w;
        hipLaunchKernelGGL(k_wrp_table, dim3((nt + 255) / 256), dim3(256), 0, st, w, p->pp_x, p->pp_y, dz, k, tab);
        hipLaunchKernelGGL(k_wrp_gather, pgrid, dim3(256), 0, st, wg, (const uint32_t*)start,
                           (const int2*)stxy, (const double*)samp, (const double2*)tab, fft, (double2*)d_wrp);
    } else {
        // no window (w = 0) or no points: the WRP and everything after it are zero
        hipLaunchKernelGGL(k_wrp_zero2, dim3((unsigned)(((int64_t)L * M + 255) / 256)), dim3(256), 0, st, fft,
                           (int64_t)L * M);
        if (d_wrp)
            hipLaunchKernelGGL(k_wrp_zero2, dim3((unsigned)((npx + 255) / 256)), dim3(256), 0, st, (double2*)d_wrp,
                               npx);
    }
    if ((st_ = check_hip(c, hipGetLastError(), "WRP accumulation"))) return st_;

    // propagation: FFT, transfer function, inverse FFT, centre crop
    if (!c->wrp_plan || c->wrp_plan_L != L || c->wrp_plan_M != M) {
        if (c->wrp_plan) hipfftDestroy((hipfftHandle)c->wrp_plan);
        c->wrp_plan = nullptr;
        hipfftHandle h;
        if (hipfftPlan2d(&h, L, M, HIPFFT_Z2Z) != HIPFFT_SUCCESS)
            return set_error(c, OHX_ERR_DEVICE, "hipfftPlan2d failed");
        c->wrp_plan = (void*)h;
        c->wrp_plan_L = L;
        c->wrp_plan_M = M;
    }
    hipfftHandle h = (hipfftHandle)c->wrp_plan;
    if (hipfftSetStream(h, st) != HIPFFT_SUCCESS) return set_error(c, OHX_ERR_DEVICE, "hipfftSetStream failed");
    if (hipfftExecZ2Z(h, (hipfftDoubleComplex*)fft, (hipfftDoubleComplex*)fft, HIPFFT_FORWARD) != HIPFFT_SUCCESS)
        return set_error(c, OHX_ERR_DEVICE, "forward FFT failed");
    hipLaunchKernelGGL(k_wrp_asm, pgrid, dim3(256), 0, st, fft, pn_x, pn_y, p->pp_x, p->pp_y, p->lambda,
                       p->propagation_distance);
    if (hipfftExecZ2Z(h, (hipfftDoubleComplex*)fft, (hipfftDoubleComplex*)fft, HIPFFT_BACKWARD) != HIPFFT_SUCCESS)
        return set_error(c, OHX_ERR_DEVICE, "inverse FFT failed");
    hipLaunchKernelGGL(k_wrp_crop, dim3(gx, gy), dim3(256), 0, st, (const double2*)fft, pn_x, pn_y,
                       (double2*)d_field);
    return check_hip(c, hipGetLastError(), "WRP propagation");
}

void wrp_release(ohx_ctx* c)
{
    if (c->wrp_plan) hipfftDestroy((hipfftHandle)c->wrp_plan);
    c->wrp_plan = nullptr;
}

}  // namespace ohx
