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
//    point order within a cell); a workgroup per 8 x 8 pixel tile stages
//    the points that reach its tile in LDS, in sorted order, and each pixel
//    sums those whose window covers it (k_wrp_gather), with the window table
//    in LDS.  The gather writes the plane samples of the zero-padded,
//    centre-shifted FFT input directly, k_wrp_pad_zero the rest.
//  * hipFFT Z2Z forward / inverse, each as two passes of contiguous 1D
//    transforms with a tiled LDS transpose between them; the input is held
//    transposed so that the pass along columns comes first and covers only
//    the columns that hold the plane (forward) or that the crop reads
//    (inverse).  k_wrp_asm multiplies by the transfer function at the
//    shifted frequency index, k_wrp_crop undoes the output shift while
//    cropping and transposing back.
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

// start[c] = first sorted position whose key is >= c, c in [0, ncells]:
// the points of cell c are [start[c], start[c + 1]) (binary search over the
// sorted keys; the off-plane key ncells sorts last)
__global__ __launch_bounds__(256) void k_wrp_cell_start(const uint32_t* __restrict__ key_s, int64_t n,
                                                       uint32_t ncells, uint32_t* __restrict__ start)
{
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c > (int64_t)ncells) return;
    int64_t lo = 0, hi = n;
    while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if (key_s[mid] < (uint32_t)c) lo = mid + 1;
        else hi = mid;
    }
    start[c] = (uint32_t)lo;
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

// Padded FFT input: sample (a, b) of the L x M grid (L = 2 pn_y, M = 2 pn_x)
// holds padded pixel (Y, X) = ((a + pn_y) mod L, (b + pn_x) mod M), i.e. WRP
// pixel (Y - pn_y/2, X - pn_x/2) when that lies on the plane, else zero.  It
// is stored transposed, T[b][a] (M rows of L samples), so that the first
// (column) pass of the FFT runs over contiguous rows.  Nonzero samples have
// a in S(L, cy) and b in S(M, cx), where S(n, c) = [n - c, n) u [0, c) and
// cx = ceil(pn_x / 2), cy = ceil(pn_y / 2); position q in [0, 2c) of S(n, c)
// is index sel(q) below.
__device__ __forceinline__ int sel(int q, int n, int c) { return q < c ? n - c + q : q - c; }

// zero samples of the rows T[b], b in S(M, cx) (the rows the first pass
// reads); k_wrp_gather writes the plane samples
__global__ __launch_bounds__(256) void k_wrp_pad_zero(int pn_x, int pn_y, double2* __restrict__ ft)
{
    const int L = 2 * pn_y, M = 2 * pn_x, cx = (pn_x + 1) / 2;
    const int a = blockIdx.x * 256 + threadIdx.x;
    const int q = blockIdx.y;
    if (a >= L || q >= 2 * cx) return;
    const int b = sel(q, M, cx);
    const int Y = a + pn_y < L ? a + pn_y : a + pn_y - L;
    const int X = b + pn_x < M ? b + pn_x : b + pn_x - M;
    const int row = Y - (pn_y >> 1), col = X - (pn_x >> 1);
    if (row >= 0 && row < pn_y && col >= 0 && col < pn_x) return;
    ft[(int64_t)b * L + a] = make_double2(0.0, 0.0);
}

// 32 x 32 tiles through LDS, 256 threads (32 x 8)
constexpr int TT = 32;

// A[a][b] = T[b][a] for b in S(M, cx), 0 elsewhere (after the first pass)
__global__ __launch_bounds__(256) void k_wrp_t2a(const double2* __restrict__ ft, double2* __restrict__ fa, int L,
                                                int M, int cx)
{
    __shared__ double2 s[TT][TT + 1];
    const int a0 = blockIdx.x * TT, b0 = blockIdx.y * TT;
    const int tx = threadIdx.x & (TT - 1), ty = threadIdx.x / TT;
    for (int i = ty; i < TT; i += 256 / TT) {
        const int b = b0 + i, a = a0 + tx;
        double2 v = make_double2(0.0, 0.0);
        if (b < M && a < L && (b >= M - cx || b < cx)) v = ft[(int64_t)b * L + a];
        s[i][tx] = v;
    }
    __syncthreads();
    for (int j = ty; j < TT; j += 256 / TT) {
        const int a = a0 + j, b = b0 + tx;
        if (a < L && b < M) fa[(int64_t)a * M + b] = s[tx][j];
    }
}

// T[b][a] = A[a][b] for b in S(M, cx), all a (before the last pass)
__global__ __launch_bounds__(256) void k_wrp_a2t(const double2* __restrict__ fa, double2* __restrict__ ft, int L,
                                                int M, int cx)
{
    __shared__ double2 s[TT][TT + 1];
    const int a0 = blockIdx.x * TT, q0 = blockIdx.y * TT;
    const int tx = threadIdx.x & (TT - 1), ty = threadIdx.x / TT;
    for (int j = ty; j < TT; j += 256 / TT) {
        const int a = a0 + j, q = q0 + tx;
        if (a < L && q < 2 * cx) s[j][tx] = fa[(int64_t)a * M + sel(q, M, cx)];
    }
    __syncthreads();
    for (int i = ty; i < TT; i += 256 / TT) {
        const int q = q0 + i, a = a0 + tx;
        if (q < 2 * cx && a < L) ft[(int64_t)sel(q, M, cx) * L + a] = s[tx][i];
    }
}

constexpr int GT = 8;           // gather tile edge: 8 x 8 pixels, one pixel per lane
constexpr int GW = 4;           // wavefronts per tile, each adding every GW-th group of 4 kept points
constexpr int GCAP = 512;       // kept points staged in LDS per pass (+ one 256-point append + padding)
constexpr int GTAB_LDS = 3600;  // window tables up to this many entries (w <= 30, 56 KB) are staged in LDS

struct alignas(16) WrpCand {
    int tx, ty;   // window centre
    double a;     // amplitude
};

// One workgroup of GW wavefronts per 8 x 8 tile of the WRP plane.  The
// points whose window centre (tx, ty) can reach the tile,
// tx in [c0 - w + 1, c0 + 7 + w] (same for y), lie in a few rows of cells;
// each row of cells is one contiguous range of the cell-sorted points.  The
// workgroup scans those ranges 256 points at a time and keeps, in scan order
// (ballot + prefix counts), the points that reach the tile.  Wavefront q adds
// the kept points of groups q, q + GW, q + 2 GW, ... (4 points a group, in
// order) for all 64 pixels; the GW partial sums are then added in q order.
// The order of every addition is fixed by the sorted point order, so the
// result is the same on every launch.  The add loop is branch-free: a point
// that misses the lane's pixel adds amplitude 0 times table entry 0 (finite,
// as r >= |dz| > 0), which leaves the partial sum bit-unchanged.  Window
// table in LDS when TL (4 w^2 <= GTAB_LDS; dynamic LDS of 4 w^2 entries),
// else read from global memory.
template<bool TL>
__global__ __launch_bounds__(64 * GW) void k_wrp_gather(WrpGeom wg, const uint32_t* __restrict__ cell_start,
                                                       const int2* __restrict__ stxy,
                                                       const double* __restrict__ samp,
                                                       const double2* __restrict__ tab,
                                                       double2* __restrict__ fft_in, double2* __restrict__ wrp_out)
{
    extern __shared__ double2 s_tab[];
    __shared__ WrpCand s_c[GCAP + 64 * GW + 4 * GW];
    __shared__ int s_wcnt[GW];
    __shared__ double2 s_part[GW - 1][64];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int c0 = blockIdx.x * GT, r0 = blockIdx.y * GT;
    const int col = c0 + (lane & (GT - 1)), row = r0 + lane / GT;
    const int w = wg.w, n2 = 2 * w;
    const int txlo = c0 - w + 1, txhi = c0 + GT - 1 + w;
    const int tylo = r0 - w + 1, tyhi = r0 + GT - 1 + w;
    // cells hold window centres by (t + w) / cell, t + w >= 0
    const int cx0 = (c0 + 1) / wg.cell, cx1 = min((c0 + GT - 1 + n2) / wg.cell, wg.cells_x - 1);
    const int cy0 = (r0 + 1) / wg.cell, cy1 = min((r0 + GT - 1 + n2) / wg.cell, wg.cells_y - 1);
    uint32_t total = 0;
    for (int cy = cy0; cy <= cy1; ++cy)
        total += cell_start[cy * wg.cells_x + cx1 + 1] - cell_start[cy * wg.cells_x + cx0];
    if (total == 0) {   // uniform over the workgroup
        if (wave == 0 && row < wg.pn_y && col < wg.pn_x) {
            const double2 z = make_double2(0.0, 0.0);
            if (wrp_out) wrp_out[(int64_t)row * wg.pn_x + col] = z;
            const int Y = row + (wg.pn_y >> 1), X = col + (wg.pn_x >> 1);
            const int a = Y >= wg.pn_y ? Y - wg.pn_y : Y + wg.pn_y;
            const int b = X >= wg.pn_x ? X - wg.pn_x : X + wg.pn_x;
            fft_in[(int64_t)b * (2 * wg.pn_y) + a] = z;
        }
        return;
    }

    if constexpr (TL)
        for (int i = tid; i < n2 * n2; i += 64 * GW) s_tab[i] = tab[i];
    double2 acc = make_double2(0.0, 0.0);
    const int cx = col + w, cy = row + w;
    auto add = [&](const WrpCand& q) {
        const int wx = cx - q.tx, wy = cy - q.ty;
        const bool hit = (unsigned)wx < (unsigned)n2 && (unsigned)wy < (unsigned)n2;
        const int ti = hit ? wy * n2 + wx : 0;
        const double am = hit ? q.a : 0.0;
        double2 k;
        if constexpr (TL) k = s_tab[ti];
        else k = tab[ti];
        acc.x = fma(am, k.x, acc.x);
        acc.y = fma(am, k.y, acc.y);
    };
    auto consume = [&](int m) {
        // pad to whole groups of 4 with points no pixel reaches
        if (tid < 4) s_c[m + tid] = WrpCand{-(1 << 29), -(1 << 29), 0.0};
        __syncthreads();
        for (int j = 4 * wave; j < m; j += 4 * GW) {
            const WrpCand q0 = s_c[j], q1 = s_c[j + 1], q2 = s_c[j + 2], q3 = s_c[j + 3];
            add(q0);
            add(q1);
            add(q2);
            add(q3);
        }
        __syncthreads();
    };
    int cnt = 0;   // kept points staged; the same value in every lane
    for (int cyr = cy0; cyr <= cy1; ++cyr) {
        const uint32_t s = cell_start[cyr * wg.cells_x + cx0], e = cell_start[cyr * wg.cells_x + cx1 + 1];
        for (uint32_t base = s; base < e; base += 64 * GW) {
            const uint32_t p = base + tid;
            bool take = false;
            int2 t = make_int2(0, 0);
            if (p < e) {
                t = stxy[p];
                take = t.x >= txlo && t.x <= txhi && t.y >= tylo && t.y <= tyhi;
            }
            const unsigned long long m = __ballot(take);
            if (lane == 0) s_wcnt[wave] = __popcll(m);
            __syncthreads();
            int off = cnt, all = 0;
            for (int k = 0; k < GW; ++k) {
                const int ck = s_wcnt[k];
                off += k < wave ? ck : 0;
                all += ck;
            }
            if (take) s_c[off + __popcll(m & ((1ull << lane) - 1ull))] = WrpCand{t.x, t.y, samp[p]};
            cnt += all;
            __syncthreads();
            if (cnt >= GCAP) {
                consume(cnt);
                cnt = 0;
            }
        }
    }
    if (cnt > 0) consume(cnt);

    // partial sums of waves 1..GW-1 to wave 0, added in wave order
    if (wave > 0) s_part[wave - 1][lane] = acc;
    __syncthreads();
    if (wave == 0) {
        for (int q = 0; q < GW - 1; ++q) {
            const double2 v = s_part[q][lane];
            acc.x += v.x;
            acc.y += v.y;
        }
        if (row < wg.pn_y && col < wg.pn_x) {
            if (wrp_out) wrp_out[(int64_t)row * wg.pn_x + col] = acc;
            const int Y = row + (wg.pn_y >> 1), X = col + (wg.pn_x >> 1);
            const int a = Y >= wg.pn_y ? Y - wg.pn_y : Y + wg.pn_y;
            const int b = X >= wg.pn_x ? X - wg.pn_x : X + wg.pn_x;
            fft_in[(int64_t)b * (2 * wg.pn_y) + a] = acc;
        }
    }
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

// out[y][x] = G[(y + pn_y/2 + pn_y) mod L][(x + pn_x/2 + pn_x) mod M], with
// G stored transposed (T[b][a] = G[a][b])
__global__ __launch_bounds__(256) void k_wrp_crop(const double2* __restrict__ ft, int pn_x, int pn_y,
                                                 double2* __restrict__ out)
{
    __shared__ double2 s[TT][TT + 1];
    const int L = 2 * pn_y, M = 2 * pn_x;
    const int x0 = blockIdx.x * TT, y0 = blockIdx.y * TT;
    const int tx = threadIdx.x & (TT - 1), ty = threadIdx.x / TT;
    for (int i = ty; i < TT; i += 256 / TT) {
        const int x = x0 + i, y = y0 + tx;
        if (x < pn_x && y < pn_y) {
            int a = y + (pn_y >> 1) + pn_y, b = x + (pn_x >> 1) + pn_x;
            if (a >= L) a -= L;
            if (b >= M) b -= M;
            s[i][tx] = ft[(int64_t)b * L + a];
        }
    }
    __syncthreads();
    for (int j = ty; j < TT; j += 256 / TT) {
        const int y = y0 + j, x = x0 + tx;
        if (x < pn_x && y < pn_y) out[(int64_t)y * pn_x + x] = s[tx][j];
    }
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
    const int cx = (pn_x + 1) / 2;   // the columns b in S(M, cx) hold the plane
    hipStream_t st = c->stream;
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
    int key_bits = 1;   // keys are in [0, ncells]
    while (key_bits < 32 && (ncells >> key_bits) != 0) ++key_bits;
    size_t sort_bytes = 0;
    if (hipcub::DeviceRadixSort::SortPairs(nullptr, sort_bytes, (uint32_t*)nullptr, (uint32_t*)nullptr,
                                           (uint32_t*)nullptr, (uint32_t*)nullptr, (int)np, 0, key_bits, st) !=
        hipSuccess)
        return set_error(c, OHX_ERR_DEVICE, "WRP sort size query failed");
    const size_t b_fft = align256((size_t)L * M * sizeof(double2));
    const size_t b_tab = align256((size_t)std::max(1, 4 * w * w) * sizeof(double2));
    const size_t b_u32 = align256((size_t)np * 4);
    const size_t b_i2 = align256((size_t)np * 8);
    const size_t b_cells = align256((size_t)(ncells + 1) * 4);
    const size_t b_tmp = align256(sort_bytes);
    const size_t total = 2 * b_fft + b_tab + 4 * b_u32 + 2 * b_i2 + 2 * b_i2 + b_cells + b_tmp;
    if ((st_ = ensure_gemm_scratch(c, total))) return st_;
    char* q = (char*)c->gemm_scratch;
    double2* ft = (double2*)q;               q += b_fft;   // transposed grid T[b][a]
    double2* fa = (double2*)q;               q += b_fft;   // grid A[a][b]
    double2* tab = (double2*)q;              q += b_tab;
    uint32_t* key = (uint32_t*)q;            q += b_u32;
    uint32_t* key_s = (uint32_t*)q;          q += b_u32;
    uint32_t* idx = (uint32_t*)q;            q += b_u32;
    uint32_t* idx_s = (uint32_t*)q;          q += b_u32;
    int2* txy = (int2*)q;                    q += b_i2;
    double* amp = (double*)q;                q += b_i2;
    int2* stxy = (int2*)q;                   q += b_i2;
    double* samp = (double*)q;               q += b_i2;
    uint32_t* start = (uint32_t*)q;          q += b_cells;
    void* tmp = (void*)q;

    if (w > 0 && n > 0) {
        const unsigned pb = (unsigned)((n + 255) / 256);
        hipLaunchKernelGGL(k_wrp_prep, dim3(pb), dim3(256), 0, st, (const double*)d_vertices, n, wg, key, idx, txy,
                           amp);
        if (hipcub::DeviceRadixSort::SortPairs(tmp, sort_bytes, key, key_s, idx, idx_s, (int)n, 0, key_bits, st) !=
            hipSuccess)
            return set_error(c, OHX_ERR_DEVICE, "WRP point sort failed");
        hipLaunchKernelGGL(k_wrp_cell_start, dim3((ncells + 1 + 255) / 256), dim3(256), 0, st,
                           (const uint32_t*)key_s, n, ncells, start);
        hipLaunchKernelGGL(k_wrp_permute, dim3(pb), dim3(256), 0, st, (const uint32_t*)idx_s, n,
                           (const int2*)txy, (const double*)amp, stxy, samp);
        const int nt = 4 * w * w;
        hipLaunchKernelGGL(k_wrp_table, dim3((nt + 255) / 256), dim3(256), 0, st, w, p->pp_x, p->pp_y, dz, k, tab);
        hipLaunchKernelGGL(k_wrp_pad_zero, dim3((unsigned)((L + 255) / 256), (unsigned)(2 * cx)), dim3(256), 0, st,
                           pn_x, pn_y, ft);
        const dim3 tgrid((unsigned)((pn_x + GT - 1) / GT), (unsigned)((pn_y + GT - 1) / GT));
        if (nt <= GTAB_LDS)
            hipLaunchKernelGGL(k_wrp_gather<true>, tgrid, dim3(64 * GW), (size_t)nt * sizeof(double2), st, wg,
                               (const uint32_t*)start, (const int2*)stxy, (const double*)samp, (const double2*)tab,
                               ft, (double2*)d_wrp);
        else
            hipLaunchKernelGGL(k_wrp_gather<false>, tgrid, dim3(64 * GW), 0, st, wg, (const uint32_t*)start,
                               (const int2*)stxy, (const double*)samp, (const double2*)tab, ft, (double2*)d_wrp);
    } else {
        // no window (w = 0) or no points: the WRP and everything after it are zero
        hipLaunchKernelGGL(k_wrp_zero2, dim3((unsigned)(((int64_t)L * M + 255) / 256)), dim3(256), 0, st, ft,
                           (int64_t)L * M);
        if (d_wrp)
            hipLaunchKernelGGL(k_wrp_zero2, dim3((unsigned)((npx + 255) / 256)), dim3(256), 0, st, (double2*)d_wrp,
                               npx);
    }
    if ((st_ = check_hip(c, hipGetLastError(), "WRP accumulation"))) return st_;

    // propagation: FFT, transfer function, inverse FFT, centre crop.  Each 2D
    // transform is two passes of contiguous 1D transforms with a transpose
    // between them, pruned to the rows that matter:
    //   forward: along a over the rows T[b], b in S(M, cx) (the others are
    //   zero); T -> A (zeros for the other b); along b over all rows of A;
    //   transfer function on A;
    //   inverse: along b over all rows of A; A -> T for b in S(M, cx) (the
    //   only columns the crop reads); along a over those rows; crop from T.
    if (!c->wrp_plan_t || c->wrp_plan_L != L || c->wrp_plan_M != M) {
        wrp_release(c);
        hipfftHandle ht, ha;
        int nl[1] = {L}, nm[1] = {M};
        if (hipfftPlanMany(&ht, 1, nl, nl, 1, L, nl, 1, L, HIPFFT_Z2Z, cx) != HIPFFT_SUCCESS)
            return set_error(c, OHX_ERR_DEVICE, "hipfftPlanMany (length 2 pn_y) failed");
        c->wrp_plan_t = (void*)ht;
        if (hipfftPlanMany(&ha, 1, nm, nm, 1, M, nm, 1, M, HIPFFT_Z2Z, L) != HIPFFT_SUCCESS)
            return set_error(c, OHX_ERR_DEVICE, "hipfftPlanMany (length 2 pn_x) failed");
        c->wrp_plan_a = (void*)ha;
        c->wrp_plan_L = L;
        c->wrp_plan_M = M;
    }
    const hipfftHandle ht = (hipfftHandle)c->wrp_plan_t, ha = (hipfftHandle)c->wrp_plan_a;
    if (hipfftSetStream(ht, st) != HIPFFT_SUCCESS || hipfftSetStream(ha, st) != HIPFFT_SUCCESS)
        return set_error(c, OHX_ERR_DEVICE, "hipfftSetStream failed");
    hipfftDoubleComplex* const t0 = (hipfftDoubleComplex*)ft;                    // rows T[0 .. cx)
    hipfftDoubleComplex* const t1 = t0 + (int64_t)(M - cx) * L;                  // rows T[M - cx .. M)
    hipfftDoubleComplex* const a0 = (hipfftDoubleComplex*)fa;
    auto tpass = [&](int dir) {
        return hipfftExecZ2Z(ht, t1, t1, dir) == HIPFFT_SUCCESS && hipfftExecZ2Z(ht, t0, t0, dir) == HIPFFT_SUCCESS;
    };
    const dim3 t2a_grid((unsigned)((L + TT - 1) / TT), (unsigned)((M + TT - 1) / TT));
    const dim3 a2t_grid((unsigned)((L + TT - 1) / TT), (unsigned)((2 * cx + TT - 1) / TT));
    if (!tpass(HIPFFT_FORWARD)) return set_error(c, OHX_ERR_DEVICE, "forward FFT failed");
    hipLaunchKernelGGL(k_wrp_t2a, t2a_grid, dim3(256), 0, st, (const double2*)ft, fa, L, M, cx);
    if (hipfftExecZ2Z(ha, a0, a0, HIPFFT_FORWARD) != HIPFFT_SUCCESS)
        return set_error(c, OHX_ERR_DEVICE, "forward FFT failed");
    hipLaunchKernelGGL(k_wrp_asm, pgrid, dim3(256), 0, st, fa, pn_x, pn_y, p->pp_x, p->pp_y, p->lambda,
                       p->propagation_distance);
    if (hipfftExecZ2Z(ha, a0, a0, HIPFFT_BACKWARD) != HIPFFT_SUCCESS)
        return set_error(c, OHX_ERR_DEVICE, "inverse FFT failed");
    hipLaunchKernelGGL(k_wrp_a2t, a2t_grid, dim3(256), 0, st, (const double2*)fa, ft, L, M, cx);
    if (!tpass(HIPFFT_BACKWARD)) return set_error(c, OHX_ERR_DEVICE, "inverse FFT failed");
    hipLaunchKernelGGL(k_wrp_crop, dim3((unsigned)((pn_x + TT - 1) / TT), (unsigned)((pn_y + TT - 1) / TT)),
                       dim3(256), 0, st, (const double2*)ft, pn_x, pn_y, (double2*)d_field);
    return check_hip(c, hipGetLastError(), "WRP propagation");
}

void wrp_release(ohx_ctx* c)
{
    if (c->wrp_plan_t) hipfftDestroy((hipfftHandle)c->wrp_plan_t);
    if (c->wrp_plan_a) hipfftDestroy((hipfftHandle)c->wrp_plan_a);
    c->wrp_plan_t = nullptr;
    c->wrp_plan_a = nullptr;
    c->wrp_plan_L = c->wrp_plan_M = 0;
}

}  // namespace ohx
