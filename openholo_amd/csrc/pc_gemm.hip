// Fresnel point-cloud sum as a complex GEMM on gfx950 matrix cores.
//
// The Fresnel term of ophGen::Fresnel_Diffraction (ophgen/src/ophGen.cpp:447-514)
// factors exactly: with alpha_p(x) = (x - x_p)^2 q_p, beta_p(y) = (y - y_p)^2 q_p,
// q_p = 1/(2 lambda z_p) and c_p = z_p/lambda (+ phase_p/2pi) in cycles,
//   H[y, x] = -j sum_p  V[y, p] U[x, p]
//   V[y, p] = A_p exp(j 2pi (beta_p(y) + c_p)) * [y in box_p],   A_p = a_p/(lambda z_p)
//   U[x, p] = exp(j 2pi alpha_p(x))           * [x in box_p]
// (the reference's anti-aliasing box is a product of an x and a y interval).
// M = rows (y), N = columns (x), K = points: a complex GEMM with K ~ 1e6.
//
// Realification: k' = 2p + t (t=0: Re V, t=1: Im V), n' = 2x + s, and
//   B'[2p, 2x] = Re U,  B'[2p+1, 2x] = -Im U,  B'[2p, 2x+1] = Im U,  B'[2p+1, 2x+1] = Re U
// so C = A' B' is [Re S, Im S] interleaved per pixel.  v_mfma_f32_32x32x16_{bf16,f16}
// then does 8 points x 16 output columns' worth of complex MACs per instruction.
//
// Operands are never stored in HBM: per 16-point stage every thread generates
// 8 V and 8 U elements (FP64 per-point constants recentred on the tile centre,
// FP32 phase offsets < ~20 cycles, v_sin/v_cos) into double-buffered LDS while
// the waves run the previous stage's MFMAs.  Precision variants:
//   BF16x1 : hi(V) hi(U)
//   BF16x3 : hi hi + hi lo + lo hi   (lo = bf16(x - hi): ~2^-17 per factor)
//   FP16x3 : same split in FP16 (~2^-22), V scaled by 1/max|A| to stay in range
// Split-K slices write FP32 partials; k_gemm_reduce sums them in FP64 in a fixed
// order (deterministic) and adds -j S into the FP64 field.
#include <algorithm>

#include "ohx_internal.h"

namespace ohx {

namespace gemm {

constexpr int BM = 128;           // tile rows (y)
constexpr int BN = 128;           // tile columns (x), complex; wave w owns 32 of them
constexpr int NT = 256;           // 4 waves
constexpr int KP = 16;            // points per stage
constexpr int KS = 2 * KP;        // realified K per stage (32)
constexpr int LDK = KS + 8;       // padded row length (bf16/f16 elements): 80 B rows
constexpr int ROWW = LDK / 2;     // row length in dwords (20)

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

enum { V_BF16X1 = 0, V_BF16X3 = 1, V_FP16X3 = 2 };

// largest phase excursion (cycles) over a tile generated with FP32 offsets
constexpr double MAX_SPAN = 64.0;

template <int VAR>
struct Traits {
    static constexpr bool three = VAR != V_BF16X1;
    static constexpr bool fp16 = VAR == V_FP16X3;
};

// pack two floats into one dword of two 16-bit values (BF16: round to nearest
// even, v_cvt_pk_bf16_f32; FP16: round toward zero, v_cvt_pkrtz_f16_f32 -- the
// 3-term split keeps the residual either way)
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
template <bool FP16>
__device__ inline unsigned pack2(float a, float b)
{
    if constexpr (FP16) {
        return __builtin_bit_cast(unsigned, __builtin_amdgcn_cvt_pkrtz(a, b));
    } else {
        const bf16x2 v = {(__bf16)a, (__bf16)b};
        return __builtin_bit_cast(unsigned, v);
    }
}
template <bool FP16>
__device__ inline float lo16(unsigned w)
{
    if constexpr (FP16) return (float)__builtin_bit_cast(_Float16, (unsigned short)(w & 0xffffu));
    return __builtin_bit_cast(float, w << 16);
}
template <bool FP16>
__device__ inline float hi16(unsigned w)
{
    if constexpr (FP16) return (float)__builtin_bit_cast(_Float16, (unsigned short)(w >> 16));
    return __builtin_bit_cast(float, w & 0xffff0000u);
}

template <int VAR>
__device__ inline f32x16 mfma(u32x4 a, u32x4 b, f32x16 c)
{
    if constexpr (Traits<VAR>::fp16)
        return __builtin_amdgcn_mfma_f32_32x32x16_f16(__builtin_bit_cast(f16x8, a), __builtin_bit_cast(f16x8, b),
                                                      c, 0, 0, 0);
    else
        return __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, a),
                                                       __builtin_bit_cast(bf16x8, b), c, 0, 0, 0);
}

// LDS: V[y][p] and U[x][p] as packed (Re, Im) 16-bit pairs, hi and (3-term) lo
// parts, double buffered; rows padded to 80 bytes (conflict-free ds_read_b128).
template <bool THREE>
struct Smem {
    static constexpr int NH = THREE ? 2 : 1;
    unsigned v[2][NH][BM][ROWW];
    unsigned u[2][NH][BN][ROWW];
};

__device__ inline double gpix_x(const Geometry& g, int col)
{
#pragma clang fp contract(off)
    return g.x0 + (double)(col - 1 + g.off_x) * g.pp_x;
}
__device__ inline double gpix_y(const Geometry& g, int row)
{
#pragma clang fp contract(off)
    return g.y0 + (double)(g.pn_y - row + g.off_y) * g.pp_y;
}

// One stage: KP points from p0 (valid below p1) into buffer b.  Thread t
// handles point t % 16, V rows and U columns t/16 + 16 i (i < 8).
template <int VAR>
__device__ inline void generate(Smem<Traits<VAR>::three>& sm, int b, const Geometry& g, const PointSoA& s,
                                int64_t p0, int64_t p1, int x0, int y0, double xc, double yc, float vscale,
                                float v0, float u0, float ppy16, float ppx16)
{
    constexpr bool FP16 = Traits<VAR>::fp16;
    constexpr bool THREE = Traits<VAR>::three;
    const int t = threadIdx.x;
    const int p = t & (KP - 1), grp = t >> 4;
    const int64_t pi = p0 + p;
    float ac = 0.f, gx = 0.f, bc = 0.f, gy = 0.f, qf = 0.f, amp = 0.f;
    double a64 = 0, b64 = 0, gx64 = 0, gy64 = 0, q64 = 0;
    bool big = false;
    int rlo = 0, rhi = 0, clo = 0, chi = 0;
    if (pi < p1) {
        const double px = s.px[pi], py = s.py[pi], q = s.q[pi];
        const int4 box = s.box[pi];
        rlo = box.z - y0;
        rhi = box.w - y0;
        clo = box.x - x0;
        chi = box.y - x0;
        const double dxc = xc - px, dyc = yc - py;
        a64 = dxc * dxc * q;
        a64 -= floor(a64);
        b64 = dyc * dyc * q + s.cyc0[pi];
        b64 -= floor(b64);
        gx64 = 2.0 * dxc * q;
        gy64 = 2.0 * dyc * q;
        q64 = q;
        ac = (float)a64;
        bc = (float)b64;
        gx = (float)gx64;
        gy = (float)gy64;
        qf = (float)q;
        amp = (float)(s.kamp[pi]) * vscale;
        // phase excursion over the tile; beyond MAX_SPAN cycles FP32 offsets
        // lose accuracy, so this point's elements are phased in FP64
        const double um = (BN / 2) * g.pp_x, vm = (BM / 2) * g.pp_y;
        big = (fabs(gx64) * um + q * um * um + fabs(gy64) * vm + q * vm * vm) > MAX_SPAN;
    }
#pragma unroll
    for (int i = 0; i < BM / 16; ++i) {
        const int r = grp + 16 * i;
        const float v = fmaf(-(float)i, ppy16, v0);          // y_r - y_c
        float f;
        if (big) {
            const double vd = (double)(BM / 2 - r) * g.pp_y;
            const double ph = b64 + vd * (gy64 + q64 * vd);
            f = (float)(ph - floor(ph));
        } else {
            f = __builtin_amdgcn_fractf(fmaf(fmaf(qf, v, gy), v, bc));
        }
        const float a = (r >= rlo && r < rhi) ? amp : 0.f;
        const float vr = a * __builtin_amdgcn_cosf(f), vi = a * __builtin_amdgcn_sinf(f);
        const unsigned h = pack2<FP16>(vr, vi);
        sm.v[b][0][r][p] = h;
        if constexpr (THREE) sm.v[b][1][r][p] = pack2<FP16>(vr - lo16<FP16>(h), vi - hi16<FP16>(h));
    }
#pragma unroll
    for (int i = 0; i < BN / 16; ++i) {
        const int c = grp + 16 * i;
        const float u = fmaf((float)i, ppx16, u0);            // x_c - x_c(tile)
        float f;
        if (big) {
            const double ud = (double)(c - BN / 2) * g.pp_x;
            const double ph = a64 + ud * (gx64 + q64 * ud);
            f = (float)(ph - floor(ph));
        } else {
            f = __builtin_amdgcn_fractf(fmaf(fmaf(qf, u, gx), u, ac));
        }
        const bool in = c >= clo && c < chi;
        const float ur = in ? __builtin_amdgcn_cosf(f) : 0.f;
        const float ui = in ? __builtin_amdgcn_sinf(f) : 0.f;
        const unsigned h = pack2<FP16>(ur, ui);
        sm.u[b][0][c][p] = h;
        if constexpr (THREE) sm.u[b][1][c][p] = pack2<FP16>(ur - lo16<FP16>(h), ui - hi16<FP16>(h));
    }
}

// B' fragments from one U fragment: Re block (Re U, -Im U), Im block (Im U, Re U)
__device__ inline u32x4 to_re(u32x4 w) { return w ^ 0x80000000u; }
__device__ inline u32x4 to_im(u32x4 w)
{
    u32x4 r;
#pragma unroll
    for (int i = 0; i < 4; ++i) r[i] = __builtin_amdgcn_alignbit(w[i], w[i], 16);
    return r;
}

template <int VAR>
__global__ __launch_bounds__(NT, 2) void k_fresnel_gemm(Geometry g, PointSoA s, float2* __restrict__ part,
                                                        int tiles_x, int tiles, int64_t kfirst, int64_t klast,
                                                        int64_t kslice, const float* __restrict__ amax)
{
    constexpr bool THREE = Traits<VAR>::three;
    __shared__ __align__(16) Smem<THREE> sm;

    const int tile = blockIdx.x % tiles;
    const int split = blockIdx.x / tiles;
    const int tx = tile % tiles_x, ty = tile / tiles_x;
    const int x0 = tx * BN, y0 = ty * BM;
    const int64_t kb = kfirst + (int64_t)split * kslice;
    const int64_t ke = min(klast, kb + kslice);
    const double xc = gpix_x(g, x0 + BN / 2), yc = gpix_y(g, y0 + BM / 2);
    const float am = amax[0];
    const float vscale = (Traits<VAR>::fp16 && am > 0.f) ? 1.0f / am : 1.0f;
    const int grp = threadIdx.x >> 4;
    const float v0 = (float)((double)(BM / 2 - grp) * g.pp_y);
    const float u0 = (float)((double)(grp - BN / 2) * g.pp_x);
    const float ppy16 = (float)(16.0 * g.pp_y), ppx16 = (float)(16.0 * g.pp_x);

    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    f32x16 acc[4][2];
#pragma unroll
    for (int i = 0; i < 4; ++i) acc[i][0] = acc[i][1] = f32x16{};

    const int nst = (int)((ke - kb + KP - 1) / KP);
    if (nst > 0) generate<VAR>(sm, 0, g, s, kb, ke, x0, y0, xc, yc, vscale, v0, u0, ppy16, ppx16);
    __syncthreads();
    const int lrow = lane & 31, koff = (lane >> 5) * 4;   // dword offset of this lane's 8 values
    const int ucol = wave * 32 + lrow;                      // this wave's 32 columns
    for (int st = 0; st < nst; ++st) {
        const int b = st & 1;
        if (st + 1 < nst)
            generate<VAR>(sm, b ^ 1, g, s, kb + (int64_t)(st + 1) * KP, ke, x0, y0, xc, yc, vscale, v0, u0,
                          ppy16, ppx16);
#pragma unroll
        for (int ks = 0; ks < 2; ++ks) {
            const int kd = ks * 8 + koff;
            u32x4 ah[4], al[4];
#pragma unroll
            for (int mb = 0; mb < 4; ++mb) {
                ah[mb] = *reinterpret_cast<const u32x4*>(&sm.v[b][0][mb * 32 + lrow][kd]);
                if constexpr (THREE) al[mb] = *reinterpret_cast<const u32x4*>(&sm.v[b][1][mb * 32 + lrow][kd]);
            }
            const u32x4 uh = *reinterpret_cast<const u32x4*>(&sm.u[b][0][ucol][kd]);
            const u32x4 bhr = to_re(uh), bhi = to_im(uh);
            u32x4 blr, bli;
            if constexpr (THREE) {
                const u32x4 ul = *reinterpret_cast<const u32x4*>(&sm.u[b][1][ucol][kd]);
                blr = to_re(ul);
                bli = to_im(ul);
            }
#pragma unroll
            for (int mb = 0; mb < 4; ++mb) {
                acc[mb][0] = mfma<VAR>(ah[mb], bhr, acc[mb][0]);
                acc[mb][1] = mfma<VAR>(ah[mb], bhi, acc[mb][1]);
                if constexpr (THREE) {
                    acc[mb][0] = mfma<VAR>(ah[mb], blr, acc[mb][0]);
                    acc[mb][1] = mfma<VAR>(ah[mb], bli, acc[mb][1]);
                    acc[mb][0] = mfma<VAR>(al[mb], bhr, acc[mb][0]);
                    acc[mb][1] = mfma<VAR>(al[mb], bhi, acc[mb][1]);
                }
            }
        }
        __syncthreads();
    }

    // FP32 partial sums S of this split, interleaved (Re, Im) per pixel
    const int col = x0 + ucol;
    if (col < g.pn_x) {
        float2* out = part + (int64_t)split * g.pn_y * (int64_t)g.pn_x;
#pragma unroll
        for (int mb = 0; mb < 4; ++mb)
#pragma unroll
            for (int v = 0; v < 16; ++v) {
                const int row = y0 + mb * 32 + 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3);
                if (row < g.pn_y) out[(int64_t)row * g.pn_x + col] = make_float2(acc[mb][0][v], acc[mb][1][v]);
            }
    }
}

// field += -j * amax * sum_split S   (FP64, fixed summation order)
__global__ __launch_bounds__(256) void k_gemm_reduce(const float2* __restrict__ part, int splits, int64_t npx,
                                                     double2* __restrict__ field, const float* __restrict__ amax,
                                                     int scaled)
{
    const double sc = (scaled && amax[0] > 0.f) ? (double)amax[0] : 1.0;
    int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int64_t stride = (int64_t)gridDim.x * 256;
    for (; i < npx; i += stride) {
        double sr = 0, si = 0;
        for (int k = 0; k < splits; ++k) {
            const float2 v = part[(int64_t)k * npx + i];
            sr += v.x;
            si += v.y;
        }
        double2 f = field[i];
        f.x += si * sc;
        f.y -= sr * sc;
        field[i] = f;
    }
}

__global__ __launch_bounds__(256) void k_amp_max(const double* __restrict__ kamp, int64_t p0, int64_t p1,
                                                 float* amax)
{
    float m = 0.f;
    for (int64_t i = p0 + (int64_t)blockIdx.x * 256 + threadIdx.x; i < p1; i += (int64_t)gridDim.x * 256)
        m = fmaxf(m, fabsf((float)kamp[i]));
    for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off));
    if ((threadIdx.x & 63) == 0 && m > 0.f) atomicMax((unsigned*)amax, __float_as_uint(m));
}

__global__ void k_amp_init(float* amax)
{
    if (threadIdx.x == 0) amax[0] = 0.f;
}

}  // namespace gemm

namespace {
struct GemmPlan {
    int tiles_x, tiles;
    int64_t splits, kslice, npx;
};
GemmPlan plan_gemm(const Geometry& g, int64_t npoints)
{
    using namespace gemm;
    GemmPlan p;
    p.tiles_x = (g.pn_x + BN - 1) / BN;
    p.tiles = p.tiles_x * ((g.pn_y + BM - 1) / BM);
    p.npx = (int64_t)g.pn_x * g.pn_y;
    // Split K so that (a) every CU gets several workgroups (two reside per CU)
    // and (b) no FP32 accumulator sums more than MAX_SLICE points: FP32
    // rounding of the running sums grows like sqrt(points per slice).
    constexpr int64_t MAX_SLICE = 65536;
    const int64_t n = std::max<int64_t>(npoints, 1);
    int64_t splits = std::max<int64_t>((1024 + p.tiles - 1) / p.tiles, (n + MAX_SLICE - 1) / MAX_SLICE);
    const int64_t max_by_k = std::max<int64_t>(1, (n + 4 * KP - 1) / (4 * KP));  // >= 4 stages per split
    splits = std::min(splits, max_by_k);
    int64_t kslice = (n + splits - 1) / splits;
    kslice = (kslice + KP - 1) / KP * KP;
    p.splits = std::max<int64_t>(1, (n + kslice - 1) / kslice);
    p.kslice = kslice;
    return p;
}
}  // namespace

size_t fresnel_gemm_scratch_bytes(const Geometry& g, int64_t npoints)
{
    const GemmPlan p = plan_gemm(g, npoints);
    return 256 + (size_t)p.splits * p.npx * sizeof(float2);
}

int launch_fresnel_gemm_range(hipStream_t stream, void* scratch, size_t scratch_bytes, const Geometry& g,
                              const PointSoA& soa, int64_t p0, int64_t p1, double* d_field, uint32_t mode)
{
    using namespace gemm;
    if (p1 <= p0) return OHX_OK;
    const int var = (mode & OHX_MODE_MFMA_FP16X3) ? V_FP16X3 : (mode & OHX_MODE_MFMA_BF16X3) ? V_BF16X3 : V_BF16X1;
    const GemmPlan pl = plan_gemm(g, p1 - p0);
    if (256 + (size_t)pl.splits * pl.npx * sizeof(float2) > scratch_bytes) return OHX_ERR_INVALID_ARG;
    float* amax = (float*)scratch;
    float2* part = (float2*)((char*)scratch + 256);
    hipLaunchKernelGGL(k_amp_init, dim3(1), dim3(64), 0, stream, amax);
    hipLaunchKernelGGL(k_amp_max, dim3(256), dim3(256), 0, stream, (const double*)soa.kamp, p0, p1, amax);
    const dim3 grid((unsigned)((int64_t)pl.tiles * pl.splits)), block(NT);
    switch (var) {
    case V_BF16X1:
        hipLaunchKernelGGL(k_fresnel_gemm<V_BF16X1>, grid, block, 0, stream, g, soa, part, pl.tiles_x, pl.tiles,
                           p0, p1, pl.kslice, (const float*)amax);
        break;
    case V_BF16X3:
        hipLaunchKernelGGL(k_fresnel_gemm<V_BF16X3>, grid, block, 0, stream, g, soa, part, pl.tiles_x, pl.tiles,
                           p0, p1, pl.kslice, (const float*)amax);
        break;
    default:
        hipLaunchKernelGGL(k_fresnel_gemm<V_FP16X3>, grid, block, 0, stream, g, soa, part, pl.tiles_x, pl.tiles,
                           p0, p1, pl.kslice, (const float*)amax);
        break;
    }
    if (hipGetLastError() != hipSuccess) return OHX_ERR_DEVICE;
    int64_t rg = (pl.npx + 255) / 256;
    if (rg > 65536) rg = 65536;
    hipLaunchKernelGGL(k_gemm_reduce, dim3((unsigned)rg), dim3(256), 0, stream, (const float2*)part, (int)pl.splits,
                       pl.npx, (double2*)d_field, (const float*)amax, var == V_FP16X3 ? 1 : 0);
    return hipGetLastError() == hipSuccess ? OHX_OK : OHX_ERR_DEVICE;
}

int launch_fresnel_gemm(ohx_ctx* ctx, const Geometry& g, const PointSoA& soa, double* d_field, uint32_t mode)
{
    const size_t bytes = fresnel_gemm_scratch_bytes(g, soa.n);
    int st = ensure_gemm_scratch(ctx, bytes);
    if (st) return st;
    st = launch_fresnel_gemm_range(ctx->stream, ctx->gemm_scratch, ctx->gemm_scratch_bytes, g, soa, 0, soa.n,
                                   d_field, mode);
    const GemmPlan pl = plan_gemm(g, soa.n);
    ctx->last.workgroups = (int64_t)pl.tiles * pl.splits;
    ctx->last.kernel_kind = 2;
    if (st) return set_error(ctx, st, "matrix-core Fresnel launch failed");
    return OHX_OK;
}

}  // namespace ohx
