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
#include "ohx_internal.h"

namespace ohx {

namespace gemm {

constexpr int BM = 128;           // tile rows (y)
constexpr int BN = 128;           // tile columns (x), complex
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

// pack two floats (round to nearest even) into one dword of two 16-bit values
template <bool FP16>
__device__ inline unsigned pack2(float a, float b)
{
    if (FP16) {
        const _Float16 ha = (_Float16)a, hb = (_Float16)b;
        return (unsigned)__builtin_bit_cast(unsigned short, ha) |
               ((unsigned)__builtin_bit_cast(unsigned short, hb) << 16);
    } else {
        const __bf16 ha = (__bf16)a, hb = (__bf16)b;
        return (unsigned)__builtin_bit_cast(unsigned short, ha) |
               ((unsigned)__builtin_bit_cast(unsigned short, hb) << 16);
    }
}
template <bool FP16>
__device__ inline float lo16(unsigned w)
{
    if (FP16) return (float)__builtin_bit_cast(_Float16, (unsigned short)(w & 0xffffu));
    return __builtin_bit_cast(float, w << 16);
}
template <bool FP16>
__device__ inline float hi16(unsigned w)
{
    if (FP16) return (float)__builtin_bit_cast(_Float16, (unsigned short)(w >> 16));
    return __builtin_bit_cast(float, w & 0xffff0000u);
}

template <int VAR>
__device__ inline f32x16 mfma(u32x4 a, u32x4 b, f32x16 c)
{
    if (Traits<VAR>::fp16)
        return __builtin_amdgcn_mfma_f32_32x32x16_f16(__builtin_bit_cast(f16x8, a), __builtin_bit_cast(f16x8, b),
                                                      c, 0, 0, 0);
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b),
                                                   c, 0, 0, 0);
}

template <bool THREE>
struct Smem {
    static constexpr int NH = THREE ? 2 : 1;
    // [buffer][hi/lo][row][dword]
    unsigned v[2][NH][BM][ROWW];
    unsigned u[2][NH][2 * BN][ROWW];
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

// generate one stage (KP points starting at p0, valid below p1) into buffer b
template <int VAR>
__device__ inline void generate(Smem<Traits<VAR>::three>& sm, int b, const Geometry& g, const PointSoA& s, int64_t p0, int64_t p1,
                                int x0, int y0, double xc, double yc, float vscale)
{
    constexpr bool FP16 = Traits<VAR>::fp16;
    constexpr bool THREE = Traits<VAR>::three;
    const int t = threadIdx.x;
    const int p = t & (KP - 1), grp = t >> 4;
    const int64_t pi = p0 + p;
    float ac = 0.f, gx = 0.f, bc = 0.f, gy = 0.f, qf = 0.f, amp = 0.f;
    double a64 = 0, b64 = 0, gx64 = 0, gy64 = 0, q64 = 0;
    bool big = false;
    int4 box = make_int4(0, 0, 0, 0);
    if (pi < p1) {
        const double px = s.px[pi], py = s.py[pi], q = s.q[pi];
        box = s.box[pi];
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
    // V rows: r = grp + 16 i ; U columns: c = grp + 16 i
#pragma unroll
    for (int i = 0; i < BM / 16; ++i) {
        const int r = grp + 16 * i;
        const int row = y0 + r;
        const double vd = -(double)(r - BM / 2) * g.pp_y;
        const float v = (float)vd;
        float f;
        if (big) {
            double ph = b64 + vd * (gy64 + q64 * vd);
            f = (float)(ph - floor(ph));
        } else {
            f = __builtin_amdgcn_fractf(fmaf(fmaf(qf, v, gy), v, bc));
        }
        const bool in = row >= box.z && row < box.w;
        const float a = in ? amp : 0.f;
        const float vr = a * __builtin_amdgcn_cosf(f), vi = a * __builtin_amdgcn_sinf(f);
        const unsigned h = pack2<FP16>(vr, vi);
        sm.v[b][0][r][p] = h;
        if constexpr (THREE) sm.v[b][1][r][p] = pack2<FP16>(vr - lo16<FP16>(h), vi - hi16<FP16>(h));
    }
#pragma unroll
    for (int i = 0; i < BN / 16; ++i) {
        const int c = grp + 16 * i;
        const int col = x0 + c;
        const double ud = (double)(c - BN / 2) * g.pp_x;
        const float u = (float)ud;
        float f;
        if (big) {
            double ph = a64 + ud * (gx64 + q64 * ud);
            f = (float)(ph - floor(ph));
        } else {
            f = __builtin_amdgcn_fractf(fmaf(fmaf(qf, u, gx), u, ac));
        }
        const bool in = col >= box.x && col < box.y;
        const float m = in ? 1.f : 0.f;
        const float ur = m * __builtin_amdgcn_cosf(f), ui = m * __builtin_amdgcn_sinf(f);
        // column 2c (Re S): (Re U, -Im U); column 2c+1 (Im S): (Im U, Re U)
        const unsigned h0 = pack2<FP16>(ur, -ui), h1 = pack2<FP16>(ui, ur);
        sm.u[b][0][2 * c][p] = h0;
        sm.u[b][0][2 * c + 1][p] = h1;
        if constexpr (THREE) {
            const float lr = ur - lo16<FP16>(h0), li = ui - lo16<FP16>(h1);
            sm.u[b][1][2 * c][p] = pack2<FP16>(lr, -li);
            sm.u[b][1][2 * c + 1][p] = pack2<FP16>(li, lr);
        }
    }
}

template <int VAR>
__global__ __launch_bounds__(NT) void k_fresnel_gemm(Geometry g, PointSoA s, float* __restrict__ part,
                                                     int tiles_x, int tiles, int64_t kslice,
                                                     const float* __restrict__ amax)
{
    constexpr bool THREE = Traits<VAR>::three;
    __shared__ __align__(16) Smem<THREE> sm;

    const int tile = blockIdx.x % tiles;
    const int split = blockIdx.x / tiles;
    const int tx = tile % tiles_x, ty = tile / tiles_x;
    const int x0 = tx * BN, y0 = ty * BM;
    const int64_t kb = (int64_t)split * kslice;
    const int64_t ke = min(s.n, kb + kslice);
    const double xc = gpix_x(g, x0 + BN / 2), yc = gpix_y(g, y0 + BM / 2);
    const float am = amax[0];
    const float vscale = (Traits<VAR>::fp16 && am > 0.f) ? 1.0f / am : 1.0f;

    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    f32x16 acc[4][2];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = f32x16{};

    const int nst = (int)((ke - kb + KP - 1) / KP);
    if (nst > 0) generate<VAR>(sm, 0, g, s, kb, ke, x0, y0, xc, yc, vscale);
    __syncthreads();
    const int arow = lane & 31, koff = (lane >> 5) * 4;   // dword offset of this lane's 8 values
    for (int st = 0; st < nst; ++st) {
        const int b = st & 1;
        if (st + 1 < nst) generate<VAR>(sm, b ^ 1, g, s, kb + (int64_t)(st + 1) * KP, ke, x0, y0, xc, yc, vscale);
#pragma unroll
        for (int ks = 0; ks < 2; ++ks) {
            const int kd = ks * 8 + koff;
            u32x4 ah[4], al[4], bh[2], bl[2];
#pragma unroll
            for (int mb = 0; mb < 4; ++mb) {
                ah[mb] = *reinterpret_cast<const u32x4*>(&sm.v[b][0][mb * 32 + arow][kd]);
                if constexpr (THREE) al[mb] = *reinterpret_cast<const u32x4*>(&sm.v[b][1][mb * 32 + arow][kd]);
            }
#pragma unroll
            for (int nb = 0; nb < 2; ++nb) {
                const int col = wave * 64 + nb * 32 + arow;
                bh[nb] = *reinterpret_cast<const u32x4*>(&sm.u[b][0][col][kd]);
                if constexpr (THREE) bl[nb] = *reinterpret_cast<const u32x4*>(&sm.u[b][1][col][kd]);
            }
#pragma unroll
            for (int mb = 0; mb < 4; ++mb)
#pragma unroll
                for (int nb = 0; nb < 2; ++nb) {
                    acc[mb][nb] = mfma<VAR>(ah[mb], bh[nb], acc[mb][nb]);
                    if constexpr (THREE) {
                        acc[mb][nb] = mfma<VAR>(ah[mb], bl[nb], acc[mb][nb]);
                        acc[mb][nb] = mfma<VAR>(al[mb], bh[nb], acc[mb][nb]);
                    }
                }
        }
        __syncthreads();
    }

    // write FP32 partial sums S (interleaved re, im) of this split
    float* out = part + (int64_t)split * g.pn_y * (int64_t)g.pn_x * 2;
#pragma unroll
    for (int mb = 0; mb < 4; ++mb)
#pragma unroll
        for (int nb = 0; nb < 2; ++nb) {
            const int np = wave * 64 + nb * 32 + (lane & 31);
            const int col = x0 + (np >> 1);
            if (col >= g.pn_x) continue;
#pragma unroll
            for (int v = 0; v < 16; ++v) {
                const int row = y0 + mb * 32 + 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3);
                if (row < g.pn_y) out[((int64_t)row * g.pn_x + col) * 2 + (np & 1)] = acc[mb][nb][v];
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

__global__ __launch_bounds__(256) void k_amp_max(const double* __restrict__ kamp, int64_t n, float* amax)
{
    float m = 0.f;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256)
        m = fmaxf(m, fabsf((float)kamp[i]));
    for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off));
    if ((threadIdx.x & 63) == 0 && m > 0.f) atomicMax((unsigned*)amax, __float_as_uint(m));
}

__global__ void k_amp_init(float* amax)
{
    if (threadIdx.x == 0) amax[0] = 0.f;
}

}  // namespace gemm

int launch_fresnel_gemm(ohx_ctx* ctx, const Geometry& g, const PointSoA& soa, double* d_field, uint32_t mode)
{
    using namespace gemm;
    const int var = (mode & OHX_MODE_MFMA_FP16X3) ? V_FP16X3 : (mode & OHX_MODE_MFMA_BF16X3) ? V_BF16X3 : V_BF16X1;
    const int tiles_x = (g.pn_x + BN - 1) / BN, tiles_y = (g.pn_y + BM - 1) / BM;
    const int tiles = tiles_x * tiles_y;
    // split K until there are at least two workgroups per CU (one resides per CU)
    int splits = (512 + tiles - 1) / tiles;
    const int64_t max_by_k = (soa.n + 4 * KP - 1) / (4 * KP);   // >= 4 stages per split
    if (splits > max_by_k) splits = (int)(max_by_k > 0 ? max_by_k : 1);
    if (splits < 1) splits = 1;
    int64_t kslice = (soa.n + splits - 1) / splits;
    kslice = (kslice + KP - 1) / KP * KP;
    splits = (int)((soa.n + kslice - 1) / kslice);
    if (splits < 1) splits = 1;
    const int64_t npx = (int64_t)g.pn_x * g.pn_y;
    const size_t part_bytes = (size_t)splits * npx * 2 * sizeof(float);
    int st = ensure_gemm_scratch(ctx, part_bytes + 256);
    if (st) return st;
    float* amax = (float*)ctx->gemm_scratch;
    float* part = (float*)((char*)ctx->gemm_scratch + 256);
    hipLaunchKernelGGL(k_amp_init, dim3(1), dim3(64), 0, ctx->stream, amax);
    hipLaunchKernelGGL(k_amp_max, dim3(256), dim3(256), 0, ctx->stream, (const double*)soa.kamp, soa.n, amax);
    const size_t shmem = 0;
    const dim3 grid((unsigned)(tiles * splits)), block(NT);
    switch (var) {
    case V_BF16X1:
        hipLaunchKernelGGL(k_fresnel_gemm<V_BF16X1>, grid, block, shmem, ctx->stream, g, soa, part, tiles_x, tiles,
                           kslice, (const float*)amax);
        break;
    case V_BF16X3:
        hipLaunchKernelGGL(k_fresnel_gemm<V_BF16X3>, grid, block, shmem, ctx->stream, g, soa, part, tiles_x, tiles,
                           kslice, (const float*)amax);
        break;
    default:
        hipLaunchKernelGGL(k_fresnel_gemm<V_FP16X3>, grid, block, shmem, ctx->stream, g, soa, part, tiles_x, tiles,
                           kslice, (const float*)amax);
        break;
    }
    st = check_hip(ctx, hipGetLastError(), "k_fresnel_gemm");
    if (st) return st;
    int64_t rg = (npx + 255) / 256;
    if (rg > 65536) rg = 65536;
    hipLaunchKernelGGL(k_gemm_reduce, dim3((unsigned)rg), dim3(256), 0, ctx->stream, (const float2*)part, splits,
                       npx, (double2*)d_field, (const float*)amax, var == V_FP16X3 ? 1 : 0);
    ctx->last.workgroups = (int64_t)tiles * splits;
    ctx->last.kernel_kind = 2;
    return check_hip(ctx, hipGetLastError(), "k_gemm_reduce");
}


}  // namespace ohx
