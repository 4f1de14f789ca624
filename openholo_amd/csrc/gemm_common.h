// Shared pieces of the matrix-core kernels (Fresnel and RS): tile shape, the
// 16-bit operand packing and split, the MFMA wrapper, the LDS operand layout
// and the pixel grid of the reference (ophgen/src/ophGen.cpp:413,422).
#pragma once

#include "ohx_internal.h"

namespace ohx {
namespace gemm {

constexpr int BM = 128;           // tile rows (y)
constexpr int BN = 128;           // tile columns (x), complex; wave w owns 32 of them
constexpr int NT = 256;           // 4 waves; two workgroups per CU (80 KB LDS each)
constexpr int KP = 16;            // points per stage
constexpr int KS = 2 * KP;        // realified K per stage (32)
constexpr int LDK = KS + 8;       // padded row length (bf16/f16 elements): 80 B rows
constexpr int ROWW = LDK / 2;     // row length in dwords (20)

// Tile shape of the Fresnel kernel: BM rows x W columns, W / 32 waves (wave w
// owns columns 32w .. 32w+31 and all BM rows), 16 points per stage with one
// point per lane of each 16-lane thread group.  W = 128: 4 waves, two
// workgroups per CU; W = 256: 8 waves, one workgroup per CU, and a quarter
// fewer generated elements per point-pixel pair.
template <int W>
struct Shape {
    static constexpr int BN = W;
    static constexpr int NT = 2 * W;             // 64 lanes per 32 columns
    static constexpr int NGRP = NT / KP;         // thread groups of 16 lanes
    static constexpr int NVT = BM / NGRP;        // V rows per thread and stage: grp + NGRP i
    static constexpr int NUT = W / NGRP;         // U columns per thread and stage: grp + NGRP j
    static constexpr int NE = NVT + NUT;         // elements per thread and stage
    static constexpr int PLANE_U = W * ROWW;     // dwords of one U plane
};

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

enum { V_BF16X1 = 0, V_BF16X3 = 1, V_FP16X3 = 2 };

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
    // the low half moved up by a byte permute: a shift would let the compiler
    // re-derive it from a second conversion of the low operand alone
    return __builtin_bit_cast(float, __builtin_amdgcn_perm(w, 0u, 0x05040c0cu));
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
template <bool THREE, int W = BN>
struct Smem {
    static constexpr int NH = THREE ? 2 : 1;
    unsigned v[2][NH][BM][ROWW];
    unsigned u[2][NH][W][ROWW];
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

// x - floor(x) in one v_fract_f64 (clamped below 1.0)
__device__ inline double frac64(double x) { return __builtin_amdgcn_fract(x); }

// Per-point records of a launch (64 bytes each) are read with buffer loads at
// a 32-bit byte offset: at most MAX_REC records per launch.
constexpr int64_t MAX_REC = int64_t(1) << 24;
using Rsrc = __amdgpu_buffer_rsrc_t;
__device__ inline Rsrc rec_rsrc(const void* base, int bytes)
{
    // raw (untyped) buffer; 0x20000: the 32-bit data format, which raw
    // loads on gfx9 need to be non-zero
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, bytes, 0x00020000);
}
__device__ inline u32x4 rec_load16(Rsrc rs, int off) { return __builtin_amdgcn_raw_buffer_load_b128(rs, off, 0, 0); }

// LDS plane layout: plane h (hi / lo part) of one buffer's V block is BM rows
// of ROWW dwords (U block: BN rows); element (row, p) at h*PLANE + row*ROWW + p.
constexpr int PLANE_V = BM * ROWW;
constexpr int PLANE_U = BN * ROWW;

// B' fragments from one U fragment: Re block (Re U, -Im U), Im block (Im U, Re U)
__device__ inline u32x4 to_re(u32x4 w) { return w ^ 0x80000000u; }
__device__ inline u32x4 to_im(u32x4 w)
{
    u32x4 r;
#pragma unroll
    for (int i = 0; i < 4; ++i) r[i] = __builtin_amdgcn_alignbit(w[i], w[i], 16);
    return r;
}

// ---- MFMA side of one pipelined stage ----------------------------------------
// C holds the stage's operand registers: acc[4][2] (Re / Im accumulator blocks
// of the wave's 128 x 32 output), A[8][2] (V fragments, hi / lo), Ub[2][2]
// (U fragments), B[2][4] (B' fragments), and the lane's fragment coordinates
// lrow, koff, ucol.
// MFMA order per (ks, mb) group: 3-term (ah,Re) (ah,Im) (ah,lo Re) (ah,lo Im)
// (al,Re) (al,Im); 1-term (ah,Re) (ah,Im).  Wave w owns output columns
// 32w .. 32w+31 of the tile (Re and Im blocks) and all 128 rows (4 blocks).
template <int VAR, class C>
__device__ inline void load_a(const unsigned* rv, C& c, int g)
{
    const unsigned* a = rv + ((g & 3) * 32 + c.lrow) * ROWW + (g >> 2) * 8 + c.koff;
    c.A[g][0] = *reinterpret_cast<const u32x4*>(a);
    if constexpr (Traits<VAR>::three) c.A[g][1] = *reinterpret_cast<const u32x4*>(a + PLANE_V);
}
template <int VAR, int PU = PLANE_U, class C>
__device__ inline void load_b(const unsigned* ru, C& c, int ks)
{
    const unsigned* b = ru + c.ucol * ROWW + ks * 8 + c.koff;
    c.Ub[ks][0] = *reinterpret_cast<const u32x4*>(b);
    if constexpr (Traits<VAR>::three) c.Ub[ks][1] = *reinterpret_cast<const u32x4*>(b + PU);
}
template <class C>
__device__ inline void xform_b(C& c, int ks, int h)
{
    c.B[ks][2 * h] = to_re(c.Ub[ks][h]);
    c.B[ks][2 * h + 1] = to_im(c.Ub[ks][h]);
}

// MFMAs per stage and wave
template <int VAR>
constexpr int mfma_per_stage()
{
    return Traits<VAR>::three ? 48 : 16;
}
// MFMAs of a stage's last (ks, mb) group, which load no operands
template <int VAR>
constexpr int tail_mfmas()
{
    return Traits<VAR>::three ? 6 : 2;
}

// The first operands of a stage, read by the previous stage right after its
// barrier, which comes within that stage's last MFMA group (late-barrier
// pipeline); the fields load_a / load_b use.
struct Carry {
    u32x4 A[8][2], Ub[2][2];
    int lrow, koff, ucol;
};

// Slot S of a stage: MFMA S and the operand loads / transforms due there.
template <int VAR, int S, int PU = PLANE_U, class C>
__device__ inline void mfma_slot(const unsigned* rv, const unsigned* ru, C& c)
{
    constexpr bool THREE = Traits<VAR>::three;
    constexpr int NTG = THREE ? 6 : 2;       // MFMAs per (ks, mb) group
    constexpr int NM = 8 * NTG;
    constexpr int g = S / NTG, t = S % NTG, ks = g >> 2, mb = g & 3;
    c.acc[mb][t & 1] = mfma<VAR>(c.A[g][t < 4 ? 0 : 1], c.B[ks][t < 4 ? t : t - 4], c.acc[mb][t & 1]);
    if constexpr (t == 0 && THREE && g + 1 < 8) load_a<VAR>(rv, c, g + 1);
    if constexpr (t == 0 && !THREE && g + 2 < 8) load_a<VAR>(rv, c, g + 2);
    if constexpr (S == NM / 4) load_b<VAR, PU>(ru, c, 1);
    if constexpr (S == NM * 3 / 8) xform_b(c, 1, 0);
    if constexpr (THREE && S == NM * 3 / 8 + 2) xform_b(c, 1, 1);
}

// Sign of split s's accumulation: the matrix cores' FP32 accumulation rounds
// with a bias toward -inf (measured: the error of long sums has a constant
// negative part in both components, the same in every split), so odd splits
// accumulate the negated products (their amplitudes negated), the split-K
// reduction subtracts their planes, and the bias of each pair of splits
// cancels.
__device__ inline float split_sign(int split) { return (split & 1) ? -1.f : 1.f; }

// FP32 partial sums of one split, interleaved (Re, Im) per pixel, into plane
// `split` of part (pn_y x pn_x float2 per plane).
__device__ inline void store_partials(const Geometry& g, const f32x16 (&acc)[4][2], float2* __restrict__ part,
                                      int split, int x0, int y0, int ucol, int lane)
{
    const int col = x0 + ucol;
    if (col >= g.pn_x) return;
    float2* out = part + (int64_t)split * g.pn_y * (int64_t)g.pn_x;
#pragma unroll
    for (int mb = 0; mb < 4; ++mb)
#pragma unroll
        for (int v = 0; v < 16; ++v) {
            const int row = y0 + mb * 32 + 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3);
            if (row < g.pn_y) out[(int64_t)row * g.pn_x + col] = make_float2(acc[mb][0][v], acc[mb][1][v]);
        }
}

}  // namespace gemm

// ---- host side: tiles and split-K plan ---------------------------------------
struct GemmPlan {
    int tiles_x, tiles;
    int64_t splits, kslice, npx;
};
// Split K so that (a) the workgroups fill the device's resident slots in whole
// rounds (256 CUs x one 8-wave workgroup of a 256-column tile, or two 4-wave
// ones of a 128-column tile): a launch's last round otherwise runs part-empty,
// and (b) no FP32 accumulator sums more than GEMM_MAX_SLICE points: FP32
// rounding of the running sums grows like sqrt(points per slice).  Among the
// split counts from one round of workgroups to about four, the plan takes the
// one with the least time, counted as rounds x points per split plus
// GEMM_SPLIT_COST points for each split's partial plane (its write and its
// reduction), the smaller count on ties; every split keeps >= 4 stages.
constexpr int64_t GEMM_MAX_SLICE = 65536;
constexpr int64_t GEMM_SPLIT_COST = 128;
inline int64_t gemm_slots(int bn) { return bn > gemm::BN ? 256 : 512; }
inline int64_t gemm_tiles(const Geometry& g, int bn)
{
    return (int64_t)((g.pn_x + bn - 1) / bn) * ((g.pn_y + gemm::BM - 1) / gemm::BM);
}
// the most splits plan_gemm chooses for any launch of at most npoints points
// (each bound below grows with the point count)
inline int64_t gemm_max_splits(const Geometry& g, int64_t npoints, int kq, int bn)
{
    const int64_t n = npoints > 1 ? npoints : 1, tiles = gemm_tiles(g, bn);
    const int64_t round1 = (gemm_slots(bn) + tiles - 1) / tiles;
    const int64_t max_by_k = std::max<int64_t>(1, (n + 4 * kq - 1) / (4 * kq));
    return std::min(std::max<int64_t>((n + GEMM_MAX_SLICE - 1) / GEMM_MAX_SLICE, 4 * round1), max_by_k);
}
// kq: K granularity of a split in points (one stage); bn: tile columns
inline GemmPlan plan_gemm(const Geometry& g, int64_t npoints, int kq, int bn = gemm::BN)
{
    GemmPlan p;
    p.tiles_x = (g.pn_x + bn - 1) / bn;
    p.tiles = (int)gemm_tiles(g, bn);
    p.npx = (int64_t)g.pn_x * g.pn_y;
    const int64_t n = npoints > 1 ? npoints : 1, slots = gemm_slots(bn);
    const int64_t round1 = (slots + p.tiles - 1) / p.tiles;
    const int64_t max_by_k = std::max<int64_t>(1, (n + 4 * kq - 1) / (4 * kq));   // >= 4 stages per split
    const int64_t lo = std::min(std::max((n + GEMM_MAX_SLICE - 1) / GEMM_MAX_SLICE, round1), max_by_k);
    const int64_t hi = std::min(std::max(lo, 4 * round1), max_by_k);
    int64_t splits = lo, best = INT64_MAX;
    for (int64_t s = lo; s <= hi; ++s) {
        const int64_t rounds = (p.tiles * s + slots - 1) / slots;
        const int64_t cost = rounds * ((n + s - 1) / s) + GEMM_SPLIT_COST * s;
        if (cost < best) {
            best = cost;
            splits = s;
        }
    }
    int64_t kslice = (n + splits - 1) / splits;
    kslice = (kslice + kq - 1) / kq * kq;
    p.splits = (n + kslice - 1) / kslice;
    if (p.splits < 1) p.splits = 1;
    p.kslice = kslice;
    return p;
}

}  // namespace ohx
