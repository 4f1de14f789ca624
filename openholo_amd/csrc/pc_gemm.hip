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
// its V rows and U columns into double-buffered LDS (4 and 8 on the 128 x 256
// tiles of 8 waves, 8 and 8 on the 128 x 128 tiles of 4 waves).  Along a
// thread's rows (and columns) the phase is a quadratic in the row index whose
// three coefficients are computed in FP64 and reduced mod 1, so the FP32
// evaluation (two FMAs, then v_sin/v_cos) only sums terms below 16 cycles for
// any point depth.  Each
// wave's generation of stage s+1 is software-pipelined with its MFMAs of
// stage s, slot by slot (a few VALU instructions after every MFMA), so that
// the VALU work issues while the matrix core runs.  Precision variants:
//   BF16x1 : hi(V) hi(U)
//   BF16x3 : hi hi + hi lo + lo hi   (lo = bf16(x - hi): ~2^-17 per factor)
//   FP16x3 : same split in FP16 (~2^-22), V scaled by 1/max|A| to stay in range
// Split-K slices write FP32 partials; k_gemm_reduce sums them in FP64 in a fixed
// order (deterministic) and adds -j S into the FP64 field.  Odd splits
// accumulate the negated products and are subtracted, which cancels the
// matrix cores' accumulation bias between pairs of splits (split_sign).
#include <algorithm>

#include "gemm_common.h"

namespace ohx {

namespace gemm {

// One point's stage data, written per launch by k_gemm_consts for one tile
// shape: position, 1/(2 lambda z), z/lambda cycles, the k^2 coefficients
// (v2, u2), the amplitude and the box; 64 bytes, read as 16-byte buffer
// loads at one 32-bit offset (the box only by the masked instances).
struct __align__(16) PointRec {
    double px, py, q, cyc0;
    float v2, u2, amp, pad;
    int4 box;
};
static_assert(sizeof(PointRec) == 64, "16-byte loads at a 64-byte stride");

// Point data of one stage as read from HBM: thread t holds point t % 16.
struct Raw {
    PointRec c;
    bool valid;
};

// Point pi of the launch's records (pi relative to the first).  Lanes past
// the slice end (pi >= p1) read the slice's last point and are marked invalid
// (zero amplitude, full box), so the loads are unconditional and in bounds.
// (The unmasked instances never read the box, and its load is dropped.)
__device__ inline void load_raw(Raw& r, Rsrc rs, int pi, int p1)
{
    r.valid = pi < p1;
    const int off = min(pi, p1 - 1) * (int)sizeof(PointRec);
    const u32x4 a = rec_load16(rs, off), b = rec_load16(rs, off + 16), c = rec_load16(rs, off + 32);
    r.c.px = __builtin_bit_cast(double, make_uint2(a.x, a.y));
    r.c.py = __builtin_bit_cast(double, make_uint2(a.z, a.w));
    r.c.q = __builtin_bit_cast(double, make_uint2(b.x, b.y));
    r.c.cyc0 = __builtin_bit_cast(double, make_uint2(b.z, b.w));
    r.c.v2 = __uint_as_float(c.x);
    r.c.u2 = __uint_as_float(c.y);
    r.c.amp = __uint_as_float(c.z);
    const u32x4 d = rec_load16(rs, off + 48);
    r.c.box = make_int4((int)d.x, (int)d.y, (int)d.z, (int)d.w);
}

// Per-(point, thread) constants of one stage.  Thread t generates V rows
// grp + NGRP i (i < NVT) and U columns grp + NGRP j (j < NUT), grp = t / 16
// (NGRP, NVT, NUT of the tile shape);
// with k = i - NVT/2 the phase (cycles) of its k-th row is exactly
// v0 + k v1 + k^2 v2  (mod 1), a quadratic in the integer k, and likewise
// u0 + k u1 + k^2 u2 (k = j - NUT/2) for its columns.  The three
// coefficients are computed in FP64 and reduced mod 1 before rounding to FP32
// (for integer k, k * frac(c) = k * c mod 1), so the FP32 evaluation only
// ever sums terms below 1, 4 and 16 cycles, whatever the phase excursion.
struct Pt {
    float v0, v1, v2, u0, u1, u2, amp;
    int rlo, rhi, clo, chi;   // box relative to the tile origin
};


// ythr / xthr: y of row grp + NGRP*(NVT/2) and x of column grp + NGRP*(NUT/2)
// of the tile (the reference element k = 0); dy = -NGRP pp_y and
// dx = NGRP pp_x (row index grows toward -y).
__device__ inline void prep_y(Pt& P, const Raw& r, double ythr, double dy)
{
    const double ey = ythr - r.c.py, ty = r.c.q * ey;
    P.v0 = (float)frac64(ty * ey + r.c.cyc0);
    P.v1 = (float)frac64(ty * (2.0 * dy));
    P.v2 = r.c.v2;
}
__device__ inline void prep_x(Pt& P, const Raw& r, double xthr, double dx)
{
    const double ex = xthr - r.c.px, tx = r.c.q * ex;
    P.u0 = (float)frac64(tx * ex);
    P.u1 = (float)frac64(tx * (2.0 * dx));
    P.u2 = r.c.u2;
}
// amplitude (times the split's sign sg) and box; returns true when the box
// does not cover the whole BM x W tile
template <int W>
__device__ inline bool prep_rest(Pt& P, const Raw& r, int x0, int y0, float sg)
{
    P.amp = r.valid ? sg * r.c.amp : 0.f;
    P.rlo = r.valid ? r.c.box.z - y0 : 0;
    P.rhi = r.valid ? r.c.box.w - y0 : BM;
    P.clo = r.valid ? r.c.box.x - x0 : 0;
    P.chi = r.valid ? r.c.box.y - x0 : W;
    return P.rlo > 0 || P.rhi < BM || P.clo > 0 || P.chi < W;
}
template <int W>
__device__ inline bool prep(Pt& P, const Raw& r, double ythr, double xthr, double dy, double dx, int x0, int y0,
                            float sg)
{
    prep_y(P, r, ythr, dy);
    prep_x(P, r, xthr, dx);
    return prep_rest<W>(P, r, x0, y0, sg);
}

// The same work in NPIECE pieces of two or three FP64 operations, spread by the
// pipelined stage over its first MFMA slots.
constexpr int NPIECE = 6;
template <int K>
__device__ inline void prep_piece(Pt& P, double (&t)[4], const Raw& r, double ythr, double xthr, double dy,
                                  double dx)
{
    if constexpr (K == 0) {
        t[0] = ythr - r.c.py;
        t[1] = r.c.q * t[0];
    } else if constexpr (K == 1) {
        P.v0 = (float)frac64(t[1] * t[0] + r.c.cyc0);
    } else if constexpr (K == 2) {
        P.v1 = (float)frac64(t[1] * (2.0 * dy));
        P.v2 = r.c.v2;
    } else if constexpr (K == 3) {
        t[2] = xthr - r.c.px;
        t[3] = r.c.q * t[2];
    } else if constexpr (K == 4) {
        P.u0 = (float)frac64(t[3] * t[2]);
    } else if constexpr (K == 5) {
        P.u1 = (float)frac64(t[3] * (2.0 * dx));
        P.u2 = r.c.u2;
    }
}

// phase (cycles, |phase| < 1 + |k| + k^2 <= 21) of element k from its
// coefficients (k is a compile-time constant wherever this is called).  No mod-1 reduction:
// v_sin_f32 / v_cos_f32 reduce a revolutions argument themselves (on gfx950
// sin(x) and sin(fract(x)) agree to 1 ulp for |x| <= 256), and the rounding
// of the FMAs above is the same with or without it.
__device__ inline float phase_at(int k, float c0, float c1, float c2)
{
    const float kf = (float)k;
    return k == 0 ? c0 : fmaf(fmaf(c2, kf, c1), kf, c0);
}

// Stores one element pair (hi and, 3-term, lo parts) of V or U.
template <int VAR>
__device__ inline void put(unsigned* hi_slot, unsigned* lo_slot, float re, float im)
{
    constexpr bool FP16 = Traits<VAR>::fp16;
    const unsigned h = pack2<FP16>(re, im);
    *hi_slot = h;
    if constexpr (Traits<VAR>::three) *lo_slot = pack2<FP16>(re - lo16<FP16>(h), im - hi16<FP16>(h));
}


// A thread's NE elements per stage, V rows spread evenly among U columns:
// element E is a V row when it raises the running V count.
template <int W>
constexpr int v_before(int E)
{
    return (E * Shape<W>::NVT) / Shape<W>::NE;
}
template <int W>
constexpr bool elem_is_v(int E)
{
    return v_before<W>(E + 1) > v_before<W>(E);
}
template <int W>
constexpr int elem_index(int E)
{
    return elem_is_v<W>(E) ? v_before<W>(E) : E - v_before<W>(E);
}

// element E of the thread's stage: phase, sincos (V: times the amplitude)
template <int W, int E>
__device__ inline float2 elem_value(const Pt& P)
{
    if constexpr (elem_is_v<W>(E)) {
        const float f = phase_at(elem_index<W>(E) - Shape<W>::NVT / 2, P.v0, P.v1, P.v2);
        return make_float2(P.amp * cos_rev(f), P.amp * sin_rev(f));
    } else {
        const float f = phase_at(elem_index<W>(E) - Shape<W>::NUT / 2, P.u0, P.u1, P.u2);
        return make_float2(cos_rev(f), sin_rev(f));
    }
}
template <int VAR, int W, int E>
__device__ inline void elem_store(unsigned* wv, unsigned* wu, float2 cs, int p, int grp)
{
    using Sh = Shape<W>;
    const int e = (grp + Sh::NGRP * elem_index<W>(E)) * ROWW + p;
    if constexpr (elem_is_v<W>(E)) put<VAR>(wv + e, wv + (Traits<VAR>::three ? PLANE_V : 0) + e, cs.x, cs.y);
    else put<VAR>(wu + e, wu + (Traits<VAR>::three ? Sh::PLANE_U : 0) + e, cs.x, cs.y);
}

template <int VAR, int W, int E>
__device__ inline void gen_elems(unsigned* wv, unsigned* wu, const Pt& P, int p, int grp)
{
    if constexpr (E < Shape<W>::NE) {
        elem_store<VAR, W, E>(wv, wu, elem_value<W, E>(P), p, grp);
        gen_elems<VAR, W, E + 1>(wv, wu, P, p, grp);
    }
}

// Generation of one stage into buffer WB, ignoring the boxes.
template <int VAR, int W, int WB>
__device__ inline void gen_stage(Smem<Traits<VAR>::three, W>& sm, const Pt& P, int p, int grp)
{
    gen_elems<VAR, W, 0>(&sm.v[WB][0][0][0], &sm.u[WB][0][0][0], P, p, grp);
}

// The boxes: zeroes the elements of rows / columns outside the point's box
// (the same thread wrote them, so program order makes the zeros win).  Run
// only by waves holding a point whose box does not cover the tile.
template <int VAR, int W, int WB>
__device__ inline void mask_stage(Smem<Traits<VAR>::three, W>& sm, const Pt& P, int p, int grp)
{
    using Sh = Shape<W>;
    constexpr int H2 = Traits<VAR>::three ? 1 : 0;
#pragma unroll
    for (int i = 0; i < Sh::NVT; ++i) {
        const int r = grp + Sh::NGRP * i;
        if (r < P.rlo || r >= P.rhi) {
            sm.v[WB][0][r][p] = 0u;
            sm.v[WB][H2][r][p] = 0u;
        }
    }
#pragma unroll
    for (int j = 0; j < Sh::NUT; ++j) {
        const int c = grp + Sh::NGRP * j;
        if (c < P.clo || c >= P.chi) {
            sm.u[WB][0][c][p] = 0u;
            sm.u[WB][H2][c][p] = 0u;
        }
    }
}

// State of one pipelined stage (see stage_body).
template <int VAR, int W>
struct StageCtx {
    f32x16 (&acc)[4][2];
    Raw& R;
    Rsrc rs;
    int pnext, ke;
    double ythr, xthr, dy, dx;
    int x0, y0;
    int p, grp, lrow, koff, ucol;
    float sg;      // split_sign of the workgroup's split
    Pt P;
    double t[4];   // FP64 temporaries shared by the prep pieces
    bool partial;
    u32x4 A[8][2], Ub[2][2], B[2][4];
    float2 cs[Shape<W>::NE];
};

// The generation schedule: 3-term, half-element h (even: value of element
// h/2, odd: its store) at slot H0 + (42 - H0) h / (2 NE); 1-term, element E
// at slot E0 + (14 - E0) E / NE (two elements share a slot when needed).
// Return the half-element / element at slot S, or -1.
// 3-term: prep piece K at slot K, amplitude/box and the next loads at slot
// NPIECE, half-elements from slot NPIECE + 1 to 41.  1-term: prep pieces two
// per slot (slots 0 .. NPIECE/2 - 1), amplitude/box at NPIECE/2, elements from
// NPIECE/2 + 1 to 13.  The generation ends before the stage's last MFMA group
// (tail_mfmas), within which the barrier comes (stage_body).
constexpr int H0 = NPIECE + 1, E0 = NPIECE / 2 + 1;
template <int W>
constexpr int half_at(int S)
{
    constexpr int NE = Shape<W>::NE, END = 48 - 6;
    for (int h = 0; h < 2 * NE; ++h)
        if (H0 + ((END - H0) * h) / (2 * NE) == S) return h;
    return -1;
}
template <int W>
constexpr int elem_at(int S, int n)   // n-th element at slot S (slots hold at most two)
{
    constexpr int NE = Shape<W>::NE, END = 16 - 2;
    for (int e = 0; e < NE; ++e)
        if (E0 + ((END - E0) * e) / NE == S && n-- == 0) return e;
    return -1;
}
// every element (1-term) and half-element (3-term) lands in exactly one slot,
// before the tail group
template <int W>
constexpr bool schedule_complete()
{
    constexpr int NE = Shape<W>::NE;
    for (int e = 0; e < NE; ++e) {
        int n = 0;
        for (int S = 0; S < 16 - 2; ++S)
            n += (elem_at<W>(S, 0) == e) + (elem_at<W>(S, 1) == e) + (elem_at<W>(S, 2) == e);
        if (n != 1) return false;
    }
    for (int S = 0; S < 16; ++S)
        if (elem_at<W>(S, 2) >= 0) return false;
    for (int h = 0; h < 2 * NE; ++h) {
        int n = 0;
        for (int S = 0; S < 48 - 6; ++S) n += half_at<W>(S) == h;
        if (n != 1) return false;
    }
    return true;
}
static_assert(schedule_complete<128>() && schedule_complete<256>(),
              "generation schedule must cover every element once, before the last MFMA group");


// Slot S of the pipelined stage: MFMA S, operand loads / transforms due at
// this slot, and this slot's share of the next stage's generation.
template <int VAR, int W, int S, int END>
__device__ inline void slot(const unsigned* __restrict__ rv, const unsigned* __restrict__ ru,
                            unsigned* __restrict__ wv, unsigned* __restrict__ wu, StageCtx<VAR, W>& c)
{
    constexpr bool THREE = Traits<VAR>::three;
    if constexpr (S < END) {
        mfma_slot<VAR, S, Shape<W>::PLANE_U>(rv, ru, c);
        if constexpr (THREE && S < NPIECE) prep_piece<S>(c.P, c.t, c.R, c.ythr, c.xthr, c.dy, c.dx);
        if constexpr (!THREE && S < NPIECE / 2) {
            prep_piece<2 * S>(c.P, c.t, c.R, c.ythr, c.xthr, c.dy, c.dx);
            prep_piece<2 * S + 1>(c.P, c.t, c.R, c.ythr, c.xthr, c.dy, c.dx);
        }
        if constexpr (S == (THREE ? NPIECE : NPIECE / 2)) {
            c.partial = prep_rest<W>(c.P, c.R, c.x0, c.y0, c.sg);
            load_raw(c.R, c.rs, c.pnext, c.ke);
        }
        if constexpr (THREE) {
            constexpr int h = half_at<W>(S);
            if constexpr (h >= 0 && (h & 1) == 0) c.cs[h >> 1] = elem_value<W, (h >> 1)>(c.P);
            if constexpr (h >= 0 && (h & 1) == 1) elem_store<VAR, W, (h >> 1)>(wv, wu, c.cs[h >> 1], c.p, c.grp);
        } else {
            constexpr int e0 = elem_at<W>(S, 0), e1 = elem_at<W>(S, 1);
            if constexpr (e0 >= 0) elem_store<VAR, W, e0>(wv, wu, elem_value<W, e0>(c.P), c.p, c.grp);
            if constexpr (e1 >= 0) elem_store<VAR, W, e1>(wv, wu, elem_value<W, e1>(c.P), c.p, c.grp);
        }
        __builtin_amdgcn_sched_barrier(0);
        slot<VAR, W, S + 1, END>(rv, ru, wv, wu, c);
    }
}

// One stage's MFMAs (from planes rv / ru) software-pipelined with the next
// stage's operand generation (into planes wv / wu), slot by slot: slot S holds
// MFMA S and a fixed share of the generation work (FP64 point constants, then
// elements), fenced by sched_barrier so that the VALU work issues while the
// matrix core runs.  The restrict-qualified planes tell the compiler that the
// generation's LDS writes never alias the operand reads.  Also loads the
// point data of the stage after next into R.  The stage's first operands come
// from cy (read by the previous stage); the barrier comes in the middle of the
// last MFMA group (whose operands are in registers, and whose first half keeps
// the matrix pipe busy while the wave waits there), after the masked
// instances' zeroing of out-of-box elements (MASK: waves holding a point whose
// box does not cover the tile); after it the next stage's first operands are
// read into cy from wv / wu, which all waves have written by then, while the
// rest of the group runs.
template <int VAR, int W, bool MASK, int RB>
__device__ inline void stage_body(Smem<Traits<VAR>::three, W>& sm, f32x16 (&acc)[4][2], Raw& R, Rsrc rs,
                                  int pnext, int ke, double ythr, double xthr, double dy, double dx, int x0, int y0,
                                  int p, int grp, int lrow, int koff, int ucol, float sg, Carry& cy)
{
    constexpr bool THREE = Traits<VAR>::three;
    constexpr int NM = mfma_per_stage<VAR>(), PU = Shape<W>::PLANE_U;
    const unsigned* __restrict__ rv = &sm.v[RB][0][0][0];
    const unsigned* __restrict__ ru = &sm.u[RB][0][0][0];
    unsigned* __restrict__ wv = &sm.v[RB ^ 1][0][0][0];
    unsigned* __restrict__ wu = &sm.u[RB ^ 1][0][0][0];
    StageCtx<VAR, W> c{acc, R, rs, pnext, ke, ythr, xthr, dy, dx, x0, y0, p, grp, lrow, koff, ucol, sg};
    c.partial = false;
    constexpr int T = NM - tail_mfmas<VAR>() / 2;   // half the last group before the barrier
    c.A[0][0] = cy.A[0][0];
    c.Ub[0][0] = cy.Ub[0][0];
    if constexpr (THREE) {
        c.A[0][1] = cy.A[0][1];
        c.Ub[0][1] = cy.Ub[0][1];
    } else {
        c.A[1][0] = cy.A[1][0];
    }
    xform_b(c, 0, 0);
    if constexpr (THREE) xform_b(c, 0, 1);
    slot<VAR, W, 0, T>(rv, ru, wv, wu, c);
    if constexpr (MASK) {
        if (__builtin_amdgcn_ballot_w64(c.partial)) mask_stage<VAR, W, RB ^ 1>(sm, c.P, p, grp);
    }
    __syncthreads();
    load_b<VAR, PU>(wu, cy, 0);
    load_a<VAR>(wv, cy, 0);
    if constexpr (!THREE) load_a<VAR>(wv, cy, 1);
    __builtin_amdgcn_sched_barrier(0);
    slot<VAR, W, T, NM>(rv, ru, wv, wu, c);
}

// One pipeline step: MFMAs of stage st (buffer RB) and generation of stage
// st+1 (buffer RB^1) from R in one basic block, so that they interleave; R
// then receives stage st+2's point data.  Past the last stage the generation
// runs on clamped, zero-amplitude data into the unused buffer: the step stays
// branch-free, which keeps the accumulators in place.  MASK: some point of
// the launch has a box smaller than the hologram, so waves holding a point
// whose box does not cover the tile zero its out-of-box elements (a ballot
// and a branch per stage, ~5 % of the kernel, paid only by such launches).
template <int VAR, int W, int RB, bool MASK>
__device__ inline void step(Smem<Traits<VAR>::three, W>& sm, f32x16 (&acc)[4][2], Raw& R, Rsrc rs, int st,
                            int kb, int ke, double ythr, double xthr, double dy, double dx, int x0, int y0, int p,
                            int grp, int lrow, int koff, int ucolx, float sg, Carry& cy)
{
    stage_body<VAR, W, MASK, RB>(sm, acc, R, rs, kb + (st + 2) * KP + p, ke, ythr, xthr, dy, dx, x0, y0, p, grp,
                                 lrow, koff, ucolx, sg, cy);
}

// One workgroup's tile and split: prologue, K loop, FP32 partials.
template <int VAR, int W, bool MASK>
__device__ __attribute__((always_inline)) inline void fresnel_tile(Smem<Traits<VAR>::three, W>& sm, const Geometry& g,
                                                                   const PointRec* pconst,
                                                                   float2* part, int tiles_x, int tiles,
                                                                   int64_t kfirst, int64_t klast, int64_t kslice)
{
    const int tile = blockIdx.x % tiles;
    const int split = blockIdx.x / tiles;
    const int tx = tile % tiles_x, ty = tile / tiles_x;
    using Sh = Shape<W>;
    const int x0 = tx * W, y0 = ty * BM;
    // point indices relative to kfirst (fewer than MAX_REC per launch)
    const int np = (int)(klast - kfirst);
    const int kb = (int)(split * kslice);
    const int ke = min(np, (int)(kb + kslice));
    const int nst = (ke - kb + KP - 1) / KP;
    const Rsrc rs = rec_rsrc(pconst, np * (int)sizeof(PointRec));
    const float sg = split_sign(split);
    const int p = threadIdx.x & (KP - 1), grp = threadIdx.x >> 4;
    const double ythr = gpix_y(g, y0 + grp + Sh::NGRP * (Sh::NVT / 2));
    const double xthr = gpix_x(g, x0 + grp + Sh::NGRP * (Sh::NUT / 2));
    const double dy = -Sh::NGRP * g.pp_y, dx = Sh::NGRP * g.pp_x;

    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int lrow = lane & 31, koff = (lane >> 5) * 4;   // dword offset of this lane's 8 values
    const int ucolx = wave * 32 + lrow;                     // this wave's 32 columns

    // prologue: stage 0 into buffer 0, stage 1's point data into R (an empty
    // split, which the planner never makes, would write zeros)
    Raw R;
    if (nst > 0) {
        load_raw(R, rs, kb + p, ke);
        Pt P;
        const bool partial = prep<W>(P, R, ythr, xthr, dy, dx, x0, y0, sg);
        gen_stage<VAR, W, 0>(sm, P, p, grp);
        if constexpr (MASK) {
            if (partial) mask_stage<VAR, W, 0>(sm, P, p, grp);
        } else {
            (void)partial;
        }
        load_raw(R, rs, kb + KP + p, ke);
    }
    __syncthreads();
    Carry cy;
    cy.lrow = lrow;
    cy.koff = koff;
    cy.ucol = ucolx;
    load_b<VAR, Sh::PLANE_U>(&sm.u[0][0][0][0], cy, 0);
    load_a<VAR>(&sm.v[0][0][0][0], cy, 0);
    if constexpr (!Traits<VAR>::three) load_a<VAR>(&sm.v[0][0][0][0], cy, 1);
    f32x16 acc[4][2];
#pragma unroll
    for (int i = 0; i < 4; ++i) acc[i][0] = acc[i][1] = f32x16{};
    // no branch between the two steps, so that the next operands' reads
    // stay ahead of the last MFMA group; for an odd stage count the second step
    // then runs one stage past the end, whose operands (generated from clamped,
    // zero-amplitude point data) contribute exact zeros.
    for (int st = 0; st < nst; st += 2) {
        step<VAR, W, 0, MASK>(sm, acc, R, rs, st, kb, ke, ythr, xthr, dy, dx, x0, y0, p, grp, lrow, koff, ucolx,
                              sg, cy);
        step<VAR, W, 1, MASK>(sm, acc, R, rs, st + 1, kb, ke, ythr, xthr, dy, dx, x0, y0, p, grp, lrow, koff,
                              ucolx, sg, cy);
    }

    store_partials(g, acc, part, split, x0, y0, ucolx, lane);
}

// pconst: the records of points kfirst .. klast-1 (k_gemm_consts; at most
// MAX_REC of them) for this instance's tile shape.
// Both MASK instances are launched; the one that does not match
// *partial_flag (!= 0: some point has a box smaller than the hologram) returns
// at once, so each keeps its own register allocation.
// Two waves per SIMD either way: two 4-wave workgroups (W = 128) or one
// 8-wave workgroup (W = 256) per CU.
template <int VAR, int W, bool MASK>
__global__ __launch_bounds__(Shape<W>::NT, 512 / Shape<W>::NT) void k_fresnel_gemm(
    Geometry g, const PointRec* __restrict__ pconst, const unsigned* __restrict__ partial_flag,
    float2* __restrict__ part, int tiles_x, int tiles, int64_t kfirst, int64_t klast, int64_t kslice)
{
    __shared__ __align__(16) Smem<Traits<VAR>::three, W> sm;
    if ((__builtin_amdgcn_readfirstlane(partial_flag[0]) != 0) != MASK) return;
    fresnel_tile<VAR, W, MASK>(sm, g, pconst, part, tiles_x, tiles, kfirst, klast, kslice);
}

// field += -j * amax * sum_split split_sign(split) S_split   (FP64, fixed
// summation order; odd splits accumulated the negated products)
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
            const double sg = split_sign(k);   // odd splits hold the negated sums
            sr = fma(sg, (double)v.x, sr);
            si = fma(sg, (double)v.y, si);
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

// amax[0] = 0 (max |A|), amax[1] = 0 as an unsigned word (the partial-box flag)
__global__ void k_amp_init(float* amax)
{
    if (threadIdx.x == 0) {
        amax[0] = 0.f;
        reinterpret_cast<unsigned*>(amax)[1] = 0u;
    }
}

// Per-point records of the stage generation, once per launch instead of once
// per (point, tile, thread): the point's position, q and cycles, the k^2
// coefficients v2 = frac(q dy^2) and u2 = frac(q dx^2) (dy = -ngrp pp_y,
// dx = ngrp pp_x, ngrp the tile shape's thread groups, 16 into pc_narrow and
// 32 into pc_wide; FP64 then reduced mod 1 and rounded) and the amplitude
// a/(lambda z), times 1/max|A| for FP16.  Also
// raises *partial_flag when a point's box is smaller than the hologram (one
// atomic per wave holding such a point, skipped once the flag reads as set).
__global__ __launch_bounds__(256) void k_gemm_consts(Geometry g, PointSoA s, int64_t p0, int64_t p1,
                                                     const float* __restrict__ amax, int scaled,
                                                     PointRec* __restrict__ pc_narrow,
                                                     PointRec* __restrict__ pc_wide, unsigned* partial_flag)
{
    const int64_t i = p0 + (int64_t)blockIdx.x * 256 + threadIdx.x;
    bool partial = false;
    if (i < p1) {
        const double q = s.q[i];
        const float am = amax[0];
        const float a = (float)s.kamp[i] * ((scaled && am > 0.f) ? 1.0f / am : 1.0f);
        const double dy = -Shape<128>::NGRP * g.pp_y, dx = Shape<128>::NGRP * g.pp_x;
        const double dy2 = q * (dy * dy), dx2 = q * (dx * dx);
        const double px = s.px[i], py = s.py[i], cyc0 = s.cyc0[i];
        const int4 b = s.box[i];
        pc_narrow[i - p0] = PointRec{px, py, q, cyc0, (float)frac64(dy2), (float)frac64(dx2), a, 0.f, b};
        pc_wide[i - p0] = PointRec{px, py, q, cyc0, (float)frac64(4.0 * dy2), (float)frac64(4.0 * dx2), a, 0.f, b};
        partial = b.x > 0 || b.y < g.pn_x || b.z > 0 || b.w < g.pn_y;
    }
    if (__builtin_amdgcn_ballot_w64(partial) && (threadIdx.x & 63) == 0 &&
        __atomic_load_n(partial_flag, __ATOMIC_RELAXED) == 0u)
        atomicOr(partial_flag, 1u);
}

}  // namespace gemm


// Tile width of each kernel instance: 256 columns (a quarter fewer generated
// elements per point-pixel pair than 128), except the masked 3-term
// instances, whose box pass does not fit the wide tile's registers; they keep
// 128-column tiles and use the same split-K plan, so the partial planes and
// their reduction do not depend on which instance runs.
template <int VAR, bool MASK>
constexpr int tile_width()
{
    return (MASK && gemm::Traits<VAR>::three) ? 128 : 256;
}

// scratch: [amax, 256 B] [per-point records for 128- and 256-column tiles,
// each padded to 256 B] [split-K partial planes]
static size_t consts_bytes(int64_t npoints)
{
    return ((size_t)npoints * sizeof(gemm::PointRec) + 255) / 256 * 256;
}

// The split-K plan of a launch (that of the 256-column tiles).
static GemmPlan fresnel_plan(const Geometry& g, int64_t npoints) { return plan_gemm(g, npoints, gemm::KP, 256); }

// Enough for any launch of at most npoints points.
size_t fresnel_gemm_scratch_bytes(const Geometry& g, int64_t npoints)
{
    const int64_t splits = gemm_max_splits(g, npoints, gemm::KP, 256);
    return 256 + 2 * consts_bytes(npoints) + (size_t)splits * g.pn_x * g.pn_y * sizeof(float2);
}

template <int VAR, bool MASK>
static void launch_instance(hipStream_t stream, const Geometry& g, const gemm::PointRec* pc_narrow,
                            const gemm::PointRec* pc_wide, const unsigned* pf, float2* part, const GemmPlan& pl,
                            int64_t p0, int64_t p1)
{
    using namespace gemm;
    constexpr int W = tile_width<VAR, MASK>();
    const int tiles_x = (g.pn_x + W - 1) / W, tiles = tiles_x * ((g.pn_y + BM - 1) / BM);
    const dim3 grid((unsigned)((int64_t)tiles * pl.splits)), block(Shape<W>::NT);
    hipLaunchKernelGGL((k_fresnel_gemm<VAR, W, MASK>), grid, block, 0, stream, g, W == 256 ? pc_wide : pc_narrow,
                       pf, part, tiles_x, tiles, p0, p1, pl.kslice);
}

// Both instances; the one that does not match the device's partial-box flag
// returns at once.
template <int VAR>
static void launch_variant(hipStream_t stream, const Geometry& g, const gemm::PointRec* pc_narrow,
                           const gemm::PointRec* pc_wide, const unsigned* pf, float2* part, const GemmPlan& pl,
                           int64_t p0, int64_t p1)
{
    launch_instance<VAR, false>(stream, g, pc_narrow, pc_wide, pf, part, pl, p0, p1);
    launch_instance<VAR, true>(stream, g, pc_narrow, pc_wide, pf, part, pl, p0, p1);
}

static int variant_of(uint32_t mode)
{
    using namespace gemm;
    return (mode & OHX_MODE_MFMA_FP16X3) ? V_FP16X3 : (mode & OHX_MODE_MFMA_BF16X3) ? V_BF16X3 : V_BF16X1;
}

// One launch sequence over at most MAX_REC points.
static int launch_piece(hipStream_t stream, void* scratch, size_t scratch_bytes, const Geometry& g,
                        const PointSoA& soa, int64_t p0, int64_t p1, double* d_field, uint32_t mode)
{
    using namespace gemm;
    const int var = variant_of(mode);
    const GemmPlan pl = fresnel_plan(g, p1 - p0);
    if (fresnel_gemm_scratch_bytes(g, p1 - p0) > scratch_bytes) return OHX_ERR_INVALID_ARG;
    float* amax = (float*)scratch;
    auto* pc_narrow = (gemm::PointRec*)((char*)scratch + 256);
    auto* pc_wide = (gemm::PointRec*)((char*)scratch + 256 + consts_bytes(p1 - p0));
    float2* part = (float2*)((char*)scratch + 256 + 2 * consts_bytes(p1 - p0));
    hipLaunchKernelGGL(k_amp_init, dim3(1), dim3(64), 0, stream, amax);
    hipLaunchKernelGGL(k_amp_max, dim3(256), dim3(256), 0, stream, (const double*)soa.kamp, p0, p1, amax);
    unsigned* partial_flag = reinterpret_cast<unsigned*>(amax) + 1;
    hipLaunchKernelGGL(k_gemm_consts, dim3((unsigned)((p1 - p0 + 255) / 256)), dim3(256), 0, stream, g, soa, p0,
                       p1, (const float*)amax, var == V_FP16X3 ? 1 : 0, pc_narrow, pc_wide, partial_flag);
    const unsigned* pf = partial_flag;
    switch (var) {
    case V_BF16X1:
        launch_variant<V_BF16X1>(stream, g, pc_narrow, pc_wide, pf, part, pl, p0, p1);
        break;
    case V_BF16X3:
        launch_variant<V_BF16X3>(stream, g, pc_narrow, pc_wide, pf, part, pl, p0, p1);
        break;
    default:
        launch_variant<V_FP16X3>(stream, g, pc_narrow, pc_wide, pf, part, pl, p0, p1);
        break;
    }
    if (hipGetLastError() != hipSuccess) return OHX_ERR_DEVICE;
    int64_t rg = (pl.npx + 255) / 256;
    if (rg > 65536) rg = 65536;
    hipLaunchKernelGGL(k_gemm_reduce, dim3((unsigned)rg), dim3(256), 0, stream, (const float2*)part, (int)pl.splits,
                       pl.npx, (double2*)d_field, (const float*)amax, var == V_FP16X3 ? 1 : 0);
    return hipGetLastError() == hipSuccess ? OHX_OK : OHX_ERR_DEVICE;
}
// Scratch sized for p1 - p0 points holds any shorter piece's.
int launch_fresnel_gemm_range(hipStream_t stream, void* scratch, size_t scratch_bytes, const Geometry& g,
                              const PointSoA& soa, int64_t p0, int64_t p1, double* d_field, uint32_t mode)
{
    if (p1 <= p0) return OHX_OK;
    if (fresnel_gemm_scratch_bytes(g, p1 - p0) > scratch_bytes) return OHX_ERR_INVALID_ARG;
    int st = OHX_OK;
    for (int64_t q0 = p0; q0 < p1 && st == OHX_OK; q0 += gemm::MAX_REC)
        st = launch_piece(stream, scratch, scratch_bytes, g, soa, q0, std::min(p1, q0 + gemm::MAX_REC), d_field,
                          mode);
    return st;
}

int launch_gemm_reduce(hipStream_t stream, const void* part, int64_t splits, int64_t npx, double* d_field)
{
    using namespace gemm;
    int64_t rg = (npx + 255) / 256;
    if (rg > 65536) rg = 65536;
    hipLaunchKernelGGL(k_gemm_reduce, dim3((unsigned)rg), dim3(256), 0, stream, (const float2*)part, (int)splits, npx,
                       (double2*)d_field, (const float*)nullptr, 0);
    return hipGetLastError() == hipSuccess ? OHX_OK : OHX_ERR_DEVICE;
}

// The split-K partial planes grow with points x pixels (16 x 66 MB for 1M points
// at 3840x2160); above MAX_GEMM_SCRATCH the points are processed in chunks, each
// reduced into the field before the next (chunk boundaries depend only on the
// point count and geometry, so results stay deterministic).
constexpr size_t MAX_GEMM_SCRATCH = size_t(4) << 30;

int launch_fresnel_gemm(ohx_ctx* ctx, const Geometry& g, const PointSoA& soa, double* d_field, uint32_t mode)
{
    int64_t chunk = soa.n;
    while (chunk > gemm::KP && (fresnel_gemm_scratch_bytes(g, chunk) > MAX_GEMM_SCRATCH || chunk > gemm::MAX_REC))
        chunk = (chunk + 1) / 2;
    const size_t bytes = fresnel_gemm_scratch_bytes(g, chunk);
    int st = ensure_gemm_scratch(ctx, bytes);
    if (st) return st;
    int64_t wgs = 0;
    for (int64_t p0 = 0; p0 < soa.n && !st; p0 += chunk) {
        const int64_t p1 = std::min<int64_t>(soa.n, p0 + chunk);
        st = launch_fresnel_gemm_range(ctx->stream, ctx->gemm_scratch, ctx->gemm_scratch_bytes, g, soa, p0, p1,
                                       d_field, mode);
        const GemmPlan pl = fresnel_plan(g, p1 - p0);   // the 256-column instance's grid
        wgs += (int64_t)pl.tiles * pl.splits;
    }
    ctx->last.workgroups = wgs;
    ctx->last.kernel_kind = 2;
    if (st) return set_error(ctx, st, "matrix-core Fresnel launch failed");
    return OHX_OK;
}

}  // namespace ohx
