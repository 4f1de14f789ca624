// RS point-cloud sum as a rank-2 complex GEMM on gfx950 matrix cores.
//
// ophGen::RS_Diffraction (ophgen/src/ophGen.cpp:361-445) adds, per point p and
// pixel (x, y) inside the anti-aliasing mask,
//   a z / (lambda r^2) (sin 2pi c - j cos 2pi c),  c = r/lambda (+ phase/2pi),
//   r = sqrt((x - x_p)^2 + (y - y_p)^2 + z^2),
// i.e. -j A e^{j 2pi c}.  Unlike the Fresnel term this does not factor into a
// row part times a column part, but on a 128 x 128 tile it is a rank-2 product
// to the accuracy of the FP32 direct kernel's separable form.  Take the tile's
// reference pixel (x_c, y_c) = pixel (x0 + 64, y0 + 64), dx_c = x_c - x_p,
// dy_c = y_c - y_p, r_c^2 = dx_c^2 + dy_c^2 + z^2 and u = x - x_c, v = y - y_c:
//   r^2 - r_c^2 = n(u) + N(v),   n = u (2 dx_c + u),   N = v (2 dy_c + v)
//   (r - r_c)/lambda = h (n + N) + k (n + N)^2 + O(e^3 r_c/lambda)
//   z/r^2 = (z/r_c^2) (1 + mu n)(1 + mu N) + O(e^2)
// with e = |n + N|/r_c^2, h = 1/(2 r_c lambda), k = -1/(8 r_c^3 lambda) and
// mu = -1/r_c^2.  Writing X(n) = h n + k n^2, Y(N) = h N + k N^2, the cross
// term 2 k n N of the phase is theta = 4 pi k n N radians.  With A, B the
// largest |n|, |N| on the tile (theta_max = |4 pi k| A B), e^{j theta} is
//   (1 - alpha n^2)(1 - beta N^2) + j theta,  alpha = (4 pi k B)^2 / 6,
//   beta = (4 pi k A)^2 / 6,
// to within theta_max^2 / 6 (the minimax choice for the real part: the plain
// 1 + j theta is off by theta_max^2 / 2 at the tile corners), so the term
// A e^{j 2pi c} is
//   U1[x] V1[y] + U2[x] V2[y]
//   V1 = (1 + mu N - beta N^2) e^{j 2pi Y(N)},                 V2 = N V1
//   U1 = A_c (1 + mu n - alpha n^2) e^{j 2pi (phi_c + X(n))},  U2 = j 4 pi k n U1
// (the products mu beta N^3, mu alpha n^3 are dropped: below e theta_max^2),
// with A_c = a z/(lambda r_c^2) and phi_c = frac(c_p + (r_c - z)/lambda): two
// K entries of the Fresnel kernel's complex GEMM per point (S = sum U V^T,
// field += -j S), 8 points per 16-entry stage.
//
// Which points and tiles: on a tile, a point's expansion holds when the
// neglected terms are small (rs_bound_ok, with r_c >= z): e^3 r_c/(16 lambda)
// <= 2e-7 cycles and e^2 <= 1e-7 (the FP32 direct kernel's bounds), and
// theta_max <= 8e-3 rad for BF16x3 (theta_max^2/6 <= 1.1e-5 of the term, the
// size of the 3-term split's own error), 3.5e-2 rad for BF16x1 (2e-4, a tenth
// of its BF16 rounding); the matrix cores may take the tile when, in
// addition, the per-pixel range test of ophGen.cpp:415-429 holds with a
// 1e-12 margin on every pixel of the tile inside the box (ophGen.cpp:384-407),
// so that the mask there is the box, a product of a row and a column
// interval.  A stable three-way partition (k_rs_classify) puts first the
// points whose box is the hologram, on all of which the range test holds and
// on every tile the expansion (the unmasked instance), then the points whose
// expansion may hold on their nearest tile (the masked instance: zero
// amplitude for rows and columns outside the box and on the tiles that
// k_rs_tilemask leaves out), then the rest.  The FP32 direct kernels then
// take the masked points' other tiles (Geometry::rs_skip_mask) and every tile
// of the rest.
//
// Operand generation, per stage and thread (point t % 8, rows and columns
// grp + 32 i, grp = t / 8): the per-(point, tile) constants in FP64 (r_c from
// an FP32 v_rsq seed and one Newton step, relative error ~2e-14), then per row
// N and Y in FP64, Y reduced mod 1 before rounding to FP32 (BF16x1: N and Y
// in FP32, then reduced), one v_sin/v_cos pair for both V1 and V2 (and for
// U1, U2).  Precision: V1 and U1 in the
// 3-term split (hi hi + hi lo + lo hi), V2 and U2 hi only (lo part zero): the
// rank-2 correction is below theta of the term, so its 2^-9 BF16 rounding
// costs below 2e-5 of the term (theta <= 8e-3).  BF16 only: A_c reaches ~1e7,
// beyond the FP16 range.
#include <algorithm>
#include <climits>

#include <hipcub/hipcub.hpp>

#include "gemm_common.h"

namespace ohx {

namespace rsgemm {
using namespace gemm;

constexpr int RP = KP / 2;       // points per stage (two K entries each)
constexpr int RG = NT / RP;      // thread groups: thread t -> point t % RP, group t / RP
constexpr int RR = BM / RG;      // V rows per thread and stage: grp + RG i
constexpr int RC = BN / RG;      // U columns per thread and stage: grp + RG j
constexpr int RU = RR + RC;      // work units (a row or a column) per thread and stage
static_assert(RR == 4 && RC == 4 && PLANE_V == PLANE_U, "unit schedule assumes 4 rows and 4 columns");

constexpr double FOUR_PI = RS_FOUR_PI;
// neglected terms on the matrix-core path (see the header comment; the
// phase and amplitude bounds are RS_MAX_PHASE3 and RS_MAX_AMP2)
constexpr double MAX_THETA_X3 = 8e-3; // rad, BF16x3
constexpr double MAX_THETA_X1 = 3.5e-2; // rad, BF16x1
constexpr double RANGE_TOL = 1e-12;   // m, margin of the range test

// Per-tile mask of the masked instance (k_rs_tilemask): the launch's point jr
// (relative to its first) has its row at bits + jr * words; the workgroup's
// tile is bit `tile`.  Null bits: every tile of every point.
struct TileMask {
    const unsigned* bits;
    int words, tile;
};

// One partitioned point (k_rs_gather): position, z, cycles, amplitude and
// box, 64 bytes, read as buffer loads at one 32-bit offset.
struct __align__(16) RsRec {
    double px, py, pz, cyc0, kamp, pad;
    int4 box;
};
static_assert(sizeof(RsRec) == 64, "16-byte loads at a 64-byte stride");

// Point data of one stage: thread t holds point t % RP.  Lanes past the slice
// end read the slice's last point and are marked invalid (zero amplitude).
// MASK: the points' boxes may be smaller than the hologram (loaded too), and
// the point's word of the tile mask (all ones without a mask).
struct Raw {
    double px, py, pz, cyc0, kamp;
    int4 box;
    unsigned mword;
    bool valid;
};
// (pi, p1 relative to the launch's first record; the tile mask holds at most
// 2^28 words, so jr * words fits 32 bits.)
template <bool MASK>
__device__ inline void load_raw(Raw& r, Rsrc rs, const TileMask& tm, int pi, int p1)
{
    r.valid = pi < p1;
    const int jr = min(pi, p1 - 1);
    if constexpr (MASK) r.mword = tm.bits ? tm.bits[jr * tm.words + (tm.tile >> 5)] : ~0u;
    const int off = jr * (int)sizeof(RsRec);
    const u32x4 a = rec_load16(rs, off), b = rec_load16(rs, off + 16);
    const uint2 c = __builtin_bit_cast(uint2, __builtin_amdgcn_raw_buffer_load_b64(rs, off + 32, 0, 0));
    r.px = __builtin_bit_cast(double, make_uint2(a.x, a.y));
    r.py = __builtin_bit_cast(double, make_uint2(a.z, a.w));
    r.pz = __builtin_bit_cast(double, make_uint2(b.x, b.y));
    r.cyc0 = __builtin_bit_cast(double, make_uint2(b.z, b.w));
    r.kamp = __builtin_bit_cast(double, c);
    if constexpr (MASK) {
        const u32x4 d = rec_load16(rs, off + 48);
        r.box = make_int4((int)d.x, (int)d.y, (int)d.z, (int)d.w);
    }
}

// Per-(point, tile) constants of a stage
struct Pt {
    double dy2, dx2;          // 2 dy_c, 2 dx_c
    double h, k;              // 1/(2 r_c lambda), -1/(8 r_c^3 lambda)
    double phic;              // frac(c_p + (r_c - z)/lambda)
    float mu, ac, acmu, g;    // -1/r_c^2, A_c (0 for invalid lanes), A_c mu, 4 pi k
    float beta, acal;         // cross-term correction: beta, A_c alpha
    unsigned box;             // MASK: box relative to the tile origin, clamped to [0, 128]:
                              // row range [b0, b1), column range [b2, b3), byte k = bk
};
struct Tmp {
    double dyc, dxc, zz, rc2, y, irc2;
};
// Tile frame of a thread: reference pixel, its rows' and columns' offsets
struct Frame {
    double yc, xc;            // reference pixel (x0 + 64, y0 + 64)
    double vb, dv, ub, du;    // v of row grp + RG i = vb + i dv; u of column grp + RG j = ub + j du
    float vbf, dvf, ubf, duf; // the same in FP32 (1-term units)
    double il, hl;            // 1/lambda, 1/(2 lambda)
    double um, vm;            // 64 pitches: the largest |u|, |v| on the tile
    TileMask tm;              // masked instance: the tiles it takes per point
    int x0, y0;               // tile origin
    float sg;                 // split_sign of the workgroup's split
};

// The constants in NPR pieces of a few dependent FP64 operations each, spread
// by the pipelined stage over its first MFMA slots.
constexpr int NPR = 5;
template <int K, bool MASK>
__device__ inline void prep_piece(Pt& P, Tmp& t, const Raw& r, const Frame& f)
{
    if constexpr (K == 0) {
        t.dyc = f.yc - r.py;
        t.dxc = f.xc - r.px;
        t.zz = r.pz * r.pz;
    } else if constexpr (K == 1) {
        t.rc2 = fma(t.dxc, t.dxc, fma(t.dyc, t.dyc, t.zz));
        t.y = (double)__builtin_amdgcn_rsqf((float)t.rc2);
    } else if constexpr (K == 2) {
        // one Newton step on 1/sqrt(r_c^2): ~1e-7 -> ~2e-14 relative
        const double e = fma(-t.rc2 * t.y, t.y, 1.0);
        t.y = fma(0.5 * t.y, e, t.y);
        t.irc2 = t.y * t.y;
    } else if constexpr (K == 3) {
        P.phic = frac64(fma(fma(t.rc2, t.y, -r.pz), f.il, r.cyc0));   // r_c = r_c^2 / r_c
        P.h = t.y * f.hl;
        P.k = (P.h * t.irc2) * -0.25;
    } else if constexpr (K == 4) {
        P.mu = (float)(-t.irc2);
        // masked instance: zero amplitude on the tiles the mask leaves to the
        // direct kernel
        bool keep = r.valid;
        if constexpr (MASK) keep = keep && ((r.mword >> (f.tm.tile & 31)) & 1u);
        P.ac = keep ? f.sg * (float)(r.kamp * t.irc2) : 0.f;
        P.acmu = P.ac * P.mu;
        P.g = (float)(FOUR_PI * P.k);
        P.dy2 = 2.0 * t.dyc;
        P.dx2 = 2.0 * t.dxc;
        // cross-term correction: with A, B the largest |n|, |N| on the tile
        // and theta = g n N, (1 - alpha n^2)(1 - beta N^2) + j theta with
        // alpha = g^2 B^2 / 6, beta = g^2 A^2 / 6 approximates e^{j theta}
        // to within theta_max^2 / 6 (the plain 1 + j theta: theta_max^2 / 2)
        const float A = (float)(f.um * (fabs(P.dx2) + f.um)), B = (float)(f.vm * (fabs(P.dy2) + f.vm));
        const float g2 = P.g * P.g * (1.f / 6.f);
        P.beta = g2 * A * A;
        P.acal = P.ac * (g2 * B * B);
        // box relative to the tile (invalid lanes: the whole tile)
        auto cl = [](int v) { return (unsigned)min(max(v, 0), 128); };
        P.box = r.valid ? cl(r.box.z - f.y0) | cl(r.box.w - f.y0) << 8 | cl(r.box.x - f.x0) << 16 |
                              cl(r.box.y - f.x0) << 24
                        : 128u << 8 | 128u << 24;
    }
}

// Unit E of a thread's stage: even E -> V row E/2, odd E -> U column E/2.
// Value: (Re, Im) of the rank-1 and the rank-2 entry.
// MASK: rows / columns outside the point's box get zero amplitude.
// F32 (the 1-term variant): the offsets, N and the phase in FP32, reduced by
// v_fract_f32.  Where the range test holds on a tile the local frequency is
// below half a cycle per pixel, so the phase stays below ~35 cycles and
// carries below 1e-5 cycles of FP32 rounding, far below that variant's BF16
// rounding of the entries (2^-9 of the term).
template <int E, bool MASK, bool F32>
__device__ inline float4 unit_value(const Pt& P, const Frame& f, int grp)
{
    constexpr int I = E >> 1;
    if constexpr (F32 && (E & 1) == 0) {
        const float v = fmaf((float)I, f.dvf, f.vbf);
        const float N = v * ((float)P.dy2 + v);
        const float ph = __builtin_amdgcn_fractf(N * fmaf((float)P.k, N, (float)P.h));
        float a = fmaf(fmaf(-P.beta, N, P.mu), N, 1.f);
        if constexpr (MASK) {
            const unsigned row = grp + RG * I;
            a = (row >= __builtin_amdgcn_ubfe(P.box, 0, 8) && row < __builtin_amdgcn_ubfe(P.box, 8, 8)) ? a : 0.f;
        }
        const float c = a * cos_rev(ph), s = a * sin_rev(ph);
        return make_float4(c, s, N * c, N * s);
    } else if constexpr (F32) {
        const float u = fmaf((float)I, f.duf, f.ubf);
        const float n = u * ((float)P.dx2 + u);
        const float ph = __builtin_amdgcn_fractf(fmaf(n, fmaf((float)P.k, n, (float)P.h), (float)P.phic));
        float a = fmaf(fmaf(-P.acal, n, P.acmu), n, P.ac);
        if constexpr (MASK) {
            const unsigned col = grp + RG * I;
            a = (col >= __builtin_amdgcn_ubfe(P.box, 16, 8) && col < __builtin_amdgcn_ubfe(P.box, 24, 8)) ? a : 0.f;
        }
        const float c = a * cos_rev(ph), s = a * sin_rev(ph);
        const float gn = P.g * n;
        return make_float4(c, s, -gn * s, gn * c);
    } else if constexpr ((E & 1) == 0) {
        // the row offsets are recomputed per stage instead of being hoisted
        // out of the K loop, which would hold eight FP64 values in registers
        double vb = f.vb;
        asm volatile("" : "+v"(vb));
        const double v = fma((double)I, f.dv, vb);
        const double N = v * (P.dy2 + v);
        const float ph = (float)frac64(N * fma(P.k, N, P.h));
        const float Nf = (float)N;
        float a = fmaf(fmaf(-P.beta, Nf, P.mu), Nf, 1.f);   // (1 + mu N)(1 - beta N^2) to first order in each
        if constexpr (MASK) {
            const unsigned row = grp + RG * I;
            a = (row >= __builtin_amdgcn_ubfe(P.box, 0, 8) && row < __builtin_amdgcn_ubfe(P.box, 8, 8)) ? a : 0.f;
        }
        const float c = a * cos_rev(ph), s = a * sin_rev(ph);
        return make_float4(c, s, Nf * c, Nf * s);
    } else {
        double ub = f.ub;
        asm volatile("" : "+v"(ub));
        const double u = fma((double)I, f.du, ub);
        const double n = u * (P.dx2 + u);
        const float ph = (float)frac64(fma(n, fma(P.k, n, P.h), P.phic));
        const float nf = (float)n;
        float a = fmaf(fmaf(-P.acal, nf, P.acmu), nf, P.ac);
        if constexpr (MASK) {
            const unsigned col = grp + RG * I;
            a = (col >= __builtin_amdgcn_ubfe(P.box, 16, 8) && col < __builtin_amdgcn_ubfe(P.box, 24, 8)) ? a : 0.f;
        }
        const float c = a * cos_rev(ph), s = a * sin_rev(ph);
        const float gn = P.g * nf;
        return make_float4(c, s, -gn * s, gn * c);
    }
}
// Store: K entries 2p (rank 1) and 2p+1 (rank 2) of the unit's row / column,
// hi plane and (3-term) lo plane, one 8-byte LDS write each.
template <int VAR, int E>
__device__ inline void unit_store(unsigned* wv, unsigned* wu, float4 w, int p, int grp)
{
    constexpr int I = E >> 1;
    unsigned* hi = ((E & 1) == 0 ? wv : wu) + (grp + RG * I) * ROWW + 2 * p;
    const unsigned h1 = pack2<false>(w.x, w.y);
    const unsigned h2 = pack2<false>(w.z, w.w);
    *reinterpret_cast<uint2*>(hi) = make_uint2(h1, h2);
    if constexpr (Traits<VAR>::three) {
        const unsigned l1 = pack2<false>(w.x - lo16<false>(h1), w.y - hi16<false>(h1));
        *reinterpret_cast<uint2*>(hi + PLANE_V) = make_uint2(l1, 0u);
    }
}

// Generation schedule.  3-term (48 MFMA slots): prep piece K at slot K, the
// next stage's point loads at slot NPR, half-unit h (even: value of unit h/2,
// odd: its store) at slot H0 + (42 - H0) h / (2 RU).  1-term (16 slots): prep
// pieces two per slot, loads at slot (NPR + 1) / 2, unit E at slot
// E0 + (14 - E0) E / RU.  Both end before the stage's last MFMA group, within
// which the barrier comes (see step).
constexpr int H0 = NPR + 1, E0 = (NPR + 1) / 2 + 1;
constexpr int half_at(int S)
{
    for (int h = 0; h < 2 * RU; ++h)
        if (H0 + ((42 - H0) * h) / (2 * RU) == S) return h;
    return -1;
}
constexpr int unit_at(int S)
{
    for (int e = 0; e < RU; ++e)
        if (E0 + ((14 - E0) * e) / RU == S) return e;
    return -1;
}
constexpr bool schedule_complete()
{
    for (int e = 0; e < RU; ++e) {
        int n = 0;
        for (int S = 0; S < 16 - 2; ++S) n += unit_at(S) == e;
        if (n != 1) return false;
    }
    for (int h = 0; h < 2 * RU; ++h) {
        int n = 0;
        for (int S = 0; S < 48 - 6; ++S) n += half_at(S) == h;
        if (n != 1) return false;
    }
    return half_at(NPR) < 0 && unit_at((NPR + 1) / 2) < 0 && E0 > (NPR + 1) / 2;
}
static_assert(schedule_complete(), "generation schedule must cover every unit once, after the prep");

template <int VAR>
struct StageCtx {
    f32x16 (&acc)[4][2];
    Raw& R;
    Rsrc rs;
    const Frame& f;
    int pnext, ke;
    int p, grp, lrow, koff, ucol;
    Pt P;
    Tmp t;
    u32x4 A[8][2], Ub[2][2], B[2][4];
    float4 cs[RU];
};

template <int VAR, bool MASK, int S, int END>
__device__ inline void slot(const unsigned* __restrict__ rv, const unsigned* __restrict__ ru,
                           unsigned* __restrict__ wv, unsigned* __restrict__ wu, StageCtx<VAR>& c)
{
    constexpr bool THREE = Traits<VAR>::three;
    if constexpr (S < END) {
        mfma_slot<VAR, S>(rv, ru, c);
        if constexpr (THREE && S < NPR) prep_piece<S, MASK>(c.P, c.t, c.R, c.f);
        if constexpr (!THREE && 2 * S < NPR) prep_piece<2 * S, MASK>(c.P, c.t, c.R, c.f);
        if constexpr (!THREE && 2 * S + 1 < NPR) prep_piece<2 * S + 1, MASK>(c.P, c.t, c.R, c.f);
        if constexpr (S == (THREE ? NPR : (NPR + 1) / 2)) load_raw<MASK>(c.R, c.rs, c.f.tm, c.pnext, c.ke);
        if constexpr (THREE) {
            constexpr int h = half_at(S);
            if constexpr (h >= 0 && (h & 1) == 0) c.cs[h >> 1] = unit_value<(h >> 1), MASK, false>(c.P, c.f, c.grp);
            if constexpr (h >= 0 && (h & 1) == 1) unit_store<VAR, (h >> 1)>(wv, wu, c.cs[h >> 1], c.p, c.grp);
        } else {
            constexpr int e = unit_at(S);
            if constexpr (e >= 0) unit_store<VAR, e>(wv, wu, unit_value<e, MASK, true>(c.P, c.f, c.grp), c.p, c.grp);
        }
        __builtin_amdgcn_sched_barrier(0);
        slot<VAR, MASK, S + 1, END>(rv, ru, wv, wu, c);
    }
}

// Whole-stage generation without MFMAs (the prologue's stage 0).
template <int VAR, bool MASK, int E>
__device__ inline void gen_units(unsigned* wv, unsigned* wu, const Pt& P, const Frame& f, int p, int grp)
{
    if constexpr (E < RU) {
        unit_store<VAR, E>(wv, wu, unit_value<E, MASK, !Traits<VAR>::three>(P, f, grp), p, grp);
        gen_units<VAR, MASK, E + 1>(wv, wu, P, f, p, grp);
    }
}
template <bool MASK>
__device__ inline void prep_all(Pt& P, const Raw& r, const Frame& f)
{
    Tmp t;
    prep_piece<0, MASK>(P, t, r, f);
    prep_piece<1, MASK>(P, t, r, f);
    prep_piece<2, MASK>(P, t, r, f);
    prep_piece<3, MASK>(P, t, r, f);
    prep_piece<4, MASK>(P, t, r, f);
}

// One pipeline step: MFMAs of stage st (buffer RB), generation of stage st+1
// (buffer RB^1) from R, then R <- stage st+2's point data.  Past the last
// stage the generation runs on clamped, zero-amplitude data into the unused
// buffer, which keeps the step branch-free.  The stage's first operands come
// from cy; its barrier comes in the middle of the last MFMA group (whose
// operands are in registers, and whose first half keeps the matrix pipe busy
// while the wave waits), after which the next stage's first operands are read
// into cy from buffer RB^1 while the rest of the group runs.
template <int VAR, bool MASK, int RB>
__device__ inline void step(Smem<Traits<VAR>::three>& sm, f32x16 (&acc)[4][2], Raw& R, Rsrc rs, const Frame& f,
                            int pnext, int ke, int p, int grp, int lrow, int koff, int ucol, Carry& cy)
{
    constexpr bool THREE = Traits<VAR>::three;
    constexpr int NM = mfma_per_stage<VAR>(), T = NM - tail_mfmas<VAR>() / 2;
    const unsigned* __restrict__ rv = &sm.v[RB][0][0][0];
    const unsigned* __restrict__ ru = &sm.u[RB][0][0][0];
    unsigned* __restrict__ wv = &sm.v[RB ^ 1][0][0][0];
    unsigned* __restrict__ wu = &sm.u[RB ^ 1][0][0][0];
    StageCtx<VAR> c{acc, R, rs, f, pnext, ke, p, grp, lrow, koff, ucol};
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
    slot<VAR, MASK, 0, T>(rv, ru, wv, wu, c);
    __syncthreads();
    load_b<VAR>(wu, cy, 0);
    load_a<VAR>(wv, cy, 0);
    if constexpr (!THREE) load_a<VAR>(wv, cy, 1);
    __builtin_amdgcn_sched_barrier(0);
    slot<VAR, MASK, T, NM>(rv, ru, wv, wu, c);
}

// S = sum_p U V^T over the np records at rec (all matrix-core eligible, at
// most MAX_REC), split-K: workgroup (tile, split) sums kslice points into
// plane `split` of part.  MASK: the points' boxes may be smaller than the
// hologram, out-of-box rows and columns get zero amplitude, and so do the
// tiles whose bit is clear in the point's row of mask (one row of
// mask_words words per record; null mask: none).
template <int VAR, bool MASK>
__global__ __launch_bounds__(NT, 2) void k_rs_gemm(Geometry g, const RsRec* __restrict__ rec,
                                                   float2* __restrict__ part, int tiles_x, int tiles, int np,
                                                   int kslice, const unsigned* __restrict__ mask, int mask_words)
{
    constexpr bool THREE = Traits<VAR>::three;
    __shared__ __align__(16) Smem<THREE> sm;

    const int tile = blockIdx.x % tiles;
    const int split = blockIdx.x / tiles;
    const int x0 = (tile % tiles_x) * BN, y0 = (tile / tiles_x) * BM;
    const int kb = split * kslice;
    const int ke = min(np, kb + kslice);
    const int nst = (ke - kb + RP - 1) / RP;
    const int p = threadIdx.x & (RP - 1), grp = threadIdx.x / RP;
    const Rsrc rs = rec_rsrc(rec, np * (int)sizeof(RsRec));
    Frame f;
    f.yc = gpix_y(g, y0 + BM / 2);
    f.xc = gpix_x(g, x0 + BN / 2);
    f.vb = -(double)(grp - BM / 2) * g.pp_y;   // row index grows toward -y
    f.dv = -(double)RG * g.pp_y;
    f.ub = (double)(grp - BN / 2) * g.pp_x;
    f.du = (double)RG * g.pp_x;
    f.vbf = (float)f.vb;
    f.dvf = (float)f.dv;
    f.ubf = (float)f.ub;
    f.duf = (float)f.du;
    f.il = g.inv_lambda;
    f.hl = 0.5 * g.inv_lambda;
    f.um = (BN / 2) * g.pp_x;
    f.vm = (BM / 2) * g.pp_y;
    f.tm = TileMask{mask, mask_words, rs_tile_index(g, x0, y0)};
    f.x0 = x0;
    f.y0 = y0;
    f.sg = split_sign(split);

    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int lrow = lane & 31, koff = (lane >> 5) * 4;
    const int ucol = wave * 32 + lrow;

    Raw R;
    if (nst > 0) {
        load_raw<MASK>(R, rs, f.tm, kb + p, ke);
        Pt P;
        prep_all<MASK>(P, R, f);
        gen_units<VAR, MASK, 0>(&sm.v[0][0][0][0], &sm.u[0][0][0][0], P, f, p, grp);
        load_raw<MASK>(R, rs, f.tm, kb + RP + p, ke);
    }
    __syncthreads();
    Carry cy;
    cy.lrow = lrow;
    cy.koff = koff;
    cy.ucol = ucol;
    load_b<VAR>(&sm.u[0][0][0][0], cy, 0);
    load_a<VAR>(&sm.v[0][0][0][0], cy, 0);
    if constexpr (!Traits<VAR>::three) load_a<VAR>(&sm.v[0][0][0][0], cy, 1);
    f32x16 acc[4][2];
#pragma unroll
    for (int i = 0; i < 4; ++i) acc[i][0] = acc[i][1] = f32x16{};
    // no branch between the two steps, so that the next operands' reads stay
    // ahead of the last MFMAs; for an odd stage count the second step runs one
    // stage past the end, whose operands (generated from clamped, zero-amplitude
    // point data) contribute exact zeros
    for (int st = 0; st < nst; st += 2) {
        step<VAR, MASK, 0>(sm, acc, R, rs, f, kb + (st + 2) * RP + p, ke, p, grp, lrow, koff, ucol, cy);
        step<VAR, MASK, 1>(sm, acc, R, rs, f, kb + (st + 3) * RP + p, ke, p, grp, lrow, koff, ucol, cy);
    }
    store_partials(g, acc, part, split, x0, y0, ucol, lane);
}

// The reference's range test (ophGen.cpp:415-429) with a RANGE_TOL margin on
// every pixel of columns [c0, c1) x rows [r0, r1): bounded with the extreme
// distances of the rectangle from the point.
__device__ inline bool range_ok_rect(const Geometry& g, double px, double py, double z, int c0, int c1, int r0,
                                     int r1)
{
    const double xl = gpix_x(g, c0), xh = gpix_x(g, c1 - 1);
    const double yh = gpix_y(g, r0), yl = gpix_y(g, r1 - 1);
    const double dxmax = fmax(fabs(xl - px), fabs(xh - px));
    const double dymax = fmax(fabs(yl - py), fabs(yh - py));
    const double dxmin = (px >= xl && px <= xh) ? 0.0 : fmin(fabs(xl - px), fabs(xh - px));
    const double dymin = (py >= yl && py <= yh) ? 0.0 : fmin(fabs(yl - py), fabs(yh - py));
    const double zz = z * z;
    return fabs(g.hx_coef) * sqrt(dymin * dymin + zz) - dxmax > RANGE_TOL &&
           fabs(g.hy_coef) * sqrt(dxmin * dxmin + zz) - dymax > RANGE_TOL;
}

// Class of point i as a scan key: 1 (unmasked matrix-core instance), 1 << 32
// (masked instance), 0 (direct kernel only).  The unmasked instance takes the
// points whose box is the hologram, on which the range test holds with margin
// (the mask is then the box) and whose expansion bound (rs_bound_ok) holds on
// every tile.  Without split, the masked instance takes the other points whose
// range test and bound hold on the whole box (out-of-box rows and columns get
// zero amplitude).  With split, it takes every point whose bound may hold on
// its nearest tile, on the tiles where k_rs_tilemask finds the range test and
// the bound both hold; the direct kernel takes the other tiles.
__global__ __launch_bounds__(256) void k_rs_classify(Geometry g, PointSoA s, double max_theta, int split,
                                                     unsigned long long* __restrict__ key)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= s.n) return;
    const double px = s.px[i], py = s.py[i], z = s.pz[i];
    const int4 b = s.box[i];
    const bool some = z > 0.0 && b.x < b.y && b.z < b.w;
    const bool full = b.x == 0 && b.y == g.pn_x && b.z == 0 && b.w == g.pn_y;
    const int bx1 = some ? b.y - 1 : 0, by1 = some ? b.w - 1 : 0, bx0 = some ? b.x : 0, by0 = some ? b.z : 0;
    const bool range = some && range_ok_rect(g, px, py, z, bx0, bx1 + 1, by0, by1 + 1);
    // the expansion bound on the farthest and on the nearest tile reference
    // pixels of the box (the bound grows with the distances)
    const double xc0 = rs_tile_xc(g, bx0), xc1 = rs_tile_xc(g, bx1);
    const double yc0 = rs_tile_yc(g, by0), yc1 = rs_tile_yc(g, by1);
    const double dxc = fmax(fabs(xc0 - px), fabs(xc1 - px));
    const double dyc = fmax(fabs(yc0 - py), fabs(yc1 - py));
    const bool all_tiles = rs_bound_ok(g, dxc, dyc, z, max_theta);
    const double dxn = (px >= fmin(xc0, xc1) && px <= fmax(xc0, xc1)) ? 0.0 : fmin(fabs(xc0 - px), fabs(xc1 - px));
    const double dyn = (py >= fmin(yc0, yc1) && py <= fmax(yc0, yc1)) ? 0.0 : fmin(fabs(yc0 - py), fabs(yc1 - py));
    const bool near_tile = rs_bound_ok(g, dxn, dyn, z, max_theta);
    key[i] = range && all_tiles && full ? 1ull
             : (split ? some && near_tile : range && all_tiles) ? 1ull << 32
                                                                : 0ull;
}

// Tile mask of the masked instance's points s[0, n): word w of row i holds
// bits for tiles 32 w .. 32 w + 31 (rs_tile_index); a bit is set when the
// tile meets the point's box and, on the part inside the box, the range test
// holds with margin (the mask there is the box) and the expansion bound holds.
__global__ __launch_bounds__(256) void k_rs_tilemask(Geometry g, PointSoA s, double max_theta, int words,
                                                     unsigned* __restrict__ mask)
{
    const int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (idx >= s.n * words) return;
    const int64_t i = idx / words;
    const int w = (int)(idx - i * words);
    const double px = s.px[i], py = s.py[i], z = s.pz[i];
    const int4 b = s.box[i];
    const int tx_n = rs_tiles_x(g), tiles = tx_n * ((g.pn_y + RS_TILE - 1) / RS_TILE);
    unsigned bits = 0u;
    for (int k = 0; k < 32; ++k) {
        const int t = w * 32 + k;
        if (t >= tiles) break;
        const int c = (t % tx_n) * RS_TILE, r = (t / tx_n) * RS_TILE;
        const int c0 = max(c, b.x), c1 = min(min(c + RS_TILE, g.pn_x), b.y);
        const int r0 = max(r, b.z), r1 = min(min(r + RS_TILE, g.pn_y), b.w);
        if (c0 >= c1 || r0 >= r1 || !(z > 0.0)) continue;
        if (rs_bound_ok(g, fabs(rs_tile_xc(g, c) - px), fabs(rs_tile_yc(g, r) - py), z, max_theta) &&
            range_ok_rect(g, px, py, z, c0, c1, r0, r1))
            bits |= 1u << k;
    }
    mask[idx] = bits;
}

// Stable three-way partition: unmasked matrix-core points to o[0, n0), masked
// ones to o[n0, n0 + n1), the others after them, each in point order; pos =
// exclusive prefix sum of key (low word: n0 before i, high word: n1 before i).
// The matrix-core kernels' records go to rec in the same order.
// count[0] = n0 + n1, count[1] = n0.
__global__ __launch_bounds__(256) void k_rs_gather(PointSoA s, const unsigned long long* __restrict__ key,
                                                   const unsigned long long* __restrict__ pos, PointSoA o,
                                                   RsRec* __restrict__ rec, unsigned long long* __restrict__ count)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int64_t n = s.n;
    if (i >= n) return;
    const unsigned long long tot = pos[n - 1] + key[n - 1];
    const int64_t n0 = (int64_t)(tot & 0xffffffffull), n1 = (int64_t)(tot >> 32);
    const unsigned long long k = key[i], ps = pos[i];
    const int64_t lo = (int64_t)(ps & 0xffffffffull), hi = (int64_t)(ps >> 32);
    const int64_t j = k == 1ull ? lo : k ? n0 + hi : n0 + n1 + (i - lo - hi);
    o.px[j] = s.px[i];
    o.py[j] = s.py[i];
    o.pz[j] = s.pz[i];
    o.cyc0[j] = s.cyc0[i];
    o.kamp[j] = s.kamp[i];
    o.q[j] = s.q[i];
    o.box[j] = s.box[i];
    rec[j] = RsRec{s.px[i], s.py[i], s.pz[i], s.cyc0[i], s.kamp[i], 0.0, s.box[i]};
    if (i == 0) {
        count[0] = (unsigned long long)(n0 + n1);
        count[1] = (unsigned long long)n0;
    }
}

}  // namespace rsgemm

namespace {
size_t al256(size_t b) { return (b + 255) & ~size_t(255); }

PointSoA soa_tail(const PointSoA& s, int64_t off)
{
    PointSoA t = s;
    t.px += off;
    t.py += off;
    t.pz += off;
    t.cyc0 += off;
    t.kamp += off;
    t.q += off;
    t.box += off;
    t.n = s.n - off;
    return t;
}

size_t scan_temp_bytes(int64_t n)
{
    size_t b = 0;
    if (hipcub::DeviceScan::ExclusiveSum(nullptr, b, (const unsigned long long*)nullptr, (unsigned long long*)nullptr,
                                         (int)n, (hipStream_t)0) != hipSuccess)
        return 0;
    return b;
}
constexpr size_t MAX_RS_PART = size_t(4) << 30;
constexpr size_t MAX_RS_MASK = size_t(1) << 30;
}  // namespace

// [count words, 256 B] [key] [pos] [scan temp] [records]; beyond INT_MAX
// points (the scan's item count) a size no budget admits
static size_t records_offset(int64_t n)
{
    return 256 + 2 * al256((size_t)n * sizeof(unsigned long long)) + al256(scan_temp_bytes(n));
}
size_t rs_partition_scratch_bytes(int64_t n)
{
    if (n > INT_MAX) return SIZE_MAX / 4;
    return records_offset(n) + al256((size_t)n * sizeof(rsgemm::RsRec)) + 256;
}
const void* rs_partition_records(const void* scratch, int64_t n)
{
    return (const char*)scratch + records_offset(n);
}

static double max_theta_of(uint32_t mode)
{
    return (mode & OHX_MODE_MFMA_BF16X3) ? rsgemm::MAX_THETA_X3 : rsgemm::MAX_THETA_X1;
}

int launch_rs_partition(hipStream_t st, void* scratch, size_t scratch_bytes, const Geometry& g, const PointSoA& soa,
                        const PointSoA& part, uint32_t mode, bool split, unsigned long long** d_count)
{
    using namespace rsgemm;
    const int64_t n = soa.n;
    if (n <= 0 || n > INT_MAX || part.n != n) return OHX_ERR_INVALID_ARG;
    size_t cub_bytes = scan_temp_bytes(n);
    if (!cub_bytes || rs_partition_scratch_bytes(n) > scratch_bytes) return OHX_ERR_INVALID_ARG;
    char* b = (char*)scratch;
    *d_count = (unsigned long long*)b;
    unsigned long long* key = (unsigned long long*)(b + 256);
    unsigned long long* pos = (unsigned long long*)(b + 256 + al256((size_t)n * sizeof(unsigned long long)));
    void* tmp = b + 256 + 2 * al256((size_t)n * sizeof(unsigned long long));
    const unsigned pgrid = (unsigned)((n + 255) / 256);
    hipLaunchKernelGGL(k_rs_classify, dim3(pgrid), dim3(256), 0, st, g, soa, max_theta_of(mode), split ? 1 : 0, key);
    if (hipcub::DeviceScan::ExclusiveSum(tmp, cub_bytes, (const unsigned long long*)key, pos, (int)n, st) !=
        hipSuccess)
        return OHX_ERR_DEVICE;
    hipLaunchKernelGGL(k_rs_gather, dim3(pgrid), dim3(256), 0, st, soa, (const unsigned long long*)key,
                       (const unsigned long long*)pos, part, (RsRec*)(b + records_offset(n)), *d_count);
    return hipGetLastError() == hipSuccess ? OHX_OK : OHX_ERR_DEVICE;
}

// Enough split-K planes for any launch of at most npoints points (the plan's
// split count, before rounding the slice to whole stages, bounds them all).
size_t rs_gemm_scratch_bytes(const Geometry& g, int64_t npoints)
{
    return (size_t)gemm_max_splits(g, npoints, rsgemm::RP, gemm::BN) * g.pn_x * g.pn_y * sizeof(float2);
}

namespace {
// One GEMM + reduce over records [p0, p1) (fewer than MAX_REC); mask: the
// tile-mask row of record mask_base.
int rs_gemm_piece(hipStream_t st, void* scratch, size_t scratch_bytes, const Geometry& g, const rsgemm::RsRec* rec,
                  int64_t p0, int64_t p1, bool masked, double* d_field, uint32_t mode, int64_t* wgs,
                  const unsigned* mask, int64_t mask_base, int mask_words)
{
    using namespace rsgemm;
    const GemmPlan pl = plan_gemm(g, p1 - p0, RP);
    if ((size_t)pl.splits * pl.npx * sizeof(float2) > scratch_bytes) return OHX_ERR_INVALID_ARG;
    const dim3 grid((unsigned)((int64_t)pl.tiles * pl.splits)), block(NT);
    float2* pp = (float2*)scratch;
    const bool x3 = (mode & OHX_MODE_MFMA_BF16X3) != 0;
    const unsigned* mk = (masked && mask) ? mask + (p0 - mask_base) * mask_words : nullptr;
    const RsRec* r = rec + p0;
    const int np = (int)(p1 - p0), ks = (int)pl.kslice;
    if (x3 && masked)
        hipLaunchKernelGGL((k_rs_gemm<V_BF16X3, true>), grid, block, 0, st, g, r, pp, pl.tiles_x, pl.tiles, np, ks,
                           mk, mask_words);
    else if (x3)
        hipLaunchKernelGGL((k_rs_gemm<V_BF16X3, false>), grid, block, 0, st, g, r, pp, pl.tiles_x, pl.tiles, np, ks,
                           mk, mask_words);
    else if (masked)
        hipLaunchKernelGGL((k_rs_gemm<V_BF16X1, true>), grid, block, 0, st, g, r, pp, pl.tiles_x, pl.tiles, np, ks,
                           mk, mask_words);
    else
        hipLaunchKernelGGL((k_rs_gemm<V_BF16X1, false>), grid, block, 0, st, g, r, pp, pl.tiles_x, pl.tiles, np, ks,
                           mk, mask_words);
    if (hipGetLastError() != hipSuccess) return OHX_ERR_DEVICE;
    if (wgs) *wgs += (int64_t)pl.tiles * pl.splits;
    return launch_gemm_reduce(st, pp, pl.splits, pl.npx, d_field);
}
int rs_gemm_launch(hipStream_t st, void* scratch, size_t scratch_bytes, const Geometry& g, const void* rec,
                   int64_t p0, int64_t p1, bool masked, double* d_field, uint32_t mode, int64_t* wgs,
                   const unsigned* mask, int64_t mask_base, int mask_words)
{
    int e = OHX_OK;
    for (int64_t q0 = p0; q0 < p1 && !e; q0 += gemm::MAX_REC)
        e = rs_gemm_piece(st, scratch, scratch_bytes, g, (const rsgemm::RsRec*)rec, q0,
                          std::min(p1, q0 + gemm::MAX_REC), masked, d_field, mode, wgs, mask, mask_base,
                          mask_words);
    return e;
}
}  // namespace

int launch_rs_gemm_range(hipStream_t st, void* scratch, size_t scratch_bytes, const Geometry& g, const void* rec,
                         int64_t p0, int64_t p1, int64_t n_full, double* d_field, uint32_t mode, int64_t* wgs,
                         const unsigned* mask, int mask_words)
{
    const int64_t m = std::min(std::max(n_full, p0), p1);
    int e = rs_gemm_launch(st, scratch, scratch_bytes, g, rec, p0, m, false, d_field, mode, wgs, nullptr, 0, 0);
    if (!e)
        e = rs_gemm_launch(st, scratch, scratch_bytes, g, rec, m, p1, true, d_field, mode, wgs, mask, n_full,
                           mask_words);
    return e;
}

int launch_rs_gemm(ohx_ctx* ctx, const Geometry& g, const PointSoA& soa, double* d_field, uint32_t mode)
{
    using namespace rsgemm;
    const int64_t n = soa.n;
    if (n <= 0) return OHX_OK;
    if (n > INT_MAX) return set_error(ctx, OHX_ERR_INVALID_ARG, "too many points for the RS matrix-core path");
    hipStream_t st = ctx->stream;
    // [partition scratch] [partitioned point constants]
    const size_t pbytes = al256(rs_partition_scratch_bytes(n));
    const size_t need = pbytes + points_bytes(n);
    if (need > ctx->rs_scratch_bytes) {
        if (ctx->rs_scratch) {
            int e = check_hip(ctx, hipStreamSynchronize(st), "hipStreamSynchronize");
            if (e) return e;
            (void)hipFree(ctx->rs_scratch);
            ctx->rs_scratch = nullptr;
            ctx->rs_scratch_bytes = 0;
        }
        int e = check_hip(ctx, hipMalloc(&ctx->rs_scratch, need), "hipMalloc(RS scratch)");
        if (e) return e;
        ctx->rs_scratch_bytes = need;
    }
    PointSoA part;
    carve_points((char*)ctx->rs_scratch + pbytes, n, &part);
    // per-tile split between the masked instance and the direct kernel, unless
    // its mask (one row per masked point) could exceed MAX_RS_MASK
    const int words = rs_mask_words(g);
    const bool split = (size_t)n * words * sizeof(unsigned) <= MAX_RS_MASK;
    unsigned long long* d_count = nullptr;
    int e = launch_rs_partition(st, ctx->rs_scratch, pbytes, g, soa, part, mode, split, &d_count);
    if (e) return set_error(ctx, e, "RS matrix-core partition failed");
    unsigned long long cnt[2] = {0, 0};
    e = check_hip(ctx, hipMemcpyAsync(cnt, d_count, sizeof(cnt), hipMemcpyDeviceToHost, st), "hipMemcpyAsync");
    if (!e) e = check_hip(ctx, hipStreamSynchronize(st), "hipStreamSynchronize");
    if (e) return e;
    const int64_t ne = (int64_t)cnt[0], n_full = (int64_t)cnt[1];

    int64_t wgs = 0;
    const int64_t nm = split ? ne - n_full : 0;   // points of the masked instance with a tile mask
    unsigned* mask = nullptr;
    if (ne > 0) {
        int64_t chunk = ne;
        while (chunk > RP && (rs_gemm_scratch_bytes(g, chunk) > MAX_RS_PART || chunk > gemm::MAX_REC))
            chunk = (chunk + 1) / 2;
        if ((e = ensure_gemm_scratch(ctx, rs_gemm_scratch_bytes(g, chunk)))) return e;
        if (nm > 0) {
            const size_t mbytes = (size_t)nm * words * sizeof(unsigned);
            if (mbytes > ctx->rs_mask_bytes) {
                if (ctx->rs_mask) {
                    if ((e = check_hip(ctx, hipStreamSynchronize(st), "hipStreamSynchronize"))) return e;
                    (void)hipFree(ctx->rs_mask);
                    ctx->rs_mask = nullptr;
                    ctx->rs_mask_bytes = 0;
                }
                if ((e = check_hip(ctx, hipMalloc(&ctx->rs_mask, mbytes), "hipMalloc(RS tile mask)"))) return e;
                ctx->rs_mask_bytes = mbytes;
            }
            mask = (unsigned*)ctx->rs_mask;
            PointSoA masked = soa_tail(part, n_full);
            masked.n = nm;
            const int64_t items = nm * words;
            hipLaunchKernelGGL(k_rs_tilemask, dim3((unsigned)((items + 255) / 256)), dim3(256), 0, st, g, masked,
                               max_theta_of(mode), words, mask);
            if ((e = check_hip(ctx, hipGetLastError(), "k_rs_tilemask"))) return e;
        }
        for (int64_t p0 = 0; p0 < ne; p0 += chunk) {
            const int64_t p1 = std::min<int64_t>(ne, p0 + chunk);
            e = launch_rs_gemm_range(st, ctx->gemm_scratch, ctx->gemm_scratch_bytes, g,
                                     rs_partition_records(ctx->rs_scratch, n), p0, p1, n_full, d_field, mode, &wgs,
                                     mask, words);
            if (e) return set_error(ctx, e, "RS matrix-core launch failed");
        }
    }
    // the masked instance took the tiles set in the mask: the direct kernel
    // takes its points' other tiles (rs_skip_mask), then every tile of the
    // points outside the matrix-core classes
    if (nm > 0) {
        Geometry gs = g;
        gs.rs_skip_mask = mask;
        gs.rs_mask_words = words;
        PointSoA masked = soa_tail(part, n_full);
        masked.n = nm;
        if ((e = launch_direct_with(ctx, gs, masked, d_field, false, ctx->minmax + 5, true))) return e;
        wgs += ctx->last.workgroups;
    }
    if (ne < n) {
        if ((e = launch_direct_with(ctx, g, soa_tail(part, ne), d_field, false, ctx->minmax + 5, true))) return e;
        wgs += ctx->last.workgroups;
    }
    ctx->last.workgroups = wgs;
    ctx->last.kernel_kind = ne > 0 ? 2 : 0;
    ctx->last.gemm_points = (int32_t)ne;
    return OHX_OK;
}

}  // namespace ohx
