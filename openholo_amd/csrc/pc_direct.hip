// Direct point-cloud diffraction sums for gfx950.
//
// Replaces the eight per-pixel kernels of ophgen/src/ophPCKernel.cu:66-872
// (RS / Fresnel x double / single x plain / round-to-zero), whose maths is
// ophGen::RS_Diffraction / Fresnel_Diffraction (ophgen/src/ophGen.cpp:361-514).
//
// Field model (per point p, pixel (x, y), cycles = radians / 2pi):
//   RS:      H += a z / (lambda r^2) * (sin 2pi c - j cos 2pi c),  c = r / lambda
//   Fresnel: H += a / (lambda z)     * (sin 2pi c - j cos 2pi c),  c = z/lambda + rho^2/(2 lambda z)
// over the reference's anti-aliasing mask (integer box of ophGen.cpp:384-407 /
// 467-490, plus the RS per-pixel range test of ophGen.cpp:415-429).
//
// Design (one workgroup = one 64 x 32 pixel tile, 4 waves, 8 rows per lane):
//  * k_pc_prep turns the 56-byte Vertex records into FP64 structure-of-arrays
//    constants, including the exact integer mask box (same FP64 operations,
//    in the same order, as the reference, contraction off) and the per-point
//    phase constant z/lambda reduced mod 1 cycle with an exact remainder.
//  * Per 128-point chunk, one thread per point classifies the (point, tile)
//    pair in FP64 -- OUT (mask empty on the tile), FAST (whole tile inside
//    the mask, with margin), MASKED (per-pixel mask needed) or EXACT (FP32
//    series not accurate enough: FP64 terms) -- and writes FP32 constants
//    *recentred on the tile centre* into LDS, compacted in point order
//    (ballot + mbcnt, deterministic).  Recentring keeps every FP32 phase
//    below ~10 cycles, so FP32 rounding costs ~1e-6 rad instead of the ~1 rad
//    a direct FP32 evaluation of k r ~ 6e6 rad would.
//  * Fresnel phase is separable: c = phi_c + u (g_x + q u) + v (g_y + q v),
//    so a term is A(col) B(row) with A = (sin, cos)(2 pi (phi_c + u (g_x + q u)))
//    (one v_sin/v_cos per lane and point) and B = amp e^{j 2pi v (g_y + q v)}
//    tabulated in LDS per chunk (zero outside the box rows; A is zeroed
//    outside the box columns): a pair is one complex multiply-add, 4 FMAs.
//  * RS uses N = r^2 - r_c^2 = (2 dx_c u + u^2) + (2 dy_c v + v^2) = n_x + N_y
//    and the series r - r_c = N/(2 r_c) - N^2/(8 r_c^3) (valid bound checked
//    per pair in FP64), amplitude z/r^2 = (z/r_c^2)(1 - N/r_c^2): no sqrt or
//    reciprocal in the loop.  The phase splits into a column part, a row
//    part and the cross term -2 c_2 n_x N_y cycles (c_2 = 1/(8 r_c^3 lambda)),
//    at most MAX_EPS rad, taken as a first-order factor 1 + j eps (the
//    amplitude's cross term is below 1e-7), folded into the column factor
//    per row (P = sa + ca eps, Q = ca - sa eps): 6 FMAs per pair, no v_sin or
//    v_cos.  Pairs with a larger cross term (FASTNS) evaluate the phase per
//    pixel: add, two FMAs, v_sin, v_cos, three FMAs.
//  * FP32 partial sums per chunk are folded into FP64 accumulators; the
//    field (interleaved FP64 complex, the reference's Complex<Real> layout)
//    is read-modify-written once per pixel at the end, 1 KB coalesced per row.
//  * MODE_DOUBLE (every term in FP64) uses the same tiles but separates the
//    phase into a column factor (one FP64 sincos per lane and point) times
//    a row factor (tabulated in LDS per chunk), times a short FP64 series
//    correction for RS; a Fresnel pair is one complex multiply-add.
#include "ohx_internal.h"

namespace ohx {

constexpr int TW = 64;            // tile width = lanes of a wave
constexpr int WAVES = 4;
constexpr int RPT = 8;            // rows per lane
constexpr int TH = WAVES * RPT;   // tile height
constexpr int NT = TW * WAVES;    // threads per workgroup
constexpr int PCH = DIRECT_CHUNK; // points per LDS chunk (three workgroups per CU fit the LDS)

// FAST: whole tile inside the mask, separable form; MASKED: per-pixel mask;
// EXACT: FP64 terms; FASTNS (RS): whole tile inside the mask, but the
// cross term of the separable form is too large, so the phase is evaluated
// per pixel.
enum { CLS_OUT = 0, CLS_FAST = 1, CLS_MASKED = 2, CLS_EXACT = 3, CLS_FASTNS = 4 };

// Largest cross-term phase (rad) of the separable RS form: e^{j eps} is taken
// as 1 + j eps, which is within eps^2/2 <= 4.5e-8.
constexpr double MAX_EPS = 3e-4;

constexpr double INV_2PI = 0.15915494309189533576888376337251436;
constexpr double TWO_PI = 6.283185307179586476925286766559;
// Largest phase excursion (cycles) over a tile that the FP32 path accepts:
// FP32 rounding then stays below ~4e-6 cycles per term; larger spans (points
// a few mm from the hologram) take the FP64 path.
constexpr double MAX_CYC = 64.0;

// ---------------------------------------------------------------------------
// reference pixel coordinates, ophGen.cpp:413 and :422 (int expression times
// pitch, then added; no contraction)
__device__ inline double pix_x(const Geometry& g, int col)
{
#pragma clang fp contract(off)
    return g.x0 + (double)(col - 1 + g.off_x) * g.pp_x;
}
__device__ inline double pix_y(const Geometry& g, int row)
{
#pragma clang fp contract(off)
    return g.y0 + (double)(g.pn_y - row + g.off_y) * g.pp_y;
}

// RS per-pixel range test, ophGen.cpp:415-429, bit-identical operation order
__device__ inline bool rs_range_ok(double px, double py, double zz, double hxc, double hyc,
                                   double xxx, double yyy)
{
#pragma clang fp contract(off)
    const double sy = sqrt((yyy - py) * (yyy - py) + zz);
    const double sx = sqrt((xxx - px) * (xxx - px) + zz);
    const double rx0 = px + fabs(hxc * sy), rx1 = px - fabs(hxc * sy);
    const double ry0 = py + fabs(hyc * sx), ry1 = py - fabs(hyc * sx);
    return ((xxx < rx0) && (xxx > rx1)) && ((yyy < ry0) && (yyy > ry1));
}

// ---------------------------------------------------------------------------
// per-point constants: Vertex (56 B) -> SoA
// boxsum (optional): adds the number of pixels inside the boxes of this
// launch's points (one atomic per workgroup), for throughput accounting.
__global__ __launch_bounds__(256) void k_pc_prep(const double* __restrict__ vtx, Geometry g, PointSoA s,
                                                 unsigned long long* __restrict__ boxsum)
{
#pragma clang fp contract(off)
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    unsigned long long area = 0;
    if (i < s.n) {
        const double* v = vtx + 7 * i;
        // genCghPointCloudCPU: scale, then distance (ophPointCloud.cpp:304-306, ophGen.cpp:379)
        const double px = v[0] * g.scale[0];
        const double py = v[1] * g.scale[1];
        const double z = v[2] * g.scale[2] + g.distance;
        const double a = v[3 + g.channel];
        double hx, hy;
        if (g.diff == OHX_DIFF_RS) {
            hx = fabs(g.hx_coef * z);
            hy = fabs(g.hy_coef * z);
        } else {
            const double operand = g.lambda * z;
            hx = fabs(operand / (2 * g.pp_x));
            hy = fabs(operand / (2 * g.pp_y));
        }
        // box bounds, ophGen.cpp:394-407 (x = -ss/2 = g.x0)
        double X0 = floor(((px + hx) - g.x0) / g.pp_x) + 1;       // exclusive upper column
        double X1 = floor(((px - hx) - g.x0) / g.pp_x) + 1;       // inclusive lower column
        double Y0 = g.pn_y - floor(((py - hy) - g.y0) / g.pp_y);  // exclusive upper row
        double Y1 = g.pn_y - floor(((py + hy) - g.y0) / g.pp_y);  // inclusive lower row
        if (X0 > g.pn_x) X0 = g.pn_x;
        if (X1 < 0) X1 = 0;
        if (Y0 > g.pn_y) Y0 = g.pn_y;
        if (Y1 < 0) Y1 = 0;
        int4 b = make_int4(0, 0, 0, 0);
        if (X1 < X0 && Y1 < Y0) b = make_int4((int)X1, (int)X0, (int)Y1, (int)Y0);

        // z / lambda mod 1 with the exact remainder of the division
        const double qz = z / g.lambda;
        const double rem = fma(-qz, g.lambda, z);
        double c = (qz - floor(qz)) + rem * g.inv_lambda;
        if (g.use_phase) c += v[6] * INV_2PI;
        c -= floor(c);

        s.px[i] = px;
        s.py[i] = py;
        s.pz[i] = z;
        s.cyc0[i] = c;
        if (g.diff == OHX_DIFF_RS) {
            s.kamp[i] = a * z / g.lambda;
            s.q[i] = 0.0;
        } else {
            s.kamp[i] = a / (g.lambda * z);
            s.q[i] = 0.5 / (g.lambda * z);
        }
        s.box[i] = b;
        area = (unsigned long long)(b.y - b.x) * (unsigned long long)(b.w - b.z);
    }
    if (boxsum) {   // uniform over the launch
        for (int off = 32; off > 0; off >>= 1) area += __shfl_xor(area, off);
        __shared__ unsigned long long part[4];
        if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = area;
        __syncthreads();
        if (threadIdx.x == 0) atomicAdd(boxsum, part[0] + part[1] + part[2] + part[3]);
    }
}

__global__ void k_zero_u64(unsigned long long* p)
{
    if (threadIdx.x == 0) *p = 0ull;
}

// ---------------------------------------------------------------------------
// sin_rev / cos_rev (ohx_internal.h): the FP32 phases (below MAX_CYC + 1
// cycles by the classification) go in unreduced.

// FP64 term (1/j) amp exp(j 2pi c) of one (point, pixel) pair, added to (re, im)
template <int DIFF>
__device__ inline void exact_term(double dx, double dy, double z, double cyc0, double kamp,
                                  double q, double inv_lambda, double& re, double& im)
{
    const double rho2 = dx * dx + dy * dy;
    double cyc, amp;
    if (DIFF == OHX_DIFF_RS) {
        const double r2 = rho2 + z * z;
        const double r = sqrt(r2);
        cyc = cyc0 + (rho2 / (r + z)) * inv_lambda;   // r/lambda = z/lambda + (r-z)/lambda
        amp = kamp / r2;                               // a z / (lambda r^2)
    } else {
        cyc = cyc0 + rho2 * q;                         // z/lambda + rho^2/(2 lambda z)
        amp = kamp;                                    // a / (lambda z)
    }
    cyc -= floor(cyc);
    double sn, cs;
    sincospi(2.0 * cyc, &sn, &cs);
    re += amp * sn;
    im -= amp * cs;
}

// LDS of one workgroup (one tile): point constants of the current chunk
template <int DIFF>
struct DirectSmem {
    float4 c0[PCH];
    float4 c1[PCH];
    float2 R[WAVES][PCH][RPT];   // row factor B(row) of the separable form
    float Ny[WAVES][PCH][DIFF == OHX_DIFF_RS ? RPT : 1];   // RS: N_y(row)
    float mu[PCH];               // RS: -1 / r_c^2
    int4 box[PCH];
    int cls[PCH];
    int cnt[2][WAVES];
};

// End of a tile: add the lane's RPT pixels into the field, or, for a point
// split other than the first (part != nullptr), store them into that split's
// partial plane, which k_split_reduce adds later in split order.
__device__ inline void store_tile(const Geometry& g, double2* __restrict__ field, double2* __restrict__ part,
                                  int col, int rowb, const double* accr, const double* acci)
{
    if (col >= g.pn_x) return;
#pragma unroll
    for (int k = 0; k < RPT; ++k) {
        const int row = rowb + k;
        if (row < g.pn_y) {
            const int64_t idx = (int64_t)row * g.pn_x + col;
            if (part) {
                part[idx] = make_double2(accr[k], acci[k]);
            } else {
                double2 f = field[idx];
                f.x += accr[k];
                f.y += acci[k];
                field[idx] = f;
            }
        }
    }
}

// What one direct_tile instance does with the (point, tile) pairs that need
// FP64 terms (EXACT class): EX_INLINE evaluates them in place (multi-job
// kernel); EX_FLAG skips them and raises *exact_flag, which keeps the FP64
// code and its registers out of the FP32 kernel (three waves per SIMD);
// EX_ONLY evaluates nothing but them (the follow-up kernel).
enum { EX_INLINE = 0, EX_FLAG = 1, EX_ONLY = 2 };

// FP64 point data of a slow-list slot (px, py, z, cyc0, kamp, q) sits in that
// slot's row-table entries of wave 0, which these slots leave unused: the
// RS slow classes use N_y only, and Fresnel fills B only for non-EXACT slots,
// reading no FP64 data for them.
__device__ inline double* slot_fp64(float2* row_entry) { return reinterpret_cast<double*>(row_entry); }

// Accumulates points [p_begin, p_end) of s into one 64 x 32 tile of field.
// Always inlined: the multi-job kernel calls all four variants, and a real
// call would pin that kernel at the callees' worst-case register count
// (one wave per SIMD) and add call overhead per tile.
template <int DIFF, int EX>
__device__ __attribute__((always_inline)) inline void direct_tile(DirectSmem<DIFF>& sm, const Geometry& g, const PointSoA& s, int64_t p_begin,
                           int64_t p_end, double2* __restrict__ field, int tile, int tiles_x,
                           unsigned long long* __restrict__ exact_flag, double2* __restrict__ part = nullptr)
{
    float4* s_c0 = sm.c0;
    float4* s_c1 = sm.c1;
    auto& s_R = sm.R;
    int4* s_box = sm.box;
    int* s_cls = sm.cls;
    auto& s_cnt = sm.cnt;

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int tx = tile % tiles_x, ty = tile / tiles_x;
    const int x0 = tx * TW, y0 = ty * TH;
    const int cx1 = min(x0 + TW, g.pn_x), ry1 = min(y0 + TH, g.pn_y);
    const int col = x0 + lane;
    const int rowb = y0 + wave * RPT;

    // tile centre and this lane's offset from it
    const double xc = pix_x(g, x0 + TW / 2), yc = pix_y(g, y0 + TH / 2);
    const float u = (float)((double)(lane - TW / 2) * g.pp_x);
    const float uu = u * u;
    const double xxx = pix_x(g, col);
    const float umax = (float)(TW / 2) * (float)g.pp_x, vmax = (float)(TH / 2) * (float)g.pp_y;
    // v = y_row - y_c of this wave's rows
    float vr[RPT];
#pragma unroll
    for (int k = 0; k < RPT; ++k) vr[k] = (float)(-(double)(wave * RPT + k - TH / 2) * g.pp_y);

    double accr[RPT], acci[RPT];
#pragma unroll
    for (int k = 0; k < RPT; ++k) accr[k] = acci[k] = 0.0;
    bool touched = false;   // EX_ONLY: some slow-list point on this tile

    for (int64_t base = p_begin; base < p_end; base += PCH) {
        // ---- classify and prepare one point per thread --------------------
        int cls = CLS_OUT;
        float4 c0 = make_float4(0, 0, 0, 0), c1 = make_float4(0, 0, 0, 0);
        float mu = 0.f;
        int4 b = make_int4(0, 0, 0, 0);
        double px = 0, py = 0, z = 0, cyc0 = 0, kamp = 0, q = 0;
        const int64_t p = base + tid;
        // RS: the tiles that the masked matrix-core instance took for this
        // point (its bit in Geometry::rs_skip_mask) are skipped
        bool skip = false;
        if (DIFF == OHX_DIFF_RS && g.rs_skip_mask && tid < PCH && p < p_end)
            skip = rs_mask_bit(g.rs_skip_mask + p * g.rs_mask_words, rs_tile_index(g, x0, y0));
        if (tid < PCH && p < p_end && !skip) {
            b = s.box[p];
            if (b.x < cx1 && b.y > x0 && b.z < ry1 && b.w > y0) {
                const bool full = b.x <= x0 && b.y >= cx1 && b.z <= y0 && b.w >= ry1;
                px = s.px[p];
                py = s.py[p];
                z = s.pz[p];
                cyc0 = s.cyc0[p];
                kamp = s.kamp[p];
                q = s.q[p];
                const double dxc = xc - px, dyc = yc - py;
                if (DIFF == OHX_DIFF_FRESNEL) {
                    // FP32 phase terms must stay below MAX_CYC cycles on the tile
                    const double span = (2.0 * fabs(dxc) * umax + (double)umax * umax +
                                         2.0 * fabs(dyc) * vmax + (double)vmax * vmax) * q;
                    cls = span > MAX_CYC ? CLS_EXACT : (full ? CLS_FAST : CLS_MASKED);
                    double ph = cyc0 + (dxc * dxc + dyc * dyc) * q;
                    ph -= floor(ph);
                    const float qf = (float)q;
                    c0 = make_float4((float)ph, (float)(2.0 * dxc * q), qf, (float)kamp);
                    c1 = make_float4((float)(2.0 * dyc * q), 0.f, 0.f, 0.f);
                } else {
                    const double zz = z * z;
                    const double rho2 = dxc * dxc + dyc * dyc;
                    const double rc2 = rho2 + zz;
                    const double rc = sqrt(rc2);
                    // bound on |N| = |r^2 - r_c^2| over the tile
                    const double nmax = 2.0 * umax * fabs(dxc) + (double)umax * umax +
                                        2.0 * vmax * fabs(dyc) + (double)vmax * vmax;
                    const double e = nmax / rc2;
                    // third-order phase term e^3 r_c/(16 lambda) and second-order
                    // amplitude term e^2 must be negligible
                    const bool series_ok = e * e * e * rc * g.inv_lambda < 16 * 2e-7 && e * e < 1e-7 &&
                                           nmax * 0.5 * g.inv_lambda / rc <= MAX_CYC;
                    if (!series_ok) {
                        cls = CLS_EXACT;
                    } else {
                        bool range_ok = false;
                        if (full) {
                            // reference range test holds on the whole tile with margin
                            const double xl = pix_x(g, x0), xh = pix_x(g, cx1 - 1);
                            const double yh = pix_y(g, y0), yl = pix_y(g, ry1 - 1);
                            const double dxmax = fmax(fabs(xl - px), fabs(xh - px));
                            const double dymax = fmax(fabs(yl - py), fabs(yh - py));
                            const double dxmin = (px >= xl && px <= xh) ? 0.0 : fmin(fabs(xl - px), fabs(xh - px));
                            const double dymin = (py >= yl && py <= yh) ? 0.0 : fmin(fabs(yl - py), fabs(yh - py));
                            const double tol = 1e-12;
                            range_ok = g.hx_coef * sqrt(dymin * dymin + zz) - dxmax > tol &&
                                       g.hy_coef * sqrt(dxmin * dxmin + zz) - dymax > tol;
                        }
                        double ph = cyc0 + (rho2 / (rc + z)) * g.inv_lambda;
                        ph -= floor(ph);
                        const double hp = 0.5 * g.inv_lambda / rc;          // 1/(2 r_c lambda)
                        const double c2 = hp / (4.0 * rc2);                  // 1/(8 r_c^3 lambda)
                        const double az = kamp / rc2;                        // a z/(lambda r_c^2)
                        const double m = az / rc2;
                        // cross term -2 c2 N_x N_y of the phase (rad) over the tile
                        const double nxm = 2.0 * umax * fabs(dxc) + (double)umax * umax;
                        const double nym = 2.0 * vmax * fabs(dyc) + (double)vmax * vmax;
                        const bool sep = 2.0 * TWO_PI * c2 * nxm * nym <= MAX_EPS;
                        cls = range_ok ? (sep ? CLS_FAST : CLS_FASTNS) : CLS_MASKED;
                        c0 = make_float4((float)ph, (float)(2.0 * dxc), (float)(-c2), (float)hp);
                        c1 = make_float4((float)(-m), (float)az, (float)(-2.0 * TWO_PI * c2), (float)(2.0 * dyc));
                        mu = (float)(-1.0 / rc2);
                    }
                }
            }
        }
        if (EX == EX_FLAG && cls == CLS_EXACT) {
            *exact_flag = 1ull;   // left to the follow-up kernel
            cls = CLS_OUT;
        }
        if (EX == EX_ONLY && cls != CLS_EXACT) cls = CLS_OUT;
        // ---- deterministic compaction (point order) ------------------------
        const unsigned long long bf = __ballot(cls == CLS_FAST);
        const unsigned long long bs = __ballot(cls >= CLS_MASKED);
        if (lane == 0) {
            s_cnt[0][wave] = __popcll(bf);
            s_cnt[1][wave] = __popcll(bs);
        }
        __syncthreads();
        int nF = 0, nS = 0, offF = 0, offS = 0;
#pragma unroll
        for (int w = 0; w < WAVES; ++w) {
            if (w < wave) {
                offF += s_cnt[0][w];
                offS += s_cnt[1][w];
            }
            nF += s_cnt[0][w];
            nS += s_cnt[1][w];
        }
        // the list lengths are workgroup-uniform: scalar loop control below
        nF = __builtin_amdgcn_readfirstlane(nF);
        nS = __builtin_amdgcn_readfirstlane(nS);
        touched |= nS > 0;
        if (cls != CLS_OUT) {
            const unsigned long long m = (cls == CLS_FAST) ? bf : bs;
            const int rank = __builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32),
                                                       __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
            const int slot = (cls == CLS_FAST) ? offF + rank : PCH - 1 - (offS + rank);
            s_box[slot] = b;
            if (cls != CLS_EXACT) {
                s_c0[slot] = c0;
                s_c1[slot] = c1;
                if (DIFF == OHX_DIFF_RS) sm.mu[slot] = mu;
            }
            if (cls >= CLS_MASKED) {
                s_cls[slot] = cls;
                if (cls == CLS_EXACT || DIFF == OHX_DIFF_RS) {
                    double* fd = slot_fp64(&s_R[0][slot][0]);
                    fd[0] = px;
                    fd[1] = py;
                    fd[2] = z;
                    fd[3] = cyc0;
                    fd[4] = kamp;
                    fd[5] = q;
                }
            }
        }
        __syncthreads();

        // ---- row table: each wave fills its own 8 rows for every listed point
        // Fresnel: B = amp (cos, sin)(2 pi v (g_y + q v)), zero outside the
        // box rows.  RS: N_y = v (2 dy_c + v) for every non-EXACT point, and
        // for the separable list B = (1 + mu N_y) (cos, sin)(2 pi N_y (h + c_z N_y)).
        for (int jj = lane; jj < nF + nS; jj += TW) {
            const int slot = jj < nF ? jj : PCH - 1 - (jj - nF);
            if (jj >= nF && s_cls[slot] == CLS_EXACT) continue;
            const float4 c = s_c0[slot], d = s_c1[slot];
            float2 o[RPT];
            if (DIFF == OHX_DIFF_FRESNEL) {
                const int4 bb = s_box[slot];
#pragma unroll
                for (int k = 0; k < RPT; ++k) {
                    const float t = vr[k] * fmaf(c.z, vr[k], d.x);
                    const int row = rowb + k;
                    const float a = (row >= bb.z && row < bb.w) ? c.w : 0.f;
                    o[k] = make_float2(a * cos_rev(t), a * sin_rev(t));
                }
            } else {
                float ny[RPT];
#pragma unroll
                for (int k = 0; k < RPT; ++k) ny[k] = vr[k] * (d.w + vr[k]);
                float4* dn = (float4*)&sm.Ny[wave][slot][0];
                dn[0] = make_float4(ny[0], ny[1], ny[2], ny[3]);
                dn[1] = make_float4(ny[4], ny[5], ny[6], ny[7]);
                if (jj >= nF) continue;
                const float m1 = sm.mu[slot];
#pragma unroll
                for (int k = 0; k < RPT; ++k) {
                    const float t = ny[k] * fmaf(c.z, ny[k], c.w);
                    const float a = fmaf(m1, ny[k], 1.f);
                    o[k] = make_float2(a * cos_rev(t), a * sin_rev(t));
                }
            }
            float4* dst = (float4*)&s_R[wave][slot][0];
#pragma unroll
            for (int i = 0; i < RPT / 2; ++i) dst[i] = make_float4(o[2 * i].x, o[2 * i].y, o[2 * i + 1].x, o[2 * i + 1].y);
        }
        __syncthreads();

        // ---- fast list: whole tile inside the mask, separable form ----------
        // Fresnel: sum amp sin/cos(f_col + t_row) = A(col) x B(row), 4 FMAs per
        // pair with one v_sin/v_cos per lane and point.  RS: the phase
        // c0 + c_w N + c_z N^2 and amplitude az (1 + mu N), N = n_x + N_y, give
        // A(col) = az (1 + mu n_x) e^{j 2pi (c0 + c_w n_x + c_z n_x^2)} and the
        // tabulated B(row); the cross term 2 c_z n_x N_y enters as
        // B (1 + j eps), eps = k_x N_y (|eps| <= MAX_EPS), and the amplitude's
        // mu^2 n_x N_y is below 1e-7: 6 FMAs per pair, no v_sin/v_cos.
        float pr[RPT], pim[RPT];
#pragma unroll
        for (int k = 0; k < RPT; ++k) pr[k] = pim[k] = 0.f;
        for (int j = 0; j < nF; ++j) {
            // RS: the point's own constants are read first and all reads are
            // issued before any arithmetic, so that the column factor (FP32
            // chain, v_sin, v_cos) waits only for its own constants while the
            // row factors are on their way (3 % faster; the lighter Fresnel
            // loop is 2 % faster left to the scheduler)
            const float4 c = s_c0[j];
            float4 d, n0, n1;   // RS only
            if constexpr (DIFF == OHX_DIFF_RS) {
                d = s_c1[j];
                __builtin_amdgcn_sched_barrier(0);
            }
            const float4* Rp = (const float4*)&s_R[wave][j][0];
            const float4 r0 = Rp[0], r1 = Rp[1], r2 = Rp[2], r3 = Rp[3];
            if constexpr (DIFF == OHX_DIFF_RS) {
                n0 = *(const float4*)&sm.Ny[wave][j][0];
                n1 = *(const float4*)&sm.Ny[wave][j][4];
                __builtin_amdgcn_sched_barrier(0);
            }
            const float Rx[RPT] = {r0.x, r0.z, r1.x, r1.z, r2.x, r2.z, r3.x, r3.z};
            const float Ry[RPT] = {r0.y, r0.w, r1.y, r1.w, r2.y, r2.w, r3.y, r3.w};
            if (DIFF == OHX_DIFF_FRESNEL) {
                const float f = fmaf(fmaf(c.z, u, c.y), u, c.x);
                const float sa = sin_rev(f), ca = cos_rev(f);
#pragma unroll
                for (int k = 0; k < RPT; ++k) {
                    pr[k] = fmaf(sa, Rx[k], fmaf(ca, Ry[k], pr[k]));
                    pim[k] = fmaf(ca, Rx[k], fmaf(-sa, Ry[k], pim[k]));
                }
            } else {
                const float Nv[RPT] = {n0.x, n0.y, n0.z, n0.w, n1.x, n1.y, n1.z, n1.w};
                const float nx = fmaf(c.y, u, uu);
                const float fa = fmaf(fmaf(c.z, nx, c.w), nx, c.x);
                const float am = fmaf(d.x, nx, d.y);
                const float kx = d.z * nx;
                const float sa = am * sin_rev(fa), ca = am * cos_rev(fa);
                // (sa, ca) against B (1 + j eps), eps = k_x N_y, regrouped per row:
                // P = sa + ca eps, Q = ca - sa eps, then pr += P Bx + Q By and
                // pim += Q Bx - P By (6 FMAs per pair)
                const float cak = ca * kx, sak = sa * kx;
#pragma unroll
                for (int k = 0; k < RPT; ++k) {
                    const float P = fmaf(cak, Nv[k], sa), Q = fmaf(-sak, Nv[k], ca);
                    pr[k] = fmaf(P, Rx[k], fmaf(Q, Ry[k], pr[k]));
                    pim[k] = fmaf(Q, Rx[k], fmaf(-P, Ry[k], pim[k]));
                }
            }
        }
        // ---- slow list: per-pixel mask and/or FP64 terms --------------------
        for (int j = 0; j < nS; ++j) {
            const int slot = PCH - 1 - j;
            const int cl = s_cls[slot];
            const int4 bb = s_box[slot];
            const double* fd = slot_fp64(&s_R[0][slot][0]);
            const bool colin = col >= bb.x && col < bb.y;
            if (cl == CLS_MASKED && DIFF == OHX_DIFF_FRESNEL) {
                // separable mask: B is zero outside the box rows, A outside its columns
                const float4 c = s_c0[slot];
                const float2* Rs = &s_R[wave][slot][0];
                const float f = fmaf(fmaf(c.z, u, c.y), u, c.x);
                const float sa = colin ? sin_rev(f) : 0.f, ca = colin ? cos_rev(f) : 0.f;
#pragma unroll
                for (int k = 0; k < RPT; ++k) {
                    const float2 bk = Rs[k];
                    pr[k] = fmaf(sa, bk.x, fmaf(ca, bk.y, pr[k]));
                    pim[k] = fmaf(ca, bk.x, fmaf(-sa, bk.y, pim[k]));
                }
            } else if (DIFF == OHX_DIFF_RS && cl == CLS_FASTNS) {
                // whole tile in the mask, phase per pixel
                const float4 c = s_c0[slot];
                const float4 d = s_c1[slot];
                const float* Ns = &sm.Ny[wave][slot][0];
                const float nx = fmaf(c.y, u, uu);
#pragma unroll
                for (int k = 0; k < RPT; ++k) {
                    const float N = nx + Ns[k];
                    const float f = fmaf(fmaf(c.z, N, c.w), N, c.x);
                    const float amp = fmaf(d.x, N, d.y);
                    pr[k] = fmaf(amp, sin_rev(f), pr[k]);
                    pim[k] = fmaf(amp, cos_rev(f), pim[k]);
                }
            } else if (cl == CLS_MASKED) {
                const float4 c = s_c0[slot];
                const float* Ns = &sm.Ny[wave][slot][0];
                const float4 d = s_c1[slot];
                const double qpx = fd[0], qpy = fd[1], qz = fd[2];
                const double zz = qz * qz;
#pragma unroll
                for (int k = 0; k < RPT; ++k) {
                    const int row = rowb + k;
                    bool in = colin && row >= bb.z && row < bb.w;
                    in = in && rs_range_ok(qpx, qpy, zz, g.hx_coef, g.hy_coef, xxx, pix_y(g, row));
                    const float N = fmaf(c.y, u, uu) + Ns[k];
                    const float f = fmaf(fmaf(c.z, N, c.w), N, c.x);
                    float amp = fmaf(d.x, N, d.y);
                    amp = in ? amp : 0.f;
                    pr[k] = fmaf(amp, sin_rev(f), pr[k]);
                    pim[k] = fmaf(amp, cos_rev(f), pim[k]);
                    __builtin_amdgcn_sched_barrier(0);  // keep one row live at a time
                }
            } else if (EX != EX_FLAG) {
                // FP64 terms, one row at a time (a rolled loop keeps the kernel's
                // register count at the fast path's; the row's term is added to
                // its accumulator by selection)
                const double qpx = fd[0], qpy = fd[1], qz = fd[2];
                const double qc = fd[3], qk = fd[4], qq = fd[5];
                const double zz = qz * qz;
#pragma unroll 1
                for (int k = 0; k < RPT; ++k) {
                    const int row = rowb + k;
                    const double yyy = pix_y(g, row);
                    bool in = colin && row >= bb.z && row < bb.w;
                    if (DIFF == OHX_DIFF_RS)
                        in = in && rs_range_ok(qpx, qpy, zz, g.hx_coef, g.hy_coef, xxx, yyy);
                    double re = 0.0, im = 0.0;
                    if (in) exact_term<DIFF>(xxx - qpx, yyy - qpy, qz, qc, qk, qq, g.inv_lambda, re, im);
#pragma unroll
                    for (int kk = 0; kk < RPT; ++kk) {
                        accr[kk] += (kk == k) ? re : 0.0;
                        acci[kk] += (kk == k) ? im : 0.0;
                    }
                }
            }
        }
#pragma unroll
        for (int k = 0; k < RPT; ++k) {
            accr[k] += (double)pr[k];
            acci[k] -= (double)pim[k];
        }
        __syncthreads();
    }

    if (EX == EX_ONLY && !touched) return;   // uniform: nothing to add
    store_tile(g, field, part, col, rowb, accr, acci);
}

// ---------------------------------------------------------------------------
// FP64 path (MODE_DOUBLE): every term in FP64, with the per-pixel sin/cos
// taken out of the loop by separating the phase over the tile.
//  * Fresnel: c = cyc0 + q dx^2 + q dy^2 exactly, so the term
//      (1/j) a e^{j 2pi c} = A(col) B(row),
//      A = a (sin, -cos)(2pi (cyc0 + q dx^2)),  B = (cos, sin)(2pi q dy^2),
//    and the mask box is separable too (A = 0 outside its columns, B = 0
//    outside its rows): a pair is one complex multiply-add (4 FMAs).
//  * RS: with N = r^2 - r_c^2 = N_x(col) + N_y(row) about the tile centre
//    and e = N / r_c^2,
//      r/lambda = r_c/lambda + h N + phi(N),  h = 1/(2 r_c lambda),
//      phi = (r_c/lambda)(-e^2/8 + e^3/16 - 5 e^4/128),
//      z/r^2 = (z/r_c^2)(1 - e + e^2 - e^3),
//    so the term is A(col) B(row) T(N), T = (z/r^2)/(z/r_c^2) e^{j 2pi phi},
//    with |2pi phi| <= 4e-3 rad evaluated by cos/sin polynomials (remainders
//    below 1e-14).  A (point, tile) pair takes this form only when the whole
//    tile is inside the mask with margin and the truncated series terms are
//    below 1e-13 cycles and 1e-14 relative amplitude; every other pair goes
//    to the per-pixel list (exact_term under the reference's mask).
constexpr int PCD = 32;   // points per chunk on the FP64 path
constexpr int NCS = 12;   // constants per point of the separable list
// separable-list constant slots
enum { F_PX = 0, F_PY, F_CYC, F_Q, F_AMP };
enum { R_C0 = 0, R_H, R_G2X, R_AMP, R_K2, R_K3, R_K4, R_M1, R_M2, R_M3, R_G2Y };

constexpr int NTAB = 256;  // angles of the FP64 sin/cos table
// FP64 tiles are 128 x 16: the two waves of a wave row share its 8 rows, so
// the row factors are computed once per row for both (the FP32 path keeps
// 64 x 32: wider tiles double its FP32 phase magnitudes at the edges).
constexpr int WX64 = 2;
constexpr int TW64 = WX64 * TW;
constexpr int TH64 = (WAVES / WX64) * RPT;

// sin and cos of 2 pi t, t in [0, 1], in FP64: the nearest table angle
// 2 pi i / NTAB plus Taylor series of the remainder |x| <= pi / NTAB
// (dropped terms below 2e-20), combined by angle addition; about 17 FP64
// instructions and one LDS read against ~40 for sincospi.
__device__ inline void sincos2pi(double t, const double2* __restrict__ tab, double& s, double& c)
{
    const double ti = rint(t * NTAB);
    const double x = fma(ti, -1.0 / NTAB, t) * TWO_PI;
    const double2 a = tab[(int)ti & (NTAB - 1)];   // (cos, sin)(2 pi i / NTAB)
    const double x2 = x * x;
    const double cb = fma(x2, fma(x2, fma(x2, -1.0 / 720.0, 1.0 / 24.0), -0.5), 1.0);
    const double sb = x * fma(x2, fma(x2, fma(x2, -1.0 / 5040.0, 1.0 / 120.0), -1.0 / 6.0), 1.0);
    s = fma(a.y, cb, a.x * sb);
    c = fma(a.x, cb, -(a.y * sb));
}

struct DirectSmem64 {
    double2 tab[NTAB];     // (cos, sin)(2 pi i / NTAB)
    double2 B[TH64][PCD];  // row factors B(row) of the separable list
    double Ny[TH64][PCD];  // RS: N_y(row)
    double c[PCD][NCS];    // separable list: per-point constants
    int4 sbox[PCD];        // separable list: mask box (Fresnel)
    double d[6][PCD];      // per-pixel list: px, py, z, cyc0, kamp, q
    int4 ebox[PCD];        // per-pixel list: mask box
    int cnt[2];
};

template <int DIFF>
__device__ __attribute__((always_inline)) inline void direct_tile64(DirectSmem64& sm, const Geometry& g,
                                                                    const PointSoA& s, int64_t p_begin, int64_t p_end,
                                                                    double2* __restrict__ field, int tile, int tiles_x,
                                                                    double2* __restrict__ part = nullptr)
{
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int cw = wave % WX64, rw = wave / WX64;   // wave (cw, rw): 64 columns x 8 rows
    const int tx = tile % tiles_x, ty = tile / tiles_x;
    const int x0 = tx * TW64, y0 = ty * TH64;
    const int cx1 = min(x0 + TW64, g.pn_x), ry1 = min(y0 + TH64, g.pn_y);
    const int col = x0 + cw * TW + lane;
    const int rowb = y0 + rw * RPT;
    const double xc = pix_x(g, x0 + TW64 / 2), yc = pix_y(g, y0 + TH64 / 2);
    const double xxx = pix_x(g, col);
    const double u = xxx - xc;
    // |u| <= 64 pitches and |v| <= 8 pitches over the tile
    const double umax = (double)(TW64 / 2) * g.pp_x, vmax = (double)(TH64 / 2) * g.pp_y;

    // sin/cos table (first read after the classification barrier below)
    for (int i = tid; i < NTAB; i += NT) {
        double sn, cs;
        sincospi(2.0 * i / NTAB, &sn, &cs);
        sm.tab[i] = make_double2(cs, sn);
    }

    double accr[RPT], acci[RPT];
#pragma unroll
    for (int k = 0; k < RPT; ++k) accr[k] = acci[k] = 0.0;

    for (int64_t base = p_begin; base < p_end; base += PCD) {
        // ---- classify one point per lane of wave 0 --------------------------
        if (wave == 0) {
            bool sep = false, ex = false;
            int4 b = make_int4(0, 0, 0, 0);
            double px = 0, py = 0, z = 0, cyc0 = 0, kamp = 0, q = 0;
            double cc[NCS];
#pragma unroll
            for (int i = 0; i < NCS; ++i) cc[i] = 0.0;
            const int64_t p = base + lane;
            if (lane < PCD && p < p_end) {
                b = s.box[p];
                if (b.x < cx1 && b.y > x0 && b.z < ry1 && b.w > y0) {
                    px = s.px[p];
                    py = s.py[p];
                    z = s.pz[p];
                    cyc0 = s.cyc0[p];
                    kamp = s.kamp[p];
                    q = s.q[p];
                    if (DIFF == OHX_DIFF_FRESNEL) {
                        sep = true;
                        cc[F_PX] = px;
                        cc[F_PY] = py;
                        cc[F_CYC] = cyc0;
                        cc[F_Q] = q;
                        cc[F_AMP] = kamp;
                    } else {
                        const bool full = b.x <= x0 && b.y >= cx1 && b.z <= y0 && b.w >= ry1;
                        const double dxc = xc - px, dyc = yc - py;
                        const double zz = z * z;
                        const double rho2 = dxc * dxc + dyc * dyc;
                        const double rc2 = rho2 + zz;
                        const double rc = sqrt(rc2);
                        const double nmax = 2.0 * umax * fabs(dxc) + umax * umax + 2.0 * vmax * fabs(dyc) + vmax * vmax;
                        const double e = nmax / rc2;
                        const double e2 = e * e;
                        const double rl = rc * g.inv_lambda;   // r_c / lambda
                        const double phimax = TWO_PI * rl * e2 * (0.125 + e * (0.0625 + e * 0.0390625));
                        bool ok = full && rl * e2 * e2 * e <= 1e-12 && e2 * e2 <= 1e-14 && phimax <= 4e-3;
                        if (ok) {
                            // reference range test holds on the whole tile with margin
                            const double xl = pix_x(g, x0), xh = pix_x(g, cx1 - 1);
                            const double yh = pix_y(g, y0), yl = pix_y(g, ry1 - 1);
                            const double dxmax = fmax(fabs(xl - px), fabs(xh - px));
                            const double dymax = fmax(fabs(yl - py), fabs(yh - py));
                            const double dxmin = (px >= xl && px <= xh) ? 0.0 : fmin(fabs(xl - px), fabs(xh - px));
                            const double dymin = (py >= yl && py <= yh) ? 0.0 : fmin(fabs(yl - py), fabs(yh - py));
                            const double tol = 1e-12;
                            ok = g.hx_coef * sqrt(dymin * dymin + zz) - dxmax > tol &&
                                 g.hy_coef * sqrt(dxmin * dxmin + zz) - dymax > tol;
                        }
                        sep = ok;
                        ex = !ok;
                        if (ok) {
                            double c0 = cyc0 + (rho2 / (rc + z)) * g.inv_lambda;
                            c0 -= floor(c0);
                            const double h = 0.5 * g.inv_lambda / rc;
                            const double ir = 1.0 / rc2;
                            cc[R_C0] = c0;
                            cc[R_H] = h;
                            cc[R_G2X] = 2.0 * dxc;
                            cc[R_AMP] = kamp * ir;
                            cc[R_K2] = -TWO_PI * 0.25 * h * ir;
                            cc[R_K3] = TWO_PI * 0.125 * h * ir * ir;
                            cc[R_K4] = -TWO_PI * 0.078125 * h * ir * ir * ir;
                            cc[R_M1] = -ir;
                            cc[R_M2] = ir * ir;
                            cc[R_M3] = -ir * ir * ir;
                            cc[R_G2Y] = 2.0 * dyc;
                        }
                    }
                }
            }
            // deterministic compaction in point order
            const unsigned long long bsep = __ballot(sep), bex = __ballot(ex);
            const unsigned long long m = sep ? bsep : bex;
            const int rank = __builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
            if (sep) {
#pragma unroll
                for (int i = 0; i < NCS; ++i) sm.c[rank][i] = cc[i];
                sm.sbox[rank] = b;
            }
            if (ex) {
                sm.d[0][rank] = px;
                sm.d[1][rank] = py;
                sm.d[2][rank] = z;
                sm.d[3][rank] = cyc0;
                sm.d[4][rank] = kamp;
                sm.d[5][rank] = q;
                sm.ebox[rank] = b;
            }
            if (lane == 0) {
                sm.cnt[0] = __popcll(bsep);
                sm.cnt[1] = __popcll(bex);
            }
        }
        __syncthreads();
        const int nS = __builtin_amdgcn_readfirstlane(sm.cnt[0]);
        const int nE = __builtin_amdgcn_readfirstlane(sm.cnt[1]);

        // ---- row factors: the two waves of a wave row split its 8 rows; lanes
        // j and j + 32 take half of a wave's share for point j
        {
            const int j = lane & (PCD - 1);
            const int r0 = rw * RPT + cw * (RPT / WX64) + (lane >> 5) * (RPT / WX64 / 2);
            if (j < nS) {
#pragma unroll
                for (int k = 0; k < RPT / WX64 / 2; ++k) {
                    const int rr = r0 + k;
                    const int row = y0 + rr;
                    const double yyy = pix_y(g, row);
                    double ph, ny = 0.0;
                    bool rin = true;
                    if (DIFF == OHX_DIFF_FRESNEL) {
                        const double dy = yyy - sm.c[j][F_PY];
                        ph = sm.c[j][F_Q] * dy * dy;
                        const int4 bb = sm.sbox[j];
                        rin = row >= bb.z && row < bb.w;
                    } else {
                        const double v = yyy - yc;
                        ny = v * (sm.c[j][R_G2Y] + v);
                        ph = sm.c[j][R_H] * ny;
                    }
                    ph -= floor(ph);
                    double sn, cs;
                    sincos2pi(ph, sm.tab, sn, cs);
                    sm.B[rr][j] = rin ? make_double2(cs, sn) : make_double2(0.0, 0.0);
                    if (DIFF == OHX_DIFF_RS) sm.Ny[rr][j] = ny;
                }
            }
        }
        __syncthreads();

        // ---- separable list --------------------------------------------------
        for (int j = 0; j < nS; ++j) {
            double Ar, Ai;
            if (DIFF == OHX_DIFF_FRESNEL) {
                const double px = sm.c[j][F_PX], cyc0 = sm.c[j][F_CYC], q = sm.c[j][F_Q];
                const int4 bb = sm.sbox[j];
                const double dx = xxx - px;
                double ph = fma(q * dx, dx, cyc0);
                ph -= floor(ph);
                double sn, cs;
                sincos2pi(ph, sm.tab, sn, cs);
                const double a = (col >= bb.x && col < bb.y) ? sm.c[j][F_AMP] : 0.0;
                Ar = a * sn;
                Ai = -a * cs;
#pragma unroll
                for (int k = 0; k < RPT; ++k) {
                    const double2 bk = sm.B[rw * RPT + k][j];
                    accr[k] = fma(Ar, bk.x, fma(-Ai, bk.y, accr[k]));
                    acci[k] = fma(Ar, bk.y, fma(Ai, bk.x, acci[k]));
                }
            } else {
                const double nx = u * (sm.c[j][R_G2X] + u);
                double ph = fma(sm.c[j][R_H], nx, sm.c[j][R_C0]);
                ph -= floor(ph);
                double sn, cs;
                sincos2pi(ph, sm.tab, sn, cs);
                const double a0 = sm.c[j][R_AMP];
                Ar = a0 * sn;
                Ai = -a0 * cs;
                const double k2 = sm.c[j][R_K2], k3 = sm.c[j][R_K3], k4 = sm.c[j][R_K4];
                const double m1 = sm.c[j][R_M1], m2 = sm.c[j][R_M2], m3 = sm.c[j][R_M3];
#pragma unroll
                for (int k = 0; k < RPT; ++k) {
                    const double2 bk = sm.B[rw * RPT + k][j];
                    const double N = nx + sm.Ny[rw * RPT + k][j];
                    const double t = fma(fma(k4, N, k3), N, k2);
                    const double ph2 = (N * N) * t;                 // 2 pi phi(N), radians
                    const double p2 = ph2 * ph2;
                    const double cr = fma(p2, fma(p2, 1.0 / 24.0, -0.5), 1.0);
                    const double ci = ph2 * fma(p2, -1.0 / 6.0, 1.0);
                    const double am = fma(fma(fma(m3, N, m2), N, m1), N, 1.0);
                    const double tr = am * cr, ti = am * ci;
                    const double sr = fma(Ar, bk.x, -Ai * bk.y), si = fma(Ar, bk.y, Ai * bk.x);
                    accr[k] = fma(sr, tr, fma(-si, ti, accr[k]));
                    acci[k] = fma(sr, ti, fma(si, tr, acci[k]));
                }
            }
        }

        // ---- per-pixel list: exact FP64 terms under the reference's mask ---
        if (DIFF == OHX_DIFF_RS) {
            for (int j = 0; j < nE; ++j) {
                const int4 bb = sm.ebox[j];
                const double qpx = sm.d[0][j], qpy = sm.d[1][j], qz = sm.d[2][j];
                const double qc = sm.d[3][j], qk = sm.d[4][j], qq = sm.d[5][j];
                const double zz = qz * qz;
                const bool colin = col >= bb.x && col < bb.y;
#pragma unroll
                for (int k = 0; k < RPT; ++k) {
                    const int row = rowb + k;
                    const double yyy = pix_y(g, row);
                    bool in = colin && row >= bb.z && row < bb.w;
                    in = in && rs_range_ok(qpx, qpy, zz, g.hx_coef, g.hy_coef, xxx, yyy);
                    if (in) exact_term<DIFF>(xxx - qpx, yyy - qpy, qz, qc, qk, qq, g.inv_lambda, accr[k], acci[k]);
                    __builtin_amdgcn_sched_barrier(0);
                }
            }
        }
        __syncthreads();
    }

    store_tile(g, field, part, col, rowb, accr, acci);
}

// LDS of the multi-job kernel, which runs both paths
union TileSmem {
    DirectSmem<OHX_DIFF_RS> rs;
    DirectSmem<OHX_DIFF_FRESNEL> fr;
    DirectSmem64 f64;
};

// FP32 kernel: three waves per SIMD (registers and LDS both fit); pairs that
// need FP64 terms raise *exact_flag and are left to k_pc_direct_exact.
// Workgroup = (tile, point split): blockIdx = split * tiles + tile; split 0
// adds into the field, the others store into their partial planes.
struct SplitArgs {
    int tiles;           // tiles per split
    int64_t split_pts;   // points per split (a multiple of the LDS chunk)
    double2* part;       // (splits - 1) partial planes of pn_x * pn_y pixels
};

__device__ inline void split_item(const Geometry& g, const PointSoA& s, const SplitArgs& sa, int& tile,
                                  int64_t& p0, int64_t& p1, double2*& part)
{
    const int split = blockIdx.x / sa.tiles;
    tile = blockIdx.x - split * sa.tiles;
    p0 = (int64_t)split * sa.split_pts;
    p1 = min(s.n, p0 + sa.split_pts);
    part = split == 0 ? nullptr : sa.part + (int64_t)(split - 1) * ((int64_t)g.pn_x * g.pn_y);
}

template <int DIFF>
__global__ __launch_bounds__(NT, 3) void k_pc_direct(Geometry g, PointSoA s, double2* __restrict__ field,
                                                      int tiles_x, unsigned long long* __restrict__ exact_flag,
                                                      SplitArgs sa)
{
    __shared__ DirectSmem<DIFF> sm;
    int tile;
    int64_t p0, p1;
    double2* part;
    split_item(g, s, sa, tile, p0, p1, part);
    direct_tile<DIFF, EX_FLAG>(sm, g, s, p0, p1, field, tile, tiles_x, exact_flag, part);
}

// field += partial planes 1..splits-1, in split order (deterministic)
__global__ __launch_bounds__(256) void k_split_reduce(double2* __restrict__ field, const double2* __restrict__ part,
                                                      int64_t npx, int nparts)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < npx; i += stride) {
        double2 f = field[i];
        for (int k = 0; k < nparts; ++k) {
            const double2 q = part[(int64_t)k * npx + i];
            f.x += q.x;
            f.y += q.y;
        }
        field[i] = f;
    }
}

// The FP64 pairs the FP32 kernel left (same classification, same tiles);
// returns at once when it left none.
template <int DIFF>
__global__ __launch_bounds__(NT) void k_pc_direct_exact(Geometry g, PointSoA s, double2* __restrict__ field,
                                                         int tiles_x, const unsigned long long* __restrict__ exact_flag)
{
    if (*exact_flag == 0ull) return;
    __shared__ DirectSmem<DIFF> sm;
    direct_tile<DIFF, EX_ONLY>(sm, g, s, 0, s.n, field, blockIdx.x, tiles_x, nullptr);
}

template <int DIFF>
__global__ __launch_bounds__(NT, 2) void k_pc_direct64(Geometry g, PointSoA s, double2* __restrict__ field,
                                                        int tiles_x, SplitArgs sa)
{
    __shared__ DirectSmem64 sm;
    int tile;
    int64_t p0, p1;
    double2* part;
    split_item(g, s, sa, tile, p0, p1, part);
    direct_tile64<DIFF>(sm, g, s, p0, p1, field, tile, tiles_x, part);
}

// Persistent multi-job form used by the serving engine: workgroups stride
// over the (job, tile) items of every job in the table; each item adds the
// job's current point slice [p0, p1) into that job's resident field.
__global__ __launch_bounds__(NT) void k_pc_direct_jobs(const DirectJob* __restrict__ jobs,
                                                       const int* __restrict__ njobs_ptr)
{
    __shared__ TileSmem sm;
    const int njobs = *njobs_ptr;
    const int total = njobs > 0 ? jobs[njobs - 1].item_base + jobs[njobs - 1].tiles : 0;
    for (int item = blockIdx.x; item < total; item += gridDim.x) {
        int j = 0;
        while (j + 1 < njobs && jobs[j + 1].item_base <= item) ++j;
        const DirectJob& jb = jobs[j];
        const int tile = item - jb.item_base;
        if (jb.g.diff == OHX_DIFF_RS) {
            if (jb.fp64) direct_tile64<OHX_DIFF_RS>(sm.f64, jb.g, jb.s, jb.p0, jb.p1, jb.field, tile, jb.tiles_x);
            else direct_tile<OHX_DIFF_RS, EX_INLINE>(sm.rs, jb.g, jb.s, jb.p0, jb.p1, jb.field, tile, jb.tiles_x, nullptr);
        } else {
            if (jb.fp64) direct_tile64<OHX_DIFF_FRESNEL>(sm.f64, jb.g, jb.s, jb.p0, jb.p1, jb.field, tile, jb.tiles_x);
            else direct_tile<OHX_DIFF_FRESNEL, EX_INLINE>(sm.fr, jb.g, jb.s, jb.p0, jb.p1, jb.field, tile, jb.tiles_x, nullptr);
        }
        __syncthreads();   // LDS reuse across items
    }
}

// ---------------------------------------------------------------------------
// FP64 accuracy probe: one thread per listed pixel, Neumaier summation
__device__ inline void neumaier(double& s, double& c, double x)
{
    const double t = s + x;
    if (fabs(s) >= fabs(x)) c += (s - t) + x;
    else c += (x - t) + s;
    s = t;
}

template <int DIFF>
__global__ __launch_bounds__(256) void k_pc_probe(Geometry g, PointSoA s, const int64_t* __restrict__ pix,
                                                  int64_t npix, double2* __restrict__ out)
{
    __shared__ double sd[6][256];
    __shared__ int4 sb[256];
    const int tid = threadIdx.x;
    const int64_t qi = (int64_t)blockIdx.x * 256 + tid;
    const int64_t npx = (int64_t)g.pn_x * g.pn_y;
    int64_t id = (qi < npix) ? pix[qi] : -1;
    const bool valid = id >= 0 && id < npx;
    const int col = valid ? (int)(id % g.pn_x) : 0;
    const int row = valid ? (int)(id / g.pn_x) : 0;
    const double xxx = pix_x(g, col), yyy = pix_y(g, row);
    double sr = 0, cr = 0, si = 0, ci = 0;
    for (int64_t base = 0; base < s.n; base += 256) {
        const int64_t p = base + tid;
        if (p < s.n) {
            sd[0][tid] = s.px[p];
            sd[1][tid] = s.py[p];
            sd[2][tid] = s.pz[p];
            sd[3][tid] = s.cyc0[p];
            sd[4][tid] = s.kamp[p];
            sd[5][tid] = s.q[p];
            sb[tid] = s.box[p];
        }
        __syncthreads();
        const int cnt = (int)min((int64_t)256, s.n - base);
        for (int j = 0; j < cnt; ++j) {
            const int4 b = sb[j];
            bool in = valid && col >= b.x && col < b.y && row >= b.z && row < b.w;
            if (DIFF == OHX_DIFF_RS && in)
                in = rs_range_ok(sd[0][j], sd[1][j], sd[2][j] * sd[2][j], g.hx_coef, g.hy_coef, xxx, yyy);
            if (in) {
                double re = 0, im = 0;
                exact_term<DIFF>(xxx - sd[0][j], yyy - sd[1][j], sd[2][j], sd[3][j], sd[4][j], sd[5][j],
                                 g.inv_lambda, re, im);
                neumaier(sr, cr, re);
                neumaier(si, ci, im);
            }
        }
        __syncthreads();
    }
    if (qi < npix) {
        const double nan = __builtin_nan("");
        out[qi] = valid ? make_double2(sr + cr, si + ci) : make_double2(nan, nan);
    }
}

// ---------------------------------------------------------------------------
int launch_prep_on(hipStream_t stream, const void* d_vertices, const Geometry& g, const PointSoA& soa,
                   unsigned long long* boxsum)
{
    const unsigned grid = (unsigned)((soa.n + 255) / 256);
    if (boxsum) hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, stream, boxsum);
    hipLaunchKernelGGL(k_pc_prep, dim3(grid), dim3(256), 0, stream, (const double*)d_vertices, g, soa, boxsum);
    return hipGetLastError() == hipSuccess ? OHX_OK : OHX_ERR_DEVICE;
}

int launch_prep(ohx_ctx* ctx, const void* d_vertices, const Geometry& g, const PointSoA& soa,
                unsigned long long* boxsum)
{
    if (launch_prep_on(ctx->stream, d_vertices, g, soa, boxsum) != OHX_OK)
        return check_hip(ctx, hipErrorLaunchFailure, "k_pc_prep");
    return OHX_OK;
}

// Workgroups the device holds at once for kernel k (occupancy x CUs).
static int resident_slots(ohx_ctx* ctx, const void* k)
{
    int per_cu = 0, cus = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k, NT, 0) != hipSuccess || per_cu < 1) per_cu = 1;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device) != hipSuccess || cus < 1)
        cus = 1;
    return per_cu * cus;
}

// Point splits per tile so that the last round of equal-cost workgroups is
// nearly full: time ~ ceil(tiles * S / slots) / S (3840x2160: 4080 tiles on
// 768 slots leave the sixth round 31 % full at S = 1; S = 3 fills 99 %).
// Each split keeps at least four LDS chunks, and the partial planes stay
// under 2 GiB.
static int choose_splits(int64_t tiles, int64_t n, int slots, int64_t npx, int chunk)
{
    int best = 1;
    double best_t = std::ceil((double)tiles / slots);
    for (int S = 2; S <= 4; ++S) {
        if (n < (int64_t)S * 4 * chunk) break;
        if ((double)(S - 1) * npx * 16.0 > 2.0 * (1ull << 30)) break;
        const double t = std::ceil((double)tiles * S / slots) / S;
        if (t < best_t * 0.97) {
            best = S;
            best_t = t;
        }
    }
    return best;
}

int launch_direct(ohx_ctx* ctx, const Geometry& g, const PointSoA& soa, double* d_field, bool fp64)
{
    return launch_direct_with(ctx, g, soa, d_field, fp64, ctx->minmax + 5, true);
}

int launch_direct_with(ohx_ctx* ctx, const Geometry& g, const PointSoA& soa, double* d_field, bool fp64,
                       unsigned long long* flag, bool allow_splits)
{
    const int tw = fp64 ? TW64 : TW, th = fp64 ? TH64 : TH;
    const int tiles_x = (g.pn_x + tw - 1) / tw, tiles_y = (g.pn_y + th - 1) / th;
    const int64_t tiles = (int64_t)tiles_x * tiles_y;
    const int64_t npx = (int64_t)g.pn_x * g.pn_y;
    if (tiles * 4 > 0x7fffffff) return set_error(ctx, OHX_ERR_INVALID_ARG, "hologram too large");
    double2* f = (double2*)d_field;
    const bool rs = g.diff == OHX_DIFF_RS;
    const void* kern = fp64 ? (rs ? (const void*)k_pc_direct64<OHX_DIFF_RS> : (const void*)k_pc_direct64<OHX_DIFF_FRESNEL>)
                            : (rs ? (const void*)k_pc_direct<OHX_DIFF_RS> : (const void*)k_pc_direct<OHX_DIFF_FRESNEL>);
    const int chunk = fp64 ? PCD : PCH;
    int splits = allow_splits ? choose_splits(tiles, soa.n, resident_slots(ctx, kern), npx, chunk) : 1;
    SplitArgs sa{(int)tiles, soa.n, nullptr};
    if (splits > 1) {
        // split boundaries on chunk boundaries: a split sums exactly as the
        // serving engine's point slices do
        const int64_t per = (soa.n + splits - 1) / splits;
        sa.split_pts = (per + chunk - 1) / chunk * chunk;
        splits = (int)((soa.n + sa.split_pts - 1) / sa.split_pts);
    }
    if (splits > 1) {
        int st = ensure_gemm_scratch(ctx, (size_t)(splits - 1) * npx * sizeof(double2));
        if (st) return st;
        sa.part = (double2*)ctx->gemm_scratch;
    }
    const dim3 grid((unsigned)(tiles * splits)), block(NT);
    const dim3 tgrid((unsigned)tiles);
    if (rs) {
        if (fp64) {
            hipLaunchKernelGGL((k_pc_direct64<OHX_DIFF_RS>), grid, block, 0, ctx->stream, g, soa, f, tiles_x, sa);
        } else {
            hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, ctx->stream, flag);
            hipLaunchKernelGGL((k_pc_direct<OHX_DIFF_RS>), grid, block, 0, ctx->stream, g, soa, f, tiles_x, flag, sa);
            hipLaunchKernelGGL((k_pc_direct_exact<OHX_DIFF_RS>), tgrid, block, 0, ctx->stream, g, soa, f, tiles_x,
                               (const unsigned long long*)flag);
        }
    } else {
        if (fp64) {
            hipLaunchKernelGGL((k_pc_direct64<OHX_DIFF_FRESNEL>), grid, block, 0, ctx->stream, g, soa, f, tiles_x, sa);
        } else {
            hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, ctx->stream, flag);
            hipLaunchKernelGGL((k_pc_direct<OHX_DIFF_FRESNEL>), grid, block, 0, ctx->stream, g, soa, f, tiles_x, flag,
                               sa);
            hipLaunchKernelGGL((k_pc_direct_exact<OHX_DIFF_FRESNEL>), tgrid, block, 0, ctx->stream, g, soa, f,
                               tiles_x, (const unsigned long long*)flag);
        }
    }
    if (splits > 1) {
        const int64_t rb = std::min<int64_t>((npx + 255) / 256, 4096);
        hipLaunchKernelGGL(k_split_reduce, dim3((unsigned)rb), dim3(256), 0, ctx->stream, f,
                           (const double2*)sa.part, npx, splits - 1);
    }
    ctx->last.workgroups = tiles * splits;
    ctx->last.kernel_kind = fp64 ? 1 : 0;
    return check_hip(ctx, hipGetLastError(), "k_pc_direct");
}

int direct_tiles(const Geometry& g, bool fp64, int* tiles_x)
{
    const int tw = fp64 ? TW64 : TW, th = fp64 ? TH64 : TH;
    *tiles_x = (g.pn_x + tw - 1) / tw;
    return *tiles_x * ((g.pn_y + th - 1) / th);
}

int launch_direct_jobs(hipStream_t stream, const DirectJob* d_jobs, const int* d_njobs, int grid)
{
    hipLaunchKernelGGL(k_pc_direct_jobs, dim3(grid), dim3(NT), 0, stream, d_jobs, d_njobs);
    return hipGetLastError() == hipSuccess ? OHX_OK : OHX_ERR_DEVICE;
}

int launch_probe(ohx_ctx* ctx, const Geometry& g, const PointSoA& soa, const int64_t* d_pix,
                 int64_t n_pix, double* d_out)
{
    const unsigned grid = (unsigned)((n_pix + 255) / 256);
    if (g.diff == OHX_DIFF_RS)
        hipLaunchKernelGGL((k_pc_probe<OHX_DIFF_RS>), dim3(grid), dim3(256), 0, ctx->stream, g, soa, d_pix, n_pix,
                           (double2*)d_out);
    else
        hipLaunchKernelGGL((k_pc_probe<OHX_DIFF_FRESNEL>), dim3(grid), dim3(256), 0, ctx->stream, g, soa, d_pix,
                           n_pix, (double2*)d_out);
    return check_hip(ctx, hipGetLastError(), "k_pc_probe");
}

}  // namespace ohx
