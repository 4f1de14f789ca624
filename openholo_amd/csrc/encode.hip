// Per-pixel encoders and 8-bit normalisation (HBM-bound).
//
// ophGen::encoding(ENCODE_FLAG) -> Phase / Amplitude / RealPart /
// ImaginaryPart (ophgen/src/ophGen.cpp:816-849, 1821-1864):
//   phase = (atan2(im, re) + pi) / (2 pi), amplitude = |H|, real, imag.
// ophGen::normalize (ophgen/src/ophGen.cpp:654-687) with minOfArr / maxOfArr
// (Openholo/src/function.h:79-99): the minimum starts at DBL_MAX and the
// maximum at DBL_MIN (the smallest *positive* normal, a reference quirk kept
// here), both taken over all channels; then
//   u8 = (uint8)(((v - min) / gap) * 255 + 0.5)   (truncating conversion).
// Min/max are reduced per block (wave shuffles) and merged with 64-bit
// atomics on an order-preserving integer image of the doubles; NaN is
// skipped as the reference's < / > comparisons skip it.
#include <algorithm>
#include <cmath>
#include <cfloat>
#include <cstdint>

#include "ohx_internal.h"

namespace ohx {

constexpr int EB = 256;
constexpr int MAXCH = 8;

__device__ inline unsigned long long order_key(double d)
{
    const unsigned long long b = __double_as_longlong(d);
    return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}
__device__ inline double key_value(unsigned long long k)
{
    const unsigned long long b = (k >> 63) ? (k & 0x7fffffffffffffffull) : ~k;
    return __longlong_as_double((long long)b);
}

template <int FLAG>
__device__ inline double encode_value(double2 h)
{
    const double pi = 3.14159265358979323846;
    const double pi2 = pi * 2;
    if (FLAG == OHX_ENCODE_PHASE) return (atan2(h.y, h.x) + pi) / pi2;
    if (FLAG == OHX_ENCODE_AMPLITUDE) return sqrt(h.x * h.x + h.y * h.y);
    if (FLAG == OHX_ENCODE_REAL) return h.x;
    return h.y;
}

template <int FLAG>
__global__ __launch_bounds__(EB) void k_encode(const double2* __restrict__ f, int64_t n, double* __restrict__ e)
{
    int64_t i = (int64_t)blockIdx.x * EB + threadIdx.x;
    const int64_t stride = (int64_t)gridDim.x * EB;
    for (; i < n; i += stride) e[i] = encode_value<FLAG>(f[i]);
}

// Single-side-band encoding (ophPointCloud::encodeHologram,
// ophgen/src/ophPointCloud.cpp:193-240).  The reference's FFT round trip
// T = s(IFFT(s(s(FFT(s(H)))))), s the centring shift out[k] = in[k + n/2]
// (Openholo.cpp:753-772), reduces per axis to T[k] = n H[k + 2h mod n] times
// exp(2 pi j (k + h) / n) for odd n (h = n/2; for even n, 2h = n and the
// factor is 1).  The carrier follows the reference's operation order
// (:225-233): y = (ss_y - pp_y) - pp_y h, Y = (pi y s_y) / pp_y,
// x = -ss_x/2 + pp_x w + pp_x/2, X = (pi x s_x) / pp_x,
// enc = Re(T) (cos X cos Y - sin X sin Y).
// X and the x factor depend on the column only, Y and the y factor on the
// row only: k_ssb_tables evaluates them once per column and row
// ({cos, sin, factor re, factor im}), k_encode_ssb combines them per pixel.
__global__ __launch_bounds__(EB) void k_ssb_tables(int pn_x, int pn_y, double pp_x, double pp_y, double sx,
                                                  double sy, double4* __restrict__ tx, double4* __restrict__ ty)
{
#pragma clang fp contract(off)
    const double pi = 3.14159265358979323846;
    const int k = blockIdx.x * EB + threadIdx.x;
    if (k < pn_x) {
        const double ss_x = pn_x * pp_x;
        const double x = -(ss_x / 2) + pp_x * k + pp_x / 2;
        const double X = (pi * x * sx) / pp_x;
        double fs = 0.0, fc = 1.0;
        if (pn_x & 1) sincospi(2.0 * (k + (pn_x >> 1)) / pn_x, &fs, &fc);
        tx[k] = make_double4(cos(X), sin(X), fc, fs);
    } else if (k < pn_x + pn_y) {
        const int h = k - pn_x;
        const double ss_y = pn_y * pp_y;
        const double y = (ss_y - pp_y) - (pp_y * h);
        const double Y = (pi * y * sy) / pp_y;
        double fs = 0.0, fc = 1.0;
        if (pn_y & 1) sincospi(2.0 * (h + (pn_y >> 1)) / pn_y, &fs, &fc);
        ty[h] = make_double4(cos(Y), sin(Y), fc, fs);
    }
}

__global__ __launch_bounds__(EB) void k_encode_ssb(const double2* __restrict__ f, int pn_x, int pn_y,
                                                  const double4* __restrict__ tx, const double4* __restrict__ ty,
                                                  double* __restrict__ e)
{
#pragma clang fp contract(off)
    const int64_t n = (int64_t)pn_x * pn_y;
    const int hx = pn_x >> 1, hy = pn_y >> 1;
    const double nn = (double)pn_x * (double)pn_y;
    int64_t i = (int64_t)blockIdx.x * EB + threadIdx.x;
    const int64_t stride = (int64_t)gridDim.x * EB;
    for (; i < n; i += stride) {
        const int w = (int)(i % pn_x), h = (int)(i / pn_x);
        int sw = w + 2 * hx, sh = h + 2 * hy;
        if (sw >= pn_x) sw -= pn_x;
        if (sh >= pn_y) sh -= pn_y;
        double2 v = f[(int64_t)sh * pn_x + sw];
        const double4 a = tx[w], b = ty[h];
        if (pn_x & 1) v = make_double2(v.x * a.z - v.y * a.w, v.x * a.w + v.y * a.z);
        if (pn_y & 1) v = make_double2(v.x * b.z - v.y * b.w, v.x * b.w + v.y * b.z);
        const double carrier = 1.0 * (a.x * b.x - a.y * b.y);
        e[i] = (nn * v.x) * carrier;
    }
}

struct ChanPtrs {
    const double* p[MAXCH];
    uint8_t* o[MAXCH];
};

__global__ void k_minmax_init(unsigned long long* mm)
{
    if (threadIdx.x == 0) {
        mm[0] = order_key(DBL_MAX);
        mm[1] = order_key(DBL_MIN);
    }
}

__device__ inline void minmax_upd(double v, unsigned long long& kmin, unsigned long long& kmax)
{
    if (v == v) {
        const unsigned long long k = order_key(v);
        kmin = k < kmin ? k : kmin;
        kmax = k > kmax ? k : kmax;
    }
}

// Block-wide min/max of the order keys (wave shuffles, then LDS), merged into
// mm[0] / mm[1] with one 64-bit atomic each per block.
__device__ inline void commit_minmax(unsigned long long kmin, unsigned long long kmax, unsigned long long* mm)
{
    for (int off = 32; off > 0; off >>= 1) {
        const unsigned long long a = __shfl_xor(kmin, off);
        const unsigned long long b = __shfl_xor(kmax, off);
        kmin = a < kmin ? a : kmin;
        kmax = b > kmax ? b : kmax;
    }
    __shared__ unsigned long long smin[EB / 64], smax[EB / 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) {
        smin[wave] = kmin;
        smax[wave] = kmax;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < EB / 64; ++w) {
            kmin = smin[w] < kmin ? smin[w] : kmin;
            kmax = smax[w] > kmax ? smax[w] : kmax;
        }
        if (kmin != ~0ull) atomicMin(&mm[0], kmin);
        if (kmax != 0ull) atomicMax(&mm[1], kmax);
    }
}

// Few blocks (MINMAX_BLOCKS) so that only a few hundred 64-bit atomics meet
// on the two result words; each thread keeps four 16-byte loads in flight.
// Channel buffers that are not 16-byte aligned take the scalar loop.
constexpr int MINMAX_BLOCKS = 512;
// Fused encode + min/max: ENCODE_MM_BLOCKS (2048) blocks, eight per CU, which
// fills every SIMD (measured at 3840x2160: 54 us per hologram against 56 us
// with 4096 and 64 us with 8192 blocks).  Each block stores its min/max keys
// to its own pair of words and one block reduces them (k_mm_final): 64-bit
// atomics on the two result words serialise at the L2 (~10 ns each; 16k of
// them cost ~160 us per launch).

__global__ __launch_bounds__(EB) void k_minmax(ChanPtrs cp, int nch, int64_t n, unsigned long long* mm)
{
    unsigned long long kmin = ~0ull, kmax = 0ull;
    const int64_t tid = (int64_t)blockIdx.x * EB + threadIdx.x;
    const int64_t stride = (int64_t)gridDim.x * EB;
    for (int c = 0; c < nch; ++c) {
        const double* __restrict__ e = cp.p[c];
        int64_t done = 0;
        if ((reinterpret_cast<uintptr_t>(e) & 15) == 0) {
            const double2* __restrict__ e2 = reinterpret_cast<const double2*>(e);
            const int64_t n2 = n / 2;
            int64_t i = tid;
            for (; i + 3 * stride < n2; i += 4 * stride) {
                const double2 a = e2[i], b = e2[i + stride], d = e2[i + 2 * stride], f = e2[i + 3 * stride];
                minmax_upd(a.x, kmin, kmax);
                minmax_upd(a.y, kmin, kmax);
                minmax_upd(b.x, kmin, kmax);
                minmax_upd(b.y, kmin, kmax);
                minmax_upd(d.x, kmin, kmax);
                minmax_upd(d.y, kmin, kmax);
                minmax_upd(f.x, kmin, kmax);
                minmax_upd(f.y, kmin, kmax);
            }
            for (; i < n2; i += stride) {
                const double2 a = e2[i];
                minmax_upd(a.x, kmin, kmax);
                minmax_upd(a.y, kmin, kmax);
            }
            done = 2 * n2;
        }
        for (int64_t i = done + tid; i < n; i += stride) minmax_upd(e[i], kmin, kmax);
    }
    commit_minmax(kmin, kmax, mm);
}

// Encoders of every channel that also take the global min/max of the encoded
// values (the normaliser's first pass), so that ohx_encode_normalize_u8 reads
// the encoded planes only once more, in k_to_u8.  Values and min/max equal
// k_encode's and k_minmax's bit for bit.
struct EncPtrs {
    const double2* f[MAXCH];
    double* e[MAXCH];
};

// Block-wide min/max of the order keys, stored to part[2 b], part[2 b + 1].
__device__ inline void store_block_minmax(unsigned long long kmin, unsigned long long kmax,
                                          unsigned long long* __restrict__ part)
{
    for (int off = 32; off > 0; off >>= 1) {
        const unsigned long long a = __shfl_xor(kmin, off);
        const unsigned long long b = __shfl_xor(kmax, off);
        kmin = a < kmin ? a : kmin;
        kmax = b > kmax ? b : kmax;
    }
    __shared__ unsigned long long smin[EB / 64], smax[EB / 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) {
        smin[wave] = kmin;
        smax[wave] = kmax;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < EB / 64; ++w) {
            kmin = smin[w] < kmin ? smin[w] : kmin;
            kmax = smax[w] > kmax ? smax[w] : kmax;
        }
        part[2 * blockIdx.x] = kmin;
        part[2 * blockIdx.x + 1] = kmax;
    }
}

// mm[0] / mm[1] = min / max over the nb blocks' keys (one block)
__global__ __launch_bounds__(EB) void k_mm_final(const unsigned long long* __restrict__ part, int nb,
                                                 unsigned long long* __restrict__ mm)
{
    unsigned long long kmin = ~0ull, kmax = 0ull;
    for (int b = threadIdx.x; b < nb; b += EB) {
        const unsigned long long a = part[2 * b], c = part[2 * b + 1];
        kmin = a < kmin ? a : kmin;
        kmax = c > kmax ? c : kmax;
    }
    for (int off = 32; off > 0; off >>= 1) {
        const unsigned long long a = __shfl_xor(kmin, off);
        const unsigned long long b = __shfl_xor(kmax, off);
        kmin = a < kmin ? a : kmin;
        kmax = b > kmax ? b : kmax;
    }
    __shared__ unsigned long long smin[EB / 64], smax[EB / 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) {
        smin[wave] = kmin;
        smax[wave] = kmax;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < EB / 64; ++w) {
            kmin = smin[w] < kmin ? smin[w] : kmin;
            kmax = smax[w] > kmax ? smax[w] : kmax;
        }
        // merged into k_minmax_init's starting keys, as k_minmax's atomics are
        const unsigned long long k0 = order_key(DBL_MAX), k1 = order_key(DBL_MIN);
        mm[0] = kmin < k0 ? kmin : k0;
        mm[1] = kmax > k1 ? kmax : k1;
    }
}

template <int FLAG>
__global__ __launch_bounds__(EB) void k_encode_mm(EncPtrs ep, int nch, int64_t n, unsigned long long* part)
{
    unsigned long long kmin = ~0ull, kmax = 0ull;
    const int64_t tid = (int64_t)blockIdx.x * EB + threadIdx.x;
    const int64_t stride = (int64_t)gridDim.x * EB;
    for (int c = 0; c < nch; ++c) {
        const double2* __restrict__ f = ep.f[c];
        double* __restrict__ e = ep.e[c];
        int64_t i = tid;
        for (; i + stride < n; i += 2 * stride) {
            const double2 h0 = f[i], h1 = f[i + stride];
            const double v0 = encode_value<FLAG>(h0), v1 = encode_value<FLAG>(h1);
            e[i] = v0;
            e[i + stride] = v1;
            minmax_upd(v0, kmin, kmax);
            minmax_upd(v1, kmin, kmax);
        }
        if (i < n) {
            const double v = encode_value<FLAG>(f[i]);
            e[i] = v;
            minmax_upd(v, kmin, kmax);
        }
    }
    store_block_minmax(kmin, kmax, part);
}

__device__ inline uint8_t to_u8(double v, double mn, double gap)
{
    const double t = ((v - mn) / gap) * 255 + 0.5;
    return (t >= 0.0 && t < 256.0) ? (uint8_t)(int)t : (uint8_t)0;
}

// Four pixels per thread (two 16-byte loads, one 4-byte store) where the
// buffers are aligned for it; the remainder and unaligned buffers go pixel by
// pixel.
__global__ __launch_bounds__(EB) void k_to_u8(ChanPtrs cp, int nch, int64_t n, const unsigned long long* mm)
{
    const double mn = key_value(mm[0]), mx = key_value(mm[1]);
    const double gap = mx - mn;
    const int64_t tid = (int64_t)blockIdx.x * EB + threadIdx.x;
    const int64_t stride = (int64_t)gridDim.x * EB;
    for (int c = 0; c < nch; ++c) {
        const double* __restrict__ e = cp.p[c];
        uint8_t* __restrict__ o = cp.o[c];
        int64_t done = 0;
        if ((reinterpret_cast<uintptr_t>(e) & 15) == 0 && (reinterpret_cast<uintptr_t>(o) & 3) == 0) {
            const int64_t n4 = n / 4;
            for (int64_t i = tid; i < n4; i += stride) {
                const double2 a = reinterpret_cast<const double2*>(e)[2 * i];
                const double2 b = reinterpret_cast<const double2*>(e)[2 * i + 1];
                const unsigned w = (unsigned)to_u8(a.x, mn, gap) | ((unsigned)to_u8(a.y, mn, gap) << 8) |
                                   ((unsigned)to_u8(b.x, mn, gap) << 16) | ((unsigned)to_u8(b.y, mn, gap) << 24);
                reinterpret_cast<unsigned*>(o)[i] = w;
            }
            done = 4 * n4;
        }
        for (int64_t i = done + tid; i < n; i += stride) o[i] = to_u8(e[i], mn, gap);
    }
}

static unsigned egrid(int64_t n)
{
    int64_t g = (n + EB - 1) / EB;
    if (g > 8192) g = 8192;
    return (unsigned)(g < 1 ? 1 : g);
}

}  // namespace ohx

using namespace ohx;

extern "C" int ohx_encode(ohx_ctx* c, const double* d_field, int64_t n, double* d_enc, uint32_t flag)
{
    if (!c) return OHX_ERR_INVALID_ARG;
    if (n < 0 || (n > 0 && (!d_field || !d_enc)))
        return set_error(c, OHX_ERR_INVALID_ARG, "null buffer or negative size");
    if (flag > OHX_ENCODE_IMAGINARY)
        return set_error(c, OHX_ERR_UNSUPPORTED, "encode flag not on the point-cloud path");
    if (n == 0) return OHX_OK;
    (void)hipSetDevice(c->device);
    const double2* f = (const double2*)d_field;
    const dim3 g(egrid(n)), b(EB);
    switch (flag) {
    case OHX_ENCODE_PHASE: hipLaunchKernelGGL(k_encode<OHX_ENCODE_PHASE>, g, b, 0, c->stream, f, n, d_enc); break;
    case OHX_ENCODE_AMPLITUDE: hipLaunchKernelGGL(k_encode<OHX_ENCODE_AMPLITUDE>, g, b, 0, c->stream, f, n, d_enc); break;
    case OHX_ENCODE_REAL: hipLaunchKernelGGL(k_encode<OHX_ENCODE_REAL>, g, b, 0, c->stream, f, n, d_enc); break;
    default: hipLaunchKernelGGL(k_encode<OHX_ENCODE_IMAGINARY>, g, b, 0, c->stream, f, n, d_enc); break;
    }
    return check_hip(c, hipGetLastError(), "k_encode");
}

extern "C" int ohx_encode_ssb(ohx_ctx* c, const double* d_field, int32_t pn_x, int32_t pn_y, double pp_x,
                              double pp_y, double shift_x, double shift_y, double* d_enc)
{
    if (!c) return OHX_ERR_INVALID_ARG;
    if (pn_x < 1 || pn_y < 1 || pn_x > (1 << 30) || pn_y > (1 << 30) || (int64_t)pn_x * pn_y > ((int64_t)1 << 31) ||
        !d_field || !d_enc)
        return set_error(c, OHX_ERR_INVALID_ARG, "null buffer or pixel number out of range");
    if (!(pp_x > 0) || !(pp_y > 0) || !std::isfinite(pp_x) || !std::isfinite(pp_y) || !std::isfinite(shift_x) ||
        !std::isfinite(shift_y))
        return set_error(c, OHX_ERR_INVALID_ARG, "pixel pitch not positive or shift not finite");
    std::lock_guard<std::mutex> lk(c->mu);
    (void)hipSetDevice(c->device);
    const int64_t n = (int64_t)pn_x * pn_y;
    // per-column and per-row tables in the context's scratch
    const size_t tx_bytes = ((size_t)pn_x * sizeof(double4) + 255) / 256 * 256;
    int st = ensure_gemm_scratch(c, tx_bytes + (size_t)pn_y * sizeof(double4));
    if (st) return st;
    double4* tx = (double4*)c->gemm_scratch;
    double4* ty = (double4*)((char*)c->gemm_scratch + tx_bytes);
    hipLaunchKernelGGL(k_ssb_tables, dim3((unsigned)((pn_x + pn_y + EB - 1) / EB)), dim3(EB), 0, c->stream, pn_x,
                       pn_y, pp_x, pp_y, shift_x, shift_y, tx, ty);
    hipLaunchKernelGGL(k_encode_ssb, dim3(egrid(n)), dim3(EB), 0, c->stream, (const double2*)d_field, pn_x, pn_y,
                       (const double4*)tx, (const double4*)ty, d_enc);
    return check_hip(c, hipGetLastError(), "k_encode_ssb");
}

extern "C" int ohx_normalize_u8(ohx_ctx* c, const double* const* d_enc, int nch, int64_t n,
                                uint8_t* const* d_out)
{
    if (!c) return OHX_ERR_INVALID_ARG;
    if (nch < 1 || nch > MAXCH || n < 0 || !d_enc || !d_out)
        return set_error(c, OHX_ERR_INVALID_ARG, "bad channel count or buffers");
    ChanPtrs cp{};
    for (int i = 0; i < nch; ++i) {
        if (n > 0 && (!d_enc[i] || !d_out[i])) return set_error(c, OHX_ERR_INVALID_ARG, "null channel buffer");
        cp.p[i] = d_enc[i];
        cp.o[i] = d_out[i];
    }
    if (n == 0) return OHX_OK;
    std::lock_guard<std::mutex> lk(c->mu);
    (void)hipSetDevice(c->device);
    hipLaunchKernelGGL(k_minmax_init, dim3(1), dim3(64), 0, c->stream, c->minmax);
    const unsigned mg = std::min<unsigned>(egrid(n), MINMAX_BLOCKS);
    hipLaunchKernelGGL(k_minmax, dim3(mg), dim3(EB), 0, c->stream, cp, nch, n, c->minmax);
    hipLaunchKernelGGL(k_to_u8, dim3(egrid((n + 3) / 4)), dim3(EB), 0, c->stream, cp, nch, n,
                       (const unsigned long long*)c->minmax);
    return check_hip(c, hipGetLastError(), "k_normalize");
}

namespace ohx {

int launch_encode_normalize_u8(hipStream_t stream, unsigned long long* mm, unsigned long long* part,
                               const double* const* d_field, int nch,
                               int64_t n, uint32_t flag, double* const* d_enc, uint8_t* const* d_out)
{
    EncPtrs ep{};
    ChanPtrs cp{};
    for (int i = 0; i < nch; ++i) {
        ep.f[i] = (const double2*)d_field[i];
        ep.e[i] = d_enc[i];
        cp.p[i] = d_enc[i];
        cp.o[i] = d_out[i];
    }
    const unsigned nb = std::min<unsigned>(egrid(n), ENCODE_MM_BLOCKS);
    const dim3 g(nb), b(EB);
    switch (flag) {
    case OHX_ENCODE_PHASE:
        hipLaunchKernelGGL(k_encode_mm<OHX_ENCODE_PHASE>, g, b, 0, stream, ep, nch, n, part);
        break;
    case OHX_ENCODE_AMPLITUDE:
        hipLaunchKernelGGL(k_encode_mm<OHX_ENCODE_AMPLITUDE>, g, b, 0, stream, ep, nch, n, part);
        break;
    case OHX_ENCODE_REAL:
        hipLaunchKernelGGL(k_encode_mm<OHX_ENCODE_REAL>, g, b, 0, stream, ep, nch, n, part);
        break;
    default:
        hipLaunchKernelGGL(k_encode_mm<OHX_ENCODE_IMAGINARY>, g, b, 0, stream, ep, nch, n, part);
        break;
    }
    hipLaunchKernelGGL(k_mm_final, dim3(1), dim3(EB), 0, stream, (const unsigned long long*)part, (int)nb, mm);
    hipLaunchKernelGGL(k_to_u8, dim3(egrid((n + 3) / 4)), dim3(EB), 0, stream, cp, nch, n,
                       (const unsigned long long*)mm);
    return hipGetLastError() == hipSuccess ? OHX_OK : OHX_ERR_DEVICE;
}

}  // namespace ohx

extern "C" int ohx_encode_normalize_u8(ohx_ctx* c, const double* const* d_field, int nch, int64_t n,
                                       uint32_t flag, double* const* d_enc, uint8_t* const* d_out)
{
    if (!c) return OHX_ERR_INVALID_ARG;
    if (nch < 1 || nch > MAXCH || n < 0 || !d_field || !d_enc || !d_out)
        return set_error(c, OHX_ERR_INVALID_ARG, "bad channel count or buffers");
    if (flag > OHX_ENCODE_IMAGINARY)
        return set_error(c, OHX_ERR_UNSUPPORTED, "encode flag not on the point-cloud path");
    for (int i = 0; i < nch; ++i)
        if (n > 0 && (!d_field[i] || !d_enc[i] || !d_out[i]))
            return set_error(c, OHX_ERR_INVALID_ARG, "null channel buffer");
    if (n == 0) return OHX_OK;
    std::lock_guard<std::mutex> lk(c->mu);
    (void)hipSetDevice(c->device);
    if (launch_encode_normalize_u8(c->stream, c->minmax, c->minmax + 8, d_field, nch, n, flag, d_enc, d_out) !=
        OHX_OK)
        return check_hip(c, hipGetLastError(), "k_encode_normalize");
    return OHX_OK;
}
