// Allocation-free variant of ohx_pc_diffract, for callers that cannot call
// hipMalloc (code under HIP graph capture, tight serving loops, harnesses
// that forbid allocation).
//
// Its scratch -- the per-point constants k_pc_prep writes for the direct
// kernels and two flag words -- lives in a file-scope device array, `scratch`,
// sized for OHX_NOALLOC_MAX_POINTS points.  Each GPU holds its own copy from
// module load until the process ends.  Host code never passes the array's
// name to a kernel: the name is a host-side stand-in, not a device address.
// ohx_ctx_create calls noalloc_init once per device with that device current;
// it takes the device address with hipGetSymbolAddress and keeps it per
// device, and every kernel argument below is derived from that address.
//
// Every call on one GPU shares that GPU's array and leaves its contents
// behind: the flags are cleared by a fill kernel before use and the point
// constants are fully rewritten by k_pc_prep.  Calls that overlap on one GPU
// (several contexts, or several host threads) race on the array; callers
// serialise them.
#include <mutex>

#include "ohx_internal.h"

namespace ohx {
namespace {

constexpr int MAX_DEVICES = 64;
// [flags: 256 B][point constants: points_bytes(OHX_NOALLOC_MAX_POINTS)]
constexpr size_t FLAGS_BYTES = 256;
constexpr size_t SCRATCH_BYTES =
    FLAGS_BYTES + (size_t)OHX_NOALLOC_MAX_POINTS * (6 * sizeof(double) + sizeof(int4)) + 8 * 256;

static __device__ __align__(256) unsigned char scratch[SCRATCH_BYTES];

std::mutex g_mu;
void* g_ptr[MAX_DEVICES] = {};
int g_status[MAX_DEVICES] = {};   // 0: not set up, 1: ok, < 0: hipGetSymbolAddress failed

}  // namespace

void noalloc_init(int device)
{
    if (device < 0 || device >= MAX_DEVICES) return;
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_status[device] != 0) return;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || p == nullptr) {
        (void)hipGetLastError();
        g_status[device] = -1;
        return;
    }
    g_ptr[device] = p;
    g_status[device] = 1;
}

}  // namespace ohx

using namespace ohx;

extern "C" {

int ohx_pc_diffract_noalloc(ohx_ctx* c, const void* d_vertices, int64_t n, double* d_field,
                            const ohx_pc_params* p, uint32_t diff, uint32_t mode)
{
    if (!c) return OHX_ERR_INVALID_ARG;
    if (n < 0 || (n > 0 && !d_vertices) || !d_field)
        return set_error(c, OHX_ERR_INVALID_ARG, "null buffer or negative point count");
    if (n > OHX_NOALLOC_MAX_POINTS)
        return set_error(c, OHX_ERR_INVALID_ARG,
                         "ohx_pc_diffract_noalloc: n_points " + std::to_string(n) + " exceeds the maximum " +
                             std::to_string((long long)OHX_NOALLOC_MAX_POINTS) + " of its static scratch");
    if (mode & (OHX_MODE_MFMA_BF16X1 | OHX_MODE_MFMA_BF16X3 | OHX_MODE_MFMA_FP16X3))
        return set_error(c, OHX_ERR_UNSUPPORTED,
                         "ohx_pc_diffract_noalloc: the matrix-core modes need split-K planes; direct modes only");
    Geometry g;
    int st = geometry_from_params(c, p, diff, &g);
    if (st) return st;
    std::lock_guard<std::mutex> lk(c->mu);
    (void)hipSetDevice(c->device);
    void* base = nullptr;
    {
        std::lock_guard<std::mutex> lk2(g_mu);
        if (c->device >= 0 && c->device < MAX_DEVICES && g_status[c->device] == 1) base = g_ptr[c->device];
    }
    if (!base) return set_error(c, OHX_ERR_DEVICE, "ohx_pc_diffract_noalloc: static scratch address unavailable");
    c->last = ohx_launch_stats{};
    c->last.pairs_nominal = n * (int64_t)g.pn_x * g.pn_y;
    c->boxsum_pending = false;
    if (n == 0) return OHX_OK;
    unsigned long long* flags = (unsigned long long*)base;   // [0] box pixels, [1] EXACT pairs left
    PointSoA soa;
    carve_points((char*)base + FLAGS_BYTES, n, &soa);
    if ((st = launch_prep(c, d_vertices, g, soa, flags))) return st;
    c->boxsum_pending = true;
    c->boxsum_src = flags;
    const bool fp64 = (mode & (OHX_MODE_FLOAT | OHX_MODE_FASTMATH)) == 0;
    return launch_direct_with(c, g, soa, d_field, fp64, flags + 1, false);
}

}  // extern "C"
