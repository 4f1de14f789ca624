// Internal declarations shared by the HIP sources of libopenholo_amd.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>
#include <string>

#include "../../include/openholo_amd.h"

struct ohx_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    // field read-back that overlaps later work on `stream` (ohx_memcpy_d2h_async_after)
    hipStream_t copy_stream = nullptr;
    hipEvent_t copy_ready = nullptr;
    std::mutex mu;                 // serialises use of the scratch buffers
    std::string last_error;
    // per-point SoA scratch (grown on demand)
    void* pts = nullptr;
    size_t pts_bytes = 0;
    // small device scratch: [0..1] min/max of the normaliser, [4] pixels in
    // the mask boxes of the last launch, [5] FP32 direct kernel left EXACT pairs,
    // [8 ..] the fused encoder's per-block min/max (ENCODE_MM_PART_WORDS)
    unsigned long long* minmax = nullptr;
    // split-K partial sums of the matrix-core Fresnel path and the point-split
    // partial planes of the direct kernels (grown on demand)
    void* gemm_scratch = nullptr;
    size_t gemm_scratch_bytes = 0;
    // RS matrix-core path: point classification, partition and the two point
    // subsets' constants (grown on demand)
    void* rs_scratch = nullptr;
    size_t rs_scratch_bytes = 0;
    // RS matrix-core path: the masked points' tile mask (grown on demand; apart
    // from gemm_scratch, which the direct kernel's point splits reuse)
    void* rs_mask = nullptr;
    size_t rs_mask_bytes = 0;
    // hipFFT plans of the WRP propagation (hipfftHandle) for an L x M grid:
    // length-L transforms of ceil(M/4) contiguous rows of the transposed
    // grid, length-M transforms of all L rows
    void* wrp_plan_t = nullptr;
    void* wrp_plan_a = nullptr;
    int wrp_plan_L = 0, wrp_plan_M = 0;
    ohx_launch_stats last{};
    bool boxsum_pending = false;   // last.pairs_in_box still on the device (*boxsum_src)
    const unsigned long long* boxsum_src = nullptr;   // minmax + 4, or the no-alloc scratch
};

namespace ohx {

// sin(2 pi x) and cos(2 pi x) of an FP32 argument in revolutions.
// v_sin_f32 / v_cos_f32 take revolutions and reduce them themselves, so the
// FP32 kernels pass unreduced phases (below 65 cycles in the direct sums,
// below 21 in the matrix-core operand generation); ohx_sincos_rev evaluates
// exactly these two functions so that tests can pin their accuracy there.
__device__ inline float sin_rev(float x) { return __builtin_amdgcn_sinf(x); }
__device__ inline float cos_rev(float x) { return __builtin_amdgcn_cosf(x); }

// Points per LDS chunk of the FP32 direct kernel: its FP32 partial sums are
// folded into FP64 at chunk ends, so a point slice that is a whole number of
// chunks sums exactly as the one-shot launch does.
constexpr int DIRECT_CHUNK = 120;

// Per-point constants produced by the prep kernel (structure of arrays).
struct PointSoA {
    double* px;     // scaled x
    double* py;     // scaled y
    double* pz;     // scaled z + distance
    double* cyc0;   // frac(z/lambda [+ phase/2pi]) in cycles
    double* kamp;   // Fresnel: a/(lambda z); RS: a z / lambda
    double* q;      // Fresnel: 1/(2 lambda z); RS: unused (0)
    int4* box;      // [x0,x1) x [y0,y1) mask box, clamped
    int64_t n;
};

// Derived launch constants, computed on the host with the reference's
// operation order (contraction off) so that box bounds are bit-identical.
struct Geometry {
    int32_t pn_x, pn_y, off_x, off_y;
    double pp_x, pp_y;
    double x0, y0;       // -ss_x/2, -ss_y/2
    double lambda, inv_lambda;
    double hx_coef, hy_coef;  // RS: tx/sqrt(1-tx^2); Fresnel: unused
    double scale[3];
    double distance;
    int32_t channel, use_phase;
    int32_t diff;
    // RS direct kernels only: non-null skips the (point, tile) pairs whose bit
    // is set in this mask (the masked matrix-core instance took them): row p
    // (the launch's point index) of rs_mask_words words holds bit t for the
    // 128 x 128 tile t (row-major over the hologram, rs_tile_index); null
    // (the default) skips nothing
    const unsigned* rs_skip_mask;
    int32_t rs_mask_words;
};

// ---- RS on the matrix cores: the expansion bound on one tile ----------------
// The rank-2 expansion of the RS term about the reference pixel (offset 64, 64)
// of a 128 x 128 tile neglects terms that are small when, with dxc, dyc the
// point's distances to that pixel and u, v <= 64 pitches, n = um (2 dxc + um),
// N = vm (2 dyc + vm), e = (n + N) / z^2 and k = 1/(8 z^3 lambda) (r_c >= z):
// third-order phase e^3 z/(16 lambda) <= RS_MAX_PHASE3 cycles, second-order
// amplitude e^2 <= RS_MAX_AMP2 and cross term 4 pi k n N <= max_theta rad.
// The classification bounds every tile of a box with the largest dxc, dyc;
// per tile, a mask (k_rs_tilemask) records where the bound and the range test
// hold: the masked matrix-core kernel takes those tiles and the direct kernel
// (Geometry::rs_skip_mask) the others.
constexpr int RS_TILE = 128;
constexpr double RS_MAX_PHASE3 = 2e-7;
constexpr double RS_MAX_AMP2 = 1e-7;
constexpr double RS_FOUR_PI = 12.566370614359172953850573533118;

// um, vm: 64 pitches ((RS_TILE / 2) * pp); il: 1 / lambda.  In FP32 (the
// limits are far from the rounding), multiplied out: e^3 z il / 16 <= P3,
// e^2 <= A2 and 4 pi (il / (8 z^3)) n N <= theta with e = (n + N) / z^2.
__host__ __device__ inline bool rs_bound_ok(double um, double vm, double il, double dxc, double dyc, double z,
                                            double max_theta)
{
#pragma clang fp contract(off)
    const float u = (float)um, v = (float)vm, l = (float)il, zf = (float)z;
    const float nx = u * (2.0f * (float)dxc + u), ny = v * (2.0f * (float)dyc + v);
    const float zz = zf * zf, sn = nx + ny;
    return sn * sn * sn * l <= (float)(16.0 * RS_MAX_PHASE3) * (zz * zz * zf) &&
           sn * sn <= (float)RS_MAX_AMP2 * (zz * zz) &&
           (float)(0.125 * RS_FOUR_PI) * l * nx * ny <= (float)max_theta * (zz * zf);
}
__host__ __device__ inline bool rs_bound_ok(const Geometry& g, double dxc, double dyc, double z, double max_theta)
{
    return rs_bound_ok((RS_TILE / 2) * g.pp_x, (RS_TILE / 2) * g.pp_y, g.inv_lambda, dxc, dyc, z, max_theta);
}
// x / y of the reference pixel of the 128 x 128 tile holding column c / row r
// (pixel coordinates as ophGen.cpp:413,422)
__host__ __device__ inline double rs_tile_xc(const Geometry& g, int c)
{
#pragma clang fp contract(off)
    return g.x0 + (double)((c / RS_TILE) * RS_TILE + RS_TILE / 2 - 1 + g.off_x) * g.pp_x;
}
__host__ __device__ inline double rs_tile_yc(const Geometry& g, int r)
{
#pragma clang fp contract(off)
    return g.y0 + (double)(g.pn_y - ((r / RS_TILE) * RS_TILE + RS_TILE / 2) + g.off_y) * g.pp_y;
}
// 128 x 128 tiles along x, tile index of the tile holding pixel (c, r), and
// 32-bit words of one mask row
__host__ __device__ inline int rs_tiles_x(const Geometry& g) { return (g.pn_x + RS_TILE - 1) / RS_TILE; }
__host__ __device__ inline int rs_tile_index(const Geometry& g, int c, int r)
{
    return (r / RS_TILE) * rs_tiles_x(g) + c / RS_TILE;
}
__host__ __device__ inline int rs_mask_words(const Geometry& g)
{
    return (rs_tiles_x(g) * ((g.pn_y + RS_TILE - 1) / RS_TILE) + 31) / 32;
}
// bit t of mask row `row`
__device__ inline bool rs_mask_bit(const unsigned* __restrict__ row, int t)
{
    return (row[t >> 5] >> (t & 31)) & 1u;
}

// One job of the persistent multi-job direct kernel (serving engine).
struct DirectJob {
    Geometry g;
    PointSoA s;
    double2* field;
    int64_t p0, p1;       // point slice of this iteration
    int tiles_x, tiles;   // tiles of the hologram (64 x 32 FP32, 128 x 16 FP64)
    int item_base;        // prefix sum of tiles over the table
    int fp64;
};

int set_error(ohx_ctx* ctx, int status, const std::string& msg);
int check_hip(ohx_ctx* ctx, hipError_t e, const char* what);
int geometry_from_params(ohx_ctx* ctx, const ohx_pc_params* p, uint32_t diff, Geometry* g);
int ensure_points(ohx_ctx* ctx, int64_t n, PointSoA* soa);
size_t points_bytes(int64_t n);
void carve_points(void* base, int64_t n, PointSoA* soa);
int ensure_gemm_scratch(ohx_ctx* ctx, size_t bytes);

// launchers (pc_direct.hip)
int launch_prep(ohx_ctx* ctx, const void* d_vertices, const Geometry& g, const PointSoA& soa,
                unsigned long long* boxsum = nullptr);
int launch_prep_on(hipStream_t stream, const void* d_vertices, const Geometry& g, const PointSoA& soa,
                   unsigned long long* boxsum = nullptr);
// fill kernel on a given stream (ohx_core.hip)
int field_zero_on(hipStream_t stream, double* f, int64_t n_px);
// encode every channel, min/max in the same pass, then 8-bit (encode.hip);
// mm: two device words (the min/max keys), part: ENCODE_MM_PART_WORDS device
// words (per-block min/max keys of the encode pass)
constexpr int ENCODE_MM_BLOCKS = 2048;
constexpr int ENCODE_MM_PART_WORDS = 2 * ENCODE_MM_BLOCKS;
int launch_encode_normalize_u8(hipStream_t stream, unsigned long long* mm, unsigned long long* part,
                               const double* const* d_field, int nch,
                               int64_t n, uint32_t flag, double* const* d_enc, uint8_t* const* d_out);
int launch_direct(ohx_ctx* ctx, const Geometry& g, const PointSoA& soa, double* d_field, bool fp64);
// flag: device word for the FP32 kernel's "EXACT pairs left" flag;
// allow_splits = false runs every tile's points in one workgroup (no partial
// planes, no occupancy query)
int launch_direct_with(ohx_ctx* ctx, const Geometry& g, const PointSoA& soa, double* d_field, bool fp64,
                       unsigned long long* flag, bool allow_splits);
// allocation-free entry point's per-device scratch (noalloc.hip): called by
// ohx_ctx_create with the device current
void noalloc_init(int device);
int launch_fresnel_gemm(ohx_ctx* ctx, const Geometry& g, const PointSoA& soa, double* d_field, uint32_t mode);
int direct_tiles(const Geometry& g, bool fp64, int* tiles_x);
int wrp_generate(ohx_ctx* ctx, const void* d_vertices, int64_t n, const ohx_wrp_params* p, double* d_field,
                 double* d_wrp);
void wrp_release(ohx_ctx* ctx);
int launch_direct_jobs(hipStream_t stream, const DirectJob* d_jobs, const int* d_njobs, int grid);
int launch_fresnel_gemm_range(hipStream_t stream, void* scratch, size_t scratch_bytes, const Geometry& g,
                              const PointSoA& soa, int64_t p0, int64_t p1, double* d_field, uint32_t mode);
size_t fresnel_gemm_scratch_bytes(const Geometry& g, int64_t npoints);
// field += -j * sum of `splits` FP32 partial planes (fixed order), pc_gemm.hip
int launch_gemm_reduce(hipStream_t stream, const void* part, int64_t splits, int64_t npx, double* d_field);
// RS on matrix cores (rs_gemm.hip): points whose separable rank-2 form holds on
// every tile go to the GEMM, the others to the FP32 direct kernel
int launch_rs_gemm(ohx_ctx* ctx, const Geometry& g, const PointSoA& soa, double* d_field, uint32_t mode);
// its pieces, for callers that slice the work: the device-side partition of
// soa into part (matrix-core points first, those whose box is the whole
// hologram before the masked ones, each group in order; (*d_count)[0] = the
// matrix-core points, (*d_count)[1] = the unmasked ones, device words inside
// scratch; split puts every point whose expansion may hold on a tile in the
// masked group, whose instance then takes only the tiles set in its tile mask
// (k_rs_tilemask), the direct kernel the others with Geometry::rs_skip_mask)
// and the GEMM over the partitioned points [p0, p1), read from their records
// rec (rs_partition_records), into d_field (points from n_full on masked,
// with mask rows from point n_full on if mask is given; wgs, if given, adds
// the workgroups launched)
size_t rs_partition_scratch_bytes(int64_t n);
// the partitioned points' records (the matrix-core kernels' input) inside
// the partition scratch of n points
const void* rs_partition_records(const void* scratch, int64_t n);
int launch_rs_partition(hipStream_t st, void* scratch, size_t scratch_bytes, const Geometry& g, const PointSoA& soa,
                        const PointSoA& part, uint32_t mode, bool split, unsigned long long** d_count);
size_t rs_gemm_scratch_bytes(const Geometry& g, int64_t npoints);
int launch_rs_gemm_range(hipStream_t st, void* scratch, size_t scratch_bytes, const Geometry& g,
                         const void* rec, int64_t p0, int64_t p1, int64_t n_full, double* d_field,
                         uint32_t mode, int64_t* wgs = nullptr, const unsigned* mask = nullptr,
                         int mask_words = 0);
int launch_probe(ohx_ctx* ctx, const Geometry& g, const PointSoA& soa, const int64_t* d_pix,
                 int64_t n_pix, double* d_out);

}  // namespace ohx
