/*
 * openholo_amd.h -- C ABI of the MI355X point-cloud hologram library
 * (libopenholo_amd.so, gfx950).
 *
 * Plain C: opaque handles, plain pointers and sizes, int status codes, no
 * C++ or framework types.  Device pointers ("d_" prefix) are HIP device
 * addresses obtained from ohx_malloc (or any hipMalloc of the same device).
 * Every call is stream-ordered on the context's stream; calls that return
 * host-visible results synchronize that stream before returning.
 *
 * Each entry point names the reference interface it replaces
 * (paths relative to the Openholo source tree).
 */
#ifndef OPENHOLO_AMD_H
#define OPENHOLO_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OHX_ABI_VERSION 1

/* Status codes.  The reference reports CUDA errors by exit(EXIT_FAILURE)
 * (ophgen/src/ophPointCloud_GPU.h:71-81); this library never exits. */
typedef enum ohx_status {
    OHX_OK = 0,
    OHX_ERR_INVALID_ARG = -1,
    OHX_ERR_OUT_OF_MEMORY = -2,
    OHX_ERR_DEVICE = -3,
    OHX_ERR_ADMISSION = -4,   /* serving: a job can never fit the memory budget */
    OHX_ERR_NOT_FOUND = -5,
    OHX_ERR_TIMEOUT = -6,
    OHX_ERR_UNSUPPORTED = -7,
    OHX_ERR_CANCELLED = -8
} ohx_status;

/* Diffraction flags: ophPointCloud::PC_DIFF_FLAG (ophgen/src/ophPointCloud.h:550-553). */
enum { OHX_DIFF_RS = 0, OHX_DIFF_FRESNEL = 1 };

/* Mode bits: MODE_GPU/MODE_FLOAT/MODE_FASTMATH (Openholo/src/define.h:155-159),
 * plus matrix-core variants of the Fresnel sum (no reference counterpart).
 *   0 (MODE_DOUBLE)         : FP64 direct sum (every term in FP64)
 *   MODE_FLOAT [|FASTMATH]   : FP32 direct sum with FP64 per-point phase constants
 *   MFMA_*                   : Fresnel: H = U diag(a) V^T on matrix cores.
 *                              RS with BF16X1 / BF16X3 (and not FP16X3): points
 *                              whose mask is their box (the range test holds
 *                              on all of it) run on matrix cores on the
 *                              128x128 tiles where their rank-2 expansion
 *                              holds and on the FP32 direct sum elsewhere; the
 *                              other points (and RS with FP16X3) run the FP32
 *                              direct sum. */
enum {
    OHX_MODE_GPU = 1,
    OHX_MODE_FLOAT = 2,
    OHX_MODE_FASTMATH = 4,
    OHX_MODE_MFMA_BF16X1 = 0x10,
    OHX_MODE_MFMA_BF16X3 = 0x20,
    OHX_MODE_MFMA_FP16X3 = 0x40
};

/* Encode flags: ophGen::ENCODE_FLAG (ophgen/src/ophGen.h:84-96), the
 * per-pixel encoders on this path. */
enum {
    OHX_ENCODE_PHASE = 0,
    OHX_ENCODE_AMPLITUDE = 1,
    OHX_ENCODE_REAL = 2,
    OHX_ENCODE_IMAGINARY = 3
};

typedef struct ohx_ctx ohx_ctx;

/* One launch's geometry.  Replaces CudaPointCloudConfig / ...RS / ...Fresnel
 * (ophgen/src/ophPointCloud_GPU.h:89-217): the same fields passed by value;
 * k, half_ss, det_tx/ty and tx/ty are derived inside as the reference does. */
typedef struct ohx_pc_params {
    int32_t pn_x, pn_y;       /* pixel_number                                  */
    int32_t offset_x, offset_y; /* start offset (OphConfig::offset)            */
    double pp_x, pp_y;        /* pixel_pitch (m)                               */
    double scale[3];          /* OphPointCloudConfig::scale                    */
    double distance;          /* OphPointCloudConfig::distance (offset_depth)  */
    double lambda;            /* wave length of this channel (m)               */
    int32_t channel;          /* index into Vertex::color (0..2)               */
    int32_t use_phase;        /* 1: amplitude = color[channel]*exp(j*phase); the
                                 reference ignores Vertex::phase (0)           */
} ohx_pc_params;

/* ---- context and memory -------------------------------------------------
 * Replaces the cudaWrapper singleton (ophgen/src/cudaWrapper.h:50-87) and the
 * per-call cudaMalloc/cudaMemcpy of genCghPointCloudGPU
 * (ophgen/src/ophPointCloud_GPU.cpp:218-285). */
int ohx_ctx_create(int device, ohx_ctx** out);
int ohx_ctx_destroy(ohx_ctx* ctx);
void* ohx_ctx_stream(ohx_ctx* ctx);            /* the hipStream_t used for launches */
int ohx_ctx_device(ohx_ctx* ctx);
const char* ohx_status_string(int status);
const char* ohx_last_error(ohx_ctx* ctx);      /* text of the last failure on ctx */
int ohx_mem_info(ohx_ctx* ctx, size_t* free_bytes, size_t* total_bytes);

int ohx_malloc(ohx_ctx* ctx, size_t bytes, void** d_ptr);
int ohx_free(ohx_ctx* ctx, void* d_ptr);
/* Page-locked host memory, zero-filled (hipHostMalloc): read-back into it
 * runs at full link rate and overlaps kernels.  The host library allocates
 * the complex field complex_H this way where the reference uses new[]
 * (ophgen/src/ophGen.cpp:153-168); OHX_ERR_DEVICE when no device is
 * present (the caller then falls back to pageable memory). */
int ohx_host_alloc(size_t bytes, void** ptr);
int ohx_host_free(void* ptr);
int ohx_memcpy_h2d(ohx_ctx* ctx, void* d_dst, const void* src, size_t bytes);
int ohx_memcpy_d2h(ohx_ctx* ctx, void* dst, const void* d_src, size_t bytes);
int ohx_memcpy_d2d(ohx_ctx* ctx, void* d_dst, const void* d_src, size_t bytes);
/* Read-back that overlaps later kernels.  Replaces the per-channel
 * cudaMemcpy of the field in ophgen/src/ophPointCloud_GPU.cpp:272.
 *
 * ohx_memcpy_d2h_async_after enqueues a device-to-host copy of d_src on the
 * context's second (copy) stream, ordered after the work enqueued on ctx's
 * stream up to this call and after nothing enqueued later, and returns at
 * once.  Work enqueued on ctx's stream afterwards (the next channel) runs
 * while the copy is in flight; it must not write d_src, and dst must stay
 * valid and untouched until ohx_copy_synchronize returns.  With page-locked
 * dst (ohx_host_alloc) the copy is asynchronous; with pageable dst the
 * runtime may complete it before returning.
 * ohx_copy_synchronize waits for every copy enqueued so far.
 * ohx_memcpy_d2h_after is the two in sequence for one copy (blocking). */
int ohx_memcpy_d2h_async_after(ohx_ctx* ctx, void* dst, const void* d_src, size_t bytes);
int ohx_copy_synchronize(ohx_ctx* ctx);
int ohx_memcpy_d2h_after(ohx_ctx* ctx, void* dst, const void* d_src, size_t bytes);
int ohx_synchronize(ohx_ctx* ctx);

/* Zero n_px complex FP64 pixels with a fill kernel (resetBuffer,
 * ophgen/src/ophGen.cpp:801-814; no memset node, safe under graph capture). */
int ohx_field_zero(ohx_ctx* ctx, double* d_field, int64_t n_px);

/* ---- the hot path --------------------------------------------------------
 * Accumulate one channel of the point-cloud hologram into d_field
 * (interleaved re,im FP64, pn_x*pn_y complex, row 0 = +y), i.e.
 * d_field += sum_p (1/j) a_p exp(j phi_p(x,y)) over the reference's
 * anti-aliasing mask.  d_vertices: n_points 56-byte Vertex records
 * (Openholo/src/struct.h:86-106) in device memory.
 * Replaces cudaPointCloud_RS and cudaPointCloud_Fresnel
 * (ophgen/src/ophPCKernel.cu:893-933, declared ophgen/src/ophPointCloud_GPU.h:225-233),
 * selected by diff_flag; mode as above. */
int ohx_pc_diffract(ohx_ctx* ctx, const void* d_vertices, int64_t n_points,
                    double* d_field, const ohx_pc_params* params,
                    uint32_t diff_flag, uint32_t mode);

/* Allocation-free variant of ohx_pc_diffract.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  It calls no hipMalloc, hipFree or allocating library routine:
 * its scratch (the per-point constants and two flag words) is a static device
 * array of about 32 MiB per GPU, held from library load until the process
 * ends.  Same arguments and result as ohx_pc_diffract for the direct modes
 * (MODE_DOUBLE, MODE_FLOAT): every tile's points run in one workgroup (no
 * partial planes), so results agree with ohx_pc_diffract to FP64 rounding of
 * the per-pixel sums.  OHX_ERR_INVALID_ARG when n_points exceeds
 * OHX_NOALLOC_MAX_POINTS; OHX_ERR_UNSUPPORTED for the matrix-core modes.
 * All calls on one GPU share its scratch: calls that overlap on one GPU
 * (several contexts or host threads) race, so callers serialise them. */
#define OHX_NOALLOC_MAX_POINTS 524288
int ohx_pc_diffract_noalloc(ohx_ctx* ctx, const void* d_vertices, int64_t n_points,
                            double* d_field, const ohx_pc_params* params,
                            uint32_t diff_flag, uint32_t mode);

/* d_dst += d_src over n_px complex FP64 pixels.
 * Replaces sum_Kernel (ophgen/src/ophPCKernel.cu:874-891). */
int ohx_field_accumulate(ohx_ctx* ctx, double* d_dst, const double* d_src, int64_t n_px);

/* FP64 field at listed pixels (d_pix: row*pn_x+col, int64), every term in FP64
 * with the phase reduced in cycles and compensated summation; d_out is
 * interleaved re,im.  The accuracy probe behind the stated relative-L2
 * errors; no reference counterpart (the reference evaluates in place). */
int ohx_pc_probe_f64(ohx_ctx* ctx, const void* d_vertices, int64_t n_points,
                     const ohx_pc_params* params, uint32_t diff_flag,
                     const int64_t* d_pix, int64_t n_pix, double* d_out);

/* Self-check of the FP32 sine/cosine the FP32 kernels use: d_sin[i] =
 * sin(2 pi x[i]), d_cos[i] = cos(2 pi x[i]) for FP32 arguments in revolutions,
 * evaluated by the same device functions (the hardware instructions, which
 * reduce the argument themselves).  No reference counterpart; the test suite
 * pins their accuracy on unreduced arguments up to the largest the kernels
 * produce. */
int ohx_sincos_rev(ohx_ctx* ctx, const float* d_x, int64_t n, float* d_sin, float* d_cos);

/* ---- WRP (wavefront recording plane) generator -----------------------------
 * One channel of ophWRP::generateHologram (ophgen/src/ophWRP.cpp:437-463):
 * autoScaling (:69-118: positions times scale, zmax), calculateWRPCPU/GPU
 * (:321-435, ophWRP_GPU.cpp:48-200: every point adds a (2w)^2 window of
 * spherical-wave terms at the common depth dz = wrp_location - zmax) and
 * ophGen::Fresnel_FFT (ophgen/src/ophGen.cpp:516-581: zero padding to
 * 2 pn_x x 2 pn_y, centred FFT, angular-spectrum transfer function, centred
 * unnormalised inverse FFT, centre crop).  FP64 throughout; the centring
 * shifts are out of place (the reference shifts in place there, which
 * corrupts its input, Openholo/src/Openholo.cpp:716); only random_phase = 0
 * is supported (the reference's random phase is not reproducible). */
typedef struct ohx_wrp_params {
    int32_t pn_x, pn_y;           /* pixel_number                                */
    double pp_x, pp_y;            /* pixel_pitch (m); the WRP uses the same grid  */
    double scale[3];              /* WRPConfig::scale                            */
    double lambda;                /* wave length of this channel (m)             */
    int32_t channel;              /* index into Vertex::color (0..2)             */
    int32_t random_phase;         /* must be 0                                   */
    double wrp_location;          /* WRPConfig::wrp_location (m)                 */
    double propagation_distance;  /* WRPConfig::propagation_distance (m)         */
} ohx_wrp_params;

/* d_field (pn_x*pn_y interleaved complex FP64) is overwritten with the
 * hologram of this channel (complex_H[ch]); d_wrp, if not null, receives the
 * WRP plane itself (ophWRP::getWRPBuff).  d_vertices are the 56-byte Vertex
 * records on the device.  OHX_ERR_INVALID_ARG for n_points > 2^31 - 1,
 * pn_x * pn_y > 2^28, a window wider than 8190 pixels, random_phase != 0. */
int ohx_wrp_generate(ohx_ctx* ctx, const void* d_vertices, int64_t n_points,
                     const ohx_wrp_params* params, double* d_field, double* d_wrp);

/* ---- encoding and 8-bit normalisation ------------------------------------
 * ophGen::encoding(ENCODE_FLAG) per channel (ophgen/src/ophGen.cpp:816-849,
 * encoders :1821-1864): d_enc[i] = f(d_field[i]). */
int ohx_encode(ohx_ctx* ctx, const double* d_field, int64_t n_px, double* d_enc,
               uint32_t encode_flag);

/* ophPointCloud::encodeHologram (ophgen/src/ophPointCloud.cpp:193-240), the
 * single-side-band encoding behind ophPointCloud::encoding(ENCODE_SSB, passband)
 * (:247-251), one channel.  The reference computes
 *   T = shift(IFFT(shift(shift(FFT(shift(H))))))      (FFTW, unnormalised)
 * with Openholo::fft2 / fftShift (Openholo/src/Openholo.cpp:559-585,698-772),
 * then d_enc[i] = Re(T[i]) * cos(X + Y), X = pi x shift_x / pp_x,
 * Y = pi y shift_y / pp_y at the pixel centre (x, y) of pixel i; band_limit is
 * unused there.  Here T is evaluated in closed form, per axis of size n with
 * h = n/2 (integer): T[k] = n H[k + 2h mod n], times exp(2 pi j (k + h) / n)
 * when n is odd.  d_field, d_enc: pn_x * pn_y interleaved complex / real FP64. */
int ohx_encode_ssb(ohx_ctx* ctx, const double* d_field, int32_t pn_x, int32_t pn_y, double pp_x, double pp_y,
                   double shift_x, double shift_y, double* d_enc);

/* ophGen::normalize (ophgen/src/ophGen.cpp:654-687): global min/max over all
 * n_channels encoded buffers, then u8 = (uint8)(((v-min)/gap)*255 + 0.5).
 * d_enc and d_out are host arrays of n_channels device pointers. */
int ohx_normalize_u8(ohx_ctx* ctx, const double* const* d_enc, int n_channels,
                     int64_t n_px, uint8_t* const* d_out);

/* ophGen::encoding(encode_flag) of every channel followed by ophGen::normalize
 * (ophgen/src/ophGen.cpp:816-849, then :654-687) in one call: the encode pass
 * also takes the global min/max, so the encoded planes are read once (for the
 * 8-bit pass) instead of twice.  d_enc and d_out receive exactly what
 * ohx_encode per channel followed by ohx_normalize_u8 writes.  d_field, d_enc,
 * d_out: host arrays of n_channels device pointers. */
int ohx_encode_normalize_u8(ohx_ctx* ctx, const double* const* d_field, int n_channels,
                            int64_t n_px, uint32_t encode_flag, double* const* d_enc,
                            uint8_t* const* d_out);

/* Shape of the last ohx_pc_diffract on ctx (for throughput accounting). */
typedef struct ohx_launch_stats {
    int64_t pairs_nominal;   /* n_points * pn_x * pn_y                        */
    int64_t workgroups;      /* workgroups of the diffraction kernel (matrix-
                                core Fresnel: of its unmasked instance, which
                                runs when every box covers the hologram)      */
    int32_t kernel_kind;     /* 0 FP32 direct, 1 FP64 direct, 2 MFMA GEMM, 3 WRP */
    int32_t gemm_points;     /* RS with an MFMA mode: points that ran on the
                                matrix cores, on some or all of their tiles
                                (the others ran the FP32 direct sum); 0
                                otherwise                                    */
    int64_t pairs_in_box;    /* point-pixel pairs inside the points' anti-
                                aliasing boxes (ophGen.cpp:384-407,467-490):
                                the pairs the kernels evaluate; the RS range
                                test may drop a few more.  Read back from the
                                device when this function is called, which
                                synchronises ctx's stream.                   */
} ohx_launch_stats;
int ohx_last_launch_stats(ohx_ctx* ctx, ohx_launch_stats* out);

/* ---- serving: continuous batching of hologram jobs ------------------------
 * No reference counterpart (the reference generates one hologram per
 * blocking call).  A server owns a device stream and a scheduler thread.
 * Jobs (one channel each) are admitted in arrival order while they fit the
 * HBM budget and a free slot; every admitted job keeps its point constants
 * and field resident.  Each scheduler iteration advances every active job
 * by one slice of its points: one persistent multi-job kernel (replayed
 * from a captured hipGraph) covers all direct-sum jobs, matrix-core Fresnel
 * jobs run their GEMM slice on the same stream.  Slice boundaries depend only
 * on the job itself, so a job's result is bit-identical alone or in a mix.
 * Finished jobs are phase-encoded, normalised to 8 bits and copied out. */
typedef struct ohx_server ohx_server;

typedef struct ohx_server_opts {
    int32_t device;
    int32_t max_active;        /* jobs in flight (0: 32)                       */
    int32_t use_graph;         /* 1: replay a captured hipGraph per iteration  */
    int32_t reserved;
    uint64_t mem_budget;       /* HBM bytes for jobs (0: free HBM - 4 GiB)     */
    double slice_pairs;        /* point-pixel pairs per job per iteration (0: 7.5e9) */
    double slice_pairs_mfma;   /* same for matrix-core jobs (0: 7.5e10; RS on
                                  matrix cores takes half of it)              */
} ohx_server_opts;

typedef struct ohx_job {
    const void* vertices;      /* host, n_points 56-byte Vertex records (copied at submit) */
    int64_t n_points;
    ohx_pc_params params;      /* geometry, wavelength, channel               */
    uint32_t diff_flag;
    uint32_t mode;
    double* out_field;         /* host, 2*pn_x*pn_y doubles, or NULL           */
    uint8_t* out_u8;           /* host, pn_x*pn_y phase bytes, or NULL         */
} ohx_job;

typedef struct ohx_job_stats {
    int32_t status;            /* OHX_OK, OHX_ERR_CANCELLED, device errors     */
    int32_t iterations;        /* scheduler iterations the job took part in    */
    double submit_ms, admit_ms, start_ms, end_ms;   /* server clock            */
    uint64_t device_bytes;     /* HBM reserved for the job                     */
} ohx_job_stats;

int ohx_server_create(const ohx_server_opts* opts, ohx_server** out);
/* OHX_ERR_ADMISSION when the job can never fit the memory budget */
int ohx_server_submit(ohx_server* srv, const ohx_job* job, uint64_t* job_id);
/* blocks until the job finished (outputs written) or timeout_s elapsed
 * (timeout_s < 0: no limit); a finished job is forgotten after its wait */
int ohx_server_wait(ohx_server* srv, uint64_t job_id, double timeout_s, ohx_job_stats* stats);
int ohx_server_cancel(ohx_server* srv, uint64_t job_id);
/* server-wide counters: jobs done, iterations, HBM reserved now / at peak */
typedef struct ohx_server_counters {
    uint64_t jobs_done, iterations, graph_replays;
    uint64_t reserved_bytes, peak_reserved_bytes, budget_bytes;
} ohx_server_counters;
int ohx_server_counters_get(ohx_server* srv, ohx_server_counters* out);
int ohx_server_destroy(ohx_server* srv);

#ifdef __cplusplus
}
#endif
#endif /* OPENHOLO_AMD_H */
