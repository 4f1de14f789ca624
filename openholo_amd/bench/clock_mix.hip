// Shader clock and issue rate under the instruction mixes of the matrix-core
// Fresnel kernel: MFMA 32x32x16 bf16 alone, FP32 FMA alone, v_sin + v_cos
// alone, and MFMAs with FMAs / sin+cos issued between them from the same
// wave.  Two 4-wave workgroups per CU (80 KB LDS each), i.e. two waves per
// SIMD as in the GEMM kernel.  Each workgroup's thread 0 reads clock64()
// (shader clock) and wall_clock64() (constant rate) around its work; the
// effective clock is d(clock64) / d(wall) * wall rate, and "cycles per slot"
// is the SIMD time per (MFMA + its VALU share) in shader cycles.
// Build: make -C openholo_amd bench   Run: openholo_amd/bin/clock_mix
#include <hip/hip_runtime.h>
#include <cstdio>
#include <vector>
#define CK(x) do { hipError_t e = (x); if (e != hipSuccess) { printf("err %s line %d\n", hipGetErrorString(e), __LINE__); return 1; } } while (0)
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

// NM MFMAs (0/1), NV FMAs (FP64 when D64) and NS sin+cos pairs per slot; 16
// slots per iteration
template <int NM, int NV, int NS, bool D64 = false>
__global__ __launch_bounds__(256) void mixk(float* out, long long* tm, int iters)
{
    __shared__ float pad[20 * 1024];
    const long long c0 = clock64(), w0 = wall_clock64();
    bf16x8 a, b;
    for (int e = 0; e < 8; ++e) {
        a[e] = (__bf16)(0.001f * (threadIdx.x + e));
        b[e] = a[e];
    }
    f32x16 d[4] = {};
    float x[8];
    double y[8];
    for (int i = 0; i < 8; ++i) {
        x[i] = 0.001f * (threadIdx.x + i);
        y[i] = 0.001 * (threadIdx.x + i);
    }
    const float c1 = out[0] + 0.9999f, c2 = out[1] + 1e-4f;
    pad[threadIdx.x] = c1;
    __syncthreads();
    for (int it = 0; it < iters; ++it) {
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            if constexpr (NM > 0) d[i & 3] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, d[i & 3], 0, 0, 0);
#pragma unroll
            for (int j = 0; j < NV; ++j) {
                if constexpr (D64) y[j & 7] = fma(y[j & 7], (double)c1, (double)c2);
                else x[j & 7] = fmaf(x[j & 7], c1, c2);
            }
#pragma unroll
            for (int j = 0; j < NS; ++j)
                x[j & 7] = __builtin_amdgcn_sinf(x[j & 7]) + __builtin_amdgcn_cosf(x[(j + 1) & 7]);
            __builtin_amdgcn_sched_barrier(0);
        }
    }
    float s = pad[(threadIdx.x + 1) & 255];
    for (int i = 0; i < 4; ++i) s += d[i][0] + d[i][15];
    for (int i = 0; i < 8; ++i) s += x[i] + (float)y[i];
    out[2 + blockIdx.x * 256 + threadIdx.x] = s;
    if (threadIdx.x == 0) {
        tm[4 * blockIdx.x] = c0;
        tm[4 * blockIdx.x + 1] = clock64();
        tm[4 * blockIdx.x + 2] = w0;
        tm[4 * blockIdx.x + 3] = wall_clock64();
    }
}

template <int NM, int NV, int NS, bool D64 = false>
int run(const char* name, float* out, long long* tm, int wall_khz)
{
    const int blocks = 256 * 8, iters = 1024;
    hipEvent_t e0, e1;
    CK(hipEventCreate(&e0));
    CK(hipEventCreate(&e1));
    float ms = 0;
    for (int rep = 0; rep < 2; ++rep) {
        CK(hipEventRecord(e0));
        hipLaunchKernelGGL((mixk<NM, NV, NS, D64>), dim3(blocks), dim3(256), 0, 0, out, tm, iters);
        CK(hipGetLastError());
        CK(hipEventRecord(e1));
        CK(hipEventSynchronize(e1));
        CK(hipEventElapsedTime(&ms, e0, e1));
    }
    std::vector<long long> h(4 * blocks);
    CK(hipMemcpy(h.data(), tm, h.size() * 8, hipMemcpyDeviceToHost));
    double sc = 0, sw = 0;
    for (int b = 0; b < blocks; ++b) {
        sc += (double)(h[4 * b + 1] - h[4 * b]);
        sw += (double)(h[4 * b + 3] - h[4 * b + 2]);
    }
    const double mhz = sc / sw * wall_khz / 1e3;
    const double waves = blocks * 4.0, slots_per_simd = waves / 1024 * iters * 16;
    printf("%-28s %7.3f ms  clock %5.0f MHz  MFMA %5.0f TFLOP/s  %5.1f cycles per slot\n", name, ms, mhz,
           waves * iters * 16 * NM * 32768.0 / ms / 1e9, ms * 1e-3 * mhz * 1e6 / slots_per_simd);
    return 0;
}

int main()
{
    int wall_khz = 0;
    CK(hipDeviceGetAttribute(&wall_khz, hipDeviceAttributeWallClockRate, 0));
    float* out;
    long long* tm;
    CK(hipMalloc(&out, (2 + 256 * 8 * 256) * 4));
    CK(hipMemset(out, 0, 8));
    CK(hipMalloc(&tm, 4 * 256 * 8 * 8));
    run<1, 0, 0>("mfma", out, tm, wall_khz);
    run<0, 8, 0>("8 fma", out, tm, wall_khz);
    run<0, 0, 2>("2 x (sin + cos)", out, tm, wall_khz);
    run<1, 6, 0>("mfma + 6 fma", out, tm, wall_khz);
    run<1, 4, 1>("mfma + 4 fma + sin + cos", out, tm, wall_khz);
    run<1, 8, 1>("mfma + 8 fma + sin + cos", out, tm, wall_khz);
    run<0, 8, 0, true>("8 fma_f64", out, tm, wall_khz);
    run<1, 4, 0>("mfma + 4 fma", out, tm, wall_khz);
    run<1, 4, 0, true>("mfma + 4 fma_f64", out, tm, wall_khz);
    return 0;
}
