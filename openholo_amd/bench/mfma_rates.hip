// gfx950 MFMA builtins: checks the operand/result lane layouts the matrix-core
// Fresnel kernel relies on (random small-integer matrices, exact in FP32) and
// measures dense 32x32x16 / 16x16x32 throughput.
// Build: make -C openholo_amd bench   Run: openholo_amd/bin/mfma_rates
#include <hip/hip_runtime.h>
#include <cstdio>
#include <cstdlib>
#include <vector>
#define CK(x) do { hipError_t e = (x); if (e != hipSuccess) { printf("err %s line %d\n", hipGetErrorString(e), __LINE__); return 1; } } while (0)
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

// Hypothesis 16x16x32: A lane l: row l%16, k = 8*(l/16)+e ; B lane l: col l%16, k = 8*(l/16)+e ; D lane l: col l%16, row 4*(l/16)+v
// Hypothesis 32x32x16: A lane l: row l%32, k = 8*(l/32)+e ; B: col l%32, k = 8*(l/32)+e ; D: col l%32, row 8*(v/4) + 4*(l/32) + v%4
template <int KIND>
__global__ void probe(const float* A, const float* B, float* D) {
  const int l = threadIdx.x;
  if (KIND == 0 || KIND == 1) {   // 16x16x32 : A[16][32], B[32][16], D[16][16]
    bf16x8 a, b; f16x8 ah, bh;
    for (int e = 0; e < 8; ++e) {
      const int k = 8 * (l / 16) + e;
      a[e] = (__bf16)A[(l % 16) * 32 + k]; b[e] = (__bf16)B[k * 16 + (l % 16)];
      ah[e] = (_Float16)A[(l % 16) * 32 + k]; bh[e] = (_Float16)B[k * 16 + (l % 16)];
    }
    f32x4 c = {0, 0, 0, 0};
    if (KIND == 0) c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
    else c = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, bh, c, 0, 0, 0);
    for (int v = 0; v < 4; ++v) D[(4 * (l / 16) + v) * 16 + (l % 16)] = c[v];
  } else {                        // 32x32x16 : A[32][16], B[16][32], D[32][32]
    bf16x8 a, b; f16x8 ah, bh;
    for (int e = 0; e < 8; ++e) {
      const int k = 8 * (l / 32) + e;
      a[e] = (__bf16)A[(l % 32) * 16 + k]; b[e] = (__bf16)B[k * 32 + (l % 32)];
      ah[e] = (_Float16)A[(l % 32) * 16 + k]; bh[e] = (_Float16)B[k * 32 + (l % 32)];
    }
    f32x16 c = {};
    if (KIND == 2) c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
    else c = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah, bh, c, 0, 0, 0);
    for (int v = 0; v < 16; ++v) D[(8 * (v / 4) + 4 * (l / 32) + (v % 4)) * 32 + (l % 32)] = c[v];
  }
}
// throughput: 8 independent accumulators per wave
template <int KIND>
__global__ __launch_bounds__(256) void thr(float* out, int iters) {
  bf16x8 a, b; f16x8 ah, bh;
  for (int e = 0; e < 8; ++e) { a[e] = (__bf16)(0.001f * (threadIdx.x + e)); b[e] = a[e]; ah[e] = (_Float16)(0.001f * e); bh[e] = ah[e]; }
  f32x4 c[8] = {}; f32x16 d[4] = {};
  for (int it = 0; it < iters; ++it) {
    if (KIND == 0) { for (int i = 0; i < 8; ++i) c[i] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c[i], 0, 0, 0); }
    else if (KIND == 1) { for (int i = 0; i < 8; ++i) c[i] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, bh, c[i], 0, 0, 0); }
    else if (KIND == 2) { for (int i = 0; i < 4; ++i) d[i] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, d[i], 0, 0, 0); }
  }
  float s = 0; for (int i = 0; i < 8; ++i) s += c[i][0] + c[i][3]; for (int i = 0; i < 4; ++i) s += d[i][0] + d[i][15];
  out[blockIdx.x * 256 + threadIdx.x] = s;
}
int main() {
  float *dA, *dB, *dD; CK(hipMalloc(&dA, 4096)); CK(hipMalloc(&dB, 4096)); CK(hipMalloc(&dD, 4096 + 4096));
  const char* nm[] = {"16x16x32 bf16", "16x16x32 f16", "32x32x16 bf16", "32x32x16 f16"};
  srand(1);
  for (int kind = 0; kind < 4; ++kind) {
    const int M = kind < 2 ? 16 : 32, N = M, K = kind < 2 ? 32 : 16;
    std::vector<float> A(M * K), B(K * N), D(M * N), R(M * N, 0.f);
    for (auto& x : A) x = (float)(rand() % 9 - 4);
    for (auto& x : B) x = (float)(rand() % 9 - 4);
    for (int i = 0; i < M; ++i) for (int j = 0; j < N; ++j) for (int k = 0; k < K; ++k) R[i * N + j] += A[i * K + k] * B[k * N + j];
    CK(hipMemcpy(dA, A.data(), A.size() * 4, hipMemcpyHostToDevice));
    CK(hipMemcpy(dB, B.data(), B.size() * 4, hipMemcpyHostToDevice));
    CK(hipMemset(dD, 0, M * N * 4));
    switch (kind) {
      case 0: hipLaunchKernelGGL(probe<0>, dim3(1), dim3(64), 0, 0, dA, dB, dD); break;
      case 1: hipLaunchKernelGGL(probe<1>, dim3(1), dim3(64), 0, 0, dA, dB, dD); break;
      case 2: hipLaunchKernelGGL(probe<2>, dim3(1), dim3(64), 0, 0, dA, dB, dD); break;
      case 3: hipLaunchKernelGGL(probe<3>, dim3(1), dim3(64), 0, 0, dA, dB, dD); break;
    }
    CK(hipGetLastError()); CK(hipDeviceSynchronize());
    CK(hipMemcpy(D.data(), dD, M * N * 4, hipMemcpyDeviceToHost));
    int bad = 0; for (int i = 0; i < M * N; ++i) bad += D[i] != R[i];
    printf("layout %-14s : %s (%d mismatches)\n", nm[kind], bad ? "MISMATCH" : "ok", bad);
  }
  float* out; CK(hipMalloc(&out, 256 * 8 * 256 * 4));
  hipEvent_t e0, e1; CK(hipEventCreate(&e0)); CK(hipEventCreate(&e1));
  const int blocks = 256 * 8, iters = 2048;
  for (int kind = 0; kind < 3; ++kind) for (int rep = 0; rep < 2; ++rep) {
    CK(hipEventRecord(e0));
    if (kind == 0) hipLaunchKernelGGL(thr<0>, dim3(blocks), dim3(256), 0, 0, out, iters);
    if (kind == 1) hipLaunchKernelGGL(thr<1>, dim3(blocks), dim3(256), 0, 0, out, iters);
    if (kind == 2) hipLaunchKernelGGL(thr<2>, dim3(blocks), dim3(256), 0, 0, out, iters);
    CK(hipGetLastError()); CK(hipEventRecord(e1)); CK(hipEventSynchronize(e1));
    float ms; CK(hipEventElapsedTime(&ms, e0, e1));
    const double flops = (double)blocks * 4 * iters * (kind < 2 ? 8.0 * 16 * 16 * 32 * 2 : 4.0 * 32 * 32 * 16 * 2);
    if (rep) printf("mfma %-14s: %.3f ms  %.1f TFLOP/s\n", nm[kind], ms, flops / ms / 1e9);
  }
  return 0;
}
