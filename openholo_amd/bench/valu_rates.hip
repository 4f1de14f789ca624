// Issue-rate microbenchmark for the VALU roofline of the direct kernels:
// FP32 FMA, v_sin, v_fract, v_rsq and mixes, plus a 1 GiB copy kernel whose
// FETCH_SIZE / WRITE_SIZE readings calibrate the HBM traffic counters.
// Build: make -C openholo_amd bench   Run: openholo_amd/bin/valu_rates
#include <hip/hip_runtime.h>
#include <cstdio>
#include <cstring>
#include <vector>
#define CK(x) do { hipError_t e = (x); if (e != hipSuccess) { printf("err %s line %d\n", hipGetErrorString(e), __LINE__); return 1; } } while (0)
constexpr int CH = 16;
template <int KIND>
__global__ __launch_bounds__(256) void k(float* out, int iters, float b, float c) {
  float a[CH], r[CH];
  for (int i = 0; i < CH; ++i) { a[i] = threadIdx.x * 1e-3f + i * 1e-2f; r[i] = 0.f; }
  for (int it = 0; it < iters; ++it) {
#pragma unroll
    for (int i = 0; i < CH; ++i) {
      if (KIND == 0) a[i] = fmaf(a[i], b, c);                                  // FMA
      else if (KIND == 1) a[i] = __builtin_amdgcn_sinf(a[i]);                   // v_sin
      else if (KIND == 2) a[i] = __builtin_amdgcn_fractf(a[i] + c);             // add+fract
      else if (KIND == 3) { a[i] = __builtin_amdgcn_sinf(a[i]) * b + c; }       // sin + fma
      else if (KIND == 4) { float s = __builtin_amdgcn_sinf(a[i]), t = __builtin_amdgcn_cosf(a[i]); a[i] = fmaf(s, b, t); } // sin+cos+fma
      else if (KIND == 5) { a[i] = fmaf(a[i], b, c); a[i] = fmaf(a[i], b, c); a[i] = fmaf(a[i], b, c); a[i] = __builtin_amdgcn_sinf(a[i]); } // 3 fma + sin
      else if (KIND == 6) a[i] = __builtin_amdgcn_rsqf(a[i] + c);               // add + rsq
      else if (KIND == 7) {  // one Fresnel pair of the direct kernel: add, sin, cos, 2 fma
        const float f = a[i] + c;
        r[i] = fmaf(b, __builtin_amdgcn_sinf(f), r[i]);
        a[i] = fmaf(b, __builtin_amdgcn_cosf(f), a[i]);
      } else if (KIND == 8) {  // one RS pair: add, 2 fma, sin, cos, fma, 2 fma
        const float N = a[i] + c;
        const float f = fmaf(fmaf(-1e-3f, N, b), N, c);
        const float amp = fmaf(-1e-4f, N, b);
        r[i] = fmaf(amp, __builtin_amdgcn_sinf(f), r[i]);
        a[i] = fmaf(amp, __builtin_amdgcn_cosf(f), a[i]);
      }
    }
  }
  float s = 0; for (int i = 0; i < CH; ++i) s += a[i] + r[i];
  out[blockIdx.x * 256 + threadIdx.x] = s;
}
// The separable pair loops of the direct kernels, per lane and point over 8
// rows: KIND 0 Fresnel (2 FMA + v_sin + v_cos per point, then 4 FMA per
// pair), KIND 1 RS (5 VALU + v_sin + v_cos + 2 MUL per point, then MUL +
// 6 FMA per pair).  The row factors sit in registers here (LDS in the kernel).
template <int KIND>
__global__ __launch_bounds__(256) void ksep(float* out, int iters, float b, float c) {
  float pr[8], pim[8], br[8], bi[8], ny[8];
  for (int k = 0; k < 8; ++k) {
    pr[k] = pim[k] = 0.f;
    br[k] = 0.5f + 0.01f * k; bi[k] = 0.25f - 0.01f * k; ny[k] = 1e-3f * k;
  }
  float x = threadIdx.x * 1e-4f;
  for (int it = 0; it < iters; ++it) {
    float sa, ca, kx = 0.f;
    if (KIND == 0) {
      const float f = fmaf(fmaf(x, b, c), x, c);
      sa = __builtin_amdgcn_sinf(f); ca = __builtin_amdgcn_cosf(f);
    } else {   // RS
      const float fa = fmaf(fmaf(c, x, b), x, c);
      const float am = fmaf(c, x, b);
      kx = b * x;
      sa = am * __builtin_amdgcn_sinf(fa); ca = am * __builtin_amdgcn_cosf(fa);
    }
    if (KIND == 2) {   // RS with the cross term folded into the column factor: 6 FMA per pair
      const float cak = ca * kx, sak = sa * kx;
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        const float P = fmaf(cak, ny[k], sa), Q = fmaf(-sak, ny[k], ca);
        pr[k] = fmaf(P, br[k], fmaf(Q, bi[k], pr[k]));
        pim[k] = fmaf(Q, br[k], fmaf(-P, bi[k], pim[k]));
      }
    } else {
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        float r = br[k], i = bi[k];
        if (KIND == 1) { const float e = kx * ny[k]; r = fmaf(-e, bi[k], br[k]); i = fmaf(e, br[k], bi[k]); }
        pr[k] = fmaf(sa, r, fmaf(ca, i, pr[k]));
        pim[k] = fmaf(ca, r, fmaf(-sa, i, pim[k]));
      }
    }
    x = fmaf(x, b, c);
  }
  float s = 0; for (int k = 0; k < 8; ++k) s += pr[k] + pim[k];
  out[blockIdx.x * 256 + threadIdx.x] = s;
}
// v_sin / v_cos on an unreduced revolutions argument against the same after
// v_fract: counts results that differ and the largest difference, for hashed
// arguments uniform in [-lim, lim].
__global__ void ksinrange(unsigned* cnt, unsigned* maxd, float lim) {
  const unsigned i = blockIdx.x * 256 + threadIdx.x;
  unsigned h = i * 2654435761u; h ^= h >> 15; h *= 2246822519u; h ^= h >> 13;
  const float x = ((float)(h >> 8) * (1.0f / 16777216.0f) * 2.0f - 1.0f) * lim;
  const float f = __builtin_amdgcn_fractf(x);
  const float s1 = __builtin_amdgcn_sinf(x), s2 = __builtin_amdgcn_sinf(f);
  const float c1 = __builtin_amdgcn_cosf(x), c2 = __builtin_amdgcn_cosf(f);
  const float d = fmaxf(fabsf(s1 - s2), fabsf(c1 - c2));
  if (s1 != s2 || c1 != c2) atomicAdd(cnt, 1u);
  if (d > 0.f) atomicMax(maxd, __float_as_uint(d));
}
// FP64 issue rate: KIND 0 fma_f64, 1 add_f64 + fract_f64, 2 cvt_f32_f64 + add_f64
template <int KIND>
__global__ __launch_bounds__(256) void k64(float* out, int iters, double b, double c) {
  double a[CH / 2];
  float f[CH / 2];
  for (int i = 0; i < CH / 2; ++i) { a[i] = threadIdx.x * 1e-3 + i * 1e-2; f[i] = 0.f; }
  for (int it = 0; it < iters; ++it) {
#pragma unroll
    for (int i = 0; i < CH / 2; ++i) {
      if (KIND == 0) a[i] = fma(a[i], b, c);
      else if (KIND == 1) a[i] = __builtin_amdgcn_fract(a[i] + c);
      else { f[i] += (float)a[i]; a[i] += c; }
    }
  }
  float s = 0; for (int i = 0; i < CH / 2; ++i) s += (float)a[i] + f[i];
  out[blockIdx.x * 256 + threadIdx.x] = s;
}
// FP32 issue rate by instruction form: all-VGPR operands (the form of the
// direct kernels' pair loops) against the SGPR-operand FMA of k<0>.
template <int KIND>
__global__ __launch_bounds__(256) void kv(float* out, int iters, float bs) {
  float a[CH], r[CH];
  const float b = threadIdx.x * 1e-7f + bs;   // a VGPR operand
  for (int i = 0; i < CH; ++i) { a[i] = threadIdx.x * 1e-3f + i * 1e-2f; r[i] = i * 1e-3f; }
  for (int it = 0; it < iters; ++it) {
#pragma unroll
    for (int i = 0; i < CH; ++i) {
      if (KIND == 0) r[i] = fmaf(a[i], b, r[i]);                 // v_fmac_f32 (VOP2)
      else if (KIND == 1) r[i] = fmaf(-a[i], b, r[i]);           // v_fma_f32, negated operand (VOP3)
      else if (KIND == 2) r[i] = r[i] * b;                       // v_mul_f32
      else { r[i] = fmaf(a[i], b, r[i]); a[i] = fmaf(-r[i], b, a[i]); }   // both FMA forms
    }
  }
  float s = 0; for (int i = 0; i < CH; ++i) s += a[i] + r[i];
  out[blockIdx.x * 256 + threadIdx.x] = s;
}
__global__ void kcopy(const float4* __restrict__ a, float4* __restrict__ b, size_t n) {
  size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x, st = (size_t)gridDim.x * blockDim.x;
  for (; i < n; i += st) b[i] = a[i];
}
int main() {
  const int blocks = 256 * 8, iters = 4096;
  float* out; CK(hipMalloc(&out, blocks * 256 * sizeof(float)));
  hipEvent_t e0, e1; CK(hipEventCreate(&e0)); CK(hipEventCreate(&e1));
  const char* names[] = {"fma (SGPR operand)", "sin", "add+fract", "sin+fma", "sin+cos+fma", "3fma+sin", "add+rsq",
                         "fresnel-pair", "rs-pair"};
  const int instr[] = {1, 1, 2, 2, 3, 4, 2, 5, 8};
  for (int kind = 0; kind < 9; ++kind) {
    for (int rep = 0; rep < 2; ++rep) {
      CK(hipEventRecord(e0));
      switch (kind) {
        case 0: hipLaunchKernelGGL(k<0>, dim3(blocks), dim3(256), 0, 0, out, iters, 0.999f, 1e-4f); break;
        case 1: hipLaunchKernelGGL(k<1>, dim3(blocks), dim3(256), 0, 0, out, iters, 0.999f, 1e-4f); break;
        case 2: hipLaunchKernelGGL(k<2>, dim3(blocks), dim3(256), 0, 0, out, iters, 0.999f, 1e-4f); break;
        case 3: hipLaunchKernelGGL(k<3>, dim3(blocks), dim3(256), 0, 0, out, iters, 0.999f, 1e-4f); break;
        case 4: hipLaunchKernelGGL(k<4>, dim3(blocks), dim3(256), 0, 0, out, iters, 0.999f, 1e-4f); break;
        case 5: hipLaunchKernelGGL(k<5>, dim3(blocks), dim3(256), 0, 0, out, iters, 0.999f, 1e-4f); break;
        case 6: hipLaunchKernelGGL(k<6>, dim3(blocks), dim3(256), 0, 0, out, iters, 0.999f, 1e-4f); break;
        // pair kinds: amplitude 1e-3 keeps the unreduced phases inside v_sin's +-256 domain
        case 7: hipLaunchKernelGGL(k<7>, dim3(blocks), dim3(256), 0, 0, out, iters, 1e-3f, 1e-4f); break;
        case 8: hipLaunchKernelGGL(k<8>, dim3(blocks), dim3(256), 0, 0, out, iters, 1e-3f, 1e-4f); break;
      }
      CK(hipGetLastError());
      CK(hipEventRecord(e1)); CK(hipEventSynchronize(e1));
      float ms; CK(hipEventElapsedTime(&ms, e0, e1));
      double lanes_ops = (double)blocks * 256 * iters * CH;   // per "element op group"
      if (rep) printf("%-12s %8.3f ms  %8.2f T elem-iter/s  %8.2f T VALU-instr/s (lane)\n", names[kind], ms,
                      lanes_ops / ms / 1e9, lanes_ops * instr[kind] / ms / 1e9);
    }
  }
  const char* nv[] = {"fmac VGPR (VOP2)", "fma neg VGPR (VOP3)", "mul VGPR", "fmac + fma neg"};
  for (int kind = 0; kind < 4; ++kind) {
    for (int rep = 0; rep < 2; ++rep) {
      CK(hipEventRecord(e0));
      if (kind == 0) hipLaunchKernelGGL(kv<0>, dim3(blocks), dim3(256), 0, 0, out, 4 * iters, 0.999f);
      if (kind == 1) hipLaunchKernelGGL(kv<1>, dim3(blocks), dim3(256), 0, 0, out, 4 * iters, 0.999f);
      if (kind == 2) hipLaunchKernelGGL(kv<2>, dim3(blocks), dim3(256), 0, 0, out, 4 * iters, 0.999f);
      if (kind == 3) hipLaunchKernelGGL(kv<3>, dim3(blocks), dim3(256), 0, 0, out, 4 * iters, 0.999f);
      CK(hipGetLastError());
      CK(hipEventRecord(e1)); CK(hipEventSynchronize(e1));
      float ms; CK(hipEventElapsedTime(&ms, e0, e1));
      const double ops = (double)blocks * 256 * 4 * iters * CH * (kind == 3 ? 2 : 1);
      if (rep) printf("%-28s %8.3f ms  %8.2f T VALU-instr/s (lane)\n", nv[kind], ms, ops / ms / 1e9);
    }
  }
  const char* nsep[] = {"fresnel-sep (8 rows)", "rs-sep, mul + 6 fma (8 rows)", "rs-sep, 6 fma (8 rows)"};
  for (int kind = 0; kind < 3; ++kind) {
    for (int rep = 0; rep < 2; ++rep) {
      CK(hipEventRecord(e0));
      if (kind == 0) hipLaunchKernelGGL(ksep<0>, dim3(blocks), dim3(256), 0, 0, out, iters, 0.999f, 1e-4f);
      else if (kind == 1) hipLaunchKernelGGL(ksep<1>, dim3(blocks), dim3(256), 0, 0, out, iters, 0.999f, 1e-4f);
      else hipLaunchKernelGGL(ksep<2>, dim3(blocks), dim3(256), 0, 0, out, iters, 0.999f, 1e-4f);
      CK(hipGetLastError());
      CK(hipEventRecord(e1)); CK(hipEventSynchronize(e1));
      float ms; CK(hipEventElapsedTime(&ms, e0, e1));
      const double pairs = (double)blocks * 256 * iters * 8;
      if (rep) printf("%-28s %8.3f ms  %8.2f T pairs/s\n", nsep[kind], ms, pairs / ms / 1e9);
    }
  }
  const char* n64[] = {"fma_f64", "add_f64+fract_f64", "cvt_f32_f64+add_f32+add_f64"};
  const int i64[] = {1, 2, 3};
  for (int kind = 0; kind < 3; ++kind) {
    for (int rep = 0; rep < 2; ++rep) {
      CK(hipEventRecord(e0));
      if (kind == 0) hipLaunchKernelGGL(k64<0>, dim3(blocks), dim3(256), 0, 0, out, iters, 0.999, 1e-4);
      if (kind == 1) hipLaunchKernelGGL(k64<1>, dim3(blocks), dim3(256), 0, 0, out, iters, 0.999, 1e-4);
      if (kind == 2) hipLaunchKernelGGL(k64<2>, dim3(blocks), dim3(256), 0, 0, out, iters, 0.999, 1e-4);
      CK(hipGetLastError());
      CK(hipEventRecord(e1)); CK(hipEventSynchronize(e1));
      float ms; CK(hipEventElapsedTime(&ms, e0, e1));
      const double el = (double)blocks * 256 * iters * (CH / 2);
      if (rep) printf("%-28s %8.3f ms  %8.2f T elem-iter/s  %8.2f T VALU-instr/s (lane)\n", n64[kind], ms, el / ms / 1e9,
                      el * i64[kind] / ms / 1e9);
    }
  }
  // copy kernel of known size for counter calibration: 1 GiB read + 1 GiB write
  size_t n = (1ull << 30) / 16;
  float4 *a, *b; CK(hipMalloc(&a, n * 16)); CK(hipMalloc(&b, n * 16));
  CK(hipMemset(a, 0, n * 16));
  for (int rep = 0; rep < 3; ++rep) {
    CK(hipEventRecord(e0));
    hipLaunchKernelGGL(kcopy, dim3(256 * 64), dim3(256), 0, 0, (const float4*)a, b, n);
    CK(hipGetLastError());
    CK(hipEventRecord(e1)); CK(hipEventSynchronize(e1));
    float ms; CK(hipEventElapsedTime(&ms, e0, e1));
    printf("copy 1 GiB -> 1 GiB: %.3f ms  %.1f GB/s (read+write)\n", ms, 2.0 * n * 16 / ms / 1e6);
  }
  hipDeviceProp_t p; CK(hipGetDeviceProperties(&p, 0));
  printf("device %s CUs %d clock %d kHz\n", p.gcnArchName, p.multiProcessorCount, p.clockRate);
  // unreduced sin/cos arguments (the direct and matrix-core kernels rely on |x| <= 256)
  unsigned* sr; CK(hipMalloc(&sr, 8));
  const float lims[] = {1.0f, 24.0f, 64.0f, 256.0f};
  for (float lim : lims) {
    CK(hipMemset(sr, 0, 8));
    hipLaunchKernelGGL(ksinrange, dim3(65536), dim3(256), 0, 0, sr, sr + 1, lim);
    CK(hipGetLastError()); CK(hipDeviceSynchronize());
    unsigned h[2]; CK(hipMemcpy(h, sr, 8, hipMemcpyDeviceToHost));
    float md; memcpy(&md, &h[1], 4);
    printf("sin/cos(x) vs sin/cos(fract(x)), |x| <= %5.0f: %u of %u differ, max |diff| %.3e\n", lim, h[0],
           65536u * 256u, md);
  }
  return 0;
}
