// Register-only throughput probes for the FP64 and FP32 arithmetic units of
// gfx950: what one MI355X sustains on v_mfma_f64_16x16x4f64, on v_fma_f64 and
// on v_mfma_f32_16x16x4f32 with no memory traffic in the loop.  A measurement
// tool for the roofline figures (libisdf_probe.so, loaded by bench.py
// --peak-probe); the ISDF library does not use it.
//
// Each wave runs `iters` rounds of NCHAIN independent accumulator chains, so
// the unit never waits for a dependent result; the accumulators are stored
// once at the end so that nothing is optimised away.

#include <hip/hip_runtime.h>

#include "mfma_gemm.h"

namespace {

constexpr int kChains = 8;

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

__global__ __launch_bounds__(256) void k_mfma_f64(double a0, double b0, int iters, double* out) {
  d4 acc[kChains];
  for (int c = 0; c < kChains; c++) acc[c] = (d4){0.0, 0.0, 0.0, 0.0};
  const double a = a0 + threadIdx.x * 1e-9, b = b0 - threadIdx.x * 1e-9;
  for (int it = 0; it < iters; it++) {
#pragma unroll
    for (int c = 0; c < kChains; c++) acc[c] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[c], 0, 0, 0);
  }
  double s = 0.0;
  for (int c = 0; c < kChains; c++) s += acc[c][0] + acc[c][1] + acc[c][2] + acc[c][3];
  out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

__global__ __launch_bounds__(256) void k_mfma_f32(float a0, float b0, int iters, double* out) {
  f4 acc[kChains];
  for (int c = 0; c < kChains; c++) acc[c] = (f4){0.f, 0.f, 0.f, 0.f};
  const float a = a0 + threadIdx.x * 1e-6f, b = b0 - threadIdx.x * 1e-6f;
  for (int it = 0; it < iters; it++) {
#pragma unroll
    for (int c = 0; c < kChains; c++) acc[c] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, acc[c], 0, 0, 0);
  }
  double s = 0.0;
  for (int c = 0; c < kChains; c++) s += acc[c][0] + acc[c][1] + acc[c][2] + acc[c][3];
  out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

// 16 independent FMA chains per lane (enough to cover the FP64 VALU latency).
__global__ __launch_bounds__(256) void k_fma_f64(double a0, double b0, int iters, double* out) {
  double acc[16];
  for (int c = 0; c < 16; c++) acc[c] = c * 1e-3;
  const double a = a0 + threadIdx.x * 1e-12, b = b0;
  for (int it = 0; it < iters; it++) {
#pragma unroll
    for (int c = 0; c < 16; c++) acc[c] = __builtin_fma(acc[c], a, b);
  }
  double s = 0.0;
  for (int c = 0; c < 16; c++) s += acc[c];
  out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

// The solve tile's inner step with no memory traffic: 64 accumulators per
// lane, two A values and two B registers, four groups of 16 v_fmac_f64 with B
// broadcast by DPP row_newbcast (isdf_hip::dpp_fma16).
__global__ __launch_bounds__(256) void k_dpp_f64(double a0, double b0, int iters, double* out) {
  double acc[64];
  for (int c = 0; c < 64; c++) acc[c] = c * 1e-3;
  const double a[2] = {a0 + threadIdx.x * 1e-12, a0 - threadIdx.x * 1e-12};
  const double b[2] = {b0 + (threadIdx.x & 63) * 1e-9, b0 - (threadIdx.x & 63) * 1e-9};
  for (int it = 0; it < iters; it++) {
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
      for (int g = 0; g < 2; g++) isdf_hip::dpp_fma16(acc + (i * 2 + g) * 16, b[g], a[i]);
  }
  double s = 0.0;
  for (int c = 0; c < 64; c++) s += acc[c];
  out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

// One dependent FP64 add chain per lane, terms from registers: the latency of
// v_add_f64 back to back (the bound of any strictly sequential FP64 sum).
__global__ __launch_bounds__(64) void k_add_chain(double a0, int iters, double* out) {
  double acc = threadIdx.x * 1e-3;
  const double a = a0 + threadIdx.x * 1e-12, b = a0 * 0.5;
  for (int it = 0; it < iters; it++) {
#pragma unroll
    for (int u = 0; u < 16; u++) acc = acc + ((u & 1) ? a : b);
  }
  out[blockIdx.x * blockDim.x + threadIdx.x] = acc;
}

// Even waves of a workgroup run the MFMA chains of k_mfma_f64, odd waves the
// FMA chains of k_fma_f64: whether the two FP64 pipes add up when both are busy.
__global__ __launch_bounds__(256) void k_mix_f64(double a0, double b0, int iters_mfma,
                                                  int iters_fma, double* out) {
  double s = 0.0;
  if (((threadIdx.x >> 6) & 1) == 0) {
    d4 acc[kChains];
    for (int c = 0; c < kChains; c++) acc[c] = (d4){0.0, 0.0, 0.0, 0.0};
    const double a = a0 + threadIdx.x * 1e-9, b = b0 - threadIdx.x * 1e-9;
    for (int it = 0; it < iters_mfma; it++) {
#pragma unroll
      for (int c = 0; c < kChains; c++) acc[c] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[c], 0, 0, 0);
    }
    for (int c = 0; c < kChains; c++) s += acc[c][0] + acc[c][1] + acc[c][2] + acc[c][3];
  } else {
    double acc[16];
    for (int c = 0; c < 16; c++) acc[c] = c * 1e-3;
    const double a = 0.999 + threadIdx.x * 1e-12, b = 1e-3;
    for (int it = 0; it < iters_fma; it++) {
#pragma unroll
      for (int c = 0; c < 16; c++) acc[c] = __builtin_fma(acc[c], a, b);
    }
    for (int c = 0; c < 16; c++) s += acc[c];
  }
  out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

}  // namespace

extern "C" {

// which: 0 v_mfma_f64_16x16x4f64, 1 v_fma_f64, 2 v_mfma_f32_16x16x4f32,
// 3 dependent v_add_f64 chain (rate of one lane, in 1e12 adds/s), 4 half the
// waves on v_mfma_f64_16x16x4f64 and half on v_fma_f64 (combined TFLOP/s),
// 5 v_fmac_f64 with B by DPP row_newbcast in the solve tile's register shape.
// Launches `blocks` workgroups of 256 threads (4 waves), each running `iters`
// rounds, twice (the first run warms the clocks up), and returns the second
// run's rate in TFLOP/s through *tflops and its time through *ms.
// Returns 0, or a hipError_t value.
int isdf_probe_peak(int which, int blocks, int iters, double* tflops, double* ms) {
  double* out = nullptr;
  hipError_t e = hipMalloc(&out, sizeof(double) * (size_t)blocks * 256);
  if (e != hipSuccess) return (int)e;
  hipEvent_t t0, t1;
  (void)hipEventCreate(&t0);
  (void)hipEventCreate(&t1);
  double flops_per_thread_iter = 0.0;
  for (int rep = 0; rep < 2; rep++) {
    (void)hipEventRecord(t0, nullptr);
    if (which == 0) {
      k_mfma_f64<<<blocks, 256>>>(1.0, 0.5, iters, out);
      // 16x16x4 f64: 2 * 16 * 16 * 4 flops per wave instruction = 32 per lane.
      flops_per_thread_iter = 32.0 * kChains;
    } else if (which == 1) {
      k_fma_f64<<<blocks, 256>>>(0.999, 1e-3, iters, out);
      flops_per_thread_iter = 2.0 * 16;
    } else if (which == 2) {
      k_mfma_f32<<<blocks, 256>>>(1.0f, 0.5f, iters, out);
      flops_per_thread_iter = 32.0 * kChains;
    } else if (which == 4) {
      // Half MFMA waves, half FMA waves; iters is the MFMA loop count and the
      // FMA loop runs 4x as many rounds (about the same time on each pipe).
      k_mix_f64<<<blocks, 256>>>(1.0, 0.5, iters, 4 * iters, out);
      flops_per_thread_iter = 0.5 * (32.0 * kChains) + 0.5 * (2.0 * 16 * 4);
    } else if (which == 5) {
      k_dpp_f64<<<blocks, 256>>>(0.999, 1e-3, iters, out);
      flops_per_thread_iter = 2.0 * 64;
    } else {
      // One wave per workgroup, `blocks` of them; *tflops then reports the
      // dependent adds per second of one lane (1e12 units), *ms the time.
      k_add_chain<<<blocks, 64>>>(1.0, iters, out);
      flops_per_thread_iter = 16.0;
    }
    (void)hipEventRecord(t1, nullptr);
  }
  e = hipEventSynchronize(t1);
  float t = 0.f;
  (void)hipEventElapsedTime(&t, t0, t1);
  if (e == hipSuccess) e = hipGetLastError();
  (void)hipEventDestroy(t0);
  (void)hipEventDestroy(t1);
  (void)hipFree(out);
  if (e != hipSuccess) return (int)e;
  *ms = t;
  *tflops = which == 3 ? flops_per_thread_iter * (double)iters / (t * 1e-3) / 1e12
                      : flops_per_thread_iter * (double)iters * blocks * 256 / (t * 1e-3) / 1e12;
  return 0;
}

}  // extern "C"
