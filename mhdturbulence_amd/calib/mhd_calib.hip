// mhd_calib.hip -- machine calibration for the step's roofline figures (executable mhd_calib).
//
// 1. HBM streams: copy, read-only sum, write-only fill and two-read-one-write over 1 GiB
//    arrays with 16-byte accesses (plain and nontemporal), the rates a streaming kernel can
//    reach on this GPU beside the 8 TB/s nominal peak.
// 2. FP64 VALU issue: each lane runs 8 independent FMA chains (or 2 reciprocal + 6 FMA
//    chains) with one block per CU (96 KiB of LDS keeps a second block off) and 1-4
//    wavefronts per SIMD.  Reported per SIMD: nanoseconds per wave64 instruction from the
//    launch's wall time, and the shader clock from s_memtime against s_memrealtime (100 MHz).
//
//     mhdturbulence_amd/mhd_calib
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

namespace {

typedef double vd2 __attribute__((ext_vector_type(2)));

#define CHECK(x)                                                                  \
  do {                                                                            \
    hipError_t e_ = (x);                                                          \
    if (e_ != hipSuccess) {                                                       \
      std::fprintf(stderr, "%s failed: %s\n", #x, hipGetErrorString(e_));         \
      std::exit(1);                                                               \
    }                                                                             \
  } while (0)

// ------------------------------------------------------------------ HBM streams
constexpr size_t NDBL = size_t(1) << 27;  // doubles per array (1 GiB)
constexpr int SGRID = 8192;

template <bool NT>
__device__ __forceinline__ vd2 ld(const vd2* p) {
  if constexpr (NT) return __builtin_nontemporal_load(p);
  else return *p;
}
template <bool NT>
__device__ __forceinline__ void st(vd2* p, vd2 v) {
  if constexpr (NT) __builtin_nontemporal_store(v, p);
  else *p = v;
}

template <bool NT>
__global__ __launch_bounds__(256) void k_copy(const vd2* __restrict__ a, vd2* __restrict__ b, size_t n) {
  for (size_t i = blockIdx.x * size_t(256) + threadIdx.x; i < n; i += size_t(gridDim.x) * 256) st<NT>(b + i, ld<NT>(a + i));
}
__global__ __launch_bounds__(256) void k_read(const vd2* __restrict__ a, double* __restrict__ out, size_t n) {
  double s = 0.0;
  for (size_t i = blockIdx.x * size_t(256) + threadIdx.x; i < n; i += size_t(gridDim.x) * 256) {
    const vd2 v = a[i];
    s += v.x + v.y;
  }
  out[blockIdx.x * 256 + threadIdx.x] = s;
}
__global__ __launch_bounds__(256) void k_fill(vd2* __restrict__ b, size_t n) {
  for (size_t i = blockIdx.x * size_t(256) + threadIdx.x; i < n; i += size_t(gridDim.x) * 256) b[i] = vd2{1.0, 2.0};
}
template <bool NT>
__global__ __launch_bounds__(256) void k_2r1w(const vd2* __restrict__ a, const vd2* __restrict__ c, vd2* __restrict__ b,
                                              size_t n) {
  for (size_t i = blockIdx.x * size_t(256) + threadIdx.x; i < n; i += size_t(gridDim.x) * 256)
    st<NT>(b + i, ld<NT>(a + i) + ld<NT>(c + i));
}

template <typename F>
double time_ms(F launch, int reps) {
  hipEvent_t e0, e1;
  CHECK(hipEventCreate(&e0));
  CHECK(hipEventCreate(&e1));
  launch();
  CHECK(hipEventRecord(e0, 0));
  for (int r = 0; r < reps; ++r) launch();
  CHECK(hipEventRecord(e1, 0));
  CHECK(hipEventSynchronize(e1));
  float ms = 0.0f;
  CHECK(hipEventElapsedTime(&ms, e0, e1));
  CHECK(hipEventDestroy(e0));
  CHECK(hipEventDestroy(e1));
  return double(ms) / reps;
}

void streams() {
  double *a, *b, *c, *out;
  CHECK(hipMalloc(&a, NDBL * 8));
  CHECK(hipMalloc(&b, NDBL * 8));
  CHECK(hipMalloc(&c, NDBL * 8));
  CHECK(hipMalloc(&out, size_t(SGRID) * 256 * 8));
  CHECK(hipMemset(a, 0, NDBL * 8));
  CHECK(hipMemset(c, 0, NDBL * 8));
  const size_t n = NDBL / 2;
  const vd2* A = reinterpret_cast<const vd2*>(a);
  const vd2* C = reinterpret_cast<const vd2*>(c);
  vd2* B = reinterpret_cast<vd2*>(b);
  const double bytes = double(NDBL) * 8;
  auto row = [&](const char* name, double nbytes, double ms) {
    std::printf("  %-22s %8.3f ms  %6.3f TB/s\n", name, ms, nbytes / (ms * 1e-3) / 1e12);
  };
  std::printf("HBM streams, 1 GiB per array, 16-byte accesses, %d blocks of 256:\n", SGRID);
  row("copy", 2 * bytes, time_ms([&] { hipLaunchKernelGGL(k_copy<false>, dim3(SGRID), dim3(256), 0, 0, A, B, n); }, 10));
  row("copy (nontemporal)", 2 * bytes,
      time_ms([&] { hipLaunchKernelGGL(k_copy<true>, dim3(SGRID), dim3(256), 0, 0, A, B, n); }, 10));
  row("read", bytes, time_ms([&] { hipLaunchKernelGGL(k_read, dim3(SGRID), dim3(256), 0, 0, A, out, n); }, 10));
  row("fill", bytes, time_ms([&] { hipLaunchKernelGGL(k_fill, dim3(SGRID), dim3(256), 0, 0, B, n); }, 10));
  row("2 read + 1 write", 3 * bytes,
      time_ms([&] { hipLaunchKernelGGL(k_2r1w<false>, dim3(SGRID), dim3(256), 0, 0, A, C, B, n); }, 10));
  row("2r + 1w (nontemporal)", 3 * bytes,
      time_ms([&] { hipLaunchKernelGGL(k_2r1w<true>, dim3(SGRID), dim3(256), 0, 0, A, C, B, n); }, 10));
  CHECK(hipFree(a));
  CHECK(hipFree(b));
  CHECK(hipFree(c));
  CHECK(hipFree(out));
}

// ------------------------------------------------------------------ FP64 issue rate
constexpr int ITER = 16384;
constexpr int NCH = 8;
constexpr int MAXTHREADS = 1024;

// MIX 0: FMA on every chain; MIX 1: v_rcp_f64 on chains 0 and 4, FMA on the others
template <int MIX>
__global__ __launch_bounds__(MAXTHREADS) void k_issue(const double* __restrict__ in, double* __restrict__ out,
                                                      long long* __restrict__ clk) {
  __shared__ double pad[12288];  // 96 KiB: one block per CU
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  pad[threadIdx.x] = in[threadIdx.x & 255];
  __syncthreads();
  const double a = pad[(threadIdx.x + 1) & 255];
  double x[NCH];
#pragma unroll
  for (int c = 0; c < NCH; ++c) x[c] = a + c;
  const long long c0 = __builtin_amdgcn_s_memtime();
  const long long r0 = __builtin_amdgcn_s_memrealtime();
  for (int it = 0; it < ITER; ++it) {
#pragma unroll
    for (int c = 0; c < NCH; ++c) {
      if (MIX == 1 && c % 4 == 0) x[c] = __builtin_amdgcn_rcp(x[c]);
      else x[c] = fma(x[c], a, 1.0e-7);
    }
  }
  double s = 0.0;
#pragma unroll
  for (int c = 0; c < NCH; ++c) s += x[c];
  const long long c1 = __builtin_amdgcn_s_memtime();
  const long long r1 = __builtin_amdgcn_s_memrealtime();
  out[t] = s;
  if ((threadIdx.x & 63) == 0) {
    clk[2 * (t >> 6)] = c1 - c0;
    clk[2 * (t >> 6) + 1] = r1 - r0;
  }
}

template <int MIX>
void issue_row(const char* name, const double* in, double* out, long long* clk, int cus, int wps) {
  const int threads = 256 * wps;
  const double ms =
      time_ms([&] { hipLaunchKernelGGL(k_issue<MIX>, dim3(cus), dim3(threads), 0, 0, in, out, clk); }, 3);
  const int waves = cus * threads / 64;
  long long* h = static_cast<long long*>(std::malloc(sizeof(long long) * 2 * waves));
  CHECK(hipMemcpy(h, clk, sizeof(long long) * 2 * waves, hipMemcpyDeviceToHost));
  double mt = 0.0, rt = 0.0;
  for (int w = 0; w < waves; ++w) {
    mt += double(h[2 * w]);
    rt += double(h[2 * w + 1]);
  }
  std::free(h);
  const double instr_per_simd = double(ITER) * NCH * wps;  // wave64 instructions each SIMD issues
  std::printf("  %-12s %d wave(s)/SIMD: %6.3f ms, %5.2f ns per wave64 instruction per SIMD, clock %.2f GHz\n", name,
              wps, ms, ms * 1e6 / instr_per_simd, mt / (rt * 10.0));
}

void issue() {
  hipDeviceProp_t p;
  CHECK(hipGetDeviceProperties(&p, 0));
  const int cus = p.multiProcessorCount;
  double *in, *out;
  long long* clk;
  CHECK(hipMalloc(&in, 256 * 8));
  CHECK(hipMalloc(&out, size_t(cus) * MAXTHREADS * 8));
  CHECK(hipMalloc(&clk, size_t(cus) * (MAXTHREADS / 64) * 2 * 8));
  double h[256];
  for (int i = 0; i < 256; ++i) h[i] = 0.999 + 1e-6 * i;
  CHECK(hipMemcpy(in, h, sizeof h, hipMemcpyHostToDevice));
  std::printf("FP64 VALU issue, %d CUs, one block per CU, 8 independent chains per lane:\n", cus);
  for (int wps = 1; wps <= 4; ++wps) {
    issue_row<0>("8 fma", in, out, clk, cus, wps);
    issue_row<1>("2 rcp + 6 fma", in, out, clk, cus, wps);
  }
  CHECK(hipFree(in));
  CHECK(hipFree(out));
  CHECK(hipFree(clk));
}

}  // namespace

int main() {
  streams();
  issue();
  return 0;
}
