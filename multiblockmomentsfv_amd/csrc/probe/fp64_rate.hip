/*
 * fp64_rate.hip -- FP64 VALU issue-rate probe for the roofline of the stage
 * kernels (libmbmprobe.so, used by `bench.py --valu-rate`; not part of the
 * solver's C ABI).
 *
 * Every lane runs NACC independent FP64 chains (no dependency between
 * consecutive instructions of a chain's neighbours), so the measured time is
 * the SIMD's issue rate for that instruction mix at the given number of
 * resident waves per SIMD, on this GPU at the clock it runs under FP64 load.
 * The clock is read as shader-clock cycles (s_memtime) over the 100 MHz
 * real-time counter (s_memrealtime) inside the kernel.
 */
#include <hip/hip_runtime.h>

#include "../mbm_phys.h"

namespace {

constexpr int NACC = 16;
constexpr int ITERS = 2048;
constexpr int WG = 256;     /* 4 waves */

/* op 0: fma, 1: mul, 2: add, 3: fma/mul/add in turn (the stencil's mix) */
template <int OP>
__global__ void __launch_bounds__(WG) rate_kernel(double a, double b, double *sink, long long *clk)
{
  double acc[NACC];
#pragma unroll
  for (int k = 0; k < NACC; k++) acc[k] = threadIdx.x * 1e-3 + k;
  const long long c0 = clock64(), t0 = wall_clock64();
  for (int it = 0; it < ITERS; it++) {
#pragma unroll
    for (int k = 0; k < NACC; k++) {
      const int sel = OP == 3 ? k % 3 : OP;
      if (sel == 0) acc[k] = __builtin_fma(acc[k], a, b);
      else if (sel == 1) acc[k] = acc[k] * a;
      else acc[k] = acc[k] + b;
    }
  }
  const long long c1 = clock64(), t1 = wall_clock64();
  double s = 0.0;
#pragma unroll
  for (int k = 0; k < NACC; k++) s += acc[k];
  const size_t gid = (size_t)blockIdx.x * WG + threadIdx.x;
  sink[gid] = s;
  if (threadIdx.x == 0) {
    clk[2 * blockIdx.x] = c1 - c0;
    clk[2 * blockIdx.x + 1] = t1 - t0;
  }
}

template <int OP>
hipError_t launch(int wgs, double *sink, long long *clk, hipStream_t st)
{
  hipLaunchKernelGGL(rate_kernel<OP>, dim3(wgs), dim3(WG), 0, st, 1.0000001, 1e-9, sink, clk);
  return hipGetLastError();
}

hipError_t launch_op(int op, int wgs, double *sink, long long *clk, hipStream_t st)
{
  switch (op) {
    case 0: return launch<0>(wgs, sink, clk, st);
    case 1: return launch<1>(wgs, sink, clk, st);
    case 2: return launch<2>(wgs, sink, clk, st);
    default: return launch<3>(wgs, sink, clk, st);
  }
}

} // namespace

/* Runs op (0 fma, 1 mul, 2 add, 3 mixed) with waves_per_simd resident waves
 * on every SIMD of the current device (1, 2 or 4), in `launches` back-to-back
 * launches of which the last half are timed (the shader clock settles under
 * sustained FP64 load within a few milliseconds; one launch takes about
 * 70 us). Returns 0 on success and sets:
 *   ns_per_instr  wall time per wave-instruction per SIMD (nanoseconds),
 *   cycles_per_instr  shader-clock cycles per wave-instruction per SIMD, as
 *                 wall time times the clock,
 *   clock_ghz     mean shader clock during the last launch. */
extern "C" int mbm_probe_fp64_rate(int op, int waves_per_simd, int launches, double *ns_per_instr,
                                   double *cycles_per_instr, double *clock_ghz)
{
  if (op < 0 || op > 3 || (waves_per_simd != 1 && waves_per_simd != 2 && waves_per_simd != 4) ||
      launches < 2 || launches > 100000)
    return 1;
  int dev = 0, cus = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
    return 2;
  const int simds = 4 * cus;
  const int wgs = simds * waves_per_simd / 4;      /* 4 waves per workgroup */
  double *sink = nullptr;
  long long *clk = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = 3;
  const int untimed = launches - launches / 2, timed = launches / 2;
  bool ok = hipMalloc(&sink, (size_t)wgs * WG * sizeof(double)) == hipSuccess &&
            hipMalloc(&clk, (size_t)2 * wgs * sizeof(long long)) == hipSuccess &&
            hipEventCreate(&e0) == hipSuccess && hipEventCreate(&e1) == hipSuccess;
  for (int i = 0; ok && i < untimed; i++) ok = launch_op(op, wgs, sink, clk, nullptr) == hipSuccess;
  ok = ok && hipEventRecord(e0, nullptr) == hipSuccess;
  for (int i = 0; ok && i < timed; i++) ok = launch_op(op, wgs, sink, clk, nullptr) == hipSuccess;
  if (ok && hipEventRecord(e1, nullptr) == hipSuccess && hipEventSynchronize(e1) == hipSuccess) {
    float ms = 0.f;
    long long *h = new long long[2 * (size_t)wgs];
    if (hipEventElapsedTime(&ms, e0, e1) == hipSuccess &&
        hipMemcpy(h, clk, (size_t)2 * wgs * sizeof(long long), hipMemcpyDeviceToHost) == hipSuccess) {
      double cyc = 0.0, rt = 0.0;
      for (int i = 0; i < wgs; i++) {
        cyc += (double)h[2 * i];
        rt += (double)h[2 * i + 1];
      }
      cyc /= wgs;
      rt /= wgs;
      const double per_simd = (double)timed * waves_per_simd * ITERS * NACC;   /* wave-instructions */
      const double ghz = rt > 0 ? cyc / (rt * 10.0) : 0.0;          /* 100 MHz real-time counter */
      const double ns = ms * 1e6 / per_simd;
      if (ns_per_instr) *ns_per_instr = ns;
      if (cycles_per_instr) *cycles_per_instr = ns * ghz;
      if (clock_ghz) *clock_ghz = ghz;
      rc = 0;
    }
    delete[] h;
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (sink) (void)hipFree(sink);
  if (clk) (void)hipFree(clk);
  return rc;
}

namespace {
__global__ void math_kernel(int fn, const double *x, double *y, long long n)
{
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = fn == 0 ? mbm::sqrtp(x[i]) : mbm::rcp(x[i]);
}
} // namespace

/* The solver's device square root (fn 0, mbm::sqrtp) and reciprocal (fn 1,
 * mbm::rcp) of n host doubles x into host y, for accuracy tests. Returns 0
 * on success. */
extern "C" int mbm_probe_math(int fn, const double *x, double *y, long long n)
{
  if ((fn != 0 && fn != 1) || !x || !y || n < 1 || n > (1ll << 28)) return 1;
  double *dx = nullptr, *dy = nullptr;
  const size_t bytes = (size_t)n * sizeof(double);
  bool ok = hipMalloc(&dx, bytes) == hipSuccess && hipMalloc(&dy, bytes) == hipSuccess &&
            hipMemcpy(dx, x, bytes, hipMemcpyHostToDevice) == hipSuccess;
  if (ok) {
    hipLaunchKernelGGL(math_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, nullptr, fn, dx,
                       dy, n);
    ok = hipGetLastError() == hipSuccess && hipDeviceSynchronize() == hipSuccess &&
         hipMemcpy(y, dy, bytes, hipMemcpyDeviceToHost) == hipSuccess;
  }
  if (dx) (void)hipFree(dx);
  if (dy) (void)hipFree(dy);
  return ok ? 0 : 2;
}
