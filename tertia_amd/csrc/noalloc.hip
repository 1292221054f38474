// noalloc.hip — qss_transform2d_noalloc: the 2-D half-integer transform of
// qss_transform2d for callers that cannot allocate device memory (code under
// HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
//
// qss_transform2d runs on a context, whose creation allocates the transform
// tables and the intermediate plane of the two GEMM passes.  This entry point
// needs no context: both tables and the intermediate live in `scratch`, a
// fixed device array of this file (3 planes of kMaxN^2 doubles, 6 MiB per
// device, held until the process ends).  Each call rebuilds the two tables in
// it (the same exact integer argument reduction as the context's host-built
// tables), runs the y pass into the intermediate and the z pass into the
// caller's output, all enqueued on the caller's stream: no allocation, no
// synchronisation.  Every call on a device shares that device's scratch, so
// calls that overlap on one device race; one stream serialises them.
//
// The array's name is not a device address in host code: its per-device
// address comes from hipGetSymbolAddress, taken once per device with that
// device current, and only that address is handed to kernels.

#include <mutex>

#include "qss_internal.h"

namespace {

constexpr int kMaxN = 512;
constexpr size_t kPlane = (size_t)kMaxN * kMaxN;
constexpr int kMaxDevices = 64;

// T_y (ny x ny), T_z (nz x nz), the y pass's result (nz x ny)
__device__ __align__(256) double scratch[3 * kPlane];

std::mutex g_mu;
double *g_scratch[kMaxDevices] = {};

// 2 cos / 2 sin (pi (2j+1)(2k+1) / (4n)), argument reduced exactly mod 2 pi
// through (2j+1)(2k+1) mod 8n (the host tables of a context, make_tables)
__global__ __launch_bounds__(256) void k_half_integer_table(double *T, int n, int kind) {
  const long i = (long)blockIdx.x * 256 + threadIdx.x;
  if (i >= (long)n * n) return;
  const int j = (int)(i / n), k = (int)(i - (long)j * n);
  const long m = ((long)(2 * j + 1) * (2 * k + 1)) % (8L * n);
  const double a = M_PI * (double)m / (4.0 * n);
  T[i] = 2.0 * (kind == QSS_KIND_C ? cos(a) : sin(a));
}

int fail(const std::string &m, int code) {
  qss_set_error("qss_transform2d_noalloc: " + m);
  return code;
}

}  // namespace

extern "C" int qss_transform2d_noalloc(int ny, int nz, int kind_z, int kind_y,
                                       const double *in_dev, double *out_dev, double scale,
                                       void *stream) {
  if (ny < 2 || nz < 2 || ny > kMaxN || nz > kMaxN)
    return fail("ny, nz must be in [2, " + std::to_string(kMaxN) + "] (the scratch holds " +
                    std::to_string(kMaxN) + "^2 planes)",
                QSS_E_ARG);
  if ((kind_z != QSS_KIND_C && kind_z != QSS_KIND_S) ||
      (kind_y != QSS_KIND_C && kind_y != QSS_KIND_S))
    return fail("unknown transform kind", QSS_E_ARG);
  if (!in_dev || !out_dev || ((uintptr_t)in_dev & 15) || ((uintptr_t)out_dev & 15))
    return fail("planes must be non-null and 16-byte aligned", QSS_E_ARG);
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return fail(hipGetErrorString(e), QSS_E_HIP);
  if (dev < 0 || dev >= kMaxDevices) return fail("device index out of range", QSS_E_ARG);
  double *base = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_scratch[dev]) {   // first call on this device, the device current
      void *p = nullptr;
      e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
      if (e != hipSuccess) return fail(hipGetErrorString(e), QSS_E_HIP);
      g_scratch[dev] = static_cast<double *>(p);
    }
    base = g_scratch[dev];
  }
  double *Ty = base, *Tz = base + kPlane, *mid = base + 2 * kPlane;
  hipStream_t s = static_cast<hipStream_t>(stream);
  hipLaunchKernelGGL(k_half_integer_table, dim3((unsigned)(((long)ny * ny + 255) / 256)),
                     dim3(256), 0, s, Ty, ny, kind_y);
  hipLaunchKernelGGL(k_half_integer_table, dim3((unsigned)(((long)nz * nz + 255) / 256)),
                     dim3(256), 0, s, Tz, nz, kind_z);
  if ((e = hipGetLastError()) != hipSuccess) return fail(hipGetErrorString(e), QSS_E_HIP);
  // the two passes of transform_batch for one FP64 plane, default tile shape
  GemmArgs g{};
  g.nbatch = 1;
  g.nw = 1;
  g.b[0].resid = -1;
  g.b[0].push = -1;
  g.M = nz; g.N = ny; g.K = ny;
  g.alpha = 1.0;
  g.b[0].A = in_dev;
  g.b[0].B = Ty;
  g.b[0].C = mid;
  if ((e = launch_gemm(g, s, QSS_PREC_FP64, 0)) != hipSuccess)
    return fail(hipGetErrorString(e), QSS_E_HIP);
  g.M = nz; g.N = ny; g.K = nz;
  g.alpha = scale;
  g.b[0].A = Tz;
  g.b[0].B = mid;
  g.b[0].C = out_dev;
  if ((e = launch_gemm(g, s, QSS_PREC_FP64, 0)) != hipSuccess)
    return fail(hipGetErrorString(e), QSS_E_HIP);
  return QSS_OK;
}
