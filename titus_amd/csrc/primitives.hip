// Stateless device-pointer primitives of the C-ABI (include/titus_hip.h):
// the Galerkin products X^T L_o X of a spatial basis held in device memory.
//
// titus_galerkin allocates the reduction's per-workgroup partial sums for the
// call and frees them when the call ends (it waits for its stream).
// titus_galerkin_noalloc runs the same kernels on a file-scope device array,
// `scratch`, so it allocates nothing and does not wait: for callers that
// cannot call hipMalloc (graph capture, serving loops).  Each device holds its
// own copy of the array; its address on a device is looked up once with
// hipGetSymbolAddress while that device is current and kept per device.
#include <hip/hip_runtime.h>

#include <cmath>
#include <mutex>
#include <string>

#include "../../include/titus_hip.h"
#include "context.h"
#include "kernels.h"

namespace {

// the partial sums of one reduction: at most tn_partial_size() doubles
// (1024 workgroups x 6 operators x 256 values, 12 MiB)
constexpr size_t kScratchDoubles = size_t(1024) * 6 * 256;
static __device__ __align__(256) double scratch[kScratchDoubles];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};

// device address of `scratch` on the current device (looked up once per device)
double* scratch_on_current_device() {
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) throw TitusError(TITUS_E_ARG, "device index beyond the scratch table");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_scratch_ptr[dev] = static_cast<double*>(p);
  }
  return g_scratch_ptr[dev];
}

titus::dev::Grid check_grid(const titus_grid_v1* gr, const double* X, int32_t ldx, int32_t r,
                            double* out) {
  require(gr != nullptr && X != nullptr && out != nullptr && gr->inv_density != nullptr,
          "null argument");
  const int nz = gr->nz <= 1 ? 1 : gr->nz;
  require(gr->nx >= 3 && gr->ny >= 3 && (nz == 1 || nz >= 3), "need at least 3 cells per direction");
  require(gr->dx > 0 && gr->dy > 0 && (nz == 1 || gr->dz > 0), "cell widths must be positive");
  require((long long)gr->nx * gr->ny * nz < (1LL << 31), "grid too large");
  const int C = gr->nx * gr->ny * nz;
  require(r >= 1 && r <= 96, "rank must be in 1..96");
  require(ldx >= C, "leading dimension below the cell count");
  titus::dev::Grid g{gr->nx, gr->ny, C, gr->dx, gr->dy, nullptr, gr->inv_density, nz,
                     nz > 1 ? gr->dz : 1.0};
  return g;
}

}  // namespace

extern "C" {

int titus_galerkin(const titus_grid_v1* grid, const double* X, int32_t ldx, int32_t r,
                   double* out, void* stream) {
  return guarded([&] {
    const titus::dev::Grid g = check_grid(grid, X, ldx, r, out);
    hipStream_t st = static_cast<hipStream_t>(stream);
    double* partial = nullptr;
    HIPCHK(hipMalloc(&partial, titus::dev::tn_partial_size() * sizeof(double)));
    try {
      titus::dev::launch_tn(st, g, X, ldx, r, X, ldx, r, titus::dev::OP_ALL, partial, out);
      HIPCHK(hipStreamSynchronize(st));
    } catch (...) {
      (void)hipStreamSynchronize(st);
      (void)hipFree(partial);
      throw;
    }
    HIPCHK(hipFree(partial));
  });
}

int titus_galerkin_noalloc(const titus_grid_v1* grid, const double* X, int32_t ldx, int32_t r,
                           double* out, void* stream) {
  return guarded([&] {
    const titus::dev::Grid g = check_grid(grid, X, ldx, r, out);
    static_assert(kScratchDoubles >= 1024 * 6 * 256, "scratch below the partial-sum bound");
    require(titus::dev::tn_partial_size() <= kScratchDoubles, "partial sums exceed the scratch array");
    double* partial = scratch_on_current_device();
    titus::dev::launch_tn(static_cast<hipStream_t>(stream), g, X, ldx, r, X, ldx, r,
                          titus::dev::OP_ALL, partial, out);
  });
}

}  // extern "C"
