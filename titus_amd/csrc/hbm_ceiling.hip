// HBM ceilings of one MI355X, measured: the denominators that tell how far a
// memory-bound kernel of the energy step is from what the memory system gives
// a kernel of its access pattern.
//
//   hbm_ceiling [nx ny nq]     (default 1025 1025 72: the large-grid stencil
//                               workload's psi planes)
//
// prints one JSON object:
//   read      -- a 1 GiB grid-stride read (16-byte loads);
//   fill      -- 1 GiB writes: hipMemsetAsync, grid-stride stores, and stores
//                where each workgroup writes its own contiguous chunk;
//   copy      -- 1 GiB copies: hipMemcpyAsync and grid-stride copies of
//                several shapes (element size, loads in flight per lane,
//                workgroups, non-temporal stores);
//   strip_copy-- a copy of nq planes of nx x ny cells with the access pattern
//                of the upwind sweep and none of its arithmetic: a wavefront
//                owns rows i0 .. i0+H-1 of 62 columns (lanes 1..62; lanes 0
//                and 63 hold the neighbouring columns and load them too),
//                walks the ordinates of its group and writes its rows of
//                every plane; workgroups numbered as the sweep numbers them
//                (column tile fastest, then strip block, then ordinate group,
//                each XCD one contiguous range).  FP64 and FP32 planes.
// Every figure is bytes moved / average duration of 10 launches after 3
// untimed ones (HIP events), in GB/s and as a fraction of 8 TB/s.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

namespace {

constexpr double kPeak = 8.0e12;
constexpr int kThreads = 256;
constexpr int kXcds = 8;
constexpr int kCols = 62;

#define CHECK(x)                                                                        \
  do {                                                                                  \
    hipError_t e_ = (x);                                                                \
    if (e_ != hipSuccess) {                                                             \
      std::fprintf(stderr, "hbm_ceiling: %s at line %d\n", hipGetErrorString(e_), __LINE__); \
      std::exit(1);                                                                     \
    }                                                                                   \
  } while (0)

typedef float f4 __attribute__((ext_vector_type(4)));

__device__ inline int xcd_block(int b, int nb) {
  const int per = nb / kXcds, rem = nb % kXcds;
  const int x = b % kXcds, k = b / kXcds;
  return x * per + min(x, rem) + k;
}

__global__ __launch_bounds__(kThreads) void k_read(const f4* __restrict__ a, size_t n, float* sink) {
  const size_t stride = (size_t)gridDim.x * kThreads;
  f4 s = {0, 0, 0, 0};
  for (size_t i = (size_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += stride) s += a[i];
  // never true for the finite data written here; keeps the loads
  if (s.x + s.y + s.z + s.w == -1.0f) sink[0] = s.x;
}

template <class V>
__global__ __launch_bounds__(kThreads) void k_fill_stride(V* __restrict__ b, size_t n) {
  const size_t stride = (size_t)gridDim.x * kThreads;
  const V z = (V)0;
  for (size_t i = (size_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += stride) b[i] = z;
}

// workgroup w writes elements [w per, (w + 1) per)
__global__ __launch_bounds__(kThreads) void k_fill_chunks(f4* __restrict__ b, size_t n, size_t per) {
  const size_t s = (size_t)blockIdx.x * per;
  const size_t e = s + per < n ? s + per : n;
  const f4 z = {0, 0, 0, 0};
  for (size_t k = s + threadIdx.x; k < e; k += kThreads) b[k] = z;
}

// grid-stride copy with U loads in flight per lane
template <class V, int U, bool NTS>
__global__ __launch_bounds__(kThreads) void k_copy_stride(const V* __restrict__ a, V* __restrict__ b, size_t n) {
  const size_t stride = (size_t)gridDim.x * kThreads;
  size_t i = (size_t)blockIdx.x * kThreads + threadIdx.x;
  for (; i + (size_t)(U - 1) * stride < n; i += (size_t)U * stride) {
    V t[U];
#pragma unroll
    for (int u = 0; u < U; ++u) t[u] = a[i + u * stride];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      if (NTS) __builtin_nontemporal_store(t[u], b + i + u * stride);
      else b[i + u * stride] = t[u];
    }
  }
  for (; i < n; i += stride) b[i] = a[i];
}

// the sweep's access pattern as a plain copy (see the file comment)
template <class T, int H>
__global__ __launch_bounds__(kThreads) void k_strip_copy(const T* __restrict__ pin, T* __restrict__ pout,
                                                          int nx, int ny, int nq, int qper, int ntj, int nsb) {
  const size_t C = (size_t)nx * ny;
  const int L = xcd_block(blockIdx.x, gridDim.x);
  const int jt = L % ntj, sb = (L / ntj) % nsb, gq = L / (ntj * nsb);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int i0 = (sb * (kThreads / 64) + wave) * H;
  const int j = jt * kCols - 1 + lane;
  const bool jv = j >= 0 && j < ny;
  const bool owner = lane >= 1 && lane <= kCols && jv;
  const int qa = gq * qper, qb = min(nq, qa + qper);
  T P[H], Pn[H] = {};
  auto load = [&](int q, T* dst) {
#pragma unroll
    for (int h = 0; h < H; ++h) {
      const int i = i0 + h;
      dst[h] = (jv && i < nx) ? pin[(size_t)q * C + (size_t)i * ny + j] : T(0);
    }
  };
  if (qa < qb) load(qa, P);
  for (int q = qa; q < qb; ++q) {
    if (q + 1 < qb) load(q + 1, Pn);
#pragma unroll
    for (int h = 0; h < H; ++h) {
      const int i = i0 + h;
      if (owner && i < nx) pout[(size_t)q * C + (size_t)i * ny + j] = P[h];
    }
#pragma unroll
    for (int h = 0; h < H; ++h) P[h] = Pn[h];
  }
}

struct Timer {
  hipEvent_t e0, e1;
  Timer() {
    CHECK(hipEventCreate(&e0));
    CHECK(hipEventCreate(&e1));
  }
  // average seconds of one call over 10 timed calls after 3 untimed ones
  template <class F>
  double operator()(F&& f) {
    for (int w = 0; w < 3; ++w) f();
    CHECK(hipGetLastError());
    CHECK(hipDeviceSynchronize());
    CHECK(hipEventRecord(e0));
    for (int r = 0; r < 10; ++r) f();
    CHECK(hipEventRecord(e1));
    CHECK(hipEventSynchronize(e1));
    CHECK(hipGetLastError());
    float ms = 0;
    CHECK(hipEventElapsedTime(&ms, e0, e1));
    return ms / 1e3 / 10;
  }
};

std::string entry(const std::string& name, double bytes, double s) {
  char buf[256];
  std::snprintf(buf, sizeof(buf), "{\"shape\": \"%s\", \"us\": %.1f, \"GB/s\": %.1f, \"frac_of_8TB/s\": %.3f}",
                name.c_str(), s * 1e6, bytes / s / 1e9, bytes / s / kPeak);
  return buf;
}

std::string join(const std::vector<std::string>& v) {
  std::string out = "[";
  for (size_t k = 0; k < v.size(); ++k) out += (k ? ", " : "") + v[k];
  return out + "]";
}

template <class T>
std::vector<std::string> strip_copies(Timer& timer, int nx, int ny, int nq) {
  const size_t C = (size_t)nx * ny, bytes = C * nq * sizeof(T);
  T *a = nullptr, *b = nullptr;
  CHECK(hipMalloc(&a, bytes));
  CHECK(hipMalloc(&b, bytes));
  CHECK(hipMemset(a, 0, bytes));
  std::vector<std::string> out;
  const int ntj = (ny + kCols - 1) / kCols;
  for (int H : {2, 4}) {
    const int nsb = (nx + 4 * H - 1) / (4 * H);
    for (int groups : {4, 8}) {
      const int qper = (nq + groups - 1) / groups, G = (nq + qper - 1) / qper;
      const dim3 grid((unsigned)(ntj * nsb * G));
      const double s = timer([&] {
        if (H == 2)
          hipLaunchKernelGGL((k_strip_copy<T, 2>), grid, dim3(kThreads), 0, 0, a, b, nx, ny, nq, qper, ntj, nsb);
        else
          hipLaunchKernelGGL((k_strip_copy<T, 4>), grid, dim3(kThreads), 0, 0, a, b, nx, ny, nq, qper, ntj, nsb);
      });
      out.push_back(entry(std::string(sizeof(T) == 8 ? "fp64" : "fp32") + " rows " + std::to_string(H) +
                              " groups " + std::to_string(G),
                          2.0 * bytes, s));
    }
  }
  CHECK(hipFree(a));
  CHECK(hipFree(b));
  return out;
}

}  // namespace

int main(int argc, char** argv) {
  int nx = 1025, ny = 1025, nq = 72;
  if (argc == 4) {
    nx = std::atoi(argv[1]);
    ny = std::atoi(argv[2]);
    nq = std::atoi(argv[3]);
  }
  if (nx < 1 || ny < 1 || nq < 1 || (size_t)nx * ny * nq * 8 > ((size_t)8 << 30)) {
    std::fprintf(stderr, "hbm_ceiling: planes must hold at most 8 GiB\n");
    return 2;
  }
  const size_t sz = (size_t)1 << 30;
  char *a = nullptr, *b = nullptr;
  float* sink = nullptr;
  CHECK(hipMalloc(&a, sz));
  CHECK(hipMalloc(&b, sz));
  CHECK(hipMalloc(&sink, sizeof(float)));
  CHECK(hipMemset(a, 0, sz));
  Timer timer;
  const double B = (double)sz;

  std::vector<std::string> read, fill, copy;
  for (unsigned g : {2048u, 8192u})
    read.push_back(entry("grid-stride 16 B loads, " + std::to_string(g) + " workgroups", B, timer([&] {
      hipLaunchKernelGGL(k_read, dim3(g), dim3(kThreads), 0, 0, (const f4*)a, sz / 16, sink);
    })));

  fill.push_back(entry("hipMemsetAsync", B, timer([&] { CHECK(hipMemsetAsync(b, 0, sz, 0)); })));
  for (unsigned g : {1024u, 8192u})
    fill.push_back(entry("grid-stride 8 B stores, " + std::to_string(g) + " workgroups", B, timer([&] {
      hipLaunchKernelGGL((k_fill_stride<double>), dim3(g), dim3(kThreads), 0, 0, (double*)b, sz / 8);
    })));
  for (size_t kib : {16, 256}) {
    const size_t per = kib * 1024 / 16;
    const unsigned g = (unsigned)(sz / 16 / per);
    fill.push_back(entry("one " + std::to_string(kib) + " KiB chunk per workgroup, 16 B stores", B, timer([&] {
      hipLaunchKernelGGL(k_fill_chunks, dim3(g), dim3(kThreads), 0, 0, (f4*)b, sz / 16, per);
    })));
  }

  copy.push_back(entry("hipMemcpyAsync", 2 * B, timer([&] {
    CHECK(hipMemcpyAsync(b, a, sz, hipMemcpyDeviceToDevice, 0));
  })));
  for (unsigned g : {512u, 1024u, 2048u}) {
    const std::string w = ", " + std::to_string(g) + " workgroups";
    copy.push_back(entry("8 B x 4 in flight" + w, 2 * B, timer([&] {
      hipLaunchKernelGGL((k_copy_stride<double, 4, false>), dim3(g), dim3(kThreads), 0, 0, (const double*)a,
                         (double*)b, sz / 8);
    })));
    copy.push_back(entry("8 B x 4 in flight, non-temporal stores" + w, 2 * B, timer([&] {
      hipLaunchKernelGGL((k_copy_stride<double, 4, true>), dim3(g), dim3(kThreads), 0, 0, (const double*)a,
                         (double*)b, sz / 8);
    })));
    copy.push_back(entry("4 B x 8 in flight, non-temporal stores" + w, 2 * B, timer([&] {
      hipLaunchKernelGGL((k_copy_stride<float, 8, true>), dim3(g), dim3(kThreads), 0, 0, (const float*)a,
                         (float*)b, sz / 4);
    })));
  }
  CHECK(hipFree(a));
  CHECK(hipFree(b));
  CHECK(hipFree(sink));

  std::vector<std::string> strips = strip_copies<double>(timer, nx, ny, nq);
  for (const std::string& s : strip_copies<float>(timer, nx, ny, nq)) strips.push_back(s);

  std::printf("{\"bytes\": %zu, \"read\": %s, \"fill\": %s, \"copy\": %s, \"strip_copy\": {\"nx\": %d, \"ny\": %d, "
              "\"nq\": %d, \"runs\": %s}}\n",
              sz, join(read).c_str(), join(fill).c_str(), join(copy).c_str(), nx, ny, nq, join(strips).c_str());
  return 0;
}
