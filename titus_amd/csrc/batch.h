// Two launch forms of one kernel: alone, or once for many parameter sets.
//
// A kernel is a `Body`: a parameter struct Body::P, a workgroup size
// Body::kThreads and a device function Body::run(const P&, Blk) that does
// the work of one workgroup given its block coordinates in its own grid.
//   * k_one<Body>   runs one grid with P in the kernel arguments;
//   * k_many<Body>  runs the grids of `count` parameter sets in one launch:
//                   workgroup (x, y) is block x of item y (blocks past item y's
//                   own grid return at once), the items are read from device
//                   memory.
// A launcher calls launch_body<Body>(stream, grid, lds, p).  When the calling
// host thread has a Recorder installed (t_recorder), the launch is recorded
// instead of issued; a batch of contexts that advance together then issues
// the recorded launches of all its contexts with one k_many launch per
// kernel (solver.cpp, titus_solve_batch_lockstep).  The arithmetic of a
// workgroup is the same in both forms, so a batched solve reproduces the
// same solve run alone bit for bit.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

namespace titus {
namespace dev {

// block coordinates of a workgroup in its own grid (nx blocks along x)
struct Blk {
  int x, y, nx;
};

template <class P>
struct alignas(16) Item {
  P p;
  int gx, gy;
};

template <class Body>
__global__ __launch_bounds__(Body::kThreads) void k_one(const typename Body::P p) {
  Body::run(p, Blk{(int)blockIdx.x, (int)blockIdx.y, (int)gridDim.x});
}

// The item is read through the constant address space, as kernel arguments
// are: its fields are uniform, so they load into scalar registers and stay
// there (a plain global read would hold them in vector registers).
template <class Body>
__global__ __launch_bounds__(Body::kThreads) void k_many(
    const Item<typename Body::P>* __restrict__ items) {
  using CItem = __attribute__((address_space(4))) const Item<typename Body::P>;
  const CItem* it = (const CItem*)(items + blockIdx.y);
  const int gx = it->gx;
  if ((int)blockIdx.x >= gx * it->gy) return;
  // word-wise copy (the constant address space has no struct copy)
  using P = typename Body::P;
  static_assert(sizeof(P) % 4 == 0, "parameter structs hold whole 32-bit words");
  using CWord = __attribute__((address_space(4))) const unsigned int;
  const CWord* src = (const CWord*)&it->p;
  P p;
  unsigned int* dst = reinterpret_cast<unsigned int*>(&p);
#pragma unroll
  for (int i = 0; i < (int)(sizeof(P) / 4); ++i) dst[i] = src[i];
  Body::run(p, Blk{(int)blockIdx.x % gx, (int)blockIdx.x / gx, gx});
}

// launch of `count` items (device memory, consecutive) of one kernel
using ManyFn = void (*)(hipStream_t st, const void* items, int count, unsigned blocks, size_t lds);

struct RecOp {
  ManyFn fn;        // the kernel (also the grouping key of a batch)
  size_t off;       // the item's bytes in Recorder::bytes
  size_t size;
  unsigned blocks;  // workgroups of the item's own grid
  size_t lds;       // dynamic LDS bytes
  // variants of a kernel with the same item layout: a batched launch of
  // items with variant indices v runs alts[max v] (e.g. the narrowest
  // template that covers every item's tile count); null: fn
  int var = 0;
  const ManyFn* alts = nullptr;
  int family = 0;   // TuneFamily of the launch (0: not tuned)
  double work = 0;  // its work in the family's unit (16-cell tiles, cell-ordinates)
};

// Tile choices of a lockstep group, set by the group's tuner between flushes
// (0, or -1 for sweep_nt: the launcher's own choice).  None of them changes
// a workgroup's arithmetic; they regroup the same tiles into workgroups.
struct BatchTuning {
  int gemm_tpw = 0;    // 16-cell tiles per wavefront, matrix-core GEMM
  int kstep_tpw = 0;   // the same, matrix-core K-step
  int sweep_rows = 0;  // rows per strip of the upwind sweep
  int sweep_nt = -1;   // non-temporal psi streaming of the sweep (0 / 1)
};
enum TuneFamily { TF_NONE = 0, TF_GEMM = 1, TF_KSTEP = 2, TF_SWEEP = 3, TF_COUNT = 3 };

// Launches of one host thread, in issue order.
struct Recorder {
  std::vector<RecOp> ops;
  std::vector<unsigned char> bytes;
  void add(ManyFn fn, const void* item, size_t size, unsigned blocks, size_t lds, int var = 0,
           const ManyFn* alts = nullptr);
  void clear() {
    ops.clear();
    bytes.clear();
  }
};

// the calling thread's recorder (null: launches are issued directly)
inline thread_local Recorder* t_recorder = nullptr;
// the tile choices of the calling thread's group (null: the launchers' own)
inline thread_local const BatchTuning* t_tune = nullptr;
// family and work of the next launch the calling thread records (set by a
// tuned launcher just before it records; consumed by Recorder::add)
inline thread_local int t_family = TF_NONE;
inline thread_local double t_work = 0.0;

inline void Recorder::add(ManyFn fn, const void* item, size_t size, unsigned blocks, size_t lds,
                          int var, const ManyFn* alts) {
  const size_t off = (bytes.size() + 15) & ~size_t(15);
  bytes.resize(off + size);
  std::memcpy(bytes.data() + off, item, size);
  ops.push_back(RecOp{fn, off, size, blocks, lds, var, alts, t_family, t_work});
  t_family = TF_NONE;
  t_work = 0.0;
}

template <class Body>
void launch_many(hipStream_t st, const void* items, int count, unsigned blocks, size_t lds) {
  hipLaunchKernelGGL(k_many<Body>, dim3(blocks, (unsigned)count), dim3(Body::kThreads), lds, st,
                     static_cast<const Item<typename Body::P>*>(items));
}

// Bodies with kBatch = false (parameters too large to hold in registers:
// coefficient blocks carried in the arguments) have no batched form.
template <class Body>
void launch_body(hipStream_t st, dim3 grid, size_t lds, const typename Body::P& p) {
  if (grid.x == 0 || grid.y == 0) {
    t_family = TF_NONE;
    return;
  }
  if (Recorder* r = t_recorder) {
    if constexpr (Body::kBatch) {
      const Item<typename Body::P> it{p, (int)grid.x, (int)grid.y};
      r->add(&launch_many<Body>, &it, sizeof(it), grid.x * grid.y, lds);
      return;
    } else {
      throw std::runtime_error("a kernel with its coefficients in the arguments has no batched form");
    }
  }
  hipLaunchKernelGGL(k_one<Body>, grid, dim3(Body::kThreads), lds, st, p);
}

// A batched launch with variants: the item is recorded under Body's kernel
// (the grouping key) with variant index `var` into `alts` (kernels whose
// items have Body's layout).
template <class Body>
void record_variant(dim3 grid, size_t lds, const typename Body::P& p, int var, const ManyFn* alts) {
  const Item<typename Body::P> it{p, (int)grid.x, (int)grid.y};
  t_recorder->add(&launch_many<Body>, &it, sizeof(it), grid.x * grid.y, lds, var, alts);
}

// launchers of kernels with no batched form refuse to run while recording
inline void no_batch(const char* what) {
  if (t_recorder) throw std::runtime_error(std::string(what) + ": no batched launch form");
}

// dst[0 : bytes) = src[0 : bytes) on the device (either side may be pinned
// host memory mapped for the device); recorded like any kernel
void launch_copy(hipStream_t st, void* dst, const void* src, size_t bytes);
// the same copy recorded into `r` (whatever the calling thread records into)
void record_copy(Recorder& r, void* dst, const void* src, size_t bytes);

}  // namespace dev
}  // namespace titus
