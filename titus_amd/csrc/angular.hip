// Angular half of the DLRA step on the device: the moment operators A_o
// (|Ax|, |Az|(, |Ay|), Ax, Az(, Ay), compressed rows) applied to the angular
// basis W (nm x k), the Galerkin blocks W_a^T (A_o W) (SolverCSD.jl:1021-1024,
// 1058-1061), the K-step coefficient blocks Z (:1027) and the L-step
// L = W S^T - dE sum_o A_o W S^T (X^T L_o X)^T (:1040).  Every block is
// column-major with its row count as leading dimension.
//
// Sizes: nm moments (<= kMaxAngRows), k, r <= 96 columns.  The products are
// tiny (a few MFLOP); what matters is that they run on the device between the
// spatial kernels, so the host does not compute them between stream syncs.
#include <hip/hip_runtime.h>

#include <stdexcept>
#include <string>

#include "angular.h"
#include "batch.h"

namespace titus {
namespace dev {
namespace {

constexpr int kAngBlock = 256;

inline void check(const char* what) {
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

__device__ inline double wave_sum(double s) {
  for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
  return s;
}

// One workgroup per (operator o, column j): column j of AW_o = A_o W into
// LDS and global memory (rows summed over the row's nonzeros in order), then
// G_o[:, j] = Wa^T AW_o[:, j] with one wavefront per output row i: lanes
// stride the moments, a shuffle tree adds the 64 partial sums.
__device__ __forceinline__ void d_ang_apply(Blk blk, 
    int nm, int k, int ka, const int* __restrict__ rowptr, const int* __restrict__ col,
    const double* __restrict__ val, const int* __restrict__ op_nz, const double* __restrict__ W,
    const double* __restrict__ Wa, double* __restrict__ AW, double* __restrict__ G) {
  __shared__ double wl[kMaxAngRows];      // column j of W
  __shared__ double colv[kMaxAngRows];    // column j of A_o W
  const int j = blk.x, o = blk.y;
  const int* rp = rowptr + (size_t)o * (nm + 1);
  const int* cl = col + op_nz[o];
  const double* vl = val + op_nz[o];
  const double* w = W + (size_t)j * nm;
  for (int m = threadIdx.x; m < nm; m += kAngBlock) wl[m] = w[m];
  __syncthreads();
  for (int m = threadIdx.x; m < nm; m += kAngBlock) {
    // the row's nonzeros in order; four index/value loads in flight
    double acc = 0.0;
    int t = rp[m];
    const int t1 = rp[m + 1];
    for (; t + 4 <= t1; t += 4) {
      const int c0 = cl[t], c1 = cl[t + 1], c2 = cl[t + 2], c3 = cl[t + 3];
      const double v0 = vl[t], v1 = vl[t + 1], v2 = vl[t + 2], v3 = vl[t + 3];
      acc += v0 * wl[c0];
      acc += v1 * wl[c1];
      acc += v2 * wl[c2];
      acc += v3 * wl[c3];
    }
    for (; t < t1; ++t) acc += vl[t] * wl[cl[t]];
    colv[m] = acc;
    AW[((size_t)o * k + j) * nm + m] = acc;
  }
  __syncthreads();
  if (!G) return;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int i = wave; i < ka; i += kAngBlock / 64) {
    const double* wa = Wa + (size_t)i * nm;
    double s = 0.0;
    for (int m = lane; m < nm; m += 64) s += wa[m] * colv[m];
    s = wave_sum(s);
    if (lane == 0) G[((size_t)o * k + j) * ka + i] = s;
  }
}

// Z = [S, S Cc, S Cx+, S Cx-, S Cy+, S Cy- (, S Cz+, S Cz-)] (r x n blocks) from
// the Galerkin blocks G_o = W^T A_o W (n x n, dissipative first):
// Cc = sum_d G_d / h_d, Cd+ = (G_{nd+d} - G_d) / (2 h_d), Cd- = (-G_{nd+d} - G_d) / (2 h_d),
// each entry of a combination formed where it is used.  One workgroup.
__device__ inline double comb_entry(const double* G, int nn, int nd, const double* h, int d, int idx) {
  if (d == 0) {
    double c = G[idx] / h[0];
    for (int dd = 1; dd < nd; ++dd) c += G[(size_t)dd * nn + idx] / h[dd];
    return c;
  }
  const int dd = (d - 1) / 2;
  const double m1 = G[(size_t)dd * nn + idx], m2 = G[(size_t)(nd + dd) * nn + idx];
  return ((d - 1) % 2 == 0) ? (m2 - m1) / (2.0 * h[dd]) : (-m2 - m1) / (2.0 * h[dd]);
}

// One workgroup per (column jj, block d): column jj of the combination C_d
// into LDS, then column jj of S C_d (block 0 also copies column jj of S).
__device__ __forceinline__ void d_ang_stream_z(Blk blk, 
    int r, int n, int nd, double hx, double hy, double hz, const double* __restrict__ S,
    const double* __restrict__ G, double* __restrict__ Z) {
  __shared__ double cc[96];
  const double h[3] = {hx, hy, hz};
  const int jj = blk.x, d = blk.y;
  const int nn = n * n, rn = r * n;
  for (int l = threadIdx.x; l < n; l += kAngBlock) cc[l] = comb_entry(G, nn, nd, h, d, jj * n + l);
  if (d == 0)
    for (int i = threadIdx.x; i < r; i += kAngBlock) Z[(size_t)jj * r + i] = S[(size_t)jj * r + i];
  __syncthreads();
  for (int i = threadIdx.x; i < r; i += kAngBlock) {
    double acc = 0.0;
    for (int l = 0; l < n; ++l) acc += S[(size_t)l * r + i] * cc[l];
    Z[(size_t)(d + 1) * rn + (size_t)jj * r + i] = acc;
  }
}

// L = W S^T - dE sum_{o in order} AW_o (S^T XLX_o^T) (nm x r): one workgroup
// per column c; v_o = S^T XLX_o[c, :]^T (r values per operator) in LDS, then
// one thread per moment.
__device__ __forceinline__ void d_ang_stream_l(Blk blk, 
    int nm, int r, int nops, const int* __restrict__ order, double dE,
    const double* __restrict__ W, const double* __restrict__ S, const double* __restrict__ AW,
    const double* __restrict__ XLX, double* __restrict__ L) {
  __shared__ double v[kMaxAngOps * 96];
  __shared__ double s_c[96];
  const int c = blk.x;
  for (int e = threadIdx.x; e < nops * r; e += kAngBlock) {
    const int t = e / r, b = e - t * r;
    const int o = order[t];
    const double* xlx = XLX + (size_t)o * r * r;
    double acc = 0.0;
    for (int a = 0; a < r; ++a) acc += S[(size_t)b * r + a] * xlx[(size_t)a * r + c];
    v[e] = acc;
  }
  for (int b = threadIdx.x; b < r; b += kAngBlock) s_c[b] = S[(size_t)b * r + c];
  __syncthreads();
  for (int m = threadIdx.x; m < nm; m += kAngBlock) {
    double ws = 0.0;
    for (int b = 0; b < r; ++b) ws += W[(size_t)b * nm + m] * s_c[b];
    double acc = 0.0;
    for (int t = 0; t < nops; ++t) {
      const int o = order[t];
      const double* aw = AW + (size_t)o * r * nm;
      double term = 0.0;
      for (int b = 0; b < r; ++b) term += aw[(size_t)b * nm + m] * v[t * r + b];
      acc += term;
    }
    L[(size_t)c * nm + m] = ws - dE * acc;
  }
}

struct AngApplyBody {
  struct P {
    int nm, k, ka; const int* rowptr; const int* col; const double* val; const int* op_nz;
    const double* W; const double* Wa; double* AW; double* G;
  };
  static constexpr int kThreads = kAngBlock;
  static constexpr bool kBatch = true;
  __device__ static void run(const P& p, Blk b) {
    d_ang_apply(b, p.nm, p.k, p.ka, p.rowptr, p.col, p.val, p.op_nz, p.W, p.Wa, p.AW, p.G);
  }
};
struct AngStreamZBody {
  struct P { int r, n, nd; double hx, hy, hz; const double* S; const double* G; double* Z; };
  static constexpr int kThreads = kAngBlock;
  static constexpr bool kBatch = true;
  __device__ static void run(const P& p, Blk b) {
    d_ang_stream_z(b, p.r, p.n, p.nd, p.hx, p.hy, p.hz, p.S, p.G, p.Z);
  }
};
struct AngStreamLBody {
  struct P {
    int nm, r, nops; const int* order; double dE; const double* W; const double* S;
    const double* AW; const double* XLX; double* L;
  };
  static constexpr int kThreads = kAngBlock;
  static constexpr bool kBatch = true;
  __device__ static void run(const P& p, Blk b) {
    d_ang_stream_l(b, p.nm, p.r, p.nops, p.order, p.dE, p.W, p.S, p.AW, p.XLX, p.L);
  }
};

}  // namespace

void launch_ang_apply(hipStream_t st, const AngOps& A, const double* W, int k, const double* Wa,
                      int ka, double* AW, double* G) {
  if (A.nm > kMaxAngRows) throw std::runtime_error("angular: more moments than the kernel holds");
  if (k < 1) return;
  launch_body<AngApplyBody>(st, dim3(k, A.nops), 0,
                            {A.nm, k, ka, A.rowptr, A.col, A.val, A.op_nz, W, Wa, AW, G});
  check("ang_apply");
}

void launch_ang_stream_z(hipStream_t st, int r, int n, int nd, double hx, double hy, double hz,
                         const double* S, const double* G, double* Z) {
  if (n > 96 || r > 96) throw std::runtime_error("angular: Z blocks beyond 96 columns");
  launch_body<AngStreamZBody>(st, dim3(n, 1 + 2 * nd), 0, {r, n, nd, hx, hy, hz, S, G, Z});
  check("ang_stream_z");
}

void launch_ang_stream_l(hipStream_t st, int nm, int r, int nops, const int* order, double dE,
                         const double* W, const double* S, const double* AW, const double* XLX,
                         double* L) {
  if (r > 96 || nops > kMaxAngOps) throw std::runtime_error("angular: L-step beyond its limits");
  launch_body<AngStreamLBody>(st, dim3(r), 0, {nm, r, nops, order, dE, W, S, AW, XLX, L});
  check("ang_stream_l");
}

}  // namespace dev
}  // namespace titus
