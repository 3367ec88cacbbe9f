// Device kernels of the DLRA energy step for gfx950 (see kernels.h).
#include "kernels.h"

#include <algorithm>
#include <stdexcept>
#include <string>

namespace titus {
namespace dev {

namespace {

constexpr int kBlock = 256;
constexpr int kMaxChunks = 256;      // blocks (row chunks) of the TN reductions
constexpr int kSub = 32;             // rows staged in LDS at a time
constexpr int kPairsPerThread = 8;   // outputs accumulated per thread
constexpr int kPairsPerGroup = kBlock * kPairsPerThread;

inline void check_launch(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess)
    throw std::runtime_error(std::string("kernel launch failed (") + what + "): " +
                             hipGetErrorString(e));
}

inline unsigned blocks_for(size_t n) { return (unsigned)((n + kBlock - 1) / kBlock); }

__device__ inline bool interior(int c, int nx, int ny) {
  const int i = c / ny, j = c - (c / ny) * ny;
  return i >= 1 && i <= nx - 2 && j >= 1 && j <= ny - 2;
}

// ---------------------------------------------------------------------------
template <class TP>
__global__ void k_bc(TP* __restrict__ psi, int C, int nq, const int* __restrict__ bcells, int nb,
                     const double* __restrict__ base, double scale) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (size_t)nb * nq) return;
  const int b = (int)(t % nb), q = (int)(t / nb);
  psi[(size_t)q * C + bcells[b]] = (TP)(base[(size_t)q * nb + b] * scale);
}

template <class TP>
__global__ void k_dose(double* __restrict__ dose, const double* __restrict__ X, int ldx, int r,
                       const double* __restrict__ v, const TP* __restrict__ psi, int C, int nq,
                       const double* __restrict__ m0, const double* __restrict__ rho,
                       double coef) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= C) return;
  double acc = 0.0;
  for (int a = 0; a < r; ++a) acc += X[(size_t)a * ldx + c] * v[a];
  double u0 = 0.0;
  for (int q = 0; q < nq; ++q) u0 += (double)psi[(size_t)q * C + c] * m0[q];
  dose[c] += coef * (acc + u0) / rho[c];
}

template <class T>
__global__ void k_sweep(const T* __restrict__ pin, T* __restrict__ pout, int nx, int ny, int nq,
                        const double* __restrict__ omega, T dx, T dy,
                        const double* __restrict__ rho, T dE, T denom,
                        const unsigned char* __restrict__ restore) {
  const int C = nx * ny;
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (size_t)C * nq) return;
  const int c = (int)(t % C), q = (int)(t / C);
  const T* p = pin + (size_t)q * C;
  const T pc = p[c];
  T out;
  if (interior(c, nx, ny)) {
    const T ox = (T)omega[q], oz = (T)omega[2 * nq + q];
    const T vc = pc / (T)rho[c];
    const T fx = (ox >= T(0)) ? (vc - p[c - ny] / (T)rho[c - ny]) : (p[c + ny] / (T)rho[c + ny] - vc);
    const T fy = (oz >= T(0)) ? (vc - p[c - 1] / (T)rho[c - 1]) : (p[c + 1] / (T)rho[c + 1] - vc);
    const T flux = fx * (ox / dx) + fy * (oz / dy);
    out = (pc - dE * flux) / denom;
  } else {
    out = (q == 0 && restore[c]) ? pc : pc / denom;
  }
  pout[t] = out;
}

// K[c,b] = X[c,:] Z0[:,b] - dE * sum_d (X[d,:]/rho[d]) Zd[:,b] (interior c)
template <class TC>
__global__ void k_stream_k(const double* __restrict__ X, int ldx, int r, int nx, int ny,
                           const double* __restrict__ rho, const double* __restrict__ Z, int n,
                           double dE, double* __restrict__ K, int ldk) {
  const int C = nx * ny;
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  const int b = blockIdx.y;
  if (c >= C) return;
  const double* z0 = Z + (size_t)b * r;
  TC acc0 = 0;
  for (int a = 0; a < r; ++a) acc0 += (TC)X[(size_t)a * ldx + c] * (TC)z0[a];
  TC out = acc0;
  if (interior(c, nx, ny)) {
    const double* zc = Z + (size_t)(1 * n + b) * r;
    const double* ze = Z + (size_t)(2 * n + b) * r;
    const double* zw = Z + (size_t)(3 * n + b) * r;
    const double* zn = Z + (size_t)(4 * n + b) * r;
    const double* zs = Z + (size_t)(5 * n + b) * r;
    TC sc = 0, se = 0, sw = 0, sn = 0, ss = 0;
    for (int a = 0; a < r; ++a) {
      const double* xa = X + (size_t)a * ldx;
      sc += (TC)xa[c] * (TC)zc[a];
      se += (TC)xa[c + ny] * (TC)ze[a];
      sw += (TC)xa[c - ny] * (TC)zw[a];
      sn += (TC)xa[c + 1] * (TC)zn[a];
      ss += (TC)xa[c - 1] * (TC)zs[a];
    }
    const TC s = sc / (TC)rho[c] + se / (TC)rho[c + ny] + sw / (TC)rho[c - ny] +
                 sn / (TC)rho[c + 1] + ss / (TC)rho[c - 1];
    out = acc0 - (TC)dE * s;
  }
  K[(size_t)b * ldk + c] = (double)out;
}

template <class TB, class TC>
__device__ inline TC stencil_value(const TB* __restrict__ B, int ldb, int b, int c, int op, int nx,
                                   int ny, TC dx, TC dy, const double* __restrict__ rho) {
  if (!interior(c, nx, ny)) return TC(0);
  const TB* col = B + (size_t)b * ldb;
  const TC h = TC(0.5);
  switch (op) {
    case OP_L1X:
      return ((TC)col[c] / (TC)rho[c] - h * ((TC)col[c - ny] / (TC)rho[c - ny]) -
              h * ((TC)col[c + ny] / (TC)rho[c + ny])) / dx;
    case OP_L1Y:
      return ((TC)col[c] / (TC)rho[c] - h * ((TC)col[c - 1] / (TC)rho[c - 1]) -
              h * ((TC)col[c + 1] / (TC)rho[c + 1])) / dy;
    case OP_L2X:
      return ((TC)col[c + ny] / (TC)rho[c + ny] - (TC)col[c - ny] / (TC)rho[c - ny]) / (TC(2) * dx);
    default:
      return ((TC)col[c + 1] / (TC)rho[c + 1] - (TC)col[c - 1] / (TC)rho[c - 1]) / (TC(2) * dy);
  }
}

// partial[chunk][o*mn + p] = sum over the chunk's rows of A[c, a] * op_o(B)[c, b],
// p = a + b*m.  NOPS = 1: the single operator `op`; NOPS = 4: the four
// stencils L1x, L1y, L2x, L2y at once (A staged once, B's images side by side).
// Staging and accumulation in TC; partials written in FP64.
template <int NOPS, class TA, class TB, class TC>
__global__ void k_tn_partial(const TA* __restrict__ A, int lda, int m, const TB* __restrict__ B,
                             int ldb, int n, int C, int op, int nx, int ny, double dx_, double dy_,
                             const double* __restrict__ rho, int rows_per_chunk,
                             double* __restrict__ partial) {
  extern __shared__ __align__(16) unsigned char lds_raw[];
  TC* As = reinterpret_cast<TC*>(lds_raw);          // m x (kSub+1)
  TC* Bs = As + (size_t)m * (kSub + 1);              // NOPS x n x (kSub+1)
  const TC dx = (TC)dx_, dy = (TC)dy_;
  const int chunk = blockIdx.x;
  const int c0 = chunk * rows_per_chunk;
  const int c1 = min(C, c0 + rows_per_chunk);
  const int mn = m * n;
  const int tid = threadIdx.x;
  int pa[kPairsPerThread], pb[kPairsPerThread];
  TC acc[NOPS][kPairsPerThread];
  for (int k = 0; k < kPairsPerThread; ++k) {
    const int p = blockIdx.y * kPairsPerGroup + k * kBlock + tid;
    pa[k] = p < mn ? p % m : -1;
    pb[k] = p < mn ? p / m : -1;
    for (int o = 0; o < NOPS; ++o) acc[o][k] = TC(0);
  }
  for (int s0 = c0; s0 < c1; s0 += kSub) {
    const int rows = min(kSub, c1 - s0);
    for (int e = tid; e < kSub * m; e += kBlock) {
      const int rr = e % kSub, a = e / kSub;
      As[a * (kSub + 1) + rr] = rr < rows ? (TC)A[(size_t)a * lda + s0 + rr] : TC(0);
    }
    for (int e = tid; e < kSub * n; e += kBlock) {
      const int rr = e % kSub, b = e / kSub;
      const int c = s0 + rr;
      if (NOPS == 1) {
        TC v = TC(0);
        if (rr < rows)
          v = (op == OP_NONE) ? (TC)B[(size_t)b * ldb + c]
                              : stencil_value<TB, TC>(B, ldb, b, c, op, nx, ny, dx, dy, rho);
        Bs[b * (kSub + 1) + rr] = v;
      } else {
        TC v[4] = {TC(0), TC(0), TC(0), TC(0)};
        if (rr < rows && interior(c, nx, ny)) {
          const TB* col = B + (size_t)b * ldb;
          const TC vc = (TC)col[c] / (TC)rho[c];
          const TC vw = (TC)col[c - ny] / (TC)rho[c - ny], ve = (TC)col[c + ny] / (TC)rho[c + ny];
          const TC vs = (TC)col[c - 1] / (TC)rho[c - 1], vn = (TC)col[c + 1] / (TC)rho[c + 1];
          const TC h = TC(0.5);
          v[OP_L1X] = (vc - h * vw - h * ve) / dx;
          v[OP_L1Y] = (vc - h * vs - h * vn) / dy;
          v[OP_L2X] = (ve - vw) / (TC(2) * dx);
          v[OP_L2Y] = (vn - vs) / (TC(2) * dy);
        }
        for (int o = 0; o < NOPS; ++o) Bs[((size_t)o * n + b) * (kSub + 1) + rr] = v[o];
      }
    }
    __syncthreads();
    for (int k = 0; k < kPairsPerThread; ++k) {
      if (pa[k] < 0) continue;
      const TC* ar = As + pa[k] * (kSub + 1);
      for (int o = 0; o < NOPS; ++o) {
        const TC* br = Bs + ((size_t)o * n + pb[k]) * (kSub + 1);
        TC s = acc[o][k];
        for (int rr = 0; rr < kSub; ++rr) s += ar[rr] * br[rr];
        acc[o][k] = s;
      }
    }
    __syncthreads();
  }
  for (int k = 0; k < kPairsPerThread; ++k)
    if (pa[k] >= 0)
      for (int o = 0; o < NOPS; ++o)
        partial[(size_t)chunk * NOPS * mn + (size_t)o * mn + pa[k] + (size_t)pb[k] * m] =
            (double)acc[o][k];
}

// One wavefront per output: lane l sums chunks l, l+64, ...; then a fixed
// shuffle tree (deterministic order).
__global__ void k_tn_reduce(const double* __restrict__ partial, int nchunks, int mn,
                            double* __restrict__ out) {
  const int lane = threadIdx.x & 63;
  const int p = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (p >= mn) return;
  double s = 0.0;
  for (int k = lane; k < nchunks; k += 64) s += partial[(size_t)k * mn + p];
  for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
  if (lane == 0) out[p] = s;
}

// Cholesky G + s I = R^T R of a k x k SPD matrix (one workgroup, LDS) and the
// inverse of R.  Shift schedule: (s = 0 first when allow_plain), then
// s = 1e-14 tr(G) growing tenfold per failed attempt, at most 12 attempts.
// status[0] = number of attempts used, or -1 when all failed.
__global__ void k_chol_inv(const double* __restrict__ G, int k, int allow_plain,
                           double* __restrict__ R, double* __restrict__ Rinv,
                           double* __restrict__ status) {
  extern __shared__ double A[];              // k x k, column-major
  __shared__ int ok;
  __shared__ double tr;
  const int tid = threadIdx.x, nt = blockDim.x;
  if (tid == 0) {
    double t = 0.0;
    for (int i = 0; i < k; ++i) t += G[(size_t)i * k + i];
    tr = t;
  }
  __syncthreads();
  double s = allow_plain ? 0.0 : (tr > 0.0 ? 1e-14 * tr : 1e-300);
  int attempt = 0;
  for (; attempt < 13; ++attempt) {
    for (int e = tid; e < k * k; e += nt) {
      const int i = e % k, j = e / k;
      A[e] = G[e] + (i == j ? s : 0.0);
    }
    if (tid == 0) ok = 1;
    __syncthreads();
    for (int j = 0; j < k; ++j) {
      if (tid == 0) {
        const double d = A[(size_t)j * k + j];
        if (!(d > 0.0)) ok = 0;
        else A[(size_t)j * k + j] = sqrt(d);
      }
      __syncthreads();
      if (!ok) break;
      const double piv = A[(size_t)j * k + j];
      for (int c = j + 1 + tid; c < k; c += nt) A[(size_t)c * k + j] /= piv;
      __syncthreads();
      const int w = k - j - 1;                // trailing block (j, k) x (j, k)
      for (int e = tid; e < w * w; e += nt) {
        const int i = j + 1 + e % w, c = j + 1 + e / w;
        if (i <= c) A[(size_t)c * k + i] -= A[(size_t)i * k + j] * A[(size_t)c * k + j];
      }
      __syncthreads();
    }
    if (ok) break;
    __syncthreads();
    s = (s == 0.0) ? (tr > 0.0 ? 1e-14 * tr : 1e-300) : 10.0 * s;
  }
  const int good = ok;
  __syncthreads();
  for (int e = tid; e < k * k; e += nt) {
    const int i = e % k, j = e / k;
    R[e] = (i <= j) ? A[e] : 0.0;
  }
  // column c of R^-1 by back substitution (one thread per column)
  for (int c = tid; c < k; c += nt) {
    for (int i = 0; i < k; ++i) Rinv[(size_t)c * k + i] = 0.0;
    Rinv[(size_t)c * k + c] = 1.0 / A[(size_t)c * k + c];
    for (int i = c - 1; i >= 0; --i) {
      double acc = 0.0;
      for (int m = i + 1; m <= c; ++m) acc += A[(size_t)m * k + i] * Rinv[(size_t)c * k + m];
      Rinv[(size_t)c * k + i] = -acc / A[(size_t)i * k + i];
    }
  }
  if (tid == 0) status[0] = good ? (double)(attempt + 1) : -1.0;
}

template <class TA, class TC>
__global__ void k_gemm(int C, const TA* __restrict__ A, int lda, int k,
                       const double* __restrict__ T, double alpha, double beta,
                       double* __restrict__ out, int ldo, const unsigned char* __restrict__ mask) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  const int b = blockIdx.y;
  if (c >= C) return;
  const double* t = T + (size_t)b * k;
  TC acc = 0;
  for (int a = 0; a < k; ++a) acc += (TC)A[(size_t)a * lda + c] * (TC)t[a];
  double* o = out + (size_t)b * ldo + c;
  double v = alpha * (double)acc;
  if (beta != 0.0) v += beta * (*o);
  if (mask && mask[c]) v = 0.0;
  *o = v;
}

__global__ void k_gather_rows(const double* __restrict__ A, int lda, int ncols,
                              const int* __restrict__ rows, int nrows, double* __restrict__ out) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (size_t)nrows * ncols) return;
  const int r = (int)(t % nrows), k = (int)(t / nrows);
  out[t] = A[(size_t)k * lda + rows[r]];
}

__global__ void k_scatter_add(double* __restrict__ A, int lda, const int* __restrict__ rows,
                              const int* __restrict__ cols, const double* __restrict__ vals,
                              int n) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  A[(size_t)cols[t] * lda + rows[t]] += vals[t];
}

template <int NOPS, class TA, class TB, class TC>
void tn_partial_launch(hipStream_t st, dim3 grid, size_t lds, const void* A, int lda, int m,
                       const void* B, int ldb, int n, const Grid& g, int op, int rpc,
                       double* partial) {
  hipLaunchKernelGGL((k_tn_partial<NOPS, TA, TB, TC>), grid, dim3(kBlock), lds, st,
                     static_cast<const TA*>(A), lda, m, static_cast<const TB*>(B), ldb, n, g.C, op,
                     g.nx, g.ny, g.dx, g.dy, g.rho, rpc, partial);
}

template <int NOPS>
void tn_dispatch(hipStream_t st, dim3 grid, size_t lds, const void* A, DT ta, int lda, int m,
                 const void* B, DT tb, int ldb, int n, const Grid& g, int op, bool fp32, int rpc,
                 double* partial) {
  if (ta != DT::F64) throw std::runtime_error("tn: the left block is always FP64");
  if (tb == DT::F64) {
    if (fp32) tn_partial_launch<NOPS, double, double, float>(st, grid, lds, A, lda, m, B, ldb, n, g, op, rpc, partial);
    else tn_partial_launch<NOPS, double, double, double>(st, grid, lds, A, lda, m, B, ldb, n, g, op, rpc, partial);
  } else {
    if (fp32) tn_partial_launch<NOPS, double, float, float>(st, grid, lds, A, lda, m, B, ldb, n, g, op, rpc, partial);
    else tn_partial_launch<NOPS, double, float, double>(st, grid, lds, A, lda, m, B, ldb, n, g, op, rpc, partial);
  }
}

}  // namespace

// ---------------------------------------------------------------------------
void launch_bc(hipStream_t st, void* psi, DT tpsi, int C, int nq, const int* bcells, int nb,
               const double* bc_base, double scale) {
  const size_t n = (size_t)nb * nq;
  if (n == 0) return;
  if (tpsi == DT::F64)
    hipLaunchKernelGGL(k_bc<double>, dim3(blocks_for(n)), dim3(kBlock), 0, st,
                       static_cast<double*>(psi), C, nq, bcells, nb, bc_base, scale);
  else
    hipLaunchKernelGGL(k_bc<float>, dim3(blocks_for(n)), dim3(kBlock), 0, st,
                       static_cast<float*>(psi), C, nq, bcells, nb, bc_base, scale);
  check_launch("bc");
}

void launch_dose(hipStream_t st, const Grid& g, double* dose, const double* X, int ldx, int r,
                 const double* v, const void* psi, DT tpsi, int nq, const double* m0, double coef) {
  if (tpsi == DT::F64)
    hipLaunchKernelGGL(k_dose<double>, dim3(blocks_for(g.C)), dim3(kBlock), 0, st, dose, X, ldx, r,
                       v, static_cast<const double*>(psi), g.C, nq, m0, g.rho, coef);
  else
    hipLaunchKernelGGL(k_dose<float>, dim3(blocks_for(g.C)), dim3(kBlock), 0, st, dose, X, ldx, r,
                       v, static_cast<const float*>(psi), g.C, nq, m0, g.rho, coef);
  check_launch("dose");
}

void launch_sweep(hipStream_t st, const Grid& g, const void* pin, void* pout, DT tpsi, int nq,
                  const double* omega, double dE, double denom, const unsigned char* restore) {
  const size_t n = (size_t)g.C * nq;
  if (tpsi == DT::F64)
    hipLaunchKernelGGL(k_sweep<double>, dim3(blocks_for(n)), dim3(kBlock), 0, st,
                       static_cast<const double*>(pin), static_cast<double*>(pout), g.nx, g.ny, nq,
                       omega, g.dx, g.dy, g.rho, dE, denom, restore);
  else
    hipLaunchKernelGGL(k_sweep<float>, dim3(blocks_for(n)), dim3(kBlock), 0, st,
                       static_cast<const float*>(pin), static_cast<float*>(pout), g.nx, g.ny, nq,
                       omega, (float)g.dx, (float)g.dy, g.rho, (float)dE, (float)denom, restore);
  check_launch("sweep");
}

void launch_stream_k(hipStream_t st, const Grid& g, const double* X, int ldx, int r,
                     const double* Z, int n, double dE, double* K, int ldk, bool fp32) {
  if (n == 0) return;
  if (fp32)
    hipLaunchKernelGGL(k_stream_k<float>, dim3(blocks_for(g.C), n), dim3(kBlock), 0, st, X, ldx, r,
                       g.nx, g.ny, g.rho, Z, n, dE, K, ldk);
  else
    hipLaunchKernelGGL(k_stream_k<double>, dim3(blocks_for(g.C), n), dim3(kBlock), 0, st, X, ldx,
                       r, g.nx, g.ny, g.rho, Z, n, dE, K, ldk);
  check_launch("stream_K");
}

static void tn_chunks(int C, int* nchunks, int* rpc) {
  int nc = std::min(kMaxChunks, (C + kSub - 1) / kSub);
  int r = (C + nc - 1) / nc;
  r = ((r + kSub - 1) / kSub) * kSub;          // whole LDS sub-chunks
  *rpc = r;
  *nchunks = (C + r - 1) / r;
}

size_t tn_partial_size(int C, int m, int n) {
  int nc, rpc;
  tn_chunks(C, &nc, &rpc);
  return (size_t)nc * m * n * 4;       // room for the four-operator form
}

void launch_tn(hipStream_t st, const Grid& g, const void* A, DT ta, int lda, int m, const void* B,
               DT tb, int ldb, int n, int op, bool fp32, double* partial, double* out) {
  if (m == 0 || n == 0) return;
  int nchunks, rpc;
  tn_chunks(g.C, &nchunks, &rpc);
  const int mn = m * n;
  const int groups = (mn + kPairsPerGroup - 1) / kPairsPerGroup;
  const int nops = (op == OP_ALL4) ? 4 : 1;
  const size_t lds = (size_t)(m + nops * n) * (kSub + 1) * (fp32 ? sizeof(float) : sizeof(double));
  if (lds > 160 * 1024) throw std::runtime_error("tn: block too wide for LDS");
  const dim3 grid(nchunks, groups);
  if (nops == 4)
    tn_dispatch<4>(st, grid, lds, A, ta, lda, m, B, tb, ldb, n, g, op, fp32, rpc, partial);
  else
    tn_dispatch<1>(st, grid, lds, A, ta, lda, m, B, tb, ldb, n, g, op, fp32, rpc, partial);
  check_launch("tn_partial");
  const int per_block = kBlock / 64;
  const int outs = nops * mn;
  hipLaunchKernelGGL(k_tn_reduce, dim3((outs + per_block - 1) / per_block), dim3(kBlock), 0, st,
                     partial, nchunks, outs, out);
  check_launch("tn_reduce");
}

void launch_chol_inv(hipStream_t st, const double* G, int k, bool allow_plain, double* R,
                     double* Rinv, double* status) {
  if (k <= 0) return;
  if (k > 96) throw std::runtime_error("chol_inv: block wider than 96 columns");
  hipLaunchKernelGGL(k_chol_inv, dim3(1), dim3(kBlock), (size_t)k * k * sizeof(double), st, G, k,
                     allow_plain ? 1 : 0, R, Rinv, status);
  check_launch("chol_inv");
}

void launch_gemm(hipStream_t st, int C, const void* A, DT ta, int lda, int k, const double* T,
                 int n, double alpha, double beta, double* out, int ldo,
                 const unsigned char* zero_mask, bool fp32) {
  if (n == 0) return;
  const dim3 grid(blocks_for(C), n);
  if (ta == DT::F64) {
    if (fp32)
      hipLaunchKernelGGL((k_gemm<double, float>), grid, dim3(kBlock), 0, st, C,
                         static_cast<const double*>(A), lda, k, T, alpha, beta, out, ldo, zero_mask);
    else
      hipLaunchKernelGGL((k_gemm<double, double>), grid, dim3(kBlock), 0, st, C,
                         static_cast<const double*>(A), lda, k, T, alpha, beta, out, ldo, zero_mask);
  } else {
    if (fp32)
      hipLaunchKernelGGL((k_gemm<float, float>), grid, dim3(kBlock), 0, st, C,
                         static_cast<const float*>(A), lda, k, T, alpha, beta, out, ldo, zero_mask);
    else
      hipLaunchKernelGGL((k_gemm<float, double>), grid, dim3(kBlock), 0, st, C,
                         static_cast<const float*>(A), lda, k, T, alpha, beta, out, ldo, zero_mask);
  }
  check_launch("gemm");
}

void launch_gather_rows(hipStream_t st, const double* A, int lda, int ncols, const int* rows,
                        int nrows, double* out) {
  const size_t n = (size_t)nrows * ncols;
  if (n == 0) return;
  hipLaunchKernelGGL(k_gather_rows, dim3(blocks_for(n)), dim3(kBlock), 0, st, A, lda, ncols, rows,
                     nrows, out);
  check_launch("gather_rows");
}

void launch_scatter_add(hipStream_t st, double* A, int lda, const int* rows, const int* cols,
                        const double* vals, int n) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_scatter_add, dim3(blocks_for(n)), dim3(kBlock), 0, st, A, lda, rows, cols,
                     vals, n);
  check_launch("scatter_add");
}

}  // namespace dev
}  // namespace titus
