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
__global__ void k_bc(double* __restrict__ psi, int C, int nq, const int* __restrict__ bcells,
                     int nb, const double* __restrict__ base, double scale) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (size_t)nb * nq) return;
  const int b = (int)(t % nb), q = (int)(t / nb);
  psi[(size_t)q * C + bcells[b]] = base[(size_t)q * nb + b] * scale;
}

__global__ void k_dose(double* __restrict__ dose, const double* __restrict__ X, int ldx, int r,
                       const double* __restrict__ v, const double* __restrict__ psi, int C,
                       int nq, const double* __restrict__ m0, const double* __restrict__ rho,
                       double coef) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= C) return;
  double acc = 0.0;
  for (int a = 0; a < r; ++a) acc += X[(size_t)a * ldx + c] * v[a];
  double u0 = 0.0;
  for (int q = 0; q < nq; ++q) u0 += psi[(size_t)q * C + c] * m0[q];
  dose[c] += coef * (acc + u0) / rho[c];
}

__global__ void k_sweep(const double* __restrict__ pin, double* __restrict__ pout, int nx, int ny,
                        int nq, const double* __restrict__ omega, double dx, double dy,
                        const double* __restrict__ rho, double dE, double denom,
                        const unsigned char* __restrict__ restore) {
  const int C = nx * ny;
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (size_t)C * nq) return;
  const int c = (int)(t % C), q = (int)(t / C);
  const double* p = pin + (size_t)q * C;
  const double pc = p[c];
  double out;
  if (interior(c, nx, ny)) {
    const double ox = omega[q], oz = omega[2 * nq + q];
    const double vc = pc / rho[c];
    const double fx = (ox >= 0.0) ? (vc - p[c - ny] / rho[c - ny]) : (p[c + ny] / rho[c + ny] - vc);
    const double fy = (oz >= 0.0) ? (vc - p[c - 1] / rho[c - 1]) : (p[c + 1] / rho[c + 1] - vc);
    const double flux = fx * (ox / dx) + fy * (oz / dy);
    out = (pc - dE * flux) / denom;
  } else {
    out = (q == 0 && restore[c]) ? pc : pc / denom;
  }
  pout[t] = out;
}

// K[c,b] = X[c,:] Z0[:,b] - dE * sum_d (X[d,:]/rho[d]) Zd[:,b] (interior c)
__global__ void k_stream_k(const double* __restrict__ X, int ldx, int r, int nx, int ny,
                           const double* __restrict__ rho, const double* __restrict__ Z, int n,
                           double dE, double* __restrict__ K, int ldk) {
  const int C = nx * ny;
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  const int b = blockIdx.y;
  if (c >= C) return;
  const double* z0 = Z + (size_t)b * r;
  double acc0 = 0.0;
  for (int a = 0; a < r; ++a) acc0 += X[(size_t)a * ldx + c] * z0[a];
  double out = acc0;
  if (interior(c, nx, ny)) {
    const double* zc = Z + (size_t)(1 * n + b) * r;
    const double* ze = Z + (size_t)(2 * n + b) * r;
    const double* zw = Z + (size_t)(3 * n + b) * r;
    const double* zn = Z + (size_t)(4 * n + b) * r;
    const double* zs = Z + (size_t)(5 * n + b) * r;
    double sc = 0.0, se = 0.0, sw = 0.0, sn = 0.0, ss = 0.0;
    for (int a = 0; a < r; ++a) {
      const double* xa = X + (size_t)a * ldx;
      sc += xa[c] * zc[a];
      se += xa[c + ny] * ze[a];
      sw += xa[c - ny] * zw[a];
      sn += xa[c + 1] * zn[a];
      ss += xa[c - 1] * zs[a];
    }
    const double s = sc / rho[c] + se / rho[c + ny] + sw / rho[c - ny] + sn / rho[c + 1] +
                     ss / rho[c - 1];
    out = acc0 - dE * s;
  }
  K[(size_t)b * ldk + c] = out;
}

__device__ inline double stencil_value(const double* __restrict__ B, int ldb, int b, int c, int op,
                                       int nx, int ny, double dx, double dy,
                                       const double* __restrict__ rho) {
  if (!interior(c, nx, ny)) return 0.0;
  const double* col = B + (size_t)b * ldb;
  switch (op) {
    case OP_L1X:
      return (col[c] / rho[c] - 0.5 * (col[c - ny] / rho[c - ny]) -
              0.5 * (col[c + ny] / rho[c + ny])) / dx;
    case OP_L1Y:
      return (col[c] / rho[c] - 0.5 * (col[c - 1] / rho[c - 1]) -
              0.5 * (col[c + 1] / rho[c + 1])) / dy;
    case OP_L2X:
      return (col[c + ny] / rho[c + ny] - col[c - ny] / rho[c - ny]) / (2.0 * dx);
    default:
      return (col[c + 1] / rho[c + 1] - col[c - 1] / rho[c - 1]) / (2.0 * dy);
  }
}

// partial[chunk][p] = sum over the chunk's rows of A[c, a] * op(B)[c, b], p = a + b*m
__global__ void k_tn_partial(const double* __restrict__ A, int lda, int m,
                             const double* __restrict__ B, int ldb, int n, int C, int op, int nx,
                             int ny, double dx, double dy, const double* __restrict__ rho,
                             int rows_per_chunk, double* __restrict__ partial) {
  extern __shared__ double lds[];
  double* As = lds;                          // m x (kSub+1)
  double* Bs = lds + (size_t)m * (kSub + 1); // n x (kSub+1)
  const int chunk = blockIdx.x;
  const int c0 = chunk * rows_per_chunk;
  const int c1 = min(C, c0 + rows_per_chunk);
  const int mn = m * n;
  const int tid = threadIdx.x;
  int pa[kPairsPerThread], pb[kPairsPerThread];
  double acc[kPairsPerThread];
  for (int k = 0; k < kPairsPerThread; ++k) {
    const int p = blockIdx.y * kPairsPerGroup + k * kBlock + tid;
    pa[k] = p < mn ? p % m : -1;
    pb[k] = p < mn ? p / m : -1;
    acc[k] = 0.0;
  }
  for (int s0 = c0; s0 < c1; s0 += kSub) {
    const int rows = min(kSub, c1 - s0);
    for (int e = tid; e < kSub * m; e += kBlock) {
      const int rr = e % kSub, a = e / kSub;
      As[a * (kSub + 1) + rr] = rr < rows ? A[(size_t)a * lda + s0 + rr] : 0.0;
    }
    for (int e = tid; e < kSub * n; e += kBlock) {
      const int rr = e % kSub, b = e / kSub;
      double v = 0.0;
      if (rr < rows) {
        const int c = s0 + rr;
        v = (op == OP_NONE) ? B[(size_t)b * ldb + c]
                            : stencil_value(B, ldb, b, c, op, nx, ny, dx, dy, rho);
      }
      Bs[b * (kSub + 1) + rr] = v;
    }
    __syncthreads();
    for (int k = 0; k < kPairsPerThread; ++k) {
      if (pa[k] < 0) continue;
      const double* ar = As + pa[k] * (kSub + 1);
      const double* br = Bs + pb[k] * (kSub + 1);
      double s = acc[k];
      for (int rr = 0; rr < kSub; ++rr) s += ar[rr] * br[rr];
      acc[k] = s;
    }
    __syncthreads();
  }
  for (int k = 0; k < kPairsPerThread; ++k)
    if (pa[k] >= 0) partial[(size_t)chunk * mn + pa[k] + (size_t)pb[k] * m] = acc[k];
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

__global__ void k_gemm(int C, const double* __restrict__ A, int lda, int k,
                       const double* __restrict__ T, double alpha, double beta,
                       double* __restrict__ out, int ldo, const unsigned char* __restrict__ mask) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  const int b = blockIdx.y;
  if (c >= C) return;
  const double* t = T + (size_t)b * k;
  double acc = 0.0;
  for (int a = 0; a < k; ++a) acc += A[(size_t)a * lda + c] * t[a];
  double* o = out + (size_t)b * ldo + c;
  double v = alpha * acc;
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

}  // namespace

// ---------------------------------------------------------------------------
void launch_bc(hipStream_t st, double* psi, int C, int nq, const int* bcells, int nb,
               const double* bc_base, double scale) {
  const size_t n = (size_t)nb * nq;
  if (n == 0) return;
  hipLaunchKernelGGL(k_bc, dim3(blocks_for(n)), dim3(kBlock), 0, st, psi, C, nq, bcells, nb,
                     bc_base, scale);
  check_launch("bc");
}

void launch_dose(hipStream_t st, const Grid& g, double* dose, const double* X, int ldx, int r,
                 const double* v, const double* psi, int nq, const double* m0, double coef) {
  hipLaunchKernelGGL(k_dose, dim3(blocks_for(g.C)), dim3(kBlock), 0, st, dose, X, ldx, r, v, psi,
                     g.C, nq, m0, g.rho, coef);
  check_launch("dose");
}

void launch_sweep(hipStream_t st, const Grid& g, const double* pin, double* pout, int nq,
                  const double* omega, double dE, double denom, const unsigned char* restore) {
  const size_t n = (size_t)g.C * nq;
  hipLaunchKernelGGL(k_sweep, dim3(blocks_for(n)), dim3(kBlock), 0, st, pin, pout, g.nx, g.ny,
                     nq, omega, g.dx, g.dy, g.rho, dE, denom, restore);
  check_launch("sweep");
}

void launch_stream_k(hipStream_t st, const Grid& g, const double* X, int ldx, int r,
                     const double* Z, int n, double dE, double* K, int ldk) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_stream_k, dim3(blocks_for(g.C), n), dim3(kBlock), 0, st, X, ldx, r, g.nx,
                     g.ny, g.rho, Z, n, dE, K, ldk);
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
  return (size_t)nc * m * n;
}

void launch_tn(hipStream_t st, const Grid& g, const double* A, int lda, int m, const double* B,
               int ldb, int n, int op, double* partial, double* out) {
  if (m == 0 || n == 0) return;
  int nchunks, rpc;
  tn_chunks(g.C, &nchunks, &rpc);
  const int mn = m * n;
  const int groups = (mn + kPairsPerGroup - 1) / kPairsPerGroup;
  const size_t lds = (size_t)(m + n) * (kSub + 1) * sizeof(double);
  if (lds > 160 * 1024) throw std::runtime_error("tn: block too wide for LDS");
  hipLaunchKernelGGL(k_tn_partial, dim3(nchunks, groups), dim3(kBlock), lds, st, A, lda, m, B,
                     ldb, n, g.C, op, g.nx, g.ny, g.dx, g.dy, g.rho, rpc, partial);
  check_launch("tn_partial");
  const int per_block = kBlock / 64;
  hipLaunchKernelGGL(k_tn_reduce, dim3((mn + per_block - 1) / per_block), dim3(kBlock), 0, st,
                     partial, nchunks, mn, out);
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

void launch_gemm(hipStream_t st, int C, const double* A, int lda, int k, const double* T, int n,
                 double alpha, double beta, double* out, int ldo, const unsigned char* zero_mask) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_gemm, dim3(blocks_for(C), n), dim3(kBlock), 0, st, C, A, lda, k, T, alpha,
                     beta, out, ldo, zero_mask);
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

void copy_block(hipStream_t st, double* dst, int ldd, const double* src, int lds, int m, int n) {
  if (m == 0 || n == 0) return;
  hipError_t e = hipMemcpy2DAsync(dst, (size_t)ldd * sizeof(double), src,
                                  (size_t)lds * sizeof(double), (size_t)m * sizeof(double), n,
                                  hipMemcpyDeviceToDevice, st);
  if (e != hipSuccess)
    throw std::runtime_error(std::string("copy_block: ") + hipGetErrorString(e));
}

}  // namespace dev
}  // namespace titus
