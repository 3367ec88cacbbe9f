// Device kernels of the angular half of the DLRA step (angular.hip): the
// moment operators applied to the angular basis, their Galerkin blocks, the
// K-step coefficients and the L-step.  Blocks are column-major with their row
// count as leading dimension.
#pragma once
#include <hip/hip_runtime.h>

namespace titus {
namespace dev {

constexpr int kMaxAngRows = 2048;   // moments (LDS column of one product)
constexpr int kMaxAngOps = 6;       // 2D: 4 operators, 3D: 6

// The operators A_o (o < nops, dissipative first, then central) in
// compressed rows: rowptr[o*(nm+1) + m] .. rowptr[o*(nm+1) + m + 1] index
// col/val from op_nz[o] on (device arrays).
struct AngOps {
  int nm = 0, nops = 0;
  const int* rowptr = nullptr;
  const int* col = nullptr;
  const double* val = nullptr;
  const int* op_nz = nullptr;
};

// AW[o] = A_o W (nm x k each, consecutive) and, when G is not null,
// G[o] = Wa^T AW[o] (ka x k each, consecutive).
void launch_ang_apply(hipStream_t st, const AngOps& A, const double* W, int k, const double* Wa,
                      int ka, double* AW, double* G);
// Z (r x (2 nd + 2) n): [S, S Cc, S C1+, S C1-, ...] of the streaming K-step
// from S (r x n) and G (2 nd blocks of n x n) (the stencil sums of
// SolverCSD.jl:1027 with the cell widths h).
void launch_ang_stream_z(hipStream_t st, int r, int n, int nd, double hx, double hy, double hz,
                         const double* S, const double* G, double* Z);
// L = W S^T - dE sum_{t} AW[order[t]] S^T XLX[order[t]]^T (nm x r), the L-step
// of SolverCSD.jl:1040 with XLX the 2 nd Galerkin blocks X^T L_o X (r x r).
void launch_ang_stream_l(hipStream_t st, int nm, int r, int nops, const int* order, double dE,
                         const double* W, const double* S, const double* AW, const double* XLX,
                         double* L);

}  // namespace dev
}  // namespace titus
