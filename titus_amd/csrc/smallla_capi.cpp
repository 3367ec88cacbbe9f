// C interface to the host dense LA for unit tests (host only, no GPU).
// Not part of the product ABI (include/titus_hip.h).
#include <cstring>
#include <exception>
#include <vector>

#include "smallla.h"

using titus::HMat;

extern "C" {

// Thin SVD of an m x n column-major matrix: U m x k, s k, V n x k, k = min(m,n).
int smallla_svd(int m, int n, const double* A, double* U, double* s, double* V) {
  try {
    HMat Am(m, n), Um, Vm;
    std::memcpy(Am.a.data(), A, sizeof(double) * m * n);
    std::vector<double> sv;
    titus::svd(Am, Um, sv, Vm);
    std::memcpy(U, Um.a.data(), sizeof(double) * Um.a.size());
    std::memcpy(V, Vm.a.data(), sizeof(double) * Vm.a.size());
    std::memcpy(s, sv.data(), sizeof(double) * sv.size());
    return 0;
  } catch (const std::exception&) {
    return -1;
  }
}

// The same SVD without right vectors (U m x k, s k).
int smallla_svd_left(int m, int n, const double* A, double* U, double* s) {
  try {
    HMat Am(m, n), Um, Vm;
    std::memcpy(Am.a.data(), A, sizeof(double) * m * n);
    std::vector<double> sv;
    titus::svd(Am, Um, sv, Vm, false);
    std::memcpy(U, Um.a.data(), sizeof(double) * Um.a.size());
    std::memcpy(s, sv.data(), sizeof(double) * sv.size());
    return 0;
  } catch (const std::exception&) {
    return -1;
  }
}

// range_basis (see smallla.h); against may be null (q = 0); candidates may be
// null (all rows).  Writes m x ncols to out and the rank to *rank.
int smallla_range_basis(int m, int k, const double* A, int ncols, int q, const double* against,
                        double scale, double tol, int greedy, int ncand, const int* cand,
                        double* out, int* rank) {
  try {
    HMat Am(m, k);
    std::memcpy(Am.a.data(), A, sizeof(double) * m * k);
    HMat Xa(m, q);
    if (q > 0) std::memcpy(Xa.a.data(), against, sizeof(double) * m * q);
    std::vector<int> c;
    if (cand) c.assign(cand, cand + ncand);
    HMat B = titus::range_basis(Am, ncols, q > 0 ? &Xa : nullptr, scale, tol, greedy != 0, c, rank);
    std::memcpy(out, B.a.data(), sizeof(double) * B.a.size());
    return 0;
  } catch (const std::exception&) {
    return -1;
  }
}

int smallla_svd_completed(int n, const double* S, double tol, double* U, double* d, double* V) {
  try {
    HMat Sm(n, n), Um, Vm;
    std::memcpy(Sm.a.data(), S, sizeof(double) * n * n);
    std::vector<double> dv;
    titus::svd_completed(Sm, tol, Um, dv, Vm);
    std::memcpy(U, Um.a.data(), sizeof(double) * n * n);
    std::memcpy(V, Vm.a.data(), sizeof(double) * n * n);
    std::memcpy(d, dv.data(), sizeof(double) * n);
    return 0;
  } catch (const std::exception&) {
    return -1;
  }
}

}  // extern "C"
