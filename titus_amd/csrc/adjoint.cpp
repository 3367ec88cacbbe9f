// Discrete adjoint of the fixed-rank energy loop (include/titus_hip.h:
// titus_set_checkpoints, titus_adjoint, titus_adjoint_batch).
//
// The reference has no adjoint.  What is differentiated is this library's
// forward map of SolveFirstCollisionSourceDLR (code/SolverCSD.jl:891-1120):
// the dose after the solved steps as a function of the density rho, with the
// energy grid, the beam source and the P_N operators held fixed.  rho enters
// as y = 1/rho in three places: the streaming stencils L_o(U) = Lhat_o (y o U)
// (code/stencils.jl:26-128), the upwind flux of y o psi (SolverCSD.jl:388-463)
// and the dose, dose += a y o (X S W[0,:]^T + psi M[0,:]^T) (:990-997).
//
// Checkpointing: with titus_set_checkpoints(every), the forward solve keeps
// (X, psi on the device; S, W on the host) before steps 1, 1 + every, ...
// titus_adjoint walks the segments backwards: it restores a checkpoint,
// recomputes the segment's steps once with the step tape on (the same
// kernels, so the same values), then runs the recorded steps in reverse.
// Device memory: one checkpoint of C (r + nq) doubles per segment, plus
// `every` step tapes of C (5 r + 2 nq) doubles.
//
// Basis updates.  X1 = range_basis(K) spans R = span(U_p) + span(E) (kept
// singular vectors, chosen canonical vectors); the downstream solve depends on
// X1 only through R, so the adjoint of the update is
//     Kbar = dU_p^*( (I - X1 X1^T) X1bar Tp^T )
// with dU_p the derivative of the kept singular subspace (its part outside
// range(K) and its coupling to the pairs dropped by the cut).  The rank
// decisions (p, chosen vectors) are locally constant.  The same rule runs on
// the host for the angular bases (titus::orth_bar).
//
// Scope: FP64, fixed-rank integrator, 2D grids.
#include "context.h"

namespace {

struct Checkpoint {
  int s = 0;
  bool saved = false;
  HMat S, W;
  double* X = nullptr;
  double* psi = nullptr;
};

}  // namespace

struct AdjointStore {
  int every = 0;
  std::vector<Checkpoint> cps;
  std::vector<StepTape> tapes;
  std::vector<double*> owned;
  std::vector<titus::SpMat> sAt;       // transposes of the angular operators
  // adjoint state and scratch (C x r unless noted)
  double *Xb = nullptr, *XmB = nullptr, *Kb = nullptr, *T1 = nullptr, *T2 = nullptr;
  double *K0 = nullptr, *Ub = nullptr;
  double *psib = nullptr, *psib2 = nullptr;                 // C x nq
  double *ybar = nullptr, *Db = nullptr, *gbuf = nullptr;   // C
  // forward state saved across the recomputation
  double *Xsave = nullptr, *psisave = nullptr, *dosesave = nullptr, *u0save = nullptr;

  double* alloc(size_t n) {
    double* p = dalloc<double>(n);
    owned.push_back(p);
    return p;
  }
  // one device block carved into buffers of the given sizes, each rounded
  // up to 64 doubles (512-byte alignment, as a block of its own has): a
  // context with checkpoints and tapes holds a few blocks instead of
  // hundreds, which keeps its creation and destruction cheap (every device
  // free waits for the device)
  std::vector<double*> alloc_many(const std::vector<size_t>& sizes) {
    size_t total = 0;
    for (size_t n : sizes) total += (n + 63) & ~size_t(63);
    double* base = alloc(total);
    std::vector<double*> out;
    size_t off = 0;
    for (size_t n : sizes) {
      out.push_back(base + off);
      off += (n + 63) & ~size_t(63);
    }
    return out;
  }
  ~AdjointStore() {
    for (double* p : owned) (void)hipFree(p);
  }
};

// Device slots of every checkpoint, allocated when checkpointing is enabled
// (so the solve itself allocates nothing); slot i holds the state before
// step 1 + i * every.
void adjoint_reserve(titus_ctx& x) {
  AdjointStore& a = *x.adj;
  const int n = (x.ne - 2) / a.every + 1;                 // steps 1 .. ne-1
  a.cps.resize(n);
  std::vector<size_t> sizes;
  for (int i = 0; i < n; ++i) {
    sizes.push_back((size_t)x.C * x.rank);
    sizes.push_back((size_t)x.C * x.nq);
  }
  const std::vector<double*> b = a.alloc_many(sizes);
  for (int i = 0; i < n; ++i) {
    a.cps[i].s = 1 + i * a.every;
    a.cps[i].X = b[2 * i];
    a.cps[i].psi = b[2 * i + 1];
  }
}

int adjoint_phase_key(const titus_ctx& x) {
  if (!x.adj || x.adj->every <= 0 || x.steps_done < 1) return 0;
  return (x.steps_done - 1) % x.adj->every + 1;
}

void adjoint_checkpoint(titus_ctx& x, int s) {
  AdjointStore& a = *x.adj;
  if (x.tape || (s - 1) % a.every != 0) return;
  const size_t i = (size_t)((s - 1) / a.every);
  if (i >= a.cps.size()) throw TitusError(TITUS_E_STATE, "adjoint: checkpoint beyond the grid");
  Checkpoint& c = a.cps[i];
  c.S = x.S;
  c.W = x.W;
  c.saved = true;
  x.dcopy(c.X, x.X, sizeof(double) * (size_t)x.C * x.rank);
  x.dcopy(c.psi, x.psi, sizeof(double) * (size_t)x.C * x.nq);
}

namespace {

HMat scaled(HMat A, double f) {
  for (double& v : A.a) v *= f;
  return A;
}
void add_to(HMat& A, const HMat& B, double f = 1.0) {
  for (size_t i = 0; i < A.a.size(); ++i) A.a[i] += f * B.a[i];
}
HMat mm(const HMat& A, const HMat& B, bool ta = false, bool tb = false) {
  return titus::matmul(A, B, ta, tb);
}
HMat rows_scaled(HMat A, const std::vector<double>& d) {
  for (int j = 0; j < A.n; ++j)
    for (int i = 0; i < A.m; ++i) A(i, j) *= d[i];
  return A;
}

// Reverse sweep over one context.
struct Reverse {
  titus_ctx& x;
  AdjointStore& a;
  const int C, r, nq, nm;
  HMat Sb, Wb;

  Reverse(titus_ctx& ctx, AdjointStore& store)
      : x(ctx), a(store), C(ctx.C), r(ctx.rank), nq(ctx.nq), nm(ctx.nm), Sb(r, r), Wb(nm, r) {}

  // ---- device helpers -------------------------------------------------------------
  // out = alpha A T + beta out (A: C x k, T: k x n on the host)
  void gemm(const double* A, int k, const HMat& T, double alpha, double beta, double* out) {
    x.gemm(A, k, T, alpha, beta, out, nullptr);
  }
  // arena offset of A^T B (C-row blocks)
  size_t tn(const double* A, int m, const double* B, int n, int op = dev::OP_NONE) {
    return x.tn(A, m, B, n, op);
  }
  void zero_bcells(double* A, int ncols) {
    dev::launch_zero_rows(x.st, A, C, ncols, x.bcells, x.nb);
  }
  // out (C x r) = sum_o L_o(U) B_o (fwd) or sum_o (Lhat_o^T U) B_o (transposed; no y)
  void stencil_blocks(const double* U, const std::vector<HMat>& B, bool transposed, double* out) {
    const int n = B[0].n;
    HMat Z = block_coefficients(B, transposed);
    const double* Zd = x.upload(Z);
    if (transposed) dev::launch_stream_kt(x.st, x.g, U, C, r, Zd, n, -1.0, out, C);
    else dev::launch_stream_k(x.st, x.g, U, C, r, Zd, n, -1.0, out, C, false);
  }
  // coefficient blocks [0, centre, x+, x-, y+, y-] of sum_o (stencil o) B_o; the
  // transposed stencils swap the neighbour roles of the central operators
  HMat block_coefficients(const std::vector<HMat>& B, bool transposed) const {
    const int n = B[0].n, nd = x.nd;
    const double h[3] = {x.dx, x.dy, x.dz};
    const int nb = 1 + 2 * nd;
    HMat Z(r, (nb + 1) * n);
    const double sgn = transposed ? -1.0 : 1.0;
    for (size_t i = 0; i < (size_t)r * n; ++i) {
      double c = B[0].a[i] / h[0];
      for (int d = 1; d < nd; ++d) c += B[d].a[i] / h[d];
      Z.a[(size_t)r * n + i] = c;
      for (int d = 0; d < nd; ++d) {
        const double m1 = B[d].a[i], m2 = sgn * B[nd + d].a[i];
        Z.a[(size_t)(2 + 2 * d) * r * n + i] = (m2 - m1) / (2.0 * h[d]);
        Z.a[(size_t)(3 + 2 * d) * r * n + i] = (-m2 - m1) / (2.0 * h[d]);
      }
    }
    return Z;
  }
  // the two Galerkin terms of XLX_o = U^T L_o(U) with adjoint XLb_o:
  // Ub += sum_o L_o(U) XLb_o^T + y o sum_o (Lhat_o^T U) XLb_o;
  // ybar += rowsum(U o sum_o (Lhat_o^T U) XLb_o)
  void galerkin_x_rev(const double* U, const std::vector<HMat>& XLb, double* Ubar) {
    std::vector<HMat> Bt;
    for (const HMat& b : XLb) Bt.push_back(titus::transpose(b));
    stencil_blocks(U, Bt, false, a.T1);
    dev::launch_axpy(x.st, (size_t)C * r, 1.0, a.T1, Ubar);
    stencil_blocks(U, XLb, true, a.T2);
    dev::launch_adj_combine(x.st, C, r, Ubar, C, a.T2, C, U, C, x.g.rho, 1.0, a.ybar);
  }
  // Kout (C x k) = adjoint of K for X1 = range_basis(K) given X1bar (device)
  void orth_bar_dev(const SpatialOrthTape& o, const double* X1, const double* X1bar, double* Kout) {
    const int k = o.k, p = o.p;
    if (p == 0) {
      // a zero block by a zero-coefficient product (no memset on the stream)
      gemm(X1bar, r, HMat(r, k), 0.0, 0.0, Kout);
      return;
    }
    gemm(X1bar, r, titus::transpose(o.Tp), 1.0, 0.0, a.Ub);            // Z = X1bar Tp^T
    x.areset();
    const size_t t1 = tn(X1, r, a.Ub, p), t2 = tn(o.Q2, k, a.Ub, p), t3 = tn(o.Q2, k, X1, r);
    x.fetch();
    const HMat T1 = x.view(t1, r, p), T2 = x.view(t2, k, p), T3 = x.view(t3, k, r);
    gemm(X1, r, T1, -1.0, 1.0, a.Ub);                                   // Ubar = Z - X1 (X1^T Z)
    // h = Q3^T Ubar = R3inv^T (Q2^T Z - (Q2^T X1)(X1^T Z))
    HMat q = T2;
    add_to(q, mm(T3, T1), -1.0);
    const HMat hq = mm(o.R3inv, q, true, false);                        // k x p
    HMat SV(p, k);                                                      // Sigma_p^-1 V_p^T
    for (int i = 0; i < p; ++i)
      for (int j = 0; j < k; ++j) SV(i, j) = o.V(j, i) / o.sv[i];
    HMat Mk = scaled(mm(hq, SV), -1.0);                                 // k x k
    if (k > p) {
      const HMat beta = mm(o.P, hq, true, false);                       // U^T Ubar (k x p)
      HMat B1(k - p, p), B2(p, k - p);
      for (int i = 0; i < p; ++i)
        for (int j = p; j < k; ++j) {
          const double si = o.sv[i], sj = o.sv[j];
          const double f = beta(j, i) / (si * si - sj * sj);
          B1(j - p, i) = f * si;
          B2(i, j - p) = f * sj;
        }
      const HMat Pd = titus::cols(o.P, p, k), Pp = titus::cols(o.P, 0, p);
      const HMat Vp = titus::cols(o.V, 0, p), Vd = titus::cols(o.V, p, k);
      add_to(Mk, mm(mm(Pd, B1), Vp, false, true));
      add_to(Mk, mm(mm(Pp, B2), Vd, false, true));
    }
    gemm(a.Ub, p, SV, 1.0, 0.0, Kout);
    gemm(o.Q2, k, mm(o.R3inv, Mk), 1.0, 1.0, Kout);
  }
  // W^T A_o W with adjoint WAb: Wbar += A_o W WAb^T + A_o^T W WAb
  void galerkin_w_rev(const HMat& W, int o, const HMat& WAb, HMat& Wbar) {
    add_to(Wbar, mm(titus::spmm(x.sA[o], W), WAb, false, true));
    add_to(Wbar, mm(titus::spmm(a.sAt[o], W), WAb));
  }
  // transpose of dose += a y (X S W[0,:]^T + psi m0)
  void dose_rev(double coef, const double* X, const HMat& S, const HMat& W, const double* psi) {
    std::vector<double> v(r, 0.0);
    for (int i = 0; i < r; ++i)
      for (int b = 0; b < r; ++b) v[i] += S(i, b) * W(0, b);
    const double* vd = x.upload_raw(v.data(), v.size());
    dev::launch_dose_adj(x.st, x.g, a.Db, coef, X, C, r, vd, psi, nq, x.m0, a.Xb, a.psib, a.ybar,
                         a.gbuf);
    x.areset();
    const size_t off = tn(X, r, a.gbuf, 1);
    x.fetch();
    const HMat Xtg = x.view(off, r, 1);
    for (int i = 0; i < r; ++i)
      for (int j = 0; j < r; ++j) Sb(i, j) += Xtg(i, 0) * W(0, j);
    for (int j = 0; j < r; ++j) {
      double t = 0.0;
      for (int i = 0; i < r; ++i) t += S(i, j) * Xtg(i, 0);
      Wb(0, j) += t;
    }
  }

  void step(const StepTape& t) {
    const int s = t.s;
    const double dE = x.dE;
    const std::vector<int> order = x.term_order();
    const double* Xs = s > 1 ? t.X1 : t.X;           // X after streaming (ghost rows are 0)
    // ---- second dose half
    dose_rev(0.5 * dE * x.stopping[s], t.X3, t.S5, t.W3, t.psi_b);
    // ---- S5 = MUp3 S3 NUp3^T + dE A3 B3, A3 = X3^T psi_b, B3 = M^T (D o W3)
    const HMat S4b = Sb;
    const HMat A3b = scaled(mm(Sb, t.B3, false, true), dE);             // r x nq
    const HMat B3b = scaled(mm(t.A3, Sb, true, false), dE);             // nq x r
    dev::launch_gemm(x.st, C, t.psi_b, dev::DT::F64, C, nq, x.upload(titus::transpose(A3b)), r, 1.0,
                     1.0, a.Xb, C, nullptr, false);                     // X3bar += psi_b A3b^T
    gemm(t.X3, r, A3b, 1.0, 1.0, a.psib);                               // psib += X3 A3b
    HMat W3b = Wb;
    add_to(W3b, rows_scaled(mm(x.M, B3b), t.Dvec));
    const HMat MUpb = mm(mm(S4b, t.NUp3), t.S3, false, true);
    HMat S3b = mm(mm(t.MUp3, S4b, true, false), t.NUp3);
    const HMat NUpb = mm(mm(S4b, t.MUp3, true, false), t.S3);
    add_to(W3b, mm(t.W2, NUpb, false, true));
    HMat W2b = mm(t.W3, NUpb);
    gemm(Xs, r, titus::transpose(MUpb), 1.0, 1.0, a.Xb);               // X3bar += Xm MUpb^T
    gemm(t.X3, r, MUpb, 1.0, 0.0, a.XmB);                               // Xmbar = X3 MUpb
    // W3 = range(Lin), Lin = W2 S3^T + dE D o (M psi_b^T Xm)
    const HMat Linb = titus::orth_bar(t.Rw3, t.W3, W3b);
    add_to(W2b, mm(Linb, t.S3));
    add_to(S3b, mm(Linb, t.W2, true, false));
    const HMat Pb = mm(x.M, scaled(rows_scaled(Linb, t.Dvec), dE), true, false);   // nq x r
    gemm(Xs, r, titus::transpose(Pb), 1.0, 1.0, a.psib);               // psib += Xm Pb^T
    dev::launch_gemm(x.st, C, t.psi_b, dev::DT::F64, C, nq, x.upload(Pb), r, 1.0, 1.0, a.XmB, C,
                     nullptr, false);                                   // Xmbar += psi_b Pb
    // X3 = range(Kin), Kin = ghost-masked (Xm S3 + dE psi_b MtDW)
    orth_bar_dev(t.Rx3, t.X3, a.Xb, a.Kb);
    zero_bcells(a.Kb, r);
    gemm(a.Kb, r, titus::transpose(t.S3), 1.0, 1.0, a.XmB);            // Xmbar += Kinb S3^T
    gemm(a.Kb, r, scaled(titus::transpose(t.MtDW), dE), 1.0, 1.0, a.psib);
    x.areset();
    const size_t o1 = tn(Xs, r, a.Kb, r), o2 = tn(a.Kb, r, t.psi_b, nq);
    x.fetch();
    add_to(S3b, x.view(o1, r, r));
    const HMat MtDWb = scaled(titus::transpose(x.view(o2, r, nq)), dE);  // nq x r
    add_to(W2b, rows_scaled(mm(x.M, MtDWb), t.Dvec));
    std::swap(a.Xb, a.XmB);
    zero_bcells(a.Xb, r);                                               // adjoint of Xs
    // ---- out-scattering: S3 = Lo^T W2, W2 = range(Lo), Lo = (Ws Ss^T) / (1 + dE Dout)
    HMat Lob = mm(t.W2, S3b, false, true);
    add_to(W2b, mm(t.Lo, S3b));
    add_to(Lob, titus::orth_bar(t.Rw2, t.W2, W2b));
    std::vector<double> dinv(nm);
    for (int i = 0; i < nm; ++i) dinv[i] = 1.0 / (1.0 + dE * t.Dout[i]);
    const HMat WSb = rows_scaled(Lob, dinv);
    Wb = mm(WSb, t.Ss);
    Sb = mm(WSb, t.Ws, true, false);
    if (s > 1) {
      // ---- S-step: S2 = S1 - dE sum_o XLX1_o S1 WAW1_o
      const HMat S2b = Sb;
      HMat S1b = S2b, W1b = Wb;
      std::vector<HMat> XLb(x.nops());
      for (int o : order) {
        const HMat& XL = t.XLX1[o];
        const HMat& WA = t.WAW1[o];
        add_to(S1b, mm(mm(XL, S2b, true, false), WA, false, true), -dE);
        XLb[o] = scaled(mm(mm(S2b, WA, false, true), t.S1, false, true), -dE);
        const HMat WAb = scaled(mm(mm(t.S1, XL, true, true), S2b), -dE);
        galerkin_w_rev(t.W1, o, WAb, W1b);
      }
      galerkin_x_rev(t.X1, XLb, a.Xb);
      // S1 = MUp1 S NUp1^T, MUp1 = X1^T X, NUp1 = W1^T W
      const HMat MUpb1 = mm(mm(S1b, t.NUp1), t.S, false, true);
      Sb = mm(mm(t.MUp1, S1b, true, false), t.NUp1);
      const HMat NUpb1 = mm(mm(S1b, t.MUp1, true, false), t.S);
      add_to(W1b, mm(t.W, NUpb1, false, true));
      Wb = mm(t.W1, NUpb1);
      gemm(t.X, r, titus::transpose(MUpb1), 1.0, 1.0, a.Xb);          // X1bar += X MUpb^T
      gemm(t.X1, r, MUpb1, 1.0, 0.0, a.XmB);                          // Xbar = X1 MUpb
      // L-step: W1 = range(L1), L1 = L0 - dE sum_o A_o L0 XLX_o^T, L0 = W S^T
      const HMat L1b = titus::orth_bar(t.Rw1, t.W1, W1b);
      HMat L0b = L1b;
      std::vector<HMat> XLb0(x.nops());
      for (int o : order) {
        add_to(L0b, mm(titus::spmm(a.sAt[o], L1b), t.XLX[o]), -dE);
        XLb0[o] = scaled(mm(L1b, titus::spmm(x.sA[o], t.L0), true, false), -dE);
      }
      galerkin_x_rev(t.X, XLb0, a.XmB);
      add_to(Wb, mm(L0b, t.S));
      add_to(Sb, mm(L0b, t.W, true, false));
      // K-step: X1 = range(K1), K1 = K0 - dE sum_o L_o(K0) WAW_o, K0 = X S
      orth_bar_dev(t.Rx1, t.X1, a.Xb, a.Kb);
      gemm(t.X, r, t.S, 1.0, 0.0, a.K0);
      x.areset();
      const size_t og = tn(a.Kb, r, a.K0, r, dev::OP_ALL);             // K1b^T L_o(K0)
      std::vector<HMat> WAt;
      for (const HMat& w : t.WAW) WAt.push_back(titus::transpose(w));
      stencil_blocks(a.Kb, WAt, true, a.T2);
      dev::launch_adj_combine(x.st, C, r, a.Kb, C, a.T2, C, a.K0, C, x.g.rho, -dE, a.ybar);
      x.fetch();
      for (int o : order) {
        const HMat WAb = scaled(titus::transpose(x.view(og + (size_t)o * r * r, r, r)), -dE);
        galerkin_w_rev(t.W, o, WAb, Wb);
      }
      gemm(a.Kb, r, titus::transpose(t.S), 1.0, 1.0, a.XmB);          // Xbar += K0b S^T
      x.areset();
      const size_t ok = tn(t.X, r, a.Kb, r);
      x.fetch();
      add_to(Sb, x.view(ok, r, r));
      std::swap(a.Xb, a.XmB);
      zero_bcells(a.Xb, r);
    }
    // ---- uncollided sweep and first dose half
    dev::launch_sweep_adj(x.st, x.g, a.psib, t.psi_a, nq, x.omega, dE, 1.0 + dE * t.sig0, x.restore,
                          a.psib2, a.ybar);
    std::swap(a.psib, a.psib2);
    dose_rev(0.5 * dE * x.stopping[s - 1], t.X, t.S, t.W, t.psi_a);
    zero_bcells(a.psib, nq);                                            // boundary overwrite
  }
};

// The forward state an adjoint overwrites and puts back.
struct AdjointRun {
  int done = 0;
  HMat S_end, W_end;
  uint64_t hash_was = 0;
  int64_t decisions_was = 0;
  bool timing_was = false;
  size_t cu = 0;
  bool begun = false;
};

// Checks, buffers, the saved forward state, the seed and zeroed adjoint
// blocks, on the context's own stream (before any shared launches).
void adjoint_begin(titus_ctx& x, const double* dose_bar, AdjointRun& run) {
  if (!x.adj) throw TitusError(TITUS_E_STATE, "adjoint: checkpoints were not enabled before the solve");
  AdjointStore& a = *x.adj;
  if (x.steps_done < 1) throw TitusError(TITUS_E_STATE, "adjoint: nothing solved yet");
  if (a.cps.empty() || !a.cps.front().saved)
    throw TitusError(TITUS_E_STATE, "adjoint: no checkpoint before the first step");
  for (int c = 0; c < x.C; ++c)
    if (!std::isfinite(dose_bar[c])) throw TitusError(TITUS_E_ARG, "adjoint: dose_bar must be finite");
  const int C = x.C, r = x.rank, nq = x.nq;
  const size_t cr = (size_t)C * r, cq = (size_t)C * nq;
  // buffers (first call)
  if (!a.Xb) {
    const std::vector<double*> b = a.alloc_many(
        {cr, cr, cr, cr, cr, cr, cr, cq, cq, (size_t)C, (size_t)C, (size_t)C, cr, cq, (size_t)C,
         (size_t)dev::kMaxSweepGroups * C});
    a.Xb = b[0]; a.XmB = b[1]; a.Kb = b[2]; a.T1 = b[3]; a.T2 = b[4]; a.K0 = b[5]; a.Ub = b[6];
    a.psib = b[7]; a.psib2 = b[8];
    a.ybar = b[9]; a.Db = b[10]; a.gbuf = b[11];
    a.Xsave = b[12]; a.psisave = b[13]; a.dosesave = b[14];
    a.u0save = b[15];
    for (const titus::SpMat& A : x.sA) {
      HMat D(A.m, A.n);
      for (int i = 0; i < A.m; ++i)
        for (int t = A.rowptr[i]; t < A.rowptr[i + 1]; ++t) D(A.col[t], i) = A.val[t];
      a.sAt.push_back(titus::SpMat::from_dense(D));
    }
  }
  // step tapes for the longest segment
  const size_t seg = (size_t)std::min(a.every, x.steps_done);
  if (a.tapes.size() < seg) {
    const size_t have = a.tapes.size(), add = seg - have;
    std::vector<size_t> sizes;
    for (size_t i = 0; i < add; ++i)
      for (size_t n : {cr, cr, cr, cq, cq, cr, cr}) sizes.push_back(n);
    const std::vector<double*> b = a.alloc_many(sizes);
    a.tapes.resize(seg);
    for (size_t i = 0; i < add; ++i) {
      StepTape& t = a.tapes[have + i];
      double* const* q = b.data() + 7 * i;
      t.X = q[0]; t.X1 = q[1]; t.X3 = q[2];
      t.psi_a = q[3]; t.psi_b = q[4];
      t.Rx1.Q2 = q[5]; t.Rx3.Q2 = q[6];
    }
  }
  hipStream_t st = x.st;
  // save the forward state the recomputation overwrites
  run.done = x.steps_done;
  run.S_end = x.S;
  run.W_end = x.W;
  HIPCHK(hipMemcpyAsync(a.Xsave, x.X, sizeof(double) * cr, hipMemcpyDeviceToDevice, st));
  HIPCHK(hipMemcpyAsync(a.psisave, x.psi, sizeof(double) * cq, hipMemcpyDeviceToDevice, st));
  HIPCHK(hipMemcpyAsync(a.dosesave, x.dose, sizeof(double) * C, hipMemcpyDeviceToDevice, st));
  run.cu = (size_t)x.u0_groups * C;
  HIPCHK(hipMemcpyAsync(a.u0save, x.u0p, sizeof(double) * run.cu, hipMemcpyDeviceToDevice, st));
  HIPCHK(hipMemcpyAsync(a.Db, dose_bar, sizeof(double) * C, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemsetAsync(a.Xb, 0, sizeof(double) * cr, st));
  HIPCHK(hipMemsetAsync(a.psib, 0, sizeof(double) * cq, st));
  HIPCHK(hipMemsetAsync(a.ybar, 0, sizeof(double) * C, st));
  run.hash_was = x.decision_hash;
  run.decisions_was = x.decisions;
  run.timing_was = x.timing.enabled;
  x.timing.enabled = false;
  run.begun = true;
}

// The segments backwards: restore a checkpoint, recompute the segment with
// the tape on, run the taped steps in reverse.  Its launches go through the
// context (alone, or recorded in a lockstep group).
void adjoint_segments(titus_ctx& x, const AdjointRun& run) {
  AdjointStore& a = *x.adj;
  const size_t cr = (size_t)x.C * x.rank, cq = (size_t)x.C * x.nq;
  const int done = run.done;
  Reverse rev(x, a);
  for (int ci = (int)a.cps.size() - 1; ci >= 0; --ci) {
    const Checkpoint& cp = a.cps[ci];
    if (cp.s > done) continue;
    if (!cp.saved) throw TitusError(TITUS_E_STATE, "adjoint: missing checkpoint");
    const int s1 = (ci + 1 < (int)a.cps.size()) ? std::min(done, a.cps[ci + 1].s - 1) : done;
    if (s1 - cp.s + 1 > (int)a.tapes.size()) throw TitusError(TITUS_E_STATE, "adjoint: checkpoint gap");
    // restore and recompute the segment with the tape on
    x.dcopy(x.X, cp.X, sizeof(double) * cr);
    x.dcopy(x.psi, cp.psi, sizeof(double) * cq);
    x.S = cp.S;
    x.W = cp.W;
    for (int s = cp.s; s <= s1; ++s) {
      x.tape = &a.tapes[s - cp.s];
      x.step_fixed(s);
      x.tape = nullptr;
    }
    for (int s = s1; s >= cp.s; --s) {
      rev.step(a.tapes[s - cp.s]);
      // the host-side adjoint coefficients carry every reverse step
      for (double v : rev.Sb.a)
        if (!std::isfinite(v)) throw TitusError(TITUS_E_NONFINITE, "adjoint: non-finite coefficients");
      for (double v : rev.Wb.a)
        if (!std::isfinite(v)) throw TitusError(TITUS_E_NONFINITE, "adjoint: non-finite coefficients");
    }
  }
}

// Puts the forward state back as it was before the adjoint (device blocks
// from their copies, host factors, counters) and, when `grad` is not null,
// writes d/drho = -y^2 d/dy.  On the context's own stream.
void adjoint_end(titus_ctx& x, const AdjointRun& run, double* grad) {
  if (!run.begun) return;
  AdjointStore& a = *x.adj;
  const int C = x.C;
  const size_t cr = (size_t)C * x.rank, cq = (size_t)C * x.nq;
  hipStream_t st = x.st;
  x.tape = nullptr;
  x.timing.enabled = run.timing_was;
  x.decision_hash = run.hash_was;
  x.decisions = run.decisions_was;
  std::vector<double> yb(grad ? C : 0), rho(grad ? C : 0);
  if (grad) {
    HIPCHK(hipMemcpyAsync(yb.data(), a.ybar, sizeof(double) * C, hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(rho.data(), x.rho, sizeof(double) * C, hipMemcpyDeviceToHost, st));
  }
  HIPCHK(hipMemcpyAsync(x.X, a.Xsave, sizeof(double) * cr, hipMemcpyDeviceToDevice, st));
  HIPCHK(hipMemcpyAsync(x.psi, a.psisave, sizeof(double) * cq, hipMemcpyDeviceToDevice, st));
  HIPCHK(hipMemcpyAsync(x.dose, a.dosesave, sizeof(double) * C, hipMemcpyDeviceToDevice, st));
  HIPCHK(hipMemcpyAsync(x.u0p, a.u0save, sizeof(double) * run.cu, hipMemcpyDeviceToDevice, st));
  x.S = run.S_end;
  x.W = run.W_end;
  x.steps_done = run.done;
  HIPCHK(hipStreamSynchronize(st));
  if (!grad) return;
  for (int c = 0; c < C; ++c) {
    if (!std::isfinite(yb[c])) throw TitusError(TITUS_E_NONFINITE, "adjoint: non-finite gradient");
    grad[c] = -yb[c] / (rho[c] * rho[c]);
  }
}

// a failed HIP call may leave the stream unusable: the restore is then best
// effort and the first error is the one reported
void adjoint_end_after_failure(titus_ctx& x, const AdjointRun& run) {
  try {
    adjoint_end(x, run, nullptr);
  } catch (...) {
  }
}

void run_adjoint(titus_ctx& x, const double* dose_bar, double* grad) {
  AdjointRun run;
  try {
    adjoint_begin(x, dose_bar, run);
    adjoint_segments(x, run);
  } catch (...) {
    adjoint_end_after_failure(x, run);
    throw;
  }
  adjoint_end(x, run, grad);
}

}  // namespace

extern "C" {

int titus_set_checkpoints(titus_ctx* ctx, int32_t every) {
  return guarded([&] {
    require(ctx != nullptr, "null context");
    require(every >= 1 && every <= 4096, "checkpoint interval out of range");
    if (ctx->integrator != TITUS_INTEGRATOR_FIXED)
      throw TitusError(TITUS_E_ARG, "adjoint: fixed-rank integrator only");
    if (ctx->precision != TITUS_PRECISION_FP64)
      throw TitusError(TITUS_E_ARG, "adjoint: FP64 precision only");
    if (ctx->nd != 2) throw TitusError(TITUS_E_ARG, "adjoint: 2D grids only");
    if (ctx->flux_order != 1) throw TitusError(TITUS_E_ARG, "adjoint: upwind uncollided flux only");
    if (ctx->steps_done != 0) throw TitusError(TITUS_E_STATE, "checkpoints must be set before the first step");
    ctx->adj = std::make_shared<AdjointStore>();
    ctx->adj->every = every;
    adjoint_reserve(*ctx);
  });
}

int titus_adjoint(titus_ctx* ctx, const double* dose_bar, double* grad_density) {
  return guarded([&] {
    require(ctx != nullptr && dose_bar != nullptr && grad_density != nullptr, "null argument");
    HIPCHK(hipSetDevice(ctx->device));
    run_adjoint(*ctx, dose_bar, grad_density);
  });
}

int titus_adjoint_batch(titus_ctx** ctxs, int32_t n, int32_t nthreads, const double* const* dose_bars,
                        double* const* grads, int32_t* status) {
  if (!ctxs || n < 0 || !dose_bars || !grads) {
    g_last_error = "bad argument";
    return TITUS_E_ARG;
  }
  if (batch_has_duplicates(ctxs, n)) {
    g_last_error = "a context appears twice in the batch";
    return TITUS_E_ARG;
  }
  int nt = nthreads > 0 ? nthreads : (int)std::max(1u, std::thread::hardware_concurrency());
  nt = std::max(1, std::min<int>(nt, n));
  std::vector<int> codes(n, TITUS_OK);
  std::vector<std::string> msgs(n);
  std::atomic<int> next{0};
  HostParallelOff hp(ctxs, n, nt > 1);
  auto worker = [&] {
    for (;;) {
      const int i = next.fetch_add(1);
      if (i >= n) return;
      codes[i] = titus_adjoint(ctxs[i], dose_bars[i], grads[i]);
      if (codes[i] != TITUS_OK) msgs[i] = g_last_error;
    }
  };
  std::vector<std::thread> pool;
  for (int t = 1; t < nt; ++t) pool.emplace_back(worker);
  worker();
  for (auto& th : pool) th.join();
  int first = TITUS_OK;
  for (int i = 0; i < n; ++i) {
    if (status) status[i] = codes[i];
    if (first == TITUS_OK && codes[i] != TITUS_OK) {
      first = codes[i];
      g_last_error = "adjoint " + std::to_string(i) + ": " + msgs[i];
    }
  }
  return first;
}

int titus_adjoint_batch_lockstep(titus_ctx** ctxs, int32_t n, int32_t width, int32_t groups,
                                 const double* const* dose_bars, double* const* grads,
                                 int32_t* status) {
  if (!ctxs || n < 0 || !dose_bars || !grads) {
    g_last_error = "bad argument";
    return TITUS_E_ARG;
  }
  if (batch_has_duplicates(ctxs, n)) {
    g_last_error = "a context appears twice in the batch";
    return TITUS_E_ARG;
  }
  for (int i = 0; i < n; ++i)
    if (!ctxs[i] || ctxs[i]->device != ctxs[0]->device || !dose_bars[i] || !grads[i]) {
      g_last_error = "adjoint lockstep batch, context " + std::to_string(i) +
                     ": null argument or contexts on different devices";
      return TITUS_E_ARG;
    }
  std::vector<AdjointRun> runs(n);
  std::vector<int> codes(n, TITUS_OK);
  std::vector<std::string> msgs(n);
  // set-up on each context's own stream (synchronised before the shared launches)
  for (int i = 0; i < n; ++i) {
    codes[i] = guarded([&] {
      HIPCHK(hipSetDevice(ctxs[i]->device));
      adjoint_begin(*ctxs[i], dose_bars[i], runs[i]);
    });
    if (codes[i] != TITUS_OK) msgs[i] = g_last_error;
  }
  std::vector<titus_ctx*> live;
  std::vector<int> idx;
  for (int i = 0; i < n; ++i)
    if (codes[i] == TITUS_OK) {
      live.push_back(ctxs[i]);
      idx.push_back(i);
    }
  std::vector<int32_t> st(live.size(), TITUS_OK);
  std::vector<std::string> lmsg(live.size());
  const int rc = lockstep_run(live.data(), (int)live.size(), width, groups, st.data(), nullptr, nullptr,
                              [&](titus_ctx& c, int j) {
                                try {
                                  adjoint_segments(c, runs[idx[j]]);
                                } catch (const std::exception& e) {
                                  lmsg[j] = e.what();
                                  throw;
                                }
                              });
  (void)rc;
  for (size_t j = 0; j < live.size(); ++j)
    if (st[j] != TITUS_OK) {
      codes[idx[j]] = st[j];
      msgs[idx[j]] = lmsg[j].empty() ? "lockstep batch failed" : lmsg[j];
    }
  // the forward states back and the gradients, on each context's stream
  for (int i = 0; i < n; ++i) {
    if (codes[i] == TITUS_OK) {
      codes[i] = guarded([&] { adjoint_end(*ctxs[i], runs[i], grads[i]); });
      if (codes[i] != TITUS_OK) msgs[i] = g_last_error;
    } else {
      adjoint_end_after_failure(*ctxs[i], runs[i]);
    }
  }
  int first = TITUS_OK;
  for (int i = 0; i < n; ++i) {
    if (status) status[i] = codes[i];
    if (first == TITUS_OK && codes[i] != TITUS_OK) {
      first = codes[i];
      g_last_error = "adjoint " + std::to_string(i) + ": " + msgs[i];
    }
  }
  return first;
}

}  // extern "C"
