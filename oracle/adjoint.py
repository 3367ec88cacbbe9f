"""FP64 CPU discrete adjoint of the fixed-rank DLRA energy loop -- TEST
INFRASTRUCTURE ONLY (the checker of the device adjoint; only ``tests/`` may
import it).

The reference has no adjoint.  What is differentiated is the forward map of
``oracle.dlra.solve_fixed`` (``SolveFirstCollisionSourceDLR``,
``code/SolverCSD.jl:891-1120``): the dose after the energy loop as a function
of the density rho, with the energy grid, the beam source and the P_N
operators held fixed.  rho enters in three places, always as y = 1/rho:

* the spatial stencils, L_o(U) = Lhat_o (y o U) (``stencils.jl:26-128``);
* the uncollided upwind flux of y o psi (``SolverCSD.jl:388-463``);
* the dose, dose += a (X S W[0,:]^T + psi M[0,:]^T) o y (``:990-997``).

Reverse mode runs the recorded steps backwards; every statement of
``solve_fixed`` has its transpose below.  The basis updates are the one
non-linear part:

    X1 = range_basis(K)  spans  R = span(U_p) + span(E),

U_p the p left singular vectors of K kept by the rank cut and E the chosen
canonical completion vectors.  Everything downstream depends on X1 only
through R (replacing X1 by X1 Q, Q orthogonal, rotates S with it and leaves
X S W^T and the dose unchanged), so only the motion of R matters:

    dX1 = (I - X1 X1^T) dU_p T_p     (mod span(X1)),

T_p the first p rows of the coefficients of X1 in the basis [U_p, E], and
dU_p the derivative of the kept singular subspace (``orth_bar``), including
its coupling to the singular pairs dropped by the cut.  The rank decisions
themselves (p and the chosen completion vectors) are locally constant; the
same holds for the angular bases.

Pins: central finite differences of the oracle's forward map
(``tests/test_adjoint_oracle.py``) and adjoint--tangent identities of each
transposed operator against its dense assembly.
"""
import numpy as np
import scipy.sparse as sp

from . import dlra


# ---------------------------------------------------------------------------
# operators as explicit sparse matrices (rho = 1): Lhat_o and the upwind flux
# ---------------------------------------------------------------------------
def stencil_matrices(P):
    """[Lhat_o] (C x C sparse) in the order of ``dlra._Ops.L``."""
    ops = dlra._Ops(P)
    C = P["C"]
    P1 = dict(P, density=np.ones(C))
    ops1 = dlra._Ops(P1)
    cols = ops1.L(np.eye(C))
    del ops
    return [sp.csr_matrix(Lo) for Lo in cols]


def flux_matrices(P):
    """[Fhat_q] (C x C sparse): upwind flux of ordinate q applied to e_c (rho = 1)."""
    C, Q = P["C"], P["Q"]
    omega = np.asarray(P["omega"])
    one = np.ones(C)
    I = np.eye(C)
    out = []
    for q in range(Q):
        om = np.repeat(omega[q:q + 1], C, axis=0)
        if int(P.get("nz", 1)) > 1:
            F = dlra.upwind_flux3(I, one, om, P["nx"], P["ny"], P["nz"], P["dx"], P["dy"], P["dz"])
        else:
            F = dlra.upwind_flux(I, one, om, P["nx"], P["ny"], P["dx"], P["dy"])
        out.append(sp.csr_matrix(F))
    return out


# ---------------------------------------------------------------------------
# basis updates and their adjoint
# ---------------------------------------------------------------------------
def orth(K, r, cands, greedy):
    """X1 = range_basis(K) and the record its adjoint needs."""
    info = {}
    X1 = dlra.range_basis(K, r, cands, greedy=greedy, info=info)
    p = info["p"]
    rec = dict(X1=X1, p=p, U=info["U"], sv=info["sv"], Vt=info["Vt"], Tp=None)
    if p > 0:
        Up = info["U"][:, :p]
        E = np.zeros((K.shape[0], len(info["chosen"])))
        for t, c in enumerate(info["chosen"]):
            E[c, t] = 1.0
        T = np.linalg.lstsq(np.hstack([Up, E]), X1, rcond=None)[0]
        rec["Tp"] = T[:p]
    return X1, rec


def orth_bar(rec, X1bar):
    """Kbar from X1bar for X1 = range_basis(K).

    X1 moves only through U_p, the kept singular subspace of K = U Sigma V^T:
    dX1 = (I - X1 X1^T) dU_p T_p (mod span X1).  The derivative of U_p
    (Golub--Van Loan) has the part outside range(K), (I - U U^T) dK V_p
    Sigma_p^-1, and for every dropped singular pair j > p the rotation
    u_j (sigma_i u_j^T dK v_i + sigma_j u_i^T dK v_j) / (sigma_i^2 - sigma_j^2).
    """
    X1, p = rec["X1"], rec["p"]
    if p == 0:
        return np.zeros((X1.shape[0], rec["Vt"].shape[1]))
    U, sv, Vt = rec["U"], rec["sv"], rec["Vt"]
    Ub = X1bar @ rec["Tp"].T
    Ub = Ub - X1 @ (X1.T @ Ub)                       # adjoint of U_p
    Up, Vp, sp_ = U[:, :p], Vt[:p].T, sv[:p]
    Y = Ub / sp_[None, :]
    Kb = (Y - U @ (U.T @ Y)) @ Vp.T
    k = len(sv)
    if k > p:
        Ud, Vd, sd = U[:, p:], Vt[p:].T, sv[p:]
        beta = Ud.T @ Ub                             # (k - p) x p
        F = 1.0 / (sp_[None, :] ** 2 - sd[:, None] ** 2)
        Bm = beta * F
        Kb = Kb + Ud @ (Bm * sp_[None, :]) @ Vp.T + Up @ (Bm * sd[:, None]).T @ Vd.T
    return Kb


# ---------------------------------------------------------------------------
# forward with a tape
# ---------------------------------------------------------------------------
def forward(P, steps=None):
    """``dlra.solve_fixed`` recording every intermediate the adjoint needs.

    Returns (dose, tape); the dose equals ``solve_fixed``'s.
    """
    ops = dlra._Ops(P)
    r = P["r"]
    X = np.array(P["X0"], dtype=np.float64)
    W = np.array(P["W0"], dtype=np.float64)
    S = np.array(P["S0"], dtype=np.float64)
    C, Q = P["C"], P["Q"]
    psi = dlra.initial_psi(P)
    dose = np.zeros(C)
    rho = ops.rho
    dE = P["dE"]
    Sst = P["Sst"]
    omega = P["omega"]
    M = ops.M
    nE = P["nE"] if steps is None else min(P["nE"], steps + 1)
    icand = ops.interior
    mcand = range(ops.Nm)
    X[ops.bcells, :] = 0.0
    tape = []
    for s in range(1, nE):
        t = dict(s=s, X=X.copy(), S=S.copy(), W=W.copy())
        sig = P["sigma"][s]
        psi = dlra._begin_step(ops, P, psi, s)
        t["psi_a"] = psi.copy()
        dose += 0.5 * dE * (X @ (S @ W[0, :]) + psi @ M[0, :]) * Sst[s - 1] / rho
        psi = dlra._sweep(ops, psi, sig[0], dE, omega)
        t["psi_b"] = psi.copy()
        Dvec = sig[ops.deg]
        Dout = sig[0] - Dvec
        t["Dvec"], t["Dout"], t["sig0"] = Dvec, Dout, sig[0]
        if s > 1:
            X[ops.bcells, :] = 0.0
            K0 = X @ S
            WAW = ops.galerkin_w(W, W)
            K1 = ops.stream_K(K0, WAW, dE)
            X1, Rx1 = orth(K1, r, icand, False)
            MUp = X1.T @ X
            L0 = W @ S.T
            XLX = ops.galerkin_x(X, X)
            L1 = ops.stream_L(L0, XLX, dE)
            W1, Rw1 = orth(L1, r, mcand, True)
            NUp = W1.T @ W
            S1 = MUp @ S @ NUp.T
            XLX1 = ops.galerkin_x(X1, X1)
            WAW1 = ops.galerkin_w(W1, W1)
            S2 = ops.stream_S(S1, XLX1, WAW1, dE)
            t.update(K0=K0, WAW=WAW, X1=X1, Rx1=Rx1, MUp1=MUp, L0=L0, XLX=XLX, W1=W1, Rw1=Rw1,
                     NUp1=NUp, S1=S1, XLX1=XLX1, WAW1=WAW1)
            X, W, S = X1, W1, S2
        t.update(Xs=X.copy(), Ws=W.copy(), Ss=S.copy())
        Lo = (W @ S.T) / (1.0 + dE * Dout)[:, None]
        W2, Rw2 = orth(Lo, r, mcand, True)
        S3 = Lo.T @ W2
        X[ops.bcells, :] = 0.0
        MtDW = M.T @ (Dvec[:, None] * W2)
        Kin = X @ S3 + dE * psi @ MtDW
        Kin[ops.bcells, :] = 0.0
        X3, Rx3 = orth(Kin, r, icand, False)
        MUp = X3.T @ X
        Lin = W2 @ S3.T + dE * Dvec[:, None] * (M @ (psi.T @ X))
        W3, Rw3 = orth(Lin, r, mcand, True)
        NUp = W3.T @ W2
        S4 = MUp @ S3 @ NUp.T
        B3 = M.T @ (Dvec[:, None] * W3)
        A3 = X3.T @ psi
        S5 = S4 + dE * A3 @ B3
        t.update(Lo=Lo, W2=W2, Rw2=Rw2, S3=S3, MtDW=MtDW, X3=X3, Rx3=Rx3, MUp3=MUp, W3=W3,
                 Rw3=Rw3, NUp3=NUp, B3=B3, A3=A3, S5=S5)
        X, W, S = X3, W3, S5
        dose += 0.5 * dE * (X @ (S @ W[0, :]) + psi @ M[0, :]) * Sst[s] / rho
        tape.append(t)
    return dose, tape


# ---------------------------------------------------------------------------
# reverse
# ---------------------------------------------------------------------------
def adjoint(P, tape, dose_bar, Lhat=None, Fhat=None):
    """Gradient of <dose_bar, dose> with respect to the density (C vector)."""
    ops = dlra._Ops(P)
    rho = ops.rho
    y = 1.0 / rho
    dE = P["dE"]
    Sst = P["Sst"]
    M = ops.M
    m0 = M[0, :]
    As = ops.As
    order = ops._order()
    bc = ops.bcells
    Lhat = stencil_matrices(P) if Lhat is None else Lhat
    Fhat = flux_matrices(P) if Fhat is None else Fhat
    Db = np.asarray(dose_bar, dtype=np.float64)
    C, Q = P["C"], P["Q"]
    r = P["r"]
    ybar = np.zeros(C)
    Xb = np.zeros((C, r))
    Sb = np.zeros((r, r))
    Wb = np.zeros((P["Nm"], r))
    psib = np.zeros((C, Q))

    def dose_rev(a, X, S, W, psi):
        nonlocal ybar, Xb, Sb, Wb, psib
        g = a * y * Db
        w0 = W[0, :]
        Sw0 = S @ w0
        ybar += a * Db * (X @ Sw0 + psi @ m0)
        Xtg = X.T @ g
        Xb += np.outer(g, Sw0)
        Sb += np.outer(Xtg, w0)
        Wb[0, :] += S.T @ Xtg
        psib += np.outer(g, m0)

    for t in reversed(tape):
        s = t["s"]
        Dvec, Dout = t["Dvec"], t["Dout"]
        psi_b = t["psi_b"]
        # ---- second dose half
        dose_rev(0.5 * dE * Sst[s], t["X3"], t["S5"], t["W3"], psi_b)
        # ---- S5 = MUp3 S3 NUp3^T + dE A3 B3
        X3, W3, W2, S3 = t["X3"], t["W3"], t["W2"], t["S3"]
        Xm = t["Xs"].copy()
        Xm[bc, :] = 0.0
        S4b = Sb
        A3b = dE * Sb @ t["B3"].T
        B3b = dE * t["A3"].T @ Sb
        X3b = Xb + psi_b @ A3b.T
        psib = psib + X3 @ A3b
        W3b = Wb + Dvec[:, None] * (M @ B3b)
        MUpb = S4b @ t["NUp3"] @ S3.T
        S3b = t["MUp3"].T @ S4b @ t["NUp3"]
        NUpb = S4b.T @ t["MUp3"] @ S3
        W3b = W3b + W2 @ NUpb.T
        W2b = W3 @ NUpb
        X3b = X3b + Xm @ MUpb.T
        Xmb = X3 @ MUpb
        # W3 = orth(Lin), Lin = W2 S3^T + dE Dvec o (M psi^T Xm)
        Linb = orth_bar(t["Rw3"], W3b)
        W2b = W2b + Linb @ S3
        S3b = S3b + Linb.T @ W2
        Pb = M.T @ (dE * Dvec[:, None] * Linb)        # Q x r, adjoint of psi^T Xm
        psib = psib + Xm @ Pb.T
        Xmb = Xmb + psi_b @ Pb
        # X3 = orth(Kin), Kin = mask(Xm S3 + dE psi MtDW)
        Kinb = orth_bar(t["Rx3"], X3b)
        Kinb[bc, :] = 0.0
        Xmb = Xmb + Kinb @ S3.T
        S3b = S3b + Xm.T @ Kinb
        psib = psib + dE * Kinb @ t["MtDW"].T
        MtDWb = dE * psi_b.T @ Kinb
        W2b = W2b + Dvec[:, None] * (M @ MtDWb)
        Xb = Xmb
        Xb[bc, :] = 0.0
        # ---- out-scattering: S3 = Lo^T W2, W2 = orth(Lo), Lo = (W S^T) / (1 + dE Dout)
        Lo = t["Lo"]
        Lob = W2 @ S3b.T
        W2b = W2b + Lo @ S3b
        Lob = Lob + orth_bar(t["Rw2"], W2b)
        WSb = Lob / (1.0 + dE * Dout)[:, None]
        Wb = WSb @ t["Ss"]
        Sb = WSb.T @ t["Ws"]
        # ---- streaming (s > 1)
        if s > 1:
            X, S, W = t["X"].copy(), t["S"], t["W"]
            X[bc, :] = 0.0
            X1, W1, S1 = t["X1"], t["W1"], t["S1"]
            # S2 = S1 - dE sum_o XLX1_o S1 WAW1_o
            S2b = Sb
            S1b = S2b.copy()
            X1b = Xb.copy()
            W1b = Wb.copy()
            for o in order:
                XL, WA = t["XLX1"][o], t["WAW1"][o]
                S1b -= dE * XL.T @ S2b @ WA.T
                XLb = -dE * S2b @ WA.T @ S1.T
                WAb = -dE * S1.T @ XL.T @ S2b
                # WAW1_o = W1^T A_o W1
                W1b += As[o] @ W1 @ WAb.T + As[o].T @ W1 @ WAb
                # XLX1_o = X1^T Lhat_o (y o X1)
                X1b += (Lhat[o] @ (y[:, None] * X1)) @ XLb.T
                T = Lhat[o].T @ (X1 @ XLb)
                X1b += y[:, None] * T
                ybar += np.einsum("ij,ij->i", X1, T)
            # S1 = MUp S NUp^T
            MUp, NUp = t["MUp1"], t["NUp1"]
            MUpb = S1b @ NUp @ S.T
            Sb = MUp.T @ S1b @ NUp
            NUpb = S1b.T @ MUp @ S
            W1b += W @ NUpb.T
            Wb = W1 @ NUpb
            X1b += X @ MUpb.T
            Xb = X1 @ MUpb
            # W1 = orth(L1), L1 = L0 - dE sum_o A_o L0 XLX_o^T, L0 = W S^T
            L1b = orth_bar(t["Rw1"], W1b)
            L0 = t["L0"]
            L0b = L1b.copy()
            for o in order:
                XL = t["XLX"][o]
                L0b -= dE * As[o].T @ L1b @ XL
                XLb = -dE * L1b.T @ As[o] @ L0
                # XLX_o = X^T Lhat_o (y o X)
                Xb += (Lhat[o] @ (y[:, None] * X)) @ XLb.T
                T = Lhat[o].T @ (X @ XLb)
                Xb += y[:, None] * T
                ybar += np.einsum("ij,ij->i", X, T)
            Wb += L0b @ S
            Sb += L0b.T @ W
            # X1 = orth(K1), K1 = K0 - dE sum_o Lhat_o (y o K0) WAW_o, K0 = X S
            K1b = orth_bar(t["Rx1"], X1b)
            K0 = t["K0"]
            K0b = K1b.copy()
            for o in order:
                WA = t["WAW"][o]
                T = Lhat[o].T @ (K1b @ WA.T)
                K0b -= dE * y[:, None] * T
                ybar -= dE * np.einsum("ij,ij->i", K0, T)
                WAb = -dE * (Lhat[o] @ (y[:, None] * K0)).T @ K1b
                Wb += As[o] @ W @ WAb.T + As[o].T @ W @ WAb
            Xb += K0b @ S.T
            Sb += X.T @ K0b
            Xb[bc, :] = 0.0
        # ---- sweep: psi_b = c (psi_a - dE Fhat (y o psi_a)), restore of ordinate 0
        psi_a = t["psi_a"]
        c = 1.0 / (1.0 + dE * t["sig0"])
        keep = psib[ops.restore, 0].copy()
        psib[ops.restore, 0] = 0.0
        pa = np.empty_like(psib)
        for q in range(Q):
            tq = Fhat[q].T @ psib[:, q]
            pa[:, q] = c * psib[:, q] - dE * c * y * tq
            ybar -= dE * c * psi_a[:, q] * tq
        pa[ops.restore, 0] += keep
        psib = pa
        # ---- first dose half, then the boundary overwrite
        dose_rev(0.5 * dE * Sst[s - 1], t["X"], t["S"], t["W"], psi_a)
        psib[bc, :] = 0.0
    # d/drho = -y^2 d/dy
    return -ybar * y * y
