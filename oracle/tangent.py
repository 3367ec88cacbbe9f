"""Forward-mode (tangent-linear) derivative of the fixed-rank DLRA loop --
TEST INFRASTRUCTURE ONLY (only ``tests/`` may import it).

The dual-number counterpart of ``oracle.adjoint``: it pushes a density
direction v through the recorded steps of ``oracle.adjoint.forward`` (the
forward map of ``oracle.dlra.solve_fixed``, ``code/SolverCSD.jl:891-1120``)
and returns d dose[v] = sum_c (d dose / d rho_c) v_c.  Each statement of the
step is differentiated in the order it runs, with the value of every operand
taken from the tape.  With y = 1/rho, dy = -y^2 v.

Basis updates X1 = range_basis(K) are differentiated as the subspace they
span: dX1 = (I - X1 X1^T) dU_p T_p, where dU_p is the derivative of the kept
singular subspace of K (Golub--Van Loan): (I - U U^T) dK V_p Sigma_p^-1 plus
u_j (sigma_i u_j^T dK v_i + sigma_j u_i^T dK v_j) / (sigma_i^2 - sigma_j^2)
for each dropped pair j.  The rotation X1 Omega within span(X1) that a
particular basis choice adds is left out: the dose does not depend on it.

``<dose_bar, tangent(v)> = <adjoint(dose_bar), v>`` for every pair (dose_bar,
v) is the dot-product test of the reverse pass; central differences of the
forward map pin the tangent itself (``tests/test_adjoint_oracle.py``).
"""
import numpy as np

from . import dlra
from .adjoint import flux_matrices, stencil_matrices


def orth_dot(rec, dK):
    """dX1 for X1 = range_basis(K) (record from ``oracle.adjoint.orth``)."""
    X1, p = rec["X1"], rec["p"]
    if p == 0:
        return np.zeros_like(X1)
    U, sv, Vt = rec["U"], rec["sv"], rec["Vt"]
    Up, Vp, sp_ = U[:, :p], Vt[:p].T, sv[:p]
    G = (dK @ Vp) / sp_[None, :]
    dUp = G - U @ (U.T @ G)
    if len(sv) > p:
        Ud, Vd, sd = U[:, p:], Vt[p:].T, sv[p:]
        A1 = Ud.T @ dK @ Vp                      # u_j^T dK v_i
        A2 = Up.T @ dK @ Vd                      # u_i^T dK v_j
        Cm = (A1 * sp_[None, :] + A2.T * sd[:, None]) / (sp_[None, :] ** 2 - sd[:, None] ** 2)
        dUp = dUp + Ud @ Cm
    dX1 = dUp @ rec["Tp"]
    return dX1 - X1 @ (X1.T @ dX1)


def tangent(P, tape, v, Lhat=None, Fhat=None):
    """d dose along the density direction v (C vector), from the tape."""
    ops = dlra._Ops(P)
    y = 1.0 / ops.rho
    dy = -y * y * np.asarray(v, dtype=np.float64)
    dE = P["dE"]
    Sst = P["Sst"]
    M = ops.M
    m0 = M[0, :]
    As = ops.As
    order = ops._order()
    bc = ops.bcells
    Lhat = stencil_matrices(P) if Lhat is None else Lhat
    Fhat = flux_matrices(P) if Fhat is None else Fhat
    C, Q, r = P["C"], P["Q"], P["r"]
    dX = np.zeros((C, r))
    dS = np.zeros((r, r))
    dW = np.zeros((P["Nm"], r))
    dpsi = np.zeros((C, Q))
    ddose = np.zeros(C)

    def L(o, U, dU):
        """d (Lhat_o (y o U)) for the tangent dU of U."""
        return Lhat[o] @ (dy[:, None] * U + y[:, None] * dU)

    def dose_dot(a, X, S, W, psi, dX, dS, dW, dpsi):
        u = X @ (S @ W[0, :]) + psi @ m0
        du = dX @ (S @ W[0, :]) + X @ (dS @ W[0, :]) + X @ (S @ dW[0, :]) + dpsi @ m0
        return a * (dy * u + y * du)

    for t in tape:
        s = t["s"]
        Dvec, Dout = t["Dvec"], t["Dout"]
        X, S, W = t["X"], t["S"], t["W"]
        psi_a, psi_b = t["psi_a"], t["psi_b"]
        # boundary overwrite, first dose half, sweep
        dpsi = dpsi.copy()
        dpsi[bc, :] = 0.0
        ddose += dose_dot(0.5 * dE * Sst[s - 1], X, S, W, psi_a, dX, dS, dW, dpsi)
        c = 1.0 / (1.0 + dE * t["sig0"])
        dpb = np.empty_like(dpsi)
        for q in range(Q):
            dpb[:, q] = c * (dpsi[:, q] - dE * (Fhat[q] @ (dy * psi_a[:, q] + y * dpsi[:, q])))
        dpb[ops.restore, 0] = dpsi[ops.restore, 0]
        dpsi = dpb
        if s > 1:
            Xm = X.copy()
            Xm[bc, :] = 0.0
            dXm = dX.copy()
            dXm[bc, :] = 0.0
            # K-step
            K0 = t["K0"]
            dK0 = dXm @ S + Xm @ dS
            dK1 = dK0.copy()
            for o in order:
                WA = t["WAW"][o]
                dWA = dW.T @ As[o] @ W + W.T @ As[o] @ dW
                dK1 -= dE * (L(o, K0, dK0) @ WA + (Lhat[o] @ (y[:, None] * K0)) @ dWA)
            X1 = t["X1"]
            dX1 = orth_dot(t["Rx1"], dK1)
            dMUp = dX1.T @ Xm + X1.T @ dXm
            # L-step
            L0 = t["L0"]
            dL0 = dW @ S.T + W @ dS.T
            dL1 = dL0.copy()
            for o in order:
                XL = t["XLX"][o]
                dXL = dXm.T @ (Lhat[o] @ (y[:, None] * Xm)) + Xm.T @ L(o, Xm, dXm)
                dL1 -= dE * (As[o] @ dL0 @ XL.T + As[o] @ L0 @ dXL.T)
            W1 = t["W1"]
            dW1 = orth_dot(t["Rw1"], dL1)
            dNUp = dW1.T @ W + W1.T @ dW
            MUp, NUp = t["MUp1"], t["NUp1"]
            S1 = t["S1"]
            dS1 = dMUp @ S @ NUp.T + MUp @ dS @ NUp.T + MUp @ S @ dNUp.T
            # S-step
            dS2 = dS1.copy()
            for o in order:
                XL, WA = t["XLX1"][o], t["WAW1"][o]
                dXL = dX1.T @ (Lhat[o] @ (y[:, None] * X1)) + X1.T @ L(o, X1, dX1)
                dWA = dW1.T @ As[o] @ W1 + W1.T @ As[o] @ dW1
                dS2 -= dE * (dXL @ S1 @ WA + XL @ dS1 @ WA + XL @ S1 @ dWA)
            dX, dW, dS = dX1, dW1, dS2
        # out-scattering
        Ws, Ss = t["Ws"], t["Ss"]
        dinv = 1.0 / (1.0 + dE * Dout)
        dLo = (dW @ Ss.T + Ws @ dS.T) * dinv[:, None]
        Lo, W2 = t["Lo"], t["W2"]
        dW2 = orth_dot(t["Rw2"], dLo)
        dS3 = dLo.T @ W2 + Lo.T @ dW2
        # in-scattering
        Xm = t["Xs"].copy()
        Xm[bc, :] = 0.0
        dXm = dX.copy()
        dXm[bc, :] = 0.0
        S3 = t["S3"]
        MtDW = t["MtDW"]
        dMtDW = M.T @ (Dvec[:, None] * dW2)
        dKin = dXm @ S3 + Xm @ dS3 + dE * (dpsi @ MtDW + psi_b @ dMtDW)
        dKin[bc, :] = 0.0
        X3 = t["X3"]
        dX3 = orth_dot(t["Rx3"], dKin)
        dMUp = dX3.T @ Xm + X3.T @ dXm
        dLin = dW2 @ S3.T + W2 @ dS3.T + dE * Dvec[:, None] * (M @ (dpsi.T @ Xm + psi_b.T @ dXm))
        W3 = t["W3"]
        dW3 = orth_dot(t["Rw3"], dLin)
        dNUp = dW3.T @ W2 + W3.T @ dW2
        MUp, NUp = t["MUp3"], t["NUp3"]
        dS4 = dMUp @ S3 @ NUp.T + MUp @ dS3 @ NUp.T + MUp @ S3 @ dNUp.T
        A3, B3 = t["A3"], t["B3"]
        dA3 = dX3.T @ psi_b + X3.T @ dpsi
        dB3 = M.T @ (Dvec[:, None] * dW3)
        dS5 = dS4 + dE * (dA3 @ B3 + A3 @ dB3)
        dX, dW, dS = dX3, dW3, dS5
        ddose += dose_dot(0.5 * dE * Sst[s], X3, t["S5"], W3, psi_b, dX, dS, dW, dpsi)
    return ddose
