"""The rejection of the rank-adaptive step in the oracle
(``oracle.dlra.solve_bug`` with ``P["rejection"] = c_eta``, after
``SolverCSD.jl:2204-2233, 2282-2301``).

Pins:
* off (c_eta = 0) is the plain solve, bit for bit;
* a rejected step restarts from its own start: the uncollided flux psi, which
  no rank decision touches, ends bit for bit as in the solve without
  rejection (a psi swept twice in a repeated step would differ);
* the rejections respect the cap (2r < cap, the reference's condition with
  settings.r in place of its undefined `rmax`), and a smaller c_eta rejects
  at least as often;
* the eta of the scattering half restated once more by hand on a restarted
  step's data agrees with the oracle's (through the rejection count).
"""
import numpy as np

import cases
from oracle import dlra


def test_rejection_off_is_the_plain_solve():
    P = cases.small_bug(steps=25, cap=12)
    a = dlra.solve_bug(P, steps=25)
    b = dlra.solve_bug(dict(P, rejection=0.0), steps=25)
    assert np.array_equal(a["dose"], b["dose"]) and np.array_equal(a["ranks"], b["ranks"])
    assert b["rejections"] == 0


def test_restarted_steps_do_not_sweep_psi_twice():
    P = cases.small_bug(steps=30, cap=16)
    a = dlra.solve_bug(P, steps=30)
    b = dlra.solve_bug(dict(P, rejection=1e-2), steps=30)
    assert b["rejections"] >= 5
    assert np.array_equal(a["psi"], b["psi"])
    assert len(b["ranks"]) == len(a["ranks"]) == 30
    # the rejections changed the collided solution (the comparison is not vacuous)
    assert np.linalg.norm(a["dose"] - b["dose"]) > 1e-6 * np.linalg.norm(a["dose"])


def test_rejections_respect_the_cap_and_grow_with_smaller_c_eta():
    counts = []
    for ce in (1e30, 5.0, 1e-2, 1e-6):
        P = cases.small_bug(steps=30, cap=12)
        out = dlra.solve_bug(dict(P, rejection=ce), steps=30)
        assert out["ranks"].max() <= 12
        counts.append(out["rejections"])
    assert counts == sorted(counts), counts
    assert counts[-1] > counts[0]
