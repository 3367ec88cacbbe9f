"""The C-ABI library loads without a GPU and exports what include/titus_hip.h
declares; the ctypes mirror of titus_problem_v1 matches the header."""
import ctypes
import os
import re

import pytest

from titus_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "titus_hip.h")


def header_functions():
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"\b(titus_[a-z_0-9]+)\s*\(", txt)))


def header_struct_fields():
    txt = open(HEADER).read()
    body = re.search(r"typedef struct titus_problem_v1 \{(.*?)\} titus_problem_v1;", txt, re.S).group(1)
    body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
    fields = []
    for decl in body.split(";"):
        decl = decl.strip()
        if not decl:
            continue
        m = re.match(r"(const\s+)?(int32_t|double)\s*(\*?)\s*(.*)", decl)
        assert m, decl
        typ = m.group(2) + ("*" if m.group(3) else "")
        for name in m.group(4).split(","):
            name = name.strip()
            if name.startswith("*"):
                typ, name = m.group(2) + "*", name[1:].strip()
            fields.append((name, typ))
    return fields


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    declared = header_functions()
    assert set(declared) == set(_lib.EXPORTED)
    for name in declared:
        assert hasattr(lib, name), name
    assert lib.titus_abi_version() == _lib.ABI_VERSION


def test_problem_struct_mirror_matches_header():
    ct = {"int32_t": ctypes.c_int32, "double": ctypes.c_double,
          "int32_t*": ctypes.POINTER(ctypes.c_int32), "double*": ctypes.POINTER(ctypes.c_double)}
    fields = header_struct_fields()
    assert [n for n, _ in fields] == [n for n, _ in _lib.ProblemV1._fields_]
    for (n, t), (n2, t2) in zip(fields, _lib.ProblemV1._fields_):
        assert ct[t] == t2, (n, t, t2)


def test_create_reports_missing_device_or_bad_args_without_compute():
    import torch
    import cases
    P = cases.small_fixed()
    if torch.cuda.is_available():
        pytest.skip("a GPU is present: covered by the gpu tests")
    with pytest.raises(_lib.TitusError) as e:
        _lib.Context(P, _lib.INTEGRATOR_FIXED)
    assert e.value.code in (-6, -2)
    # argument validation runs before any device call
    st, keep = _lib.make_problem(P, _lib.INTEGRATOR_FIXED)
    st.version = 99
    h = ctypes.c_void_p()
    assert _lib.load().titus_create(ctypes.byref(h), ctypes.byref(st), 0) == -1
    assert b"version" in _lib.load().titus_last_error()


def test_batches_refuse_a_repeated_context():
    """Two host threads must never drive one context: a batch naming a
    context twice is refused before any context is touched."""
    lib = _lib.load()
    dummy = ctypes.create_string_buffer(64)               # never dereferenced
    ptrs = (ctypes.c_void_p * 2)(ctypes.addressof(dummy), ctypes.addressof(dummy))
    status = (ctypes.c_int32 * 2)()
    ctx_array = ctypes.cast(ptrs, ctypes.POINTER(ctypes.c_void_p))
    assert lib.titus_solve_batch(ctx_array, 2, 2, status) == -1
    assert b"twice" in lib.titus_last_error()
    P = ctypes.POINTER(ctypes.c_double)
    bufs = [(ctypes.c_double * 4)() for _ in range(2)]
    bars = (P * 2)(*[ctypes.cast(b, P) for b in bufs])
    grads = (P * 2)(*[ctypes.cast(b, P) for b in bufs])
    assert lib.titus_adjoint_batch(ctx_array, 2, 2, bars, grads, status) == -1
    assert b"twice" in lib.titus_last_error()


def _refused(mutate, expect):
    """titus_create refuses the mutated problem with TITUS_E_ARG before any
    device call, naming `expect` in its message."""
    import numpy as np
    import cases
    P = cases.small_fixed(n=12, N=3, r=2, steps=4)
    st, keep = _lib.make_problem(P, _lib.INTEGRATOR_FIXED)
    extra = mutate(st, keep, np)
    h = ctypes.c_void_p()
    code = _lib.load().titus_create(ctypes.byref(h), ctypes.byref(st), 0)
    msg = _lib.load().titus_last_error().decode()
    del extra
    assert code == -1, (code, msg)
    assert expect in msg, msg
    assert not h.value


def _set_array(st, keep, np, name, key, fn):
    a = keep[key].copy()
    fn(a)
    keep[key + "_mut"] = a
    setattr(st, name, a.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))


@pytest.mark.parametrize("case", ["zero_density", "negative_density", "nan_density", "inf_density"])
def test_create_refuses_bad_density(case):
    val = {"zero_density": 0.0, "negative_density": -1.0, "nan_density": float("nan"),
           "inf_density": float("inf")}[case]

    def mutate(st, keep, np):
        _set_array(st, keep, np, "density", "density", lambda a: a.__setitem__(7, val))
    _refused(mutate, "density")


def test_create_refuses_nonfinite_initial_flux():
    def mutate(st, keep, np):
        psi0 = np.zeros((st.nx * st.ny, st.nq), order="F")
        psi0[5, 0] = float("nan")
        keep["psi0_mut"] = psi0
        st.psi0 = psi0.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        return psi0
    _refused(mutate, "initial flux")


@pytest.mark.parametrize("field", ["nb", "nrestore", "ncand"])
def test_create_refuses_negative_counts(field):
    def mutate(st, keep, np):
        setattr(st, field, -3)
    _refused(mutate, "negative")


def test_create_refuses_missing_restore_cells():
    def mutate(st, keep, np):
        st.restore_cells = ctypes.POINTER(ctypes.c_int32)()
    _refused(mutate, "restore_cells")


def test_create_refuses_bad_energy_step_and_cells():
    def step(st, keep, np):
        st.dE = float("nan")
    _refused(step, "energy step")

    def cell(st, keep, np):
        a = keep["bcells"].copy()
        a[0] = st.nx * st.ny
        keep["bc_mut"] = a
        st.bcells = a.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))
    _refused(cell, "ghost cell")


def test_cfl_number_of_the_reference_grids():
    """The CFL number the library reports is the one the reference prints,
    dE / min(dx, dy) * max(1 / rho) (SolverCSD.jl:963); checked here on the
    host formula for the phantom and slab (the library returns the same
    value after titus_create, covered by the gpu tests)."""
    import numpy as np
    import cases
    P = cases.phantom_bug(n=64, N=5, cap=8)
    cfl = P["dE"] / min(P["dx"], P["dy"]) * np.max(1.0 / P["density"])
    assert 0.1 < cfl < 1.0


def test_galerkin_primitive_refuses_bad_arguments_before_any_device_call():
    lib = _lib.load()
    g = _lib.GridV1(8, 8, 1, 0.1, 0.1, 1.0, None)
    buf = ctypes.create_string_buffer(8)
    p = ctypes.cast(buf, ctypes.c_void_p)
    assert lib.titus_galerkin(ctypes.byref(g), p, 64, 4, p, None) == -1     # no density
    assert b"null" in lib.titus_last_error()
    g.inv_density = p
    assert lib.titus_galerkin(ctypes.byref(g), p, 64, 0, p, None) == -1     # rank 0
    assert lib.titus_galerkin(ctypes.byref(g), p, 64, 97, p, None) == -1    # rank above 96
    assert lib.titus_galerkin(ctypes.byref(g), p, 10, 4, p, None) == -1     # ldx < C
    assert b"leading" in lib.titus_last_error()
    g.nx = 2
    assert lib.titus_galerkin(ctypes.byref(g), p, 64, 4, p, None) == -1
