"""Extract the tabulated material data of the reference into JSON.

The reference keeps its stopping-power and elastic-scattering tables as Julia
array literals inside constructor functions:

* electrons: ``code/MaterialParameters.jl:14-137`` (``S_tab``, ``E_tab``,
  ``E_sigmaTab``, ``sigma_tab``), water, ESTAR-derived stopping power and
  Legendre moments of the elastic cross-section;
* protons: ``code/MaterialParametersProtons.jl:19-60`` (``S_tab_PSTAR``,
  ``E_tab_PSTAR``, ``E_sigmaTab``, ``sigmaEl_OInt_ICRU``).

These are physics data, not code. The literals are parsed as text (no Julia,
no code execution) and written to ``electrons.json`` / ``protons.json`` in this
directory with a provenance record.  Run once, in a container where the
reference tree is mounted:

    python titus_amd/data/extract_tables.py /root/reference/code
"""
import json
import os
import re
import sys

import numpy as np


def _literal(text, name, start_line=None):
    """Return the numbers of the Julia literal ``name = [ ... ]`` as an array.

    Julia semantics: commas separate vector entries; ``;`` or a newline ends a
    matrix row and whitespace separates columns.
    """
    pat = re.compile(r"(^|\s)" + re.escape(name) + r"\s*=\s*\[", re.M)
    m = pat.search(text)
    if m is None:
        raise KeyError(name)
    i = m.end()
    j = text.index("]", i)
    body = text[i:j]
    # strip Julia comments
    body = re.sub(r"#[^\n]*", "", body)
    if "," in body:
        vals = [float(t) for t in re.split(r"[,\s;]+", body.strip()) if t]
        return np.array(vals)
    rows = []
    for row in re.split(r"[;\n]", body):
        toks = [t for t in row.split() if t]
        if toks:
            rows.append([float(t) for t in toks])
    if all(len(r) == 1 for r in rows):
        return np.array([r[0] for r in rows])
    if len(rows) == 1:
        # a 1xn row literal (transposed to a vector by the reference)
        return np.array(rows[0])
    ncol = len(rows[0])
    if any(len(r) != ncol for r in rows):
        # a row-vector literal wrapped over lines: flatten
        return np.array([v for r in rows for v in r])
    return np.array(rows)


def extract(code_dir):
    out_dir = os.path.dirname(os.path.abspath(__file__))
    with open(os.path.join(code_dir, "MaterialParameters.jl")) as f:
        txt = f.read()
    el = {
        "S_tab": _literal(txt, "S_tab"),
        "E_tab": _literal(txt, "E_tab"),
        "E_sigmaTab": _literal(txt, "E_sigmaTab"),
        "sigma_tab": _literal(txt, "sigma_tab"),
    }
    assert el["S_tab"].shape == (111,) and el["E_tab"].shape == (111,)
    assert el["E_sigmaTab"].shape == (96,) and el["sigma_tab"].shape == (96, 40)
    doc = {
        "provenance": "CSMMLab/TITUS code/MaterialParameters.jl:14-137 "
                      "(water; stopping power in MeV cm^2/g vs kinetic energy in MeV; "
                      "elastic Legendre moments sigma_tab[energy, l] for l = 0..39)",
        **{k: v.tolist() for k, v in el.items()},
    }
    with open(os.path.join(out_dir, "electrons.json"), "w") as f:
        json.dump(doc, f)

    with open(os.path.join(code_dir, "MaterialParametersProtons.jl")) as f:
        txt = f.read()
    pr = {
        "S_tab_PSTAR": _literal(txt, "S_tab_PSTAR"),
        "E_tab_PSTAR": _literal(txt, "E_tab_PSTAR"),
        "E_sigmaTab": _literal(txt, "E_sigmaTab"),
        "sigmaEl_OInt_ICRU": _literal(txt, "sigmaEl_OInt_ICRU"),
    }
    assert pr["S_tab_PSTAR"].shape == (90,) and pr["E_tab_PSTAR"].shape == (90,)
    assert pr["E_sigmaTab"].shape == (32,)
    assert pr["sigmaEl_OInt_ICRU"].shape == (32, 40), pr["sigmaEl_OInt_ICRU"].shape
    doc = {
        "provenance": "CSMMLab/TITUS code/MaterialParametersProtons.jl:19-93 "
                      "(PSTAR stopping power of water vs kinetic energy in MeV; "
                      "O-16 elastic Legendre moments on E_sigmaTab (kinetic MeV))",
        **{k: v.tolist() for k, v in pr.items()},
    }
    with open(os.path.join(out_dir, "protons.json"), "w") as f:
        json.dump(doc, f)


if __name__ == "__main__":
    extract(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/code")
