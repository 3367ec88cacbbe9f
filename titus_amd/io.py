"""Dose outputs in the reference's file formats.

The reference's drivers write the dose twice (``code/main.jl:189-193``,
``code/plot.jl:53-56``, ``code/runLung.jl:228-231``):

* ``writedlm(path, dose)`` -- the nx x ny dose matrix as text, one matrix row
  (fixed i) per line, tab-separated, each value in Julia's shortest
  round-trip form (``DelimitedFiles`` of Julia 1.8.1);
* ``vtk_grid(path, xMid, yMid)`` with point arrays "dose", "dose_normalized"
  (dose / max) and optionally "u", saved by ``vtk_save`` -- a VTK XML
  RectilinearGrid (``.vtr``), appended raw binary, UInt64 block headers, each
  array zlib-compressed (level 6) as one block (WriteVTK.jl 1.16.0 over
  CodecZlib 0.7.0, the versions pinned in ``code/Manifest.toml``).

Cell c = i*ny + j of the solver's dose vector is matrix entry (i, j); VTK
point arrays run with i fastest.  These writers are byte-compatible with the
files in the reference's ``code/output/`` (``tests/test_io.py``).
"""
import math
import re
import struct
import zlib

import numpy as np

ZLIB_LEVEL = 6


# ---------------------------------------------------------------------------
# text (writedlm / readdlm)
# ---------------------------------------------------------------------------
def format_julia(x):
    """Julia's ``print`` of a Float64: the shortest digits that round-trip,
    plain notation for 1e-4 <= |x| < 1e6 (and 0), scientific ``d.ddde±x``
    otherwise, always with a fractional part."""
    x = float(x)
    if math.isnan(x):
        return "NaN"
    if math.isinf(x):
        return "Inf" if x > 0 else "-Inf"
    if x == 0.0:
        return "-0.0" if math.copysign(1.0, x) < 0 else "0.0"
    sign = "-" if x < 0 else ""
    digits, exp10 = _shortest_digits(abs(x))   # value = 0.d1d2... * 10^exp10
    e = exp10 - 1                               # scientific exponent
    if -4 <= e < 6:
        if exp10 <= 0:
            s = "0." + "0" * (-exp10) + digits
        elif exp10 >= len(digits):
            s = digits + "0" * (exp10 - len(digits)) + ".0"
        else:
            s = digits[:exp10] + "." + digits[exp10:]
        return sign + s
    mant = digits[0] + "." + (digits[1:] if len(digits) > 1 else "0")
    return f"{sign}{mant}e{e}"


def _shortest_digits(v):
    """(digit string, exponent) of the shortest round-trip decimal of v > 0,
    from Python's repr (also shortest round-trip)."""
    r = repr(v)
    m = re.fullmatch(r"(\d+)(?:\.(\d*))?(?:e([+-]?\d+))?", r)
    ip, fp, ex = m.group(1), m.group(2) or "", int(m.group(3) or 0)
    digits = (ip + fp).lstrip("0")
    lead = len(ip) - (len(ip + fp) - len((ip + fp).lstrip("0")))
    exp10 = lead + ex
    digits = digits.rstrip("0") or "0"
    return digits, exp10


def write_dose_txt(path, dose, nx, ny):
    """``writedlm(path, Vec2Mat(nx, ny, dose))``: line i holds (i, 0..ny-1)."""
    m = np.asarray(dose, dtype=np.float64).reshape(nx, ny)
    with open(path, "w") as f:
        for row in m:
            f.write("\t".join(format_julia(v) for v in row) + "\n")


def read_dose_txt(path):
    """``readdlm(path, Float64)`` of a whitespace-delimited matrix."""
    with open(path) as f:
        return np.array([[float(t) for t in line.split()] for line in f if line.strip()])


# ---------------------------------------------------------------------------
# VTK XML RectilinearGrid (WriteVTK vtk_grid / vtk_save)
# ---------------------------------------------------------------------------
def _block(arr):
    raw = np.ascontiguousarray(arr, dtype="<f8").tobytes()
    comp = zlib.compress(raw, ZLIB_LEVEL)
    # one block: [nblocks, block size, last block size, compressed sizes...]
    return struct.pack("<4Q", 1, len(raw), len(raw), len(comp)) + comp


def write_vtr(path, x, y, point_data):
    """A 2D rectilinear grid with coordinates x (nx), y (ny), z = [0] and
    Float64 point arrays (name -> nx x ny matrix or i-fastest vector).
    ``path`` gets the ".vtr" suffix when it lacks one, as vtk_grid does."""
    if not str(path).endswith(".vtr"):
        path = f"{path}.vtr"
    x = np.asarray(x, dtype=np.float64)
    y = np.asarray(y, dtype=np.float64)
    nx, ny = len(x), len(y)
    arrays = [("x", x), ("y", y), ("z", np.zeros(1))]
    pdata = []
    for name, a in point_data.items():
        a = np.asarray(a, dtype=np.float64)
        if a.ndim == 2:
            if a.shape != (nx, ny):
                raise ValueError(f"point array {name!r} is {a.shape}, grid is {(nx, ny)}")
            a = a.ravel(order="F")                     # i fastest
        elif a.size != nx * ny:
            raise ValueError(f"point array {name!r} has {a.size} values for {nx * ny} points")
        pdata.append((name, a))
    blobs, offsets, off = [], [], 0
    for _, a in arrays + pdata:
        b = _block(a)
        offsets.append(off)
        blobs.append(b)
        off += len(b)

    def entry(name, o):
        return (f'        <DataArray type="Float64" Name="{name}" NumberOfComponents="1" '
                f'format="appended" offset="{o}"/>\n')

    ext = f"0 {nx - 1} 0 {ny - 1} 0 0"
    head = ('<?xml version="1.0" encoding="utf-8"?>\n'
            '<VTKFile type="RectilinearGrid" version="1.0" byte_order="LittleEndian" '
            'header_type="UInt64" compressor="vtkZLibDataCompressor">\n'
            f'  <RectilinearGrid WholeExtent="{ext}">\n'
            f'    <Piece Extent="{ext}">\n'
            '      <Coordinates>\n')
    head += "".join(entry(n, o) for (n, _), o in zip(arrays, offsets[:3]))
    head += '      </Coordinates>\n      <PointData>\n'
    head += "".join(entry(n, o) for (n, _), o in zip(pdata, offsets[3:]))
    head += ('      </PointData>\n    </Piece>\n  </RectilinearGrid>\n'
             '  <AppendedData encoding="raw">\n_')
    tail = '\n  </AppendedData>\n</VTKFile>'
    with open(path, "wb") as f:
        f.write(head.encode())
        for b in blobs:
            f.write(b)
        f.write(tail.encode())
    return path


def read_vtr(path):
    """(x, y, z, {name: i-fastest vector}) of a file written as above."""
    data = open(path, "rb").read()
    mark = b'<AppendedData encoding="raw">\n_'
    i = data.index(mark) + len(mark)
    head = data[:i].decode()
    app = data[i:]
    out = {}
    for name, off in re.findall(r'Name="([^"]+)"[^>]*offset="(\d+)"', head):
        o = int(off)
        nb, _, _ = struct.unpack("<3Q", app[o:o + 24])
        sizes = struct.unpack(f"<{nb}Q", app[o + 24:o + 24 + 8 * nb])
        p = o + 24 + 8 * nb
        raw = b""
        for s in sizes:
            raw += zlib.decompress(app[p:p + s])
            p += s
        out[name] = np.frombuffer(raw, "<f8").copy()
    x, y, z = out.pop("x"), out.pop("y"), out.pop("z")
    return x, y, z, out


def write_dose_outputs(prefix, settings, dose, u=None):
    """The reference driver's outputs for a solved dose (``main.jl:189-193``):
    ``{prefix}.txt`` (writedlm of the nx x ny dose) and ``{prefix}.vtr`` with
    "dose", "dose_normalized" and, when given, "u" (the scalar flux already
    scaled as ``u/0.5/sqrt(gamma[1])``)."""
    nx, ny = settings.NCellsX, settings.NCellsY
    d = np.asarray(dose, dtype=np.float64).reshape(nx, ny)
    write_dose_txt(f"{prefix}.txt", d, nx, ny)
    fields = {"dose": d, "dose_normalized": d / d.max()}
    if u is not None:
        fields["u"] = np.asarray(u, dtype=np.float64).reshape(nx, ny)
    return write_vtr(prefix, settings.xMid, settings.yMid, fields)
