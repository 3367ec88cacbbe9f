"""Dose outputs in the reference's formats (``titus_amd/io.py``).

Fixtures: two of the reference's own output pairs,
``code/output/dose_csd_nx{20,50}ny{20,50}.{txt,vtr}`` (``tests/golden/``),
written by ``writedlm`` (DelimitedFiles, Julia 1.8.1) and ``vtk_grid`` /
``vtk_save`` (WriteVTK 1.16.0) in ``code/main.jl:189-193`` / ``code/plot.jl``.
The pin is byte identity: reading a fixture and writing it back must give the
same file, which fixes the number format, the delimiter, the XML layout, the
block headers, the array order (i fastest) and the zlib level at once.
"""
import math
import os

import numpy as np
import pytest

from titus_amd import io as tio

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


@pytest.mark.parametrize("n", [20, 50])
def test_dose_text_round_trip_is_byte_identical(n, tmp_path):
    src = os.path.join(GOLDEN, f"dose_csd_nx{n}ny{n}.txt")
    m = tio.read_dose_txt(src)
    assert m.shape == (n, n)
    out = tmp_path / "d.txt"
    tio.write_dose_txt(out, m.reshape(-1), n, n)
    assert out.read_bytes() == open(src, "rb").read()


@pytest.mark.parametrize("n", [20, 50])
def test_vtr_round_trip_is_byte_identical(n, tmp_path):
    src = os.path.join(GOLDEN, f"dose_csd_nx{n}ny{n}.vtr")
    x, y, z, fields = tio.read_vtr(src)
    assert len(x) == len(y) == n and np.array_equal(z, [0.0])
    p = tio.write_vtr(tmp_path / "d", x, y, fields)
    assert open(p, "rb").read() == open(src, "rb").read()


def test_vtr_dose_is_the_text_matrix_i_fastest():
    """The 20 x 20 pair was written from one dose (plot.jl): the VTK array is
    the text matrix with i fastest, and dose_normalized = dose / max."""
    m = tio.read_dose_txt(os.path.join(GOLDEN, "dose_csd_nx20ny20.txt"))
    _, _, _, f = tio.read_vtr(os.path.join(GOLDEN, "dose_csd_nx20ny20.vtr"))
    assert np.array_equal(f["dose"], m.ravel(order="F"))
    assert np.array_equal(f["dose_normalized"], m.ravel(order="F") / m.max())


@pytest.mark.parametrize("x, s", [
    (0.0, "0.0"), (-0.0, "-0.0"), (1.0, "1.0"), (100000.0, "100000.0"),
    (1e6, "1.0e6"), (999999.9999999999, "999999.9999999999"), (1e-4, "0.0001"),
    (9.999999999999999e-5, "9.999999999999999e-5"), (5e-324, "5.0e-324"),
    (1.5e300, "1.5e300"), (-2.5e-7, "-2.5e-7"), (0.1, "0.1"), (123.456, "123.456"),
    (1e16, "1.0e16"), (12345678.9, "1.23456789e7"),
])
def test_julia_float_printing(x, s):
    assert tio.format_julia(x) == s
    assert float(s) == x and math.copysign(1, float(s)) == math.copysign(1, x)


def test_driver_outputs_from_a_solver_dose(tmp_path):
    """write_dose_outputs: the solver's c = i*ny + j dose vector to the
    driver's text and VTK files, read back exactly."""
    from titus_amd.settings import water_slab_pencil_beam
    s = water_slab_pencil_beam(n=11, N=3, r=4, steps=5)
    nx, ny = s.NCellsX, s.NCellsY
    rng = np.random.default_rng(0)
    dose = rng.standard_normal(nx * ny) * 1e3
    u = rng.standard_normal(nx * ny)
    p = tio.write_dose_outputs(str(tmp_path / "dose_csd"), s, dose, u=u)
    m = tio.read_dose_txt(str(tmp_path / "dose_csd.txt"))
    assert np.array_equal(m, dose.reshape(nx, ny))
    x, y, _, f = tio.read_vtr(p)
    assert np.array_equal(x, s.xMid) and np.array_equal(y, s.yMid)
    assert np.array_equal(f["dose"], dose.reshape(nx, ny).ravel(order="F"))
    assert np.array_equal(f["u"], u.reshape(nx, ny).ravel(order="F"))
    assert list(f) == ["dose", "dose_normalized", "u"]


# ---------------------------------------------------------------------------
# image phantoms (settings.jl:266-310)
# ---------------------------------------------------------------------------
def test_image_density_index_map():
    """An 8 x 6 image on a 4 x 3 grid: floor(i/4*8) = 2, 4, 6, 8 and
    floor(j/3*6) = 2, 4, 6 (1-based), i.e. rows 1, 3, 5, 7 and columns 1, 3, 5."""
    from titus_amd.settings import image_density
    img = np.arange(48, dtype=np.float64).reshape(8, 6) / 100.0
    d = image_density(img, 4, 3, 0.05)
    want = np.maximum(1.85 * img[np.ix_([1, 3, 5, 7], [1, 3, 5])], 0.05)
    assert np.array_equal(d, want)
    assert d[0, 0] == 1.85 * img[1, 1] and d[3, 2] == 1.85 * img[7, 5]


def test_lung_preset_reads_the_image():
    from titus_amd.settings import Settings, image_gray
    path = os.path.join(GOLDEN, "Lung.png")
    g = image_gray(path)
    assert g.shape == (1069, 1068)                     # rows x columns, as load() gives
    k = np.rint(g * 255.0)
    assert np.array_equal(g, k / 255.0)
    s = Settings(72, 72, 20, "lung", "Electrons", phantomImage=path)
    assert s.density.shape == (71, 71)
    assert s.density.min() >= 0.05 and s.density.max() <= 1.85
    vals = np.unique(s.density)
    ok = np.isclose(vals[:, None], np.maximum(1.85 * (np.arange(256) / 255.0), 0.05)[None, :],
                    rtol=0, atol=0).any(axis=1)
    assert ok.all()
    assert s.b == s.d == 14.5 and s.eMax == math.sqrt(21.0 ** 2 + 0.5 ** 2)
    assert s.dE == s.cfl * min(s.dx, s.dy) * s.density.min()
