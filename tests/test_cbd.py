"""CBD snapshot files (src/fileio.jl), restating test/output_test.jl:1-50 for
one process: write in float32 and float64, read back, compare the header and
the values; read a sub-range of levels."""

import numpy as np
import pytest

from conftest import load_package

B = load_package()
cbd = B.cbd


@pytest.mark.parametrize("T", [np.float32, np.float64])
@pytest.mark.parametrize("T_file", [np.float32, np.float64])
@pytest.mark.parametrize("ns", ["C", "I"])
def test_file_io(tmp_path, T, T_file, ns):
    rng = np.random.default_rng(73330459108)
    n1, n2, n3 = 6, 8, 12
    ds = rng.random(3)
    x1 = np.linspace(0, ds[0], n1 + 1)[:n1]
    x2 = np.linspace(0, ds[1], n2 + 1)[:n2]
    x3 = np.linspace(0, ds[2], 2 * n3 + 1)[(1 if ns == "C" else 2):-1:2]
    vel = (x1[:, None, None] / ds[0] + x2[None, :, None] / ds[1] + x3[None, None, :] / ds[2]).astype(T)
    fn = str(tmp_path / f"vel-{np.dtype(T_file).name}-{ns}.cbd")
    cbd.write_cbd(T_file, fn, vel, x1, x2, x3, (0, 0, 0), tuple(ds))
    xmin, xmax, x1r, x2r, x3r, data = cbd.read_cbd(fn)
    assert xmin == (0, 0, 0) and np.array_equal(xmax, ds)
    assert np.array_equal(x1r, x1) and np.array_equal(x2r, x2) and np.array_equal(x3r, x3)
    assert data.dtype == np.dtype(T_file)
    if T == T_file:
        assert np.array_equal(data, vel)
    else:
        assert np.allclose(data, vel, rtol=1e-6)


def test_header_layout_and_subrange(tmp_path):
    vals = np.arange(2 * 3 * 4, dtype=np.float64).reshape((2, 3, 4), order="F")
    fn = str(tmp_path / "a.cbd")
    cbd.write_cbd(np.float64, fn, vals, [0, 0.5], [0, 1, 2], [0.1, 0.2, 0.3, 0.4], (0, 0, 0), (1, 3, 0.5))
    raw = open(fn, "rb").read()
    assert raw[:8] == bytes([0xCB, 0xDF, 1, 0, 0, 0, 0, 8])
    assert len(raw) == 8 + 24 + 48 + 8 * (2 + 3 + 4) + 8 * 24
    _, _, _, _, x3, data = cbd.read_cbd(fn, (0.2, 0.3))
    assert np.array_equal(x3, [0.2, 0.3])
    assert np.array_equal(data, vals[:, :, 1:3])
    with open(fn, "r+b") as f:
        f.write(b"\x00")
    with pytest.raises(ValueError, match="magic"):
        cbd.read_cbd(fn)


def test_positions():
    assert np.allclose(cbd.hrange(2.0, 4, False), [0, 0.5, 1.0, 1.5])
    assert np.allclose(cbd.hrange(2.0, 4, True), [0.25, 0.75, 1.25, 1.75])
