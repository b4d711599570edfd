"""Host parts of the StepTimer and ProgressMonitor outputs (src/logging.jl:
81-308) that need no GPU: the JSON file of the timer and the zero-padded
layer transform of the monitor's wall-stress sample."""

import json
import math

import numpy as np

from conftest import load_package

B = load_package()


def test_step_timer_json(tmp_path):
    path = tmp_path / "a" / "b" / "timestamps.json"
    st = B.StepTimer(path=str(path), frequency=4)
    for t in (0.0, 0.5, 1.25):
        st.stamp(t)
    st.flush()
    text = path.read_text()
    assert text.startswith("{\n  \"frequency\": 4,\n  \"wallTime\": [")
    d = json.loads(text)
    assert d["frequency"] == 4 and d["simulationTime"] == [0.0, 0.5, 1.25]
    assert len(d["wallTime"]) == 3 and all(w >= 0 for w in d["wallTime"])


def test_layer_physical_matches_the_fourier_series():
    """Coefficients of 2 + cos(x) - 0.5 sin(2y) + 0.25 cos(x + y) (complex
    amplitudes of the unnormalised series, negative k2 at the end) on a larger
    grid equal the series evaluated there."""
    K1, K2 = 3, 5  # k1 0..2, k2 -2..2
    fd = np.zeros((K1, K2), dtype=np.complex128)
    fd[0, 0] = 2
    fd[1, 0] = 0.5  # cos(x) = (e^{ix} + e^{-ix}) / 2, only k1 >= 0 stored
    fd[0, 2] = 0.25j  # -0.5 sin(2y) = 0.25j e^{2iy} - 0.25j e^{-2iy}
    fd[0, 3] = -0.25j
    fd[1, 1] = 0.125  # 0.25 cos(x + y)
    for n1, n2 in ((6, 6), (9, 8), (12, 10)):
        x = 2 * math.pi * np.arange(n1) / n1
        y = 2 * math.pi * np.arange(n2) / n2
        X, Y = np.meshgrid(x, y, indexing="ij")
        want = 2 + np.cos(X) - 0.5 * np.sin(2 * Y) + 0.25 * np.cos(X + Y)
        np.testing.assert_allclose(B.model._layer_physical(fd, n1, n2), want, rtol=0, atol=1e-14)
