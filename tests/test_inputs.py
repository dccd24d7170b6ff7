"""CPU tests of the reference's UVW text input (grid.cu:338-366): every
value equals the reference's float arithmetic, x = (double)(float)(u *
(float)N) / 2048, restated here with numpy float32."""
import os

import numpy as np
import pytest

import migrid

FIX = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "uvw_in_sample.dat")


def numpy_parse(path, N):
    rows = []
    for line in open(path):
        if not line.strip():
            continue
        ints, floats = line.split(":")
        rows.append([np.float32(x) for x in floats.split(",")])
    a = np.array(rows, dtype=np.float32)
    uv = np.stack([(a[:, 0] * np.float32(N)).astype(np.float64) / 2048.0,
                   (a[:, 1] * np.float32(N)).astype(np.float64) / 2048.0], 1)
    return uv, a[:, 2].astype(np.float64)


@pytest.mark.parametrize("N", [2048, 4096, 1000])
def test_uvw_text_matches_reference_arithmetic(N):
    uv, w = migrid.read_uvw_text(FIX, N)
    want_uv, want_w = numpy_parse(FIX, N)
    assert uv.shape == (40, 2)
    np.testing.assert_array_equal(uv, want_uv)
    np.testing.assert_array_equal(w, want_w)


def test_uvw_text_limits_and_errors(tmp_path):
    uv, _ = migrid.read_uvw_text(FIX, 2048, max_rows=5)
    assert uv.shape == (5, 2)
    bad = tmp_path / "bad.dat"
    bad.write_text("0,1,0: 1.0, 2.0, 3.0\n1,2: 1.0, 2.0\n")
    with pytest.raises(migrid.MigridError) as e:
        migrid.read_uvw_text(str(bad), 2048)
    assert "line 2" in str(e.value)
    # the reported line counts blank lines too
    bad.write_text("0,1,0: 1.0, 2.0, 3.0\n\n\n1,2: 1.0, 2.0\n")
    with pytest.raises(migrid.MigridError) as e:
        migrid.read_uvw_text(str(bad), 2048)
    assert "line 4" in str(e.value)
    with pytest.raises(migrid.MigridError):
        migrid.read_uvw_text(str(tmp_path / "missing.dat"), 2048, max_rows=3)
