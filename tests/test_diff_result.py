"""rpd_diff_result (the report of the reference's utils/diffResult,
main.cpp:63-119) on two synthetic result files, against the same metrics
computed here with numpy.  CPU only."""
import os
import re
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TOOL = os.path.join(ROOT, "gpu-restricted-power-diagrams_amd", "rpd_diff_result")


def _write(path, a):
    with open(path, "w") as f:
        f.write("%d\n" % a.shape[0])
        for r in a:
            f.write("%.17g %.17g %.17g %.17g\n" % tuple(r))


def _nums(line):
    return [float(v) for v in re.findall(r"[-+]?\d+\.?\d*(?:[eE][-+]?\d+)?", line.split(":", 1)[1])]


def test_diff_result_metrics(tmp_path):
    rng = np.random.default_rng(3)
    n = 500
    v = rng.uniform(0.5, 2.0, n)
    a = np.column_stack([rng.uniform(0, 1000, (n, 3)) * v[:, None], v])
    b = a * (1 + rng.normal(0, 1e-6, a.shape))
    b[7, 3] = 0.0  # a cell without volume in one file: compared for volume only
    _write(tmp_path / "a.xyzw", a)
    _write(tmp_path / "b.xyzw", b)
    out = subprocess.run([TOOL, str(tmp_path / "a.xyzw"), str(tmp_path / "b.xyzw")], capture_output=True, text=True,
                         check=True).stdout.splitlines()
    vols = _nums(out[1])
    assert np.allclose(vols, [a[:, 3].sum(), b[:, 3].sum(), b[:, 3].sum() - a[:, 3].sum()], rtol=1e-12, atol=1e-12)
    ok = (a[:, 3] > 0) & (b[:, 3] > 0)
    ba, bb = a[ok, :3] / a[ok, 3:], b[ok, :3] / b[ok, 3:]
    assert np.isclose(_nums(out[4])[0], np.abs(b[:, 3] - a[:, 3]).max(), rtol=1e-12)
    assert np.allclose(_nums(out[5]), np.abs(bb - ba).max(axis=0), rtol=1e-9)
    assert np.allclose(_nums(out[6]), np.abs(b[ok, :3] - a[ok, :3]).max(axis=0), rtol=1e-9)
    assert np.isclose(_nums(out[8])[0], np.abs(b[ok, 3] - a[ok, 3]).mean(), rtol=1e-9)
    assert np.allclose(_nums(out[9]), np.abs(bb - ba).mean(axis=0), rtol=1e-9)


def test_diff_result_rejects_mismatched_files(tmp_path):
    _write(tmp_path / "a.xyzw", np.ones((3, 4)))
    _write(tmp_path / "b.xyzw", np.ones((4, 4)))
    r = subprocess.run([TOOL, str(tmp_path / "a.xyzw"), str(tmp_path / "b.xyzw")], capture_output=True, text=True)
    assert r.returncode != 0 and "same numbers" in r.stderr
