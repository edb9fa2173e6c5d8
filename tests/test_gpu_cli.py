"""test_voronoi-compatible driver (rpd_voronoi): the same options and the
mg / xyz outputs (test_voronoi.cpp:158-214), checked against a solve through
the Python binding on the same inputs."""
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT

pytestmark = pytest.mark.gpu

CLI = os.path.join(ROOT, "gpu-restricted-power-diagrams_amd", "rpd_voronoi")


def test_cli_mg_output_matches_binding(rpd, ctx, tmp_path):
    n = 5000
    x = rpd.generate_seeds(rpd.SEEDS_UNIFORM, n, seed=21)
    pts = tmp_path / "pts.xyz"
    with open(pts, "w") as f:
        f.write("%d\n" % n)
        for p in x:
            f.write("%.17g %.17g %.17g\n" % (p[0], p[1], p[2]))
    grid = rpd.Grid.shape(rpd.SHAPE_SPHERE, 24, (0.5, 0.5, 0.5, 0.45))
    gfile = str(tmp_path / "sphere.grid")
    grid.save(gfile)
    r = subprocess.run([CLI, "--K=40", "--P=40", "--V=60", "--T=100", "--precision=fp64", "--grid=" + gfile,
                        "--output=mg", "--output=xyz", "--output=geogram", "--printInvariants", str(pts)],
                       cwd=str(tmp_path), capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert 'GTime["stat"]' in r.stdout
    assert "Sum of volumes" in r.stderr
    mg = np.loadtxt(tmp_path / "gridGPU_mg.xyzw", skiprows=1)
    assert mg.shape == (n, 4)
    loaded = rpd.Grid.load(gfile)
    dom = ctx.domain(loaded)
    ctx.set_seeds(x)
    ctx.solve(dom, K=40, P=40, V=60, T=100, precision=rpd.PREC_FP64)
    ref = ctx.moments()
    # geogram ascii (Utils.cpp:148-170) of the output cells: with a grid and
    # the default final keep distance 0 (test_voronoi.cpp:38, Options.cpp:42)
    # the used cells (volume >= keepVolume), in ascending seed index
    used = np.nonzero(ref[:, 3] >= 0.001)[0]
    m = used.size
    assert 0 < m < n
    geo = open(tmp_path / "gridGPU.geogram_ascii").read().split("\n")
    assert geo[:6] == ["[HEAD]", '"GEOGRAM"', '"1.0"', "[ATTS]", '"GEO::Mesh::vertices"', str(m)]
    w0 = 12 + 3 * m + 6
    assert geo[12 + 3 * m:w0] == ["[ATTR]", '"GEO::Mesh::vertices"', '"w"', '"double"', "8", "1"]
    gw = np.array([float(v) for v in geo[w0:w0 + m]])
    assert np.allclose(gw, ref[used, 3], rtol=1e-15, atol=0)
    dom.close()
    ref[ref[:, 3] <= 0.001] = 0.0
    d = np.abs(mg - ref)
    print("cli vs binding: max abs diff", d.max(axis=0), "cells differing", int((d.max(axis=1) > 0).sum()))
    assert np.allclose(mg, ref, rtol=1e-15, atol=1e-9)
    xyz = np.loadtxt(tmp_path / "gridGPU.xyz", skiprows=1)
    keep = ref[:, 3] >= 0.001
    assert xyz.shape[0] == keep.sum()
    assert np.allclose(xyz, ref[keep, :3] / ref[keep, 3:4], rtol=1e-5)


def test_cli_seed_formats_counters_and_volume(rpd, ctx, tmp_path):
    """Seed files: 4D ascii (--points4D, Utils.cpp:36-79) saved as binary
    (--saveBinaryPoints, Utils.cpp:100-113) and read back
    (--loadBinaryPoints, Utils.cpp:81-98) give the same diagram; --output=
    volume prints the sum of volumes (test_voronoi.cpp:168-182); the debug
    counter options add counter0..4 and counter10 to GTime["stat"]."""
    n = 4000
    h = (1.0 / n) ** (1 / 3)
    x = rpd.generate_seeds(rpd.SEEDS_UNIFORM, n, seed=23, wmax=0.25 * h * h * 1e6)
    pts = tmp_path / "pts4.xyzw"
    with open(pts, "w") as f:
        f.write("%d\n" % n)
        for p in x:
            f.write("%.17g %.17g %.17g %.17g\n" % tuple(p))
    grid = rpd.Grid.shape(rpd.SHAPE_TORUS, 32, (0.5, 0.5, 0.5, 0.3, 0.12))
    gfile = str(tmp_path / "torus.grid")
    grid.save(gfile)
    binf = str(tmp_path / "pts.bin")
    common = ["--K=50", "--P=40", "--V=60", "--precision=fp64", "--grid=" + gfile, "--output=mg", "--output=volume"]
    r1 = subprocess.run([CLI] + common + ["--points4D", "--saveBinaryPoints=" + binf, "--cVoroDebugCounterMax",
                                          "--cVoroDebugCounterHisto", "--cVoroDebugCounterStatus", str(pts)],
                        cwd=str(tmp_path), capture_output=True, text=True, timeout=120)
    assert r1.returncode == 0, r1.stderr
    mg1 = np.loadtxt(tmp_path / "gridGPU_mg.xyzw", skiprows=1)
    for c in ("counter0", "counter1", "counter2", "counter3", "counter4", "counter10"):
        assert '"%s":[' % c in r1.stdout, c
    # the binary file: size_t n, then 4n doubles
    raw = open(binf, "rb").read()
    assert len(raw) == 8 + 32 * n and np.frombuffer(raw[:8], np.uint64)[0] == n
    assert np.array_equal(np.frombuffer(raw[8:], np.float64).reshape(-1, 4), x)
    r2 = subprocess.run([CLI] + common + ["--loadBinaryPoints", binf], cwd=str(tmp_path), capture_output=True,
                        text=True, timeout=120)
    assert r2.returncode == 0, r2.stderr
    mg2 = np.loadtxt(tmp_path / "gridGPU_mg.xyzw", skiprows=1)
    assert np.array_equal(mg1, mg2)
    vols = [float(line.split("Sum of volumes:")[1].split()[0]) for line in r2.stderr.splitlines()
            if "Sum of volumes:" in line]
    loaded = rpd.Grid.load(gfile)
    dom = ctx.domain(loaded)
    ctx.set_seeds(x)
    ctx.solve(dom, K=50, P=40, V=60, precision=rpd.PREC_FP64)
    ref = ctx.moments()
    dom.close()
    want = ref[ref[:, 3] > 0.001, 3].sum()
    assert len(vols) == 1 and abs(vols[0] - want) <= 1e-9 * want
    assert abs(want - loaded.volume()) <= 1e-9 * want
    ref[ref[:, 3] <= 0.001] = 0.0
    assert np.allclose(mg2, ref, rtol=1e-15, atol=1e-9)
