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
    # geogram ascii (Utils.cpp:148-170): n vertices, barycentres then w
    geo = open(tmp_path / "gridGPU.geogram_ascii").read().split("\n")
    assert geo[:6] == ["[HEAD]", '"GEOGRAM"', '"1.0"', "[ATTS]", '"GEO::Mesh::vertices"', str(n)]
    w0 = 12 + 3 * n + 6
    assert geo[12 + 3 * n:w0] == ["[ATTR]", '"GEO::Mesh::vertices"', '"w"', '"double"', "8", "1"]
    gw = np.array([float(v) for v in geo[w0:w0 + n]])
    loaded = rpd.Grid.load(gfile)
    dom = ctx.domain(loaded)
    ctx.set_seeds(x)
    ctx.solve(dom, K=40, P=40, V=60, T=100, precision=rpd.PREC_FP64)
    ref = ctx.moments()
    assert np.allclose(gw, ref[:, 3], rtol=1e-15, atol=0)
    dom.close()
    ref[ref[:, 3] <= 0.001] = 0.0
    d = np.abs(mg - ref)
    print("cli vs binding: max abs diff", d.max(axis=0), "cells differing", int((d.max(axis=1) > 0).sum()))
    assert np.allclose(mg, ref, rtol=1e-15, atol=1e-9)
    xyz = np.loadtxt(tmp_path / "gridGPU.xyz", skiprows=1)
    keep = ref[:, 3] >= 0.001
    assert xyz.shape[0] == keep.sum()
    assert np.allclose(xyz, ref[keep, :3] / ref[keep, 3:4], rtol=1e-5)
