"""Golden fixtures (tests/golden, written by tests/golden/make_golden.py):
the oracle must reproduce them on the CPU, the HIP path on the GPU."""
import gzip
import hashlib
import json
import os

import numpy as np
import pytest

from conftest import ROOT, cell_errors, moment_errors, oracle_grid

GOLD = os.path.join(ROOT, "tests", "golden")


def _ref_seeds(name):
    with gzip.open(os.path.join(GOLD, name), "rt") as f:
        p = np.loadtxt(f, skiprows=1)
    x = np.zeros((p.shape[0], 4))
    x[:, :3] = p
    return x


def _c1(rpd):
    x = rpd.generate_seeds(rpd.SEEDS_UNIFORM, 10000, seed=0)
    grid = rpd.Grid.shape(rpd.SHAPE_SPHERE, 32, (0.5, 0.5, 0.5, 0.45))
    return x, grid, np.load(os.path.join(GOLD, "golden_c1.npz"))


def _orig_lists(perm, kn):
    out = np.full(kn.shape, -1, np.int32)
    ok = kn >= 0
    out[perm] = np.where(ok, perm[np.where(ok, kn, 0)], -1)
    return out


def test_oracle_reproduces_bruteforce_golden(rpd, orc):
    z = np.load(os.path.join(GOLD, "golden_bruteforce.npz"))
    x, g = z["seeds"], int(z["g"])
    grid = rpd.Grid.shape(rpd.SHAPE_SPHERE, g, tuple(z["sphere"]))
    out, st, _, _ = orc.solve(x, oracle_grid(orc, grid), K=x.shape[0] + 4, P=200, V=250, T=4000)
    ev, em = moment_errors(out, z["moments"], x.shape[0])
    assert ev < 1e-11 and em < 1e-11


def test_oracle_reproduces_c1_golden(rpd, orc):
    x, grid, z = _c1(rpd)
    for name, prec in (("fp32", 0), ("fp64", 1)):
        m, st, _, _ = orc.solve(x, oracle_grid(orc, grid), K=50, P=30, V=40, T=100, precision=prec, vhs0=2, vhs1=2)
        assert np.array_equal(st, z["status_" + name].astype(np.int32))
        ev, em = moment_errors(m, z["moments_" + name], x.shape[0])
        assert ev < 1e-13 and em < 1e-13


def test_oracle_reproduces_knn_golden(orc):
    want = json.load(open(os.path.join(GOLD, "golden_knn.json")))
    for key, h in want.items():
        fname, _, prec = key.split(":")
        x = _ref_seeds(fname)
        perm, kn, _ = orc.knn(x, 20, precision=int(prec[-1]))
        assert hashlib.sha256(_orig_lists(perm, kn).tobytes()).hexdigest() == h, key


@pytest.mark.gpu
def test_gpu_c1_golden(rpd, orc, ctx):
    """The CPU configuration against its committed fixtures (of the mean cell
    volume), and per cell against the FP64 oracle in each seed's frame."""
    x, grid, z = _c1(rpd)
    dom = ctx.domain(grid)
    ctx.set_seeds(x)
    ref1, _, _, _ = orc.solve(x, oracle_grid(orc, grid), K=50, P=30, V=40, T=100, frame=1)
    for prec, ref, tol in ((rpd.PREC_FP64, "moments_fp64", 1e-12), (rpd.PREC_MIXED, "moments_fp64", 1e-6),
                           (rpd.PREC_FP32, "moments_fp32", 1e-4)):
        ctx.solve(dom, K=50, P=30, V=40, T=100, precision=prec)
        got = ctx.moments()
        ev, em = moment_errors(got, z[ref], x.shape[0])
        assert ev < tol and em < tol, (prec, ev, em)
        if prec != rpd.PREC_FP32:
            cv, cb, nc = cell_errors(got, ref1, x.shape[0])
            print("C1", prec, "per cell", cv, cb, nc, flush=True)
            assert cv <= tol and cb <= tol and nc > 1000, (prec, cv, cb, nc)
    dom.close()


@pytest.mark.gpu
def test_gpu_knn_golden_on_reference_seed_files(ctx):
    want = json.load(open(os.path.join(GOLD, "golden_knn.json")))
    for key, h in want.items():
        fname, _, prec = key.split(":")
        x = _ref_seeds(fname)
        ctx.set_seeds(x)
        kn = ctx.knn(20, precision=int(prec[-1]))
        got = _orig_lists(ctx.perm(), kn)
        assert hashlib.sha256(got.tobytes()).hexdigest() == h, key
