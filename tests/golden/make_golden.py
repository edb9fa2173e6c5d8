"""Writes the golden fixtures of tests/golden (run from the repository root:
python tests/golden/make_golden.py).  The reference ships no expected outputs,
so the fixtures are:

- golden_bruteforce.npz: 40 power-weighted seeds, sphere domain at 8^3 --
  restricted volumes and first moments from brute-force polytope clipping
  (tests/polytope.py), independent of the oracle;
- golden_c1.npz: the CPU-runnable configuration (10,000 uniform seeds, seed 0,
  sphere r 0.45 at 32^3, K50 P30 V40, w = 0) solved by the oracle in its FP32
  (VHSAlgo 2) and FP64 forms: moments and status per seed;
- golden_knn.json: sha256 of the oracle's canonical neighbour lists (original
  ids, int32) for the reference's own seed files data/100K-blue.xyz and
  data/100K-perturbed-grid.xyz (kept here gzipped), K = 20, FP32 and FP64.
"""
import gzip
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

import __graft_entry__ as ge  # noqa: E402
import polytope as pt  # noqa: E402


def knn_original(orc, x, K, prec):
    perm, kn, _ = orc.knn(x, K, precision=prec)
    out = np.full(kn.shape, -1, np.int32)
    ok = kn >= 0
    out[perm] = np.where(ok, perm[np.where(ok, kn, 0)], -1)
    return out


def load_ref_seeds(name):
    with gzip.open(os.path.join(HERE, name), "rt") as f:
        p = np.loadtxt(f, skiprows=1)
    x = np.zeros((p.shape[0], 4))
    x[:, :3] = p
    return x


def main():
    rpd = ge.load_package()
    orc = ge.load_oracle()

    n, g = 40, 8
    prm = (0.5, 0.5, 0.5, 0.42)
    x = rpd.generate_seeds(rpd.SEEDS_UNIFORM, n, seed=3, wmax=0.25 * (1.0 / n) ** (2 / 3) * 1e6)
    mask = rpd.voxel_mask(rpd.SHAPE_SPHERE, g, prm)
    bf = np.array([pt.restricted_cell(pt.power_cell(x, i), mask, g) for i in range(n)])
    np.savez_compressed(os.path.join(HERE, "golden_bruteforce.npz"), seeds=x, g=g, sphere=np.array(prm), moments=bf)

    n = 10000
    x = rpd.generate_seeds(rpd.SEEDS_UNIFORM, n, seed=0)
    grid = rpd.Grid.shape(rpd.SHAPE_SPHERE, 32, (0.5, 0.5, 0.5, 0.45))
    og = orc.Grid(grid.size, grid.in_domain, grid.points, grid.triangles, grid.tri_list, grid.offsets, grid.centre)
    out = {}
    for name, prec in (("fp32", 0), ("fp64", 1)):
        m, st, _, _ = orc.solve(x, og, K=50, P=30, V=40, T=100, precision=prec, vhs0=2, vhs1=2)
        out["moments_" + name] = m
        out["status_" + name] = st.astype(np.int8)
    np.savez_compressed(os.path.join(HERE, "golden_c1.npz"), **out)

    hashes = {}
    for fname in ("ref_100K-blue.xyz.gz", "ref_100K-perturbed-grid.xyz.gz"):
        x = load_ref_seeds(fname)
        for prec in (0, 1):
            lists = knn_original(orc, x, 20, prec)
            hashes["%s:K20:prec%d" % (fname, prec)] = hashlib.sha256(lists.tobytes()).hexdigest()
    with open(os.path.join(HERE, "golden_knn.json"), "w") as f:
        json.dump(hashes, f, indent=1, sort_keys=True)
    print("golden fixtures written", sorted(os.listdir(HERE)))


if __name__ == "__main__":
    main()
