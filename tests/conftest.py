import importlib.util
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through librpd's HIP kernels)")


def _load(name, path):
    if name in sys.modules:
        return sys.modules[name]
    spec = importlib.util.spec_from_file_location(name, path)
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


@pytest.fixture(scope="session")
def rpd():
    return _load("rpd_amd", os.path.join(ROOT, "gpu-restricted-power-diagrams_amd", "__init__.py"))


@pytest.fixture(scope="session")
def orc():
    return _load("rpd_oracle", os.path.join(ROOT, "oracle", "oracle.py"))


@pytest.fixture(scope="session")
def ctx(rpd):
    c = rpd.Context(0)
    yield c
    c.close()


def oracle_grid(orc, grid):
    return orc.Grid(grid.size, grid.in_domain, grid.points, grid.triangles, grid.tri_list, grid.offsets,
                    grid.centre)


def moment_errors(got, ref, n):
    """Max errors normalised by the mean cell volume 1e9/n (volume) and by
    mean cell volume x frame size 1000 (first moments)."""
    v = 1e9 / n
    return (float(np.abs(got[:, 3] - ref[:, 3]).max() / v),
            float(np.abs(got[:, :3] - ref[:, :3]).max() / (v * 1000.0)))


def cell_errors(got, ref, n, floor=0.01):
    """Per-result errors over the cells whose reference volume exceeds floor
    x the mean cell volume 1e9/n: (largest relative volume error, largest
    barycentre displacement over the cell's size (cube root of its volume),
    number of cells compared).  Below the floor a cell is a sliver of a
    restricted cell, whose relative error grows as the absolute rounding of
    its pieces (a few ulps of the whole cell's volume) over its own volume."""
    v = 1e9 / n
    big = ref[:, 3] > floor * v
    if not big.any():
        return 0.0, 0.0, 0
    m = ref[big, 3]
    rv = float((np.abs(got[big, 3] - m) / m).max())
    bg = got[big, :3] / got[big, 3:4]
    br = ref[big, :3] / ref[big, 3:4]
    rb = float((np.abs(bg - br).max(1) / np.cbrt(m)).max())
    return rv, rb, int(big.sum())
