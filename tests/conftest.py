"""Shared test setup: import paths, the `gpu` marker, meshes and the RHS tolerance."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "pihm-cuda_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


_MESHES = {}


def mesh_cached(key, factory):
    if key not in _MESHES:
        _MESHES[key] = factory()
    return _MESHES[key]


@pytest.fixture(scope="session")
def c0():
    from pihm_amd.meshgen import make_watershed
    return mesh_cached("c0", lambda: make_watershed(535, seed=1, n_seg=20, network="valley"))


@pytest.fixture(scope="session")
def tree_small():
    from pihm_amd.meshgen import make_watershed
    return mesh_cached("tree", lambda: make_watershed(6000, seed=7, n_seg=150, network="tree"))


def assert_rhs_close(got, ref, mag, rtol=1e-12, what="ydot"):
    """Per-component parity bound for RHS values:
        |got - ref| <= rtol * max(|ref|, sum of |terms| assembled into the component).
    Pure relative error is ill-posed where the assembled fluxes cancel to ~0, so the
    bound is relative to the magnitude of the terms (mag), as computed by the oracle."""
    got = np.asarray(got)
    ref = np.asarray(ref)
    scale = np.maximum(np.abs(ref), np.asarray(mag))
    err = np.abs(got - ref)
    bad = ~(err <= rtol * scale)
    bad &= ~((err == 0) & (scale == 0))
    bad &= ~(np.isnan(got) & np.isnan(ref))      # NaN where the reference has NaN
    if bad.any():
        i = np.nonzero(bad)[0]
        raise AssertionError(f"{what}: {len(i)} components outside {rtol} (first {i[:5]}: "
                             f"got {got[i[:5]]} ref {ref[i[:5]]} scale {scale[i[:5]]})")


GOLDEN = os.path.join(ROOT, "tests", "golden", "example")


@pytest.fixture(scope="session")
def example():
    """The reference's own 535-element / 20-segment example watershed
    (benchmark/input/example), read from the committed copy of its input files."""
    from pihm_amd.inputs import read_project
    return mesh_cached("example", lambda: read_project(GOLDEN, "example", os.path.join(GOLDEN, "vegprmt.tbl")))
