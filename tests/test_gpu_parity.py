"""GPU parity: the HIP insertion path against the FP64 oracle.

Bar: per voxel |gpu - oracle| <= 1e-5 of the oracle cube's peak, and total
flux within 1e-6, at workload sizes the oracle finishes in seconds.
"""

import numpy as np
import pytest

from martini_amd import synthetic
from tests.helpers import FLUX_TOL, VOXEL_TOL, gpu_cube, oracle_cube, parity

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("method", ["plain", "mfma"])
def test_disc1e4_64px(method):
    w = synthetic.WORKLOADS["disc1e4_64px"]
    ref, P = oracle_cube(w)
    gpu, m = gpu_cube(w, method=method)
    voxel, flux = parity(gpu, ref)
    print(f"disc1e4_64px {method}: voxel {voxel:.3e} flux {flux:.3e}")
    assert voxel <= VOXEL_TOL
    assert flux <= FLUX_TOL


@pytest.mark.parametrize("method", ["plain", "mfma"])
def test_disc1e6_256px(method):
    w = synthetic.WORKLOADS["disc1e6_256px"]
    ref, P = oracle_cube(w)
    gpu, m = gpu_cube(w, method=method)
    voxel, flux = parity(gpu, ref)
    print(f"disc1e6_256px {method}: voxel {voxel:.3e} flux {flux:.3e}")
    assert voxel <= VOXEL_TOL
    assert flux <= FLUX_TOL


@pytest.mark.parametrize("method", ["plain", "mfma"])
def test_deterministic(method):
    w = synthetic.WORKLOADS["disc1e4_64px"]
    a, _ = gpu_cube(w, method=method)
    b, _ = gpu_cube(w, method=method)
    assert np.array_equal(a, b)
