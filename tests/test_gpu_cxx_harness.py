"""The C++ drop-in surface on the GPU: build/engine/flip2_tester is a harness
written against the reference's Particles member names (as the friend class
ParticleSystemTester, engine/particles.hu:94), the CudaVec method set on
DeviceVec (host mirror, stream upload/download, zeroDeviceAsync, getMax with
the reference's sampled semantics, swapDevicePtr) and the algorithm free
functions with the reference's signatures (algorithms.hpp).  It checks each
result against a host computation (binning with the float-narrowed floor,
stable order, key counts, node segments, row table, the sampled maximum over
every tile, std::stable_sort, std::partial_sum) or against the Particles
stage that computes the same thing, runs frames and reports the device
pool's growth; exit status 0 means every check held."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

EXE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "build", "engine", "flip2_tester")


@pytest.mark.parametrize("n,cells,seed", [(20000, 24, 7), (3000, 16, 11), (40000, 28, 3)])
def test_cxx_harness(n, cells, seed):
    assert os.path.exists(EXE), "build/engine/flip2_tester is not built"
    r = subprocess.run([EXE, str(n), str(cells), str(seed)], capture_output=True, text=True, timeout=170)
    print(r.stdout, r.stderr)
    assert r.returncode == 0, r.stderr
    assert "tester ok" in r.stdout
