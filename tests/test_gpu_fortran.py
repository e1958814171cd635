"""The Fortran driver (namelist case file -> ISO_C_BINDING -> HIP kernels)
against the oracle: same case, same steps, bit-identical final state."""
import os
import subprocess

import numpy as np
import pytest

import _oracle as O
from fvmgpu import cases

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "fvm-3d-f90_amd", "bin", "fvm_driver")
CASE = os.path.join(ROOT, "tests", "cases", "riemann2d_101.nml")


def test_fortran_driver_riemann2d(tmp_path):
    assert os.path.exists(DRIVER), "build the Fortran driver: make -C fvm-3d-f90_amd"
    out = tmp_path / "q.bin"
    text = open(CASE).read().replace('"gpurun_out/fortran_riemann2d_101.bin"', f'"{out}"')
    nml = tmp_path / "case.nml"
    nml.write_text(text)
    print("case fortran driver riemann2d 101x101, 20 steps", flush=True)
    r = subprocess.run([DRIVER, f"--input={nml}"], capture_output=True, text=True, timeout=120)
    print(r.stdout[-2000:], r.stderr[-2000:])
    assert r.returncode == 0
    raw = open(out, "rb").read()
    hdr = np.frombuffer(raw[:24], dtype=np.int32, count=5)
    ni, nj, nk, ng, steps = (int(v) for v in hdr)
    assert (ni, nj, nk, ng, steps) == (101, 101, 1, 3, 20)
    t_end = np.frombuffer(raw[20:28], dtype=np.float64)[0]
    assert abs(t_end - 20 * 2e-4) < 1e-15
    q = np.frombuffer(raw[28:], dtype=np.float64).reshape(5, 1, nj + 6, ni + 6)
    p, dx = cases.riemann2d(101, seed=0)
    o = O.oracle_solve(p, 2, (dx, dx, 1.0), dt=2e-4, nsteps=20, p_from_q=1)
    assert np.array_equal(q[:, 0, 3:-3, 3:-3], o["q"][:, 0, 3:-3, 3:-3])
    # ghost faces hold the zeroth-order extrapolation
    np.testing.assert_array_equal(q[:, 0, 3:-3, 0], q[:, 0, 3:-3, 3])
