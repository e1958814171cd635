"""The Fortran driver (namelist case file -> ISO_C_BINDING -> HIP kernels)
against the oracle: same case, same steps, bit-identical final state."""
import os
import subprocess

import numpy as np
import pytest

import _oracle as O
import fvmgpu
from fvmgpu import cases

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "fvm-3d-f90_amd", "bin", "fvm_driver")
CASE = os.path.join(ROOT, "tests", "cases", "riemann2d_101.nml")


def test_fortran_driver_riemann2d(tmp_path):
    assert os.path.exists(DRIVER), "build the Fortran driver: make -C fvm-3d-f90_amd"
    out = tmp_path / "q.bin"
    text = open(CASE).read().replace('"fortran_riemann2d_101.bin"', f'"{out}"')
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


def test_fortran_driver_output_snapshots(tmp_path):
    """&Output variables at the output times of &Time_Stepping (the reference's
    write_flowfield, src/solve/class_output.f90:204-237): the initial field and
    every output_time, each equal to the oracle's state at that step."""
    prefix = tmp_path / "flow"
    text = open(CASE).read()
    text = text.replace('"fortran_riemann2d_101.bin"', '""')
    text = text.replace("output_time  = 0.01", "output_time  = 1.0E-3")
    text = text.replace("print_every = 5", f'print_every = 5\n  snapshot_prefix = "{prefix}"')
    nml = tmp_path / "case.nml"
    nml.write_text(text)
    print("case fortran driver snapshots every 5 steps", flush=True)
    r = subprocess.run([DRIVER, f"--input={nml}"], capture_output=True, text=True, timeout=120)
    print(r.stdout[-1500:], r.stderr[-1500:])
    assert r.returncode == 0
    p, dx = cases.riemann2d(101, seed=0)
    for step in (0, 5, 10, 15, 20):
        f = fvmgpu.read_snapshot_file(f"{prefix}_{step:06d}.fvmsnap")
        assert f["step"] == step and f["codes"] == [201, 202, 203, 301]
        (n, h, arr), = f["blocks"]
        assert n == (101, 101, 1)
        if step:
            o = O.oracle_solve(p, 2, (dx, dx, 1.0), dt=2e-4, nsteps=step, p_from_q=1)
            q = o["q"][:, :1, 3:-3, 3:-3]
        else:
            q = cases.to_conserved(p)
        pv = np.moveaxis(cases.to_primitive(q), -1, 0)
        assert np.array_equal(arr, pv[[0, 1, 2, 4]])


def test_fortran_serving_equals_python_serving(tmp_path):
    """The Fortran serving example (bind(C) fvm_request, initial states
    computed on the device) gives, bit for bit, the states the Python binding
    gets for the same requests submitted with uploaded initial states."""
    exe = os.path.join(ROOT, "fvm-3d-f90_amd", "bin", "fvm_serve")
    assert os.path.exists(exe), "build the Fortran examples: make -C fvm-3d-f90_amd"
    prefix = tmp_path / "served"
    print("case fortran serving example, 3 Riemann requests", flush=True)
    r = subprocess.run([exe, str(prefix)], capture_output=True, text=True, timeout=120)
    print(r.stdout[-1000:], r.stderr[-1000:])
    assert r.returncode == 0
    with fvmgpu.Server(max_active=16) as srv:
        ids = {}
        for n in (41, 64, 101):
            p, dx = cases.riemann2d(n, seed=0)
            ids[n] = srv.submit(cases.to_conserved(p), (dx, dx, 1.0), 0.01, 0.5, recon=O.MUSCL,
                                limiter=O.LIM_MINMOD, flux=O.ROE, rk=O.RK_SSP3)
        while srv.poll(16, 1)[2] < 3:
            pass
        for n, i in ids.items():
            q, info = srv.result(i)
            assert info["status"] == 2
            f = np.fromfile(f"{prefix}_{n}.bin", dtype=np.float64).reshape(5, 1, n, n)
            assert np.array_equal(f, q)


def test_fortran_checkpoint_restart(tmp_path):
    """The driver's checkpoint_file / restart_file: 10 steps, a checkpoint,
    then a second run restarting from it to step 20, equal bit for bit to one
    run of 20 steps (time and step count carried over)."""
    base = open(CASE).read()

    def run(tag, **gpu):
        text = base.replace('"fortran_riemann2d_101.bin"', f'"{tmp_path / tag}.bin"')
        text = text.replace("max_step    = 20", f"max_step    = {gpu.pop('max_step')}")
        extra = "".join(f'\n  {k} = {v}' for k, v in gpu.items())
        text = text.replace("print_every = 5", "print_every = 5" + extra)
        nml = tmp_path / f"{tag}.nml"
        nml.write_text(text)
        r = subprocess.run([DRIVER, f"--input={nml}"], capture_output=True, text=True, timeout=120)
        print(r.stdout[-600:], r.stderr[-600:])
        assert r.returncode == 0
        raw = open(tmp_path / f"{tag}.bin", "rb").read()
        steps = int(np.frombuffer(raw[:20], dtype=np.int32)[4])
        t = float(np.frombuffer(raw[20:28], dtype=np.float64)[0])
        return steps, t, np.frombuffer(raw[28:], dtype=np.float64)

    print("case fortran checkpoint and restart", flush=True)
    base = base.replace("output_time  = 0.01", "output_time  = 1.0E-3")   # every 5 steps
    ck = tmp_path / "state.fvmsnap"
    s1, t1, _ = run("first", max_step=10, checkpoint_file=f'"{ck}"')
    assert s1 == 10
    s2, t2, q2 = run("second", max_step=20, restart_file=f'"{ck}"',
                     snapshot_prefix=f'"{tmp_path / "second"}"')
    s3, t3, q3 = run("whole", max_step=20, snapshot_prefix=f'"{tmp_path / "whole"}"')
    assert s2 == s3 == 20 and t2 == t3
    assert np.array_equal(q2, q3)
    # output snapshots after the restart continue at the uninterrupted run's
    # output times (steps 15 and 20), and equal its snapshots
    for step in (15, 20):
        a = fvmgpu.read_snapshot_file(str(tmp_path / f"second_{step:06d}.fvmsnap"))
        b = fvmgpu.read_snapshot_file(str(tmp_path / f"whole_{step:06d}.fvmsnap"))
        assert a["time"] == b["time"] and a["step"] == b["step"] == step
        assert np.array_equal(a["blocks"][0][2], b["blocks"][0][2])
    for step in (0, 5, 10):   # already written before the checkpoint
        assert not (tmp_path / f"second_{step:06d}.fvmsnap").exists()


@pytest.mark.parametrize("dim,ncut", [("2D", "3, 2, 1"), ("2D", "2, 4, 1"), ("3D", "2, 2, 2")])
def test_fortran_driver_cut_equals_single_block(tmp_path, dim, ncut):
    """The driver's ncut option: the grid run as the reference's uniform cut
    into blocks joined across their shared vertex planes, each initialised at
    its offset, gathered back into one array: equal bit for bit to the
    single-block run (vertices and the ghost layers of the grid's faces)."""
    base = open(CASE).read()
    if dim == "3D":
        base = base.replace('dimension     = "2D"', 'dimension     = "3D"')
        base = base.replace("gridSize  = 101, 101", "gridSize  = 33, 29, 21")

    def run(tag, extra):
        text = base.replace('"fortran_riemann2d_101.bin"', f'"{tmp_path / tag}.bin"')
        text = text.replace("print_every = 5", "print_every = 5" + extra)
        nml = tmp_path / f"{tag}.nml"
        nml.write_text(text)
        r = subprocess.run([DRIVER, f"--input={nml}"], capture_output=True, text=True, timeout=120)
        print(r.stdout[-600:], r.stderr[-600:])
        assert r.returncode == 0
        raw = open(tmp_path / f"{tag}.bin", "rb").read()
        ni, nj, nk, ng, steps = (int(v) for v in np.frombuffer(raw[:20], dtype=np.int32))
        kk = nk + 2 * ng if nk > 1 else 1
        return steps, np.frombuffer(raw[28:], dtype=np.float64).reshape(5, kk, nj + 2 * ng, ni + 2 * ng)

    print(f"case fortran driver {dim} ncut={ncut}", flush=True)
    s1, q1 = run("whole", "")
    s2, q2 = run("cut", f"\n  ncut = {ncut}")
    assert s1 == s2 == 20
    kin = slice(None) if dim == "2D" else slice(3, -3)
    assert np.array_equal(q1[:, kin, 3:-3, :], q2[:, kin, 3:-3, :])
    assert np.array_equal(q1[:, kin, :, 3:-3], q2[:, kin, :, 3:-3])
    if dim == "3D":
        assert np.array_equal(q1[:, :, 3:-3, 3:-3], q2[:, :, 3:-3, 3:-3])


def test_fortran_driver_3d_equals_oracle(tmp_path):
    """The driver's 3-D path (dimension = "3D", gridSize of three entries, dz
    = L / (nk - 1)): the four-quadrant regions (unbounded in z) initialised on
    the device, 20 steps, bit for bit the oracle on the same grid."""
    ni, nj, nk = 33, 29, 21
    out = tmp_path / "q3.bin"
    text = open(CASE).read().replace('"fortran_riemann2d_101.bin"', f'"{out}"')
    text = text.replace('dimension     = "2D"', 'dimension     = "3D"')
    text = text.replace("gridSize  = 101, 101", f"gridSize  = {ni}, {nj}, {nk}")
    nml = tmp_path / "case3.nml"
    nml.write_text(text)
    print(f"case fortran driver 3-D {ni}x{nj}x{nk} vs oracle", flush=True)
    r = subprocess.run([DRIVER, f"--input={nml}"], capture_output=True, text=True, timeout=120)
    print(r.stdout[-600:], r.stderr[-600:])
    assert r.returncode == 0
    raw = open(out, "rb").read()
    assert tuple(int(v) for v in np.frombuffer(raw[:20], dtype=np.int32)) == (ni, nj, nk, 3, 20)
    q = np.frombuffer(raw[28:], dtype=np.float64).reshape(5, nk + 6, nj + 6, ni + 6)
    (dx, x), (dy, y), (dz, _) = (cases.nodal_coords(n, 1.0) for n in (ni, nj, nk))
    p = np.zeros((nk, nj, ni, 5))
    X, Y = np.meshgrid(x, y, indexing="xy")
    for (xa, xb), (ya, yb), state in cases.REF_QUADRANTS:
        p[:, (X >= xa) & (X <= xb) & (Y >= ya) & (Y <= yb)] = state
    o = O.oracle_solve(p, 3, (dx, dy, dz), dt=2e-4, nsteps=20, p_from_q=1)
    assert np.array_equal(q[:, 3:-3, 3:-3, 3:-3], o["q"][:, 3:-3, 3:-3, 3:-3])


@pytest.mark.parametrize("names,codes", [
    (("WENO", "none_", "Roe", "Runge-Kutta", "fixed"), (O.WENO, O.LIM_NONE, O.ROE, O.RK_SSP3)),
    (("MUSCL", "VanLeer_", "VanLeer", "Runge-Kutta", "CFL"), (O.MUSCL, O.LIM_VANLEER, O.VANLEER, O.RK_SSP3)),
    (("MUSCL", "vanAlbada_", "AUSM+", "RK4", "CFL"), (O.MUSCL, O.LIM_VANALBADA, O.AUSMPLUS, O.RK4)),
    (("MUSCL", "minmod_", "AUSM", "Runge-Kutta", "fixed"), (O.MUSCL, O.LIM_MINMOD, O.AUSM, O.RK_SSP3)),
    (("simple", "none_", "Roe", "Runge-Kutta", "CFL"), (O.SIMPLE, O.LIM_NONE, O.ROE, O.RK_SSP3)),
], ids=["weno_roe_fixed", "muscl_vanleer_cfl", "muscl_ausmplus_rk4_cfl", "muscl_ausm_fixed",
        "simple_roe_cfl"])
def test_fortran_driver_scheme_names(tmp_path, names, codes):
    """The case file's scheme names (&Inviscid_Flux reconstruct / limiter /
    flux_split, &Time_Stepping time_scheme / timestep_set, spelled as in
    src/defs/type_defs_scheme.f90) select the oracle's scheme: 20 steps of the
    101^2 Riemann case, fixed dt or CFL 0.5, bit for bit."""
    recon, lim, flux, rk, tset = names
    out = tmp_path / "q.bin"
    text = open(CASE).read().replace('"fortran_riemann2d_101.bin"', f'"{out}"')
    for old, new in (('reconstruct   = "MUSCL"', f'reconstruct   = "{recon}"'),
                     ('limiter       = "minmod_"', f'limiter       = "{lim}"'),
                     ('flux_split    = "Roe"', f'flux_split    = "{flux}"'),
                     ('time_scheme  = "Runge-Kutta"', f'time_scheme  = "{rk}"'),
                     ('timestep_set = "fixed"', f'timestep_set = "{tset}"')):
        assert old in text
        text = text.replace(old, new)
    nml = tmp_path / "case.nml"
    nml.write_text(text)
    print(f"case fortran driver scheme names {names}", flush=True)
    r = subprocess.run([DRIVER, f"--input={nml}"], capture_output=True, text=True, timeout=120)
    print(r.stdout[-600:], r.stderr[-600:])
    assert r.returncode == 0
    raw = open(out, "rb").read()
    assert int(np.frombuffer(raw[:20], dtype=np.int32)[4]) == 20
    t_end = float(np.frombuffer(raw[20:28], dtype=np.float64)[0])
    q = np.frombuffer(raw[28:], dtype=np.float64).reshape(5, 1, 107, 107)
    p, dx = cases.riemann2d(101, seed=0)
    c_recon, c_lim, c_flux, c_rk = codes
    kw = dict(cfl=0.5, dt=0.0) if tset == "CFL" else dict(dt=2e-4)
    o = O.oracle_solve(p, 2, (dx, dx, 1.0), recon=c_recon, limiter=c_lim, flux=c_flux, rk=c_rk,
                       nsteps=20, p_from_q=1, **kw)
    assert np.array_equal(q[:, 0, 3:-3, 3:-3], o["q"][:, 0, 3:-3, 3:-3])
    if tset == "CFL":
        assert t_end == pytest.approx(float(np.sum(o["dt"])), rel=1e-13)


def test_integration_snippet_reference_layout(tmp_path):
    """The integration notes' snippet, compiled verbatim (bin/integration_example),
    on the reference's own 2-D case: ClassBlockData%q allocated as the
    reference allocates it, q(n+6, n+6, 2+6, 5), the case file's scheme values
    through ClassSpaceDiscret%input, ten fixed steps.  Every real plane equals
    the oracle bit for bit; the ghost planes hold the real vertices."""
    from _parity import NG, reference_q_2d
    exe = os.path.join(ROOT, "fvm-3d-f90_amd", "bin", "integration_example")
    assert os.path.exists(exe), "build it: make -C fvm-3d-f90_amd"
    n, nsteps, dt = 41, 10, 2e-4
    p, dx = cases.riemann2d(n, seed=0)
    H = reference_q_2d(p)
    fin, fout = tmp_path / "q0.bin", tmp_path / "q1.bin"
    H.tofile(fin)
    print(f"case integration snippet riemann2d {n}x{n}x2, {nsteps} steps", flush=True)
    r = subprocess.run([exe, str(n), str(nsteps), repr(dt), str(fin), str(fout)],
                       capture_output=True, text=True, timeout=120)
    print(r.stdout[-1500:], r.stderr[-1500:])
    assert r.returncode == 0
    q = np.fromfile(fout, dtype=np.float64).reshape(H.shape)
    o = O.oracle_solve(p, 2, (dx, dx, 1.0), dt=dt, nsteps=nsteps, p_from_q=1)
    want = o["q"][:, 0, NG:-NG, NG:-NG]
    for kk in range(2 + 2 * NG):
        assert np.array_equal(q[:, kk, NG:-NG, NG:-NG], want), f"plane {kk}"
    norms = [float(x) for x in r.stdout.split("norms")[1].split()]
    np.testing.assert_allclose(norms, o["res"][-1], rtol=1e-12, atol=1e-300)
