"""Python host side of the MI355X residual + RK path.

A thin ctypes mirror of include/fvm_gpu.h.  The compute runs in
lib/libfvmgpu.so (HIP, gfx950); there is no CPU fallback: importing this
module on a machine without the built library raises, and so does creating a
context without a GPU.

Scheme codes are the reference's (src/defs/type_defs_scheme.f90:11-36).
"""
import ctypes
import os

import numpy as np

from . import cases  # noqa: F401  (re-export)

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("FVM_LIB") or os.path.join(os.path.dirname(_HERE), "lib",
                                                     "libfvmgpu.so")

RK_SSP3, RK4 = 101, 1001
SIMPLE, MUSCL, WENO = 200, 201, 202
LIM_NONE, LIM_MINMOD, LIM_VANLEER, LIM_VANALBADA = 250, 251, 252, 253
ROE, VANLEER, AUSM, AUSMPLUS = 300, 301, 303, 1303
BC_EXTRAP, BC_PERIODIC = 0, 1
IC_REGIONS, IC_VORTEX, IC_TGV = 1, 2, 3
IC_MAX_PARAMS = 176
# the reference's BC codes (src/defs/type_defs_bc.f90:11-30)
BC_INFLOW, BC_INFLOW_SUBSONIC, BC_INFLOW_SUPERSONIC = 10, 11, 12
BC_OUTFLOW, BC_OUTFLOW_SUBSONIC, BC_OUTFLOW_SUPERSONIC = 20, 21, 22
BC_WALL, BC_WALL_INVISCID, BC_FARFIELD, BC_EXTRAPOLATE, BC_SYMMETRY = 30, 31, 40, 41, 50
BC_WALL_VISCOUS_ISOTHERMAL, BC_WALL_VISCOUS = 32, 33
NUMERICS_EXACT, NUMERICS_FAST = 0, 1
LAUNCH_PER_BLOCK, LAUNCH_BATCHED, LAUNCH_PERSISTENT = 0, 1, 2
# snapshot variables: the reference's output variables
# (src/defs/type_defs_variables.f90:25-30) and the conserved components
VAR_DENSITY, VAR_VELOCITY_X, VAR_VELOCITY_Y, VAR_VELOCITY_Z, VAR_PRESSURE = 201, 202, 203, 204, 301
VAR_RHO, VAR_RHOU, VAR_RHOV, VAR_RHOW, VAR_RHOE = 1, 2, 3, 4, 5
PRIMITIVE_VARS = (VAR_DENSITY, VAR_VELOCITY_X, VAR_VELOCITY_Y, VAR_VELOCITY_Z, VAR_PRESSURE)
CONSERVED_VARS = (VAR_RHO, VAR_RHOU, VAR_RHOV, VAR_RHOW, VAR_RHOE)

# scheme names as the case file spells them (type_defs_scheme.f90:41-84)
SCHEME_INDEX = {
    "Runge-Kutta": RK_SSP3, "RK4": RK4,
    "simple": SIMPLE, "MUSCL": MUSCL, "WENO": WENO,
    "none_": LIM_NONE, "minmod_": LIM_MINMOD, "VanLeer_": LIM_VANLEER,
    "vanAlbada_": LIM_VANALBADA,
    "Roe": ROE, "VanLeer": VANLEER, "AUSM": AUSM, "AUSM+": AUSMPLUS,
}

EXPORTS = ("fvm_last_error", "fvm_version", "fvm_ctx_create", "fvm_ctx_destroy",
           "fvm_scheme_set", "fvm_block_create", "fvm_block_destroy", "fvm_block_set_bc",
           "fvm_upload", "fvm_download", "fvm_step", "fvm_norms", "fvm_sync",
           "fvm_timing_enable", "fvm_timing_read", "fvm_set_numerics", "fvm_autotune",
           "fvm_block_launch_shape", "fvm_block_connect", "fvm_set_launch_mode",
           "fvm_server_create", "fvm_server_destroy", "fvm_server_submit", "fvm_server_poll",
           "fvm_server_result", "fvm_set_time", "fvm_time", "fvm_viscous_set",
           "fvm_set_graph", "fvm_conn_add", "fvm_snapshot_begin", "fvm_snapshot_ready",
           "fvm_snapshot_read", "fvm_snapshot_save", "fvm_snapshot_release", "fvm_restart_load",
           "fvm_block_set_bc_state", "fvm_block_init", "fvm_cut_uniform",
           "fvm_block_launch_tile", "fvm_block_init_at", "fvm_upload_planes",
           "fvm_download_planes", "fvm_stage", "fvm_step_noalloc")


class FvmError(RuntimeError):
    pass


class Request(ctypes.Structure):
    """fvm_request of include/fvm_gpu.h."""
    _fields_ = [("n", ctypes.c_int64 * 3), ("spacing", ctypes.c_double * 3),
                ("recon", ctypes.c_int), ("limiter", ctypes.c_int), ("flux", ctypes.c_int),
                ("rk", ctypes.c_int), ("kappa", ctypes.c_double),
                ("entropy_fix", ctypes.c_double), ("gamma", ctypes.c_double),
                ("ref_compat", ctypes.c_int), ("numerics", ctypes.c_int),
                ("bc", ctypes.c_int * 6), ("t_end", ctypes.c_double), ("cfl", ctypes.c_double),
                ("q0", ctypes.c_void_p), ("borrow", ctypes.c_int),
                ("ic_kind", ctypes.c_int), ("ic_nparams", ctypes.c_int),
                ("ic_params", ctypes.c_double * IC_MAX_PARAMS), ("ic_origin", ctypes.c_double * 3)]


_lib = None


def load_library(path=LIB_PATH):
    """Load libfvmgpu.so and declare its signatures (no GPU call is made)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise FvmError(f"{path} is not built: run `make -C fvm-3d-f90_amd` "
                       "(or __graft_entry__.build())")
    lib = ctypes.CDLL(path)
    c_void_pp = ctypes.POINTER(ctypes.c_void_p)
    i64, ci, cd, vp = ctypes.c_int64, ctypes.c_int, ctypes.c_double, ctypes.c_void_p
    sig = {
        "fvm_last_error": ([], ctypes.c_char_p),
        "fvm_version": ([], ctypes.c_char_p),
        "fvm_ctx_create": ([ci, c_void_pp], ci),
        "fvm_ctx_destroy": ([vp], ci),
        "fvm_scheme_set": ([vp, ci, ci, ci, ci, cd, cd, cd, ci], ci),
        "fvm_block_create": ([vp, i64, i64, i64, cd, cd, cd, c_void_pp], ci),
        "fvm_block_destroy": ([vp], ci),
        "fvm_block_set_bc": ([vp, ci, ci], ci),
        "fvm_upload": ([vp, vp, ci], ci),
        "fvm_download": ([vp, vp, ci], ci),
        "fvm_upload_planes": ([vp, vp, ci, i64, i64], ci),
        "fvm_download_planes": ([vp, vp, ci, i64], ci),
        "fvm_stage": ([vp, ci, cd], ci),
        "fvm_step_noalloc": ([vp, ci, cd, cd, ctypes.POINTER(cd)], ci),
        "fvm_step": ([vp, ci, cd, cd, ctypes.POINTER(cd)], ci),
        "fvm_norms": ([vp, ctypes.POINTER(cd), ctypes.POINTER(ci)], ci),
        "fvm_sync": ([vp], ci),
        "fvm_timing_enable": ([vp, ci], ci),
        "fvm_set_numerics": ([vp, ci], ci),
        "fvm_autotune": ([vp, ci], ci),
        "fvm_block_connect": ([vp, ci, vp, ci], ci),
        "fvm_set_launch_mode": ([vp, ci], ci),
        "fvm_set_time": ([vp, cd, cd], ci),
        "fvm_viscous_set": ([vp, cd, cd, cd, cd, cd], ci),
        "fvm_set_graph": ([vp, ci], ci),
        "fvm_conn_add": ([vp, ctypes.POINTER(i64), ctypes.POINTER(i64), vp, ctypes.POINTER(i64),
                          ctypes.POINTER(i64), ctypes.POINTER(ci)], ci),
        "fvm_time": ([vp, ctypes.POINTER(cd), ctypes.POINTER(i64)], ci),
        "fvm_snapshot_begin": ([vp, ci, ctypes.POINTER(ci), ctypes.POINTER(i64)], ci),
        "fvm_snapshot_ready": ([vp, i64, ctypes.POINTER(ci)], ci),
        "fvm_snapshot_read": ([vp, i64, vp, vp, ctypes.POINTER(cd), ctypes.POINTER(i64)], ci),
        "fvm_snapshot_save": ([vp, i64, ctypes.c_char_p], ci),
        "fvm_snapshot_release": ([vp, i64], ci),
        "fvm_restart_load": ([vp, ctypes.c_char_p], ci),
        "fvm_block_set_bc_state": ([vp, ci, ci, ctypes.POINTER(cd)], ci),
        "fvm_block_init": ([vp, ci, ci, ctypes.POINTER(cd), ctypes.POINTER(cd)], ci),
        "fvm_block_init_at": ([vp, ci, ci, ctypes.POINTER(cd), ctypes.POINTER(i64)], ci),
        "fvm_cut_uniform": ([i64, i64, ci, ctypes.POINTER(i64)], ci),
        "fvm_block_launch_tile": ([vp, ctypes.POINTER(ci)], ci),
        "fvm_server_create": ([ci, cd, ci, c_void_pp], ci),
        "fvm_server_destroy": ([vp], ci),
        "fvm_server_submit": ([vp, ctypes.POINTER(Request), ctypes.POINTER(i64)], ci),
        "fvm_server_poll": ([vp, ci, ci, ctypes.POINTER(ci), ctypes.POINTER(ci), ctypes.POINTER(ci)],
                            ci),
        "fvm_server_result": ([vp, i64, vp, ctypes.POINTER(cd)], ci),
        "fvm_block_launch_shape": ([vp, ctypes.POINTER(ci), ctypes.POINTER(ci),
                                    ctypes.POINTER(ci)], ci),
        "fvm_timing_read": ([vp, ctypes.POINTER(cd), ctypes.POINTER(i64), ctypes.POINTER(cd)], ci),
    }
    for name, (args, res) in sig.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    _lib = lib
    return lib


def _check(rc):
    if rc != 0:
        raise FvmError(f"fvm error {rc}: {_lib.fvm_last_error().decode()}")


class Block:
    """One structured block: ni x nj x nk real vertices (nk == 1: 2-D)."""

    def __init__(self, ctx, shape, spacing):
        self.ctx = ctx
        self.ni, self.nj, self.nk = (int(s) for s in shape)
        self.dim = 2 if self.nk == 1 else 3
        h = ctypes.c_void_p()
        _check(_lib.fvm_block_create(ctx._h, self.ni, self.nj, self.nk, float(spacing[0]),
                                     float(spacing[1]), float(spacing[2]), ctypes.byref(h)))
        self._h = h

    def set_bc(self, face, kind):
        _check(_lib.fvm_block_set_bc(self._h, face, kind))

    def set_bc_state(self, face, kind, prim):
        """Inflow / farfield face: ghosts hold the primitive state (rho, u, v, w, p)."""
        _check(_lib.fvm_block_set_bc_state(self._h, face, kind,
                                           (ctypes.c_double * 5)(*(float(x) for x in prim))))

    def connect(self, face, other, shared=True):
        """Face `face` of this block abuts the opposite face of `other`."""
        _check(_lib.fvm_block_connect(self._h, face, other._h, int(bool(shared))))

    def conn_add(self, range_min, range_max, donor, donor_range_min, donor_range_max,
                 donor_type=(1, 2, 3)):
        """One 1-to-1 connection patch (1-based inclusive vertex ranges, CGNS
        Conn-1to1 point range, donor range and transform): the ghosts of this
        block's face range are filled from `donor`."""
        def i3(v):
            return (ctypes.c_int64 * 3)(*(int(x) for x in v))
        _check(_lib.fvm_conn_add(self._h, i3(range_min), i3(range_max), donor._h,
                                 i3(donor_range_min), i3(donor_range_max),
                                 (ctypes.c_int * 3)(*(int(x) for x in donor_type))))

    def set_periodic(self, dims=(0, 1, 2)):
        for d in dims:
            if d < self.dim:
                self.set_bc(2 * d, BC_PERIODIC)

    def host_shape(self, ng):
        nz = self.nk + 2 * ng if self.dim == 3 else 1
        return (5, nz, self.nj + 2 * ng, self.ni + 2 * ng)

    def upload(self, q, ng=0):
        """q: (5, nk(+2ng), nj+2ng, ni+2ng) float64 C-order, i.e. the memory
        image of the Fortran array q(i, j, k, eq)."""
        q = np.ascontiguousarray(q, dtype=np.float64)
        if q.shape != self.host_shape(ng):
            raise FvmError(f"upload shape {q.shape} != {self.host_shape(ng)}")
        _check(_lib.fvm_upload(self._h, q.ctypes.data, ng))

    def init(self, kind, params, origin=None):
        """Initial state computed on the device (IC_REGIONS / IC_VORTEX / IC_TGV)."""
        pv = (ctypes.c_double * max(1, len(params)))(*(float(x) for x in params))
        ov = None if origin is None else (ctypes.c_double * 3)(*(float(x) for x in origin))
        _check(_lib.fvm_block_init(self._h, kind, len(params), pv, ov))

    def init_at(self, kind, params, first=(0, 0, 0)):
        """Initial state of a block whose first vertex is vertex `first` of a
        larger grid (coordinates dx * (first + i), one product)."""
        pv = (ctypes.c_double * max(1, len(params)))(*(float(x) for x in params))
        fv = (ctypes.c_int64 * 3)(*(int(x) for x in first))
        _check(_lib.fvm_block_init_at(self._h, kind, len(params), pv, fv))

    def launch_tile(self):
        """Launch-shape id (the FVM_TILE / FVM_TILE2 value) of this block."""
        v = ctypes.c_int()
        _check(_lib.fvm_block_launch_tile(self._h, ctypes.byref(v)))
        return v.value

    def launch_shape(self):
        tx, ty, ch = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        _check(_lib.fvm_block_launch_shape(self._h, ctypes.byref(tx), ctypes.byref(ty),
                                           ctypes.byref(ch)))
        return tx.value, ty.value, ch.value

    def download(self, ng=0):
        q = np.zeros(self.host_shape(ng))
        _check(_lib.fvm_download(self._h, q.ctypes.data, ng))
        return q

    def upload_planes(self, q, ng, nk_host, k_plane=1):
        """q: (5, nk_host + 2ng, nj + 2ng, ni + 2ng), the reference's own
        allocation of ClassBlockData%q; a 2-D block reads real plane k_plane."""
        q = np.ascontiguousarray(q, dtype=np.float64)
        want = (5, nk_host + 2 * ng, self.nj + 2 * ng, self.ni + 2 * ng)
        if q.shape != want:
            raise FvmError(f"upload shape {q.shape} != {want}")
        _check(_lib.fvm_upload_planes(self._h, q.ctypes.data, ng, int(nk_host), int(k_plane)))

    def download_planes(self, ng, nk_host):
        q = np.zeros((5, nk_host + 2 * ng, self.nj + 2 * ng, self.ni + 2 * ng))
        _check(_lib.fvm_download_planes(self._h, q.ctypes.data, ng, int(nk_host)))
        return q


class Context:
    """A device context: one stream and the blocks created on it."""

    def __init__(self, device=0):
        load_library()
        h = ctypes.c_void_p()
        _check(_lib.fvm_ctx_create(device, ctypes.byref(h)))
        self._h = h
        self.blocks = []
        self._snap_nvars = {}       # snapshot id -> number of variables

    def close(self):
        if self._h:
            _lib.fvm_ctx_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def scheme(self, recon=MUSCL, limiter=LIM_MINMOD, flux=ROE, rk=RK_SSP3, kappa=0.33333,
               entropy_fix=0.2, gamma=1.4, ref_compat=1):
        _check(_lib.fvm_scheme_set(self._h, recon, limiter, flux, rk, kappa, entropy_fix, gamma,
                                   ref_compat))

    def numerics(self, mode):
        """NUMERICS_EXACT (bit-identical to the reference arithmetic) or
        NUMERICS_FAST (FMA contraction + reciprocals, tolerance-checked)."""
        _check(_lib.fvm_set_numerics(self._h, mode))

    def block(self, shape, spacing):
        b = Block(self, shape, spacing)
        self.blocks.append(b)
        return b

    def cut_blocks(self, q, spacing, ncut, periodic=(False, False, False)):
        """The conserved state q (5, nk, nj, ni) as ncut = (ci, cj, ck) blocks
        of a uniform cut (fvm_cut_uniform), uploaded and connected across their
        shared planes; periodic directions wrap across the outer faces.
        Returns {(a, b, c): (block, (i0, j0, k0))}."""
        nk, nj, ni = q.shape[1:]
        ncut = tuple(ncut) + (1,) * (3 - len(ncut))
        # (a 2-D grid's single k plane is not cut)
        ranges = [cut_uniform(n, c) if n > 1 else [(0, 0)] for n, c in zip((ni, nj, nk), ncut)]
        blocks = {}
        for a, (x0, x1) in enumerate(ranges[0]):
            for b, (y0, y1) in enumerate(ranges[1]):
                for c, (z0, z1) in enumerate(ranges[2]):
                    blk = self.block((x1 - x0 + 1, y1 - y0 + 1, z1 - z0 + 1), spacing)
                    blk.upload(q[:, z0:z1 + 1, y0:y1 + 1, x0:x1 + 1])
                    blocks[(a, b, c)] = (blk, (x0, y0, z0))
        for idx, (blk, _) in blocks.items():
            for d in range(3):
                nb = list(idx)
                if idx[d] + 1 < ncut[d]:
                    nb[d] += 1
                    blk.connect(2 * d + 1, blocks[tuple(nb)][0], shared=True)
                elif periodic[d] and (d < 2 or nk > 1):
                    nb[d] = 0
                    if ncut[d] == 1:
                        blk.set_bc(2 * d, BC_PERIODIC)
                    else:
                        blk.connect(2 * d + 1, blocks[tuple(nb)][0], shared=False)
        return blocks

    def launch_mode(self, mode):
        """LAUNCH_BATCHED (default), LAUNCH_PERSISTENT or LAUNCH_PER_BLOCK."""
        _check(_lib.fvm_set_launch_mode(self._h, mode))

    def autotune(self, reps=3):
        """Pick each block's tile shape and k-chunk by timing stage launches."""
        _check(_lib.fvm_autotune(self._h, int(reps)))

    def step(self, nsteps=1, dt=0.0, cfl=0.0):
        """Advance every block; dt > 0 fixed, else CFL.  Returns the last dt."""
        out = ctypes.c_double(0.0)
        _check(_lib.fvm_step(self._h, int(nsteps), float(dt), float(cfl), ctypes.byref(out)))
        return out.value

    def step_noalloc(self, nsteps=1, dt=0.0, cfl=0.0):
        """fvm_step_noalloc: step() with no allocation (static scratch)."""
        out = ctypes.c_double(0.0)
        _check(_lib.fvm_step_noalloc(self._h, int(nsteps), float(dt), float(cfl),
                                     ctypes.byref(out)))
        return out.value

    def stage(self, s, dt):
        """Runge-Kutta stage s (1-based) of the step in progress, fixed dt."""
        _check(_lib.fvm_stage(self._h, int(s), float(dt)))

    def viscous(self, mu_ref, t_ref=1.0, sutherland_s=0.0, prandtl=0.72, r_gas=1.0):
        """Central viscous fluxes with Sutherland's law; mu_ref <= 0: off."""
        _check(_lib.fvm_viscous_set(self._h, float(mu_ref), float(t_ref), float(sutherland_s),
                                    float(prandtl), float(r_gas)))

    def graph(self, enable=True):
        """Replay whole time steps from a captured hipGraph."""
        _check(_lib.fvm_set_graph(self._h, int(bool(enable))))

    def set_time(self, t=0.0, t_end=0.0):
        """Current time and end time (t_end <= t: none).  CFL steps clip to
        t_end and skip steps past it on the device."""
        _check(_lib.fvm_set_time(self._h, float(t), float(t_end)))

    def time(self):
        """(time, steps run) after the last step."""
        t, n = ctypes.c_double(), ctypes.c_int64()
        _check(_lib.fvm_time(self._h, ctypes.byref(t), ctypes.byref(n)))
        return t.value, n.value

    def norms(self):
        out = (ctypes.c_double * 6)()
        flag = ctypes.c_int(0)
        _check(_lib.fvm_norms(self._h, out, ctypes.byref(flag)))
        return np.array(out[:]), bool(flag.value)

    def sync(self):
        _check(_lib.fvm_sync(self._h))

    def snapshot(self, variables=PRIMITIVE_VARS):
        """Start an asynchronous snapshot of u^n of every block; returns its id."""
        v = (ctypes.c_int * len(variables))(*(int(x) for x in variables))
        sid = ctypes.c_int64()
        _check(_lib.fvm_snapshot_begin(self._h, len(variables), v, ctypes.byref(sid)))
        self._snap_nvars[sid.value] = len(variables)
        return sid.value

    def snapshot_ready(self, sid):
        r = ctypes.c_int()
        _check(_lib.fvm_snapshot_ready(self._h, sid, ctypes.byref(r)))
        return bool(r.value)

    def snapshot_read(self, sid, blk):
        """(array (nvars, nk, nj, ni), time, steps) of block `blk` in the snapshot."""
        out = np.empty((self._snap_nvars[sid], blk.nk, blk.nj, blk.ni))
        t, n = ctypes.c_double(), ctypes.c_int64()
        _check(_lib.fvm_snapshot_read(self._h, sid, blk._h, out.ctypes.data, ctypes.byref(t),
                                      ctypes.byref(n)))
        return out, t.value, n.value

    def snapshot_save(self, sid, path):
        """Write the snapshot to `path` on a background thread."""
        _check(_lib.fvm_snapshot_save(self._h, sid, os.fsencode(path)))

    def snapshot_release(self, sid):
        """Free the slot (waits for a pending save and reports its status)."""
        self._snap_nvars.pop(sid, None)
        _check(_lib.fvm_snapshot_release(self._h, sid))

    def restart(self, path):
        """Load a conserved-state snapshot file into this context's blocks."""
        _check(_lib.fvm_restart_load(self._h, os.fsencode(path)))

    def timing(self, every=1):
        """Time the stage launches of every `every`-th step (0 / False: off)."""
        _check(_lib.fvm_timing_enable(self._h, int(every)))

    def timing_read(self):
        ms, n, cu = ctypes.c_double(), ctypes.c_int64(), ctypes.c_double()
        _check(_lib.fvm_timing_read(self._h, ctypes.byref(ms), ctypes.byref(n), ctypes.byref(cu)))
        return ms.value, n.value, cu.value


def cut_uniform(n, ncut):
    """Vertex ranges (0-based, inclusive) of a uniform cut of n vertices into
    ncut pieces sharing their interface planes (fvm_cut_uniform)."""
    load_library()
    pos = (ctypes.c_int64 * (ncut + 1))()
    _check(_lib.fvm_cut_uniform(1, int(n), int(ncut), pos))
    return [(pos[c] - 1, pos[c + 1] - 1) for c in range(ncut)]


def region_params(regions, zrange=(-1e300, 1e300)):
    """[((xmin, xmax), (ymin, ymax), (rho, u, v, w, p)), ...] (2-D regions as
    cases.REF_QUADRANTS spells them) -> the flat IC_REGIONS parameter list."""
    out = []
    for (xa, xb), (ya, yb), st in regions:
        out += [xa, xb, ya, yb, zrange[0], zrange[1], *st]
    return out


def read_snapshot_file(path):
    """Snapshot file -> dict(codes, time, step, blocks=[(shape (ni, nj, nk),
    spacing, array (nvars, nk, nj, ni))])."""
    with open(path, "rb") as f:
        if f.read(8) != b"FVMSNAP1":
            raise FvmError(f"{path} is not a snapshot file")
        nb, nv = np.frombuffer(f.read(8), "<i4")
        codes = np.frombuffer(f.read(4 * int(nv)), "<i4").tolist()
        time = float(np.frombuffer(f.read(8), "<f8")[0])
        step = int(np.frombuffer(f.read(8), "<i8")[0])
        heads = []
        for _ in range(int(nb)):
            n = np.frombuffer(f.read(24), "<i8")
            h = np.frombuffer(f.read(24), "<f8")
            heads.append((tuple(int(x) for x in n), tuple(float(x) for x in h)))
        blocks = []
        for n, h in heads:
            cnt = int(nv) * n[0] * n[1] * n[2]
            a = np.frombuffer(f.read(8 * cnt), "<f8")
            if a.size != cnt:
                raise FvmError(f"{path}: truncated")
            blocks.append((n, h, a.reshape(int(nv), n[2], n[1], n[0]).copy()))
    return dict(codes=codes, time=time, step=step, blocks=blocks)


class Server:
    """Continuous-batching server of independent solves (fvm_server_*)."""

    def __init__(self, device=0, hbm_budget=0.0, max_active=64):
        load_library()
        h = ctypes.c_void_p()
        _check(_lib.fvm_server_create(device, float(hbm_budget), int(max_active), ctypes.byref(h)))
        self._h = h
        self.shapes = {}
        self._hold = {}     # borrowed initial states, kept until the request is done

    def close(self):
        if self._h:
            _lib.fvm_server_destroy(self._h)
            self._h = None
        self._hold.clear()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def submit(self, q0, spacing, t_end, cfl, recon=MUSCL, limiter=LIM_MINMOD, flux=ROE,
               rk=RK_SSP3, kappa=0.33333, entropy_fix=0.2, gamma=1.4, ref_compat=1,
               numerics=NUMERICS_EXACT, periodic=(False, False, False), borrow=True, ic=None):
        """q0: (5, nk, nj, ni) conserved state.  Returns the request id.
        borrow: the server's staging thread copies q0 after submit returns
        (the array is held here until the result is taken); otherwise submit
        copies it.  ic = (shape (ni, nj, nk), kind, params[, origin]): the
        state is computed on the device at admission instead (q0 is None)."""
        ic_kind, ic_np = 0, 0
        ic_params = (ctypes.c_double * IC_MAX_PARAMS)()
        ic_origin = (ctypes.c_double * 3)()
        if ic is not None:
            shape, ic_kind, params = ic[0], ic[1], list(ic[2])
            if len(params) > IC_MAX_PARAMS:
                raise FvmError("too many initial-state parameters")
            ic_np = len(params)
            for k, v in enumerate(params):
                ic_params[k] = float(v)
            for k, v in enumerate(ic[3] if len(ic) > 3 else (0.0, 0.0, 0.0)):
                ic_origin[k] = float(v)
            ni, nj, nk = (int(x) for x in shape)
            ptr, borrow = None, False
        else:
            q0 = np.ascontiguousarray(q0, dtype=np.float64)
            _, nk, nj, ni = q0.shape
            ptr = q0.ctypes.data
        bc = [0] * 6
        for d in range(3):
            if periodic[d]:
                bc[2 * d] = bc[2 * d + 1] = BC_PERIODIC
        r = Request((ni, nj, nk), tuple(float(x) for x in spacing), recon, limiter, flux, rk,
                    kappa, entropy_fix, gamma, ref_compat, numerics, tuple(bc), float(t_end),
                    float(cfl), ptr, 1 if borrow else 0, ic_kind, ic_np, ic_params, ic_origin)
        rid = ctypes.c_int64()
        _check(_lib.fvm_server_submit(self._h, ctypes.byref(r), ctypes.byref(rid)))
        self.shapes[rid.value] = (5, nk, nj, ni)
        if borrow:
            self._hold[rid.value] = q0
        return rid.value

    def poll(self, iterations=1, steps_per_sync=1):
        """Scheduler iterations; returns (pending, active, done) counts."""
        p, a, d = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        _check(_lib.fvm_server_poll(self._h, int(iterations), int(steps_per_sync), ctypes.byref(p),
                                    ctypes.byref(a), ctypes.byref(d)))
        return p.value, a.value, d.value

    def result(self, rid):
        """(final state, info dict) of a finished request."""
        info = (ctypes.c_double * 8)()
        if rid not in self.shapes:            # the library reports unknown ids
            _check(_lib.fvm_server_result(self._h, rid, None, info))
        q = np.zeros(self.shapes[rid])
        _check(_lib.fvm_server_result(self._h, rid, q.ctypes.data, info))
        self._hold.pop(rid, None)
        keys = ("t_submit", "t_admit", "t_done", "steps", "cell_updates", "t_final", "status",
                "bytes")
        return q, dict(zip(keys, info[:]))
