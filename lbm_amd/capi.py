"""ctypes binding of the liblbm_hip C ABI (include/lbm_hip.h).

This is the Python view of the same entry points the C++ host calls.  It
loads ``lbm_amd/liblbm_hip.so`` built in-tree and raises if the library is
missing: there is no CPU fallback.
"""
import ctypes as ct
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("LBM_HIP_LIB") or os.path.join(HERE, "liblbm_hip.so")

LBM_OK, LBM_EINVAL, LBM_EHIP, LBM_ENOMEM, LBM_DIVERGED = 0, -1, -2, -3, -4
D2Q9, D3Q19 = 9, 19
STORE = {"f32": 0, "f64": 1, "f16": 2, "bf16": 3}
BYTES = {"f32": 4, "f64": 8, "f16": 2, "bf16": 2}
FACE = {"periodic": 0, "wall": 1, "pressure_abb": 2, "velocity_zh": 3, "pressure_zh": 4, "pressure_neem": 5, "velocity_neem": 6,
        "periodic_pressure": 7, "wall_eq": 8, "wall_neem": 9, "wall_nebb": 10, "wall_nebb0": 11}
FACE_IDS = {"-x": 0, "+x": 1, "-y": 2, "+y": 3, "-z": 4, "+z": 5}


class GridDesc(ct.Structure):
    _fields_ = [("dim", ct.c_int32), ("nx", ct.c_int32), ("ny", ct.c_int32), ("nz", ct.c_int32),
                ("solid", ct.POINTER(ct.c_uint8))]


class Params(ct.Structure):
    _fields_ = [("lattice", ct.c_int32), ("storage", ct.c_int32), ("omega", ct.c_double),
                ("force", ct.c_double * 3)]


class BcDesc(ct.Structure):
    _fields_ = [("face", ct.c_int32), ("kind", ct.c_int32), ("value", ct.c_double * 3), ("rho", ct.c_double),
                ("priority", ct.c_int32)]


class Sums(ct.Structure):
    _fields_ = [("mass", ct.c_double), ("momentum", ct.c_double * 3), ("abs_diff", ct.c_double * 4),
                ("diff_sq", ct.c_double), ("norm_sq", ct.c_double), ("res_rel", ct.c_double),
                ("nonfinite", ct.c_double), ("fluid_cells", ct.c_double)]

    def as_dict(self):
        return {"mass": self.mass, "momentum": list(self.momentum), "abs_diff": list(self.abs_diff),
                "diff_sq": self.diff_sq, "norm_sq": self.norm_sq, "res_rel": self.res_rel,
                "nonfinite": self.nonfinite, "fluid_cells": self.fluid_cells}


# every symbol declared in include/lbm_hip.h
EXPORTS = ["lbm_create", "lbm_set_bc", "lbm_init_fields", "lbm_step", "lbm_step_check", "lbm_step_ex", "lbm_step_ex_noalloc",
           "lbm_get_macros",
           "lbm_reduce", "lbm_get_populations", "lbm_set_populations", "lbm_time_steps", "lbm_time_step_groups", "lbm_set_launch",
           "lbm_set_halo", "lbm_set_pressure_forcing", "lbm_set_index_width", "lbm_set_storage", "lbm_info", "lbm_synchronize", "lbm_destroy", "lbm_last_error", "lbm_copy_bench"]

_lib = None


def load():
    """Load the in-tree library (raises when it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} not found: build it first (python -c 'import __graft_entry__ as g; g.build()')")
    lib = ct.CDLL(LIB_PATH)
    P = ct.c_void_p
    dp = ct.POINTER(ct.c_double)
    lib.lbm_create.argtypes = [ct.POINTER(GridDesc), ct.POINTER(Params), ct.POINTER(P)]
    lib.lbm_set_bc.argtypes = [P, ct.POINTER(BcDesc), ct.c_int32]
    lib.lbm_init_fields.argtypes = [P, dp, dp]
    lib.lbm_step.argtypes = [P, ct.c_int32]
    lib.lbm_step_check.argtypes = [P, ct.POINTER(Sums)]
    lib.lbm_step_ex.argtypes = [P, ct.c_int32, ct.POINTER(Sums)]
    lib.lbm_step_ex_noalloc.argtypes = [P, ct.c_int32, ct.POINTER(Sums)]
    lib.lbm_get_macros.argtypes = [P, dp, dp]
    lib.lbm_reduce.argtypes = [P, ct.POINTER(Sums)]
    lib.lbm_get_populations.argtypes = [P, dp]
    lib.lbm_set_populations.argtypes = [P, dp]
    lib.lbm_time_steps.argtypes = [P, ct.c_int32, dp]
    lib.lbm_time_step_groups.argtypes = [P, ct.POINTER(ct.c_int32), ct.c_int32, dp]
    lib.lbm_set_launch.argtypes = [P, ct.c_int32, ct.c_int32, ct.c_int32]
    lib.lbm_set_index_width.argtypes = [P, ct.c_int32]
    lib.lbm_set_halo.argtypes = [P, ct.c_int32]
    lib.lbm_set_pressure_forcing.argtypes = [P, ct.c_int32, ct.c_double, ct.c_double]
    lib.lbm_set_storage.argtypes = [P, ct.c_int32]
    lib.lbm_info.argtypes = [P, ct.POINTER(ct.c_int64), ct.POINTER(ct.c_int32), ct.POINTER(ct.c_int32)]
    lib.lbm_synchronize.argtypes = [P]
    lib.lbm_destroy.argtypes = [P]
    lib.lbm_last_error.argtypes = [P]
    lib.lbm_last_error.restype = ct.c_char_p
    lib.lbm_copy_bench.argtypes = [ct.c_int64, ct.c_int32, dp]
    for name in EXPORTS:
        getattr(lib, name)
        if name != "lbm_last_error":
            getattr(lib, name).restype = ct.c_int
    _lib = lib
    return lib


class LBMError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"liblbm_hip error {code}: {msg}")
        self.code = code


def face_priority(kind):
    """Default corner rule of the structured API: walls win over open faces."""
    return 2 if kind == "wall" else 1


def _dptr(a):
    return a.ctypes.data_as(ct.POINTER(ct.c_double))


class Lattice:
    """One structured lattice on the GPU (cell index x + nx*(y + ny*z)).

    faces: dict like {"-y": ("wall", (0,0,0), 1.0), "+x": ("pressure_zh", None, 1.0)}; an optional
    fourth entry is the face priority (default: 2 for walls, 1 otherwise).  Unlisted faces are periodic.
    """

    def __init__(self, shape, lattice=None, storage="f32", omega=1.0, force=(0.0, 0.0, 0.0), faces=None,
                 solid=None):
        lib = load()
        self.lib = lib
        shape = tuple(int(s) for s in shape)
        if len(shape) == 2:
            shape = shape + (1,)
        self.nx, self.ny, self.nz = shape
        self.dim = 2 if self.nz == 1 else 3
        self.Q = lattice or (D2Q9 if self.dim == 2 else D3Q19)
        self.storage = storage
        self.ncell = self.nx * self.ny * self.nz
        self._solid = None
        g = GridDesc(self.dim, self.nx, self.ny, self.nz, None)
        if solid is not None:
            self._solid = np.ascontiguousarray(np.asarray(solid, dtype=np.uint8).reshape(-1))
            assert self._solid.size == self.ncell
            g.solid = self._solid.ctypes.data_as(ct.POINTER(ct.c_uint8))
        p = Params(self.Q, STORE[storage], float(omega), (ct.c_double * 3)(*[float(f) for f in force]))
        h = ct.c_void_p()
        self._check(lib.lbm_create(ct.byref(g), ct.byref(p), ct.byref(h)), None)
        self.h = h
        if faces:
            self.set_faces(faces)

    def _check(self, rc, h="self"):
        if rc != LBM_OK:
            msg = self.lib.lbm_last_error(self.h if h == "self" else h)
            raise LBMError(rc, (msg or b"").decode())

    def set_faces(self, faces):
        arr = (BcDesc * len(faces))()
        for k, (name, spec) in enumerate(faces.items()):
            kind, val, rho = spec[:3]
            prio = spec[3] if len(spec) > 3 else face_priority(kind)
            val = tuple(val) if val is not None else (0.0, 0.0, 0.0)
            val = val + (0.0,) * (3 - len(val))
            arr[k] = BcDesc(FACE_IDS[name], FACE[kind], (ct.c_double * 3)(*val), float(rho), int(prio))
        self._check(self.lib.lbm_set_bc(self.h, arr, len(faces)))

    def init(self, rho, u):
        rho = np.ascontiguousarray(np.asarray(rho, dtype=np.float64).reshape(-1))
        u = np.ascontiguousarray(np.asarray(u, dtype=np.float64).reshape(-1))
        assert rho.size == self.ncell and u.size == self.ncell * self.dim
        self._check(self.lib.lbm_init_fields(self.h, _dptr(rho), _dptr(u)))

    def step(self, n=1):
        self._check(self.lib.lbm_step(self.h, int(n)))

    def step_check(self):
        s = Sums()
        rc = self.lib.lbm_step_check(self.h, ct.byref(s))
        if rc not in (LBM_OK, LBM_DIVERGED):
            self._check(rc)
        return s, rc

    def step_ex(self, mode):
        s = Sums()
        rc = self.lib.lbm_step_ex(self.h, int(mode), ct.byref(s))
        if rc not in (LBM_OK, LBM_DIVERGED):
            self._check(rc)
        return s, rc

    def step_ex_noalloc(self, mode):
        """lbm_step_ex through the allocation-free entry point."""
        s = Sums()
        rc = self.lib.lbm_step_ex_noalloc(self.h, int(mode), ct.byref(s))
        if rc not in (LBM_OK, LBM_DIVERGED):
            self._check(rc)
        return s, rc

    def macros(self):
        rho = np.empty(self.ncell)
        u = np.empty(self.ncell * self.dim)
        self._check(self.lib.lbm_get_macros(self.h, _dptr(rho), _dptr(u)))
        return rho, u.reshape(self.ncell, self.dim)

    def reduce(self):
        s = Sums()
        self._check(self.lib.lbm_reduce(self.h, ct.byref(s)))
        return s

    def populations(self):
        f = np.empty(self.Q * self.ncell)
        self._check(self.lib.lbm_get_populations(self.h, _dptr(f)))
        return f.reshape(self.Q, self.ncell)

    def set_populations(self, f):
        f = np.ascontiguousarray(np.asarray(f, dtype=np.float64).reshape(-1))
        assert f.size == self.Q * self.ncell
        self._check(self.lib.lbm_set_populations(self.h, _dptr(f)))

    def time_steps(self, n):
        ms = ct.c_double()
        self._check(self.lib.lbm_time_steps(self.h, int(n), ct.byref(ms)))
        return ms.value

    def time_step_groups(self, groups):
        """Time consecutive groups of steps in one uninterrupted run; returns
        the milliseconds of each group."""
        g = np.ascontiguousarray(np.asarray(groups, dtype=np.int32))
        ms = np.zeros(g.size)
        self._check(self.lib.lbm_time_step_groups(self.h, g.ctypes.data_as(ct.POINTER(ct.c_int32)), int(g.size),
                                                  _dptr(ms)))
        return ms

    def set_launch(self, bx=0, by=0, nt=-1):
        self._check(self.lib.lbm_set_launch(self.h, int(bx), int(by), int(nt)))

    def set_halo(self, lds):
        """0: gather from global memory (L2 reuse); 1: LDS-staged x halo."""
        self._check(self.lib.lbm_set_halo(self.h, int(lds)))

    def set_pressure_forcing(self, on, p_inlet=1.0, p_outlet=1.0):
        """The reference's 'forcing' correction on the x faces (D2Q9)."""
        self._check(self.lib.lbm_set_pressure_forcing(self.h, int(on), float(p_inlet), float(p_outlet)))

    def set_index_width(self, bits):
        self._check(self.lib.lbm_set_index_width(self.h, int(bits)))

    def set_storage(self, storage):
        """Convert the populations to another storage form on the device."""
        self._check(self.lib.lbm_set_storage(self.h, STORE[storage]))
        self.storage = storage

    def sync(self):
        self._check(self.lib.lbm_synchronize(self.h))

    def close(self):
        if getattr(self, "h", None):
            self.lib.lbm_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def copy_bench(nbytes, reps=20):
    lib = load()
    ms = ct.c_double()
    rc = lib.lbm_copy_bench(int(nbytes), int(reps), ct.byref(ms))
    if rc != LBM_OK:
        raise LBMError(rc, lib.lbm_last_error(None).decode())
    return ms.value


# ---------------------------------------------------------------- batches
class DomainDesc(ct.Structure):
    _fields_ = [("nx", ct.c_int32), ("ny", ct.c_int32), ("nz", ct.c_int32), ("omega", ct.c_double),
                ("solid", ct.POINTER(ct.c_uint8)), ("faces", BcDesc * 6), ("nfaces", ct.c_int32)]


class DomainResult(ct.Structure):
    _fields_ = [("steps", ct.c_int64), ("res_rel", ct.c_double), ("mass", ct.c_double)]


BATCH_EXPORTS = ["lbm_batch_create", "lbm_batch_tune", "lbm_batch_set_tiling", "lbm_batch_set_launch",
                 "lbm_batch_init_fields",
                 "lbm_batch_run", "lbm_batch_get_macros", "lbm_batch_get_populations", "lbm_batch_info",
                 "lbm_batch_last_error", "lbm_batch_destroy"]
_batch_ready = False


def _load_batch():
    global _batch_ready
    lib = load()
    if _batch_ready:
        return lib
    P = ct.c_void_p
    dp = ct.POINTER(ct.c_double)
    lib.lbm_batch_create.argtypes = [ct.c_int32, ct.c_int32, ct.POINTER(DomainDesc), ct.c_int32, ct.POINTER(P)]
    lib.lbm_batch_tune.argtypes = [P, ct.c_int32, ct.c_char_p, ct.c_int32]
    lib.lbm_batch_set_tiling.argtypes = [P, ct.c_int32, ct.c_int32]
    lib.lbm_batch_set_launch.argtypes = [P, ct.c_int32]
    lib.lbm_batch_init_fields.argtypes = [P, dp, dp]
    lib.lbm_batch_run.argtypes = [P, ct.c_int32, ct.c_double, ct.c_int64, ct.POINTER(DomainResult), dp]
    lib.lbm_batch_get_macros.argtypes = [P, ct.c_int32, dp, dp]
    lib.lbm_batch_get_populations.argtypes = [P, ct.c_int32, dp]
    lib.lbm_batch_info.argtypes = [P, ct.POINTER(ct.c_int64), ct.POINTER(ct.c_int64), ct.POINTER(ct.c_int32),
                                   ct.POINTER(ct.c_int32)]
    lib.lbm_batch_last_error.argtypes = [P]
    lib.lbm_batch_last_error.restype = ct.c_char_p
    lib.lbm_batch_destroy.argtypes = [P]
    for n in BATCH_EXPORTS:
        if n != "lbm_batch_last_error":
            getattr(lib, n).restype = ct.c_int
    _batch_ready = True
    return lib


class Batch:
    """Many uneven structured domains advanced together (include/lbm_batch.h).

    domains: list of dicts {shape: (nx, ny[, nz]), omega, solid (optional flat bool array),
    faces: {"-x": (kind, u, rho[, priority]), ...}}."""

    def __init__(self, domains, storage="f32"):
        lib = _load_batch()
        self.lib = lib
        self.storage = storage
        shapes = [tuple(d["shape"]) + ((1,) if len(d["shape"]) == 2 else ()) for d in domains]
        self.dim = 2 if shapes[0][2] == 1 else 3
        self.Q = D2Q9 if self.dim == 2 else D3Q19
        self.shapes = shapes
        arr = (DomainDesc * len(domains))()
        self._keep = []
        for k, d in enumerate(domains):
            nx, ny, nz = shapes[k]
            arr[k].nx, arr[k].ny, arr[k].nz = nx, ny, nz
            arr[k].omega = float(d["omega"])
            if d.get("solid") is not None:
                s = np.ascontiguousarray(np.asarray(d["solid"], dtype=np.uint8).reshape(-1))
                assert s.size == nx * ny * nz
                self._keep.append(s)
                arr[k].solid = s.ctypes.data_as(ct.POINTER(ct.c_uint8))
            faces = d.get("faces") or {}
            for i, (name, spec) in enumerate(faces.items()):
                kind, val, rho = spec[:3]
                prio = spec[3] if len(spec) > 3 else face_priority(kind)
                val = tuple(val) if val is not None else (0.0, 0.0, 0.0)
                val = val + (0.0,) * (3 - len(val))
                arr[k].faces[i] = BcDesc(FACE_IDS[name], FACE[kind], (ct.c_double * 3)(*val), float(rho), int(prio))
            arr[k].nfaces = len(faces)
        h = ct.c_void_p()
        rc = lib.lbm_batch_create(self.Q, STORE[storage], arr, len(domains), ct.byref(h))
        if rc != LBM_OK:
            raise LBMError(rc, lib.lbm_batch_last_error(None).decode())
        self.h = h
        self.ndom = len(domains)
        self.ncells = [s[0] * s[1] * s[2] for s in shapes]

    def _check(self, rc):
        if rc != LBM_OK:
            raise LBMError(rc, self.lib.lbm_batch_last_error(self.h).decode())

    def tune(self, steps=20):
        """Times every tiling and launch form; returns (tile_x, blocks,
        one_launch, ms per step) of the fastest, which stays set."""
        buf = ct.create_string_buffer(64)
        self._check(self.lib.lbm_batch_tune(self.h, int(steps), buf, 64))
        tx, blocks, one, ms = buf.value.decode().split(",")
        return int(tx), int(blocks), int(one), float(ms)

    def set_tiling(self, tx, blocks):
        self._check(self.lib.lbm_batch_set_tiling(self.h, int(tx), int(blocks)))

    def set_launch(self, one_launch):
        self._check(self.lib.lbm_batch_set_launch(self.h, int(one_launch)))

    def init(self, rho, u):
        rho = np.ascontiguousarray(np.asarray(rho, dtype=np.float64).reshape(-1))
        u = np.ascontiguousarray(np.asarray(u, dtype=np.float64).reshape(-1))
        assert rho.size == sum(self.ncells) and u.size == rho.size * self.dim
        self._check(self.lib.lbm_batch_init_fields(self.h, _dptr(rho), _dptr(u)))

    def run(self, interval, tol, max_steps):
        res = (DomainResult * self.ndom)()
        ms = ct.c_double()
        self._check(self.lib.lbm_batch_run(self.h, int(interval), float(tol), int(max_steps), res, ct.byref(ms)))
        steps = np.array([r.steps for r in res], dtype=np.int64)
        rr = np.array([r.res_rel for r in res])
        mass = np.array([r.mass for r in res])
        return steps, rr, mass, ms.value

    def macros(self, dom):
        n = self.ncells[dom]
        rho = np.empty(n)
        u = np.empty(n * self.dim)
        self._check(self.lib.lbm_batch_get_macros(self.h, int(dom), _dptr(rho), _dptr(u)))
        return rho, u.reshape(n, self.dim)

    def populations(self, dom):
        n = self.ncells[dom]
        f = np.empty(self.Q * n)
        self._check(self.lib.lbm_batch_get_populations(self.h, int(dom), _dptr(f)))
        return f.reshape(self.Q, n)

    def info(self):
        nc, nt = ct.c_int64(), ct.c_int64()
        tx, bl = ct.c_int32(), ct.c_int32()
        self._check(self.lib.lbm_batch_info(self.h, ct.byref(nc), ct.byref(nt), ct.byref(tx), ct.byref(bl)))
        return dict(cells=nc.value, tiles=nt.value, tile_x=tx.value, blocks=bl.value)

    def close(self):
        if getattr(self, "h", None):
            self.lib.lbm_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
