"""ctypes binding of the Poisson step library (include/lbm_poisson.h), part of
liblbm_hip.so.

    p = Poisson(lattice=5, nx=128, ny=128, omega=1.0, source=s)
    p.set_faces({"-x": ("dirichlet", values_along_y, 2), "+x": ("neumann", None, 0), ...})
    p.init(phi0)                 # phi0[ny][nx] or flat, x fastest
    p.step(100); d, rc = p.step_check(); phi = p.phi()
"""
import ctypes as ct

import numpy as np

from . import capi

EXPORTS = ["lbm_poisson_create", "lbm_poisson_set_faces", "lbm_poisson_init", "lbm_poisson_step",
           "lbm_poisson_step_check", "lbm_poisson_get", "lbm_poisson_destroy", "lbm_poisson_last_error"]

DIRICHLET, NEUMANN = 1, 2
FACE_IDS = {"-x": 0, "+x": 1, "-y": 2, "+y": 3}


class Params(ct.Structure):
    _fields_ = [("lattice", ct.c_int32), ("nx", ct.c_int32), ("ny", ct.c_int32), ("omega", ct.c_double),
                ("source", ct.c_double)]


class Face(ct.Structure):
    _fields_ = [("face", ct.c_int32), ("kind", ct.c_int32), ("priority", ct.c_int32),
                ("value", ct.POINTER(ct.c_double))]


def _bind(lib):
    if getattr(lib, "_poisson_bound", False):
        return lib
    P = ct.c_void_p
    dp = ct.POINTER(ct.c_double)
    lib.lbm_poisson_create.argtypes = [ct.POINTER(Params), ct.POINTER(P)]
    lib.lbm_poisson_set_faces.argtypes = [P, ct.POINTER(Face), ct.c_int32]
    lib.lbm_poisson_init.argtypes = [P, dp]
    lib.lbm_poisson_step.argtypes = [P, ct.c_int32]
    lib.lbm_poisson_step_check.argtypes = [P, dp]
    lib.lbm_poisson_get.argtypes = [P, dp]
    lib.lbm_poisson_destroy.argtypes = [P]
    lib.lbm_poisson_last_error.argtypes = [P]
    lib.lbm_poisson_last_error.restype = ct.c_char_p
    for n in EXPORTS:
        if n != "lbm_poisson_last_error":
            getattr(lib, n).restype = ct.c_int
    lib._poisson_bound = True
    return lib


class PoissonError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"lbm_poisson error {code}: {msg}")
        self.code = code


def _dp(a):
    return a.ctypes.data_as(ct.POINTER(ct.c_double))


class Poisson:
    def __init__(self, lattice, nx, ny, omega, source):
        self.lib = _bind(capi.load())
        self.Q, self.nx, self.ny = int(lattice), int(nx), int(ny)
        p = Params(self.Q, self.nx, self.ny, float(omega), float(source))
        h = ct.c_void_p()
        rc = self.lib.lbm_poisson_create(ct.byref(p), ct.byref(h))
        if rc != 0:
            raise PoissonError(rc, self.lib.lbm_poisson_last_error(None).decode())
        self.h = h

    def _check(self, rc):
        if rc != 0:
            raise PoissonError(rc, self.lib.lbm_poisson_last_error(self.h).decode())

    def set_faces(self, faces):
        """faces: {"-x"|"+x"|"-y"|"+y": (kind "dirichlet"/"neumann", values along the face or None, priority)}."""
        arr = (Face * len(faces))()
        self._vals = []
        for k, (name, (kind, values, prio)) in enumerate(faces.items()):
            arr[k].face = FACE_IDS[name]
            arr[k].kind = DIRICHLET if kind == "dirichlet" else NEUMANN
            arr[k].priority = int(prio)
            if values is not None:
                v = np.ascontiguousarray(np.asarray(values, dtype=np.float64).reshape(-1))
                self._vals.append(v)
                arr[k].value = _dp(v)
        self._check(self.lib.lbm_poisson_set_faces(self.h, arr, len(faces)))

    def init(self, phi0):
        phi0 = np.ascontiguousarray(np.asarray(phi0, dtype=np.float64).reshape(-1))
        assert phi0.size == self.nx * self.ny
        self._check(self.lib.lbm_poisson_init(self.h, _dp(phi0)))

    def step(self, n=1):
        self._check(self.lib.lbm_poisson_step(self.h, int(n)))

    def step_check(self):
        d = ct.c_double()
        rc = self.lib.lbm_poisson_step_check(self.h, ct.byref(d))
        if rc not in (0, capi.LBM_DIVERGED):
            self._check(rc)
        return d.value, rc

    def phi(self):
        out = np.empty(self.nx * self.ny)
        self._check(self.lib.lbm_poisson_get(self.h, _dp(out)))
        return out

    def close(self):
        if getattr(self, "h", None):
            self.lib.lbm_poisson_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
