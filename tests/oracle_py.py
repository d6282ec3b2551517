"""ctypes wrapper of the CPU oracle (oracle/d3q19_oracle.c).

Test infrastructure: the checker for the HIP library.  Only tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg use it.
"""
import ctypes as C
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
BUILD = os.path.join(ORACLE_DIR, "_build")

F_C, F_S, U, TAU, GPF, PARKIN, PARFORCE, PARFLUID, POINTS = range(9)
_libs = {}


def ensure_built():
    for sfx in ("f32", "f64"):
        if not os.path.exists(os.path.join(BUILD, "liboracle_%s.so" % sfx)):
            subprocess.run(["make", "-C", ORACLE_DIR], check=True, capture_output=True)
            break


def load(sfx, path=None):
    """The oracle library of precision sfx ("f32" / "f64"); `path` loads another build of the same
    source (the pin tests load deliberately broken builds to show that their checks fail)."""
    key = (sfx, path)
    if key in _libs:
        return _libs[key]
    if path is None:
        ensure_built()
    L = C.CDLL(path or os.path.join(BUILD, "liboracle_%s.so" % sfx))
    V = C.c_void_p
    getattr(L, "orc_create_" + sfx).restype = V
    getattr(L, "orc_create_" + sfx).argtypes = [V, V, V]
    for n in ("destroy", "step", "collide_stream", "reset_forces", "boundary_periodic", "ibm",
              "particle_dynamics", "particle_forces", "swap"):
        getattr(L, "orc_%s_%s" % (n, sfx)).argtypes = [V]
    getattr(L, "orc_boundary_velocity_" + sfx).argtypes = [V, C.c_int, C.c_int]
    getattr(L, "orc_run_" + sfx).argtypes = [V, C.c_int]
    getattr(L, "orc_ptr_" + sfx).restype = V
    getattr(L, "orc_ptr_" + sfx).argtypes = [V, C.c_int]
    getattr(L, "orc_len_" + sfx).restype = C.c_long
    getattr(L, "orc_len_" + sfx).argtypes = [V, C.c_int]
    getattr(L, "orc_moment_matrix_" + sfx).argtypes = [V, V, V]
    getattr(L, "orc_feq_" + sfx).argtypes = [C.c_double, V, V]
    getattr(L, "orc_guo_" + sfx).argtypes = [V, V, V]
    getattr(L, "orc_tau_of_" + sfx).argtypes = [V, C.c_double]
    getattr(L, "orc_tau_of_" + sfx).restype = C.c_double
    getattr(L, "orc_peskin_" + sfx).argtypes = [C.c_double]
    getattr(L, "orc_peskin_" + sfx).restype = C.c_double
    getattr(L, "orc_set_s_override_" + sfx).argtypes = [V, V]
    getattr(L, "orc_zone_lists_" + sfx).argtypes = [V, V, V, V, V]
    getattr(L, "orc_zone_lists_" + sfx).restype = C.c_int
    getattr(L, "orc_steps_done_" + sfx).argtypes = [V]
    getattr(L, "orc_steps_done_" + sfx).restype = C.c_int
    getattr(L, "orc_pair_force_of_" + sfx).argtypes = [V, C.c_int, C.c_int]
    _libs[key] = L
    return L


class Oracle:
    """One oracle simulation; arrays are numpy views on the oracle's own memory."""

    def __init__(self, case, precision="f32", opt=None, lib_path=None):
        self.sfx = precision
        self.L = load(precision, lib_path)
        self.dtype = np.float32 if precision == "f32" else np.float64
        o = opt if opt is not None else case.opt
        self.h = getattr(self.L, "orc_create_" + precision)(C.byref(case.ip), C.byref(case.fp), C.byref(o))
        self.case = case

    def fn(self, name):
        return getattr(self.L, "orc_%s_%s" % (name, self.sfx))

    def arr(self, which):
        p = self.fn("ptr")(self.h, which)
        n = self.fn("len")(self.h, which)
        return np.ctypeslib.as_array(C.cast(p, C.POINTER(C.c_float if self.dtype == np.float32 else C.c_double)), shape=(n,))

    def load_case(self):
        c = self.case
        self.arr(F_C)[:] = c.f
        self.arr(F_S)[:] = c.f
        self.arr(U)[:] = c.u
        self.arr(TAU)[:] = c.tau
        if c.np > 0:
            self.arr(PARKIN)[:] = c.parKin[: 16 * c.np]
            self.arr(PARFORCE)[:] = c.parForce[: 8 * c.np]
            self.arr(PARFLUID)[:] = c.parFluid[: len(self.arr(PARFLUID))]
            self.arr(POINTS)[:] = c.points[: 4 * c.ip.PointsPerParticle]
        return self

    def step(self, n=1):
        if n == 1:
            self.fn("step")(self.h)
        else:
            self.fn("run")(self.h, n)

    def call(self, name, *args):
        self.fn(name)(self.h, *args)

    def zone_lists(self):
        """(pars_zone, num_in_zone, members[nzt, np], neigh[nzt, 28]) from the current positions."""
        c = self.case
        nzt = int(np.prod(list(c.ip.NumZones)))
        pz, nz = np.zeros(max(c.np, 1), np.int32), np.zeros(nzt, np.int32)
        mem, nb = np.zeros(nzt * max(c.np, 1), np.int32), np.zeros(28 * nzt, np.int32)
        self.fn("zone_lists")(self.h, *[a.ctypes.data_as(C.c_void_p) for a in (pz, nz, mem, nb)])
        return pz[: c.np], nz, mem.reshape(nzt, max(c.np, 1)), nb.reshape(nzt, 28)

    def close(self):
        if self.h:
            self.fn("destroy")(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def moment_matrix(self):
        M = np.zeros(361)
        Mi = np.zeros(361)
        self.fn("moment_matrix")(self.h, M.ctypes.data_as(C.c_void_p), Mi.ctypes.data_as(C.c_void_p))
        return M.reshape(19, 19), Mi.reshape(19, 19)

    def tau_of(self, srt):
        return self.fn("tau_of")(self.h, srt)

    def set_s_override(self, s):
        if s is None:
            self.fn("set_s_override")(self.h, None)
        else:
            s = np.ascontiguousarray(s, dtype=np.float64)
            self.fn("set_s_override")(self.h, s.ctypes.data_as(C.c_void_p))
            self._keep = s


def feq(rho, u, precision="f64"):
    out = np.zeros(19)
    u = np.ascontiguousarray(u, dtype=np.float64)
    getattr(load(precision), "orc_feq_" + precision)(rho, u.ctypes.data_as(C.c_void_p), out.ctypes.data_as(C.c_void_p))
    return out


def guo(u, g, precision="f64"):
    out = np.zeros(19)
    u = np.ascontiguousarray(u, dtype=np.float64)
    g = np.ascontiguousarray(g, dtype=np.float64)
    getattr(load(precision), "orc_guo_" + precision)(u.ctypes.data_as(C.c_void_p), g.ctypes.data_as(C.c_void_p), out.ctypes.data_as(C.c_void_p))
    return out


def peskin(r, precision="f64"):
    return getattr(load(precision), "orc_peskin_" + precision)(r)
