"""Python mirror of the C-ABI in include/d3q19_abi.h (ctypes).

This is the host-side binding the tests and the benchmark use; it adds no behaviour of its own.
The library it loads is lib/libd3q19.so, built from csrc/ for gfx950.  There is no CPU
fallback: loading fails loudly when the library is missing.
"""
import ctypes as C
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIBDIR = os.path.join(PKG, "lib")
DATA = os.path.join(PKG, "data", "unit_sphere_partitions")


class IntParams(C.Structure):
    """int_param_struct (struct_header_host.h:4-31), 332 bytes."""
    _fields_ = [
        ("MaxIterations", C.c_int32),
        ("BasisVel", (C.c_int32 * 3) * 19),
        ("LatticeSize", C.c_int32 * 3),
        ("SystemSize", C.c_int32 * 3),
        ("BufferSize", C.c_int32 * 3),
        ("BoundaryConds", C.c_int32 * 3),
        ("NumZones", C.c_int32 * 3),
        ("MaintainShear", C.c_int32),
        ("ViscosityModel", C.c_int32),
        ("NumParticles", C.c_int32),
        ("ParForceModel", C.c_int32),
        ("MaxSurfPointsPerNode", C.c_int32),
        ("InterpOrderIBM", C.c_int32),
        ("PointsPerParticle", C.c_int32),
        ("PointsPerWorkGroup", C.c_int32),
        ("TotalSurfPoints", C.c_int32),
        ("WorkGroupsPerParticle", C.c_int32),
    ]


class FlpParams(C.Structure):
    """flp_param_struct (struct_header_host.h:34-57), 176 bytes."""
    _fields_ = [
        ("ConstBodyForce", C.c_float * 3),
        ("EqWeights", C.c_float * 19),
        ("VelUpper", C.c_float * 3),
        ("VelLower", C.c_float * 3),
        ("NewtonianTau", C.c_float),
        ("ViscosityParams", C.c_float * 4),
        ("ZoneWidth", C.c_float * 3),
        ("ParticleDiam", C.c_float),
        ("ParticleMass", C.c_float),
        ("PointArea", C.c_float),
        ("ParticleMomInertia", C.c_float),
        ("ParForceParams", C.c_float * 2),
        ("ParticleZBuffer", C.c_float),
        ("DirectForcingCoeff", C.c_float),
    ]


class Options(C.Structure):
    _fields_ = [
        ("collision", C.c_int32),
        ("variable_viscosity", C.c_int32),
        ("tau_min", C.c_float),
        ("tau_max", C.c_float),
        ("ibm_kernel", C.c_int32),
        ("storage", C.c_int32),
        ("calc_rho", C.c_int32),
        ("tangential_bc", C.c_int32 * 3),
        ("block", C.c_int32 * 3),
        ("contact_detection", C.c_int32),
        ("rebuild_freq", C.c_int32),
        ("rounding", C.c_int32),
        ("overlap_slabs", C.c_int32),
        ("ibm_layout", C.c_int32),
        ("reserved", C.c_int32 * 1),
    ]


class HostParams(C.Structure):
    """host_param_struct of host/d3q19_host.h."""
    _fields_ = [
        ("ConsolePrintFreq", C.c_int),
        ("InitialDist", C.c_char * 128),
        ("InitialVel", C.c_float * 3),
        ("InitialParticleDistribution", C.c_int),
        ("ParticleBuffer", C.c_float),
        ("InterpOrderIBM", C.c_int),
        ("DomainDecomp", C.c_int * 3),
        ("ParticleDensity", C.c_float),
        ("RebuildFreq", C.c_int),
        ("VideoFreq", C.c_int),
        ("FluidOutputSpacing", C.c_int),
        ("TangentialVelBC", C.c_int * 3),
        ("ShearStressFreq", C.c_int),
        ("RandParticleShift", C.c_float),
        ("MarkerCount", C.c_int),
        ("PlacementSeed", C.c_longlong),
        ("PlacementGap", C.c_float),
        ("CollisionModel", C.c_int),
        ("VariableViscosity", C.c_int),
        ("TauClamp", C.c_float * 2),
        ("IbmKernel", C.c_int),
        ("Storage", C.c_int),
        ("MassModel", C.c_int),
        ("ContactDetection", C.c_int),
        ("Rounding", C.c_int),
        ("DataDir", C.c_char * 256),
    ]


class OutputData(C.Structure):
    _fields_ = [("ShearStressCount", C.c_int), ("ShearStressAvg", C.c_float), ("ActualShearRate", C.c_float)]


class Observables(C.Structure):
    _fields_ = [("wall_stress", C.c_double), ("wall_stress_lower", C.c_double), ("wall_stress_upper", C.c_double),
                ("shear_rate", C.c_double), ("apparent_viscosity", C.c_double), ("fluid_viscosity", C.c_double),
                ("relative_viscosity", C.c_double), ("drag_coefficient", C.c_double), ("mean_fluid_force", C.c_double),
                ("wall_axis", C.c_int32), ("flow_axis", C.c_int32), ("reserved", C.c_double * 4)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "reserved"}


class BatchOptions(C.Structure):
    _fields_ = [("storage", C.c_int32), ("collision", C.c_int32), ("ibm_kernel", C.c_int32), ("max_cells", C.c_int32),
                ("arena_bytes", C.c_uint64), ("reserved", C.c_int32 * 8)]


class Tuning(C.Structure):
    _fields_ = [("tile", C.c_int32), ("grid_collide", C.c_int32), ("grid_sphere", C.c_int32), ("grid_ibm", C.c_int32),
                ("collide_ms", C.c_double), ("span", C.c_int32), ("reserved", C.c_int32 * 3)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "reserved"}


class SolveParams(C.Structure):
    _fields_ = [("tolerance", C.c_double), ("check_interval", C.c_int32), ("min_steps", C.c_int32),
                ("max_steps", C.c_int32), ("reserved", C.c_int32 * 5)]


class Result(C.Structure):
    _fields_ = [("id", C.c_int64), ("status", C.c_int32), ("steps", C.c_int32), ("obs", Observables),
                ("t_submit", C.c_double), ("t_admit", C.c_double), ("t_finish", C.c_double)]


COLL_MRT, COLL_BGK = 0, 1
IBM_LINEAR, IBM_PESKIN4 = 0, 1
IBM_LAYOUT_AUTO, IBM_LAYOUT_WAVE, IBM_LAYOUT_ROWS, IBM_LAYOUT_ROWS_LARGE = 0, 1, 2, 3
STORE_FP32, STORE_FP16, STORE_BF16, STORE_FP64 = 0, 1, 2, 3
CONTACT_ZONES, CONTACT_ALL_PAIRS = 0, 1
ROUND_STOCHASTIC, ROUND_NEAREST = 0, 1
STEP_WRITE_U = 1
E_INVALID, E_DEVICE, E_NOMEM, E_UNSUPPORTED, E_STATE, E_NONFINITE, E_CAPACITY = -1, -2, -3, -4, -5, -6, -7

assert C.sizeof(IntParams) == 332 and C.sizeof(FlpParams) == 176

_host = None
_lib = None


def host_lib():
    """libd3q19host.so: parser / set-up / writers of the C host (no GPU)."""
    global _host
    if _host is None:
        path = os.path.join(LIBDIR, "libd3q19host.so")
        if not os.path.exists(path):
            raise RuntimeError("libd3q19host.so not built; run build.py")
        h = C.CDLL(path)
        P = C.POINTER
        h.host_initialize_data.argtypes = [C.c_char_p, P(IntParams), P(FlpParams), P(HostParams)]
        h.host_process_input_line.argtypes = [C.c_char_p, P(IntParams), P(FlpParams), P(HostParams)]
        h.host_parameter_checking.argtypes = [P(IntParams), P(FlpParams), P(HostParams)]
        h.host_system_size.argtypes = [P(IntParams)]
        h.host_sphere_discretization.argtypes = [P(IntParams), P(FlpParams), P(HostParams), C.c_void_p]
        h.host_workgroup_split.argtypes = [P(IntParams), C.c_int]
        h.host_initialize_lattice_fields.argtypes = [P(HostParams), P(IntParams), P(FlpParams), C.c_void_p, C.c_void_p, C.c_void_p]
        h.host_initialize_particle_fields.argtypes = [P(HostParams), P(IntParams), P(FlpParams), C.c_void_p, C.c_void_p, C.c_void_p]
        h.host_zone_geometry.argtypes = [P(HostParams), P(IntParams), P(FlpParams)]
        h.host_equilibrium.argtypes = [C.c_float, C.c_void_p, C.c_void_p]
        h.host_compute_shear_stress.argtypes = [P(OutputData), P(HostParams), P(IntParams), P(FlpParams), C.c_void_p, C.c_char_p, C.c_int]
        h.host_write_lattice_field.argtypes = [C.c_void_p, P(IntParams), C.c_char_p]
        h.host_fill_options.argtypes = [P(HostParams), P(Options)]
        _host = h
    return _host


def lib():
    """libd3q19.so, the HIP library.  Raises if it is not built."""
    global _lib
    if _lib is None:
        path = os.environ.get("D3Q19_LIBRARY") or os.path.join(LIBDIR, "libd3q19.so")
        if not os.path.exists(path):
            raise RuntimeError("libd3q19.so not built; run build.py (no CPU fallback exists)")
        L = C.CDLL(path)
        P, V = C.POINTER, C.c_void_p
        sim_pp = P(V)
        L.d3q19_device_count.argtypes = [P(C.c_int)]
        L.d3q19_device_name.argtypes = [C.c_int, C.c_char_p, C.c_size_t]
        L.d3q19_create.argtypes = [C.c_int, P(IntParams), P(FlpParams), P(Options), V, sim_pp]
        L.d3q19_destroy.argtypes = [V]
        for n in ("write_f", "read_f", "write_f64", "read_f64", "write_u", "read_u", "write_tau", "read_tau", "read_force"):
            getattr(L, "d3q19_" + n).argtypes = [V, V]
        L.d3q19_write_particles.argtypes = [V, V, V, V]
        L.d3q19_read_particles.argtypes = [V, V, V, V]
        L.d3q19_read_particles_f64.argtypes = [V, V, V, V, V, V]
        L.d3q19_collide_stream.argtypes = [V, C.c_int]
        for n in ("particle_dynamics", "particle_fluid_forces", "particle_particle_forces", "sync", "check_finite"):
            getattr(L, "d3q19_" + n).argtypes = [V]
        L.d3q19_step.argtypes = [V, C.c_int, C.c_int]
        L.d3q19_run.argtypes = [V, C.c_int, C.c_int, C.c_int]
        L.d3q19_plane_mean_u.argtypes = [V, V]
        L.d3q19_plane_mean_u_noalloc.argtypes = [V, V]
        L.d3q19_moments.argtypes = [V, V]
        L.d3q19_time_collide.argtypes = [V, C.c_int, C.c_int, P(C.c_double)]
        L.d3q19_time_steps.argtypes = [V, C.c_int, C.c_int, P(C.c_double), P(C.c_double)]
        L.d3q19_footprint.argtypes = [P(IntParams), P(Options), P(C.c_size_t)]
        L.d3q19_observables_get.argtypes = [V, P(Observables)]
        L.d3q19_checkpoint_size.argtypes = [V, P(C.c_size_t)]
        L.d3q19_checkpoint_save.argtypes = [V, V, C.c_size_t]
        L.d3q19_checkpoint_load.argtypes = [V, V, C.c_size_t]
        L.d3q19_checkpoint_write.argtypes = [V, C.c_char_p]
        L.d3q19_checkpoint_read.argtypes = [V, C.c_char_p]
        L.d3q19_batch_create.argtypes = [C.c_int, P(BatchOptions), P(V)]
        L.d3q19_batch_destroy.argtypes = [V]
        L.d3q19_batch_add.argtypes = [V, P(IntParams), P(FlpParams), P(Options), V, V, P(C.c_int)]
        L.d3q19_batch_remove.argtypes = [V, C.c_int]
        L.d3q19_batch_cell_bytes.argtypes = [V, P(IntParams), P(C.c_size_t)]
        L.d3q19_batch_memory.argtypes = [V, P(C.c_size_t), P(C.c_size_t), P(C.c_int)]
        L.d3q19_batch_run.argtypes = [V, C.c_int, C.c_int]
        L.d3q19_batch_sync.argtypes = [V]
        L.d3q19_batch_observables.argtypes = [V, P(C.c_int), C.c_int, P(Observables)]
        L.d3q19_batch_read_u.argtypes = [V, C.c_int, V]
        L.d3q19_batch_read_f.argtypes = [V, C.c_int, V]
        L.d3q19_batch_read_particles_f64.argtypes = [V, C.c_int, V, V, V, V, V]
        L.d3q19_batch_autotune.argtypes = [V, P(Tuning)]
        L.d3q19_batch_get_tuning.argtypes = [V, P(Tuning)]
        L.d3q19_batch_set_tuning.argtypes = [V, P(Tuning)]
        L.d3q19_batch_time_steps.argtypes = [V, C.c_int, P(C.c_double), P(C.c_double)]
        L.d3q19_server_create.argtypes = [C.c_int, P(BatchOptions), P(V)]
        L.d3q19_server_destroy.argtypes = [V]
        L.d3q19_server_submit.argtypes = [V, P(IntParams), P(FlpParams), P(Options), V, V, P(SolveParams), P(C.c_int64)]
        L.d3q19_server_progress.argtypes = [V, P(Result), C.c_int, P(C.c_int)]
        L.d3q19_server_status.argtypes = [V, P(C.c_int), P(C.c_int), P(C.c_size_t)]
        L.d3q19_server_get_batch.argtypes = [V, P(V)]
        L.d3q19_solve.argtypes = [C.c_int, P(IntParams), P(FlpParams), P(Options), V, V, P(SolveParams), P(Result)]
        L.d3q19_clock.restype = C.c_double
        L.d3q19_stream_probe.argtypes = [C.c_int] * 9 + [P(C.c_double), P(C.c_double)]
        L.d3q19_strerror.argtypes = [C.c_int]
        L.d3q19_strerror.restype = C.c_char_p
        L.d3q19_last_error.restype = C.c_char_p
        _lib = L
    return _lib


class D3Q19Error(RuntimeError):
    def __init__(self, code, where):
        L = lib()
        msg = L.d3q19_strerror(code).decode() + ": " + (L.d3q19_last_error() or b"").decode()
        super().__init__("%s failed (%d): %s" % (where, code, msg))
        self.code = code


def check(code, where):
    if code != 0:
        raise D3Q19Error(code, where)


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


class Case:
    """A configured run: parameter structs, host options and initial arrays, built with the C
    host's own parser and set-up functions."""

    def __init__(self, lines=None, path=None, max_wg=256, data_dir=DATA):
        h = host_lib()
        self.ip, self.fp, self.hp = IntParams(), FlpParams(), HostParams()
        import tempfile
        if path is None:
            fd, tmp = tempfile.mkstemp(suffix=".txt")
            with os.fdopen(fd, "w") as f:
                f.write("\n".join(lines or []) + "\n")
            path = tmp
        else:
            tmp = None
        try:
            rc = h.host_initialize_data(path.encode(), C.byref(self.ip), C.byref(self.fp), C.byref(self.hp))
            if rc != 0:
                raise RuntimeError("cannot read input file " + path)
            # marker folder: keep the file's value unless it is the default relative name
            if self.hp.DataDir == b"unit_sphere_partitions":
                self.hp.DataDir = data_dir.encode()
            if h.host_parameter_checking(C.byref(self.ip), C.byref(self.fp), C.byref(self.hp)) > 0:
                raise ValueError("parameter checking failed")
            h.host_system_size(C.byref(self.ip))
            self.points = np.zeros(4 * 1024, dtype=np.float32)
            if h.host_sphere_discretization(C.byref(self.ip), C.byref(self.fp), C.byref(self.hp), _ptr(self.points)) != 0:
                raise RuntimeError("sphere discretization failed")
            h.host_workgroup_split(C.byref(self.ip), max_wg)
            self.points = self.points[: 4 * self.ip.PointsPerParticle].copy()
            nn = self.nn
            self.f = np.empty(19 * nn, dtype=np.float32)
            self.u = np.empty(3 * nn, dtype=np.float32)
            self.tau = np.empty(nn, dtype=np.float32)
            h.host_initialize_lattice_fields(C.byref(self.hp), C.byref(self.ip), C.byref(self.fp), _ptr(self.f), _ptr(self.u), _ptr(self.tau))
            np_ = max(self.ip.NumParticles, 0)
            nfa = max(self.ip.WorkGroupsPerParticle, 1)
            self.parKin = np.zeros(16 * max(np_, 1), dtype=np.float32)
            self.parForce = np.zeros(8 * max(np_, 1), dtype=np.float32)
            self.parFluid = np.zeros(8 * max(np_, 1) * nfa, dtype=np.float32)
            if h.host_initialize_particle_fields(C.byref(self.hp), C.byref(self.ip), C.byref(self.fp), _ptr(self.parKin), _ptr(self.parForce), _ptr(self.parFluid)) != 0:
                raise RuntimeError("particle placement failed")
            h.host_zone_geometry(C.byref(self.hp), C.byref(self.ip), C.byref(self.fp))
            self.opt = Options()
            h.host_fill_options(C.byref(self.hp), C.byref(self.opt))
        finally:
            if tmp:
                os.unlink(tmp)

    @property
    def N(self):
        return tuple(self.ip.LatticeSize)

    @property
    def nn(self):
        N = self.N
        return N[0] * N[1] * N[2]

    @property
    def np(self):
        return self.ip.NumParticles


class Sim:
    """One simulation on the GPU through the C-ABI."""

    def __init__(self, case, device=0, opt=None):
        L = lib()
        self.case = case
        self.L = L
        self.h = C.c_void_p()
        o = opt if opt is not None else case.opt
        pts = case.points if case.np > 0 else None
        check(L.d3q19_create(device, C.byref(case.ip), C.byref(case.fp), C.byref(o), _ptr(pts), C.byref(self.h)), "d3q19_create")
        self.nn = case.nn
        self.np = case.np

    def close(self):
        if self.h:
            self.L.d3q19_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def write_f(self, f):
        f = np.ascontiguousarray(f, dtype=np.float32)
        check(self.L.d3q19_write_f(self.h, _ptr(f)), "write_f")

    def read_f(self):
        f = np.empty(19 * self.nn, dtype=np.float32)
        check(self.L.d3q19_read_f(self.h, _ptr(f)), "read_f")
        return f

    def write_f64(self, f):
        f = np.ascontiguousarray(f, dtype=np.float64)
        check(self.L.d3q19_write_f64(self.h, _ptr(f)), "write_f64")

    def read_f64(self):
        f = np.empty(19 * self.nn, dtype=np.float64)
        check(self.L.d3q19_read_f64(self.h, _ptr(f)), "read_f64")
        return f

    def write_u(self, u):
        u = np.ascontiguousarray(u, dtype=np.float32)
        check(self.L.d3q19_write_u(self.h, _ptr(u)), "write_u")

    def read_u(self):
        u = np.empty(3 * self.nn, dtype=np.float32)
        check(self.L.d3q19_read_u(self.h, _ptr(u)), "read_u")
        return u

    def write_tau(self, t):
        t = np.ascontiguousarray(t, dtype=np.float32)
        check(self.L.d3q19_write_tau(self.h, _ptr(t)), "write_tau")

    def read_tau(self):
        t = np.empty(self.nn, dtype=np.float32)
        check(self.L.d3q19_read_tau(self.h, _ptr(t)), "read_tau")
        return t

    def read_force(self):
        g = np.empty(3 * self.nn, dtype=np.float32)
        check(self.L.d3q19_read_force(self.h, _ptr(g)), "read_force")
        return g

    def write_particles(self, parKin=None, parForce=None, parFluid=None):
        a = [None if x is None else np.ascontiguousarray(x, dtype=np.float32) for x in (parKin, parForce, parFluid)]
        check(self.L.d3q19_write_particles(self.h, _ptr(a[0]), _ptr(a[1]), _ptr(a[2])), "write_particles")

    def read_particles(self):
        np_ = max(self.np, 1)
        nfa = max(self.case.ip.WorkGroupsPerParticle, 1)
        pk = np.zeros(16 * np_, dtype=np.float32)
        pf = np.zeros(8 * np_, dtype=np.float32)
        pff = np.zeros(8 * np_ * nfa, dtype=np.float32)
        check(self.L.d3q19_read_particles(self.h, _ptr(pk), _ptr(pf), _ptr(pff)), "read_particles")
        return pk, pf, pff

    def read_particles_f64(self):
        n = max(self.np, 1)
        out = [np.zeros(3 * n) for _ in range(5)]
        check(self.L.d3q19_read_particles_f64(self.h, *[_ptr(x) for x in out]), "read_particles_f64")
        return out

    def collide_stream(self, flags=0):
        check(self.L.d3q19_collide_stream(self.h, flags), "collide_stream")

    def particle_dynamics(self):
        check(self.L.d3q19_particle_dynamics(self.h), "particle_dynamics")

    def particle_fluid_forces(self):
        check(self.L.d3q19_particle_fluid_forces(self.h), "particle_fluid_forces")

    def particle_particle_forces(self):
        check(self.L.d3q19_particle_particle_forces(self.h), "particle_particle_forces")

    def step(self, t, flags=0):
        check(self.L.d3q19_step(self.h, t, flags), "step")

    def run(self, t_first, n, flags=0):
        check(self.L.d3q19_run(self.h, t_first, n, flags), "run")

    def sync(self):
        check(self.L.d3q19_sync(self.h), "sync")

    def check_finite(self):
        return self.L.d3q19_check_finite(self.h)

    def plane_mean_u(self):
        out = np.zeros(3 * self.case.N[2])
        check(self.L.d3q19_plane_mean_u(self.h, _ptr(out)), "plane_mean_u")
        return out.reshape(3, -1)

    def plane_mean_u_noalloc(self):
        """d3q19_plane_mean_u_noalloc: the same plane means, staged in the library's device scratch."""
        out = np.zeros(3 * self.case.N[2])
        check(self.L.d3q19_plane_mean_u_noalloc(self.h, _ptr(out)), "plane_mean_u_noalloc")
        return out.reshape(3, -1)

    def moments(self):
        out = np.zeros(4)
        check(self.L.d3q19_moments(self.h, _ptr(out)), "moments")
        return out

    def time_collide(self, n, flags=0):
        ms = C.c_double()
        check(self.L.d3q19_time_collide(self.h, n, flags, C.byref(ms)), "time_collide")
        return ms.value

    def time_steps(self, t_first, n):
        """n real steps, each kernel timed: (total ms, mean step-kernel ms, mean IBM-pass ms)."""
        tot, k = C.c_double(), (C.c_double * 2)()
        check(self.L.d3q19_time_steps(self.h, t_first, n, C.byref(tot), k), "time_steps")
        return tot.value, k[0], k[1]

    def observables(self):
        o = Observables()
        check(self.L.d3q19_observables_get(self.h, C.byref(o)), "observables_get")
        return o

    def checkpoint(self):
        """The whole simulation state as bytes (d3q19_checkpoint_save)."""
        n = C.c_size_t()
        check(self.L.d3q19_checkpoint_size(self.h, C.byref(n)), "checkpoint_size")
        buf = C.create_string_buffer(n.value)
        check(self.L.d3q19_checkpoint_save(self.h, buf, n.value), "checkpoint_save")
        return buf.raw

    def restore(self, image):
        """Load a state written by checkpoint() (d3q19_checkpoint_load)."""
        buf = C.create_string_buffer(bytes(image), len(image))
        check(self.L.d3q19_checkpoint_load(self.h, buf, len(image)), "checkpoint_load")

    def checkpoint_write(self, path):
        check(self.L.d3q19_checkpoint_write(self.h, str(path).encode()), "checkpoint_write")

    def checkpoint_read(self, path):
        check(self.L.d3q19_checkpoint_read(self.h, str(path).encode()), "checkpoint_read")

    def upload_case(self):
        c = self.case
        self.write_f(c.f)
        self.write_u(c.u)
        self.write_tau(c.tau)
        if c.np > 0:
            self.write_particles(c.parKin, c.parForce, c.parFluid)


def _case_arrays(case):
    pts = case.points if case.np > 0 else None
    kin = case.parKin if case.np > 0 else None
    return pts, kin


class Batch:
    """A batched ensemble of shear cells (one arena, persistent kernels, graph-replayed steps)."""

    def __init__(self, storage=STORE_FP32, collision=COLL_BGK, ibm=IBM_PESKIN4, arena_bytes=0, max_cells=0, device=0):
        self.L = lib()
        self.h = C.c_void_p()
        bo = BatchOptions()
        bo.storage, bo.collision, bo.ibm_kernel, bo.max_cells, bo.arena_bytes = storage, collision, ibm, max_cells, arena_bytes
        check(self.L.d3q19_batch_create(device, C.byref(bo), C.byref(self.h)), "batch_create")
        self.cases = {}

    def close(self):
        if self.h:
            self.L.d3q19_batch_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add(self, case, opt=None):
        cid = C.c_int(-1)
        pts, kin = _case_arrays(case)
        o = opt if opt is not None else case.opt
        check(self.L.d3q19_batch_add(self.h, C.byref(case.ip), C.byref(case.fp), C.byref(o), _ptr(pts), _ptr(kin), C.byref(cid)),
              "batch_add")
        self.cases[cid.value] = case
        return cid.value

    def remove(self, cid):
        check(self.L.d3q19_batch_remove(self.h, cid), "batch_remove")
        self.cases.pop(cid, None)

    def cell_bytes(self, case):
        b = C.c_size_t()
        check(self.L.d3q19_batch_cell_bytes(self.h, C.byref(case.ip), C.byref(b)), "batch_cell_bytes")
        return b.value

    def memory(self):
        a, u, n = C.c_size_t(), C.c_size_t(), C.c_int()
        check(self.L.d3q19_batch_memory(self.h, C.byref(a), C.byref(u), C.byref(n)), "batch_memory")
        return a.value, u.value, n.value

    def run(self, n, flags=0):
        check(self.L.d3q19_batch_run(self.h, n, flags), "batch_run")

    def sync(self):
        check(self.L.d3q19_batch_sync(self.h), "batch_sync")

    def observables(self, ids):
        ids = list(ids)
        arr = (C.c_int * len(ids))(*ids)
        out = (Observables * len(ids))()
        check(self.L.d3q19_batch_observables(self.h, arr, len(ids), out), "batch_observables")
        return list(out)

    def _nn(self, cid):
        c = self.cases.get(cid)
        return c.nn if c is not None else 1  # unknown ids are rejected by the library

    def read_u(self, cid):
        u = np.empty(3 * self._nn(cid), dtype=np.float32)
        check(self.L.d3q19_batch_read_u(self.h, cid, _ptr(u)), "batch_read_u")
        return u

    def read_f(self, cid):
        f = np.empty(19 * self._nn(cid), dtype=np.float32)
        check(self.L.d3q19_batch_read_f(self.h, cid, _ptr(f)), "batch_read_f")
        return f

    def read_particles_f64(self, cid):
        c = self.cases.get(cid)
        n = max(c.np if c is not None else 1, 1)
        out = [np.zeros(3 * n) for _ in range(5)]
        check(self.L.d3q19_batch_read_particles_f64(self.h, cid, *[_ptr(x) for x in out]), "batch_read_particles_f64")
        return out

    def autotune(self):
        t = Tuning()
        check(self.L.d3q19_batch_autotune(self.h, C.byref(t)), "batch_autotune")
        return t

    def tuning(self):
        t = Tuning()
        check(self.L.d3q19_batch_get_tuning(self.h, C.byref(t)), "batch_get_tuning")
        return t

    def set_tuning(self, t):
        check(self.L.d3q19_batch_set_tuning(self.h, C.byref(t)), "batch_set_tuning")

    def time_steps(self, n):
        tot, col = C.c_double(), C.c_double()
        check(self.L.d3q19_batch_time_steps(self.h, n, C.byref(tot), C.byref(col)), "batch_time_steps")
        return tot.value, col.value


def solve_params(tolerance=1e-4, check_interval=100, min_steps=0, max_steps=100000):
    sp = SolveParams()
    sp.tolerance, sp.check_interval, sp.min_steps, sp.max_steps = tolerance, check_interval, min_steps, max_steps
    return sp


class Server:
    """Continuous-batching solver of shear-cell requests (admission by device-memory footprint)."""

    def __init__(self, storage=STORE_FP32, collision=COLL_BGK, ibm=IBM_PESKIN4, arena_bytes=0, max_cells=0, device=0):
        self.L = lib()
        self.h = C.c_void_p()
        bo = BatchOptions()
        bo.storage, bo.collision, bo.ibm_kernel, bo.max_cells, bo.arena_bytes = storage, collision, ibm, max_cells, arena_bytes
        check(self.L.d3q19_server_create(device, C.byref(bo), C.byref(self.h)), "server_create")

    def close(self):
        if self.h:
            self.L.d3q19_server_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit(self, case, sp, opt=None):
        rid = C.c_int64()
        pts, kin = _case_arrays(case)
        o = opt if opt is not None else case.opt
        check(self.L.d3q19_server_submit(self.h, C.byref(case.ip), C.byref(case.fp), C.byref(o), _ptr(pts), _ptr(kin),
                                         C.byref(sp), C.byref(rid)), "server_submit")
        return rid.value

    def progress(self, cap=256):
        out = (Result * cap)()
        n = C.c_int()
        check(self.L.d3q19_server_progress(self.h, out, cap, C.byref(n)), "server_progress")
        return list(out[: n.value])

    def status(self):
        q, r, u = C.c_int(), C.c_int(), C.c_size_t()
        check(self.L.d3q19_server_status(self.h, C.byref(q), C.byref(r), C.byref(u)), "server_status")
        return q.value, r.value, u.value

    def batch_handle(self):
        b = C.c_void_p()
        check(self.L.d3q19_server_get_batch(self.h, C.byref(b)), "server_get_batch")
        return b


def solve(case, sp, opt=None, device=0):
    """Plain sequential solve of one request (single simulation, same convergence rule)."""
    L = lib()
    res = Result()
    pts, kin = _case_arrays(case)
    o = opt if opt is not None else case.opt
    check(L.d3q19_solve(device, C.byref(case.ip), C.byref(case.fp), C.byref(o), _ptr(pts), _ptr(kin), C.byref(sp), C.byref(res)),
          "solve")
    return res


def clock():
    return lib().d3q19_clock()


def stream_probe(nx, ny, nz, streams=19, width=1, pull=True, lds_bytes=0, launches=20, device=0, nontemporal=False):
    """Attainable bandwidth of the collide's access pattern without its arithmetic
    (d3q19_stream_probe): returns (GB/s counting every byte read and written, ms per launch).
    nontemporal: non-temporal loads and stores (chunk layouts, width -64 / -256)."""
    g, ms = C.c_double(), C.c_double()
    flags = (1 if pull else 0) | (2 if nontemporal else 0)
    check(lib().d3q19_stream_probe(device, nx, ny, nz, streams, width, flags, lds_bytes, launches, C.byref(g), C.byref(ms)),
          "stream_probe")
    return g.value, ms.value
