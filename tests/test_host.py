"""C host and C-ABI checks that run without a GPU.

The host's parser, set-up and writers (host/params.c, host/output.c) are compared with the
reference's own host file (D3Q19-OpenCL.c, compiled by oracle/Makefile into oracle/_ref when the
reference tree is present) on the reference's input files; the ABI library is checked to load and
export every entry point its header declares.
"""
import ctypes as C
import os
import re
import shutil
import subprocess
import tempfile

import numpy as np
import pytest

import d3q19

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
GOLD = os.path.join(HERE, "golden")
REFLIB = os.path.join(ROOT, "oracle", "_ref", "libd3q19_refhost.so")

needs_ref = pytest.mark.skipif(not os.path.exists(REFLIB), reason="reference host library not built (no reference tree)")


class RefHostParams(C.Structure):
    """host_param_struct of the reference (D3Q19-OpenCL_header.h:75-94)."""
    _fields_ = [
        ("ConsolePrintFreq", C.c_int), ("InitialDist", C.c_char * 128), ("InitialVel", C.c_float * 3),
        ("InitialParticleDistribution", C.c_int), ("ParticleBuffer", C.c_float), ("InterpOrderIBM", C.c_int),
        ("DomainDecomp", C.c_int * 3), ("ParticleDensity", C.c_float), ("WorkItemSizes", C.c_size_t * 3),
        ("MaxWorkGroupSize", C.c_size_t), ("RebuildFreq", C.c_int), ("VideoFreq", C.c_int),
        ("FluidOutputSpacing", C.c_int), ("TangentialVelBC", C.c_int * 3), ("ShearStressFreq", C.c_int),
        ("RandParticleShift", C.c_float),
    ]


class RefOutput(C.Structure):
    _fields_ = [("ShearStressCount", C.c_int), ("ShearStressAvg", C.c_float), ("ActualShearRate", C.c_float)]


_ref = None


def ref():
    global _ref
    if _ref is None:
        _ref = C.CDLL(REFLIB)
    return _ref


class cwd:
    def __init__(self, d):
        self.d = d

    def __enter__(self):
        self.old = os.getcwd()
        os.chdir(self.d)

    def __exit__(self, *a):
        os.chdir(self.old)


def ref_setup(input_text, tmp):
    """Run the reference's parser and set-up in `tmp` (it reads input_file.txt and
    unit_sphere_partitions/ from the working directory)."""
    with open(os.path.join(tmp, "input_file.txt"), "w") as f:
        f.write(input_text)
    if not os.path.exists(os.path.join(tmp, "unit_sphere_partitions")):
        os.symlink(d3q19.DATA, os.path.join(tmp, "unit_sphere_partitions"))
    ip, fp, hp = d3q19.IntParams(), d3q19.FlpParams(), RefHostParams()
    with cwd(tmp):
        assert ref().initialize_data(C.byref(ip), C.byref(fp), C.byref(hp)) == 0
    return ip, fp, hp


def my_setup(input_text, tmp):
    path = os.path.join(tmp, "mine.txt")
    with open(path, "w") as f:
        f.write(input_text)
    ip, fp, hp = d3q19.IntParams(), d3q19.FlpParams(), d3q19.HostParams()
    assert d3q19.host_lib().host_initialize_data(path.encode(), C.byref(ip), C.byref(fp), C.byref(hp)) == 0
    return ip, fp, hp


INPUTS = ["input_file.txt", "input_file_power_law.txt", "input_file_casson.txt"]
QUIRKS = """iterations 7
ovito_xyz_video_freq 33
fluid_ouput_spacing 2
particle_collision_params 2.5 9.0
cpu_only_mode 1
not_a_keyword 5
total_lattice_size 20 22 24
boundary_conditions_xyz 0 0 1
viscosity_model 4
viscosity_params 0.1 0.2
velocity_bc_upper 0.01 0.0 0.0
num_particles 3
initial_particle_distribution 1
"""


@needs_ref
@pytest.mark.parametrize("name", INPUTS + ["quirks"])
def test_parser_matches_reference(name):
    text = QUIRKS if name == "quirks" else open(os.path.join(GOLD, "inputs", name)).read()
    with tempfile.TemporaryDirectory() as tmp:
        ri, rf, rh = ref_setup(text, tmp)
        mi, mf, mh = my_setup(text, tmp)
    for fld in ("MaxIterations", "ViscosityModel", "NumParticles", "ParForceModel", "MaxSurfPointsPerNode", "MaintainShear"):
        assert getattr(ri, fld) == getattr(mi, fld), fld
    for fld in ("LatticeSize", "BufferSize", "BoundaryConds"):
        assert list(getattr(ri, fld)) == list(getattr(mi, fld)), fld
    assert [list(r) for r in ri.BasisVel] == [list(r) for r in mi.BasisVel]
    # a vector keyword given fewer values than its length makes the reference copy uninitialised
    # stack values into the remaining entries (D3Q19-OpenCL.c:642-653); compare the given ones
    given = {}
    for ln in text.splitlines():
        toks = ln.split()
        if toks:
            given[toks[0]] = len(toks) - 1
    for fld, key in (("ConstBodyForce", "constant_body_force"), ("VelUpper", "velocity_bc_upper"),
                     ("VelLower", "velocity_bc_lower"), ("ViscosityParams", "viscosity_params"), ("EqWeights", None)):
        n = given.get(key, 99) if key else 99
        assert list(getattr(rf, fld))[:n] == list(getattr(mf, fld))[:n], fld
    for fld in ("NewtonianTau", "ParticleZBuffer", "ParticleDiam", "DirectForcingCoeff"):
        assert getattr(rf, fld) == getattr(mf, fld), fld
    assert rf.ParForceParams[0] == mf.ParForceParams[0]
    for fld in ("ConsolePrintFreq", "InitialParticleDistribution", "InterpOrderIBM", "RebuildFreq", "VideoFreq",
                "FluidOutputSpacing", "ShearStressFreq", "ParticleBuffer", "ParticleDensity", "RandParticleShift"):
        assert getattr(rh, fld) == getattr(mh, fld), fld
    assert rh.InitialDist == mh.InitialDist
    for fld in ("InitialVel", "DomainDecomp", "TangentialVelBC"):
        assert list(getattr(rh, fld)) == list(getattr(mh, fld)), fld


@needs_ref
@pytest.mark.parametrize("diam", [4.0, 8.0, 10.0, 12.0, 16.0])
def test_sphere_discretization_matches_reference(diam):
    text = "particle_diameter %r\nnum_particles 2\n" % diam
    with tempfile.TemporaryDirectory() as tmp:
        ri, rf, rh = ref_setup(text, tmp)
        pts = C.c_void_p()
        with cwd(tmp):
            ref().sphere_discretization(C.byref(ri), C.byref(rf), C.byref(pts))
        n = ri.PointsPerParticle
        rp = np.ctypeslib.as_array(C.cast(pts, C.POINTER(C.c_float)), shape=(4 * n,)).copy()
        mi, mf, mh = my_setup(text, tmp)
        mh.DataDir = d3q19.DATA.encode()
        mp = np.zeros(4 * 1024, dtype=np.float32)
        assert d3q19.host_lib().host_sphere_discretization(C.byref(mi), C.byref(mf), C.byref(mh), mp.ctypes.data_as(C.c_void_p)) == 0
    assert mi.PointsPerParticle == n and mi.TotalSurfPoints == ri.TotalSurfPoints
    assert mf.PointArea == rf.PointArea
    assert np.array_equal(mp[: 4 * n], rp)


@needs_ref
@pytest.mark.parametrize("name", INPUTS)
def test_lattice_fields_match_reference(name):
    text = open(os.path.join(GOLD, "inputs", name)).read().replace("total_lattice_size              96 96 96", "total_lattice_size 20 18 16")
    with tempfile.TemporaryDirectory() as tmp:
        ri, rf, rh = ref_setup(text, tmp)
        mi, mf, mh = my_setup(text, tmp)
    nn = ri.LatticeSize[0] * ri.LatticeSize[1] * ri.LatticeSize[2]
    rfld = [np.zeros(19 * nn, np.float32), np.zeros(3 * nn * ri.MaxSurfPointsPerNode, np.float32), np.zeros(3 * nn, np.float32),
            np.zeros(nn, np.float32), np.zeros(nn, np.int32)]
    ref().initialize_lattice_fields(C.byref(rh), C.byref(ri), C.byref(rf), *[a.ctypes.data_as(C.c_void_p) for a in rfld])
    mfld = [np.zeros(19 * nn, np.float32), np.zeros(3 * nn, np.float32), np.zeros(nn, np.float32)]
    d3q19.host_lib().host_initialize_lattice_fields(C.byref(mh), C.byref(mi), C.byref(mf), *[a.ctypes.data_as(C.c_void_p) for a in mfld])
    assert list(ri.SystemSize) == list(mi.SystemSize)
    assert np.array_equal(rfld[0], mfld[0]) and np.array_equal(rfld[2], mfld[1]) and np.array_equal(rfld[3], mfld[2])


@needs_ref
@pytest.mark.parametrize("dist,np_,shift", [(1, 8, 0.0), (1, 5, 0.0), (2, 27, 0.0), (2, 12, 1.5), (3, 1, 0.0)])
def test_particle_fields_match_reference(dist, np_, shift):
    text = "total_lattice_size 40 40 41\nboundary_conditions_xyz 0 0 1\nnum_particles %d\ninitial_particle_distribution %d\n" \
           "random_particle_shift %r\ninitial_particle_buffer 6.0\nz_wall_particle_buffer 4.0\nparticle_diameter 6.0\n" \
           "particle_density 1.3\nviscosity_model 1\n" % (np_, dist, shift)
    libc = C.CDLL("libc.so.6")
    with tempfile.TemporaryDirectory() as tmp:
        ri, rf, rh = ref_setup(text, tmp)
        mi, mf, mh = my_setup(text, tmp)
        pts = C.c_void_p()
        with cwd(tmp):
            ref().sphere_discretization(C.byref(ri), C.byref(rf), C.byref(pts))
    ri.WorkGroupsPerParticle = 1
    mi.WorkGroupsPerParticle = 1
    nn = ri.LatticeSize[0] * ri.LatticeSize[1] * ri.LatticeSize[2]
    scratch = [np.zeros(19 * nn, np.float32), np.zeros(3 * nn * 8, np.float32), np.zeros(3 * nn, np.float32), np.zeros(nn, np.float32), np.zeros(nn, np.int32)]
    ref().initialize_lattice_fields(C.byref(rh), C.byref(ri), C.byref(rf), *[a.ctypes.data_as(C.c_void_p) for a in scratch])
    d3q19.host_lib().host_system_size(C.byref(mi))
    rk, rfo, rff = np.zeros(16 * np_, np.float32), np.zeros(8 * np_, np.float32), np.zeros(8 * np_, np.float32)
    mk, mfo, mff = np.zeros(16 * np_, np.float32), np.zeros(8 * np_, np.float32), np.zeros(8 * np_, np.float32)
    libc.srand(7)
    ref().initialize_particle_fields(C.byref(rh), C.byref(ri), C.byref(rf), *[a.ctypes.data_as(C.c_void_p) for a in (rk, rfo, rff)])
    libc.srand(7)
    assert d3q19.host_lib().host_initialize_particle_fields(C.byref(mh), C.byref(mi), C.byref(mf), *[a.ctypes.data_as(C.c_void_p) for a in (mk, mfo, mff)]) == 0
    assert rf.ParticleMass == mf.ParticleMass and rf.ParticleMomInertia == mf.ParticleMomInertia
    assert np.array_equal(rk, mk)
    # zone geometry of initialize_particle_zones
    hp_ref = rh
    parsZone, threadMembers, numParInThread = np.zeros(np_, np.int32), np.zeros(8 * np_, np.int32), np.zeros(8, np.int32)
    zm, nz, zn = C.c_void_p(), C.c_void_p(), C.c_void_p()
    ref().initialize_particle_zones(C.byref(hp_ref), C.byref(ri), C.byref(rf), rk.ctypes.data_as(C.c_void_p), parsZone.ctypes.data_as(C.c_void_p),
                                    C.byref(zm), C.byref(nz), threadMembers.ctypes.data_as(C.c_void_p), numParInThread.ctypes.data_as(C.c_void_p), C.byref(zn))
    d3q19.host_lib().host_zone_geometry(C.byref(mh), C.byref(mi), C.byref(mf))
    assert list(ri.NumZones) == list(mi.NumZones) and list(rf.ZoneWidth) == list(mf.ZoneWidth)


@needs_ref
def test_shear_stress_and_writers_match_reference():
    text = open(os.path.join(GOLD, "inputs", "input_file.txt")).read().replace("96 96 96", "14 12 17")
    rng = np.random.default_rng(3)
    with tempfile.TemporaryDirectory() as tmp:
        ri, rf, rh = ref_setup(text, tmp)
        mi, mf, mh = my_setup(text, tmp)
        d3q19.host_lib().host_system_size(C.byref(mi))
        ri.SystemSize[:] = list(mi.SystemSize)
        nn = 14 * 12 * 17
        u = (0.05 * rng.standard_normal(3 * nn)).astype(np.float32)
        tau = np.ones(nn, np.float32)
        os.makedirs(os.path.join(tmp, "ref", "matlab_postproc"))
        os.makedirs(os.path.join(tmp, "mine", "matlab_postproc"))
        ro, mo = RefOutput(), d3q19.OutputData()
        for k in range(2):
            with cwd(os.path.join(tmp, "ref")):
                ref().compute_shear_stress(C.byref(ro), C.byref(rh), C.byref(ri), C.byref(rf), u.ctypes.data_as(C.c_void_p),
                                           tau.ctypes.data_as(C.c_void_p), 100 + k)
            d3q19.host_lib().host_compute_shear_stress(C.byref(mo), C.byref(mh), C.byref(mi), C.byref(mf), u.ctypes.data_as(C.c_void_p),
                                                       os.path.join(tmp, "mine").encode(), 1)
            u[: nn] *= 1.1
        assert (ro.ShearStressCount, ro.ShearStressAvg, ro.ActualShearRate) == (mo.ShearStressCount, mo.ShearStressAvg, mo.ActualShearRate)
        with cwd(os.path.join(tmp, "ref")):
            ref().write_lattice_field(u.ctypes.data_as(C.c_void_p), C.byref(ri))
        d3q19.host_lib().host_write_lattice_field(u.ctypes.data_as(C.c_void_p), C.byref(mi), os.path.join(tmp, "mine").encode())
        for f in ("velocity_profile_z.txt", "velocity_field_final.txt", "matlab_postproc/velocity_field_centerline.txt"):
            a = open(os.path.join(tmp, "ref", f)).read()
            b = open(os.path.join(tmp, "mine", f)).read()
            assert a == b, f


def test_abi_library_exports_every_declared_entry_point():
    hdr = open(os.path.join(ROOT, "include", "d3q19_abi.h")).read()
    names = sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(d3q19_\w+)\s*\(", hdr, re.M)))
    assert len(names) >= 30
    L = d3q19.lib()
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    assert C.sizeof(d3q19.IntParams) == 332 and C.sizeof(d3q19.FlpParams) == 176


def test_kernel_parameters_carry_no_address_space():
    # The global-address-space qualifier exists in the device compilation only; on a kernel
    # parameter it changes the device-side mangled name, and the launch then finds no kernel.
    csrc = os.path.join(ROOT, "dispersion-shear_d3q19-opencl_amd", "csrc")
    bad = []
    for f in sorted(os.listdir(csrc)):
        if not f.endswith((".hip", ".h")):
            continue
        src = open(os.path.join(csrc, f)).read()
        for m in re.finditer(r"__global__[^{;]*?\b(\w+)\s*\(([^)]*)\)\s*\{", src):
            if "D3Q19_GLOBAL" in m.group(2):
                bad.append("%s:%s" % (f, m.group(1)))
    assert not bad, bad


def test_footprint_and_error_strings():
    L = d3q19.lib()
    c = d3q19.Case(lines=["total_lattice_size 66 66 67", "boundary_conditions_xyz 0 0 1"])
    b = C.c_size_t()
    assert L.d3q19_footprint(C.byref(c.ip), C.byref(c.opt), C.byref(b)) == 0
    assert b.value >= 2 * 19 * 4 * 64 * 64 * 65
    for code in range(0, -8, -1):
        assert L.d3q19_strerror(code)


def test_invalid_parameters_rejected_before_device_use():
    L = d3q19.lib()
    c = d3q19.Case(lines=["total_lattice_size 12 12 13", "boundary_conditions_xyz 0 0 1"])
    c.ip.SystemSize[2] = 99  # inconsistent with LatticeSize and walls
    h = C.c_void_p()
    rc = L.d3q19_create(0, C.byref(c.ip), C.byref(c.fp), C.byref(c.opt), None, C.byref(h))
    assert rc == -1 and b"SystemSize" in L.d3q19_last_error()
    c = d3q19.Case(lines=["total_lattice_size 12 12 13", "boundary_conditions_xyz 0 0 1", "population_storage 1",
                          "population_rounding 2"])
    assert c.opt.storage == 1 and c.opt.rounding == 2
    rc = L.d3q19_create(0, C.byref(c.ip), C.byref(c.fp), C.byref(c.opt), None, C.byref(h))
    assert rc == -1 and b"rounding" in L.d3q19_last_error()


def test_population_plane_above_4_gib_rejected_before_device_use():
    """Kernels address a population plane by a 32-bit byte offset, so a lattice whose plane (interior
    nodes x bytes per stored value) reaches 4 GiB is refused: 1024^3 interior nodes in fp32, 512 x
    1024 x 1024 in fp64."""
    L = d3q19.lib()
    h = C.c_void_p()
    for nx, storage in ((1024, 0), (512, 3)):
        # a small case, its lattice sizes then raised (the host-side arrays stay small)
        c = d3q19.Case(lines=["total_lattice_size 12 12 13", "boundary_conditions_xyz 0 0 1"])
        for d, n in enumerate((nx, 1024, 1024)):
            c.ip.LatticeSize[d] = n + 2 + (d == 2)
            c.ip.SystemSize[d] = n
        c.opt.storage = storage
        rc = L.d3q19_create(0, C.byref(c.ip), C.byref(c.fp), C.byref(c.opt), None, C.byref(h))
        assert rc == -4 and b"4 GiB" in L.d3q19_last_error()


def test_parameter_checking_rejects_two_wall_axes():
    with pytest.raises(ValueError):
        d3q19.Case(lines=["boundary_conditions_xyz 1 0 1"])


def test_malformed_marker_file_rejected():
    with tempfile.TemporaryDirectory() as tmp:
        with open(os.path.join(tmp, "unit_sphere_partition_128.txt"), "w") as f:
            f.write("0.5,0.5,0.5\n" * 128)
        with pytest.raises(RuntimeError):
            d3q19.Case(lines=["particle_diameter 4.0", "num_particles 1", "marker_count 128"], data_dir=tmp)


def test_host_program_reports_missing_gpu(gpu_available):
    if gpu_available:
        pytest.skip("a GPU is present")
    exe = os.path.join(ROOT, "dispersion-shear_d3q19-opencl_amd", "bin", "d3q19_sim")
    with tempfile.TemporaryDirectory() as tmp:
        shutil.copy(os.path.join(GOLD, "inputs", "input_file.txt"), tmp)
        r = subprocess.run([exe], cwd=tmp, capture_output=True, text=True, timeout=120)
    assert r.returncode != 0


def test_marker_count_above_buffer_capacity_rejected():
    """The marker buffer holds 4 x 1024 floats (the reference's largest automatic set); a larger
    marker_count is an input error, not a buffer overrun."""
    with pytest.raises(ValueError):
        d3q19.Case(lines=["particle_diameter 4.0", "num_particles 1", "marker_count 2048"])
    with pytest.raises(ValueError):
        d3q19.Case(lines=["particle_diameter 4.0", "num_particles 1", "marker_count -3"])


def test_sphere_beyond_brick_marking_capacity_rejected_before_device_use():
    """The brick marking lists at most 64 bricks of 4 nodes per axis for one sphere; d3q19_create
    refuses a sphere whose stencils could span more (diameter above about 240 lattice units) instead
    of marking it incompletely: at d = 250 the markers' floor coordinates span 253 nodes, their
    padded stencils 261 nodes, which can touch 67 bricks.  d = 200 is accepted up to the device
    check."""
    L = d3q19.lib()
    c = d3q19.Case(lines=["total_lattice_size 40 40 41", "boundary_conditions_xyz 0 0 1", "num_particles 1",
                          "particle_diameter 8.0", "marker_count 128", "initial_particle_distribution 3"])
    h = C.c_void_p()
    for diam, want_reject in ((250.0, True), (200.0, False)):
        c.fp.ParticleDiam = diam
        pts = (c.points.reshape(-1, 4) * (diam / 8.0)).astype(np.float32).ravel()
        rc = L.d3q19_create(0, C.byref(c.ip), C.byref(c.fp), C.byref(c.opt), pts.ctypes.data_as(C.c_void_p), C.byref(h))
        if want_reject:
            assert rc == d3q19.E_UNSUPPORTED and b"brick marking" in L.d3q19_last_error()
        else:
            assert rc != d3q19.E_UNSUPPORTED
            if rc == 0:
                L.d3q19_destroy(h)


def test_build_refuses_an_object_whose_kernels_lack_device_code(tmp_path):
    """build.py checks every compiled object: each kernel that its host half registers by name must
    be defined in its gfx950 code object, so a library that would abort at launch with "Cannot find
    Symbol" (a host half paired with device code of another source) fails the build instead.  An
    object whose .hip_fatbin is replaced by another object's device code is refused."""
    import build as pkg_build
    srcs = {"a": "k_alpha", "b": "k_beta"}
    for name, kern in srcs.items():
        with open(os.path.join(tmp_path, name + ".hip"), "w") as f:
            f.write("#include <hip/hip_runtime.h>\n__global__ void %s(int* p) { *p = 1; }\n"
                    "void launch_%s(int* p) { hipLaunchKernelGGL(%s, dim3(1), dim3(1), 0, 0, p); }\n" % (kern, name, kern))
        subprocess.run([pkg_build.HIPCC, "--offload-arch=gfx950", "-c", "-fPIC", os.path.join(tmp_path, name + ".hip"), "-o",
                        os.path.join(tmp_path, name + ".o")], check=True, capture_output=True)
    pkg_build.check_kernel_registration(os.path.join(tmp_path, "a.o"))
    fb, bad = os.path.join(tmp_path, "b.fatbin"), os.path.join(tmp_path, "bad.o")
    subprocess.run([os.path.join(pkg_build.LLVM, "llvm-objcopy"), "--dump-section", ".hip_fatbin=" + fb, os.path.join(tmp_path, "b.o"),
                    os.path.join(tmp_path, "discard.o")], check=True)
    subprocess.run([os.path.join(pkg_build.LLVM, "llvm-objcopy"), "--update-section", ".hip_fatbin=" + fb, os.path.join(tmp_path, "a.o"),
                    bad], check=True)
    with pytest.raises(RuntimeError, match="k_alpha"):
        pkg_build.check_kernel_registration(bad)
