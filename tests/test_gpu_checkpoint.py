"""GPU: checkpoint / restart continues bit for bit.

A run of n + m steps is compared with a run of n steps, a checkpoint, and m more steps in a fresh
simulation that loaded it: populations, velocity, tau and the FP64 sphere state must be bitwise
equal.  Cases cover spheres with contacts through the zone lists (rebuilt inside the resumed
part), per-node tau, fp16-shifted storage, a restart through a file, and the rejection of an image
written with other parameters.
"""
import numpy as np
import pytest

import d3q19
from cases import case

pytestmark = pytest.mark.gpu


def state(s, c):
    s.sync()
    out = [s.read_f(), s.read_u()]
    if c.opt.variable_viscosity:
        out.append(s.read_tau())
    if c.np:
        out += [np.ravel(a) for a in s.read_particles_f64()]
    return out


def cases():
    from test_zones import zone_case
    return {
        "zones-powerlaw": zone_case((20, 30, 16), (0, 0, 1), 10, 6.0, 3),
        "fp16-peskin": case(size=(24, 20, 22), np_=4, diam=6.0, markers=128, dist=4, seed=8, ibm=1, uw=0.02, storage=1,
                            visc=2, vparams=(0.02, 0.7, 0, 0), local=1, clamp=(0.51, 5.0), tau=0.9),
        "mrt-linear-fp64": case(size=(16, 14, 12), coll=0, ibm=0, np_=1, diam=6.0, markers=128, tau=0.8, uw=0.02, storage=3),
    }


@pytest.mark.parametrize("name", ["zones-powerlaw", "fp16-peskin", "mrt-linear-fp64"])
def test_restart_is_bitwise(name, tmp_path):
    c = cases()[name]
    n, m = 7, 9
    a = d3q19.Sim(c)
    a.upload_case()
    a.run(1, n + m, d3q19.STEP_WRITE_U)
    want = state(a, c)
    a.close()

    b = d3q19.Sim(c)
    b.upload_case()
    b.run(1, n)
    img = b.checkpoint()
    b.close()
    r = d3q19.Sim(c)
    if name == "fp16-peskin":
        p = tmp_path / "cell.ck"
        p.write_bytes(img)
        r.checkpoint_read(p)
    else:
        r.restore(img)
    r.run(n + 1, m, d3q19.STEP_WRITE_U)
    got = state(r, c)
    r.close()
    print("%s: image %d bytes, %d arrays compared" % (name, len(img), len(want)), flush=True)
    for x, y in zip(got, want):
        assert x.shape == y.shape
        assert np.array_equal(x.view(np.uint8), y.view(np.uint8))


def test_image_of_other_parameters_is_rejected():
    c1 = case(size=(16, 14, 12), np_=1, diam=6.0, markers=128, tau=0.8, uw=0.02)
    c2 = case(size=(16, 14, 12), np_=1, diam=6.0, markers=128, tau=0.9, uw=0.02)
    a = d3q19.Sim(c1)
    a.upload_case()
    img = a.checkpoint()
    a.close()
    b = d3q19.Sim(c2)
    with pytest.raises(d3q19.D3Q19Error):
        b.restore(img)
    with pytest.raises(d3q19.D3Q19Error):
        b.restore(img[:100])
    b.close()
