"""GPU: population storage precision against the fp64 build on a Casson suspension.

Scaled-down form of the 256x256x128 precision workload: a 64x64x32 box, 20 spheres, 300 steps.
fp32, fp16-shifted and bf16-shifted storage (all fp32 arithmetic) are compared with the fp64 build
of the same kernels on the same seeded input.

The 16-bit formats round stochastically (the default); they are also run with round-to-nearest
stores, which are reported but not held to bars.  Bars (about 4x the measured errors; the runs
are deterministic):
- fp32: u within 1e-4 relative L-inf (measured 2.7e-5); drag coefficient within 2e-4 absolute
  (measured 3.5e-5, C_d ~ 5).
- fp16-shifted: u within 2e-3 relative L-inf (measured 4.2e-4; 5.6e-4 rounding to nearest); drag
  coefficient within 1e-3 relative (measured 2.1e-7; 2.9e-4 rounding to nearest), the target for
  reduced-precision storage.
- bf16-shifted: a 7-bit mantissa on f - w: u within 2e-2 (measured 4.5e-3) and drag within 2e-3
  relative (measured 4.0e-4; 1.0e-3 rounding to nearest).
"""
import json
import math

import numpy as np
import pytest

import d3q19
from workloads import c3_deck, deck, precision_study

pytestmark = pytest.mark.gpu


def small_c3(storage=0):
    return c3_deck(storage=storage, size=(64, 64, 32), spheres=20)


def test_storage_precision_against_fp64_build():
    r = precision_study(small_c3, 300)
    print(json.dumps(r, indent=1), flush=True)
    assert r["fp32"]["u_linf_rel"] < 1e-4
    assert r["fp32"]["drag_abs_err"] < 2e-4
    assert r["fp16_shift"]["u_linf_rel"] < 2e-3
    assert r["fp16_shift"]["drag_rel_err"] < 1e-3
    for k in ("u_l2_rel", "u_linf_rel", "drag_abs_err"):
        assert math.isfinite(r["bf16_shift"][k])
    assert r["bf16_shift"]["u_linf_rel"] < 2e-2
    assert r["bf16_shift"]["drag_rel_err"] < 2e-3
    for k in ("fp16_shift_nearest", "bf16_shift_nearest"):
        assert math.isfinite(r[k]["u_linf_rel"]) and math.isfinite(r[k]["drag_abs_err"])


def uniform_flow(storage, rounding):
    """32^3 periodic box at rest density with the uniform velocity (0.01, 0.005, 0): every node
    holds the same populations, so round-to-nearest gives every node the same rounding error."""
    return deck(size=32, walls=(0, 0, 0), tau=0.8, uw=0.0, init="constant", u0=(0.01, 0.005, 0.0), storage=storage,
                rounding=rounding)


def one_step_from(storage, rounding, f_in=None):
    c = d3q19.Case(lines=uniform_flow(storage, rounding))
    s = d3q19.Sim(c)
    s.upload_case()
    if f_in is not None:
        s.write_f64(f_in)
    f0 = s.read_f64()
    s.run(1, 1)
    f1 = s.read_f64().reshape(19, c.N[2], c.N[1], c.N[0])[:, 1:-1, 1:-1, 1:-1].reshape(19, -1)
    s.close()
    return f0, f1


@pytest.mark.parametrize("storage,mant", [(1, 10), (2, 7)])
def test_stochastic_rounding_is_unbiased(storage, mant):
    """One collide from a uniform state stored in 16 bits, against the same step with fp32
    storage from the same (decoded) input.  Per population, the mean over the 32768 nodes of the
    stored value minus the fp32 value, in units of the 16-bit format's spacing there:
    - stochastic rounding: within 0.03 (the standard error of the mean is at most
      0.5/sqrt(32768) = 0.003), and every node holds one of the two neighbouring 16-bit values;
    - round to nearest: every node makes the same error, the mean error of some population
      exceeds 0.05 (a check that this test can tell the two apart).
    Populations whose 16-bit spacing is below 64 float32 ulps of f are left out: there the fp32
    arithmetic of the two runs differs by a sizeable part of a spacing (the rest population and
    those with c.u = 0 hold f - w ~ -1.5 u^2 w, which the shifted formats resolve more finely than
    float32 resolves f)."""
    w = np.array([1.0 / 3.0] + [1.0 / 18.0] * 6 + [1.0 / 36.0] * 12)  # rest, axis, diagonal
    f_in, f_sr = one_step_from(storage, d3q19.ROUND_STOCHASTIC)
    _, f_rn = one_step_from(storage, d3q19.ROUND_NEAREST)
    _, f32 = one_step_from(0, 0, f_in)
    ref = f32.mean(axis=1)
    assert np.abs(f32 - ref[:, None]).max() <= 1e-7 * np.abs(ref).max()  # uniform after the step
    # spacing of the stored quantity at the reference value, in population units
    if storage == 1:
        y = (ref - w) / w * 16384.0
        spacing = 2.0 ** (np.floor(np.log2(np.abs(y))) - mant) * w / 16384.0
    else:
        y = ref - w
        spacing = 2.0 ** (np.floor(np.log2(np.abs(y))) - mant)
    keep = spacing >= 64.0 * 2.0 ** -24 * np.abs(ref)
    assert keep.sum() >= 8
    bias_sr = ((f_sr.mean(axis=1) - ref) / spacing)[keep]
    bias_rn = ((f_rn.mean(axis=1) - ref) / spacing)[keep]
    spread = (np.abs(f_sr - ref[:, None]).max(axis=1) / spacing)[keep]
    print("storage %d, %d populations: max |mean error| stochastic %.4f, nearest %.4f (spacings); max node error %.3f" % (
        storage, keep.sum(), np.abs(bias_sr).max(), np.abs(bias_rn).max(), spread.max()), flush=True)
    assert np.abs(bias_sr).max() < 0.03
    assert spread.max() < 1.01  # the two runs' fp32 arithmetic differs by ~0.002 spacings
    assert np.abs(bias_rn).max() > 0.05


def test_stochastic_rounding_is_deterministic():
    """Same input, same bits: two fp16 runs of 20 steps on the sphere workload agree bitwise, and
    differ from round-to-nearest."""
    out = []
    for rnd in (0, 0, 1):
        c = d3q19.Case(lines=small_c3(storage=1) + ["population_rounding %d" % rnd])
        s = d3q19.Sim(c)
        s.write_particles(c.parKin)
        s.run(1, 20)
        out.append(s.read_f())
        s.close()
    assert np.array_equal(out[0], out[1])
    assert not np.array_equal(out[0], out[2])
