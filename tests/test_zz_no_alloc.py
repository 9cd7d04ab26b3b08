"""phpfdtd_set_dft_noalloc (the DFT set-up that allocates nothing: weights and
accumulator in the library's static per-device workspace) against
phpfdtd_set_dft through the C ABI.

Bar: the same bin and the same accumulated DFT bit for bit (both paths
accumulate the same FP64 weights into a zeroed accumulator with the same
kernel), on two calls in a row on one context, the second with a smaller plane
and fewer samples, in both numerical forms; a no-alloc DFT after an allocating
one on the same context; sizes beyond the workspace (65536 samples, 1048576
plane cells) rejected with PHPFDTD_EINVAL.
"""
import ctypes

import numpy as np
import pytest

import golden_cases
from php_fdtd_amd import _abi, workloads
from php_fdtd_amd.sim import Simulation

pytestmark = pytest.mark.gpu

# (component, axis, position, interval, samples, steps): the second call has
# the smaller plane (x plane 20 x 16 against the 20 x 20 z plane) and fewer samples
CALLS = (("Ez", "z", 7, 5, 40, 200), ("Hy", "x", 8, 4, 30, 120))


def set_dft(sim, noalloc, comp, axis, pos, interval, nsamp):
    fn = sim.lib.phpfdtd_set_dft_noalloc if noalloc else sim.lib.phpfdtd_set_dft
    k = ctypes.c_int()
    rc = fn(sim.h, _abi.COMPONENTS[comp], axis.encode(), pos, 1, interval, nsamp,
            float(sim.spec["grid"]["f_source"]), ctypes.byref(k))
    _abi.check(sim.lib, rc)
    sim._dft_shape = sim._plane_shape(axis)
    return k.value


def two_calls(spec, med, prec, noalloc, first_alloc=False):
    sim = Simulation(spec, med, prec)
    out = []
    for q, (comp, axis, pos, interval, nsamp, steps) in enumerate(CALLS):
        k = set_dft(sim, noalloc and not (first_alloc and q == 0), comp, axis, pos, interval, nsamp)
        sim.run(steps)
        x, n = sim.dft()
        out.append((k, n, x))
    sim.close()
    return out


@pytest.mark.parametrize("precision", [_abi.FP64, _abi.FP32])
def test_noalloc_dft_matches_allocating(precision):
    c = golden_cases.load("hbn")
    spec = workloads.with_steps(c["spec"], 400)
    ref = two_calls(spec, c["medium"], precision, False)
    got = two_calls(spec, c["medium"], precision, True)
    for (k0, n0, x0), (k1, n1, x1), call in zip(ref, got, CALLS):
        print("call", call, "bin", k0, "samples", n0, flush=True)
        assert k1 == k0 and n1 == n0 == call[4]
        assert np.abs(x0).max() > 0
        assert x1.shape == x0.shape
        np.testing.assert_array_equal(x1, x0)
    assert got[1][2].size < got[0][2].size
    # a no-alloc DFT replacing an allocating one on the same context
    mixed = two_calls(spec, c["medium"], precision, True, first_alloc=True)
    for (k0, n0, x0), (k1, n1, x1) in zip(ref, mixed):
        assert k1 == k0 and n1 == n0
        np.testing.assert_array_equal(x1, x0)


def test_noalloc_dft_limits():
    c = golden_cases.load("hbn")
    sim = Simulation(c["spec"], c["medium"], _abi.FP64)
    with pytest.raises(_abi.PhpFdtdError) as e:
        set_dft(sim, True, "Ez", "z", 7, 1, 65537)
    assert e.value.code == _abi.EINVAL and "workspace" in str(e.value)
    sim.close()
    # a 1040 x 1024 z plane has more cells than the workspace holds
    f = 3e13
    g = workloads._grid((1040, 1024, 2), 20e-9, 4, 0, f)
    spec = dict(name="wide", grid=g, materials=workloads.builtin_table(f), sources=[])
    med = np.zeros((1040, 1024, 2), np.int32)
    sim = Simulation(spec, med, _abi.FP32)
    with pytest.raises(_abi.PhpFdtdError) as e:
        set_dft(sim, True, "Ez", "z", 0, 1, 4)
    assert e.value.code == _abi.EINVAL
    assert set_dft(sim, False, "Ez", "z", 0, 1, 4) >= 0  # the allocating form takes it
    sim.close()
