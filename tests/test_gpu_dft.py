"""Near-field single-bin DFT (phpfdtd_set_dft / phpfdtd_read_dft) against the
FFT frequency_analysis.c takes of the slice files.

frequency_analysis.c (FFTW forward transform of every pixel's slice time
series, divided by N, bin nearest the source frequency among i < N/2 with
fs = (N-1)/T; frequency_analysis.c:362-390, 548-569) is restated here with
numpy.fft.  Bars:
  * bin index equal to that choice;
  * device DFT of the FP64 plain form vs numpy FFT of the reference's own
    float32 slice files (the golden fixtures): relative L2 <= 1e-6 (the slices
    agree to 1 float32 ulp);
  * vs numpy FFT of the library's own slices at the same steps:
    relative L2 <= 1e-12, for the source bin and for a bin with non-zero
    phase weights;
  * FP32 precision form vs FP64: amplitude relative L2 <= 1e-3.
"""
import numpy as np
import pytest

import golden_cases
from php_fdtd_amd import _abi
from php_fdtd_amd.sim import Simulation

pytestmark = pytest.mark.gpu


def rel_l2(got, ref):
    nr = np.linalg.norm(ref)
    return np.linalg.norm(got - ref) / nr if nr else np.linalg.norm(got)


def analysis_bin(steps, dt, f_hz):
    """frequency_analysis.c:362-390: first argmin over i < N/2 of |i fs/N - f|."""
    N = len(steps)
    T = steps[-1] * dt - steps[0] * dt
    fs = 1.0 / (T / (N - 1))
    best, k = 1e100, -1
    for i in range(N // 2):
        d = abs(i * fs / N - f_hz)
        if d < best:
            best, k = d, i
    return k, fs


def source_freq(spec):
    # frequency_analysis.c:103-105 takes the last source's frequency
    srcs = spec.get("sources", [])
    return srcs[-1]["frequency"] if srcs else spec["grid"]["f_source"]


@pytest.mark.parametrize("name", golden_cases.CASES)
def test_dft_matches_fft_of_reference_slices(name):
    c = golden_cases.load(name)
    spec, out, g = c["spec"], c["spec"]["output"], c["spec"]["grid"]
    steps = sorted(c["slices"])
    N = len(steps)
    assert steps == [n * out["interval"] for n in range(N)]
    f = source_freq(spec)
    k, fs = analysis_bin(steps, g["dt"], f)
    series = np.stack([c["slices"][t].astype(np.float64) for t in steps])
    x_ref = np.fft.fft(series, axis=0)[k] / N
    print("case", name, "N", N, "bin", k, flush=True)
    sim = Simulation(spec, c["medium"], _abi.FP64)
    kk = sim.set_dft(out["component"], out["axis"], out["pos"], out["interval"], N, f, ref_compat=True)
    assert kk == k
    sim.run(g["tsteps"])
    x, ns = sim.dft()
    sim.close()
    assert ns == N
    assert x.shape == x_ref.shape
    assert rel_l2(x, x_ref) <= 1e-6, rel_l2(x, x_ref)


@pytest.mark.parametrize("name", ["hbn", "moo3", "ade"])
def test_dft_matches_fft_of_own_slices(name):
    c = golden_cases.load(name)
    spec, out, g = c["spec"], c["spec"]["output"], c["spec"]["grid"]
    interval = 3
    N = g["tsteps"] // interval
    steps = [n * interval for n in range(N)]
    f = source_freq(spec)
    k0, fs = analysis_bin(steps, g["dt"], f)
    for k in sorted({k0, N // 2 - 1}):
        print("case", name, "bin", k, flush=True)
        sim = Simulation(spec, c["medium"], _abi.FP64)
        assert sim.set_dft(out["component"], out["axis"], out["pos"], interval, N, k * fs / N) == k
        series = []
        for t in range(g["tsteps"]):
            sim.run(1)
            if t % interval == 0 and len(series) < N:
                series.append(sim.slice(out["component"], out["axis"], out["pos"]).astype(np.float64))
        x, ns = sim.dft()
        sim.close()
        assert ns == N
        x_ref = np.fft.fft(np.stack(series), axis=0)[k] / N
        assert rel_l2(x, x_ref) <= 1e-12, (k, rel_l2(x, x_ref))


def test_dft_precision_form_error():
    c = golden_cases.load("hbn")
    spec, out, g = c["spec"], c["spec"]["output"], c["spec"]["grid"]
    N = g["tsteps"] // 2
    f = source_freq(spec)
    res = {}
    for prec in (_abi.FP64, _abi.FP32):
        sim = Simulation(spec, c["medium"], prec)
        sim.set_dft("Ez", "z", out["pos"], 2, N, f, ref_compat=False)
        sim.run(g["tsteps"])
        res[prec] = sim.dft()[0]
        sim.close()
    amp = rel_l2(np.abs(res[_abi.FP32]), np.abs(res[_abi.FP64]))
    print("fp32 amplitude rel L2", amp, flush=True)
    assert amp <= 1e-3


def test_dft_errors():
    c = golden_cases.load("hbn")
    sim = Simulation(c["spec"], c["medium"], _abi.FP64)
    with pytest.raises(_abi.PhpFdtdError) as e:
        sim._dft_shape = (1,)
        sim.dft()
    assert e.value.code == _abi.ESTATE
    for bad in [dict(interval=0), dict(nsamples=1)]:
        kw = dict(interval=2, nsamples=10)
        kw.update(bad)
        with pytest.raises(_abi.PhpFdtdError):
            sim.set_dft("Ez", "z", 3, kw["interval"], kw["nsamples"], 1e13)
    with pytest.raises(_abi.PhpFdtdError):
        sim.set_dft("Ez", "q", 3, 2, 10, 1e13)
    with pytest.raises(_abi.PhpFdtdError):
        sim.set_dft("Ez", "z", 99, 2, 10, 1e13)
    sim.close()
