"""tb3_energy_grad_dev_noalloc, the entry point that allocates nothing,
against tb3_energy_grad_dev on the same positions through the C ABI.

P3BCA cutoff 2, density 2: two calls in a row, the second with a smaller
input (the static scratch keeps the first call's contents), in FP64 and
FP32-mixed.  Bars: identical triplet counts; energy within 1e-12 (FP64) or
2e-6 (FP32-mixed) relative, per-particle dE/dx within 1e-12 (FP64) or 3e-6
(FP32-mixed) of the other call's with |dE/dx| floored at its RMS.  Both
calls sum in an order that varies from run to run, and in FP32-mixed which
particle of a triplet owns it (and so the FP32 rounding of its terms)
varies with it: each FP32-mixed result is within ~1.3e-6 (floored) of the
FP64 reference, so two of them agree to twice that, and two runs of either
entry point differ by up to ~9e-7 in dE/dx and ~8e-7 in E.  Sizes beyond the scratch and unsupported configurations
are refused on the host before any launch.
"""
import ctypes
import sys

import numpy as np
import pytest

import threebodymdalgorithms_amd as tb3
from threebodymdalgorithms_amd._abi import TB3Error
from conftest import uniform_particles

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def floored(a, b):
    num = np.linalg.norm(a - b, axis=1)
    den = np.linalg.norm(b, axis=1)
    return float((num / np.maximum(den, np.sqrt(np.mean(den ** 2)))).max())


def both(eng, P):
    x = torch.tensor(P, dtype=torch.float64, device="cuda")
    E1 = torch.zeros((), dtype=torch.float64, device="cuda")
    g1 = torch.empty_like(x)
    E2 = torch.zeros((), dtype=torch.float64, device="cuda")
    g2 = torch.empty_like(x)
    s = torch.cuda.current_stream().cuda_stream
    c1 = eng.energy_grad_dev(len(P), x.data_ptr(), 1.0, E1.data_ptr(), g1.data_ptr(), s, want_count=True)
    c2 = eng.energy_grad_dev_noalloc(len(P), x.data_ptr(), 1.0, E2.data_ptr(), g2.data_ptr(), s, want_count=True)
    torch.cuda.synchronize()
    return (c1, float(E1), g1.cpu().numpy()), (c2, float(E2), g2.cpu().numpy())


@pytest.mark.parametrize("precision,tol,etol", [("fp64", 1e-12, 1e-12), ("fp32", 3e-6, 2e-6)])
def test_noalloc_matches_energy_grad_dev(precision, tol, etol):
    n1 = 16000
    P = uniform_particles(n1, (n1 / 2.0) ** (1 / 3), 4242)
    with tb3.Engine(0) as eng:
        eng.set_particles(P)
        eng.configure("p3bca", cutoff=2.0, precision=precision)
        for n in (n1, 5000):
            say(f"noalloc {precision} n={n}")
            (c1, E1, g1), (c2, E2, g2) = both(eng, P[:n])
            assert c1 == c2 and c1 > 0
            assert E2 == pytest.approx(E1, rel=etol)
            e = floored(g2, g1)
            say(f"  count {c2} energy {E2:.10e} vs {E1:.10e} gradient {e:.2e} (floored)")
            assert e <= tol


def test_noalloc_refuses_what_it_cannot_serve():
    P = uniform_particles(300, 5.3, 7)
    x = torch.tensor(P, dtype=torch.float64, device="cuda")
    E = torch.zeros((), dtype=torch.float64, device="cuda")
    g = torch.empty_like(x)
    with tb3.Engine(0) as eng:
        eng.set_particles(P)
        eng.configure("p3bca", cutoff=2.0, precision="fp64")
        # far more particles than the scratch array holds: refused on the
        # host, before any kernel would touch the (small) buffers
        with pytest.raises(TB3Error) as ei:
            eng.energy_grad_dev_noalloc(10_000_000, x.data_ptr(), 1.0, E.data_ptr(), g.data_ptr())
        assert "bytes of scratch" in str(ei.value)
        eng.configure("p3bca", cutoff=2.0, precision="exact")
        with pytest.raises(TB3Error):
            eng.energy_grad_dev_noalloc(len(P), x.data_ptr(), 1.0, E.data_ptr(), g.data_ptr())
        eng.configure("auta", precision="fp64")
        with pytest.raises(TB3Error):
            eng.energy_grad_dev_noalloc(len(P), x.data_ptr(), 1.0, E.data_ptr(), g.data_ptr())
        # the context still serves its other calls
        eng.configure("p3bca", cutoff=2.0, precision="fp64")
        c = eng.energy_grad_dev_noalloc(len(P), x.data_ptr(), 1.0, E.data_ptr(), g.data_ptr(), want_count=True)
        assert c == eng.energy()[1]
