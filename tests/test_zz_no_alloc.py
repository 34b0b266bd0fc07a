"""d2q9_mass_fp64_static (no workspace, no allocation) against
d2q9_mass_fp64 with a caller workspace, through the C ABI: the same
kernels, so the FP64 sums must agree bit for bit.  Two calls in a row, the
second on a smaller lattice, then the first lattice again (the static
scratch keeps the previous call's partial sums, which must not leak into
the next result)."""
import ctypes

import numpy as np
import pytest
import torch

from pylabolt_amd import cases

pytestmark = pytest.mark.gpu


def _engine(sim, storage):
    from pylabolt_amd.solvers.fluidLB import Solver
    s = Solver(simulation=sim, storage=storage, verbose=False)
    s.set_backend()
    s.advance(7)
    return s.engine


def _both(eng):
    from pylabolt_amd import _native as nt
    lib = nt.lib()
    work = torch.full((6 * 2048,), float("nan"), dtype=torch.float64,
                      device=eng.device)
    out = torch.zeros(2, dtype=torch.float64, device=eng.device)
    f = eng.f[eng.i_pop]
    nt.check(lib.d2q9_mass_fp64(eng._Lp, nt.ptr(f), nt.ptr(work),
                                ctypes.c_void_p(out.data_ptr()),
                                eng._stream_ptr), "d2q9_mass_fp64")
    nt.check(lib.d2q9_mass_fp64_static(
        eng._Lp, nt.ptr(f), ctypes.c_void_p(out.data_ptr() + 8),
        eng._stream_ptr), "d2q9_mass_fp64_static")
    eng.synchronize()
    a, b = out.cpu().numpy()
    return float(a), float(b), float(eng.pop().sum())


def test_static_mass_matches_workspace_mass():
    big = _engine(cases.cylinder(300, 200, diameter=20, xc=60), "f64")
    small = _engine(cases.lid_driven_cavity(40, 30), "f32")
    for label, eng in (("cylinder 300x200 f64", big),
                       ("cavity 40x30 f32", small),
                       ("cylinder 300x200 f64 again", big)):
        a, b, host = _both(eng)
        print(f"[no-alloc] {label}: workspace {a!r} static {b!r} host {host!r}",
              flush=True)
        assert np.isfinite(a) and a == b, (label, a, b)
        assert abs(a - host) <= 1e-11 * abs(host)


def test_static_mass_rejects_bad_descriptor():
    from pylabolt_amd import _native as nt
    lib = nt.lib()
    out = torch.zeros(1, dtype=torch.float64, device="cuda")
    rc = lib.d2q9_mass_fp64_static(None, None, ctypes.c_void_p(out.data_ptr()),
                                   None)
    assert rc == -1
    assert b"null lattice" in lib.d2q9_last_error()
