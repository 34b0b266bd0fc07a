"""Adjoint gradients (GPU): against torch autograd through an FP64 torch
restatement of the steps (small lattice, 1e-9 relative), and against central
finite differences on the full adjoint workload (1024 x 256, 2000 steps,
32 seeded cells near the outlet; bar 1e-6 relative).  CPU: the torch
restatement reproduces the oracle."""
import numpy as np
import pytest
import torch

from oracle import d2q9_oracle as O
from pylabolt_amd import cases
from tests import torch_reference as TR
from tests.helpers import oracle_case


def _loss_u(u, target):
    return 0.5 * ((u[-1, :, 0] - target) ** 2).sum()


@pytest.mark.parametrize("forcing", ["guo_linear", "guo_second_order"])
def test_torch_reference_matches_oracle(forcing):
    sim, F, u0, _ = cases.adjoint_channel(12, 8, steps=15, forcing=forcing)
    rng = np.random.default_rng(5)
    u0 = 1e-3 * rng.standard_normal(u0.shape)
    nx, ny = 12, 8
    Fpad = np.zeros((nx + 2, ny + 2, 2))
    Fpad[1:-1, 1:-1] = F * 1e3
    case = oracle_case(sim, force_field=Fpad.reshape(-1, 2))
    case.velocity.reshape(nx + 2, ny + 2, 2)[1:-1, 1:-1] = u0
    O.initialize_pop(case)
    O.run(case, 14)
    O.compute_density(case)
    O.compute_velocity(case)
    want = case.velocity.reshape(nx + 2, ny + 2, 2)[1:-1, 1:-1]
    om = float(O.omega_from_visc(sim.transport_dict["kin_visc"]))
    got = TR.rollout(torch.tensor(F * 1e3), torch.tensor(u0), 15, om, forcing)
    assert np.max(np.abs(got.numpy() - want)) <= 1e-15


@pytest.mark.gpu
@pytest.mark.parametrize("forcing", ["guo_linear", "guo_second_order"])
@pytest.mark.parametrize("mode", ["sqrt", "moments"])
def test_adjoint_matches_autograd(forcing, mode):
    from pylabolt_amd.adjoint import FlowRollout, LBMFlow
    nx, ny, steps = 20, 12, 50
    sim, F, u0, target = cases.adjoint_channel(nx, ny, steps=steps,
                                               forcing=forcing)
    rng = np.random.default_rng(7)
    F = F * 1e3                     # larger force: nonlinear terms matter
    u0 = 1e-3 * rng.standard_normal(u0.shape)
    target = target * 10
    om = float(O.omega_from_visc(sim.transport_dict["kin_visc"]))
    # torch autograd through the steps (CPU, FP64)
    Ft = torch.tensor(F, requires_grad=True)
    ut = torch.tensor(u0, requires_grad=True)
    lt = _loss_u(TR.rollout(Ft, ut, steps, om, forcing), torch.tensor(target))
    lt.backward()
    # hand-written adjoint on the GPU
    ro = FlowRollout(sim, steps, checkpoint=mode)
    Fg = torch.tensor(F, device="cuda", requires_grad=True)
    ug = torch.tensor(u0, device="cuda", requires_grad=True)
    u = LBMFlow.apply(Fg, ug, ro)
    lg = _loss_u(u, torch.tensor(target, device="cuda"))
    lg.backward()
    assert abs(lg.item() - lt.item()) <= 1e-12 * abs(lt.item())
    for name, a, b in (("force", Fg.grad.cpu(), Ft.grad),
                       ("u0", ug.grad.cpu(), ut.grad)):
        err = (a - b).abs().max().item() / b.abs().max().item()
        print(f"[adjoint] {forcing} {mode} d/d{name}: max rel err {err:.2e}")
        assert err <= 1e-9, name
    ro.close()


@pytest.mark.gpu
def test_adjoint_vs_finite_differences_full_workload():
    from pylabolt_amd.adjoint import FlowRollout
    sim, F, u0, target = cases.adjoint_channel()
    nx, ny = F.shape[:2]
    ro = FlowRollout(sim, 2000, checkpoint="sqrt")
    Ft = torch.tensor(F, device="cuda")
    ut = torch.tensor(u0, device="cuda")
    tt = torch.tensor(target, device="cuda")

    def loss(force):
        return _loss_u(ro.forward(force, ut), tt).item()

    u = ro.forward(Ft, ut)
    grad_u = torch.zeros_like(u)
    grad_u[-1, :, 0] = u[-1, :, 0] - tt
    gF, gu0 = ro.backward(grad_u)
    rng = np.random.default_rng(11)
    cells = [(nx - 1 - int(rng.integers(0, 12)), int(rng.integers(0, ny)))
             for _ in range(32)]
    worst = 0.0
    # fourth-order central difference; h = 5e-7 keeps the loss difference
    # far above the rounding noise of a 2000-step FP64 rollout (~1e-17)
    # while the response to one cell's force stays linear to ~1e-10
    h = 5e-7

    def at(i, j, dh):
        Fp = Ft.clone()
        Fp[i, j, 0] += dh
        return loss(Fp)

    for (i, j) in cells:
        fd = (-at(i, j, 2 * h) + 8 * at(i, j, h) - 8 * at(i, j, -h)
              + at(i, j, -2 * h)) / (12 * h)
        ad = gF[i, j, 0].item()
        rel = abs(ad - fd) / max(abs(fd), 1e-300)
        worst = max(worst, rel)
        print(f"[fd] cell ({i},{j}) adjoint {ad:.10e} fd {fd:.10e} rel {rel:.2e}")
    print(f"[fd] worst relative difference {worst:.2e}")
    assert worst <= 1e-6
    ro.close()
