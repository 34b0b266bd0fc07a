"""Differentiable FP64 reference of the time step, written with torch ops so
that torch.autograd gives the gradient through all steps (test
infrastructure for the hand-written adjoint).

Same step as the reference fluidLB (pylabolt/solvers/fluidLB.py:206-253)
for the cases the adjoint supports: BGK, Guo forcing (linear or second
order) with a per-cell force field, x-periodic, bounce-back walls on the
bottom and top rows (halfway, parallel/cpu/fluid_boundary_kernels.py:5-25).
Arrays are interior (nx, ny, 9); the streaming is torch.roll along x and a
zero-filled shift along y overwritten by bounce-back on the wall rows.
"""
import numpy as np
import torch

CX = torch.tensor([0, 1, 0, -1, 0, 1, -1, -1, 1], dtype=torch.float64)
CY = torch.tensor([0, 0, 1, 0, -1, 1, 1, -1, -1], dtype=torch.float64)
INV = [0, 3, 4, 1, 2, 7, 8, 5, 6]


def constants():
    cs = np.float64(1 / np.sqrt(3))
    ics2 = 1.0 / (cs * cs)
    ics4 = ics2 * ics2
    w = torch.tensor([4/9, 1/9, 1/9, 1/9, 1/9, 1/36, 1/36, 1/36, 1/36],
                     dtype=torch.float64)
    return float(ics2), float(ics4), w, float(np.finfo(np.float64).eps)


def feq(rho, ux, uy):
    ics2, ics4, w, _ = constants()
    cu = ux[..., None] * CX + uy[..., None] * CY
    u2 = (ux * ux + uy * uy)[..., None]
    return (w * rho[..., None]) * (1 + ics2 * cu + 0.5 * ics4 * cu * cu
                                   - 0.5 * ics2 * u2)


def moments(g, F):
    _, _, _, eps = constants()
    rho = g.sum(-1)
    inv = 1 / (rho + eps)
    ux = (g * CX).sum(-1) * inv + 0.5 * F[..., 0] * inv
    uy = (g * CY).sum(-1) * inv + 0.5 * F[..., 1] * inv
    return rho, ux, uy


def collide(g, F, omega, forcing):
    ics2, ics4, w, _ = constants()
    rho, ux, uy = moments(g, F)
    out = (1 - omega) * g + omega * feq(rho, ux, uy)
    if forcing == "guo_linear":
        ft = w * (CX * F[..., 0:1] + CY * F[..., 1:2]) * ics2
        out = out + (1 - 0.5 * omega) * ft
    elif forcing == "guo_second_order":
        cu = ux[..., None] * CX + uy[..., None] * CY
        cx_ = (CX - ux[..., None]) * ics2 + cu * CX * ics4
        cy_ = (CY - uy[..., None]) * ics2 + cu * CY * ics4
        ft = w * (cx_ * F[..., 0:1] + cy_ * F[..., 1:2])
        out = out + (1 - 0.5 * omega) * ft
    return out


def stream(p):
    """x periodic, y walls with halfway bounce-back."""
    nx, ny, _ = p.shape
    cols = []
    for k in range(9):
        q = torch.roll(p[..., k], shifts=int(CX[k]), dims=0)
        cy = int(CY[k])
        if cy == 1:
            q = torch.cat([torch.zeros_like(q[:, :1]), q[:, :-1]], dim=1)
        elif cy == -1:
            q = torch.cat([q[:, 1:], torch.zeros_like(q[:, :1])], dim=1)
        cols.append(q)
    g = torch.stack(cols, dim=-1)
    bottom = {2: 4, 5: 7, 6: 8}
    top = {4: 2, 7: 5, 8: 6}
    g_b = g[:, 0, :].clone()
    for d, o in bottom.items():
        g_b[:, d] = p[:, 0, o]
    g_t = g[:, -1, :].clone()
    for d, o in top.items():
        g_t[:, d] = p[:, -1, o]
    return torch.cat([g_b[:, None], g[:, 1:-1], g_t[:, None]], dim=1)


def rollout(F, u0, steps, omega, forcing, rho0=1.0):
    """Velocity of the last step after ``steps`` reference steps."""
    rho = torch.full(u0.shape[:2], rho0, dtype=torch.float64)
    g = feq(rho, u0[..., 0], u0[..., 1])
    for n in range(1, steps + 1):
        if n == steps:
            _, ux, uy = moments(g, F)
            return torch.stack([ux, uy], dim=-1)
        g = stream(collide(g, F, omega, forcing))
