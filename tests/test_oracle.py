"""Pins of the oracle (CPU only).

The oracle restates the reference algorithm; these tests check it against
what the reference's own test suite fixes and against what the mathematics
fixes, chosen so that plausible mistakes (a dropped term, a wrong sign or
index, a transposed operand, a wrong corner order) fail at least one:

* constants and the D2Q9 set: reference tests/unit/test_lattice.py:94-135;
* periodic wrap of edges and corners: tests/unit/test_mpi_operator.py
  :639-1397 (edge = periodic image, corner = diagonal image when both
  directions wrap, the ghost row's own value otherwise);
* equilibrium moments, omega = 1 / omega = 0 collisions, rest state;
* pure streaming as an exact shift (omega = 0, doubly periodic);
* mass conservation of the closed cavity (bounce-back + moving lid);
* MRT with S = omega * I equals BGK (1e-14);
* Guo-forced plane Poiseuille start-up against the transient series;
* the lid-driven cavity at Re = 100 against Ghia, Ghia & Shin (1982);
* NumPy and C++ restatements agree bit for bit.
Planted mistakes (in a copy of the oracle, on the CPU) show the pins bite.
"""
import numpy as np
import pytest

from oracle import d2q9_oracle as O
from pylabolt_amd import cases
from tests.helpers import oracle_case


# reference tests/unit/test_lattice.py:94-135
REF_CX = np.array([0, 1, 0, -1, 0, 1, -1, -1, 1])
REF_CY = np.array([0, 0, 1, 0, -1, 1, 1, -1, -1])
REF_INV = np.array([0, 3, 4, 1, 2, 7, 8, 5, 6])


def test_lattice_constants_match_reference_fixture():
    c = O.lattice_constants()
    assert list(O.CX) == list(REF_CX)
    assert list(O.CY) == list(REF_CY)
    assert list(O.INV) == list(REF_INV)
    assert np.all(c["w"] == np.array([4/9, 1/9, 1/9, 1/9, 1/9,
                                      1/36, 1/36, 1/36, 1/36]))
    cs = np.float64(1 / np.sqrt(3))
    assert c["inv_cs_2"] == 1 / (cs * cs)
    # the reference's values in double, not 3 and 9
    assert c["inv_cs_2"] == 2.999999999999999
    assert c["inv_cs_4"] == 8.999999999999995
    assert c["float_min"] == np.finfo(np.float64).eps
    for k in range(9):
        assert O.CX[O.INV[k]] == -O.CX[k] and O.CY[O.INV[k]] == -O.CY[k]


def _wrap_case(xp, yp, nx=5, ny=4):
    sim = cases.lid_driven_cavity(nx, ny)
    case = oracle_case(sim)
    case.x_periodic, case.y_periodic = xp, yp
    rng = np.random.default_rng(0)
    case.pop[:] = rng.random(case.pop.shape)
    return case


def _check_wrap(wrap, xp, yp):
    case = _wrap_case(xp, yp)
    nxp, nyp = case.shape
    before = case.pop.reshape(nxp, nyp, 9).copy()
    wrap(case)
    after = case.pop.reshape(nxp, nyp, 9)
    nx, ny = nxp - 2, nyp - 2
    inner = slice(1, nyp - 1)
    if xp:
        assert np.array_equal(after[0, inner], before[nx, inner])
        assert np.array_equal(after[nx + 1, inner], before[1, inner])
    else:
        assert np.array_equal(after[0, inner], before[0, inner])
        assert np.array_equal(after[nx + 1, inner], before[nx + 1, inner])
    if yp:
        assert np.array_equal(after[1:nx + 1, 0], before[1:nx + 1, ny])
        assert np.array_equal(after[1:nx + 1, ny + 1], before[1:nx + 1, 1])
    corners = {(0, 0): (nx, ny), (0, ny + 1): (nx, 1),
               (nx + 1, 0): (1, ny), (nx + 1, ny + 1): (1, 1)}
    for (ci, cj), (di, dj) in corners.items():
        if xp and yp:
            expect = before[di, dj]            # diagonal image
        elif xp:
            expect = before[di, cj]            # column copy of the ghost row
        elif yp:
            expect = before[ci, dj]
        else:
            expect = before[ci, cj]
        assert np.array_equal(after[ci, cj], expect), (ci, cj)


@pytest.mark.parametrize("xp,yp", [(True, True), (True, False),
                                   (False, True), (False, False)])
def test_periodic_wrap_edges_and_corners(xp, yp):
    """Single-domain form of reference test_mpi_operator.py:639-1397."""
    _check_wrap(O.periodic_wrap, xp, yp)


def test_equilibrium_moments_closed_form():
    c = O.lattice_constants()
    rng = np.random.default_rng(1)
    rho = 1 + 0.1 * rng.standard_normal(100)
    ux, uy = 0.05 * rng.standard_normal((2, 100))
    f = np.stack([O.equilibrium(rho, ux, uy, k, c) for k in range(9)], -1)
    cs2 = 1 / c["inv_cs_2"]
    assert np.allclose(f.sum(-1), rho, rtol=0, atol=1e-14)
    assert np.allclose(f @ REF_CX, rho * ux, atol=1e-15)
    assert np.allclose(f @ REF_CY, rho * uy, atol=1e-15)
    assert np.allclose(f @ (REF_CX * REF_CX), rho * (cs2 + ux * ux), atol=1e-14)
    assert np.allclose(f @ (REF_CX * REF_CY), rho * ux * uy, atol=1e-15)
    assert np.allclose(f @ (REF_CY * REF_CY), rho * (cs2 + uy * uy), atol=1e-14)


def _box(omega, u=(0.0, 0.0), rough=True, n=(9, 7)):
    """Doubly periodic box without walls."""
    sim = cases.poiseuille_channel(n[0], n[1], model="BGK", forcing=None)
    nx, ny = n
    sim.boundary_dict = {
        "options": {},
        "x": {"wall": False, "segments": [[[0, 0], [0, ny - 1]],
                                          [[nx - 1, 0], [nx - 1, ny - 1]]],
              "fluid": {"type": "periodic"}},
        "y": {"wall": False, "segments": [[[0, 0], [nx - 1, 0]],
                                          [[0, ny - 1], [nx - 1, ny - 1]]],
              "fluid": {"type": "periodic"}},
    }
    if rough:
        sim.initial_fields_dict["default"]["fluid"]["velocity"] = {
            "type": "func", "func": lambda i, j: (0.02 * np.sin(1.7 * i + j),
                                                  0.03 * np.cos(i - 2.3 * j))}
    case = oracle_case(sim)
    case.omega = np.float64(omega)
    return case


def test_collision_omega_one_is_equilibrium_and_zero_is_identity():
    case = _box(1.0)
    O.initialize_pop(case)
    rng = np.random.default_rng(2)
    fl = case.fluid
    case.pop_new[fl] *= 1 + 0.01 * rng.standard_normal(case.pop_new[fl].shape)
    O.compute_density(case)
    O.compute_velocity(case)
    O.bgk_collide(case)
    feq = np.stack([O.equilibrium(case.density[fl], case.velocity[fl, 0],
                                  case.velocity[fl, 1], k, case.const)
                    for k in range(9)], -1)
    assert np.array_equal(case.pop[fl], feq)
    case.omega = np.float64(0.0)
    O.bgk_collide(case)
    assert np.array_equal(case.pop[fl], case.pop_new[fl])


def test_rest_state_is_a_fixed_point():
    case = _box(0.7, rough=False)
    O.initialize_pop(case)
    w = case.const["w"]
    assert np.array_equal(case.pop_new[case.fluid], np.tile(w, (case.fluid.sum(), 1)))
    O.run(case, 5)
    assert np.allclose(case.pop_new[case.fluid], w, rtol=0, atol=4e-16)


def test_streaming_is_an_exact_periodic_shift():
    case = _box(0.0)
    O.initialize_pop(case)
    nxp, nyp = case.shape
    nx, ny = nxp - 2, nyp - 2
    start = case.pop_new.reshape(nxp, nyp, 9)[1:-1, 1:-1].copy()
    steps = 5
    O.run(case, steps)
    got = case.pop_new.reshape(nxp, nyp, 9)[1:-1, 1:-1]
    for k in range(9):
        want = np.roll(np.roll(start[:, :, k], steps * REF_CX[k], 0),
                       steps * REF_CY[k], 1)
        assert np.array_equal(got[:, :, k], want), k


def test_closed_cavity_conserves_mass():
    case = oracle_case(cases.lid_driven_cavity(20, 14))
    O.initialize_pop(case)
    m0 = case.pop_new[case.fluid].sum()
    O.run(case, 300)
    m1 = case.pop_new[case.fluid].sum()
    assert abs(m1 - m0) <= 1e-12 * m0


@pytest.mark.parametrize("forcing", [None, "guo_linear", "guo_second_order"])
def test_mrt_with_uniform_rates_equals_bgk(forcing):
    case = _box(0.8)
    case.forcing = forcing
    case.gravity = np.array([1e-4, -2e-4])
    O.initialize_pop(case)
    rng = np.random.default_rng(3)
    case.pop_new[case.fluid] *= 1 + 0.01 * rng.standard_normal(
        case.pop_new[case.fluid].shape)
    O.compute_density(case)
    O.compute_force(case)
    O.compute_velocity(case)
    O.bgk_collide(case)
    bgk = case.pop.copy()
    O.mrt_collide_general(case, rates=np.full(9, case.omega))
    assert np.max(np.abs(case.pop - bgk)) <= 1e-14


def _poiseuille_series(y, H, nu, g, t, terms=200):
    """Start-up of plane Poiseuille flow between walls at y = 0 and H."""
    u_max = g * H * H / (8 * nu)
    u = 4 * u_max * (y / H) * (1 - y / H)
    for n in range(terms):
        m = 2 * n + 1
        u -= (32 * u_max / (np.pi ** 3 * m ** 3)) * np.sin(m * np.pi * y / H) \
            * np.exp(-(m * np.pi / H) ** 2 * nu * t)
    return u


def _channel_profile(forcing, steps, ny=24, nu=1/6):
    sim = cases.poiseuille_channel(4, ny, model="BGK", forcing=forcing)
    sim.transport_dict["kin_visc"] = nu
    g = 1e-6
    sim.forcing_dict["gravity"] = [g, 0.0]
    case = oracle_case(sim)
    O.initialize_pop(case)
    O.run(case, steps)
    nxp, nyp = case.shape
    ux = case.velocity.reshape(nxp, nyp, 2)[2, 1:-1, 0]
    y = np.arange(1, ny + 1) - 0.5          # halfway walls at 0 and ny
    return ux, _poiseuille_series(y, ny, nu, g, steps), g * ny * ny / (8 * nu)


@pytest.mark.parametrize("forcing", ["guo_linear", "guo_second_order"])
def test_guo_forced_channel_follows_startup_series(forcing):
    ux, ref, umax = _channel_profile(forcing, 1500)
    # tau = 1: halfway bounce-back puts the wall at y = 0 to O(1/H^2)
    assert np.max(np.abs(ux - ref)) <= 5e-3 * umax


# Ghia, Ghia & Shin (1982), Re = 100, u on the vertical centre line.
GHIA_Y = np.array([1.0000, 0.9766, 0.9688, 0.9609, 0.9531, 0.8516, 0.7344,
                   0.6172, 0.5000, 0.4531, 0.2813, 0.1719, 0.1016, 0.0703,
                   0.0625, 0.0547, 0.0000])
GHIA_U = np.array([1.00000, 0.84123, 0.78871, 0.73722, 0.68717, 0.23151,
                   0.00332, -0.13641, -0.20581, -0.21090, -0.15662, -0.10150,
                   -0.06434, -0.04775, -0.04192, -0.03717, 0.00000])


def _cavity_centreline(ref_lib, steps, n=64, case_fn=None):
    sim = cases.lid_driven_cavity(n, n)
    case = oracle_case(sim)
    if case_fn:
        case_fn(case)
    r = ref_lib.RefRunner(case)
    r.initialize()
    r.run(steps)
    nxp, nyp = case.shape
    u = case.velocity.reshape(nxp, nyp, 2)
    # vertical centre line between columns n/2 and n/2+1 (x = 0.5)
    uc = 0.5 * (u[n // 2, 1:-1, 0] + u[n // 2 + 1, 1:-1, 0]) / 0.1
    y = (np.arange(1, n + 1) - 0.5) / n
    prof = np.interp(GHIA_Y, np.concatenate([[0], y, [1]]),
                     np.concatenate([[0], uc, [1]]))
    return prof


def test_cavity_re100_matches_ghia(ref_lib):
    prof = _cavity_centreline(ref_lib, 30000)
    err = np.max(np.abs(prof - GHIA_U))
    print("cavity 64^2 Re=100 max |u - Ghia| =", err)
    assert err <= 0.03


SMALL = {
    "cavity": lambda: cases.lid_driven_cavity(12, 9),
    "channel_mrt": lambda: cases.poiseuille_channel(10, 7),
    "channel_bgk1": lambda: cases.poiseuille_channel(9, 8, model="BGK",
                                                     forcing="guo_linear"),
    "cylinder": lambda: cases.cylinder(40, 20, diameter=8, xc=12),
}


@pytest.mark.parametrize("name", sorted(SMALL))
def test_numpy_and_cpp_restatements_bitwise(name, ref_lib):
    sim = SMALL[name]()
    if name.startswith("channel"):
        sim.forcing_dict["gravity"] = [2e-5, 3e-6]
    a = oracle_case(sim)
    b = oracle_case(sim)
    O.initialize_pop(a)
    r = ref_lib.RefRunner(b)
    r.initialize()
    O.run(a, 25)
    r.run(25, 2)
    for n in ("pop", "pop_new", "density", "velocity"):
        assert np.array_equal(getattr(a, n), getattr(b, n)), n


# ---- planted mistakes: each must be caught by one of the pins above -------

def test_planted_dropped_force_term_is_caught(monkeypatch):
    """Drop the (1 - omega/2) Guo factor: the channel leaves the series."""
    def bad(case):
        m = case.fluid
        c = case.const
        om = case.omega
        rho = case.density[m]
        ux, uy = case.velocity[m, 0], case.velocity[m, 1]
        fx, fy = case.force_field[m, 0], case.force_field[m, 1]
        pn = case.pop_new[m]
        out = np.empty_like(pn)
        for k in range(9):
            feq = O.equilibrium(rho, ux, uy, k, c)
            ft = O.guo_source(k, ux, uy, fx, fy, case.forcing, c)
            out[:, k] = (1 - om) * pn[:, k] + om * feq + ft
        case.pop[m] = out
    monkeypatch.setattr(O, "bgk_collide", bad)
    ux, ref, umax = _channel_profile("guo_linear", 1500)
    assert np.max(np.abs(ux - ref)) > 5e-3 * umax


def test_planted_wall_sign_is_caught(ref_lib):
    """Flip the lid velocity: the centre line leaves Ghia's profile."""
    def flip(case):
        for e in case.elements:
            if e["type"] == "fixed_velocity":
                e["vector"] = [-v for v in e["vector"]]
    prof = _cavity_centreline(ref_lib, 3000, n=32, case_fn=flip)
    assert np.max(np.abs(prof - GHIA_U)) > 0.03


def test_planted_transposed_velocity_set_is_caught(monkeypatch):
    """Swap cx and cy in the streaming pass: the exact shift breaks."""
    monkeypatch.setattr(O, "CX", O.CY.copy())
    with pytest.raises(AssertionError):
        test_streaming_is_an_exact_periodic_shift()


def test_planted_corner_skip_is_caught():
    """Wrap only the interior rows and columns (ghost corners left alone):
    the corner rule of the halo test fails."""
    def interior_only(case):
        nxp, nyp = case.shape
        p = case.pop.reshape(nxp, nyp, 9)
        if case.x_periodic:
            p[0, 1:-1] = p[nxp - 2, 1:-1].copy()
            p[nxp - 1, 1:-1] = p[1, 1:-1].copy()
        if case.y_periodic:
            p[1:-1, 0] = p[1:-1, nyp - 2].copy()
            p[1:-1, nyp - 1] = p[1:-1, 1].copy()
    with pytest.raises(AssertionError):
        _check_wrap(interior_only, True, True)
