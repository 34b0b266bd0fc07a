"""FP64 CPU restatement of the reference fluidLB D2Q9 time step.

TEST INFRASTRUCTURE ONLY.  This module is the checker for the HIP engine in
``pylabolt_amd``: only ``tests/``, ``__graft_entry__.smoke()`` and the CPU
baseline leg of ``bench.py`` may import it.  The product never calls it.

It restates, pass by pass and in the reference's operation order, the
multi-pass step of ``pylabolt/solvers/fluidLB.py:206-253`` over the
reference's AoS layout ``pop[size, 9]`` with the padded index
``ind = i * Ny_pad + j``:

1. density   parallel/cpu/compute_fields_kernels.py:5-28
2. force     parallel/cpu/force_field_kernels.py:5-25
3. velocity  parallel/cpu/compute_fields_kernels.py:31-65
4. collide   parallel/cpu/collision_kernels.py:5-153 (BGK x 3 forcings)
5. periodic  parallel/gpu/MPI_kernels.py:34-99 (single-rank wrap of the
             post-collision populations, x columns first, then y rows)
6. stream    parallel/cpu/streaming_kernels.py:5-47
7. boundary  parallel/cpu/fluid_boundary_kernels.py:5-62,100-152, one
             element after another in dict/segment order

NumPy applies each elementwise operation with one IEEE rounding, so writing
every expression with the reference's association (Python evaluates left to
right) gives the same bits the reference's scalar loops give without
fast-math.

Parts with no counterpart in the reference are defined here and in the HIP
engine identically (parity for them is against this definition):

* MRT: ``f* = f - M^-1 [ S (M f - M feq) - (I - S/2) M Fbar ]`` with the
  reference's moment matrix M and ``S = (1,1,1,1,1,1,1,omega,omega)``
  (base/collision_operator.py:136-165, whose own setup is broken); Fbar is
  the Guo source of the selected forcing without its (1 - omega/2) factor.
  With unit rates on rows 0..6 it reduces to ``feq + 0.5 Fbar`` plus the
  relaxed stress moments (p_xx, p_xy) of ``f - feq + 0.5 Fbar``; that
  reduced form, in the operation order ``mrt_collide`` spells out, is the
  definition the engine follows.  ``mrt_collide_general`` evaluates the
  operator through the full moment map (``pair_moments``,
  ``pair_inverse``) for any rates; the tests pin the reduced form to it.
* zero_gradient: the unknown directions of a boundary node take the values
  the streaming pass produced in the same directions at the inward
  neighbour (before any boundary element ran).
* force_source "field": the force field is a fixed per-cell input instead
  of ``rho * gravity``.
* fixed_pressure uses the integer inward normal (the reference indexes with
  the float normal, fluid_boundary_kernels.py:133-135).

Pins (tests/test_oracle.py, tests/test_oracle_pins.py): every function
here is checked against the reference's own fixtures, closed forms
(equilibrium and Guo-source moments, fixed points, exact shifts),
analytic flows (start-up Poiseuille series for BGK and MRT, plane
Poiseuille between pressure ends, Ghia's cavity) or an identity (MRT
reduced form = full moment map; zero-gradient = periodic for fields
independent of y), each with a planted mistake that fails.  The one
assumption no check here can pin is Numba's rounding: one IEEE rounding
per operation in source order, as without fastmath.

The reference's CPU single-rank halo exchange sends and receives with its
own rank (MPI_operator.py:261-369), which copies column 1 into ghost column
0 instead of wrapping; the single-GPU path and the multi-rank path wrap
periodically, and so does this restatement.
"""
import numpy as np

CX = np.array([0, 1, 0, -1, 0, 1, -1, -1, 1], dtype=np.int64)
CY = np.array([0, 0, 1, 0, -1, 1, 1, -1, -1], dtype=np.int64)
INV = np.array([0, 3, 4, 1, 2, 7, 8, 5, 6], dtype=np.int64)

# Reference moment matrix (base/collision_operator.py:147-157).
M_MRT = np.array([
    [1, 1, 1, 1, 1, 1, 1, 1, 1],
    [-4, -1, -1, -1, -1, 2, 2, 2, 2],
    [4, -2, -2, -2, -2, 1, 1, 1, 1],
    [0, 1, 0, -1, 0, 1, -1, -1, 1],
    [0, -2, 0, 2, 0, 1, -1, -1, 1],
    [0, 0, 1, 0, -1, 1, 1, -1, -1],
    [0, 0, -2, 0, 2, 1, 1, -1, -1],
    [0, 1, -1, 1, -1, 0, 0, 0, 0],
    [0, 0, 0, 0, 0, 1, -1, 1, -1],
], dtype=np.int64)
# The rows of M are orthogonal: M^-1 = M^T diag(1 / |row|^2).
MRT_ROW_NORM = np.array([9, 36, 36, 6, 12, 6, 12, 4, 4], dtype=np.int64)


def lattice_constants():
    """Constants computed as base/lattice.py:41-59, base/control.py:49."""
    cs = np.float64(1 / np.sqrt(3))
    cs_2 = cs * cs
    inv_cs_2 = 1.0 / cs_2
    inv_cs_4 = inv_cs_2 * inv_cs_2
    w = np.array([4/9, 1/9, 1/9, 1/9, 1/9, 1/36, 1/36, 1/36, 1/36],
                 dtype=np.float64)
    float_min = np.finfo(np.float64).eps
    return dict(cs=cs, cs_2=cs_2, inv_cs_2=np.float64(inv_cs_2),
                inv_cs_4=np.float64(inv_cs_4), w=w,
                float_min=np.float64(float_min))


def omega_from_visc(kin_visc):
    """base/collision_operator.py:89-91."""
    c = lattice_constants()
    tau = np.float64(kin_visc) * c["inv_cs_2"] + 0.5
    return np.float64(1 / tau)


class OracleCase:
    """Everything one run needs, as plain arrays.

    ``elements`` is the ordered list of boundary elements; each is a dict
    with keys ``type`` ("bounce_back", "fixed_velocity", "fixed_pressure",
    "zero_gradient", "periodic"), ``nodes`` (int array of flat indices),
    ``out``/``inv`` (direction lists), ``vector`` (2 floats), ``scalar``
    (float) and ``normal`` (2 ints).
    """

    def __init__(self, shape, solid, ghost, density, velocity, elements,
                 x_periodic, y_periodic, omega, model="BGK", forcing=None,
                 gravity=(0.0, 0.0), force_source="gravity",
                 force_field=None):
        self.shape = (int(shape[0]), int(shape[1]))
        self.size = self.shape[0] * self.shape[1]
        self.solid = np.asarray(solid, dtype=bool).copy()
        self.ghost = np.asarray(ghost, dtype=bool).copy()
        self.fluid = ~self.solid & ~self.ghost
        self.density = np.asarray(density, dtype=np.float64).copy()
        self.velocity = np.asarray(velocity, dtype=np.float64).copy()
        self.elements = elements
        self.x_periodic = bool(x_periodic)
        self.y_periodic = bool(y_periodic)
        self.omega = np.float64(omega)
        self.model = model
        self.forcing = forcing
        self.gravity = np.asarray(gravity, dtype=np.float64)
        self.force_source = force_source
        self.force_field = np.zeros((self.size, 2), dtype=np.float64)
        if force_field is not None:
            self.force_field[:] = force_field
        self.pop = np.zeros((self.size, 9), dtype=np.float64)
        self.pop_new = np.zeros((self.size, 9), dtype=np.float64)
        self.const = lattice_constants()
        self.time_step = 0


def equilibrium(rho, ux, uy, k, c):
    """feq_k (parallel/cpu/equilibrium_kernels.py:6-35), vectorised."""
    inv_cs_2, inv_cs_4 = c["inv_cs_2"], c["inv_cs_4"]
    u2 = ux * ux + uy * uy
    cu = CX[k] * ux + CY[k] * uy
    return c["w"][k] * rho * (
        1 + inv_cs_2 * cu + 0.5 * inv_cs_4 * cu * cu - 0.5 * inv_cs_2 * u2
    )


def initialize_pop(case):
    """pop = pop_new = feq(rho0, u0) on fluid nodes
    (parallel/cpu/equilibrium_kernels.py:38-78)."""
    m = case.fluid
    rho = case.density[m]
    ux, uy = case.velocity[m, 0], case.velocity[m, 1]
    for k in range(9):
        feq = equilibrium(rho, ux, uy, k, case.const)
        case.pop[m, k] = feq
        case.pop_new[m, k] = feq


def compute_density(case):
    m = case.fluid
    s = np.zeros(int(m.sum()), dtype=np.float64)
    pn = case.pop_new[m]
    for k in range(9):
        s += pn[:, k]
    case.density[m] = s


def compute_force(case):
    if case.force_source != "gravity":
        return
    m = case.fluid
    rho = case.density[m]
    case.force_field[m, 0] = rho * case.gravity[0]
    case.force_field[m, 1] = rho * case.gravity[1]


def compute_velocity(case):
    m = case.fluid
    pn = case.pop_new[m]
    sx = np.zeros(pn.shape[0], dtype=np.float64)
    sy = np.zeros(pn.shape[0], dtype=np.float64)
    for k in range(9):
        sx += CX[k] * pn[:, k]
        sy += CY[k] * pn[:, k]
    inv = 1 / (case.density[m] + case.const["float_min"])
    fx = case.force_field[m, 0]
    fy = case.force_field[m, 1]
    case.velocity[m, 0] = sx * inv + 0.5 * fx * inv
    case.velocity[m, 1] = sy * inv + 0.5 * fy * inv


def guo_source(k, ux, uy, fx, fy, forcing, c):
    """Guo source term without the (1 - omega/2) factor
    (collision_kernels.py:85-87 and 138-144)."""
    inv_cs_2, inv_cs_4 = c["inv_cs_2"], c["inv_cs_4"]
    if forcing == "guo_linear":
        return c["w"][k] * (CX[k] * fx + CY[k] * fy) * inv_cs_2
    cu = CX[k] * ux + CY[k] * uy
    const_x = (CX[k] - ux) * inv_cs_2 + cu * CX[k] * inv_cs_4
    const_y = (CY[k] - uy) * inv_cs_2 + cu * CY[k] * inv_cs_4
    return c["w"][k] * (const_x * fx + const_y * fy)


def bgk_collide(case):
    m = case.fluid
    c = case.const
    om = case.omega
    rho = case.density[m]
    ux, uy = case.velocity[m, 0], case.velocity[m, 1]
    fx, fy = case.force_field[m, 0], case.force_field[m, 1]
    pn = case.pop_new[m]
    out = np.empty_like(pn)
    for k in range(9):
        feq = equilibrium(rho, ux, uy, k, c)
        if case.forcing is None:
            out[:, k] = (1 - om) * pn[:, k] + om * feq
        else:
            ft = guo_source(k, ux, uy, fx, fy, case.forcing, c)
            out[:, k] = (1 - om) * pn[:, k] + om * feq + (1 - 0.5 * om) * ft
    case.pop[m] = out


# opposite pairs (k, inv k) in the order the MRT form below visits them
PAIRS = ((1, 3), (2, 4), (5, 7), (6, 8))


def pair_moments(x0, S, D):
    """Rows of M applied to a population-like vector given as its rest value
    ``x0`` and the sums ``S[n]`` and differences ``D[n]`` of the pairs
    ``PAIRS[n]`` (the columns of k and inv k agree on rows 0, 1, 2, 7, 8
    and are opposite on rows 3..6), in this operation order."""
    s12 = S[0] + S[1]
    s34 = S[2] + S[3]
    a = D[2] - D[3]
    b = D[2] + D[3]
    return [(x0 + s12) + s34,
            ((-4.0 * x0) - s12) + 2.0 * s34,
            ((4.0 * x0) + s34) - 2.0 * s12,
            D[0] + a,
            a - 2.0 * D[0],
            D[1] + b,
            b - 2.0 * D[1],
            S[0] - S[1],
            S[2] - S[3]]


def pair_inverse(gm):
    """``M^T gm`` by columns k = 0..8, in this operation order (with
    ``gm = diag(1/|row|^2) p`` it is ``M^-1 p``)."""
    out0 = gm[0] + 4.0 * (gm[2] - gm[1])
    base13 = (gm[0] - gm[1]) - 2.0 * gm[2]
    sym13 = base13 + gm[7]
    sym24 = base13 - gm[7]
    asym13 = gm[3] - 2.0 * gm[4]
    asym24 = gm[5] - 2.0 * gm[6]
    base57 = (gm[0] + gm[2]) + 2.0 * gm[1]
    sym57 = base57 + gm[8]
    sym68 = base57 - gm[8]
    p34 = gm[3] + gm[4]
    p56 = gm[5] + gm[6]
    asym57 = p34 + p56
    asym68 = p56 - p34
    return [out0, sym13 + asym13, sym24 + asym24, sym13 - asym13,
            sym24 - asym24, sym57 + asym57, sym68 + asym68, sym57 - asym57,
            sym68 - asym68]


def mrt_collide_general(case, rates=None):
    """The MRT operator with any rates, through the full moment map, in the
    opposite-pair form:

      e_k = f_k - feq_k,  Fbar_k = Guo source without (1 - omega/2)
      me = M e, mf = M Fbar         (pair_moments of the pair sums and
                                      differences e_k + e_kb, e_k - e_kb)
      p_i = (1 - s_i) me_i + (1 - 0.5 s_i) mf_i   (no forcing: first term)
      f*_k = feq_k + (M^-1 p)_k     (pair_inverse of p_i * (1/|row_i|^2))

    which is f - M^-1 [S (M f - M feq) - (I - S/2) M Fbar] in exact
    arithmetic; ``s = rates`` (default (1,1,1,1,1,1,1,omega,omega)).  The
    time step uses ``mrt_collide``, the same operator for the reference's
    rates in fewer operations; this form is the one the pins compare it
    with (and the one that can take other rates)."""
    m_ = case.fluid
    c = case.const
    om = case.omega
    if rates is None:
        rates = np.array([1, 1, 1, 1, 1, 1, 1, om, om], dtype=np.float64)
    rates = np.asarray(rates, dtype=np.float64)
    rho = case.density[m_]
    ux, uy = case.velocity[m_, 0], case.velocity[m_, 1]
    fx, fy = case.force_field[m_, 0], case.force_field[m_, 1]
    pn = case.pop_new[m_]
    forced = case.forcing is not None
    feq = [equilibrium(rho, ux, uy, k, c) for k in range(9)]
    fb = [guo_source(k, ux, uy, fx, fy, case.forcing, c) if forced else None
          for k in range(9)]
    e0 = pn[:, 0] - feq[0]
    eS, eD, fS, fD = [], [], [], []
    for k, kb in PAIRS:
        ek = pn[:, k] - feq[k]
        ekb = pn[:, kb] - feq[kb]
        eS.append(ek + ekb)
        eD.append(ek - ekb)
        if forced:
            fS.append(fb[k] + fb[kb])
            fD.append(fb[k] - fb[kb])
    me = pair_moments(e0, eS, eD)
    mf = pair_moments(fb[0], fS, fD) if forced else None
    gm = []
    for i in range(9):
        s = rates[i]
        p = (1.0 - s) * me[i]
        if forced:
            p = p + (1.0 - 0.5 * s) * mf[i]
        gm.append(p * np.float64(1.0 / MRT_ROW_NORM[i]))
    out = pair_inverse(gm)
    res = np.empty_like(pn)
    for k in range(9):
        res[:, k] = feq[k] + out[k]
    case.pop[m_] = res


def mrt_collide(case, stress_rates=None):
    """Build-defined MRT with Guo forcing and the reference's rates
    S = (1,1,1,1,1,1,1,omega,omega) (base/collision_operator.py:158-162);
    ``stress_rates`` = (s7, s8) replaces the two omegas (the engine takes
    any rates on the stress rows).

    With unit rates on rows 0..6 only the stress rows 7 (p_xx) and 8 (p_xy)
    keep a part of f - feq, and M^-1 = M^T diag(1/|row|^2) with
    |row 7|^2 = |row 8|^2 = 4, so the operator of ``mrt_collide_general``
    becomes f*_k = feq_k + 0.5 Fbar_k + sigma_k (1 - s) / 4 (p + 0.5 p^F),
    with p the stress moment of f - feq and p^F that of Fbar.  The stress
    moments of feq and Fbar have closed forms (sums over the equilibrium
    and source formulas, with the lattice's own w and inv_cs_4):
      p_xx(feq) = w1 ics4 rho (ux^2 - uy^2),  p_xy(feq) = 4 w5 ics4 rho ux uy,
      p_xx(Fbar) = 2 w1 ics4 (ux fx - uy fy),
      p_xy(Fbar) = 4 w5 ics4 (ux fy + uy fx)   (second-order Guo),
      p(Fbar) = 0                               (linear Guo, no forcing).
    In this operation order:

      q_k = feq_k + 0.5 Fbar_k            (feq_k alone without forcing)
      a7 = ((f1 + f3) - (f2 + f4)) - ((w1 ics4) rho) (ux ux - uy uy)
      a8 = ((f5 + f7) - (f6 + f8)) - (((4 w5) ics4) rho) (ux uy)
        second order:  a7 = a7 + (w1 ics4) (ux fx - uy fy)
                       a8 = a8 + ((2 w5) ics4) (ux fy + uy fx)
      r7 = ((1 - s7) * 0.25) * a7,  r8 = ((1 - s8) * 0.25) * a8
      f*_0 = q_0;  f*_k = q_k + r7 (k = 1, 3), - r7 (2, 4),
                          + r8 (5, 7), - r8 (6, 8)

    (the signs are the entries of rows 7 and 8 of M)."""
    m_ = case.fluid
    c = case.const
    om = case.omega
    rho = case.density[m_]
    ux, uy = case.velocity[m_, 0], case.velocity[m_, 1]
    fx, fy = case.force_field[m_, 0], case.force_field[m_, 1]
    pn = case.pop_new[m_]
    w1, w5, ics4 = c["w"][1], c["w"][5], c["inv_cs_4"]
    q = [equilibrium(rho, ux, uy, k, c) for k in range(9)]
    a7 = ((pn[:, 1] + pn[:, 3]) - (pn[:, 2] + pn[:, 4])) - \
        ((w1 * ics4) * rho) * (ux * ux - uy * uy)
    a8 = ((pn[:, 5] + pn[:, 7]) - (pn[:, 6] + pn[:, 8])) - \
        (((4.0 * w5) * ics4) * rho) * (ux * uy)
    if case.forcing is not None:
        if case.forcing == "guo_second_order":
            a7 = a7 + (w1 * ics4) * (ux * fx - uy * fy)
            a8 = a8 + ((2.0 * w5) * ics4) * (ux * fy + uy * fx)
        fb = [guo_source(k, ux, uy, fx, fy, case.forcing, c)
              for k in range(9)]
        q = [q[k] + 0.5 * fb[k] for k in range(9)]
    s7, s8 = (om, om) if stress_rates is None else stress_rates
    r7 = ((1.0 - np.float64(s7)) * 0.25) * a7
    r8 = ((1.0 - np.float64(s8)) * 0.25) * a8
    res = np.empty_like(pn)
    res[:, 0] = q[0]
    for k, r in ((1, r7), (3, r7), (5, r8), (7, r8)):
        res[:, k] = q[k] + r
    for k, r in ((2, r7), (4, r7), (6, r8), (8, r8)):
        res[:, k] = q[k] - r
    case.pop[m_] = res


def periodic_wrap(case):
    """Single-rank wrap of post-collision populations
    (parallel/gpu/MPI_kernels.py:34-99): columns over all j, then rows over
    all i, so ghost corners receive the diagonal image."""
    nx, ny = case.shape
    p = case.pop.reshape(nx, ny, 9)
    if case.x_periodic:
        left = p[nx - 2, :, :].copy()
        right = p[1, :, :].copy()
        p[0, :, :] = left
        p[nx - 1, :, :] = right
    if case.y_periodic:
        bottom = p[:, ny - 2, :].copy()
        top = p[:, 1, :].copy()
        p[:, 0, :] = bottom
        p[:, ny - 1, :] = top


def stream(case):
    """Pull streaming with moving-wall bounce-back at solid neighbours
    (parallel/cpu/streaming_kernels.py:5-47)."""
    c = case.const
    nx, ny = case.shape
    idx = np.flatnonzero(case.fluid)
    x = idx // ny
    y = idx - x * ny
    rho = case.density[idx]
    case.pop_new[idx, 0] = case.pop[idx, 0]
    for k in range(1, 9):
        nb = (x - CX[k]) * ny + (y - CY[k])
        sol = case.solid[nb]
        val = case.pop[nb, k].copy()
        if np.any(sol):
            ki = INV[k]
            s_idx = idx[sol]
            temp = 2 * c["w"][ki] * rho[sol] * c["inv_cs_2"] * (
                CX[ki] * case.velocity[nb[sol], 0] +
                CY[ki] * case.velocity[nb[sol], 1]
            )
            val[sol] = case.pop[s_idx, ki] - temp
        case.pop_new[idx, k] = val


def apply_boundaries(case, streamed):
    """Boundary elements in order (fluid_boundary_kernels.py:5-152).
    ``streamed`` is pop_new as the streaming pass left it (used by
    zero_gradient only)."""
    c = case.const
    nx, ny = case.shape
    for e in case.elements:
        t = e["type"]
        if t == "periodic":
            continue
        nodes = np.asarray(e["nodes"], dtype=np.int64)
        if nodes.size == 0:
            continue
        nodes = nodes[~case.solid[nodes]]
        outl, invl = e["out"], e["inv"]
        if t == "bounce_back":
            for o, v in zip(outl, invl):
                case.pop_new[nodes, v] = case.pop[nodes, o]
        elif t == "fixed_velocity":
            uw = np.asarray(e["vector"], dtype=np.float64)
            rho = case.density[nodes]
            for o, v in zip(outl, invl):
                temp = 2 * c["w"][o] * rho * c["inv_cs_2"] * (
                    CX[o] * uw[0] + CY[o] * uw[1]
                )
                case.pop_new[nodes, v] = case.pop[nodes, o] - temp
        elif t == "fixed_pressure":
            rho_b = np.float64(e["scalar"])
            nvec = e["normal"]
            nb = nodes + int(nvec[0]) * ny + int(nvec[1])
            vx = case.velocity[nodes, 0]
            vy = case.velocity[nodes, 1]
            ex = vx + 0.5 * (vx - case.velocity[nb, 0])
            ey = vy + 0.5 * (vy - case.velocity[nb, 1])
            u2 = ex * ex + ey * ey
            for o, v in zip(outl, invl):
                cu = CX[o] * ex + CY[o] * ey
                temp = 2 * c["w"][o] * rho_b * (
                    1 + 0.5 * c["inv_cs_4"] * cu * cu -
                    0.5 * c["inv_cs_2"] * u2
                )
                case.pop_new[nodes, v] = -case.pop[nodes, o] + temp
        elif t == "zero_gradient":
            nvec = e["normal"]
            nb = nodes + int(nvec[0]) * ny + int(nvec[1])
            for v in invl:
                case.pop_new[nodes, v] = streamed[nb, v]
        else:
            raise ValueError("unknown boundary type " + str(t))


def single_time_step(case):
    """One reference time step (solvers/fluidLB.py:206-253)."""
    compute_density(case)
    compute_force(case)
    compute_velocity(case)
    if case.model == "BGK":
        bgk_collide(case)
    elif case.model == "MRT":
        mrt_collide(case)
    else:
        raise ValueError("unknown collision model " + str(case.model))
    periodic_wrap(case)
    stream(case)
    needs_snapshot = any(e["type"] == "zero_gradient" for e in case.elements)
    streamed = case.pop_new.copy() if needs_snapshot else None
    apply_boundaries(case, streamed)
    case.time_step += 1


def run(case, steps):
    for _ in range(steps):
        single_time_step(case)
    return case


def case_from_state(state, model, forcing, omega, gravity,
                    force_source="gravity", force_field=None):
    """Build an ``OracleCase`` from a case-state object.

    ``state`` is duck-typed: it needs ``domain.shape``, ``fields.solid``,
    ``fields.ghost_node``, ``fields.density``, ``fields.velocity`` and
    ``boundary.{boundary_elements, x_periodic, y_periodic}`` with element
    attributes ``type_fluid, boundary_nodes, out_list, inv_list,
    vector_fluid, scalar_fluid`` and ``surface_normals`` -- the reference's
    names (base/boundary.py:7-84).
    """
    elements = []
    for e in state.boundary.boundary_elements:
        elements.append(dict(
            type=e.type_fluid,
            nodes=np.asarray(e.boundary_nodes, dtype=np.int64),
            out=[int(v) for v in e.out_list],
            inv=[int(v) for v in e.inv_list],
            vector=[float(v) for v in e.vector_fluid],
            scalar=float(e.scalar_fluid),
            normal=[int(round(float(v))) for v in e.surface_normals],
        ))
    f = state.fields
    return OracleCase(state.domain.shape, f.solid, f.ghost_node, f.density,
                      f.velocity, elements, state.boundary.x_periodic,
                      state.boundary.y_periodic, omega, model=model,
                      forcing=forcing, gravity=gravity,
                      force_source=force_source, force_field=force_field)


# ---------------------------------------------------------------------------
# Momentum exchange (forces on obstacles and walls)
# ---------------------------------------------------------------------------

def fluid_boundary_nodes(shape, solid, ghost, solid_id):
    """compute_obstacle_boundary, fluid branch
    (parallel/cpu/obstacle_kernels.py:338-348): a non-ghost fluid node with
    a solid neighbour at ind + c_k is a fluid boundary node and takes the
    solid_id of the first such neighbour.  Returns (fluid_boundary, node
    solid_id) arrays over the padded lattice."""
    nyp = shape[1]
    fb = np.zeros(shape[0] * nyp, dtype=bool)
    sid = np.asarray(solid_id).copy()
    for ind in range(shape[0] * nyp):
        if ghost[ind] or solid[ind]:
            continue
        x, y = divmod(ind, nyp)
        for k in range(9):
            nb = (x + CX[k]) * nyp + (y + CY[k])
            if solid[nb]:
                fb[ind] = True
                sid[ind] = sid[nb]
                break
    return fb, sid


def obstacle_boundary_fields(shape, solid, ghost, solid_id, circles,
                             x_periodic, y_periodic):
    """compute_obstacle_boundary (parallel/cpu/obstacle_kernels.py:307-348,
    both branches) then compute_normals_circle (:97-143) for each circle
    ``(id, center)``, sequential, single rank (offset 0).  Returns
    (solid_boundary, fluid_boundary, solid_id, surface_normals) over the
    padded lattice."""
    nxp, nyp = shape
    Nx, Ny = nxp - 2, nyp - 2
    size = nxp * nyp
    sb = np.zeros(size, dtype=bool)
    fb = np.zeros(size, dtype=bool)
    sid = np.asarray(solid_id).copy()
    for ind in range(size):
        if ghost[ind]:
            continue
        x, y = divmod(ind, nyp)
        for k in range(9):
            nb = (x + CX[k]) * nyp + (y + CY[k])
            if solid[ind] and not solid[nb]:
                sb[ind] = True
                break
            if not solid[ind] and solid[nb]:
                fb[ind] = True
                sid[ind] = sid[nb]
                break
    normals = np.zeros((size, 2), dtype=np.float64)
    for cid, center in circles:
        for ind in range(size):
            if sid[ind] != cid or not (sb[ind] or fb[ind]):
                continue
            x, y = divmod(ind, nyp)
            rx = (x - 1) - center[0]
            ry = (y - 1) - center[1]
            rxm, rym = rx, ry
            if x_periodic:
                if abs(rx + Nx) < abs(rxm):
                    rxm = rx + Nx
                if abs(rx - Nx) < abs(rxm):
                    rxm = rx - Nx
            if y_periodic:
                if abs(ry + Ny) < abs(rym):
                    rym = ry + Ny
                if abs(ry - Ny) < abs(rym):
                    rym = ry - Ny
            mag = np.sqrt(rxm * rxm + rym * rym)
            normals[ind, 0] = rxm / mag
            normals[ind, 1] = rym / mag
    return sb, fb, sid, normals


def obstacle_force_torque(case, solid_id, fluid_boundary, ref_point,
                          current_id, pop=None, pop_new=None):
    """compute_force_torque_single_phase
    (parallel/cpu/force_torque_kernels.py:11-95), sequential, with the
    single-rank offset 0: over fluid boundary nodes of obstacle
    ``current_id``, for every direction k whose neighbour is solid,
    value = pop[k] c_k - pop_new[inv k] c_{inv k}; torque about
    ``ref_point`` with the minimum image across periodic sides."""
    pop = case.pop if pop is None else pop
    pop_new = case.pop_new if pop_new is None else pop_new
    nxp, nyp = case.shape
    Nx, Ny = nxp - 2, nyp - 2
    fx = fy = tq = 0.0
    for ind in range(case.size):
        if case.ghost[ind] or not fluid_boundary[ind]:
            continue
        if solid_id[ind] != current_id:
            continue
        x, y = divmod(ind, nyp)
        rx = (x - 1) - ref_point[0]
        ry = (y - 1) - ref_point[1]
        rxm, rym = rx, ry
        if case.x_periodic:
            if abs(rx + Nx) < abs(rxm):
                rxm = rx + Nx
            if abs(rx - Nx) < abs(rxm):
                rxm = rx - Nx
        if case.y_periodic:
            if abs(ry + Ny) < abs(rym):
                rym = ry + Ny
            if abs(ry - Ny) < abs(rym):
                rym = ry - Ny
        for k in range(9):
            nb = (x + CX[k]) * nyp + (y + CY[k])
            if case.solid[nb]:
                kb = INV[k]
                vx = pop[ind, k] * CX[k] - pop_new[ind, kb] * CX[kb]
                vy = pop[ind, k] * CY[k] - pop_new[ind, kb] * CY[kb]
                fx += vx
                fy += vy
                tq += rxm * vy - rym * vx
    return np.array([fx, fy, tq])


def boundary_force(case, element, pop=None, pop_new=None):
    """compute_boundary_force_single_phase
    (parallel/cpu/force_torque_kernels.py:98-141) for one element."""
    pop = case.pop if pop is None else pop
    pop_new = case.pop_new if pop_new is None else pop_new
    fx = fy = 0.0
    for ind in element["nodes"]:
        if case.solid[ind]:
            continue
        for ko, ki in zip(element["out"], element["inv"]):
            fx += pop[ind, ko] * CX[ko] - pop_new[ind, ki] * CX[ki]
            fy += pop[ind, ko] * CY[ko] - pop_new[ind, ki] * CY[ki]
    return np.array([fx, fy])
