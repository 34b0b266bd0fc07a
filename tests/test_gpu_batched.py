"""Batched sweep on the GPU: many lattices of uneven size, both collision
models and every boundary type advanced by one persistent launch per step.

* FP32: after N batched steps every lattice's populations, streamed
  populations and moments are bitwise equal to the same lattice advanced on
  its own (the per-lattice sweep runs the same cell code), for every queue
  policy, tile shape, grid and claim size, across HIP-graph replays.
* FP64: the batch against the oracle, bar 1e-12 max relative error.
* The autotuner leaves the state as it found it and caches its choice.
"""
import numpy as np
import pytest

from oracle import d2q9_oracle as O
from pylabolt_amd import cases
from tests.helpers import max_rel, oracle_case

pytestmark = pytest.mark.gpu


def _mixed_sims(precision="single", obstacle=True):
    sims = cases.batch_cases(n=6, seed=5, nx_range=(32, 96),
                             ny_range=(64, 320), precision=precision)
    if obstacle:
        cyl = cases.cylinder(48, 140, diameter=12, xc=16)
        if precision == "single":
            cyl.control_dict["precision"] = "single"
        sims.append(cyl)
    return sims


def _solo(sim, storage):
    from pylabolt_amd.solvers.fluidLB import Solver
    s = Solver(simulation=sim, storage=storage, verbose=False)
    s.set_backend(verbose=False)
    return s.engine


def _assert_same(a, b, label):
    pa, pb = a.pop(), b.pop()
    na, nb = a.pop_new(), b.pop_new()
    ra, ua = a.moments_host()
    rb, ub = b.moments_host()
    for name, x, y in (("pop", pa, pb), ("pop_new", na, nb),
                       ("density", ra, rb), ("velocity", ua, ub)):
        assert np.array_equal(x, y), f"{label}: {name} differs"


CONFIGS = [
    (4, "static", None, 1, 0, True),
    (4, "dynamic", None, 1, 0, True),
    (4, "dynamic", 64, 4, 6, True),
    (2, "dynamic", 24, 3, 8, True),
    (1, "static", 8, 1, 6, True),
    (2, "static", 4096, 1, 8, True),
    # no obstacle: the build without the general-tile path
    (4, "dynamic", None, 4, 0, False),
    (1, "static", 16, 1, 6, False),
]


@pytest.mark.parametrize("tile_i,policy,grid,chunk,waves,obstacle", CONFIGS)
def test_batched_fp32_bitwise_equals_per_lattice(tile_i, policy, grid, chunk,
                                                 waves, obstacle):
    from pylabolt_amd.batch import BatchConfig, BatchedLattices
    cfg = BatchConfig.from_dict({"tile_i": tile_i, "policy": policy,
                                 "grid": grid, "chunk": chunk,
                                 "waves": waves})
    sims = _mixed_sims(obstacle=obstacle)
    print(f"[batched] f32 {cfg} lattices "
          f"{[tuple(s.mesh_dict['grid']) for s in sims]}", flush=True)
    batch = BatchedLattices(sims, storage="f32", config=cfg, graph_steps=8)
    assert batch.has_general == int(any(e.tables.n_general > 0
                                        for e in batch.engines))
    assert any((e.tables.tile_class != 0).any()
               for e in batch.engines) == obstacle
    solo = [_solo(s, "f32") for s in _mixed_sims(obstacle=obstacle)]
    done = 0
    for n in (1, 1, 19, 17):
        batch.advance(n)
        for e in solo:
            e.advance(n)
        done += n
        for l, (a, b) in enumerate(zip(batch.engines, solo)):
            _assert_same(a, b, f"{cfg} lattice {l} step {done}")
    batch.close()


def test_batched_fp64_matches_oracle():
    from pylabolt_amd.batch import BatchConfig, BatchedLattices
    sims = [cases.lid_driven_cavity(12, 9),
            cases.poiseuille_channel(10, 7),
            cases.lid_driven_cavity(9, 192, model="MRT"),
            cases.poiseuille_channel(12, 128, model="BGK",
                                     forcing="guo_linear"),
            cases.cylinder(44, 130, diameter=10, xc=14)]
    for s in sims:
        if "gravity" in s.forcing_dict:
            s.forcing_dict["gravity"] = [2e-5, 3e-6]
    batch = BatchedLattices(sims, storage="f64",
                            config=BatchConfig(4, 1, None, 2), graph_steps=4)
    ocs = [oracle_case(s) for s in sims]
    for c in ocs:
        O.initialize_pop(c)
    done = 0
    for n in (1, 12, 7):
        batch.advance(n)
        for c in ocs:
            O.run(c, n)
        done += n
        for l, (e, c) in enumerate(zip(batch.engines, ocs)):
            fl = c.fluid
            rho, vel = e.moments_host()
            errs = [max_rel(e.pop()[fl], c.pop[fl]),
                    max_rel(e.pop_new()[fl], c.pop_new[fl]),
                    max_rel(rho[fl], c.density[fl]),
                    max_rel(vel[fl], c.velocity[fl])]
            print(f"[batched] f64 lattice {l} step {done}: {errs}",
                  flush=True)
            assert max(errs) <= 1e-12
    batch.close()


def test_autotuner_keeps_state_and_caches(tmp_path):
    from pylabolt_amd.batch import BatchAutotuner, BatchConfig, BatchedLattices
    path = str(tmp_path / "tune.json")
    tuner = BatchAutotuner(cache_path=path, reps=1)
    tuned = BatchedLattices(_mixed_sims(), storage="f32", tuner=tuner)
    assert len(tuner.trials) == len(tuner.candidates(tuned))
    plain = BatchedLattices(_mixed_sims(), storage="f32",
                            config=BatchConfig(4, 0, None, 1))
    # tuning before the first step, and again after some steps
    for n in (3, 10):
        tuned.advance(n)
        plain.advance(n)
        tuner.tune(tuned)
        tuned.set_config(tuner.config_for(tuned))
    tuned.advance(5)
    plain.advance(5)
    for l, (a, b) in enumerate(zip(tuned.engines, plain.engines)):
        _assert_same(a, b, f"tuned lattice {l}")
    again = BatchAutotuner(cache_path=path)
    assert again.signature(tuned) in again.cache
    cfg = again.config_for(tuned)
    assert cfg.as_dict() == BatchConfig.from_dict(
        again.cache[again.signature(tuned)]["config"]).as_dict()
    tuned.close()
    plain.close()


def test_batched_rejects_bad_arguments():
    from pylabolt_amd import _native as nt
    from pylabolt_amd.batch import BatchConfig, BatchedLattices
    with pytest.raises(ValueError, match="FP64 or FP32"):
        BatchedLattices(_mixed_sims(), storage="f16dev",
                        config=BatchConfig())
    b = BatchedLattices(_mixed_sims()[:2], storage="f32",
                        config=BatchConfig(4, 1, None, 1))
    b._sync_entries()
    lib = nt.lib()
    # grid not a multiple of 8, chunk 0, FP16 storage, waves 5: refused on
    # the host
    for args in ((0, nt.STORE_F32, 4, 12, 1, 1, 0), (0, nt.STORE_F32, 4, 16, 1, 0, 0),
                 (0, nt.STORE_F16DEV, 4, 16, 1, 1, 0),
                 (0, nt.STORE_F32, 4, 16, 1, 1, 5)):
        p, st, ti, g, pol, ch, wv = args
        rc = lib.d2q9_batched_step(nt.ptr(b.entries), nt.ptr(b.items),
                                   nt.ptr(b.bounds), nt.ptr(b.counters),
                                   p, st, ti, g, pol, ch, wv, 1, 0, b._sp)
        assert rc != 0
    b.close()
