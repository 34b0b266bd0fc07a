"""Host side of the batched sweep (CPU only): work items cover every tile
and fix-up cell of every lattice exactly once, the XCD queues partition the
item list by whole lattices, and the seeded batch workload has the stated
shape."""
from types import SimpleNamespace

import numpy as np
import pytest

from pylabolt_amd import cases
from pylabolt_amd.base.state import State
from pylabolt_amd.batch import N_XCD, pack_queues, work_items
from pylabolt_amd.engine import CaseTables, elements_from_boundary


def _fake_engine(sim, B, tile_i):
    from pylabolt_amd.engine import geometry
    st = State(sim)
    nx, ny = (int(v) for v in sim.mesh_dict["grid"])
    t = CaseTables(nx, ny, st.fields.solid, elements_from_boundary(st.boundary),
                   st.boundary.x_periodic, st.boundary.y_periodic, B, tile_i)
    return SimpleNamespace(nx=nx, ny=ny, tables=t, pitch=geometry(nx, ny)[0])


@pytest.mark.parametrize("tile_i", [1, 2, 4])
def test_items_cover_tiles_and_fix_cells_once(tile_i):
    sims = cases.batch_cases(n=12, seed=3, nx_range=(64, 256),
                             ny_range=(64, 320))
    engines = [_fake_engine(s, 128, tile_i) for s in sims]
    per = work_items(engines, tile_i)
    items, bounds, owner = pack_queues(engines, per)
    assert bounds[0] == 0 and bounds[-1] == items.shape[0]
    assert np.all(np.diff(bounds) >= 0) and len(bounds) == N_XCD + 1
    # the pieces partition every lattice's item list
    for l in range(len(engines)):
        spans = sorted((a, b) for q in owner for (m, a, b) in q if m == l)
        assert spans[0][0] == 0 and spans[-1][1] == per[l].shape[0]
        assert all(spans[k][1] == spans[k + 1][0]
                   for k in range(len(spans) - 1))
    for q in range(N_XCD):
        lat = set(items[bounds[q]:bounds[q + 1], 0].tolist())
        assert lat == {m for (m, _, _) in owner[q]}
    for l, e in enumerate(engines):
        t = e.tables
        codes = items[items[:, 0] == l, 1]
        tiles = np.sort(codes[codes >= 0])
        assert np.array_equal(tiles, np.arange(t.n_bulk + t.n_general))
        chunks = np.sort(-codes[codes < 0] - 1)
        nchunk = -(-t.n_fix // (64 * tile_i))
        assert np.array_equal(chunks, np.arange(nchunk))
        # the tile list covers every tile of the lattice once
        assert np.array_equal(np.sort(t.tile_list),
                              np.arange(t.nti * t.ntj))
        # tiles in row-band order; a fix-up chunk follows the band of the
        # row after its last row (or comes at the end)
        own = per[l][:, 1]
        bands = [int(t.tile_list[c]) // t.ntj for c in own if c >= 0]
        assert bands == sorted(bands)


@pytest.mark.parametrize("sizes", [
    [5000],                                  # one large lattice
    [40, 7, 300, 1200, 64, 64, 9, 500, 33],  # uneven
    [100] * 64,                              # many equal
])
def test_queues_balanced(sizes):
    engines = [SimpleNamespace(nx=1, ny=n, tables=None) for n in sizes]
    per = [np.stack([np.full(n, l), np.arange(n)], axis=1).astype(np.int32)
           for l, n in enumerate(sizes)]
    items, bounds, owner = pack_queues(engines, per)
    load = np.diff(bounds)
    total = sum(sizes)
    cap = -(-total // (N_XCD * 4))
    # LPT over pieces of at most cap items: spread at most one piece
    assert load.max() - load.min() <= cap
    assert load.sum() == total
    # every item exactly once
    got = sorted(map(tuple, items.tolist()))
    assert got == sorted((l, k) for l, n in enumerate(sizes)
                         for k in range(n))


def test_batch_workload_shape():
    sims = cases.batch_cases()
    assert len(sims) == 256
    nx = np.array([s.mesh_dict["grid"][0] for s in sims])
    ny = np.array([s.mesh_dict["grid"][1] for s in sims])
    assert nx.min() >= 64 and nx.max() <= 1024 and np.all(nx % 32 == 0)
    assert ny.min() >= 64 and ny.max() <= 512 and np.all(ny % 32 == 0)
    for k, s in enumerate(sims):
        c = s.collision_dict["fluid"]
        if k % 2 == 0:
            assert c["model"] == "BGK" and c["forcing_model"] is None
            assert "lid" in s.boundary_dict
        else:
            assert c["model"] == "MRT"
            assert c["forcing_model"] == "guo_second_order"
            assert s.forcing_dict["gravity"][0] > 0
    # same seed, same batch
    again = cases.batch_cases()
    assert [s.mesh_dict["grid"] for s in again] == [s.mesh_dict["grid"]
                                                    for s in sims]
