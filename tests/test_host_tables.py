"""Host-built cell tables: flags, boundary rules, tile classes, fix-up list
and bulk masks (no GPU)."""
import numpy as np
import pytest

from pylabolt_amd import cases
from pylabolt_amd import engine as E
from pylabolt_amd.base.state import State

CASES = {
    "cylinder": lambda: cases.cylinder(200, 300, diameter=30, xc=60),
    "cavity": lambda: cases.lid_driven_cavity(70, 140),
    "channel": lambda: cases.poiseuille_channel(33, 260),
}


def _tables(name, B):
    st = State(CASES[name]())
    nx, ny = (int(v) for v in st.mesh.grid_global_shape)
    t = E.CaseTables(nx, ny, st.fields.solid,
                     E.elements_from_boundary(st.boundary),
                     st.boundary.x_periodic, st.boundary.y_periodic, B, 4)
    return st, t, nx, ny


@pytest.mark.parametrize("name", sorted(CASES))
@pytest.mark.parametrize("B", [64, 128, 256])
def test_bulk_masks_match_fix_cells(name, B):
    """Masked bits inside the lattice are exactly the fix-up cells; every
    cell past the lattice's last column or row is masked."""
    st, t, nx, ny = _tables(name, B)
    pitch, _ = E.geometry(nx, ny)
    fix = set(int(c) for c in t.fix_cells[:t.n_fix])
    seen = set()
    ntj = t.ntj
    masks = t.bulk_masks.reshape(-1, 4, B // 64) if t.n_bulk else None
    for p in range(t.n_bulk):
        tid = int(t.tile_list[p])
        tj, ti = tid % ntj, tid // ntj
        for r in range(4):
            for w in range(B // 64):
                word = int(masks[p, r, w])
                for b in range(64):
                    i = 1 + ti * 4 + r
                    j = 1 + tj * B + 64 * w + b
                    inside = i <= nx and j <= ny
                    if not inside:
                        assert (word >> b) & 1, (p, r, w, b)
                    elif (word >> b) & 1:
                        seen.add(E.ORIGIN + i * pitch + j)
    assert seen == fix


@pytest.mark.parametrize("name", sorted(CASES))
def test_tile_classes_partition(name):
    st, t, nx, ny = _tables(name, 64)
    assert t.n_bulk + t.n_general == t.ntj * t.nti
    assert sorted(t.tile_list.tolist()) == list(range(t.ntj * t.nti))
    fluid = t.flags_pad[1:-1, 1:-1] & 1
    solid = (t.flags_pad[1:-1, 1:-1] & 2) != 0
    for p in range(t.n_bulk):
        tid = int(t.tile_list[p])
        tj, ti = tid % t.ntj, tid // t.ntj
        # the part of the tile inside the lattice is all fluid
        blk = fluid[ti * 4:ti * 4 + 4, tj * 64:tj * 64 + 64]
        assert blk.size and blk.all()
    for p in range(t.n_bulk, t.n_bulk + t.n_general):
        tid = int(t.tile_list[p])
        tj, ti = tid % t.ntj, tid // t.ntj
        assert solid[ti * 4:ti * 4 + 4, tj * 64:tj * 64 + 64].any()


def test_cavity_corner_rules_last_element_wins():
    """The lid is listed last: its moving-wall rule owns directions 4,7,8 of
    the top corners, the side walls keep their own directions."""
    st, t, nx, ny = _tables("cavity", 64)
    from pylabolt_amd import _native as nt
    top_left = (1, ny)
    slot = t.bc_slot_pad[top_left]
    rules = t.bc_nodes[slot, :9]
    kinds = [int(r) & 15 for r in rules]
    assert kinds[4] == nt.RULE_VEL and kinds[7] == nt.RULE_VEL
    assert kinds[8] == nt.RULE_VEL       # lid wins over the left wall
    assert kinds[1] == nt.RULE_BB and kinds[5] == nt.RULE_BB


@pytest.mark.parametrize("name", ["cylinder", "cavity", "channel"])
@pytest.mark.parametrize("B", [64, 128, 256])
def test_wall_rows_swept_inline(name, B):
    """Straight-wall cells whose only rules are one wall's (zero-gradient
    rows of the cylinder, bounce-back walls of the cavity and channel) are
    swept inside their bulk tiles (wall bits), the other special cells go
    to the fix-up list; together they are all special cells of bulk tiles.
    Corners, the moving lid, stored-moment and periodic-edge cells stay in
    the list."""
    st, t, nx, ny = _tables(name, B)
    pitch, _ = E.geometry(nx, ny)
    fix = set(int(c) for c in t.fix_cells[:t.n_fix])
    inner = t.flags_pad[1:-1, 1:-1]
    fluid = (inner & 1) != 0
    special = (inner != 1) & fluid
    if t.x_periodic:
        special[0, :] |= fluid[0, :]
        special[-1, :] |= fluid[-1, :]
    in_bulk = E.in_bulk_cells(t.tile_class.reshape(t.nti, t.ntj) == 0,
                              t.tile_i, t.tile_j, nx, ny)
    wall = set()
    zm = t.wall_masks.reshape(-1, 4, 6, B // 64)
    assert t.wall_index.size == t.n_bulk
    assert zm.shape[0] == int((t.wall_index >= 0).sum())
    for p in range(t.n_bulk):
        tid = int(t.tile_list[p])
        tj, ti = tid % t.ntj, tid // t.ntj
        slot = int(t.wall_index[p])
        if slot < 0:
            continue
        for r in range(4):
            for s in range(6):
                for w in range(B // 64):
                    word = int(zm[slot, r, s, w])
                    for b in range(64):
                        if (word >> b) & 1:
                            i = ti * 4 + r          # interior index
                            j = tj * B + 64 * w + b
                            assert t.wall_side_pad[i + 1, j + 1] == s
                            # the side matches the wall the cell lies on
                            edge = {0: j == 0, 1: j == ny - 1, 2: j == 0,
                                    3: j == ny - 1, 4: i == 0,
                                    5: i == nx - 1}[s]
                            assert edge, (s, i, j)
                            wall.add(E.ORIGIN + (i + 1) * pitch + (j + 1))
    want = set(E.ORIGIN + (i + 1) * pitch + (j + 1)
               for i, j in zip(*np.nonzero(special & in_bulk)))
    assert not (wall & fix)
    assert wall | fix == want
    assert len(wall) == t.n_wall and len(wall) > 0
    if name == "cavity":
        # the lid (moving wall) and the corners stay in the list
        lid = set(E.ORIGIN + (i + 1) * pitch + ny for i in range(nx))
        assert lid <= fix


@pytest.mark.parametrize("B", [64, 128])
def test_solid_tiles_as_fix_cells(B):
    """"cells" mode lists no general tile: the fix-up list gains every fluid
    cell of the tiles holding a solid, the bulk tiles and their masks are
    unchanged; "auto" takes "cells" while at most 1/32 of the tiles are
    general."""
    st = State(CASES["cylinder"]())
    nx, ny = (int(v) for v in st.mesh.grid_global_shape)
    args = (nx, ny, st.fields.solid, E.elements_from_boundary(st.boundary),
            st.boundary.x_periodic, st.boundary.y_periodic, B, 4)
    sweep = E.CaseTables(*args, solid_tiles="sweep")
    cells = E.CaseTables(*args, solid_tiles="cells")
    assert sweep.n_general > 0 and cells.n_general == 0
    assert cells.n_bulk == sweep.n_bulk
    assert np.array_equal(cells.tile_list, sweep.tile_list[:sweep.n_bulk])
    assert np.array_equal(cells.bulk_masks, sweep.bulk_masks)
    pitch, _ = E.geometry(nx, ny)
    general = np.zeros((nx, ny), dtype=bool)
    for tid in sweep.tile_list[sweep.n_bulk:]:
        tj, ti = int(tid) % sweep.ntj, int(tid) // sweep.ntj
        general[ti * 4:ti * 4 + 4, tj * B:tj * B + B] = True
    fluid = (sweep.flags_pad[1:-1, 1:-1] & 1) != 0
    ii, jj = np.nonzero(general & fluid)
    extra = set((E.ORIGIN + (ii + 1) * pitch + (jj + 1)).tolist())
    before = set(sweep.fix_cells[:sweep.n_fix].tolist())
    after = cells.fix_cells[:cells.n_fix]
    assert np.all(np.diff(after) > 0)          # sorted, no duplicates
    assert set(after.tolist()) == before | extra
    assert not (before & extra)
    auto = E.CaseTables(*args, solid_tiles="auto")
    n_tiles = sweep.n_bulk + sweep.n_general
    want = "cells" if 32 * sweep.n_general <= n_tiles else "sweep"
    assert auto.solid_tiles == want
    with pytest.raises(ValueError):
        E.CaseTables(*args, solid_tiles="tiles")


DEEP_CASES = {
    "cylinder": lambda: cases.cylinder(120, 300, diameter=20, xc=40),
    "cylinder_walls_both": lambda: cases.cylinder(100, 256, diameter=16,
                                                  xc=50),
    "cavity": lambda: cases.lid_driven_cavity(100, 200),
    "channel": lambda: cases.poiseuille_channel(80, 256, model="BGK"),
    "channel_rows_multiple": lambda: cases.poiseuille_channel(72, 192),
    "channel_rows_odd": lambda: cases.poiseuille_channel(77, 130),
    "periodic_box": lambda: _periodic_box(),
}


def _periodic_box(nx=60, ny=300):
    box = cases.poiseuille_channel(nx, ny, model="BGK", forcing="guo_linear")
    box.boundary_dict = {
        "options": {},
        "x": {"wall": False,
              "segments": [[[0, 0], [0, ny - 1]], [[nx - 1, 0], [nx - 1, ny - 1]]],
              "fluid": {"type": "periodic"}},
        "y": {"wall": False,
              "segments": [[[0, 0], [nx - 1, 0]], [[0, ny - 1], [nx - 1, ny - 1]]],
              "fluid": {"type": "periodic"}},
    }
    return box


@pytest.mark.parametrize("name", sorted(DEEP_CASES))
def test_deep_split_covers_and_feeds_the_rim(name):
    """Deep tiles of the two-step sweep: block-aligned, plain fluid one cell
    around but for a flagged straight wall column (zero-gradient or
    bounce-back wall cells swept inline), disjoint but for the last row of
    tiles of a lattice periodic in x (moved up to end at its last row, as
    the kernel moves it; rows past the lattice's ends are their periodic
    images); every other cell lies in a listed rim tile; the first rim step
    (list A) computes every cell the second (list B) pulls from."""
    from pylabolt_amd import _native as nt
    st = State(DEEP_CASES[name]())
    nx, ny = (int(v) for v in st.mesh.grid_global_shape)
    t = E.CaseTables(nx, ny, st.fields.solid,
                     E.elements_from_boundary(st.boundary),
                     st.boundary.x_periodic, st.boundary.y_periodic, 64, 4)
    rows = 12
    d = t.deep_split(rows)
    assert d.n_deep > 0
    inner = t.flags_pad[1:-1, 1:-1]
    fluid = (inner & nt.CELL_FLUID) != 0
    plain = inner == nt.CELL_FLUID
    wrap = t.x_periodic and nx >= 2 * rows
    if t.x_periodic and not wrap:
        plain[[0, -1], :] = False
    if t.y_periodic:
        plain[:, [0, -1]] = False
    ntj = -(-ny // 64)
    cov = np.zeros((nx, ny), dtype=int)
    side = t.wall_side_pad[1:-1, 1:-1]
    inl = t.wall_inline_wrap if wrap else t.wall_inline
    for tid, wf in zip(d.deep_ids[:d.n_deep], d.deep_walls[:d.n_deep]):
        ti, tj = divmod(int(tid), ntj)
        lo, hi = int(wf) & 3, (int(wf) >> 2) & 3
        a, b = ti * rows, tj * 64
        if a + rows > nx:                    # moved up, as the kernel does
            a = nx - rows
            assert wrap and nx % 4 == 0
        if wrap:
            assert 0 <= a and a + rows <= nx
        else:
            assert a + rows <= nx - 1 and a >= 1
        grown = np.arange(a - 1, a + rows + 1) % nx   # periodic images
        c0 = b - 1 if not lo else b + 1      # the wall column apart
        c1 = b + 64 if not hi else b + 62
        if lo:
            assert b == 0
            # the wall column: inline wall cells of the flagged kind
            want = 0 if lo == 1 else 2
            assert inl[grown, 0].all()
            assert (side[grown, 0] == want).all()
        if hi:
            assert b + 64 == ny
            want = 1 if hi == 1 else 3
            assert inl[grown, ny - 1].all()
            assert (side[grown, ny - 1] == want).all()
        assert c0 >= 0 and c1 <= ny - 1
        assert plain[grown][:, c0:c1 + 1].all()
        cov[a:a + rows, b:b + 64] += 1
    assert cov.max() <= (2 if wrap else 1)
    assert np.array_equal(cov >= 1, d.covered)
    if wrap and nx % 4 == 0 and ny % 64 == 0 and name.startswith("channel"):
        # a channel periodic in x with straight walls at y: no rim at all
        assert d.covered.all()
        assert d.n_bulk + d.n_general == 0 and d.n_fix_a == 0 and d.n_fix_b == 0
    # every table handed to the library is a real (non-empty) array, also
    # for an empty rim
    for a in (d.deep_ids, d.deep_walls, d.tile_list, d.bulk_masks, d.fix_a,
              d.fix_b) + ((d.wall_index,) if d.wall_index is not None else ()):
        assert a.size >= 1
    # rim tiles: exactly the tiles not inside a deep tile
    listed = np.zeros(t.nti * ntj, dtype=int)
    for tid in d.tile_list[:d.n_bulk + d.n_general]:
        listed[int(tid)] += 1
    cpad = np.zeros((t.nti * 4, ntj * 64), dtype=bool)
    cpad[:nx, :ny] = d.covered
    inside = cpad.reshape(t.nti, 4, ntj, 64).all(axis=(1, 3)).ravel()
    assert np.array_equal(listed, (~inside).astype(int))
    # fix-up lists: B = the full list outside the deep tiles; A adds the deep
    # cells that rim cells pull from
    pitch, _ = E.geometry(nx, ny)

    def ij(c):
        c = np.asarray(c, dtype=np.int64) - E.ORIGIN
        return c // pitch - 1, c % pitch - 1

    full = set(int(c) for c in t.fix_cells[:t.n_fix])
    fb = set(int(c) for c in d.fix_b[:d.n_fix_b])
    fi, fj = ij(sorted(full))
    assert fb == set(c for c, i, j in zip(sorted(full), fi, fj)
                     if not d.covered[i, j])
    fa = set(int(c) for c in d.fix_a[:d.n_fix_a])
    assert fb <= fa
    ai, aj = ij(sorted(fa - fb))
    assert d.covered[ai, aj].all()
    cx = [0, 1, 0, -1, 0, 1, -1, -1, 1]
    cy = [0, 0, 1, 0, -1, 1, 1, -1, -1]
    ri, rj = np.nonzero(fluid & ~d.covered)
    for k in range(1, 9):
        si, sj = ri - cx[k], rj - cy[k]
        if t.x_periodic:
            si %= nx
        if t.y_periodic:
            sj %= ny
        ok = (si >= 0) & (si < nx) & (sj >= 0) & (sj < ny)
        si, sj = si[ok], sj[ok]
        src_deep = d.covered[si, sj]
        c = E.ORIGIN + (si[src_deep] + 1) * pitch + (sj[src_deep] + 1)
        assert set(c.tolist()) <= fa, k
