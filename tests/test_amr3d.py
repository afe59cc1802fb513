"""The refined 3D hierarchy of equal blocks: the oracle's invariants (CPU) and the
GPU path against the oracle (FP64: bit for bit)."""

import itertools

import numpy as np
import pytest

from oracle.amr3d import OracleAMR3D
from oracle.mhd3d import Oracle3D
from wellbalanced_amd import amr3d_maps, initial
from wellbalanced_amd.amr3d_maps import Hierarchy3D
from wellbalanced_amd.hierarchy import AMRHierarchy

TWO_PI = 2.0 * np.pi
B0 = (0.3, 0.2, 0.1)


def _state():
    return dict(
        density=lambda x, y, z: 1.0 + 0.2 * np.sin(TWO_PI * x) * np.sin(TWO_PI * y) * np.cos(TWO_PI * z),
        velocity=lambda x, y, z: (0.3 + 0.1 * np.cos(TWO_PI * y), 0.2 + 0.0 * x, 0.1 * np.sin(TWO_PI * x)),
        pressure=lambda x, y, z: 1.0 + 0.1 * np.cos(TWO_PI * z) + 0.05 * np.sin(TWO_PI * x),
        vector_potential=(lambda x, y, z: 0.05 * np.sin(TWO_PI * y) * np.cos(TWO_PI * z),
                          lambda x, y, z: 0.05 * np.sin(TWO_PI * z),
                          lambda x, y, z: 0.05 * np.sin(TWO_PI * x) + 0.02 * np.cos(TWO_PI * y)),
    )


def _blast_state():
    """A pressure bump off the block corners, smooth enough for 2 steps."""
    return dict(
        density=1.0,
        velocity=(0.1, -0.05, 0.02),
        pressure=lambda x, y, z: 1.0 + 4.0 * np.exp(-((x - 0.3) ** 2 + (y - 0.35) ** 2 + (z - 0.3) ** 2) / 0.01),
        vector_potential=(lambda x, y, z: 0.0 * x,
                          lambda x, y, z: 0.1 * np.sin(TWO_PI * z),
                          lambda x, y, z: 0.1 * np.sin(TWO_PI * x)),
    )


def _oracle(h, state, limiter="minmod"):
    t = amr3d_maps.build(h)
    buf = amr3d_maps.pack_state(h, t, 5.0 / 3.0, B0, state)
    o = OracleAMR3D(h, t, buf, mean_field=B0, limiter=limiter)
    o.prepare()
    return t, o


def _two_levels():
    h = Hierarchy3D(32)
    h.refine(0, [(0, 0, 0)])
    return h


def _three_levels():
    h = Hierarchy3D(32)
    h.refine(0, [(0, 0, 0)])
    h.refine(1, [(1, 1, 1)])
    return h


def _l_shape():
    """Three of the eight base blocks refined: coarse cells in the concave corner have
    coarse-fine faces in two directions."""
    h = Hierarchy3D(32)
    h.refine(0, [(0, 0, 0), (1, 0, 0), (0, 1, 0)])
    return h


def _hierarchy(kind):
    return {2: _two_levels, 3: _three_levels, "L": _l_shape}[kind]()


# -- host tables -----------------------------------------------------------------------------
@pytest.mark.parametrize("kind", [2, 3, "L"])
def test_reflux_cells_list_every_face_once_in_direction_order(kind):
    t = amr3d_maps.build(_hierarchy(kind))
    items, cells = t.reflux_items, t.reflux_cells
    assert [int((items["pad"] == d).sum()) for d in range(6)] == [len(r) for r in t.reflux]
    seen = set()
    for c in cells:
        n = int(c["n"])
        assert 1 <= n <= 3 and (c["item"][n:] == -1).all()
        dirs = [int(items["pad"][c["item"][f]]) for f in range(n)]
        assert dirs == sorted(set(dirs))
        # the faces of one cell name the same coarse cell
        where = set()
        for f in range(n):
            it = items[c["item"][f]]
            axis, lane = int(it["pad"]) // 2, int(c["lane"][f])
            q = [int(it["ci"]), int(it["cj"]), int(it["ck"])]
            q[1 if axis == 0 else 0] += lane // 8
            q[1 if axis == 2 else 2] += lane % 8
            where.add((int(it["coarse"]), *q))
            seen.add((int(c["item"][f]), lane))
        assert len(where) == 1
    assert len(seen) == 64 * len(items) == int(cells["n"].sum())
    assert (np.diff(cells["n"]) <= 0).all()
    if kind == "L":
        assert (cells["n"] == 2).any()


def _shell_fill_sets(ids, b=16):
    """The ghost cells and (component, edge) positions the restricted shell walk of the
    fill sets for a block whose 27 same-level neighbours are `ids` (-1: none)."""
    X, XN = b + 4, b + 5
    rel = lambda u: -1 if u < 0 else (1 if u >= b else 0)
    region = lambda q: ((rel(q[0] - 2) + 1) * 3 + rel(q[1] - 2) + 1) * 3 + rel(q[2] - 2) + 1
    own = lambda q, lim: all(2 <= q[a] < lim for a in range(3))
    cells = {p for p in itertools.product(range(X), repeat=3)
             if not own(p, b + 2) and ids[region(p)] < 0}
    edges = set()
    for c in range(3):
        ext = [X if a == c else XN for a in range(3)]
        for p in itertools.product(*(range(e) for e in ext)):
            if own(p, b + 2):
                continue
            for s in range(4):
                q, bit = list(p), 0
                for a in range(3):
                    if a != c:
                        q[a] -= (s >> bit) & 1
                        bit += 1
                if min(q) < 0 or max(q) >= X or own(q, b + 2):
                    continue
                if ids[region(q)] < 0:
                    edges.add((c, p))
                    break
    return cells, edges


def _region_fill_sets(ids, b=16):
    """The same for the (block, region) tasks: each missing region's cells and the
    edges of those cells outside the block's own range."""
    cells, edges = set(), set()
    for r in range(27):
        if r == 13 or ids[r] >= 0:
            continue
        rel = (r // 9 - 1, (r // 3) % 3 - 1, r % 3 - 1)
        lo = [0 if x < 0 else (2 if x == 0 else b + 2) for x in rel]
        n = [b if x == 0 else 2 for x in rel]
        cells |= set(itertools.product(*(range(lo[a], lo[a] + n[a]) for a in range(3))))
        for c in range(3):
            box = (range(lo[a], lo[a] + n[a] + (a != c)) for a in range(3))
            edges |= {(c, p) for p in itertools.product(*box)
                      if not all(2 <= p[a] < b + 2 for a in range(3))}
    return cells, edges


def test_fill_regions_set_what_the_restricted_shell_fill_sets():
    """The per-region fill tasks cover exactly the ghost cells and edges the shell walk
    of the restricted fill sets (same per-element code, so the same values)."""
    t = amr3d_maps.build(_l_shape())
    missing = np.flatnonzero(t.missing)
    assert len(missing) and set(t.fill_regions[:, 0]) == set(missing)
    # every missing region once: whole (part 0) or as its four parts; the parent named
    assert (t.fill_regions[:, 2] >= 0).all()
    for k in missing:
        got = t.fill_regions[t.fill_regions[:, 0] == k, 1]
        regions = sorted(set(int(v) & 31 for v in got))
        assert regions == sorted(int(r) for r in np.flatnonzero(t.hoods[k] == -1))
        for r in regions:
            parts = sorted(int(v) >> 5 for v in got if int(v) & 31 == r)
            assert parts in ([0], [1, 2, 3, 4])
    rng = np.random.default_rng(3)
    patterns = [t.hoods[k] for k in missing[:: max(1, len(missing) // 3)]]
    patterns += [np.where(rng.random(27) < 0.4, -1, 1) for _ in range(2)]
    for ids in patterns:
        ids = np.array(ids)
        ids[13] = 1
        assert _shell_fill_sets(ids) == _region_fill_sets(ids)


def test_nesting_adds_the_parent_neighbourhood():
    h = _three_levels()
    assert len(h.levels) == 3 and len(h.levels[2]) == 8
    for b in h.levels[2]:
        parent = tuple(v // 2 for v in b)
        for d in np.ndindex(3, 3, 3):
            q = tuple((parent[a] + d[a] - 1) % h.nb(1) for a in range(3))
            assert q in h.levels[1]


def test_a_refined_level_covering_the_domain_is_the_uniform_run():
    """All 8 level-1 blocks over the periodic domain evolve exactly as one uniform
    32^3 grid: same-level ghost cells and shared edges are exact copies."""
    h = Hierarchy3D(16)
    h.refine(0, [(0, 0, 0)])
    t, amr = _oracle(h, _state())
    uh = AMRHierarchy((32, 32, 32), ((0.0, 1.0),) * 3, periodic=(True, True, True))
    cons, pots = initial.host_state_3d(uh, uh.patches[0], 5.0 / 3.0, B0, **_state())
    uniform = Oracle3D(cons, pots, uh.spacing(0), mean_field=B0, limiter="minmod")
    for _ in range(2):
        dt = uniform.stable_step()
        uniform.step(dt)
        amr.step(dt)
    for k, row in enumerate(t.blocks):
        if row["level"] != 1:
            continue
        cells, bp = amr.views(amr.buf, k)
        o = [16 * int(row[a]) for a in ("bx", "by", "bz")]
        want = uniform.interior()[:, o[0] : o[0] + 16, o[1] : o[1] + 16, o[2] : o[2] + 16]
        np.testing.assert_array_equal(cells[:, 2:18, 2:18, 2:18], want)
        for c in range(3):
            sl = tuple(slice(2, 18 + (a != c)) for a in range(3))
            usl = tuple(slice(2 + o[a], 18 + o[a] + (a != c)) for a in range(3))
            np.testing.assert_array_equal(bp[c][sl], uniform.pot[c][usl])


@pytest.mark.parametrize("levels", [2, 3])
def test_refined_blocks_conserve_and_keep_div_b_zero(levels):
    h = _hierarchy(levels)
    t, amr = _oracle(h, _blast_state())
    before = amr.leaf_totals()
    for _ in range(2):
        amr.step()
    after = amr.leaf_totals()
    np.testing.assert_allclose(after, before, rtol=1e-13, atol=1e-14)
    for k in range(len(t.blocks)):
        o = Oracle3D(*amr.views(amr.buf, k), amr.spacing(k), mean_field=B0)
        assert o.divergence() < 1e-13
    # the coarse edges of refined blocks are the mean of the fine edges along them
    for row in t.children:
        _, cp = amr.views(amr.buf, row[0])
        kids = [amr.views(amr.buf, k)[1] for k in row[1:]]
        np.testing.assert_array_equal(cp[0][2, 2, 2], 0.5 * (kids[0][0][2, 2, 2] + kids[0][0][3, 2, 2]))


# -- GPU ----------------------------------------------------------------------------------
@pytest.mark.gpu
@pytest.mark.parametrize("levels,direct", [(2, True), (2, False), (3, True), ("L", True)])
def test_gpu_hierarchy_3d_is_bitwise_the_oracle(gpu, levels, direct):
    """Bit for bit the oracle, whose fill copies every ghost: with direct reads the
    stage and the refluxing read the cells around a block in place from the blocks of
    its level, and only the ghosts the level has no block for are filled."""
    from wellbalanced_amd.amr3d import AMRMHD3D

    h = _hierarchy(levels)
    t, amr = _oracle(h, _blast_state())
    model = AMRMHD3D(h, mean_field=B0, limiter="minmod")
    model.direct = direct
    assert 0 < len(model._missing) < model.nblocks
    model.set_state(**_blast_state())
    np.testing.assert_array_equal(model.buffer(), amr.buf)
    for _ in range(2):
        dt = amr.stable_step()
        assert model.stable_step() == dt
        model.step(dt)
        amr.step(dt)
    got = model.buffer()
    for k in range(len(t.blocks)):
        gc, gp = amr.views(got, k)
        wc, wp = amr.views(amr.buf, k)
        np.testing.assert_array_equal(gc[:, 2:18, 2:18, 2:18], wc[:, 2:18, 2:18, 2:18])
        for c in range(3):
            sl = tuple(slice(2, 18 + (a != c)) for a in range(3))
            np.testing.assert_array_equal(gp[c][sl], wp[c][sl])


@pytest.mark.gpu
@pytest.mark.parametrize("kind", [3, "L"])
def test_gpu_region_fill_is_the_restricted_shell_fill(gpu, kind):
    """wb_amr3d_fill_regions and the restricted shell walk of wb_amr3d_fill_list leave
    the same buffer, bit for bit."""
    import torch

    from wellbalanced_amd import _lib
    from wellbalanced_amd.amr3d import AMRMHD3D

    model = AMRMHD3D(_hierarchy(kind), mean_field=B0, limiter="minmod")
    model.set_state(**_blast_state())
    model.advance(1)
    stream = _lib.stream_handle(model.device)
    a = model._buf[0].clone()
    b = model._buf[0].clone()
    _lib.call("wb_amr3d_fill_regions", _lib.ptr(a), _lib.ptr(model._blocks),
              _lib.ptr(model._hoods), _lib.ptr(model._regions), model.nregions, stream)
    _lib.call("wb_amr3d_fill_list", _lib.ptr(b), _lib.ptr(model._blocks),
              _lib.ptr(model._missing), len(model._missing), 1, _lib.ptr(model._lookup),
              _lib.ptr(model._levels), stream)
    # the same with each face region's cells and edge components as separate tasks
    t = model.tables
    parent = np.full(len(t.blocks), -1, np.int32)
    for (L, *bc), k in t.index.items():
        if L > 0:
            parent[k] = t.index[(L - 1, *(v // 2 for v in bc))]
    split = amr3d_maps.fill_regions(t.hoods, parent, split=True)
    assert len(split) > model.nregions
    tasks = torch.as_tensor(split.reshape(-1), device=model.device)
    c = model._buf[0].clone()
    _lib.call("wb_amr3d_fill_regions", _lib.ptr(c), _lib.ptr(model._blocks),
              _lib.ptr(model._hoods), _lib.ptr(tasks), len(split), stream)
    torch.cuda.synchronize()
    assert model.nregions > 0
    assert torch.equal(a, b)
    assert torch.equal(c, b)


def _block_views(buf, t, k, X=20):
    """Block k's 5 cell fields (X^3 each) and its 3 potential components in a packed
    hierarchy buffer."""
    off = int(t.blocks["off"][k])
    cells = buf[off:off + 5 * X**3].reshape(5, X, X, X)
    o, pots = off + 5 * X**3, []
    for shape in [(X, X + 1, X + 1), (X + 1, X, X + 1), (X + 1, X + 1, X)]:
        m = int(np.prod(shape))
        pots.append(buf[o:o + m].reshape(shape))
        o += m
    return cells, pots


@pytest.mark.gpu
def test_gpu_mixed_block_stage_shapes_keep_the_bits(gpu):
    """The mixed block stage's launch shapes (whole blocks at 2 waves per SIMD, half
    blocks at 3; direct reads or filled ghosts) leave the same cells and potentials,
    bit for bit, over steps at one step size."""
    from wellbalanced_amd.amr3d import AMRMHD3D

    h = _hierarchy(3)
    results = {}
    for waves, direct in [(3, True), (3, False), (2, True), (2, False)]:
        model = AMRMHD3D(h, mean_field=B0, limiter="minmod", precision="mixed")
        model.waves, model.direct = waves, direct
        model.set_state(**_blast_state())
        dt = 0.5 * model.stable_step()
        for _ in range(3):
            model.step(dt)
        model.check()
        results[waves, direct] = (model, model.buffer())
    ref_model, ref = results[3, True]
    t = ref_model.tables
    for key, (model, got) in results.items():
        for k in range(len(t.blocks)):
            gc, gp = _block_views(got, t, k)
            wc, wp = _block_views(ref, t, k)
            np.testing.assert_array_equal(gc[:, 2:18, 2:18, 2:18], wc[:, 2:18, 2:18, 2:18],
                                          err_msg=str(key))
            for c in range(3):
                sl = tuple(slice(2, 18 + (a != c)) for a in range(3))
                np.testing.assert_array_equal(gp[c][sl], wp[c][sl], err_msg=str(key))


@pytest.mark.gpu
def test_gpu_three_level_mixed_conserves(gpu):
    from wellbalanced_amd.amr3d import AMRMHD3D

    h = _three_levels()
    model = AMRMHD3D(h, mean_field=B0, limiter="minmod", precision="mixed")
    model.set_state(**_blast_state())
    before = model.leaf_totals()
    model.advance(4)
    model.check()
    np.testing.assert_allclose(model.leaf_totals(), before, rtol=1e-12, atol=1e-13)
    assert model.divergence() < 1e-13
    assert model.composite_divergence() < 1e-12


@pytest.mark.gpu
def test_gpu_hierarchy_from_a_uniform_run(gpu):
    """A hierarchy started from a uniform run: the base blocks that are not refined
    keep the uniform state bit for bit, every block is divergence-free, and it steps."""
    from wellbalanced_amd import MHD3D
    from wellbalanced_amd.amr3d import AMRMHD3D

    problem = initial.blast_turbulence_3d(32, seed=2)
    uni = MHD3D.from_problem(problem)
    uni.advance(2)
    h = _two_levels()
    amr = AMRMHD3D.from_uniform(uni, h, precision="mixed")
    cons, _ = uni.arrays()
    got, _ = amr.level_state(0)
    refined = {tuple(int(v) for v in row[["bx", "by", "bz"]]) for row in amr.tables.blocks[amr.tables.children[:, 0]]}
    for b in h.levels[0]:
        if b in refined:
            continue
        sl = (slice(None),) + tuple(slice(2 + 16 * v, 18 + 16 * v) for v in b)
        np.testing.assert_array_equal(got[sl], cons[sl])
    assert amr.divergence() < 1e-13
    assert amr.composite_divergence() < 1e-12
    before = amr.leaf_totals()
    amr.advance(2)
    amr.check()
    np.testing.assert_allclose(amr.leaf_totals(), before, rtol=1e-12, atol=1e-13)
