"""A refined 3D hierarchy of equal blocks and the device tables of its step.

Every level is a set of blocks of ``B^3`` cells (``B = 16``) with a halo of 2; level
``L + 1`` refines level ``L`` by 2, and a refined level-``L`` block has all 8 children
(the refinement is an octree over blocks).  Nesting: the 26 neighbours (periodic)
of the parent of every block exist, so the ghost cells and edges of a block come
either from a block of its own level or by interpolation from the level below, and
the parent level covers every cell such an interpolation reads.

Each block's state is a 3D state of ``B^3`` cells as the uniform grid stores it
(5 cell planes ``(B+4)^3`` then ``A_x``, ``A_y``, ``A_z`` on the edges), the blocks
back to back in one buffer.  The tables:

* ``blocks``: per block (offset, level, block coordinates, spacing);
* ``lookup``: per level, the block id at each block coordinate (-1: none), so that
  kernels find the block holding any cell or edge of a level;
* ``children``: per refined block the 8 children (average-down and restriction);
* ``reflux``: per direction, the coarse-fine block faces (the fine block, the coarse
  block across the face and the origin of the 8 x 8 coarse cells there);
  ``reflux_items`` / ``reflux_cells``: all of them in one table and, per coarse cell
  with such faces, its faces in direction order (the one-launch refluxing);
* ``hoods``: per block, the offsets of the 27 blocks of its level around it (-1:
  none), through which the stage and the refluxing read the cells around a block in
  place;
* ``missing``: per block, whether a block of its level around it is missing, and
  ``fill_regions``: the ghost regions of such blocks that their level has no block for,
  the only ghosts filled (by interpolation) when the cells around a block are read in
  place.

Domains are periodic on every axis (the refined 3D workloads).
"""

from __future__ import annotations

import itertools
from dataclasses import dataclass

import numpy as np

HALO = 2
B = 16
NCELL = 5
EDGE = (B + 4, B + 5, B + 5)  # A_x extents; A_y, A_z are the cyclic permutations


def block_size(b: int = B) -> int:
    """Doubles of one block's state."""
    X = b + 4
    return NCELL * X**3 + 3 * X * (X + 1) ** 2


BLOCK_DTYPE = np.dtype([("off", np.int64), ("level", np.int32), ("bx", np.int32),
                        ("by", np.int32), ("bz", np.int32), ("dx", np.float64),
                        ("dy", np.float64), ("dz", np.float64)])
REFLUX_DTYPE = np.dtype([("fine", np.int32), ("coarse", np.int32), ("ci", np.int32),
                         ("cj", np.int32), ("ck", np.int32), ("pad", np.int32)])
# a coarse cell with coarse-fine faces: its n faces as (item, lane) in direction order
REFLUX_CELL_DTYPE = np.dtype([("n", np.int32), ("item", np.int32, 3), ("lane", np.int32, 3),
                              ("pad", np.int32)])


@dataclass
class Hierarchy3D:
    """Block sets per level over the periodic box ``bounds`` with ``n0^3`` base cells."""

    n0: int
    bounds: tuple = ((0.0, 1.0), (0.0, 1.0), (0.0, 1.0))
    block: int = B

    def __post_init__(self):
        if self.n0 % self.block:
            raise ValueError("the base grid must be a whole number of blocks")
        nb = self.n0 // self.block
        self.levels = [set(itertools.product(range(nb), repeat=3))]

    # -- geometry ------------------------------------------------------------------------
    def nb(self, level: int) -> int:
        return (self.n0 // self.block) * 2**level

    def level_shape(self, level: int) -> tuple:
        return (self.n0 * 2**level,) * 3

    def spacing(self, level: int) -> tuple:
        n = self.n0 * 2**level
        return tuple((hi - lo) / n for lo, hi in self.bounds)

    @property
    def max_level(self) -> int:
        return len(self.levels) - 1

    def blocks(self, level: int) -> list:
        return sorted(self.levels[level])

    # -- refinement ---------------------------------------------------------------------
    def refine(self, level: int, parents) -> None:
        """Refine the level-``level`` blocks ``parents`` (all 8 children each), then
        add blocks on every level until the nesting rule holds."""
        parents = {tuple(int(v) for v in p) for p in parents}
        missing = parents - self.levels[level]
        if missing:
            raise ValueError(f"blocks {sorted(missing)[:3]} are not on level {level}")
        while len(self.levels) <= level + 1:
            self.levels.append(set())
        for p in parents:
            for d in itertools.product((0, 1), repeat=3):
                self.levels[level + 1].add(tuple(2 * p[a] + d[a] for a in range(3)))
        self._nest()

    def _nest(self) -> None:
        """Finest level first: the parent of every block and its 26 neighbours must be
        blocks of the level below; a missing one is made by refining its own parent."""
        changed = True
        while changed:
            changed = False
            for L in range(len(self.levels) - 1, 0, -1):
                nbp = self.nb(L - 1)
                need = set()
                for b in self.levels[L]:
                    parent = tuple(v // 2 for v in b)
                    for d in itertools.product((-1, 0, 1), repeat=3):
                        need.add(tuple((parent[a] + d[a]) % nbp for a in range(3)))
                for q in need:
                    if q not in self.levels[L - 1]:
                        if L - 1 == 0:
                            raise AssertionError("level 0 covers the domain")
                        self._refine_one(L - 2, tuple(v // 2 for v in q))
                        changed = True

    def _refine_one(self, level: int, block) -> None:
        if block not in self.levels[level]:
            self._refine_one(level - 1, tuple(v // 2 for v in block))
        for d in itertools.product((0, 1), repeat=3):
            self.levels[level + 1].add(tuple(2 * block[a] + d[a] for a in range(3)))

    def cells(self) -> int:
        return sum(len(s) for s in self.levels) * self.block**3

    def leaf_cells(self) -> int:
        covered = sum(len(s) for s in self.levels[1:]) // 8 if len(self.levels) > 1 else 0
        return self.cells() - covered * self.block**3

    def refine_flagged(self, level: int, flags: np.ndarray) -> None:
        """Refine every level-``level`` block holding a flagged cell (``flags`` over the
        level's cells; cells outside its blocks are ignored)."""
        b = self.block
        nb = self.nb(level)
        f = np.asarray(flags, bool).reshape(nb, b, nb, b, nb, b).any(axis=(1, 3, 5))
        parents = [tuple(p) for p in np.argwhere(f) if tuple(p) in self.levels[level]]
        if parents:
            self.refine(level, parents)


@dataclass
class Tables3D:
    blocks: np.ndarray        # BLOCK_DTYPE per block
    index: dict               # (level, bx, by, bz) -> block id
    lookup: np.ndarray        # int32, every level's nb^3 ids back to back
    level_nb: list
    level_base: list
    children: np.ndarray      # int32 (nrefined, 9): parent id, 8 children (x slowest)
    reflux: list              # per direction (axis, side): REFLUX_DTYPE rows
    total: int                # doubles of the buffer
    size: int                 # doubles per block
    missing: np.ndarray = None  # bool per block: a block of its level around it is missing
    hoods: np.ndarray = None    # int64 (nblocks, 27): offsets of the blocks of its level around
    reflux_items: np.ndarray = None  # REFLUX_DTYPE, every direction, pad = 2 axis + side
    reflux_cells: np.ndarray = None  # REFLUX_CELL_DTYPE per refluxed coarse cell
    fill_regions: np.ndarray = None  # int32 (n, 4): block, region + 32 part, parent, 0


def build(h: Hierarchy3D) -> Tables3D:
    size = block_size(h.block)
    rows, index = [], {}
    for L in range(len(h.levels)):
        dx, dy, dz = h.spacing(L)
        for bc in h.blocks(L):
            k = len(rows)
            index[(L, *bc)] = k
            rows.append((k * size, L, bc[0], bc[1], bc[2], dx, dy, dz))
    blocks = np.array(rows, BLOCK_DTYPE)
    level_nb, level_base, parts = [], [], []
    base = 0
    for L in range(len(h.levels)):
        nb = h.nb(L)
        lk = np.full((nb, nb, nb), -1, np.int32)
        for bc in h.levels[L]:
            lk[bc] = index[(L, *bc)]
        level_nb.append(nb)
        level_base.append(base)
        parts.append(lk.reshape(-1))
        base += nb**3
    lookup = np.concatenate(parts).astype(np.int32)
    children = []
    for L in range(len(h.levels) - 1):
        for bc in h.blocks(L):
            kids = [index.get((L + 1, 2 * bc[0] + a, 2 * bc[1] + b, 2 * bc[2] + c))
                    for a in (0, 1) for b in (0, 1) for c in (0, 1)]
            if all(k is not None for k in kids):
                children.append([index[(L, *bc)], *kids])
    children = np.array(children, np.int32).reshape(-1, 9)
    reflux = []
    half = h.block // 2
    for axis in range(3):
        for side in (0, 1):
            # side 0: the coarse cells at the fine block's low face (their +axis face);
            # side 1: those at its high face (their -axis face)
            items = []
            for L in range(1, len(h.levels)):
                nb, nbc = h.nb(L), h.nb(L - 1)
                for bc in h.blocks(L):
                    step = -1 if side == 0 else 1
                    nbr = list(bc)
                    nbr[axis] = (nbr[axis] + step) % nb
                    if tuple(nbr) in h.levels[L]:
                        continue
                    # coarse cells across the face: level-(L-1) cell coordinates
                    lo = [bc[a] * h.block // 2 for a in range(3)]
                    lo[axis] = (lo[axis] - 1) % (nbc * h.block) if side == 0 else \
                        (lo[axis] + half) % (nbc * h.block)
                    cb = tuple(lo[a] // h.block for a in range(3))
                    coarse = index[(L - 1, *cb)]
                    items.append((index[(L, *bc)], coarse, *(lo[a] % h.block for a in range(3)), 0))
            reflux.append(np.array(items, REFLUX_DTYPE))
    # the blocks with a block of their level missing around them
    missing = np.zeros(len(rows), bool)
    around = list(itertools.product((-1, 0, 1), repeat=3))
    for L in range(len(h.levels)):
        nb = h.nb(L)
        for bc in h.levels[L]:
            missing[index[(L, *bc)]] = not all(
                tuple((bc[a] + d[a]) % nb for a in range(3)) in h.levels[L] for d in around)
    # the 27 blocks of its level around each block (x slowest), -1 where there is none
    hoods = np.full((len(rows), 27), -1, np.int64)
    for (L, *bc), k in index.items():
        nb = h.nb(L)
        for r, d in enumerate(around):
            j = index.get((L, *((bc[a] + d[a]) % nb for a in range(3))))
            if j is not None:
                hoods[k, r] = blocks["off"][j]
    items, cells = reflux_cells(reflux, h.block)
    parent = np.full(len(rows), -1, np.int32)
    for (L, *bc), k in index.items():
        if L > 0:
            parent[k] = index[(L - 1, *(v // 2 for v in bc))]
    return Tables3D(blocks, index, lookup, level_nb, level_base, children, reflux,
                    len(rows) * size, size, missing, hoods, items, cells,
                    fill_regions(hoods, parent))


def fill_regions(hoods: np.ndarray, parent: np.ndarray, split: bool = False) -> np.ndarray:
    """The fill tasks of the 26 ghost regions around each block (region (x + 1) * 9 +
    (y + 1) * 3 + z + 1) that its level has no block for: (block, region + 32 part,
    parent block, 0), part 0 the whole region, or (`split`) for a face region four tasks,
    its cells (part 1) and the edges of each component (parts 2-4); the largest tasks
    first."""
    b, r = np.nonzero(hoods == -1)
    rel = np.stack([r // 9 - 1, (r // 3) % 3 - 1, r % 3 - 1], axis=1)
    face = ((rel == 0).sum(axis=1) == 2) & split
    zeros = (rel == 0).sum(axis=1)  # 2: a face, 1: an edge, 0: a corner
    tb = np.concatenate([np.repeat(b[face], 4), b[~face]])
    tr = np.concatenate([np.repeat(r[face], 4) + 32 * np.tile(np.arange(1, 5), face.sum()),
                         r[~face]])
    tz = np.concatenate([np.full(4 * face.sum(), 2), zeros[~face]])
    order = np.lexsort((tr, tb, -tz))
    tb, tr = tb[order], tr[order]
    return np.stack([tb, tr, parent[tb], np.zeros_like(tb)], axis=1).astype(np.int32)


def reflux_cells(reflux: list, b: int = B) -> tuple[np.ndarray, np.ndarray]:
    """Every direction's refluxing items in one table (pad = 2 axis + side) and, per
    coarse cell with coarse-fine faces, its faces as (item, lane) in direction order;
    cells with more faces first (a wave then runs one face count)."""
    parts = []
    for d, rows in enumerate(reflux):
        r = rows.copy()
        r["pad"] = d
        parts.append(r)
    items = np.concatenate(parts) if parts else np.zeros(0, REFLUX_DTYPE)
    if len(items) == 0:
        return items, np.zeros(0, REFLUX_CELL_DTYPE)
    lane = np.tile(np.arange(64), len(items))
    it = np.repeat(np.arange(len(items)), 64)
    axis = items["pad"][it] // 2
    c = np.stack([items["ci"][it], items["cj"][it], items["ck"][it]], axis=1)
    a1 = np.where(axis == 0, 1, 0)
    a2 = np.where(axis == 2, 1, 2)
    rows_ = np.arange(len(it))
    c[rows_, a1] += lane // 8
    c[rows_, a2] += lane % 8
    key = (items["coarse"][it].astype(np.int64) * b + c[:, 0]) * b * b + c[:, 1] * b + c[:, 2]
    order = np.argsort(key, kind="stable")  # items are in direction order
    key, it, lane = key[order], it[order], lane[order]
    first = np.flatnonzero(np.r_[True, key[1:] != key[:-1]])
    count = np.diff(np.r_[first, len(key)])
    if count.max() > 3:
        raise ValueError("a coarse cell with more than three coarse-fine faces")
    cells = np.zeros(len(first), REFLUX_CELL_DTYPE)
    cells["n"] = count
    cells["item"] = -1
    cells["lane"] = -1
    for f in range(3):
        has = count > f
        cells["item"][has, f] = it[first[has] + f]
        cells["lane"][has, f] = lane[first[has] + f]
    return items, cells[np.argsort(-count, kind="stable")]


def face_slots(nblocks: int, items: np.ndarray):
    """The block faces whose fluxes the stage records for the refluxing: per block 6
    slots (face 2 axis + side, side 0 the low face; -1: not recorded) and their count.
    An item (a fine block's face over coarse cells, pad = 2 axis + side) needs the fine
    block's face 2 axis + side and the coarse block's face 2 axis + (1 - side)."""
    slots = np.full((nblocks, 6), -1, np.int32)
    n = 0
    for it in items:
        axis, side = int(it["pad"]) >> 1, int(it["pad"]) & 1
        for blk, face in ((int(it["fine"]), 2 * axis + side), (int(it["coarse"]), 2 * axis + 1 - side)):
            if slots[blk, face] < 0:
                slots[blk, face] = n
                n += 1
    return slots, n


# -- host views of a block ----------------------------------------------------------------
def block_views(buf: np.ndarray, off: int, b: int = B):
    """(cells (5, X, X, X), [A_x, A_y, A_z]) views of the block at ``off``."""
    X = b + 4
    cells = buf[off : off + NCELL * X**3].reshape(NCELL, X, X, X)
    o = off + NCELL * X**3
    shapes = [(X, X + 1, X + 1), (X + 1, X, X + 1), (X + 1, X + 1, X)]
    pots = []
    for s in shapes:
        n = int(np.prod(s))
        pots.append(buf[o : o + n].reshape(s))
        o += n
    return cells, pots


def _wrapped(fn, bounds, n):
    """``fn`` evaluated with node coordinates at the upper end of the periodic box moved
    to the lower end (the same node, and the bit-identical coordinate ``lo + 0 * dx``)."""
    if not callable(fn):
        return fn

    def wrapped(*xs):
        out = []
        for a, x in enumerate(xs):
            lo, hi = bounds[a]
            dx = (hi - lo) / n
            idx = np.rint((np.asarray(x) - lo) / dx)
            out.append(np.where(idx >= n, lo + (idx - n) * dx, x))
        return fn(*out)

    return wrapped


def pack_state(h: Hierarchy3D, t: Tables3D, gamma: float, mean_field, state: dict) -> np.ndarray:
    """The initial buffer: every block evaluated from ``state`` (as `initial.host_state_3d`
    takes it) on its own level; halos are left to the first ghost fill."""
    from . import initial
    from .hierarchy import AMRHierarchy, Patch

    buf = np.zeros(t.total)
    for k, row in enumerate(t.blocks):
        L = int(row["level"])
        n = h.n0 * 2**L
        lvl = AMRHierarchy((n, n, n), h.bounds, periodic=(True, True, True))
        box = tuple((int(row[a]) * h.block, (int(row[a]) + 1) * h.block) for a in ("bx", "by", "bz"))
        patch = Patch(0, box)
        st = dict(state)
        pots = st.get("vector_potential", (0.0, 0.0, 0.0))
        st["vector_potential"] = tuple(_wrapped(f, h.bounds, n) for f in pots)
        cons, potentials = initial.host_state_3d(lvl, patch, gamma, mean_field, **st)
        cells, views = block_views(buf, int(row["off"]), h.block)
        cells[...] = cons
        for v, p in zip(views, potentials):
            v[...] = p
    return buf
