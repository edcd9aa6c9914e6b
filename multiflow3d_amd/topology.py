"""Block topology: a conforming grid of rectangular blocks, described the way
the reference describes its ``dom(ib)`` entries.

The reference finds face neighbours by matching block coordinates
(alloc_dom.for:456-616) and periodic partners by matching the global domain
ends (alloc_dom.for:972-1150); a block is never its own periodic partner, so
a single block along a periodic axis wraps onto itself instead
(exchangepp.for:4801-4811).  Here a block grid is a tensor product of
per-axis extents: block (a, b, c) has extent (ex[a], ey[b], ez[c]) and index
``a + nbx*(b + nby*c)``.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import List, Sequence

PERIODIC = 5      # the reference's bc code for periodic (mgsolver.for:192-203)
NEUMANN = 4       # any code other than 5 is a Neumann wall for pressure

FACES = ("W", "E", "S", "N", "B", "T")


@dataclass
class Block:
    """The multigrid-relevant fields of one ``dom(ib)`` (module_multidata.for:16-28,36,116)."""
    n: List[int]                      # interior cells along i, j, k
    h: List[float]                    # dx, dy, dz
    nbr: List[int] = field(default_factory=lambda: [-1] * 6)   # iprev,inext,jprev,jnext,kprev,knext
    per: List[int] = field(default_factory=lambda: [-1] * 6)   # per_ip,per_in,per_jp,per_jn,per_kp,per_kn
    bc: List[int] = field(default_factory=lambda: [NEUMANN] * 6)
    origin: List[int] = field(default_factory=lambda: [0, 0, 0])  # global cell offset
    rdiv: int = 1                     # refinement ratio (infodom.cin column 2); n and h are
                                      # at the block's own resolution

    @property
    def cells(self) -> int:
        return self.n[0] * self.n[1] * self.n[2]


@dataclass
class BlockGrid:
    ex: List[int]
    ey: List[int]
    ez: List[int]
    bc: List[int]
    h: List[float]
    blocks: List[Block]

    @property
    def rdiv(self) -> List[int]:
        return [b.rdiv for b in self.blocks]

    @property
    def refined(self) -> bool:
        return any(b.rdiv > 1 for b in self.blocks)

    @property
    def shape(self):
        return (len(self.ex), len(self.ey), len(self.ez))

    @property
    def global_n(self):
        return (sum(self.ex), sum(self.ey), sum(self.ez))

    @property
    def cells(self) -> int:
        gx, gy, gz = self.global_n
        return gx * gy * gz


def block_grid(ex: Sequence[int], ey: Sequence[int], ez: Sequence[int],
               bc: Sequence[int], h: Sequence[float]) -> BlockGrid:
    """Tensor grid of blocks with global boundary codes ``bc`` (W,E,S,N,B,T)."""
    ex, ey, ez = list(ex), list(ey), list(ez)
    bc = list(bc)
    if len(bc) != 6:
        raise ValueError("bc needs 6 codes (W,E,S,N,B,T)")
    for d, (lo, hi) in enumerate(((0, 1), (2, 3), (4, 5))):
        if (bc[lo] == PERIODIC) != (bc[hi] == PERIODIC):
            raise ValueError(f"axis {d}: periodic on one side only")
    nb = (len(ex), len(ey), len(ez))

    def bid(a, b, c):
        if 0 <= a < nb[0] and 0 <= b < nb[1] and 0 <= c < nb[2]:
            return a + nb[0] * (b + nb[1] * c)
        return -1

    blocks = []
    oz = 0
    for c in range(nb[2]):
        oy = 0
        for b in range(nb[1]):
            ox = 0
            for a in range(nb[0]):
                blk = Block(n=[ex[a], ey[b], ez[c]], h=list(map(float, h)), bc=list(bc),
                            origin=[ox, oy, oz])
                blk.nbr = [bid(a - 1, b, c), bid(a + 1, b, c), bid(a, b - 1, c),
                           bid(a, b + 1, c), bid(a, b, c - 1), bid(a, b, c + 1)]
                per = [-1] * 6
                if bc[0] == PERIODIC and a == 0 and nb[0] > 1:
                    per[0] = bid(nb[0] - 1, b, c)
                if bc[1] == PERIODIC and a == nb[0] - 1 and nb[0] > 1:
                    per[1] = bid(0, b, c)
                if bc[2] == PERIODIC and b == 0 and nb[1] > 1:
                    per[2] = bid(a, nb[1] - 1, c)
                if bc[3] == PERIODIC and b == nb[1] - 1 and nb[1] > 1:
                    per[3] = bid(a, 0, c)
                if bc[4] == PERIODIC and c == 0 and nb[2] > 1:
                    per[4] = bid(a, b, nb[2] - 1)
                if bc[5] == PERIODIC and c == nb[2] - 1 and nb[2] > 1:
                    per[5] = bid(a, b, 0)
                blk.per = per
                blocks.append(blk)
                ox += ex[a]
            oy += ey[b]
        oz += ez[c]
    return BlockGrid(ex=ex, ey=ey, ez=ez, bc=bc, h=list(map(float, h)), blocks=blocks)


def refined_grid(ex: Sequence[int], ey: Sequence[int], ez: Sequence[int], bc: Sequence[int],
                 h: Sequence[float], rdiv: Sequence[int]) -> BlockGrid:
    """The tensor grid of block_grid with block i refined rdiv[i] times
    (localparameters.for:255-257): extents x rdiv, spacing / rdiv.  ex, ey,
    ez, h and the blocks' origins stay in unrefined cells."""
    g = block_grid(ex, ey, ez, bc, h)
    if len(rdiv) != len(g.blocks):
        raise ValueError("one rdiv per block")
    for blk, r in zip(g.blocks, rdiv):
        r = int(r)
        if r < 1 or r & (r - 1):
            raise ValueError("rdiv must be a power of 2")
        blk.rdiv = r
        blk.n = [n * r for n in blk.n]
        blk.h = [x / r for x in blk.h]
    return g


def single_block(n, h, bc) -> BlockGrid:
    n = list(n) if hasattr(n, "__len__") else [n, n, n]
    return block_grid([n[0]], [n[1]], [n[2]], bc, h)


# Boundary conditions of every workload here: periodic x and z, Neumann walls in y.
CHANNEL_BC = [PERIODIC, PERIODIC, NEUMANN, NEUMANN, PERIODIC, PERIODIC]


def max_levels(grid: BlockGrid) -> int:
    """Largest global level count ngrd_gl such that every block extent
    divides by 2**(own levels - 1), a block of rdiv r having log2(r) levels
    more than an unrefined one (localparameters.for:33-56,201-206)."""
    top = max(b.rdiv for b in grid.blocks).bit_length() - 1
    L = top + 1
    while all((n % (1 << (L - top + (b.rdiv.bit_length() - 1)))) == 0 for b in grid.blocks for n in b.n):
        L += 1
    return L
