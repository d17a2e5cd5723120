"""Block-sparse tensors in the "flat tensor" exchange format of the C ABI.

A tensor with quantum-number (QN) sectors, as GraceQ/tensor's GQTensor
(used throughout include/gqmps2/, e.g. the leg conventions of
two_site_update_finite_vmps_impl.h:117-129), is exchanged as

* ``meta`` (int32): rank, qn_width, nblocks; per leg: dir (0 IN, 1 OUT),
  nsect, then per sector qn[qn_width] and dim; then per block the sector
  coordinate of every leg.  Blocks are sorted lexicographically.
* ``data`` (float64): the blocks in ``meta`` order, each dense row-major
  over its sector dimensions; complex tensors (GQTEN_Complex) store each
  element as two float64 (real, imaginary), interleaved.

The divergence of a block is sum(qn of OUT legs) - sum(qn of IN legs).
This module is plain numpy: it builds inputs, densifies outputs and
compares tensors in the tests; it performs no product computation.
"""
from __future__ import annotations

import itertools
from dataclasses import dataclass, field

import numpy as np

IN, OUT = 0, 1


@dataclass
class Leg:
    dir: int
    sectors: list  # list of (qn tuple, dim)

    def inverse(self) -> "Leg":
        return Leg(1 - self.dir, list(self.sectors))

    @property
    def dim(self) -> int:
        return sum(d for _, d in self.sectors)

    def offsets(self):
        out, o = [], 0
        for _, d in self.sectors:
            out.append(o)
            o += d
        return out

    def key(self):
        return (self.dir, tuple((tuple(q), d) for q, d in self.sectors))


@dataclass
class BTensor:
    legs: list
    qw: int = 1
    blocks: dict = field(default_factory=dict)  # coord tuple -> ndarray

    @property
    def rank(self):
        return len(self.legs)

    @property
    def is_complex(self) -> bool:
        return any(np.iscomplexobj(b) for b in self.blocks.values())

    def block_shape(self, c):
        return tuple(self.legs[i].sectors[c[i]][1] for i in range(self.rank))

    def block_div(self, c):
        q = np.zeros(self.qw, dtype=np.int64)
        for i, leg in enumerate(self.legs):
            s = 1 if leg.dir == OUT else -1
            q += s * np.asarray(leg.sectors[c[i]][0], dtype=np.int64)
        return tuple(int(x) for x in q)

    # ------------------------------------------------------------ exchange
    def to_flat(self):
        meta = [self.rank, self.qw, len(self.blocks)]
        for leg in self.legs:
            meta += [leg.dir, len(leg.sectors)]
            for q, d in leg.sectors:
                meta += list(q) + [d]
        coords = sorted(self.blocks)
        for c in coords:
            meta += list(c)
        if self.is_complex:
            data = [np.asarray(self.blocks[c], dtype=np.complex128).ravel().view(np.float64) for c in coords]
        else:
            data = [np.asarray(self.blocks[c], dtype=np.float64).ravel() for c in coords]
        data = np.concatenate(data) if data else np.zeros(0)
        return np.asarray(meta, dtype=np.int32), np.ascontiguousarray(data, dtype=np.float64)

    @staticmethod
    def from_flat(meta, data, is_complex=False) -> "BTensor":
        meta = [int(x) for x in meta]
        p = 0
        rank, qw, nblk = meta[0], meta[1], meta[2]
        p = 3
        legs = []
        for _ in range(rank):
            d, ns = meta[p], meta[p + 1]
            p += 2
            secs = []
            for _ in range(ns):
                q = tuple(meta[p:p + qw])
                p += qw
                secs.append((q, meta[p]))
                p += 1
            legs.append(Leg(d, secs))
        t = BTensor(legs, qw)
        coords = []
        for _ in range(nblk):
            coords.append(tuple(meta[p:p + rank]))
            p += rank
        off = 0
        w = 2 if is_complex else 1
        for c in coords:
            shp = t.block_shape(c)
            n = int(np.prod(shp)) if shp else 1
            x = np.asarray(data[off:off + w * n], dtype=np.float64)
            if is_complex:
                x = x.view(np.complex128)
            t.blocks[c] = x.reshape(shp).copy()
            off += w * n
        return t

    # ------------------------------------------------------------ helpers
    def dense(self) -> np.ndarray:
        shp = tuple(l.dim for l in self.legs)
        out = np.zeros(shp, dtype=np.complex128 if self.is_complex else np.float64)
        offs = [l.offsets() for l in self.legs]
        for c, b in self.blocks.items():
            sl = tuple(slice(offs[i][c[i]], offs[i][c[i]] + b.shape[i]) for i in range(self.rank))
            out[sl] += b
        return out

    def allowed_coords(self, div):
        div = tuple(div)
        out = []
        for c in itertools.product(*[range(len(l.sectors)) for l in self.legs]):
            if self.block_div(c) == div:
                out.append(c)
        return out

    @staticmethod
    def random(legs, div, rng, qw=1, scale=1.0) -> "BTensor":
        t = BTensor(list(legs), qw)
        for c in t.allowed_coords(div):
            t.blocks[c] = scale * rng.uniform(-1.0, 1.0, size=t.block_shape(c))
        return t

    def dag(self) -> "BTensor":
        return BTensor([l.inverse() for l in self.legs], self.qw,
                       {c: np.conj(b).copy() for c, b in self.blocks.items()})

    def norm(self) -> float:
        return float(np.sqrt(sum(float(np.sum(np.abs(b) ** 2)) for b in self.blocks.values())))

    @staticmethod
    def random_complex(legs, div, rng, qw=1, scale=1.0) -> "BTensor":
        t = BTensor(list(legs), qw)
        for c in t.allowed_coords(div):
            shp = t.block_shape(c)
            t.blocks[c] = scale * (rng.uniform(-1.0, 1.0, size=shp) + 1j * rng.uniform(-1.0, 1.0, size=shp))
        return t

    def nnz(self) -> int:
        return sum(b.size for b in self.blocks.values())


def leg(dir, *sectors):
    """leg(OUT, ((1,), 1), ((-1,), 1)) -> Leg."""
    return Leg(dir, [(tuple(q) if isinstance(q, (tuple, list)) else (q,), d) for q, d in sectors])
