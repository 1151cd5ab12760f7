"""MOLSIM configuration files (.cnf), as IOCnf reads and writes them
(Src/molsim.F90:840-963).

A .cnf file is a Fortran sequential unformatted file: every record is its
payload between two little-endian int32 byte counts.  IOCnf('write') writes

  1. istep1, iseed (int32), am (float64), ix, iy (int32), boxlen(3) (float64)
     (files before 2017 hold nstep1done, iseed, boxlen(3): 32 bytes)
  2. np, iaux (int32), raux, raux (float64), laux, laux (logical = int32)
  3. (ro(1:3,ip), qua(0:3,ip), ip = 1, np)                      float64
  4. if lclink:      nbondcl(1:np), bondcl(1:maxvalnbondcl,1:np) int32
  5. if lweakcharge: laz(1:np)                                   logical
  6. if lmd:         (rod(1:3,ip), quad(0:3,ip), ip = 1, np)     float64

and other modules may append their own records, which are kept as bytes.
The count record may hold np = 0 (older writers); the particle count then
follows from the position record.  Reading and writing a file reproduces
it byte for byte.
"""
from __future__ import annotations

import struct
from dataclasses import dataclass, field

import numpy as np

__all__ = ["Cnf", "read_records", "write_records", "read_cnf", "write_cnf"]


def read_records(path) -> list[bytes]:
    """The payloads of a Fortran sequential unformatted file."""
    data = open(path, "rb").read()
    out, o = [], 0
    while o < len(data):
        if o + 4 > len(data):
            raise ValueError(f"{path}: truncated record marker at byte {o}")
        (n,) = struct.unpack_from("<i", data, o)
        if n < 0 or o + 8 + n > len(data):
            raise ValueError(f"{path}: bad record length {n} at byte {o}")
        (n2,) = struct.unpack_from("<i", data, o + 4 + n)
        if n2 != n:
            raise ValueError(f"{path}: record markers {n} and {n2} differ at byte {o}")
        out.append(data[o + 4:o + 4 + n])
        o += 8 + n
    return out


def write_records(path, records) -> None:
    with open(path, "wb") as fh:
        for r in records:
            fh.write(struct.pack("<i", len(r)))
            fh.write(r)
            fh.write(struct.pack("<i", len(r)))


@dataclass
class Cnf:
    """One configuration: header, counts, positions and orientations, and
    the optional bond, charge-flag and velocity records."""
    istep: int
    iseed: int
    boxlen: np.ndarray                       # (3,)
    ro: np.ndarray                           # (np, 3)
    qua: np.ndarray                          # (np, 4) quaternions
    am: float | None = None                  # Random's state (None: old header)
    ix: int = 0
    iy: int = 0
    counts: tuple = (0, 0, 0.0, 0.0, 0, 0)   # record 2 as stored
    nbondcl: np.ndarray | None = None        # (np,)
    bondcl: np.ndarray | None = None         # (np, maxvalnbondcl)
    laz: np.ndarray | None = None            # (np,) bool
    rod: np.ndarray | None = None            # (np, 3)
    quad: np.ndarray | None = None           # (np, 4)
    extra: list = field(default_factory=list)

    @property
    def np(self) -> int:
        return self.ro.shape[0]


def read_cnf(path, lclink=False, maxvalnbondcl=0, lweakcharge=False, lmd=None) -> Cnf:
    """IOCnf('read').  The optional records follow the input's flags, as in
    the reference; lmd=None takes a velocity record when the next record has
    the size of one."""
    recs = read_records(path)
    if len(recs) < 3:
        raise ValueError(f"{path}: {len(recs)} records, a .cnf file has at least 3")
    h = recs[0]
    if len(h) == 48:
        istep, iseed, am, ix, iy, *bl = struct.unpack("<iidii3d", h)
    elif len(h) == 32:
        istep, iseed, *bl = struct.unpack("<ii3d", h)
        am, ix, iy = None, 0, 0
    else:
        raise ValueError(f"{path}: header record of {len(h)} bytes")
    if len(recs[1]) != 32:
        raise ValueError(f"{path}: count record of {len(recs[1])} bytes")
    counts = struct.unpack("<iiddii", recs[1])
    if len(recs[2]) % 56:
        raise ValueError(f"{path}: position record of {len(recs[2])} bytes")
    n = len(recs[2]) // 56
    if counts[0] not in (0, n):
        raise ValueError(f"{path}: np {counts[0]} but {n} positions")
    pq = np.frombuffer(recs[2], "<f8").reshape(n, 7)
    c = Cnf(istep=istep, iseed=iseed, boxlen=np.array(bl), ro=pq[:, :3].copy(),
            qua=pq[:, 3:].copy(), am=am, ix=ix, iy=iy, counts=counts)
    k = 3
    if lclink:
        r = np.frombuffer(recs[k], "<i4")
        if r.size != n * (1 + maxvalnbondcl):
            raise ValueError(f"{path}: bond record of {len(recs[k])} bytes")
        c.nbondcl = r[:n].copy()
        c.bondcl = r[n:].reshape(n, maxvalnbondcl).copy()
        k += 1
    if lweakcharge:
        r = np.frombuffer(recs[k], "<i4")
        if r.size != n:
            raise ValueError(f"{path}: charge-flag record of {len(recs[k])} bytes")
        c.laz = r != 0
        k += 1
    if lmd or (lmd is None and k < len(recs) and len(recs[k]) == 56 * n):
        r = np.frombuffer(recs[k], "<f8").reshape(n, 7)
        c.rod, c.quad = r[:, :3].copy(), r[:, 3:].copy()
        k += 1
    c.extra = recs[k:]
    return c


def write_cnf(path, c: Cnf) -> None:
    """IOCnf('write'): the records in the order above (the header in the
    form it was read in)."""
    if c.am is None:
        recs = [struct.pack("<ii3d", c.istep, c.iseed, *map(float, c.boxlen))]
    else:
        recs = [struct.pack("<iidii3d", c.istep, c.iseed, c.am, c.ix, c.iy,
                            *map(float, c.boxlen))]
    recs.append(struct.pack("<iiddii", *c.counts))
    recs.append(np.ascontiguousarray(np.hstack([c.ro, c.qua]), "<f8").tobytes())
    if c.nbondcl is not None:
        recs.append(np.concatenate([c.nbondcl, c.bondcl.reshape(-1)]).astype("<i4").tobytes())
    if c.laz is not None:
        recs.append(np.asarray(c.laz, "<i4").tobytes())
    if c.rod is not None:
        recs.append(np.ascontiguousarray(np.hstack([c.rod, c.quad]), "<f8").tobytes())
    recs.extend(c.extra)
    write_records(path, recs)
