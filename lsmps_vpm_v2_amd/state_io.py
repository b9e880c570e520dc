"""Particle-state CSV files of the reference (2-D), read and written.

Format (reference src/save_data/state_data.cpp:335-392, ``type`` 0): a header
line, then one line per particle with the columns

    xp,yp,gpz,vor,up,vp,sp,active,chi[,pressure]

written with 12 significant digits (``precision(12)``, i.e. ``%.12g``), the
velocity written relative to the body (``u - ubody``, ``v - vbody``) and the
active flag as 0 / 1.  The neighbour columns (``ngh_num``, ``ngh_ID``) belong
to the LSMPS neighbour search, which this package does not have; they are
neither written nor read.

Reading follows ``initialization::read_2d_particle``
(src/initialization/initialization.cpp:561-691): the header line is skipped,
each field is parsed as a double, column 9 is the pressure (kept only when
asked for), the ``chi`` column is not restored (the reference's switch has no
case for it; it is returned separately here), and each particle's level comes
from its size:

    lvl = max_level;  mul = int((sp + sigma/100) / sigma);
    while mul != 1: lvl -= 1; mul //= 2

(:660-666).  A size below 0.99 sigma gives ``mul = 0``, for which that loop
never ends in the reference; here it raises ``ValueError``.
"""
from __future__ import annotations

import os

import numpy as np

from .solver import Particle

COLUMNS = ("xp", "yp", "gpz", "vor", "up", "vp", "sp", "active", "chi")


def particle_level(sp, max_level: int, sigma: float) -> np.ndarray:
    """Level of each particle from its size (initialization.cpp:660-666)."""
    sp = np.atleast_1d(np.asarray(sp, dtype=np.float64))
    mul = ((sp + sigma / 100) / sigma).astype(np.int64)  # C++ int() truncation
    if np.any(mul < 1):
        i = int(np.argmax(mul < 1))
        raise ValueError(f"particle {i}: size {sp[i]!r} is below 0.99 sigma ({sigma!r}); "
                         "the reference's level loop would not end")
    # the loop halves mul until it is 1: lvl = max_level - floor(log2(mul))
    steps = np.zeros_like(mul)
    m = mul.copy()
    while np.any(m != 1):
        more = m != 1
        steps[more] += 1
        m[more] //= 2
    return (max_level - steps).astype(np.int32)


def read_2d_particle(path: str, max_level: int, sigma: float, pressure: bool = False):
    """Read a 2-D particle state (initialization.cpp:561-691).

    Returns ``(particle, extra)``: a :class:`Particle` with x, y, s,
    vorticity, isActive, u, v, and a dict with ``gz``, ``level``, ``chi`` and,
    when ``pressure`` is set, ``P``.  A missing file raises ``RuntimeError``
    with the reference's message; a malformed number raises ``ValueError``
    (the reference's ``std::stod`` throws ``std::invalid_argument``)."""
    if not os.path.isfile(path):
        raise RuntimeError("Error: The intended file is not existed!")
    with open(path, "r", newline="") as f:
        f.readline()  # header
        rows = []
        for line in f:
            line = line.rstrip("\n")
            vals = [float(e) for e in line.split(",")]
            rows.append(vals)
    width = max((len(r) for r in rows), default=len(COLUMNS))
    if any(len(r) < 8 for r in rows):
        raise ValueError("a particle line has fewer than the 8 required fields")
    a = np.full((len(rows), width), np.nan)
    for i, r in enumerate(rows):
        a[i, :len(r)] = r
    n = len(rows)
    col = lambda k: a[:, k].copy() if n else np.zeros(0)  # noqa: E731
    p = Particle(x=col(0), y=col(1), s=col(6), vorticity=col(3),
                 isActive=(a[:, 7] != 0).astype(np.uint8) if n else np.zeros(0, np.uint8),
                 u=col(4), v=col(5))
    extra = {"gz": col(2), "level": particle_level(p.s, max_level, sigma) if n else
             np.zeros(0, np.int32),
             "chi": col(8) if width > 8 else np.zeros(n)}
    if pressure:
        if width <= 9:
            raise ValueError("no pressure column in the file")
        extra["P"] = col(9)
    return p, extra


def _g12(v: float) -> str:
    return "%.12g" % v


def save_par_state_2D(p: Particle, path: str, gz=None, chi=None, P=None,  # noqa: N802
                      ubody: float = 0.0, vbody: float = 0.0) -> None:
    """Write a 2-D particle state (state_data.cpp:335-392, type 0).

    ``gz`` defaults to ``vorticity * s**2`` (the circulation the reference
    keeps beside the vorticity), ``chi`` to zeros; the pressure column is
    written when ``P`` is given."""
    n = p.num
    gz = np.asarray(p.vorticity * p.s ** 2 if gz is None else gz, dtype=np.float64)
    chi = np.zeros(n) if chi is None else np.asarray(chi, dtype=np.float64)
    head = ",".join(COLUMNS) + (",pressure" if P is not None else "")
    lines = [head]
    for i in range(n):
        f = [_g12(p.x[i]), _g12(p.y[i]), _g12(gz[i]), _g12(p.vorticity[i]),
             _g12(p.u[i] - ubody), _g12(p.v[i] - vbody), _g12(p.s[i]),
             "1" if p.isActive[i] else "0", _g12(chi[i])]
        if P is not None:
            f.append(_g12(P[i]))
        lines.append(",".join(f))
    with open(path, "w", newline="") as fo:
        fo.write("\n".join(lines) + "\n")
