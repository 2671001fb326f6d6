"""On-disk formats of the Bulk_3D_FFT drivers: input files, restart dumps and result files.

Every reader follows the Fortran statement that reads the file (list-directed ``read(u,*)``:
one READ consumes one record and ignores what is left of it, ``d``/``D`` exponents, ``r*c``
repeat counts, ``,`` or blanks as separators, ``/`` ends the list); every writer produces text
that the driver's own READ accepts and that the readers here parse back exactly (17
significant digits, the round-trip precision of a double).

Arrays use the layout of the rest of the package: a field set ``field(gridx,gridy,gridz,k)``
of the Fortran code is the C-ordered ``(k, nz, ny, nx)`` numpy array with the same memory.
"""
from __future__ import annotations

import dataclasses
import os
from typing import Iterable, List, Sequence

import numpy as np

# ---------------------------------------------------------------------------
# list-directed values
# ---------------------------------------------------------------------------


def _split_values(record: str) -> List[str]:
    """Values of one list-directed record (a ``/`` ends the list, ``r*c`` repeats c)."""
    out: List[str] = []
    cur = []
    for ch in record:
        if ch in " \t,\r\n":
            if cur:
                out.append("".join(cur))
                cur = []
        elif ch == "/":
            break
        else:
            cur.append(ch)
    if cur:
        out.append("".join(cur))
    expanded: List[str] = []
    for tok in out:
        head, star, tail = tok.partition("*")
        if star and head.isdigit():
            expanded.extend([tail] * int(head))
        else:
            expanded.append(tok)
    return expanded


def parse_real(tok: str) -> float:
    """A Fortran real constant: ``1.0d0``, ``2.5D-3``, ``5.0E-002``, ``3.``, ``-7``."""
    t = tok.strip().replace("d", "e").replace("D", "e")
    return float(t)


def parse_int(tok: str) -> int:
    return int(tok.strip())


def parse_logical(tok: str) -> bool:
    """A list-directed logical: an optional ``.`` then ``t``/``f`` (``.true.``, ``T``, ``.f``)."""
    t = tok.strip().lower().lstrip(".")
    if t.startswith("t"):
        return True
    if t.startswith("f"):
        return False
    raise ValueError(f"not a Fortran logical: {tok!r}")


_PARSERS = {float: parse_real, int: parse_int, bool: parse_logical}


class RecordReader:
    """Sequential list-directed reads from one text file, as ``read(u,*) a, b, ...`` does:
    each call starts at a new record, takes values across records until it has all it
    asked for, and drops the rest of the last record it touched."""

    def __init__(self, path: str):
        with open(path) as f:
            self.records = f.read().split("\n")
        self.pos = 0
        self.path = path

    def read(self, *types):
        """``read(u,*)`` of len(types) items (no types: skip one record)."""
        if not types:
            self._next()
            return None
        vals: List[str] = []
        while len(vals) < len(types):
            vals.extend(_split_values(self._next()))
        out = [_PARSERS[t](v) for t, v in zip(types, vals)]
        return out[0] if len(types) == 1 else tuple(out)

    def _next(self) -> str:
        if self.pos >= len(self.records):
            raise EOFError(f"{self.path}: end of file after {self.pos} records")
        r = self.records[self.pos]
        self.pos += 1
        return r


def format_real(v: float) -> str:
    """17 significant digits with an ``E`` exponent (exact round trip of a double)."""
    return f"{float(v):.16E}"


# ---------------------------------------------------------------------------
# input files
# ---------------------------------------------------------------------------


@dataclasses.dataclass
class DiblockInput:
    """Main_Diblock_Melt/input_general (main_diblock.f90:235-247)."""
    sizex: float
    sizey: float
    sizez: float
    C: float
    mu: float
    ds: float
    itermax_scf: int
    accuracy_scf: float
    lam: float


@dataclasses.dataclass
class BlendInput:
    """Main_Blend/input_general (main_blend.f90:287-305): per polymer type the number of
    chains, mu and chain length."""
    sizex: float
    sizey: float
    sizez: float
    number_of_chains: List[float]
    mu: List[float]
    chain_length: List[float]
    ds: float
    number_of_timesteps: int
    timestep: float
    qread: bool
    qinit: int
    qepd: bool


@dataclasses.dataclass
class MultiblockInput:
    """Main_Multiblock/input_general (main_multiblock.f90:407-425)."""
    sizex: float
    sizey: float
    sizez: float
    C: float
    mu: float
    ds: float
    start_time: float
    number_of_timesteps: int
    timestep: float
    qread: bool
    qinit: int
    qddft: bool
    wcutoff: float


def read_input_general(path: str, driver: str, polymer_types: int = 2):
    """input_general of ``driver`` in ('diblock', 'blend', 'multiblock')."""
    r = RecordReader(path)
    if driver == "diblock":
        sx, sy, sz = r.read(float), r.read(float), r.read(float)
        r.read()
        C, mu = r.read(float), r.read(float)
        r.read()
        ds = r.read(float)
        r.read()
        itmax, acc = r.read(int), r.read(float)
        r.read()
        lam = r.read(float)
        return DiblockInput(sx, sy, sz, C, mu, ds, itmax, acc, lam)
    if driver == "blend":
        sx, sy, sz = r.read(float), r.read(float), r.read(float)
        r.read()
        nch, mus, lens = [], [], []
        for _ in range(polymer_types):
            nch.append(r.read(float))
            mus.append(r.read(float))
            lens.append(r.read(float))
            r.read()
        ds = r.read(float)
        r.read()
        nts, dt = r.read(int), r.read(float)
        r.read()
        qread, qinit, qepd = r.read(bool), r.read(int), r.read(bool)
        return BlendInput(sx, sy, sz, nch, mus, lens, ds, nts, dt, qread, qinit, qepd)
    if driver == "multiblock":
        sx, sy, sz = r.read(float), r.read(float), r.read(float)
        r.read()
        C, mu = r.read(float), r.read(float)
        r.read()
        ds = r.read(float)
        r.read()
        t0, nts, dt = r.read(float), r.read(int), r.read(float)
        r.read()
        qread, qinit, qddft = r.read(bool), r.read(int), r.read(bool)
        r.read()
        wcut = r.read(float)
        return MultiblockInput(sx, sy, sz, C, mu, ds, t0, nts, dt, qread, qinit, qddft, wcut)
    raise ValueError(f"unknown driver {driver!r}")


def _fmt_logical(b: bool) -> str:
    return ".true." if b else ".false."


def write_input_general(path: str, inp) -> None:
    """Write an input_general that the driver's read_input_parameters reads back as ``inp``
    (one value per record, blank separator records where the driver skips one)."""
    R = format_real
    if isinstance(inp, DiblockInput):
        lines = [R(inp.sizex), R(inp.sizey), R(inp.sizez), "", R(inp.C), R(inp.mu), "", R(inp.ds), "",
                 str(inp.itermax_scf), R(inp.accuracy_scf), "", R(inp.lam)]
    elif isinstance(inp, BlendInput):
        lines = [R(inp.sizex), R(inp.sizey), R(inp.sizez), ""]
        for n, m, l in zip(inp.number_of_chains, inp.mu, inp.chain_length):
            lines += [R(n), R(m), R(l), ""]
        lines += [R(inp.ds), "", str(inp.number_of_timesteps), R(inp.timestep), "", _fmt_logical(inp.qread),
                  str(inp.qinit), _fmt_logical(inp.qepd)]
    elif isinstance(inp, MultiblockInput):
        lines = [R(inp.sizex), R(inp.sizey), R(inp.sizez), "", R(inp.C), R(inp.mu), "", R(inp.ds), "",
                 R(inp.start_time), str(inp.number_of_timesteps), R(inp.timestep), "", _fmt_logical(inp.qread),
                 str(inp.qinit), _fmt_logical(inp.qddft), "", R(inp.wcutoff)]
    else:
        raise TypeError(type(inp))
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


@dataclasses.dataclass
class Interactions:
    """input_interactions (Interaction_Modules/mod_FloryHuggins.f90:37-66): kappa, one record
    per chi(c, c2) with c < c2 in row order, the monomer volumes.  chi is the full symmetric
    matrix with a zero diagonal."""
    kappa: float
    chi: np.ndarray
    volume: List[float]


def read_interactions(path: str, components: int = 2) -> Interactions:
    r = RecordReader(path)
    kappa = r.read(float)
    r.read()
    chi = np.zeros((components, components))
    for c in range(components):
        for c2 in range(c + 1, components):
            chi[c, c2] = chi[c2, c] = r.read(float)
    r.read()
    vols = [r.read(float) for _ in range(components)]
    return Interactions(kappa, chi, vols)


def write_interactions(path: str, inter: Interactions) -> None:
    chi = np.asarray(inter.chi)
    n = len(inter.volume)
    lines = [format_real(inter.kappa), ""]
    lines += [format_real(chi[c, c2]) for c in range(n) for c2 in range(c + 1, n)]
    lines += [""] + [format_real(v) for v in inter.volume]
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


NQMAX = 1000  # main_multiblock.f90:447


@dataclasses.dataclass
class Mobilities:
    """input_mobilities (main_multiblock.f90:441-480).  table[t2, t1, iq-1] is the Fortran
    mobility_table(iq, t1, t2), i.e. the C view of mobility_table(nq, 2, 2)."""
    nq: int
    dq: float
    table: np.ndarray


def read_mobilities(path: str) -> Mobilities:
    """nq, dq, three skipped records, then nq records ``q, L_AA, L_AB, L_BB``; q must equal
    iq*dq within 1e-6 and nq may not exceed 1000 (the reference stops on either);
    L_BA = L_AB (main_multiblock.f90:456-477)."""
    r = RecordReader(path)
    nq = r.read(int)
    if nq > NQMAX:
        raise ValueError("too many entries in mobility file -> adjust nqmax")
    dq = r.read(float)
    r.read()
    r.read()
    r.read()
    table = np.empty((2, 2, nq))
    for iq in range(1, nq + 1):
        q, l11, l12, l22 = r.read(float, float, float, float)
        if abs(q - iq * dq) >= 1e-6:
            raise ValueError(f"q values in mobility are not consistent: {q} {iq * dq}")
        table[0, 0, iq - 1] = l11
        table[1, 0, iq - 1] = l12  # mobility_table(iq, 1, 2)
        table[1, 1, iq - 1] = l22
        table[0, 1, iq - 1] = l12  # mobility_table(:,2,1) = mobility_table(:,1,2)
    return Mobilities(nq, dq, table)


def write_mobilities(path: str, mob: Mobilities) -> None:
    R = format_real
    lines = [f"{mob.nq}     ! number of entries nq", f"{R(mob.dq)}    ! delta q", "",
             "! Description of entries: q, LAA, LAB, LBB", ""]
    for iq in range(1, mob.nq + 1):
        t = mob.table[:, :, iq - 1]
        lines.append(" ".join(R(v) for v in (iq * mob.dq, t[0, 0], t[1, 0], t[1, 1])))
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


# ---------------------------------------------------------------------------
# restart dumps: cnf.in / cnf.out
# ---------------------------------------------------------------------------


def read_cnf(path: str, shape: Sequence[int]) -> np.ndarray:
    """``read(33,*) field`` (main_multiblock.f90:169, main_blend.f90:134): all
    nx*ny*nz*k values in Fortran order.  shape: the C shape (k, nz, ny, nx)."""
    n = int(np.prod(shape))
    vals: List[float] = []
    with open(path) as f:
        for rec in f:
            for tok in _split_values(rec):
                vals.append(parse_real(tok))
                if len(vals) == n:
                    return np.array(vals).reshape(shape)
    raise EOFError(f"{path}: {len(vals)} values, {n} needed")


def write_cnf(path: str, field: np.ndarray, per_record: int = 4) -> None:
    """``write(33,*) field`` (main_multiblock.f90:302, 370; main_blend.f90:252): the values
    in Fortran order, several per record."""
    v = np.ascontiguousarray(field, dtype=np.float64).ravel()
    with open(path, "w") as f:
        for i in range(0, v.size, per_record):
            f.write(" " + " ".join(format_real(x) for x in v[i:i + per_record]) + "\n")


# ---------------------------------------------------------------------------
# result files
# ---------------------------------------------------------------------------


def _rows(path: str, width: int) -> np.ndarray:
    """The file's values as rows of ``width``.  List-directed output may continue one WRITE
    on the next record (flang starts a new record when an item would pass column 80), so
    values are taken in sequence across records; records that start with ``#`` (headers)
    and blank records carry no row values."""
    vals: List[float] = []
    with open(path) as f:
        for rec in f:
            t = _split_values(rec)
            if t and not t[0].startswith("#"):
                vals.extend(parse_real(x) for x in t)
    if len(vals) % width:
        raise ValueError(f"{path}: {len(vals)} values do not make rows of {width}")
    return np.array(vals).reshape(-1, width)


def read_profile(path: str, shape: Sequence[int], order: str, index_columns: int = 3) -> np.ndarray:
    """A ``x y z value_1 .. value_k`` profile with a blank record after each sweep of the
    fastest loop.  order gives the loop nesting the driver writes with, fastest last:
    'zyx' for x fastest (saddle_dens.dat, main_diblock.f90:164-172; dens.dat,
    main_blend.f90:236-250), 'xyz' for z fastest (initial_/final_dens.dat,
    main_multiblock.f90:232-250, 345-367, whose '# Energy' header is skipped; profiles.dat,
    main_multiblock.f90:288-298, with the time in front).
    index_columns: 3 (coordinates) or 4 (profiles.dat, with the time first).
    Returns (k, nz, ny, nx)."""
    k, nz, ny, nx = shape
    a = _rows(path, index_columns + k)
    if a.shape != (nx * ny * nz, index_columns + k):
        raise ValueError(f"{path}: {a.shape} rows, expected {(nx * ny * nz, index_columns + k)}")
    vals = a[:, index_columns:]
    if order == "zyx":
        return np.ascontiguousarray(vals.T.reshape(k, nz, ny, nx))
    if order == "xyz":
        return np.ascontiguousarray(vals.T.reshape(k, nx, ny, nz).transpose(0, 3, 2, 1))
    raise ValueError(order)


def write_profile(path: str, values: np.ndarray, coords: Iterable[np.ndarray], order: str,
                  header: Sequence[str] = (), leading: Sequence[float] = ()) -> None:
    """Write a profile as the drivers do: one record per point with the coordinates (or
    indices) of coords = (x, y, z) arrays per axis, then the k values, and a blank record
    after each sweep of the fastest loop (order as in read_profile)."""
    v = np.asarray(values)
    k, nz, ny, nx = v.shape
    cx, cy, cz = (np.asarray(c) for c in coords)
    lead = " ".join(format_real(t) for t in leading)
    lines = list(header)
    if order == "zyx":
        for z in range(nz):
            for y in range(ny):
                for x in range(nx):
                    lines.append(" ".join(filter(None, [lead, format_real(cx[x]), format_real(cy[y]),
                                                        format_real(cz[z])] + [format_real(v[t, z, y, x])
                                                                               for t in range(k)])))
                lines.append("")
    elif order == "xyz":
        for x in range(nx):
            for y in range(ny):
                for z in range(nz):
                    lines.append(" ".join(filter(None, [lead, format_real(cx[x]), format_real(cy[y]),
                                                        format_real(cz[z])] + [format_real(v[t, z, y, x])
                                                                               for t in range(k)])))
            lines.append("")
    else:
        raise ValueError(order)
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def read_energy_rows(path: str, width: int) -> np.ndarray:
    """energy_*.dat (main_diblock.f90:182: sizex sizey sizez energy rho, width 5) and
    energy.dat (main_multiblock.f90:275: time energy rho mu, width 4)."""
    return _rows(path, width)


def exists_all(dirpath: str, names: Iterable[str]) -> bool:
    return all(os.path.exists(os.path.join(dirpath, n)) for n in names)
