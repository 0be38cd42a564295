"""AMBER prmtop / inpcrd readers and DCD reader for the oracle (test infrastructure only).

Restates geodesic-io/src/prmtop.rs:182-344, inpcrd.rs:12-84 and
fixed_width.rs:4-14 (paths relative to the reference repository root), plus
the H-bond constraint promotion of geodesic-engine/src/constraint.rs:10-40 and
a reader for the DCD layout written by geodesic-io/src/dcd.rs:94-137.

One documented extension: ``pairs14`` returns the proper-dihedral 1-4 pairs,
which prmtop.rs:132-137 discards together with the sign flags.
"""
from __future__ import annotations

import struct
from dataclasses import dataclass, field

import numpy as np

AMBER_CHARGE_TO_ELEMENTARY = 18.2223  # prmtop.rs:9
AMBER_VELOCITY_TO_ANG_PER_PS = 20.455  # inpcrd.rs:7

ELEM_H, ELEM_C, ELEM_N, ELEM_O, ELEM_S, ELEM_UNKNOWN = range(6)


def _lines(text: str):
    """Rust str::lines: split on '\\n', drop one trailing '\\r' per line."""
    out = text.split("\n")
    if out and out[-1] == "":
        out.pop()
    return [ln[:-1] if ln.endswith("\r") else ln for ln in out]


def chunk_fields(raw: str, width: int, count: int):
    """fixed_width.rs:4-14: `count` fields of exactly `width` chars."""
    fields = []
    for i in range(count):
        start = i * width
        end = min(start + width, len(raw))
        fields.append(raw[start:end] if start < len(raw) else "")
    return fields


def _sections(text: str) -> dict:
    """prmtop.rs:16-47: %FLAG sections -> concatenated data text."""
    sections: dict = {}
    current = None
    for line in _lines(text):
        if line.startswith("%FLAG "):
            flag = line[len("%FLAG "):].strip()
            sections.setdefault(flag, "")
            current = flag
            continue
        if line.startswith("%"):
            continue
        if current is not None:
            sections[current] += line
    return sections


def _ints(sec, flag, width, count):
    if flag not in sec:
        raise ValueError(f"missing %FLAG {flag}")
    return [int(f.strip()) for f in chunk_fields(sec[flag], width, count)]


def _floats(sec, flag, width, count):
    if flag not in sec:
        raise ValueError(f"missing %FLAG {flag}")
    return [float(f.strip()) for f in chunk_fields(sec[flag], width, count)]


def _infer_element(name: str) -> int:
    """prmtop.rs:102-117: first alphabetic character decides."""
    for ch in name:
        if ch.isascii() and ch.isalpha():
            return {"H": ELEM_H, "C": ELEM_C, "N": ELEM_N, "O": ELEM_O, "S": ELEM_S}.get(
                ch.upper(), ELEM_UNKNOWN)
    return ELEM_UNKNOWN


@dataclass
class Topology:
    """SoA mirror of geodesic-core BondedTopology (topology.rs:4-37)."""
    bond_i: np.ndarray
    bond_j: np.ndarray
    bond_k: np.ndarray
    bond_r0: np.ndarray
    angle_i: np.ndarray
    angle_j: np.ndarray
    angle_k: np.ndarray
    angle_kth: np.ndarray
    angle_th0: np.ndarray
    dihed_i: np.ndarray
    dihed_j: np.ndarray
    dihed_k: np.ndarray
    dihed_l: np.ndarray
    dihed_kphi: np.ndarray
    dihed_n: np.ndarray
    dihed_delta: np.ndarray
    excl_i: np.ndarray
    excl_j: np.ndarray
    constr_i: np.ndarray = field(default_factory=lambda: np.zeros(0, np.uint32))
    constr_j: np.ndarray = field(default_factory=lambda: np.zeros(0, np.uint32))
    constr_dsq: np.ndarray = field(default_factory=lambda: np.zeros(0, np.float64))
    pair14_i: np.ndarray = field(default_factory=lambda: np.zeros(0, np.uint32))
    pair14_j: np.ndarray = field(default_factory=lambda: np.zeros(0, np.uint32))


@dataclass
class Atoms:
    """SoA mirror of geodesic-core AtomData (atoms.rs:3-12)."""
    sigma: np.ndarray
    epsilon: np.ndarray
    mass: np.ndarray
    charge: np.ndarray
    element: np.ndarray
    names: list


def parse_prmtop(text: str):
    """prmtop.rs:182-344.  Returns (Atoms, Topology)."""
    sec = _sections(text)
    ptr = _ints(sec, "POINTERS", 8, 20)
    natom, ntypes, nbonh, ntheth, nphih = ptr[0], ptr[1], ptr[2], ptr[4], ptr[6]
    nnb, nres, nbona, ntheta, nphia = ptr[10], ptr[11], ptr[12], ptr[13], ptr[14]
    numbnd, numang, nptra = ptr[15], ptr[16], ptr[17]

    names = chunk_fields(sec["ATOM_NAME"], 4, natom)
    names = [(n + "    ")[:4] for n in names]
    charge = np.array(_floats(sec, "CHARGE", 16, natom)) / AMBER_CHARGE_TO_ELEMENTARY
    mass = np.array(_floats(sec, "MASS", 16, natom), dtype=np.float64)
    type_index = _ints(sec, "ATOM_TYPE_INDEX", 8, natom)
    n_excluded = _ints(sec, "NUMBER_EXCLUDED_ATOMS", 8, natom)
    nb_parm_index = _ints(sec, "NONBONDED_PARM_INDEX", 8, ntypes * ntypes)
    n_lj = ntypes * (ntypes + 1) // 2
    acoef = _floats(sec, "LENNARD_JONES_ACOEF", 16, n_lj)
    bcoef = _floats(sec, "LENNARD_JONES_BCOEF", 16, n_lj)
    _ints(sec, "RESIDUE_POINTER", 8, nres)
    excluded_list = _ints(sec, "EXCLUDED_ATOMS_LIST", 8, nnb)

    sigma = np.zeros(natom)
    epsilon = np.zeros(natom)
    for i, ty in enumerate(type_index):
        ico = nb_parm_index[ntypes * (ty - 1) + ty - 1]
        if ico <= 0:
            raise ValueError(f"atom {i} uses a 10-12 H-bond LJ type ({ico}); not supported")
        a, b = acoef[ico - 1], bcoef[ico - 1]
        if a == 0.0 or b == 0.0:
            sigma[i] = 0.0
            epsilon[i] = 0.0
        else:
            sigma[i] = (a / b) ** (1.0 / 6.0)  # prmtop.rs:234
            epsilon[i] = (b * b) / (4.0 * a)  # prmtop.rs:235
    atoms = Atoms(sigma, epsilon, mass, charge,
                  np.array([_infer_element(n) for n in names], dtype=np.int32), names)

    def chunks(raw, w):
        return [raw[k:k + w] for k in range(0, len(raw), w)]

    bonds = chunks(_ints(sec, "BONDS_INC_HYDROGEN", 8, 3 * nbonh), 3) + \
        chunks(_ints(sec, "BONDS_WITHOUT_HYDROGEN", 8, 3 * nbona), 3)
    angles = chunks(_ints(sec, "ANGLES_INC_HYDROGEN", 8, 4 * ntheth), 4) + \
        chunks(_ints(sec, "ANGLES_WITHOUT_HYDROGEN", 8, 4 * ntheta), 4)
    diheds = chunks(_ints(sec, "DIHEDRALS_INC_HYDROGEN", 8, 5 * nphih), 5) + \
        chunks(_ints(sec, "DIHEDRALS_WITHOUT_HYDROGEN", 8, 5 * nphia), 5)
    bk = _floats(sec, "BOND_FORCE_CONSTANT", 16, numbnd)
    br0 = _floats(sec, "BOND_EQUIL_VALUE", 16, numbnd)
    ak = _floats(sec, "ANGLE_FORCE_CONSTANT", 16, numang)
    ath0 = _floats(sec, "ANGLE_EQUIL_VALUE", 16, numang)
    dk = _floats(sec, "DIHEDRAL_FORCE_CONSTANT", 16, nptra)
    dper = _floats(sec, "DIHEDRAL_PERIODICITY", 16, nptra)
    dph = _floats(sec, "DIHEDRAL_PHASE", 16, nptra)

    def trunc3(v):  # Rust i64 division truncates toward zero
        return int(abs(v) // 3) * (1 if v >= 0 else -1)

    u32 = lambda a: np.array(a, dtype=np.uint32)  # noqa: E731
    f64 = lambda a: np.array(a, dtype=np.float64)  # noqa: E731

    def rust_round(v):  # half away from zero
        return int(np.floor(abs(v) + 0.5)) * (1 if v >= 0 else -1)

    excl_i, excl_j, cursor = [], [], 0
    for i, cnt in enumerate(n_excluded):  # prmtop.rs:303-317
        for _ in range(cnt):
            raw = excluded_list[cursor]
            cursor += 1
            if raw == 0:
                continue
            excl_i.append(i)
            excl_j.append(raw - 1)

    p14 = set()
    for c in diheds:  # extension: proper dihedrals whose 1-4 flag is not set
        if c[2] >= 0 and c[3] >= 0:
            a, b = abs(c[0]) // 3, abs(c[3]) // 3
            p14.add((min(a, b), max(a, b)))
    p14 = sorted(p14)

    topo = Topology(
        bond_i=u32([trunc3(c[0]) for c in bonds]), bond_j=u32([trunc3(c[1]) for c in bonds]),
        bond_k=f64([bk[c[2] - 1] for c in bonds]), bond_r0=f64([br0[c[2] - 1] for c in bonds]),
        angle_i=u32([trunc3(c[0]) for c in angles]), angle_j=u32([trunc3(c[1]) for c in angles]),
        angle_k=u32([trunc3(c[2]) for c in angles]),
        angle_kth=f64([ak[c[3] - 1] for c in angles]), angle_th0=f64([ath0[c[3] - 1] for c in angles]),
        dihed_i=u32([abs(c[0]) // 3 for c in diheds]), dihed_j=u32([abs(c[1]) // 3 for c in diheds]),
        dihed_k=u32([abs(c[2]) // 3 for c in diheds]), dihed_l=u32([abs(c[3]) // 3 for c in diheds]),
        dihed_kphi=f64([dk[c[4] - 1] for c in diheds]),
        dihed_n=u32([rust_round(dper[c[4] - 1]) for c in diheds]),
        dihed_delta=f64([dph[c[4] - 1] for c in diheds]),
        excl_i=u32(excl_i), excl_j=u32(excl_j),
        pair14_i=u32([p[0] for p in p14]), pair14_j=u32([p[1] for p in p14]),
    )
    return atoms, topo


def parse_inpcrd(text: str, expected_n_atoms: int, has_box: bool = False):
    """inpcrd.rs:12-84.  Returns (x, y, z, vx, vy, vz) float64 arrays."""
    lines = _lines(text)
    if len(lines) < 2:
        raise ValueError("inpcrd has fewer than 2 lines")
    natom = int(lines[1].split()[0])
    if natom != expected_n_atoms:
        raise ValueError(f"inpcrd declares {natom} atoms but prmtop declares {expected_n_atoms}")
    nvals = natom * 3
    ncl = (nvals + 5) // 6
    data = lines[2:]
    if len(data) < ncl:
        raise ValueError("too few coordinate lines")
    coords = np.array([float(f.strip()) for f in chunk_fields("".join(data[:ncl]), 12, nvals)])
    x, y, z = coords[0::3].copy(), coords[1::3].copy(), coords[2::3].copy()
    vx, vy, vz = np.zeros(natom), np.zeros(natom), np.zeros(natom)
    rem = data[ncl:]
    avail = max(0, len(rem) - (1 if has_box else 0))
    if avail >= ncl and ncl > 0:
        v = np.array([float(f.strip()) for f in chunk_fields("".join(rem[:ncl]), 12, nvals)])
        vx, vy, vz = (v[0::3] * AMBER_VELOCITY_TO_ANG_PER_PS, v[1::3] * AMBER_VELOCITY_TO_ANG_PER_PS,
                      v[2::3] * AMBER_VELOCITY_TO_ANG_PER_PS)
    return x, y, z, vx, vy, vz


def promote_hydrogen_bonds(topo: Topology, atoms: Atoms) -> None:
    """constraint.rs:10-40: X-H bonds become constraints with d0^2 = r0*r0."""
    keep, ci, cj, dsq = [], [], [], []
    for idx in range(len(topo.bond_i)):
        i, j = int(topo.bond_i[idx]), int(topo.bond_j[idx])
        r0 = float(topo.bond_r0[idx])
        if atoms.element[i] == ELEM_H or atoms.element[j] == ELEM_H:
            ci.append(i)
            cj.append(j)
            dsq.append(r0 * r0)
        else:
            keep.append(idx)
    keep = np.array(keep, dtype=np.int64)
    topo.constr_i = np.concatenate([topo.constr_i, np.array(ci, np.uint32)])
    topo.constr_j = np.concatenate([topo.constr_j, np.array(cj, np.uint32)])
    topo.constr_dsq = np.concatenate([topo.constr_dsq, np.array(dsq, np.float64)])
    topo.bond_i = topo.bond_i[keep]
    topo.bond_j = topo.bond_j[keep]
    topo.bond_k = topo.bond_k[keep]
    topo.bond_r0 = topo.bond_r0[keep]


def read_dcd(data: bytes):
    """Reads the layout of dcd.rs:94-137.  Returns (header dict, frames[nf,3,natom] f32)."""
    (m0,) = struct.unpack_from("<i", data, 0)
    if m0 != 84 or data[4:8] != b"CORD":
        raise ValueError("not a CORD DCD")
    icntrl = struct.unpack_from("<20i", data, 8)
    delta = struct.unpack_from("<f", data, 8 + 9 * 4)[0]
    off = 8 + 80 + 4
    (tlen,) = struct.unpack_from("<i", data, off)
    off += 4 + tlen + 4
    (_, natom, _) = struct.unpack_from("<3i", data, off)
    off += 12
    nset = icntrl[0]
    frames = np.zeros((nset, 3, natom), dtype=np.float32)
    for f in range(nset):
        for c in range(3):
            (nb,) = struct.unpack_from("<i", data, off)
            off += 4
            frames[f, c] = np.frombuffer(data, dtype="<f4", count=natom, offset=off)
            off += nb + 4
    return {"nset": nset, "nsavc": icntrl[2], "delta": delta, "natom": natom}, frames
