"""ctypes front-end to the f64 CPU oracle (oracle/gm_oracle.c).

TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py as the checker.  The product library never calls
into it.  Function-level citations of the reference live in gm_oracle.c.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "libgm_oracle.so")

_lib = None

_P = C.c_void_p
_I64 = C.c_int64
_D = C.c_double


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library not built: {LIB_PATH} (run `make -C oracle`)")
        L = C.CDLL(LIB_PATH)
        L.or_wrap_positions.argtypes = [_I64, _P, _P, _P, _P]
        L.or_wrap_positions.restype = None
        for name in ("or_neighbor_build_brute", "or_neighbor_build_cells"):
            f = getattr(L, name)
            f.argtypes = [_I64, _P, _P, _P, _P, _D, _I64, _P, _P, _P, _P, _I64]
            f.restype = _I64
        L.or_sur_features.argtypes = [_I64, _P, _P, _P, _P, _P, _D, _D, _D, _P]
        L.or_sur_features.restype = None
        L.or_sur_forward.argtypes = [_I64, _P, _P, _P, _P, C.c_int, _P]
        L.or_sur_forward.restype = None
        L.or_sur_loss_grad.argtypes = [_I64, _P, _P, _P, _P, _P, _P, C.c_int, _P]
        L.or_sur_loss_grad.restype = _D
        L.or_sur_shell.argtypes = [_I64, _P, _P, _P, _P, _P, _I64, _P, _P, _P, _D, _D, _D, _D, _D, _P, _P, _P, _P]
        L.or_sur_shell.restype = None
        L.or_sur_nparam.argtypes = []
        L.or_sur_nparam.restype = _I64
        L.or_needs_rebuild.argtypes = [_I64, _P, _P, _P, _P, _P, _P, _D, _D]
        L.or_needs_rebuild.restype = C.c_int
        L.or_pair_forces.argtypes = [_P, _P, _P, _P, _P, _I64, _P, _P, _D, _D, _P, _P, _P, _P]
        L.or_pair_forces.restype = _D
        L.or_bond_forces.argtypes = [_P, _P, _P, _I64, _P, _P, _P, _P, C.c_int, _P, _P, _P, _P]
        L.or_bond_forces.restype = _D
        L.or_angle_forces.argtypes = [_P, _P, _P, _I64, _P, _P, _P, _P, _P, C.c_int, _P, _P, _P, _P]
        L.or_angle_forces.restype = _D
        L.or_dihedral_forces.argtypes = [_P, _P, _P, _I64, _P, _P, _P, _P, _P, _P, _P, C.c_int,
                                         _P, _P, _P, _P]
        L.or_dihedral_forces.restype = _D
        L.or_lj14_forces.argtypes = [_P, _P, _P, _P, _P, _I64, _P, _P, _D, _P, _P, _P, _P]
        L.or_lj14_forces.restype = _D
        L.or_cpu_backend_compute.argtypes = [_P, _P, _P, _P, _P]
        L.or_cpu_backend_compute.restype = C.c_int
        L.or_half_kick.argtypes = [_I64, _P, _P, _P, _P, _P, _P, _P, _D]
        L.or_half_kick.restype = None
        L.or_drift_and_constrain.argtypes = [_I64, _P, _P, _P, _P, _P, _P, _I64, _P, _P, _P, _P,
                                             _D, C.c_int, _D]
        L.or_drift_and_constrain.restype = C.c_int
        L.or_constrain_velocities.argtypes = [_I64, _P, _P, _P, _P, _P, _P, _I64, _P, _P, _P,
                                              C.c_int, _D]
        L.or_constrain_velocities.restype = C.c_int
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None and a.size else None


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _u32(a):
    return np.ascontiguousarray(a, dtype=np.uint32)


def _box(box):
    return np.ascontiguousarray(np.broadcast_to(np.asarray(box, np.float64), (3,)))


def wrap_positions(x, y, z, box):
    """neighbor.rs:78-88, in place on float64 arrays."""
    b = _box(box)
    lib().or_wrap_positions(len(x), _p(x), _p(y), _p(z), _p(b))


def neighbor_build(x, y, z, box, r_skin, excl_i, excl_j, method="auto"):
    """Half pair list (pair_i, pair_j), i-major, as neighbor.rs:25-43 builds it
    from the given (already wrapped, if periodic) positions."""
    x, y, z = _f64(x), _f64(y), _f64(z)
    ei, ej, b = _u32(excl_i), _u32(excl_j), _box(box)
    n = len(x)
    if method == "auto":
        method = "brute" if n <= 4096 else "cells"
    f = lib().or_neighbor_build_brute if method == "brute" else lib().or_neighbor_build_cells
    cnt = f(n, _p(x), _p(y), _p(z), _p(b), float(r_skin), len(ei), _p(ei), _p(ej), None, None, 0)
    if cnt < 0:
        raise MemoryError("oracle neighbour build failed")
    pi = np.zeros(max(cnt, 1), np.uint32)
    pj = np.zeros(max(cnt, 1), np.uint32)
    cnt2 = f(n, _p(x), _p(y), _p(z), _p(b), float(r_skin), len(ei), _p(ei), _p(ej), _p(pi), _p(pj), cnt)
    assert cnt2 == cnt
    return pi[:cnt].copy(), pj[:cnt].copy()


def needs_rebuild(x, y, z, ref_x, ref_y, ref_z, r_skin, r_cutoff):
    a = [_f64(v) for v in (x, y, z, ref_x, ref_y, ref_z)]
    return bool(lib().or_needs_rebuild(len(a[0]), *[_p(v) for v in a], float(r_skin), float(r_cutoff)))


def pair_forces(x, y, z, sigma, epsilon, pi, pj, r_cutoff, r_switch, box, fx=None, fy=None, fz=None):
    """force/nonbonded.rs:20-82.  Returns (energy, fx, fy, fz); accumulates into given arrays."""
    n = len(x)
    fx = np.zeros(n) if fx is None else fx
    fy = np.zeros(n) if fy is None else fy
    fz = np.zeros(n) if fz is None else fz
    x, y, z, s, e = _f64(x), _f64(y), _f64(z), _f64(sigma), _f64(epsilon)
    pi, pj, b = _u32(pi), _u32(pj), _box(box)
    en = lib().or_pair_forces(_p(x), _p(y), _p(z), _p(s), _p(e), len(pi), _p(pi), _p(pj),
                              float(r_cutoff), float(r_switch), _p(b), _p(fx), _p(fy), _p(fz))
    return en, fx, fy, fz


def _zeros3(n, fx, fy, fz):
    return (np.zeros(n) if fx is None else fx, np.zeros(n) if fy is None else fy,
            np.zeros(n) if fz is None else fz)


def bond_forces(x, y, z, topo, min_image=False, box=(1e3, 1e3, 1e3), fx=None, fy=None, fz=None):
    fx, fy, fz = _zeros3(len(x), fx, fy, fz)
    x, y, z = _f64(x), _f64(y), _f64(z)
    a = [_u32(topo.bond_i), _u32(topo.bond_j), _f64(topo.bond_k), _f64(topo.bond_r0)]
    e = lib().or_bond_forces(_p(x), _p(y), _p(z), len(a[0]), *[_p(v) for v in a], int(min_image),
                             _p(_box(box)), _p(fx), _p(fy), _p(fz))
    return e, fx, fy, fz


def angle_forces(x, y, z, topo, min_image=False, box=(1e3, 1e3, 1e3), fx=None, fy=None, fz=None):
    fx, fy, fz = _zeros3(len(x), fx, fy, fz)
    x, y, z = _f64(x), _f64(y), _f64(z)
    a = [_u32(topo.angle_i), _u32(topo.angle_j), _u32(topo.angle_k), _f64(topo.angle_kth),
         _f64(topo.angle_th0)]
    e = lib().or_angle_forces(_p(x), _p(y), _p(z), len(a[0]), *[_p(v) for v in a], int(min_image),
                              _p(_box(box)), _p(fx), _p(fy), _p(fz))
    return e, fx, fy, fz


def dihedral_forces(x, y, z, topo, min_image=False, box=(1e3, 1e3, 1e3), fx=None, fy=None, fz=None):
    fx, fy, fz = _zeros3(len(x), fx, fy, fz)
    x, y, z = _f64(x), _f64(y), _f64(z)
    a = [_u32(topo.dihed_i), _u32(topo.dihed_j), _u32(topo.dihed_k), _u32(topo.dihed_l),
         _f64(topo.dihed_kphi), _u32(topo.dihed_n), _f64(topo.dihed_delta)]
    e = lib().or_dihedral_forces(_p(x), _p(y), _p(z), len(a[0]), *[_p(v) for v in a],
                                 int(min_image), _p(_box(box)), _p(fx), _p(fy), _p(fz))
    return e, fx, fy, fz


def lj14_forces(x, y, z, sigma, epsilon, pi, pj, scale, box, fx=None, fy=None, fz=None):
    fx, fy, fz = _zeros3(len(x), fx, fy, fz)
    x, y, z, s, e = _f64(x), _f64(y), _f64(z), _f64(sigma), _f64(epsilon)
    pi, pj = _u32(pi), _u32(pj)
    en = lib().or_lj14_forces(_p(x), _p(y), _p(z), _p(s), _p(e), len(pi), _p(pi), _p(pj),
                              float(scale), _p(_box(box)), _p(fx), _p(fy), _p(fz))
    return en, fx, fy, fz


class _OrSystem(C.Structure):
    _fields_ = [
        ("n_atoms", _I64), ("x", _P), ("y", _P), ("z", _P), ("sigma", _P), ("epsilon", _P),
        ("box", _D * 3), ("periodic", C.c_int), ("r_cutoff", _D), ("r_switch", _D),
        ("nb", _I64), ("bi", _P), ("bj", _P), ("bk", _P), ("br0", _P),
        ("na", _I64), ("ai", _P), ("aj", _P), ("ak", _P), ("akth", _P), ("ath0", _P),
        ("nd", _I64), ("di", _P), ("dj", _P), ("dk", _P), ("dl", _P), ("dkphi", _P), ("dn", _P),
        ("ddelta", _P),
        ("n14", _I64), ("p14i", _P), ("p14j", _P), ("scale14", _D),
        ("bonded_min_image", C.c_int),
        ("np", _I64), ("pi", _P), ("pj", _P),
        ("n_threads", C.c_int),
    ]


@dataclass
class Terms:
    bond: float
    angle: float
    dihedral: float
    lj: float
    lj14: float
    total: float


def cpu_backend_compute(x, y, z, sigma, epsilon, topo, pi, pj, box, r_cutoff, r_switch,
                        n_threads=1, scale14=0.0, bonded_min_image=False, periodic=True):
    """CpuBackend::compute_forces (cpu_backend.rs:92-182) over a prebuilt half
    list.  Returns (Terms, fx, fy, fz)."""
    keep = []
    arr = lambda a, f: keep.append(f(a)) or keep[-1]  # noqa: E731
    s = _OrSystem()
    n = len(x)
    s.n_atoms = n
    s.x, s.y, s.z = (_p(arr(v, _f64)) for v in (x, y, z))
    s.sigma, s.epsilon = _p(arr(sigma, _f64)), _p(arr(epsilon, _f64))
    for k in range(3):
        s.box[k] = float(_box(box)[k])
    s.periodic = int(periodic)
    s.r_cutoff, s.r_switch = float(r_cutoff), float(r_switch)
    s.nb = len(topo.bond_i)
    s.bi, s.bj = _p(arr(topo.bond_i, _u32)), _p(arr(topo.bond_j, _u32))
    s.bk, s.br0 = _p(arr(topo.bond_k, _f64)), _p(arr(topo.bond_r0, _f64))
    s.na = len(topo.angle_i)
    s.ai, s.aj, s.ak = (_p(arr(v, _u32)) for v in (topo.angle_i, topo.angle_j, topo.angle_k))
    s.akth, s.ath0 = _p(arr(topo.angle_kth, _f64)), _p(arr(topo.angle_th0, _f64))
    s.nd = len(topo.dihed_i)
    s.di, s.dj, s.dk, s.dl = (_p(arr(v, _u32)) for v in
                              (topo.dihed_i, topo.dihed_j, topo.dihed_k, topo.dihed_l))
    s.dkphi, s.dn, s.ddelta = (_p(arr(topo.dihed_kphi, _f64)), _p(arr(topo.dihed_n, _u32)),
                               _p(arr(topo.dihed_delta, _f64)))
    s.n14 = len(topo.pair14_i)
    s.p14i, s.p14j = _p(arr(topo.pair14_i, _u32)), _p(arr(topo.pair14_j, _u32))
    s.scale14 = float(scale14)
    s.bonded_min_image = int(bonded_min_image)
    s.np = len(pi)
    s.pi, s.pj = _p(arr(pi, _u32)), _p(arr(pj, _u32))
    s.n_threads = int(n_threads)
    fx, fy, fz = np.zeros(n), np.zeros(n), np.zeros(n)
    terms = np.zeros(6)
    rc = lib().or_cpu_backend_compute(C.byref(s), _p(fx), _p(fy), _p(fz), _p(terms))
    if rc != 0:
        raise RuntimeError("oracle cpu backend failed")
    return Terms(*[float(t) for t in terms]), fx, fy, fz


def half_kick(vx, vy, vz, fx, fy, fz, mass, dt_half):
    lib().or_half_kick(len(vx), _p(vx), _p(vy), _p(vz), _p(_f64(fx)), _p(_f64(fy)), _p(_f64(fz)),
                       _p(_f64(mass)), float(dt_half))


def drift_and_constrain(x, y, z, vx, vy, vz, topo, mass, dt, max_iter, tol):
    ci, cj, dsq = _u32(topo.constr_i), _u32(topo.constr_j), _f64(topo.constr_dsq)
    return lib().or_drift_and_constrain(len(x), _p(x), _p(y), _p(z), _p(vx), _p(vy), _p(vz),
                                        len(ci), _p(ci), _p(cj), _p(dsq), _p(_f64(mass)),
                                        float(dt), int(max_iter), float(tol))


def constrain_velocities(x, y, z, vx, vy, vz, topo, mass, max_iter, tol):
    ci, cj = _u32(topo.constr_i), _u32(topo.constr_j)
    return lib().or_constrain_velocities(len(x), _p(_f64(x)), _p(_f64(y)), _p(_f64(z)), _p(vx),
                                         _p(vy), _p(vz), len(ci), _p(ci), _p(cj), _p(_f64(mass)),
                                         int(max_iter), float(tol))


# ---- shell-force surrogate (no reference counterpart) -------------------------

def sur_features(x, y, z, box, cls, r_lo=7.5, r_hi=12.0, fade=0.5):
    """Brute-force f64 features, n x 96."""
    x, y, z = _f64(x), _f64(y), _f64(z)
    cls = np.ascontiguousarray(cls, dtype=np.uint8)
    out = np.zeros((len(x), 96))
    lib().or_sur_features(len(x), _p(x), _p(y), _p(z), _p(_box(box)), _p(cls), float(r_lo), float(r_hi),
                          float(fade), _p(out))
    return out


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def sur_forward(feat, fmean, finv, params, round_bf16=True):
    feat, fmean, finv, params = _f32(feat), _f32(fmean), _f32(finv), _f32(params)
    out = np.zeros((len(feat), 3))
    lib().or_sur_forward(len(feat), _p(feat), _p(fmean), _p(finv), _p(params), int(bool(round_bf16)), _p(out))
    return out


def sur_loss_grad(feat, fmean, finv, params, basis, target, round_bf16=True):
    feat, fmean, finv, params = _f32(feat), _f32(fmean), _f32(finv), _f32(params)
    basis, target = _f32(basis), _f32(target)
    g = np.zeros(int(lib().or_sur_nparam()))
    loss = lib().or_sur_loss_grad(len(feat), _p(feat), _p(fmean), _p(finv), _p(params), _p(basis), _p(target),
                                  int(bool(round_bf16)), _p(g))
    return float(loss), g


def sur_shell(x, y, z, sigma, epsilon, pi, pj, box, r_switch=7.5, r_cutoff=12.0, r_in_switch=6.5,
              r_in_cutoff=7.5):
    """Shell force (3 x n) and shell Hessian K (n x 6) over a pair list."""
    x, y, z, sigma, epsilon = _f64(x), _f64(y), _f64(z), _f64(sigma), _f64(epsilon)
    pi, pj = _u32(pi), _u32(pj)
    n = len(x)
    fx, fy, fz, K = np.zeros(n), np.zeros(n), np.zeros(n), np.zeros((n, 6))
    lib().or_sur_shell(n, _p(x), _p(y), _p(z), _p(sigma), _p(epsilon), len(pi), _p(pi), _p(pj), _p(_box(box)),
                       float(r_switch), float(r_cutoff - r_switch), float(r_in_switch),
                       float(r_in_cutoff - r_in_switch), float(r_cutoff), _p(fx), _p(fy), _p(fz), _p(K))
    return np.stack([fx, fy, fz]), K
