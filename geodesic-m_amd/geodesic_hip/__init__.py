"""Python front-end of libgeodesic_hip.so (the gfx950 force engine).

Thin ctypes layer over the C-ABI in include/geodesic_hip.h, plus `GpuBackend`,
a mirror of the reference's `ComputeBackend` implementation
(geodesic-core/src/backend.rs:11-23, geodesic-gpu/src/gpu_backend.rs:11-100;
paths relative to the reference repository root) so host code reads like the
reference's own backend calls.

There is no CPU fallback: if the library or a gfx950 device is missing, every
entry point raises.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("GEODESIC_HIP_LIB") or os.path.join(os.path.dirname(_HERE), "lib",
                                                                "libgeodesic_hip.so")
HEADER_PATH = os.path.join(os.path.dirname(os.path.dirname(_HERE)), "include", "geodesic_hip.h")

GM_OK = 0
GM_ERR_INVALID_ARG = 5
GM_ERR_CONVERGENCE = 9
STATUS_NAMES = {
    0: "GM_OK", 1: "GM_ERR_DEVICE_LOST", 2: "GM_ERR_KERNEL_LOAD", 3: "GM_ERR_OOM",
    4: "GM_ERR_NO_DEVICE", 5: "GM_ERR_INVALID_ARG", 6: "GM_ERR_KERNEL", 7: "GM_ERR_NOT_BUILT",
    8: "GM_ERR_CAPACITY", 9: "GM_ERR_CONVERGENCE", 10: "GMH_ERR_IO", 11: "GMH_ERR_CONFIG",
    12: "GM_ERR_NUMERICAL",
}
GM_ERR_NUMERICAL = 12

_P = C.c_void_p
_I64 = C.c_int64
_D = C.c_double


class GmError(RuntimeError):
    """A non-zero gm_status.  `.status` holds the code (see STATUS_NAMES);
    codes 1-4 correspond to geodesic-core BackendError (error.rs:82-91)."""

    def __init__(self, status, msg):
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {msg}")
        self.status = status
        self.result = None  # GM_ERR_NUMERICAL: the outputs, written all the same


def _check_numerical(rc, result):
    """_check, except that GM_ERR_NUMERICAL (a non-finite force; outputs are
    written, those atoms hold NaN) raises with the outputs attached."""
    if rc == GM_ERR_NUMERICAL:
        err = GmError(rc, lib().gm_last_error().decode())
        err.result = result
        raise err
    _check(rc)


class SystemDesc(C.Structure):
    _fields_ = [
        ("n_atoms", _I64), ("sigma", _P), ("epsilon", _P), ("box", _D * 3), ("periodic", C.c_int32),
        ("r_cutoff", _D), ("r_switch", _D), ("r_skin", _D), ("scale_14", _D),
        ("bonded_min_image", C.c_int32),
        ("n_bonds", _I64), ("bond_i", _P), ("bond_j", _P), ("bond_k", _P), ("bond_r0", _P),
        ("n_angles", _I64), ("angle_i", _P), ("angle_j", _P), ("angle_k", _P), ("angle_kth", _P),
        ("angle_th0", _P),
        ("n_dihedrals", _I64), ("dihed_i", _P), ("dihed_j", _P), ("dihed_k", _P), ("dihed_l", _P),
        ("dihed_kphi", _P), ("dihed_n", _P), ("dihed_delta", _P),
        ("n_excl", _I64), ("excl_i", _P), ("excl_j", _P),
        ("n_pairs14", _I64), ("pair14_i", _P), ("pair14_j", _P),
    ]


class EnergyTerms(C.Structure):
    _fields_ = [("bond", _D), ("angle", _D), ("dihedral", _D), ("lj", _D), ("lj14", _D), ("total", _D)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class Profile(C.Structure):
    _fields_ = [("n_evals", _I64), ("n_untimed", _I64), ("ms_nonbonded_total", _D),
                ("ms_nonbonded_min", _D), ("ms_nonbonded_max", _D), ("ms_eval_total", _D),
                ("ms_prep_total", _D)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class Stats(C.Structure):
    _fields_ = [("n_atoms", _I64), ("n_superclusters", _I64), ("n_clusters", _I64),
                ("n_entries", _I64), ("n_tiles", _I64), ("n_tiles_cutoff", _I64), ("n_list_pairs", _I64),
                ("n_pairs_cutoff", _I64),
                ("n_work_items", _I64), ("ms_nonbonded", C.c_float), ("ms_eval", C.c_float),
                ("ms_build", C.c_float), ("n_partner_entries", _I64), ("pair_layout", C.c_int32),
                ("n_lj_types", C.c_int32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_lib = None


class SurrogateDesc(C.Structure):
    _fields_ = [("r_inner_cutoff", _D), ("r_inner_switch", _D), ("r_inner_skin", _D),
                ("r_lo", _D), ("r_hi", _D), ("fade", _D), ("atom_class", _P), ("seed", C.c_uint64)]


class TrainDesc(C.Structure):
    _fields_ = [("n_frames", _I64), ("mu", _D), ("sigma", _D), ("cap", _D), ("seed", C.c_uint64),
                ("lr", C.c_float), ("batch", _I64), ("update", C.c_int32)]


class TrainReport(C.Structure):
    _fields_ = [("loss_first", _D), ("loss_last", _D), ("loss_net", _D), ("loss_stale", _D),
                ("loss_taylor", _D), ("ms_total", _D), ("frames", _I64), ("steps", _I64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class MdDesc(C.Structure):
    _fields_ = [("n_steps", _I64), ("dt", _D), ("frame_interval", C.c_int32), ("max_iter", C.c_int32),
                ("tol", _D), ("first_step", _I64), ("resume", C.c_int32)]


class MdFailure(C.Structure):
    """gm_md_failure: kind 0 none, 1 NanInPos, 2 NanInForce; component 0..2 = X, Y, Z."""
    _fields_ = [("kind", C.c_int32), ("component", C.c_int32), ("step", _I64), ("atom", _I64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


SUR_NPARAM = 29315
SUR_NFEAT = 96
MTS_HELD, MTS_NET, MTS_TAYLOR = 0, 1, 2
# Brownian trajectories at 300 K: kT = 0.5962 kcal/mol, 0.02 A of noise per
# coordinate per step, mobility mu = sigma^2 / (2 kT), steps capped at 0.1 A.
BD_SIGMA = 0.02
BD_MU = BD_SIGMA ** 2 / (2.0 * 0.5962)
BD_CAP = 0.1


def lib():
    """Load libgeodesic_hip.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"libgeodesic_hip.so not built at {LIB_PATH}; run __graft_entry__.build()")
        L = C.CDLL(LIB_PATH)
        sig = {
            "gm_init": ([C.c_int, C.POINTER(_P)], C.c_int),
            "gm_ctx_destroy": ([_P], C.c_int),
            "gm_last_error": ([], C.c_char_p),
            "gm_device_count": ([], C.c_int),
            "gm_system_create": ([_P, C.POINTER(SystemDesc), C.POINTER(_P)], C.c_int),
            "gm_system_destroy": ([_P], C.c_int),
            "gm_build_neighbor_list": ([_P, _P, _P, _P], C.c_int),
            "gm_needs_rebuild": ([_P, _P, _P, _P, C.POINTER(C.c_int32)], C.c_int),
            "gm_compute_forces": ([_P, _P, _P, _P, _P, _P, _P, C.POINTER(EnergyTerms)], C.c_int),
            "gm_get_pair_list": ([_P, _P, _P, _I64, C.POINTER(_I64)], C.c_int),
            "gm_build_neighbor_list_device": ([_P, _P, _P, _P], C.c_int),
            "gm_compute_forces_device": ([_P, _P, _P, _P, _P, _P, _P], C.c_int),
            "gm_sync": ([_P, C.POINTER(EnergyTerms)], C.c_int),
            "gm_get_stats": ([_P, C.POINTER(Stats)], C.c_int),
            "gm_profile_begin": ([_P, C.c_int32], C.c_int),
            "gm_profile_end": ([_P, C.POINTER(Profile)], C.c_int),
            "gm_set_work_chunk": ([_P, C.c_int32], C.c_int),
            "gm_set_tile_pruning": ([_P, C.c_int32], C.c_int),
            "gm_set_pair_layout": ([_P, C.c_int32], C.c_int),
            "gm_set_positions": ([_P, _P, _P, _P], C.c_int),
            "gm_compute_forces_resident": ([_P], C.c_int),
            "gm_get_forces": ([_P, _P, _P, _P], C.c_int),
            "gm_device_alloc": ([_P, C.c_size_t, C.POINTER(_P)], C.c_int),
            "gm_device_free": ([_P, _P], C.c_int),
            "gm_copy_to_device": ([_P, _P, _P, C.c_size_t], C.c_int),
            "gm_copy_to_host": ([_P, _P, _P, C.c_size_t], C.c_int),
            "gm_batch_create": ([_P, _P, C.c_int32, C.POINTER(_P)], C.c_int),
            "gm_batch_destroy": ([_P], C.c_int),
            "gm_batch_replica": ([_P, C.c_int32, C.POINTER(_P)], C.c_int),
            "gm_batch_build_neighbor_lists": ([_P, _P, _P, _P], C.c_int),
            "gm_batch_set_positions": ([_P, _P, _P, _P], C.c_int),
            "gm_batch_compute": ([_P], C.c_int),
            "gm_batch_sync": ([_P, _P, C.POINTER(C.c_float)], C.c_int),
            "gm_batch_get_forces": ([_P, _P, _P, _P], C.c_int),
            "gm_batch_set_shape": ([_P, C.c_int32, C.c_int32], C.c_int),
            "gm_batch_set_pair_layout": ([_P, C.c_int32], C.c_int),
            "gm_batch_get_pair_layout": ([_P, C.POINTER(C.c_int32)], C.c_int),
            "gm_batch_autotune": ([_P, C.POINTER(C.c_int32), C.POINTER(C.c_int32), C.POINTER(C.c_float)],
                                  C.c_int),
            "gm_surrogate_create": ([_P, C.POINTER(SystemDesc), C.POINTER(SurrogateDesc), C.POINTER(_P)],
                                    C.c_int),
            "gm_surrogate_destroy": ([_P], C.c_int),
            "gm_surrogate_system": ([_P, C.c_int32, C.POINTER(_P)], C.c_int),
            "gm_surrogate_build": ([_P, _P, _P, _P], C.c_int),
            "gm_surrogate_features": ([_P, _P, _P, _P, _P], C.c_int),
            "gm_surrogate_set_params": ([_P, _P, _P, _P, _P], C.c_int),
            "gm_surrogate_get_params": ([_P, _P, _P, _P, _P], C.c_int),
            "gm_surrogate_infer": ([_P, _P, _I64, _P], C.c_int),
            "gm_surrogate_loss_grad": ([_P, _P, _P, _P, _I64, C.POINTER(C.c_double), _P], C.c_int),
            "gm_surrogate_train": ([_P, _P, _P, _P, C.POINTER(TrainDesc), _P, C.POINTER(TrainReport)], C.c_int),
            "gm_mts_reset": ([_P], C.c_int),
            "gm_mts_step": ([_P, _P, _P, _P, _P, _P, _P, C.c_int32, C.POINTER(C.c_int32)], C.c_int),
            "gm_surrogate_random_walk": ([_P, _P, _P, _P, C.c_double, C.c_uint64, C.c_uint64], C.c_int),
            "gm_surrogate_bd_step": ([_P, _P, _P, _P, _P, _P, _P, C.c_double, C.c_double, C.c_double,
                                      C.c_uint64, C.c_uint64], C.c_int),
            "gm_surrogate_minimize": ([_P, _P, _P, _P, C.c_int32, C.c_double, C.c_double,
                                       C.POINTER(C.c_double)], C.c_int),
            "gm_surrogate_maybe_rebuild": ([_P, _P, _P, _P, C.POINTER(C.c_int32)], C.c_int),
            "gm_surrogate_shell_forces": ([_P, _P, _P, _P, _P, _P, _P], C.c_int),
            "gm_needs_rebuild_device": ([_P, _P, _P, _P, C.POINTER(C.c_int32)], C.c_int),
            "gm_surrogate_get_hessian": ([_P, _P], C.c_int),
            "gm_dynamics_set": ([_P, _P, _I64, _P, _P, _P], C.c_int),
            "gm_set_constraint_groups": ([_P, C.c_int32, C.c_int32, _P], C.c_int),
            "gm_half_kick_device": ([_P, _P, _P, _P, _P, _P, _P, C.c_double], C.c_int),
            "gm_drift_and_constrain_device": ([_P, _P, _P, _P, _P, _P, _P, C.c_double, C.c_int32, C.c_double,
                                               _I64, C.POINTER(C.c_int32)], C.c_int),
            "gm_constrain_velocities_device": ([_P, _P, _P, _P, _P, _P, _P, C.c_int32, C.c_double, _I64,
                                                C.POINTER(C.c_int32)], C.c_int),
            "gm_md_run": ([_P, _P, _P, _P, _P, _P, _P, C.POINTER(MdDesc), _P, _P, _P, C.POINTER(_I64)], C.c_int),
            "gm_jacobi_factor": ([_P, C.c_double, C.POINTER(C.c_double)], C.c_int),
            "gm_md_last_failure": ([_P, C.POINTER(MdFailure)], C.c_int),
            "gm_compute_forces_device_noalloc": ([_P, _P, _P, _P, _P, _P, _P], C.c_int),
        }
        for name, (args, res) in sig.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        _lib = L
    return _lib


def declared_symbols():
    """Function names declared in include/geodesic_hip.h."""
    import re
    with open(HEADER_PATH) as f:
        text = f.read()
    return sorted(set(re.findall(r"^\s*(?:gm_status|const char \*|int)\s+(gm_\w+)\s*\(", text, re.M)))


def _check(rc):
    if rc != GM_OK:
        raise GmError(rc, lib().gm_last_error().decode())


def device_count():
    return int(lib().gm_device_count())


def _p(a):
    return a.ctypes.data_as(_P) if a is not None and a.size else None


class Context:
    """gm_ctx: one per device (device.rs:8-36 try_new)."""

    def __init__(self, device=0):
        h = _P()
        _check(lib().gm_init(int(device), C.byref(h)))
        self.h = h

    def close(self):
        if self.h:
            lib().gm_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DeviceArray:
    """A float64 array in device memory owned through the C-ABI helpers."""

    def __init__(self, n, ctx: "Context | None" = None):
        self.ctx = ctx or default_context()
        self.n = int(n)
        h = _P()
        _check(lib().gm_device_alloc(self.ctx.h, 8 * self.n, C.byref(h)))
        self.ptr = h.value

    @classmethod
    def from_host(cls, a, ctx=None):
        a = np.ascontiguousarray(a, np.float64)
        d = cls(len(a), ctx)
        _check(lib().gm_copy_to_device(d.ctx.h, _P(d.ptr), _p(a), 8 * d.n))
        return d

    def to_host(self):
        out = np.empty(self.n)
        _check(lib().gm_copy_to_host(self.ctx.h, _p(out), _P(self.ptr), 8 * self.n))
        return out

    def free(self):
        if self.ptr:
            lib().gm_device_free(self.ctx.h, _P(self.ptr))
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


_default_ctx = None


def default_context():
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = Context(0)
    return _default_ctx


@dataclass
class Params:
    """The SimParams fields the force path reads (params.rs:4-23) + extensions."""
    box: tuple
    periodic: bool
    r_cutoff: float
    r_switch: float
    r_skin: float
    scale_14: float = 0.0
    bonded_min_image: bool = False


def make_desc(sigma, epsilon, topo, params: Params, keep: list) -> SystemDesc:
    """gm_system_desc for the arrays given; `keep` holds the converted arrays
    alive while the descriptor is in use."""

    def arr(a, dt):
        a = np.ascontiguousarray(a, dtype=dt)
        keep.append(a)
        return a

    d = SystemDesc()
    sig = arr(sigma, np.float64)
    eps = arr(epsilon, np.float64)
    d.n_atoms = len(sig)
    d.sigma, d.epsilon = _p(sig), _p(eps)
    box = np.broadcast_to(np.asarray(params.box, np.float64), (3,))
    for k in range(3):
        d.box[k] = float(box[k])
    d.periodic = int(bool(params.periodic))
    d.r_cutoff, d.r_switch, d.r_skin = params.r_cutoff, params.r_switch, params.r_skin
    d.scale_14 = params.scale_14
    d.bonded_min_image = int(bool(params.bonded_min_image))
    u32, f64 = np.uint32, np.float64
    g = lambda name, dt: arr(getattr(topo, name), dt)  # noqa: E731
    bi, bj, bk, br0 = g("bond_i", u32), g("bond_j", u32), g("bond_k", f64), g("bond_r0", f64)
    d.n_bonds = len(bi)
    d.bond_i, d.bond_j, d.bond_k, d.bond_r0 = _p(bi), _p(bj), _p(bk), _p(br0)
    ai, aj, ak = g("angle_i", u32), g("angle_j", u32), g("angle_k", u32)
    akth, ath0 = g("angle_kth", f64), g("angle_th0", f64)
    d.n_angles = len(ai)
    d.angle_i, d.angle_j, d.angle_k, d.angle_kth, d.angle_th0 = _p(ai), _p(aj), _p(ak), _p(akth), _p(ath0)
    di, dj, dk, dl = g("dihed_i", u32), g("dihed_j", u32), g("dihed_k", u32), g("dihed_l", u32)
    dkp, dn, dd = g("dihed_kphi", f64), g("dihed_n", u32), g("dihed_delta", f64)
    d.n_dihedrals = len(di)
    d.dihed_i, d.dihed_j, d.dihed_k, d.dihed_l = _p(di), _p(dj), _p(dk), _p(dl)
    d.dihed_kphi, d.dihed_n, d.dihed_delta = _p(dkp), _p(dn), _p(dd)
    ei, ej = g("excl_i", u32), g("excl_j", u32)
    d.n_excl = len(ei)
    d.excl_i, d.excl_j = _p(ei), _p(ej)
    p14i = arr(getattr(topo, "pair14_i", np.zeros(0)), u32)
    p14j = arr(getattr(topo, "pair14_j", np.zeros(0)), u32)
    d.n_pairs14 = len(p14i)
    d.pair14_i, d.pair14_j = _p(p14i), _p(p14j)
    return d


class System:
    """gm_sys: one simulated system on the device."""

    def __init__(self, sigma, epsilon, topo, params: Params, ctx: Context | None = None):
        ctx = ctx or default_context()
        self.ctx = ctx
        self._owned = True
        keep = []
        d = make_desc(sigma, epsilon, topo, params, keep)
        self.n = int(d.n_atoms)
        self.box = tuple(float(d.box[k]) for k in range(3))
        h = _P()
        _check(lib().gm_system_create(ctx.h, C.byref(d), C.byref(h)))
        self.h = h
        self.params = params

    @classmethod
    def borrowed(cls, handle, n, params):
        """A view of a system owned elsewhere (a batch replica)."""
        self = cls.__new__(cls)
        self.ctx = None
        self._owned = False
        self.h = handle
        self.n = int(n)
        self.params = params
        self.box = tuple(float(b) for b in np.broadcast_to(np.asarray(params.box, np.float64), (3,)))
        return self

    def close(self):
        if getattr(self, "h", None) and getattr(self, "_owned", False):
            lib().gm_system_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def build_neighbor_list(self, x, y, z):
        """Wraps x, y, z (float64, C-contiguous) in place when periodic."""
        for a in (x, y, z):
            assert a.dtype == np.float64 and a.flags.c_contiguous and len(a) == self.n
        _check(lib().gm_build_neighbor_list(self.h, _p(x), _p(y), _p(z)))

    def needs_rebuild(self, x, y, z):
        out = C.c_int32(0)
        x, y, z = (np.ascontiguousarray(a, np.float64) for a in (x, y, z))
        _check(lib().gm_needs_rebuild(self.h, _p(x), _p(y), _p(z), C.byref(out)))
        return bool(out.value)

    def compute_forces(self, x, y, z):
        """Returns (EnergyTerms, fx, fy, fz)."""
        x, y, z = (np.ascontiguousarray(a, np.float64) for a in (x, y, z))
        fx, fy, fz = np.empty(self.n), np.empty(self.n), np.empty(self.n)
        e = EnergyTerms()
        rc = lib().gm_compute_forces(self.h, _p(x), _p(y), _p(z), _p(fx), _p(fy), _p(fz), C.byref(e))
        _check_numerical(rc, (e, fx, fy, fz))
        return e, fx, fy, fz

    def pair_list(self):
        n = _I64(0)
        _check(lib().gm_get_pair_list(self.h, None, None, 0, C.byref(n)))
        pi = np.zeros(max(n.value, 1), np.uint32)
        pj = np.zeros(max(n.value, 1), np.uint32)
        _check(lib().gm_get_pair_list(self.h, _p(pi), _p(pj), n.value, C.byref(n)))
        return pi[:n.value], pj[:n.value]

    # device-resident variants: integer device pointers (e.g. tensor.data_ptr())
    def build_neighbor_list_device(self, dx, dy, dz):
        _check(lib().gm_build_neighbor_list_device(self.h, _P(dx), _P(dy), _P(dz)))

    def compute_forces_device_noalloc(self, dx, dy, dz, dfx, dfy, dfz):
        """gm_compute_forces_device with the library's static scratch array."""
        _check(lib().gm_compute_forces_device_noalloc(self.h, _P(dx), _P(dy), _P(dz), _P(dfx), _P(dfy), _P(dfz)))

    def compute_forces_device(self, dx, dy, dz, dfx, dfy, dfz):
        _check(lib().gm_compute_forces_device(self.h, _P(dx), _P(dy), _P(dz), _P(dfx), _P(dfy), _P(dfz)))

    def set_positions(self, x, y, z):
        x, y, z = (np.ascontiguousarray(a, np.float64) for a in (x, y, z))
        _check(lib().gm_set_positions(self.h, _p(x), _p(y), _p(z)))

    def compute_forces_resident(self):
        """Asynchronous evaluation at the resident positions; see sync()."""
        _check(lib().gm_compute_forces_resident(self.h))

    def get_forces(self):
        fx, fy, fz = np.empty(self.n), np.empty(self.n), np.empty(self.n)
        _check(lib().gm_get_forces(self.h, _p(fx), _p(fy), _p(fz)))
        return fx, fy, fz

    def sync(self):
        e = EnergyTerms()
        _check_numerical(lib().gm_sync(self.h, C.byref(e)), e)
        return e

    # ---- device-resident dynamics (integrator.rs, constraint.rs, lib.rs run loop)

    def dynamics_set(self, mass, constr_i=(), constr_j=(), constr_dsq=()):
        m = np.ascontiguousarray(mass, np.float64)
        ci = np.ascontiguousarray(constr_i, np.uint32)
        cj = np.ascontiguousarray(constr_j, np.uint32)
        dsq = np.ascontiguousarray(constr_dsq, np.float64)
        _check(lib().gm_dynamics_set(self.h, _p(m), len(ci), _p(ci), _p(cj), _p(dsq)))

    def set_constraint_groups(self, enable=True, iteration_cap=0):
        """SHAKE / RATTLE by constraint groups (enable) or per-iteration launches;
        iteration_cap > 0 hands over to the latter after that many group
        iterations.  Returns the number of groups (0: none fit)."""
        ng = C.c_int32(0)
        _check(lib().gm_set_constraint_groups(self.h, 1 if enable else 0, int(iteration_cap), C.byref(ng)))
        return ng.value

    def half_kick_device(self, dvx, dvy, dvz, dfx, dfy, dfz, dt_half):
        _check(lib().gm_half_kick_device(self.h, *(_P(a.ptr) for a in (dvx, dvy, dvz, dfx, dfy, dfz)),
                                         float(dt_half)))

    def drift_and_constrain_device(self, dx, dy, dz, dvx, dvy, dvz, dt, max_iter, tol, step=0):
        it = C.c_int32(0)
        _check(lib().gm_drift_and_constrain_device(self.h, *(_P(a.ptr) for a in (dx, dy, dz, dvx, dvy, dvz)),
                                                   float(dt), int(max_iter), float(tol), int(step), C.byref(it)))
        return it.value

    def constrain_velocities_device(self, dx, dy, dz, dvx, dvy, dvz, max_iter, tol, step=0):
        it = C.c_int32(0)
        _check(lib().gm_constrain_velocities_device(self.h, *(_P(a.ptr) for a in (dx, dy, dz, dvx, dvy, dvz)),
                                                    int(max_iter), float(tol), int(step), C.byref(it)))
        return it.value

    def md_run(self, dx, dy, dz, dvx, dvy, dvz, n_steps, dt, frame_interval=1, max_iter=100, tol=1e-8,
               first_step=0, resume=False):
        """Returns (frames [n_frames, 3, n] f32, potential, kinetic).  On a
        failure the GmError carries the frames taken before it in .result."""
        md = MdDesc(n_steps=int(n_steps), dt=float(dt), frame_interval=int(frame_interval),
                    max_iter=int(max_iter), tol=float(tol), first_step=int(first_step), resume=1 if resume else 0)
        cap = int(n_steps) // int(frame_interval) + 2
        frames = np.empty((cap, 3, self.n), np.float32)
        epot, ekin = np.empty(cap), np.empty(cap)
        nf = _I64(0)
        rc = lib().gm_md_run(self.h, *(_P(a.ptr) for a in (dx, dy, dz, dvx, dvy, dvz)), C.byref(md),
                             _p(frames), _p(epot), _p(ekin), C.byref(nf))
        k = nf.value
        if rc != GM_OK:
            err = GmError(rc, lib().gm_last_error().decode())
            err.result = (frames[:k], epot[:k], ekin[:k])
            raise err
        return frames[:k], epot[:k], ekin[:k]

    def md_last_failure(self):
        f = MdFailure()
        _check(lib().gm_md_last_failure(self.h, C.byref(f)))
        return f.as_dict()

    def jacobi_factor(self, total_energy):
        out = C.c_double(0.0)
        _check(lib().gm_jacobi_factor(self.h, float(total_energy), C.byref(out)))
        return out.value

    def profile_begin(self, every=1):
        """Time the next evaluation and every `every`-th one after it with HIP events."""
        _check(lib().gm_profile_begin(self.h, int(every)))

    def profile_end(self):
        p = Profile()
        _check(lib().gm_profile_end(self.h, C.byref(p)))
        return p.as_dict()

    def set_work_chunk(self, entries_per_item):
        _check(lib().gm_set_work_chunk(self.h, int(entries_per_item)))

    def set_tile_pruning(self, enable):
        """Skip tiles provably beyond r_cutoff (default on); results are
        bit-identical either way."""
        _check(lib().gm_set_tile_pruning(self.h, 1 if enable else 0))

    LAYOUTS = {"tiles": 0, "partners": 1, "auto": 2}

    def set_pair_layout(self, layout):
        """Nonbonded list layout: "auto" (the default: partner lists unless they
        need 32-bit slot numbers, then tiles), "partners" (per-atom partner
        lists) or "tiles" (8x8 cluster tiles).  Systems with more than 256
        distinct (sigma, epsilon) pairs always use tiles."""
        _check(lib().gm_set_pair_layout(self.h, self.LAYOUTS[layout]))

    def stats(self):
        s = Stats()
        _check(lib().gm_get_stats(self.h, C.byref(s)))
        return s.as_dict()


class Batch:
    """gm_batch: independent replicas evaluated by one nonbonded launch from a
    captured hipGraph (partner lists of every replica, or the persistent tile
    kernel).  `members` is a list of (sigma, epsilon, topo, params); positions
    and forces are concatenated in member order."""

    def __init__(self, members, ctx: Context | None = None):
        ctx = ctx or default_context()
        self.ctx = ctx
        keep = []
        descs = (SystemDesc * len(members))()
        for k, (sig, eps, topo, params) in enumerate(members):
            descs[k] = make_desc(sig, eps, topo, params, keep)
        self.sizes = [int(descs[k].n_atoms) for k in range(len(members))]
        self.offsets = np.concatenate([[0], np.cumsum(self.sizes)]).astype(np.int64)
        self.n = int(self.offsets[-1])
        self.params = [m[3] for m in members]
        h = _P()
        _check(lib().gm_batch_create(ctx.h, C.cast(descs, _P), len(members), C.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().gm_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __len__(self):
        return len(self.sizes)

    def replica(self, r):
        h = _P()
        _check(lib().gm_batch_replica(self.h, int(r), C.byref(h)))
        return System.borrowed(h, self.sizes[r], self.params[r])

    def build_neighbor_lists(self, x, y, z):
        for a in (x, y, z):
            assert a.dtype == np.float64 and a.flags.c_contiguous and len(a) == self.n
        _check(lib().gm_batch_build_neighbor_lists(self.h, _p(x), _p(y), _p(z)))

    def set_positions(self, x, y, z):
        x, y, z = (np.ascontiguousarray(a, np.float64) for a in (x, y, z))
        _check(lib().gm_batch_set_positions(self.h, _p(x), _p(y), _p(z)))

    def compute(self):
        _check(lib().gm_batch_compute(self.h))

    def sync(self):
        """Returns (list of EnergyTerms per replica, device ms of the last compute)."""
        e = (EnergyTerms * len(self.sizes))()
        ms = C.c_float(0.0)
        rc = lib().gm_batch_sync(self.h, C.cast(e, _P), C.byref(ms))
        _check_numerical(rc, ([e[k] for k in range(len(self.sizes))], ms.value))
        return [e[k] for k in range(len(self.sizes))], ms.value

    def get_forces(self):
        fx, fy, fz = np.empty(self.n), np.empty(self.n), np.empty(self.n)
        _check(lib().gm_batch_get_forces(self.h, _p(fx), _p(fy), _p(fz)))
        return fx, fy, fz

    def set_pair_layout(self, layout):
        """"auto" (the default: partner lists when every replica's lists hold
        16-bit codes, else tiles for all), "partners" or "tiles"; applied at
        once when built, else at the build."""
        _check(lib().gm_batch_set_pair_layout(self.h, System.LAYOUTS[layout]))

    def pair_layout(self):
        """The layout the current plan evaluates: "partners" or "tiles"."""
        v = C.c_int32(-1)
        _check(lib().gm_batch_get_pair_layout(self.h, C.byref(v)))
        return {0: "tiles", 1: "partners"}[v.value]

    def set_shape(self, entries_per_item, waves_per_cu):
        """Tile layout work shape (speed only: every shape gives the same bytes)."""
        _check(lib().gm_batch_set_shape(self.h, int(entries_per_item), int(waves_per_cu)))

    def autotune(self):
        """Tile layout: the fastest (entries per item, waves per CU) and its ms,
        kept.  Partner lists have no shape: (0, 0, ms of the step)."""
        c, w, ms = C.c_int32(0), C.c_int32(0), C.c_float(0.0)
        _check(lib().gm_batch_autotune(self.h, C.byref(c), C.byref(w), C.byref(ms)))
        return c.value, w.value, ms.value


class Surrogate:
    """gm_surrogate: the shell-force network and the 4:1 multiple-time-step
    driver over an exact model (`params`) and an inner model (same atoms,
    LJ cut at r_inner_cutoff).  Positions and forces passed to the device
    methods are DeviceArray objects."""

    def __init__(self, sigma, epsilon, topo, params: Params, atom_class, r_inner=(7.5, 6.5, 9.0),
                 window=(7.5, 12.0, 0.5), seed=7, ctx: Context | None = None):
        ctx = ctx or default_context()
        self.ctx = ctx
        keep = []
        d = make_desc(sigma, epsilon, topo, params, keep)
        self.n = int(d.n_atoms)
        cls = np.ascontiguousarray(atom_class, dtype=np.uint8)
        keep.append(cls)
        sd = SurrogateDesc()
        sd.r_inner_cutoff, sd.r_inner_switch, sd.r_inner_skin = (float(v) for v in r_inner)
        sd.r_lo, sd.r_hi, sd.fade = (float(v) for v in window)
        sd.atom_class = _p(cls)
        sd.seed = int(seed)
        self.params = params
        self.inner_params = Params(box=params.box, periodic=params.periodic, r_cutoff=float(r_inner[0]),
                                   r_switch=float(r_inner[1]), r_skin=float(r_inner[2]),
                                   scale_14=params.scale_14, bonded_min_image=params.bonded_min_image)
        h = _P()
        _check(lib().gm_surrogate_create(ctx.h, C.byref(d), C.byref(sd), C.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().gm_surrogate_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def system(self, which):
        """Borrowed System: 0 = exact model, 1 = inner model."""
        h = _P()
        _check(lib().gm_surrogate_system(self.h, int(which), C.byref(h)))
        return System.borrowed(h, self.n, self.params if which == 0 else self.inner_params)

    def build(self, dx, dy, dz):
        _check(lib().gm_surrogate_build(self.h, _P(dx.ptr), _P(dy.ptr), _P(dz.ptr)))

    def features(self, dx, dy, dz):
        out = np.empty((self.n, SUR_NFEAT), np.float32)
        _check(lib().gm_surrogate_features(self.h, _P(dx.ptr), _P(dy.ptr), _P(dz.ptr), _p(out)))
        return out

    def hessian(self):
        """Shell Hessian K (n x 6) of the last features() call."""
        out = np.empty((self.n, 6), np.float32)
        _check(lib().gm_surrogate_get_hessian(self.h, _p(out)))
        return out

    def set_params(self, params=None, fmean=None, finv=None, scales=None):
        conv = lambda a, dt: None if a is None else np.ascontiguousarray(a, dtype=dt)  # noqa: E731
        p, m, v, sc = conv(params, np.float32), conv(fmean, np.float32), conv(finv, np.float32), conv(scales, np.float64)
        _check(lib().gm_surrogate_set_params(self.h, _p(p), _p(m), _p(v), _p(sc)))

    def get_params(self):
        p = np.empty(SUR_NPARAM, np.float32)
        m, v = np.empty(SUR_NFEAT, np.float32), np.empty(SUR_NFEAT, np.float32)
        sc = np.empty(3, np.float64)
        _check(lib().gm_surrogate_get_params(self.h, _p(p), _p(m), _p(v), _p(sc)))
        return p, m, v, sc

    def infer(self, feat):
        feat = np.ascontiguousarray(feat, np.float32)
        out = np.empty((len(feat), 3), np.float32)
        _check(lib().gm_surrogate_infer(self.h, _p(feat), len(feat), _p(out)))
        return out

    def loss_grad(self, feat, basis, target):
        feat = np.ascontiguousarray(feat, np.float32)
        basis = np.ascontiguousarray(basis, np.float32)
        target = np.ascontiguousarray(target, np.float32)
        loss = C.c_double(0.0)
        g = np.empty(SUR_NPARAM, np.float32)
        _check(lib().gm_surrogate_loss_grad(self.h, _p(feat), _p(basis), _p(target), len(feat), C.byref(loss),
                                            _p(g)))
        return loss.value, g

    def train(self, dx, dy, dz, n_frames, mu=BD_MU, sigma=BD_SIGMA, cap=BD_CAP, seed=11, lr=1e-3, batch=8192,
              update=True):
        """Brownian trajectory from (dx, dy, dz), one training frame per inner
        step.  Returns (loss history n_frames x 3 [network, o = 0, Taylor], report)."""
        td = TrainDesc(n_frames=int(n_frames), mu=float(mu), sigma=float(sigma), cap=float(cap), seed=int(seed),
                       lr=float(lr), batch=int(batch), update=int(bool(update)))
        hist = np.empty((max(int(n_frames), 1), 3), np.float32)
        rep = TrainReport()
        _check(lib().gm_surrogate_train(self.h, _P(dx.ptr), _P(dy.ptr), _P(dz.ptr), C.byref(td), _p(hist),
                                        C.byref(rep)))
        return hist[:int(n_frames)], rep.as_dict()

    def mts_reset(self):
        _check(lib().gm_mts_reset(self.h))

    def mts_step(self, dx, dy, dz, dfx, dfy, dfz, mode=MTS_NET):
        """mode: MTS_HELD (shell force held), MTS_NET (network), MTS_TAYLOR."""
        was = C.c_int32(0)
        _check(lib().gm_mts_step(self.h, _P(dx.ptr), _P(dy.ptr), _P(dz.ptr), _P(dfx.ptr), _P(dfy.ptr),
                                 _P(dfz.ptr), int(mode), C.byref(was)))
        return bool(was.value)

    def bd_step(self, dx, dy, dz, dfx, dfy, dfz, mu=BD_MU, sigma=BD_SIGMA, cap=BD_CAP, seed=0, step=0):
        _check(lib().gm_surrogate_bd_step(self.h, _P(dx.ptr), _P(dy.ptr), _P(dz.ptr), _P(dfx.ptr), _P(dfy.ptr),
                                          _P(dfz.ptr), float(mu), float(sigma), float(cap), int(seed), int(step)))

    def minimize(self, dx, dy, dz, steps=300, mu=1e-3, cap=0.05):
        """Steepest descent on the exact model; returns the largest force component after."""
        m = C.c_double(0.0)
        _check(lib().gm_surrogate_minimize(self.h, _P(dx.ptr), _P(dy.ptr), _P(dz.ptr), int(steps), float(mu),
                                           float(cap), C.byref(m)))
        return m.value

    def maybe_rebuild(self, dx, dy, dz):
        r = C.c_int32(0)
        _check(lib().gm_surrogate_maybe_rebuild(self.h, _P(dx.ptr), _P(dy.ptr), _P(dz.ptr), C.byref(r)))
        return bool(r.value)

    def shell_forces(self, dx, dy, dz, dfx, dfy, dfz):
        _check(lib().gm_surrogate_shell_forces(self.h, _P(dx.ptr), _P(dy.ptr), _P(dz.ptr), _P(dfx.ptr),
                                               _P(dfy.ptr), _P(dfz.ptr)))

    def random_walk(self, dx, dy, dz, sigma, seed, step):
        _check(lib().gm_surrogate_random_walk(self.h, _P(dx.ptr), _P(dy.ptr), _P(dz.ptr), float(sigma),
                                              int(seed), int(step)))

    def sync(self):
        """Waits for the surrogate's stream (shared by both systems)."""
        self.system(0).sync()


# ---- mirror of the reference's backend API ------------------------------------

@dataclass
class ForceBuffer:
    """geodesic-core ForceBuffer (buffers.rs:5-9)."""
    fx: np.ndarray
    fy: np.ndarray
    fz: np.ndarray


class GpuBackend:
    """ComputeBackend over the gfx950 engine (gpu_backend.rs:11-100).

    `try_new(atoms, topology, params)` raises GmError with the BackendError
    code on failure, like GpuBackend::try_new.  `state` objects need float64
    `pos_x/pos_y/pos_z` (and, for geodesic_drift, `vel_*` and `step`) arrays
    (SimState, state.rs:3-23).  max_constr_iter / constr_tol are the SimParams
    fields the constraint solver reads (params.rs:13-14)."""

    def __init__(self, system: System, atoms=None, topology=None, max_constr_iter=100, constr_tol=1e-10):
        self._sys = system
        self._reduced = None
        self._potential = 0.0
        self._terms = None
        self._atoms, self._topo = atoms, topology
        self._max_iter, self._tol = max_constr_iter, constr_tol
        self._dyn = False

    @classmethod
    def try_new(cls, atoms, topology, params: Params, ctx: Context | None = None, max_constr_iter=100,
                constr_tol=1e-10):
        return cls(System(atoms.sigma, atoms.epsilon, topology, params, ctx), atoms, topology, max_constr_iter,
                   constr_tol)

    def build_neighbor_list(self, state, params=None):
        self._sys.build_neighbor_list(state.pos_x, state.pos_y, state.pos_z)

    def compute_forces(self, state):
        e, fx, fy, fz = self._sys.compute_forces(state.pos_x, state.pos_y, state.pos_z)
        self._reduced = ForceBuffer(fx, fy, fz)
        self._potential = e.total
        self._terms = e
        return self._reduced

    def reduce_forces(self):
        r = self._reduced
        return ForceBuffer(r.fx.copy(), r.fy.copy(), r.fz.copy())

    def potential_energy(self):
        return self._potential

    def energy_terms(self):
        """Per-term breakdown (EnergyReport, lib.rs:37-45), an inherent addition."""
        return self._terms.as_dict()

    def needs_rebuild(self, state):
        return self._sys.needs_rebuild(state.pos_x, state.pos_y, state.pos_z)

    def n_threads(self):
        return 1

    def geodesic_drift(self, state, dt):
        """integrator.rs:40-71 (drift, SHAKE, velocity resync) on the device, in
        place on `state`; GmError with status 9 (ConvergenceError) when the
        solver does not converge."""
        if self._atoms is None:
            raise GmError(GM_ERR_INVALID_ARG, "geodesic_drift needs the backend built by try_new")
        if not self._dyn:
            t = self._topo
            self._sys.dynamics_set(self._atoms.mass, getattr(t, "constr_i", ()), getattr(t, "constr_j", ()),
                                   getattr(t, "constr_dsq", ()))
            self._dyn = True
        names = ("pos_x", "pos_y", "pos_z", "vel_x", "vel_y", "vel_z")
        d = [DeviceArray.from_host(np.ascontiguousarray(getattr(state, k), np.float64)) for k in names]
        self._sys.drift_and_constrain_device(*d, dt, self._max_iter, self._tol, step=getattr(state, "step", 0))
        for k, a in zip(names, d):
            getattr(state, k)[:] = a.to_host()
