"""ctypes front-end of libgeodesic_host.so (include/geodesic_host.h).

The C++ host layer above the force engine: AMBER prmtop / inpcrd reading
(geodesic-io/src/prmtop.rs, inpcrd.rs; paths relative to the reference
repository root), the DCD and CSV writers (dcd.rs, export.rs), `geodesic
energy` and `geodesic run` (geodesic/src/lib.rs:113-171, 187-330) and the
ComputeBackend trait over the GPU backend (backend.rs:11-23)."""
import ctypes as C
import os

import numpy as np

from . import GmError, _P, lib as _engine_lib

_HERE = os.path.dirname(os.path.abspath(__file__))
HOST_LIB_PATH = os.environ.get("GEODESIC_HOST_LIB") or os.path.join(os.path.dirname(_HERE), "lib",
                                                                    "libgeodesic_host.so")
HOST_HEADER_PATH = os.path.join(os.path.dirname(os.path.dirname(_HERE)), "include", "geodesic_host.h")

GMH_ERR_IO, GMH_ERR_CONFIG, GMH_ERR_NUMERICAL = 10, 11, 12

# field ids of gmh_topology_field, in header order, with their element types
FIELDS = [("sigma", np.float64), ("epsilon", np.float64), ("mass", np.float64), ("charge", np.float64),
          ("element", np.uint8), ("residue", np.uint32),
          ("bond_i", np.uint32), ("bond_j", np.uint32), ("bond_k", np.float64), ("bond_r0", np.float64),
          ("angle_i", np.uint32), ("angle_j", np.uint32), ("angle_k", np.uint32), ("angle_kth", np.float64),
          ("angle_th0", np.float64),
          ("dihed_i", np.uint32), ("dihed_j", np.uint32), ("dihed_k", np.uint32), ("dihed_l", np.uint32),
          ("dihed_kphi", np.float64), ("dihed_n", np.uint32), ("dihed_delta", np.float64),
          ("constr_i", np.uint32), ("constr_j", np.uint32), ("constr_dsq", np.float64),
          ("excl_i", np.uint32), ("excl_j", np.uint32),
          ("pair14_i", np.uint32), ("pair14_j", np.uint32)]

_D, _I32, _I64 = C.c_double, C.c_int32, C.c_int64


class EnergyReport(C.Structure):
    _fields_ = [("bond", _D), ("angle", _D), ("dihedral", _D), ("nonbonded", _D), ("total", _D),
                ("lj14", _D), ("r_switch", _D), ("r_cutoff", _D)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class RunConfig(C.Structure):
    _fields_ = [("n_steps", _I64), ("frame_interval", _I32), ("backend", C.c_char_p), ("n_threads", _I32),
                ("prmtop", C.c_char_p), ("inpcrd", C.c_char_p), ("box_size", _D * 3), ("periodic", _I32),
                ("dt", _D), ("total_energy", _D), ("r_cutoff", _D), ("r_skin", _D), ("r_switch", _D),
                ("max_constr_iter", _I32), ("constr_tol", _D), ("constrain_hydrogens", _I32),
                ("trajectory", C.c_char_p), ("energy_log", C.c_char_p), ("monitoring", _I32),
                ("energy_drift_threshold_kcal", _D), ("energy_drift_stop", _I32), ("fidelity", _I32),
                ("device", _I32)]


class RunSummary(C.Structure):
    _fields_ = [("n_steps", _I64), ("n_frames", _I64), ("final_potential", _D), ("final_kinetic", _D)]


_lib = None


def host_lib():
    """Load libgeodesic_host.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(HOST_LIB_PATH):
            raise RuntimeError(f"libgeodesic_host.so not built at {HOST_LIB_PATH}; run __graft_entry__.build()")
        _engine_lib()  # the engine first (the host library links it)
        L = C.CDLL(HOST_LIB_PATH)
        pp, cp = C.POINTER(_P), C.c_char_p
        sig = {
            "gmh_last_error": ([], C.c_char_p),
            "gmh_prmtop_read": ([cp, _I32, pp], C.c_int),
            "gmh_topology_free": ([_P], C.c_int),
            "gmh_topology_field": ([_P, _I32, C.POINTER(C.c_void_p), C.POINTER(_I64)], C.c_int),
            "gmh_topology_scales": ([_P, C.POINTER(_D), C.POINTER(_D)], C.c_int),
            "gmh_topology_promote_hydrogen_bonds": ([_P], C.c_int),
            "gmh_inpcrd_read": ([cp, _I64, _I32, _P, _P], C.c_int),
            "gmh_dcd_write": ([cp, _I64, _I32, _D, _I64, _P], C.c_int),
            "gmh_energy_log_write": ([cp, _I64, _I64, _P, _P, _P, _P], C.c_int),
            "gmh_energy_from_files": ([cp, cp, _I32, _I32, C.POINTER(EnergyReport)], C.c_int),
            "gmh_run": ([C.POINTER(RunConfig), C.POINTER(RunSummary)], C.c_int),
            "gmh_backend_create": ([_P, C.POINTER(_D), _I32, _D, _D, _D, _I32, _D, _I32, _I32, pp], C.c_int),
            "gmh_backend_free": ([_P], C.c_int),
            "gmh_backend_set_state": ([_P, _P, _P, C.c_uint64], C.c_int),
            "gmh_backend_get_state": ([_P, _P, _P], C.c_int),
            "gmh_backend_build_neighbor_list": ([_P], C.c_int),
            "gmh_backend_compute_forces": ([_P, _P, _P], C.c_int),
            "gmh_backend_needs_rebuild": ([_P, C.POINTER(_I32)], C.c_int),
            "gmh_backend_geodesic_drift": ([_P, _D], C.c_int),
            "gmh_backend_n_threads": ([_P, C.POINTER(_I64)], C.c_int),
        }
        for name, (args, res) in sig.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        _lib = L
    return _lib


def _check(rc):
    if rc != 0:
        raise GmError(rc, host_lib().gmh_last_error().decode())


def _enc(s):
    return os.fsencode(s)


def _p(a):
    return a.ctypes.data_as(_P)


class Topology:
    """A parsed prmtop (AtomData + BondedTopology) held by the library."""

    def __init__(self, path, fidelity=False):
        self.h = None
        h = _P()
        _check(host_lib().gmh_prmtop_read(_enc(path), 1 if fidelity else 0, C.byref(h)))
        self.h = h

    def field(self, name):
        k = [f for f, _ in FIELDS].index(name)
        dt = FIELDS[k][1]
        ptr, n = C.c_void_p(), _I64()
        _check(host_lib().gmh_topology_field(self.h, k, C.byref(ptr), C.byref(n)))
        if n.value == 0:
            return np.zeros(0, dt)
        buf = (C.c_char * (n.value * np.dtype(dt).itemsize)).from_address(ptr.value)
        return np.frombuffer(buf, dtype=dt).copy()

    def as_dict(self):
        return {name: self.field(name) for name, _ in FIELDS}

    def scales(self):
        s14, scee = _D(), _D()
        _check(host_lib().gmh_topology_scales(self.h, C.byref(s14), C.byref(scee)))
        return s14.value, scee.value

    def promote_hydrogen_bonds(self):
        _check(host_lib().gmh_topology_promote_hydrogen_bonds(self.h))

    def close(self):
        if self.h:
            host_lib().gmh_topology_free(self.h)
            self.h = None

    __del__ = close


def read_inpcrd(path, n_atoms, has_box=False):
    """(pos, vel): float64 arrays of shape (3, n)."""
    pos, vel = np.zeros((3, n_atoms)), np.zeros((3, n_atoms))
    _check(host_lib().gmh_inpcrd_read(_enc(path), n_atoms, 1 if has_box else 0, _p(pos), _p(vel)))
    return pos, vel


def write_dcd(path, frames, frame_interval, dt_ps):
    """frames: float32 (n_frames, 3, natom)."""
    f = np.ascontiguousarray(frames, np.float32)
    _check(host_lib().gmh_dcd_write(_enc(path), f.shape[2], frame_interval, dt_ps, f.shape[0], _p(f)))


def write_energy_log(path, n_atoms, step, time_ps, potential, kinetic):
    step = np.ascontiguousarray(step, np.uint64)
    arrs = [np.ascontiguousarray(a, np.float64) for a in (time_ps, potential, kinetic)]
    _check(host_lib().gmh_energy_log_write(_enc(path), n_atoms, len(step), _p(step), *[_p(a) for a in arrs]))


def energy_from_files(prmtop, inpcrd, fidelity=False, device=0):
    """`geodesic energy` through the GPU backend: per-term energies (dict)."""
    r = EnergyReport()
    _check(host_lib().gmh_energy_from_files(_enc(prmtop), _enc(inpcrd), 1 if fidelity else 0, device, C.byref(r)))
    return r.as_dict()


def run(prmtop, inpcrd, trajectory, energy_log, n_steps, frame_interval=1, dt=0.001, box_size=(1000.0,) * 3,
        periodic=False, r_cutoff=12.0, r_skin=14.0, r_switch=10.0, max_constr_iter=100, constr_tol=1e-8,
        constrain_hydrogens=True, total_energy=0.0, backend="gpu", n_threads=1, monitoring=None,
        fidelity=False, device=0):
    """`geodesic run` with a parsed config (config.rs:102-128).  monitoring:
    None or (threshold_kcal, "warn" | "stop").  Returns the RunSummary dict."""
    c = RunConfig()
    c.n_steps, c.frame_interval, c.backend, c.n_threads = n_steps, frame_interval, backend.encode(), n_threads
    c.prmtop, c.inpcrd = _enc(prmtop), _enc(inpcrd)
    for k in range(3):
        c.box_size[k] = box_size[k]
    c.periodic, c.dt, c.total_energy = 1 if periodic else 0, dt, total_energy
    c.r_cutoff, c.r_skin, c.r_switch = r_cutoff, r_skin, r_switch
    c.max_constr_iter, c.constr_tol, c.constrain_hydrogens = max_constr_iter, constr_tol, 1 if constrain_hydrogens else 0
    c.trajectory, c.energy_log = _enc(trajectory), _enc(energy_log)
    if monitoring is not None:
        c.monitoring, c.energy_drift_threshold_kcal = 1, monitoring[0]
        c.energy_drift_stop = 1 if monitoring[1] == "stop" else 0
    c.fidelity, c.device = 1 if fidelity else 0, device
    s = RunSummary()
    _check(host_lib().gmh_run(C.byref(c), C.byref(s)))
    return {k: getattr(s, k) for k, _ in RunSummary._fields_}


class Backend:
    """The ComputeBackend trait (backend.rs:11-23) over the C++ GpuBackend,
    holding one SimState (positions, velocities, step)."""

    def __init__(self, topology: Topology, box, periodic, r_cutoff, r_skin, r_switch, max_constr_iter=100,
                 constr_tol=1e-8, bonded_min_image=False, device=0):
        self.h = None
        h = _P()
        b = (_D * 3)(*box)
        _check(host_lib().gmh_backend_create(topology.h, b, 1 if periodic else 0, r_cutoff, r_skin, r_switch,
                                             max_constr_iter, constr_tol, 1 if bonded_min_image else 0, device,
                                             C.byref(h)))
        self.h = h
        self.n = len(topology.field("mass"))

    def set_state(self, pos, vel, step=0):
        pos, vel = (np.ascontiguousarray(a, np.float64).reshape(3, self.n) for a in (pos, vel))
        _check(host_lib().gmh_backend_set_state(self.h, _p(pos), _p(vel), step))

    def state(self):
        pos, vel = np.zeros((3, self.n)), np.zeros((3, self.n))
        _check(host_lib().gmh_backend_get_state(self.h, _p(pos), _p(vel)))
        return pos, vel

    def build_neighbor_list(self):
        _check(host_lib().gmh_backend_build_neighbor_list(self.h))

    def compute_forces(self):
        from . import EnergyTerms
        f, e = np.zeros((3, self.n)), EnergyTerms()
        _check(host_lib().gmh_backend_compute_forces(self.h, _p(f), C.byref(e)))
        return f, e

    def needs_rebuild(self):
        o = _I32()
        _check(host_lib().gmh_backend_needs_rebuild(self.h, C.byref(o)))
        return bool(o.value)

    def geodesic_drift(self, dt):
        _check(host_lib().gmh_backend_geodesic_drift(self.h, dt))

    def n_threads(self):
        o = _I64()
        _check(host_lib().gmh_backend_n_threads(self.h, C.byref(o)))
        return o.value

    def close(self):
        if self.h:
            host_lib().gmh_backend_free(self.h)
            self.h = None

    __del__ = close


def declared_host_symbols():
    import re
    with open(HOST_HEADER_PATH) as f:
        text = f.read()
    return sorted(set(re.findall(r"\b(gmh_[a-z0-9_]+)\s*\(", text)))
