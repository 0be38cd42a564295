"""Shared helpers for the test suite: fixture loading and comparison metrics."""
import os

import numpy as np

from oracle import amber

FIX = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "fixtures")


def load(name):
    """Mirror of geodesic-gpu/tests/common/mod.rs:8-14: prmtop + inpcrd."""
    with open(os.path.join(FIX, f"{name}.prmtop")) as f:
        atoms, topo = amber.parse_prmtop(f.read())
    with open(os.path.join(FIX, f"{name}.inpcrd")) as f:
        x, y, z, vx, vy, vz = amber.parse_inpcrd(f.read(), len(atoms.mass))
    return (x, y, z), atoms, topo


def force_rms_rel(fx, fy, fz, gx, gy, gz):
    """sqrt(sum|F_gpu - F_ref|^2 / sum|F_ref|^2)."""
    num = np.sum((fx - gx) ** 2 + (fy - gy) ** 2 + (fz - gz) ** 2)
    den = np.sum(gx ** 2 + gy ** 2 + gz ** 2)
    return float(np.sqrt(num / den)) if den > 0 else float(np.sqrt(num))


def rel(a, b):
    return abs(a - b) / max(abs(b), 1e-300)
