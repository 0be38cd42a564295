"""The C++ host layer's readers and writers (libgeodesic_host.so), no GPU.

Pins, with paths relative to the reference repository root:
  - prmtop / inpcrd reading equals the oracle's restatement of prmtop.rs and
    inpcrd.rs field for field, bit for bit, on every fixture; the fidelity
    mode's 1-4 pairs equal the oracle's list of proper dihedrals with the 1-4
    flag clear, scaled 1/2.0 (no SCNB section in the fixtures);
  - DcdWriter re-writes the reference's committed golden trajectory
    (geodesic/tests/golden/ala_dipeptide_ref.dcd) byte for byte from its own
    frames (dcd.rs header, title, frame records, NSET patch);
  - EnergyLogWriter reproduces geodesic-io/tests/export_roundtrip.rs:4-25
    (header, "0", "0.000", 300.00 K);
  - error behaviour: IoError for a missing file, ConfigError for a truncated
    prmtop, a 10-12 H-bond type and an inpcrd/prmtop atom-count mismatch.
"""
import os

import numpy as np
import pytest

import geodesic_hip as gh
from geodesic_hip import host
from oracle import amber

from _common import FIX

NAMES = ["lj_pair", "harmonic_dimer", "water_box_4", "ala_dipeptide"]


def test_host_library_exports_every_declared_symbol():
    names = host.declared_host_symbols()
    assert len(names) >= 20
    L = host.host_lib()
    for n in names:
        assert hasattr(L, n), n


@pytest.mark.parametrize("name", NAMES)
def test_prmtop_matches_oracle(name):
    path = os.path.join(FIX, f"{name}.prmtop")
    with open(path) as f:
        atoms, topo = amber.parse_prmtop(f.read())
    t = host.Topology(path)
    d = t.as_dict()
    for k in ("sigma", "epsilon", "mass", "charge"):
        assert d[k].tobytes() == getattr(atoms, k).astype(np.float64).tobytes(), k
    assert np.array_equal(d["element"], atoms.element.astype(np.uint8))
    for k in ("bond_i", "bond_j", "bond_k", "bond_r0", "angle_i", "angle_j", "angle_k", "angle_kth",
              "angle_th0", "dihed_i", "dihed_j", "dihed_k", "dihed_l", "dihed_kphi", "dihed_n", "dihed_delta",
              "excl_i", "excl_j"):
        ref = getattr(topo, k)
        assert d[k].dtype == ref.dtype and d[k].tobytes() == ref.tobytes(), k
    # the reference's reading keeps no 1-4 pairs
    assert len(d["pair14_i"]) == 0 and t.scales()[0] == 0.0


def test_prmtop_fidelity_keeps_the_1_4_flags():
    path = os.path.join(FIX, "ala_dipeptide.prmtop")
    with open(path) as f:
        _, topo = amber.parse_prmtop(f.read())
    t = host.Topology(path, fidelity=True)
    assert np.array_equal(t.field("pair14_i"), topo.pair14_i)
    assert np.array_equal(t.field("pair14_j"), topo.pair14_j)
    assert len(topo.pair14_i) == 41  # the 41 proper 1-4 pairs of the capped dipeptide
    s14, scee = t.scales()
    assert s14 == 0.5 and scee == 1.2  # AMBER defaults without SCNB/SCEE sections
    # every 1-4 pair is also an exclusion (prmtop EXCLUDED_ATOMS_LIST)
    ex = set(zip(np.minimum(topo.excl_i, topo.excl_j), np.maximum(topo.excl_i, topo.excl_j)))
    assert all((a, b) in ex for a, b in zip(topo.pair14_i, topo.pair14_j))


def test_promote_hydrogen_bonds_matches_oracle():
    path = os.path.join(FIX, "ala_dipeptide.prmtop")
    with open(path) as f:
        atoms, topo = amber.parse_prmtop(f.read())
    amber.promote_hydrogen_bonds(topo, atoms)
    t = host.Topology(path)
    t.promote_hydrogen_bonds()
    for k in ("bond_i", "bond_j", "bond_k", "bond_r0", "constr_i", "constr_j", "constr_dsq"):
        assert t.field(k).tobytes() == getattr(topo, k).tobytes(), k
    assert len(topo.constr_i) == 12


@pytest.mark.parametrize("name", NAMES)
def test_inpcrd_matches_oracle(name):
    with open(os.path.join(FIX, f"{name}.prmtop")) as f:
        atoms, _ = amber.parse_prmtop(f.read())
    n = len(atoms.mass)
    path = os.path.join(FIX, f"{name}.inpcrd")
    with open(path) as f:
        ref = amber.parse_inpcrd(f.read(), n)
    pos, vel = host.read_inpcrd(path, n)
    assert np.stack(ref[:3]).tobytes() == pos.tobytes()
    assert np.stack(ref[3:]).tobytes() == vel.tobytes()


def test_dcd_writer_reproduces_the_golden_file(tmp_path):
    with open(os.path.join(FIX, "ala_dipeptide_ref.dcd"), "rb") as f:
        golden = f.read()
    hdr, frames = amber.read_dcd(golden)
    assert hdr["natom"] == 22 and hdr["nset"] == 100 and hdr["nsavc"] == 1
    out = tmp_path / "out.dcd"
    host.write_dcd(str(out), frames, frame_interval=1, dt_ps=0.001)  # golden_reference.rs:40-71 config
    assert out.read_bytes() == golden


def test_energy_log_format_and_temperature(tmp_path):
    """geodesic-io/tests/export_roundtrip.rs:4-25."""
    p = tmp_path / "e.csv"
    host.write_energy_log(str(p), 100, [0], [0.0], [-100.0], [0.59616123 * 150.0])
    lines = p.read_text().splitlines()
    assert lines[0] == "step,time_ps,potential_kcal,kinetic_kcal,total_kcal,temperature_K"
    f = lines[1].split(",")
    assert f[0] == "0" and f[1] == "0.000" and f[5] == "300.00"
    assert f[2] == "-100.000" and f[3] == "89.424" and f[4] == "-10.576"


def test_errors(tmp_path):
    with pytest.raises(gh.GmError) as ei:
        host.Topology(str(tmp_path / "missing.prmtop"))
    assert ei.value.status == host.GMH_ERR_IO
    text = open(os.path.join(FIX, "lj_pair.prmtop")).read()
    bad = tmp_path / "trunc.prmtop"
    bad.write_text(text[: text.index("%FLAG MASS")])
    with pytest.raises(gh.GmError) as ei:
        host.Topology(str(bad))
    assert ei.value.status == host.GMH_ERR_CONFIG and "MASS" in str(ei.value)
    # a 10-12 H-bond type: NONBONDED_PARM_INDEX entry <= 0 (prmtop.rs:221-227)
    lines = text.splitlines(keepends=True)
    k = next(i for i, ln in enumerate(lines) if ln.startswith("%FLAG NONBONDED_PARM_INDEX"))
    first = lines[k + 2]
    lines[k + 2] = f"{-1:8d}" + first[8:]
    hb = tmp_path / "hbond.prmtop"
    hb.write_text("".join(lines))
    with pytest.raises(gh.GmError) as ei:
        host.Topology(str(hb))
    assert ei.value.status == host.GMH_ERR_CONFIG and "10-12" in str(ei.value)
    with pytest.raises(gh.GmError) as ei:
        host.read_inpcrd(os.path.join(FIX, "lj_pair.inpcrd"), 3)
    assert ei.value.status == host.GMH_ERR_CONFIG and "declares 2 atoms" in str(ei.value)


HOST_LAYOUT_C = r"""
#include <stdio.h>
#include <stddef.h>
#include "geodesic_host.h"
#define F(T, m) printf(#T "." #m " %zu\n", offsetof(T, m));
int main(void) {
  printf("gmh_energy_report %zu\n", sizeof(gmh_energy_report));
  printf("gmh_run_config %zu\n", sizeof(gmh_run_config));
  printf("gmh_run_summary %zu\n", sizeof(gmh_run_summary));
  F(gmh_energy_report, nonbonded) F(gmh_energy_report, lj14) F(gmh_energy_report, r_cutoff)
  F(gmh_run_config, backend) F(gmh_run_config, box_size) F(gmh_run_config, dt) F(gmh_run_config, max_constr_iter)
  F(gmh_run_config, constr_tol) F(gmh_run_config, trajectory) F(gmh_run_config, monitoring)
  F(gmh_run_config, energy_drift_threshold_kcal) F(gmh_run_config, fidelity) F(gmh_run_config, device)
  F(gmh_run_summary, final_kinetic)
  printf("GMH_F_COUNT %d\n", (int)GMH_F_COUNT);
  printf("GMH_ERR_NUMERICAL %d\n", (int)GMH_ERR_NUMERICAL);
  return 0;
}
"""


def test_host_ctypes_structs_match_c_layout(tmp_path):
    import ctypes as C
    import subprocess
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    src = tmp_path / "layout.c"
    src.write_text(HOST_LAYOUT_C)
    exe = tmp_path / "layout"
    subprocess.check_call(["gcc", "-std=c11", "-I", os.path.join(root, "include"), str(src), "-o", str(exe)])
    out = dict(line.rsplit(" ", 1) for line in subprocess.check_output([str(exe)]).decode().splitlines())
    classes = {"gmh_energy_report": host.EnergyReport, "gmh_run_config": host.RunConfig,
               "gmh_run_summary": host.RunSummary}
    for name, cls in classes.items():
        assert int(out[name]) == C.sizeof(cls), name
    for key, val in out.items():
        if "." in key:
            t, m = key.split(".")
            assert getattr(classes[t], m).offset == int(val), key
    assert int(out["GMH_F_COUNT"]) == len(host.FIELDS) and int(out["GMH_ERR_NUMERICAL"]) == host.GMH_ERR_NUMERICAL
