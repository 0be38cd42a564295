"""Host-side checks of the C-ABI that need no GPU: the library loads, exports
every entry point include/geodesic_hip.h declares, reports a missing device
as GM_ERR_NO_DEVICE (BackendError::NoAdapter, error.rs:82-91), and the ctypes
mirror of every struct matches the C layout."""
import ctypes as C
import os
import subprocess

import pytest

import geodesic_hip as gh

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_declared_symbol():
    names = gh.declared_symbols()
    assert len(names) >= 16
    L = gh.lib()
    for n in names:
        assert hasattr(L, n), n


def test_missing_device_maps_to_no_adapter():
    if gh.device_count() > 0:
        pytest.skip("a GPU is visible")
    with pytest.raises(gh.GmError) as ei:
        gh.Context(0)
    assert ei.value.status == 4 and "no HIP device" in str(ei.value)


def test_last_error_is_a_string():
    assert isinstance(gh.lib().gm_last_error(), bytes)


LAYOUT_C = r"""
#include <stdio.h>
#include <stddef.h>
#include "geodesic_hip.h"
#define F(T, m) printf(#T "." #m " %zu\n", offsetof(T, m));
int main(void) {
  printf("gm_system_desc %zu\n", sizeof(gm_system_desc));
  printf("gm_energy_terms %zu\n", sizeof(gm_energy_terms));
  printf("gm_stats %zu\n", sizeof(gm_stats));
  printf("gm_profile %zu\n", sizeof(gm_profile));
  F(gm_system_desc, box) F(gm_system_desc, periodic) F(gm_system_desc, r_cutoff)
  F(gm_system_desc, scale_14) F(gm_system_desc, bonded_min_image) F(gm_system_desc, n_bonds)
  F(gm_system_desc, n_angles) F(gm_system_desc, n_dihedrals) F(gm_system_desc, dihed_delta)
  F(gm_system_desc, n_excl) F(gm_system_desc, n_pairs14) F(gm_system_desc, pair14_j)
  F(gm_stats, n_tiles_cutoff) F(gm_stats, n_pairs_cutoff) F(gm_stats, ms_nonbonded) F(gm_stats, ms_build)
  F(gm_stats, n_partner_entries) F(gm_stats, pair_layout) F(gm_stats, n_lj_types)
  printf("gm_surrogate_desc %zu\n", sizeof(gm_surrogate_desc));
  printf("gm_train_desc %zu\n", sizeof(gm_train_desc));
  printf("gm_train_report %zu\n", sizeof(gm_train_report));
  F(gm_surrogate_desc, fade) F(gm_surrogate_desc, atom_class) F(gm_surrogate_desc, seed)
  F(gm_train_desc, mu) F(gm_train_desc, cap) F(gm_train_desc, lr) F(gm_train_desc, batch) F(gm_train_desc, update)
  F(gm_train_report, loss_taylor) F(gm_train_report, ms_total) F(gm_train_report, steps)
  printf("gm_md_desc %zu\n", sizeof(gm_md_desc));
  F(gm_md_desc, dt) F(gm_md_desc, frame_interval) F(gm_md_desc, max_iter) F(gm_md_desc, tol)
  F(gm_md_desc, first_step) F(gm_md_desc, resume)
  printf("gm_md_failure %zu\n", sizeof(gm_md_failure));
  F(gm_md_failure, component) F(gm_md_failure, step) F(gm_md_failure, atom)
  printf("GM_ERR_NUMERICAL %d\n", (int)GM_ERR_NUMERICAL);
  printf("GM_ERR_CONVERGENCE %d\n", (int)GM_ERR_CONVERGENCE);
  printf("GM_SUR_NPARAM %d\n", (int)GM_SUR_NPARAM);
  printf("GM_SUR_NFEAT %d\n", (int)GM_SUR_NFEAT);
  return 0;
}
"""


def test_ctypes_structs_match_c_layout(tmp_path):
    src = tmp_path / "layout.c"
    src.write_text(LAYOUT_C)
    exe = tmp_path / "layout"
    subprocess.check_call(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)])
    out = dict(line.rsplit(" ", 1) for line in subprocess.check_output([str(exe)]).decode().splitlines())
    assert int(out["gm_system_desc"]) == C.sizeof(gh.SystemDesc)
    assert int(out["gm_energy_terms"]) == C.sizeof(gh.EnergyTerms)
    assert int(out["gm_stats"]) == C.sizeof(gh.Stats)
    assert int(out["gm_profile"]) == C.sizeof(gh.Profile)
    assert int(out["gm_surrogate_desc"]) == C.sizeof(gh.SurrogateDesc)
    assert int(out["gm_train_desc"]) == C.sizeof(gh.TrainDesc)
    assert int(out["gm_train_report"]) == C.sizeof(gh.TrainReport)
    assert int(out["GM_SUR_NPARAM"]) == gh.SUR_NPARAM and int(out["GM_SUR_NFEAT"]) == gh.SUR_NFEAT
    assert int(out["gm_md_desc"]) == C.sizeof(gh.MdDesc) and int(out["GM_ERR_CONVERGENCE"]) == 9
    assert int(out["gm_md_failure"]) == C.sizeof(gh.MdFailure)
    assert int(out["GM_ERR_NUMERICAL"]) == gh.GM_ERR_NUMERICAL
    classes = {"gm_system_desc": gh.SystemDesc, "gm_stats": gh.Stats, "gm_surrogate_desc": gh.SurrogateDesc,
               "gm_train_desc": gh.TrainDesc, "gm_train_report": gh.TrainReport, "gm_md_desc": gh.MdDesc,
               "gm_md_failure": gh.MdFailure}
    for key, val in out.items():
        if "." in key:
            t, m = key.split(".")
            assert getattr(classes[t], m).offset == int(val), key
