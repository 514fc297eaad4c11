"""The oracle against the reference's own golden vectors (CPU).

Runs step 1 of each recorded preset tape (the 120x120x40 supercell, the
300x60x40 squall line, the 144x144x40 supercell on a stretched vertical grid)
through the oracle and compares every reproducible region-hash slot with the
hashes the reference recorded on its GPU (tests/golden/<preset>/hashes.ref.bin).
"""
import numpy as np
import pytest

import golden_check
import oracle_lib as ol
import tapeplan


# reproducible slots per preset (supercell 825, squall line 815, stretched 52)
MIN_SLOTS = {"supercell": 825, "squall_line": 815, "supercell_stretched": 52}


@pytest.mark.parametrize("preset", tapeplan.PRESETS)
def test_preset_step1_region_hashes(preset):
    P = golden_check.plan(preset)
    have = set(ol.kernel_names()) | {"_copy"}
    names, A = P.args()
    prog = ol.Program(names, A)
    M = P.memory()
    ok, bad, skipped, mism = golden_check.check(
        P, lambda i: prog.run(M, i, i), lambda off, n: ol.region_hash(M, off, n), have)
    assert bad == 0, mism[:10]
    # every dynamics / turbulence / acoustic kernel of the step is pinned
    # (on the stretched grid only what precedes turb_km_kh's pow: see
    # golden_check.UNPINNED)
    assert ok >= MIN_SLOTS[preset], ok
    assert skipped <= 16, skipped


def test_kernels_of_the_path_are_all_in_the_tape():
    """Every kernel the oracle implements is exercised by the golden tape,
    except the options the supercell run does not use (stability_max runs
    only with adapt_dt = 1; its pins are in test_adaptive.py)."""
    P = golden_check.plan()
    used = {o["name"] for o in P.ops}
    have = set(ol.kernel_names())
    unused = have - used
    allowed = {"update_diagnostics_dry", "bc_periodic_x", "bc_periodic_y", "divergence3d",
               "adv_flux_x", "adv_flux_y", "adv_flux_z", "adv_flux_x_stag", "adv_flux_y_stag",
               "adv_flux_z_stag", "adv_flux_z_weno_w", "adv_flux_x_weno_uv", "adv_flux_y_weno_uv",
               "tke_epst", "diabatic_unapply", "pdef_dum", "pdef_cell", "pdef_mass", "pdef_apply",
               "velocity_to_massflux", "stage_state_reset", "turb_sfcstress", "update_diagnostics",
               "adv_flux_z_weno_uv", "stability_max"}
    assert unused <= allowed, sorted(unused - allowed)
