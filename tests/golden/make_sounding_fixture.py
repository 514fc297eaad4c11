"""Copies the reference's bundled CM1 input_sounding file (a data file) into
tests/golden/soundings/, unchanged, with a provenance note.

Run in the build container (the reference tree is not present on GPU hosts):

    python tests/golden/make_sounding_fixture.py [/root/reference]
"""
import os
import shutil
import sys

ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "soundings")
os.makedirs(dst, exist_ok=True)
name = "hurricane_3d.input_sounding"
shutil.copyfile(os.path.join(ref, "data", "presets", name), os.path.join(dst, name))
with open(os.path.join(dst, "PROVENANCE.txt"), "w") as fh:
    fh.write(
        "hurricane_3d.input_sounding is a data file of ThunderEngine (defds7721-cyber/thunderengine-web),\n"
        "copyright (C) 2026 vvas2002, licensed AGPL-3.0-only: data/presets/hurricane_3d.input_sounding,\n"
        "the sounding of its hurricane_3d preset (CM1 isnd = 7 format: p_sfc mb, theta_sfc K, qv_sfc g/kg;\n"
        "then z m, theta K, qv g/kg, u m/s, v m/s per level), copied unchanged by\n"
        "tests/golden/make_sounding_fixture.py.\n")
