"""The programs over the C ABI, without a GPU: the LUT tool writes the
shipped LUT file byte for byte (create_LUT.c's output for the 40-beta set),
and the simulation program's argument handling and error exits."""
import hashlib
import json
import os
import subprocess

import numpy as np
import pytest

from conftest import GOLDEN, ROOT

PKG = os.path.join(ROOT, "ea-lcd-mc_amd")
EA_TUBE = os.path.join(PKG, "ea_tube")
EA_LUT = os.path.join(PKG, "ea_make_lut")

needs_build = pytest.mark.skipif(
    not (os.path.exists(EA_TUBE) and os.path.exists(EA_LUT)),
    reason="programs not built (make -C ea-lcd-mc_amd)")


@needs_build
def test_lut_tool_writes_shipped_file(tmp_path):
    info = json.load(open(os.path.join(GOLDEN, "lut_nb40.json")))
    (tmp_path / "betas").write_text("".join("%.17g\n" % b for b in info["betas"]))
    r = subprocess.run([EA_LUT, "betas"], cwd=tmp_path, capture_output=True,
                       text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    out = tmp_path / "LUT_for_PRNG_nbits12_NB40.bin"
    blob = out.read_bytes()
    assert len(blob) == info["bytes"]
    assert hashlib.sha256(blob).hexdigest() == info["sha256"]


@needs_build
def test_lut_tool_refuses_unsorted_betas(tmp_path):
    (tmp_path / "b").write_text("0.5\n0.9\n")
    r = subprocess.run([EA_LUT, "b"], cwd=tmp_path, capture_output=True,
                       text=True, timeout=60)
    assert r.returncode == 1 and "ordenadas" in r.stderr


@needs_build
def test_program_help_and_usage():
    r = subprocess.run([EA_TUBE, "-h"], capture_output=True, text=True,
                       timeout=60)
    assert r.returncode == 0 and "isample nbits beta.dat" in r.stdout
    r = subprocess.run([EA_TUBE, "1", "2"], capture_output=True, text=True,
                       timeout=60)
    assert r.returncode == 1 and "Usage:" in r.stderr
    r = subprocess.run([EA_TUBE, "--nr", "3", "0", "1", "b", "i", "l", "0"],
                       capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "power of two" in r.stderr


@needs_build
def test_program_argument_errors(tmp_path):
    r = subprocess.run([EA_TUBE, "0", "40", "b", "i", "l", "0"],
                       capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "nbits out of range" in r.stderr
    r = subprocess.run([EA_TUBE, "-1", "1", "b", "i", "l", "0"],
                       capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "sample out of range" in r.stderr
    r = subprocess.run([EA_TUBE, "20000", "1", "b", "i", "l", "0"],
                       capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "cannot accomodate" in r.stderr
