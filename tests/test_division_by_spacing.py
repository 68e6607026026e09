"""The division by a grid spacing of the small-grid loop (div_h in
pseudotransientstokes.jl_amd/csrc/pts_step2d.hip): Markstein's correction of
x·(1/h) must give x/h rounded to nearest, bit for bit, so that the loop stays
bit-identical to the per-kernel path and the oracle (which divide).  A C
program (the host's fma and division) checks random numerators and those
whose quotients lie next to a midpoint between doubles, for the spacings of
the tested grids.  The GPU tests confirm the device code on real states."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc")
def test_division_by_spacing_is_correctly_rounded(tmp_path):
    exe = str(tmp_path / "div_by_spacing_check")
    subprocess.run(["gcc", "-O2", "-ffp-contract=off", "-o", exe, os.path.join(HERE, "div_by_spacing_check.c"),
                    "-lm"], check=True)
    r = subprocess.run([exe, "2000000", "500000"], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout
    assert r.stdout.strip().startswith("0 of "), r.stdout


def test_a_plausible_mistake_is_caught(tmp_path):
    # the same check against the product x·(1/h) alone (no correction step)
    # must find differences: the check can tell the two apart
    if shutil.which("gcc") is None:
        pytest.skip("needs gcc")
    src = open(os.path.join(HERE, "div_by_spacing_check.c")).read()
    src = src.replace("return fma(-r, rh, q);", "(void)r;\n    return q;")
    p = tmp_path / "plain_product.c"
    p.write_text(src)
    exe = str(tmp_path / "plain_product")
    subprocess.run(["gcc", "-O2", "-ffp-contract=off", "-o", exe, str(p), "-lm"], check=True)
    r = subprocess.run([exe, "20000", "5000"], capture_output=True, text=True, timeout=300)
    assert r.returncode != 0, r.stdout
