"""The drop-in called from Fortran, as the reference's users call boxfgt.

``tests/fortran/drop_in_driver.f90`` (built with the library by
``fgt_amd.build``) calls ``boxfgt`` through the reference's implicit
interface (symbol ``boxfgt_``, test/bfgt/test_boxfgt_all.f:366-369) and
``fgt_boxfgt`` / ``fgt_last_status`` through the ``fgt_amd`` module
(include/fgt_amd.f90).  The driver reads the problem from a raw binary file
written here from a golden fixture and writes ``pot`` back.
"""
from __future__ import annotations

import os
import subprocess

import numpy as np
import pytest

from conftest import load_case, rel_l2

from fgt_amd.build import FORTRAN_DRIVER

needs_driver = pytest.mark.skipif(not os.path.exists(FORTRAN_DRIVER),
                                  reason="Fortran driver not built (no Fortran compiler)")


def write_input(c, path):
    with open(path, "wb") as f:
        np.array([c.nd, c.ndim, c.ipoly, c.iperiod, c.norder, c.npbox, c.nboxes, c.nlevels, c.ltree],
                 np.int32).tofile(f)
        np.array([c.delta, c.eps], np.float64).tofile(f)
        np.ascontiguousarray(c.itree, np.int32).tofile(f)
        np.ascontiguousarray(c.iptr, np.int32).tofile(f)
        np.ascontiguousarray(c.centers, np.float64).tofile(f)
        np.ascontiguousarray(c.boxsize, np.float64).tofile(f)
        np.ascontiguousarray(c.fvals, np.float64).tofile(f)


@needs_driver
def test_fortran_caller_gets_invalid_option_status(tmp_path):
    """An invalid ifnewtree (2) through the module: status 2, message,
    outputs untouched (no kernel runs, so this needs no GPU)."""
    c = load_case("c1_small")
    inp = tmp_path / "in.bin"
    write_input(c, inp)
    r = subprocess.run([FORTRAN_DRIVER, str(inp), str(tmp_path / "out.bin"), "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    lines = r.stdout.split("\n")
    assert any(ln.split() == ["status", "2"] for ln in lines), r.stdout
    assert "ifnewtree" in r.stdout
    pot_line = [ln for ln in lines if ln.startswith("pot max")][0]
    assert float(pot_line.split()[-1]) == 0.0


@needs_driver
@pytest.mark.gpu
@pytest.mark.parametrize("name", ["c1_small", "p3_n8"])
def test_fortran_caller_matches_reference(tmp_path, name):
    """boxfgt called from Fortran with the reference's argument list: the
    potential matches the reference's own output (relative l2 <= 1e-12)."""
    print(f"[gpu] Fortran drop-in: {name}", flush=True)
    c = load_case(name)
    inp, out = tmp_path / "in.bin", tmp_path / "out.bin"
    write_input(c, inp)
    r = subprocess.run([FORTRAN_DRIVER, str(inp), str(out)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "status 0" in r.stdout, r.stdout
    pot = np.fromfile(out, np.float64).reshape(c.fvals.shape)
    err = rel_l2(pot, c.pot_ref)
    print(f"[gpu]   rel l2 = {err:.3e}", flush=True)
    assert err < 1e-12
