"""GPU parity of each STEP 3 form against the reference's output.

The cutoff-level translation runs on target groups (form 0), on sibling
bricks as entry streams (form 1, the default where a tree allows bricks) or
on sibling bricks as a separable three-pass stencil with a second sweep
that takes the leaf-leaf pairs back out (form 2).  The form is read from
FGT_M2L_BRICK once per process, so each form runs in a child process of its
own over 3-D fixtures (free space and periodic, Legendre and Chebyshev,
norder 8 / 12 / 16, nd = 2) and must match the reference's potential to a
relative l2 of 1e-12.
"""
from __future__ import annotations

import json
import os
import subprocess
import sys

import pytest

from conftest import ROOT

pytestmark = pytest.mark.gpu

CASES = ["c1_small", "p3_n8", "p3_n16", "p2_cheb_n12", "nd2_f3_n8", "c3_n12_e6_d4", "f1_n8"]

CHILD = r"""
import json, sys
import numpy as np
sys.path.insert(0, {root!r})
sys.path.insert(0, {tests!r})
from conftest import load_case, rel_l2
import fgt_amd
out = {{}}
for name in {cases!r}:
    c = load_case(name)
    if c.ndim != 3:
        continue
    print(f"[gpu child] form {form}: {{name}}", file=sys.stderr, flush=True)
    pot = np.zeros_like(c.fvals)
    fgt_amd.boxfgt(c.nd, c.ndim, c.delta, c.eps, c.ipoly, c.iperiod, c.norder, c.npbox, c.nboxes, c.nlevels,
                   c.ltree, c.itree, c.iptr, c.centers, c.boxsize, c.fvals, 1, pot)
    s = fgt_amd.Solver(c.nd, c.ndim, c.delta, c.eps, c.ipoly, c.iperiod, c.norder, *c.tree_args())
    out[name] = {{"err": float(rel_l2(pot, c.pot_ref)), "form": int(s.stats()["m2l_form"])}}
print(json.dumps(out))
"""


@pytest.mark.parametrize("form", [0, 1, 2])
def test_m2l_form_matches_reference(form):
    print(f"[gpu] STEP 3 form {form} in a child process", flush=True)
    env = dict(os.environ, FGT_M2L_BRICK=str(form))
    code = CHILD.format(root=ROOT, tests=os.path.join(ROOT, "tests"), cases=CASES, form=form)
    r = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True, timeout=240)
    sys.stderr.write(r.stderr[-2000:])
    assert r.returncode == 0, r.stderr[-2000:]
    res = json.loads(r.stdout.strip().splitlines()[-1])
    assert res, "no 3-D case ran"
    forms = set()
    for name, v in res.items():
        print(f"[gpu]   {name}: form {v['form']} rel l2 {v['err']:.2e}", flush=True)
        assert v["err"] < 1e-12, name
        forms.add(v["form"])
    # the forced form runs wherever the tree allows it (bricks need a
    # 3-D tree without two sources at one brick position)
    assert form in forms if form > 0 else forms == {0}
