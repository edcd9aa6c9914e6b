"""End to end from the reference's own checkpoint: the tecbin files its
tecbin wrote are read by the library (mf3d_tecbin_read), the pressure and
velocities restart the velocity form of newsolv_mg on the GPU, and the
result is bitwise the reference's run restarted from the same checkpoint
(tests/golden/make_golden_ckpt.py)."""
import warnings

import numpy as np
import pytest

from multiflow3d_amd import NotConverged, PressureMultigrid
from multiflow3d_amd import checkpoint as C

from .test_checkpoint import MAN, load_ckpt

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", sorted(MAN))
def test_restart_from_reference_checkpoint_matches_reference(name):
    g, m, d, e = load_ckpt(name)
    print(f"[checkpoint restart] {name}", flush=True)
    blocks = C.load_checkpoint(d, g, pl=1)
    us = [b["u"] for b in blocks]
    vs = [b["v"] for b in blocks]
    ws = [b["w"] for b in blocks]
    p = [b["p"].copy() for b in blocks]
    u, v, w = [x.copy() for x in us], [x.copy() for x in vs], [x.copy() for x in ws]
    su = [np.zeros_like(x) for x in p]
    with PressureMultigrid(g, m["levels"], maxcy=m["cycles"]) as mg:
        with warnings.catch_warnings():
            warnings.simplefilter("ignore", NotConverged)
            res = mg.newsolv(us, vs, ws, u, v, w, p, m["dt"], m["alfapr"], 0.0, su=su)
    assert res.iters == e["iters"]
    assert np.array_equal(res.hist, e["rmax"])
    for got, key in ((p, "r_p"), (u, "r_u"), (v, "r_v"), (w, "r_w")):
        for x, y in zip(got, e[key]):
            assert np.array_equal(x, y), key
    for x, y in zip(su, e["r_su"]):
        assert np.array_equal(x[1:-1, 1:-1, 1:-1], y)
