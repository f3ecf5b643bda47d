"""FormBCS's three passes (bcs.c:4483-5872) on faces that are only walls,
periodic pairs or an inlet: the second and third passes rewrite the values
the first left, bit for bit.

Why: each pass recomputes Ucat from Ucont (unchanged unless a far-field,
slip or outlet face corrects Ucont inside the pass), then the ghosts.  Wall
ghosts and the periodic plane copies take only interior values; the edge
averages take those face ghosts, and every corner's last write (the edge
sweeps along k, which run last, or the periodic full-plane copies after
them) takes edges of the same pass.  So with no far-field, slip or outlet face the ghost state
after one pass is a fixed point of the pass, and the device path makes one
pass there.  This test checks that claim on the oracle, which keeps the
reference's three passes by default: one pass against three, on random
Ucont, Ucat (ghosts included) and P around an immersed sphere (whose IB
cells keep their Ucat), over the face types the claim covers.
The far-field, slip and outlet cases are the control: there the passes do
change the state, so the comparison can fail."""
import numpy as np
import pytest

from curvib_old_amd import cases
from oracle import pyoracle as po

FIXED_POINT = [(7, 7, 1, 1, 7, 7), (1, 1, 1, 1, 1, 1), (7, 7, 1, 1, 1, 1), (1, 1, 1, 1, 7, 7),
               (7, 7, 1, 1, 5, 1), (1, 1, 1, 1, 5, 1)]
MOVING = [(7, 7, 1, 6, 6, 6), (3, 3, 1, 3, 3, 3), (7, 7, 1, 1, 5, 4)]


def formbcs_state(bctype, passes, seed=5):
    inlet = 7 if bctype[4] == 5 else 2
    case = cases.ChannelCase(cells=(10, 9, 8), amp=0.1, sphere=(0.5, 0.5, 0.5, 0.35), seed=3, bctype=bctype,
                             inlet=inlet, inlet_center=(0.0, 0.5, 0.0))
    xyz, nv, uc = cases.build(case)
    o = po.Oracle(case.nodes, case.bctype, 1, case.re, case.dt)
    try:
        o.set_inlet(case.inlet, case.inlet_center)
        o.set_grid(xyz)
        o.set_nvert(nv)
        rng = np.random.default_rng(seed)
        o.put("ucont", uc + 0.05 * rng.standard_normal(uc.shape))
        o.put("ucat", rng.standard_normal(uc.shape))
        o.put("p", rng.standard_normal(uc.shape[:3]))
        o.call("o_inflowflux")
        o.set_formbcs_passes(passes)
        o.call("o_formbcs")
        return {f: o.get(f) for f in ("ucont", "ucat", "p", "nvert")}
    finally:
        o.close()


@pytest.mark.parametrize("bctype", FIXED_POINT)
def test_one_formbcs_pass_equals_three_without_far_slip_or_outlet(bctype):
    one, three = formbcs_state(bctype, 1), formbcs_state(bctype, 3)
    for f in one:
        np.testing.assert_array_equal(one[f], three[f], err_msg=f"{bctype} {f}")


@pytest.mark.parametrize("bctype", MOVING)
def test_far_field_slip_and_outlet_passes_change_the_state(bctype):
    one, three = formbcs_state(bctype, 1), formbcs_state(bctype, 3)
    assert any(not np.array_equal(one[f], three[f]) for f in one), bctype
