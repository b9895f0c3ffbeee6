"""The reference-named Python surface drives the same GPU run as the direct binding:
a Sod script written the reference's way (two NodeLists, reflecting planes, controller)
gives bit-identical state to engine.Run on the same inputs."""
import numpy as np
import pytest

from test_mirror import noh_script, sod_script
from spheral_amd import Spheral1d as S1

pytestmark = pytest.mark.gpu


def test_sod_script_equals_direct_run(gpu):
    print("[case] mirror sod1d", flush=True)
    cfg, st, integ, lists = sod_script()
    control = S1.SpheralController(integ, kernel=None, iterateInitialH=False)
    control.advance(0.02)
    r = gpu.make_run(cfg, st)
    r.initialize()
    k, t = r.advance(0.02, 100000)
    assert control.totalSteps == k and integ.currentTime == t
    x = np.concatenate([nl.positions() for nl in lists])
    rho = np.concatenate([nl.massDensity() for nl in lists])
    eps = np.concatenate([nl.specificThermalEnergy() for nl in lists])
    assert np.array_equal(x, r.field("position"))
    assert np.array_equal(rho, r.field("mass density"))
    assert np.array_equal(eps, r.field("specific thermal energy"))


def test_noh_script_reaches_the_golden(gpu):
    """The Noh script through the reference-named surface (controller with iterateInitialH,
    step(5), advance(0.6)) gives the direct run's state bit for bit and the reference's norms."""
    import noh_planar as NP
    ref = NP.reference_norms()
    print("[case] mirror noh", flush=True)
    cfg, st, integ, hydro, nodes1 = noh_script()
    control = S1.SpheralController(integ, kernel=None)
    control.step(5)
    control.advance(ref["goal_time"])
    r = gpu.make_run(cfg, st)
    r.initialize()
    r.iterate_ideal_h()
    NP.drive(r, ref["goal_time"])
    assert np.array_equal(nodes1.positions(), r.field("position"))
    assert np.array_equal(nodes1.massDensity(), r.field("mass density"))
    ours = NP.norms(nodes1.positions()[:, 0], nodes1.massDensity(), r.field("pressure"), nodes1.velocity()[:, 0],
                    nodes1.specificThermalEnergy(), 1.0/nodes1.Hfield()[:, 0, 0], integ.currentTime, 0.0135,
                    ref["rmin"], ref["rmax"])
    assert NP.within(ours, ref["norms"], ref["tol"]) == []
