"""The reference-named Python surface drives the same GPU run as the direct binding:
a Sod script written the reference's way (two NodeLists, reflecting planes, controller)
gives bit-identical state to engine.Run on the same inputs."""
import numpy as np
import pytest

from test_mirror import sod_script
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
