"""Error paths of the entry points added for the Limited Q, NBSpline, iterateIdealH and the
TreeGravity-only SynchronousRK2: each misuse returns its SPX_ERR_* code (no device work)."""
import pytest

from spheral_amd import abi, problems
from spheral_amd.engine import SpxError

pytestmark = pytest.mark.gpu


UNSUP, ARG = -4, -1


@pytest.mark.parametrize("change,code", [
    (dict(integrator=abi.SPX_INTEGRATOR_SYNCHRONOUS_RK2, gravity_only=1), UNSUP),  # needs gravity
    (dict(gravity_only=1), UNSUP),                                              # needs gravity + SynchronousRK2
    (dict(gravity_only=1, gravity=1), UNSUP),                                   # needs SynchronousRK2
    (dict(kernel=abi.SPX_KERNEL_NBSPLINE, kernel_order=11), ARG),
    (dict(viscosity=7), ARG),
    (dict(viscosity=abi.SPX_Q_LIMITED_MONAGHAN_GINGOLD, eta_fold_frac=0.0), ARG),
    (dict(integrator=5), ARG),
    # TreeGravity's DynamicalTime vote is refused, not replaced by AccelerationRatio
    (dict(gravity=1, G=1.0, softening=0.01, opening=0.5, ftimestep=0.1,
          gravity_time_step=abi.SPX_GRAVITY_DT_DYNAMICAL_TIME), UNSUP),
])
def test_bad_configurations_are_rejected(gpu, change, code):
    cfg = abi.default_config(3)
    for k, v in change.items():
        setattr(cfg, k, v)
    print(f"[case] {change}", flush=True)
    with pytest.raises(SpxError) as err:
        gpu.Run(cfg, 8, 0)
    assert err.value.code == code


def test_iterate_ideal_h_needs_initialize(gpu):
    cfg, st = problems.noh_planar_1d()
    g = gpu.make_run(cfg, st)
    with pytest.raises(SpxError) as err:
        g.iterate_ideal_h()
    assert err.value.code == -3
    g.initialize()
    assert g.iterate_ideal_h(50, 1e-4) > 1


def test_gravity_only_batch_is_rejected(gpu):
    """A TreeGravity-only problem is stepped as a single run; a batch of them is refused."""
    cfg, _, _ = problems.collisionless_sphere_3d(nr=3)
    with pytest.raises(SpxError) as err:
        gpu.Batch(cfg, 1000, 1)
    assert err.value.code == UNSUP
