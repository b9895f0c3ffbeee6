"""The reference-named Python surface (spheral_amd.Spheral{1,2,3}d): a script written the way
the reference's Sod test builds its problem maps onto the same engine configuration as the
direct ABI binding, and unsupported options raise instead of silently differing."""
import numpy as np
import pytest

from spheral_amd import Spheral1d as S1
from spheral_amd import Spheral3d as S3
from spheral_amd import abi, problems


def sod_script(n=400, two_lists=True):
    """Sod set-up in the style of tests/functional/Hydro/Sod/Sod-planar-1d.py (two NodeLists)."""
    cfg_ref, st = problems.sod_1d(n)
    eos = S1.GammaLawGas(1.4, 2.0)
    WT = S1.TableKernel(S1.BSplineKernel(), 1000)
    left = st["x"][:, 0] < 0.0
    lists = []
    parts = [left, ~left] if two_lists else [np.ones(n, bool)]
    for k, sel in enumerate(parts):
        nl = S1.makeFluidNodeList(f"nodes{k + 1}", eos, numInternal=int(sel.sum()), hmin=1e-20, hmax=1e20,
                                  nPerh=1.51, kernelExtent=WT.kernelExtent, rhoMin=1e-10)
        nl.positions()[...] = st["x"][sel]
        nl.velocity()[...] = st["v"][sel]
        nl.mass()[...] = st["m"][sel]
        nl.massDensity()[...] = st["rho"][sel]
        nl.specificThermalEnergy()[...] = st["eps"][sel]
        nl.Hfield()[...] = st["H"][sel]
        lists.append(nl)
    db = S1.DataBase()
    for nl in lists:
        db.appendNodeList(nl)
    q = S1.MonaghanGingoldViscosity(1.0, 1.0)
    hydro = S1.SPH(dataBase=db, W=WT, Q=q, cfl=0.25, compatibleEnergyEvolution=True, gradhCorrection=True,
                   XSPH=False, correctVelocityGradient=True)
    for bc in (S1.ReflectingBoundary(S1.Plane(S1.Vector(-0.5), S1.Vector(1.0))),
               S1.ReflectingBoundary(S1.Plane(S1.Vector(0.5), S1.Vector(-1.0)))):
        hydro.appendBoundary(bc)
    integ = S1.CheapSynchronousRK2Integrator(db)
    integ.appendPhysicsPackage(hydro)
    integ.lastDt = 1e-6
    integ.dtMin = 1e-8
    return cfg_ref, st, integ, lists


FIELDS = [name for name, _ in abi.SpxConfig._fields_]


def test_script_config_matches_problem_config():
    cfg_ref, _, integ, _ = sod_script()
    cfg = integ.config()
    for name in FIELDS:
        a, b = getattr(cfg, name), getattr(cfg_ref, name)
        if hasattr(a, "__len__"):
            assert list(a) == list(b), name
        else:
            assert a == b, name


def test_synchronous_rk2_script_selects_the_integrator():
    """The same Sod script under SynchronousRK2Integrator (as Sod-spherical-1d.py selects it) maps
    onto the problem's configuration with the SynchronousRK2 step and nothing else changed."""
    cfg_ref, _, cheap, _ = sod_script()
    integ = S1.SynchronousRK2Integrator(cheap.dataBase)
    for p in cheap.physicsPackages():
        integ.appendPhysicsPackage(p)
    integ.lastDt, integ.dtMin = cheap.lastDt, cheap.dtMin
    cfg = integ.config()
    cfg_ref.integrator = abi.SPX_INTEGRATOR_SYNCHRONOUS_RK2
    for name in FIELDS:
        a, b = getattr(cfg, name), getattr(cfg_ref, name)
        assert (list(a) == list(b)) if hasattr(a, "__len__") else a == b, name


def test_nodelists_concatenate_in_order():
    _, st, integ, lists = sod_script()
    g = integ._gather_state()
    for k in ("x", "v", "m", "rho", "eps", "H"):
        assert np.array_equal(g[k], st[k]), k
    assert [(a, b) for _, a, b in integ._slices] == [(0, lists[0].numInternalNodes), (lists[0].numInternalNodes, 400)]


def test_reference_defaults_that_the_gpu_path_lacks_raise():
    _, _, integ, lists = sod_script()
    db = integ.dataBase
    WT = S1.TableKernel(S1.BSplineKernel(), 1000)
    q = S1.MonaghanGingoldViscosity(1.0, 1.0)
    with pytest.raises(NotImplementedError):
        S1.SPH(dataBase=db, W=WT, Q=q)                     # factory default XSPH=True
    assert isinstance(S1.SPH(dataBase=db, W=WT, XSPH=False).Q, S1.LimitedMonaghanGingoldViscosity)
    with pytest.raises(NotImplementedError):
        S1.SPH(dataBase=db, W=WT, Q=q, XSPH=False, evolveTotalEnergy=True)
    with pytest.raises(NotImplementedError):
        S1.SPH(dataBase=db, W=WT, Q=q, XSPH=False, densityUpdate=S1.IntegrateDensity)
    with pytest.raises(NotImplementedError):
        S1.MonaghanGingoldViscosity(1.0, 1.0, linearInExpansion=True)._check()
    with pytest.raises(NotImplementedError):
        S1.Plane(S1.Vector(0.0), S1.Vector(0.5)).axis_side(1)


def test_materials_must_agree_across_nodelists():
    _, _, integ, lists = sod_script()
    lists[1].eos = S1.GammaLawGas(5.0/3.0, 2.0)
    with pytest.raises(NotImplementedError):
        integ.config()


def test_periodic_box_and_gravity_map_to_config():
    cfg_ref, st = problems.taylor_green_3d(8)
    WT = S3.TableKernel(S3.WendlandC4Kernel(), 1000)
    nl = S3.makeFluidNodeList("fluid", S3.GammaLawGas(5.0/3.0, 2.0), numInternal=len(st["m"]), nPerh=2.0,
                              kernelExtent=WT.kernelExtent)
    db = S3.DataBase()
    db.appendNodeList(nl)
    hydro = S3.SPH(dataBase=db, W=WT, Q=S3.MonaghanGingoldViscosity(1.0, 1.0), XSPH=False)
    for a in range(3):
        lo = [0.0, 0.0, 0.0]
        hi = [0.0, 0.0, 0.0]
        hi[a] = 1.0
        n = [0.0, 0.0, 0.0]
        n[a] = 1.0
        m = [-c for c in n]
        hydro.appendBoundary(S3.PeriodicBoundary(S3.Plane(S3.Vector(*lo), S3.Vector(*n)),
                                                 S3.Plane(S3.Vector(*hi), S3.Vector(*m))))
    grav = S3.OctTreeGravity(G=1.0, softeningLength=0.01, opening=0.5, ftimestep=0.1)
    integ = S3.CheapSynchronousRK2Integrator(db)
    integ.appendPhysicsPackage(hydro)
    integ.appendPhysicsPackage(grav)
    cfg = integ.config()
    assert list(cfg.bc_lo) == [abi.SPX_BC_PERIODIC]*3 and list(cfg.bc_hi) == [abi.SPX_BC_PERIODIC]*3
    assert list(cfg.xmin) == [0.0]*3 and list(cfg.xmax) == [1.0]*3
    assert cfg.kernel == abi.SPX_KERNEL_WENDLANDC4 and cfg.neighbor_extent == 1.0 and cfg.nperh == 2.0
    assert cfg.gravity == 1 and cfg.G == 1.0 and cfg.softening == 0.01 and cfg.opening == 0.5


def noh_script():
    """Planar Noh set-up written as tests/functional/Hydro/Noh/Noh-planar-1d.py builds it with its
    defaults: NBSplineKernel(5), the SPH() factory's default Q, a reflecting plane at x = 0."""
    cfg_ref, st = problems.noh_planar_1d()
    eos = S1.GammaLawGasMKS(5.0/3.0, 1.0)
    WT = S1.TableKernel(S1.NBSplineKernel(5), 1000)
    nodes1 = S1.makeFluidNodeList("nodes1", eos, numInternal=len(st["m"]), hmin=0.0001, hmax=0.1, nPerh=1.35,
                                  kernelExtent=WT.kernelExtent)
    for attr, key in (("positions", "x"), ("velocity", "v"), ("mass", "m"), ("massDensity", "rho"),
                      ("specificThermalEnergy", "eps"), ("Hfield", "H")):
        getattr(nodes1, attr)()[...] = st[key]
    db = S1.DataBase()
    db.appendNodeList(nodes1)
    hydro = S1.SPH(dataBase=db, W=WT, cfl=0.5, useVelocityMagnitudeForDt=False, compatibleEnergyEvolution=True,
                   evolveTotalEnergy=False, gradhCorrection=True, XSPH=False, correctVelocityGradient=True,
                   densityUpdate=S1.RigorousSumDensity, epsTensile=0.0, nTensile=4.0, HUpdate=S1.IdealH, ASPH=False)
    hydro.appendBoundary(S1.ReflectingBoundary(S1.Plane(S1.Vector(0.0), S1.Vector(1.0))))
    integ = S1.CheapSynchronousRK2Integrator(db)
    integ.appendPhysicsPackage(hydro)
    integ.lastDt = 0.0001
    integ.dtMin = 1.0e-5
    integ.dtMax = 0.1
    integ.dtGrowth = 2.0
    return cfg_ref, st, integ, hydro, nodes1


def test_noh_script_config_matches_problem_config():
    """The factory's default Q is LimitedMonaghanGingold with Cl = extent, Cq = extent^2/2 (3, 4.5
    for NBSpline(5)), and the script maps onto the same configuration as problems.noh_planar_1d."""
    cfg_ref, _, integ, hydro, _ = noh_script()
    assert isinstance(hydro.Q, S1.LimitedMonaghanGingoldViscosity)
    assert (hydro.Q.Cl, hydro.Q.Cq) == (3.0, 4.5) and hydro.Q.requireVelocityGradient()
    cfg = integ.config()
    for name in FIELDS:
        a, b = getattr(cfg, name), getattr(cfg_ref, name)
        if hasattr(a, "__len__"):
            assert list(a) == list(b), name
        else:
            assert a == b, name


def test_unsupported_q_options_raise():
    _, _, integ, hydro, _ = noh_script()
    hydro.Q.balsaraShearCorrection = True
    with pytest.raises(NotImplementedError):
        integ.config()


def test_nbody_script_maps_to_gravity_only_config():
    """CollisionlessSphereCollapse.py's package list (OctTreeGravity alone, SynchronousRK2Integrator)
    maps onto the gravity-only configuration of problems.collisionless_sphere_3d."""
    cfg_ref, st, _ = problems.collisionless_sphere_3d()
    eos = S3.GammaLawGasMKS(5.0/3.0, 1.0)
    nodes = S3.makeFluidNodeList("nodes", eos, numInternal=len(st["m"]))
    for attr, key in (("positions", "x"), ("velocity", "v"), ("mass", "m"), ("massDensity", "rho"),
                      ("specificThermalEnergy", "eps"), ("Hfield", "H")):
        getattr(nodes, attr)()[...] = st[key]
    db = S3.DataBase()
    db.appendNodeList(nodes)
    gravity = S3.OctTreeGravity(G=cfg_ref.G, softeningLength=cfg_ref.softening, opening=1.0, ftimestep=0.1,
                                timeStepChoice=S3.AccelerationRatio)
    integ = S3.SynchronousRK2Integrator(db)
    integ.appendPhysicsPackage(gravity)
    integ.lastDt, integ.dtMin, integ.dtMax, integ.dtGrowth = (cfg_ref.last_dt, cfg_ref.dt_min, cfg_ref.dt_max,
                                                              cfg_ref.dt_growth)
    cfg = integ.config()
    for name in ("gravity", "gravity_only", "integrator", "G", "softening", "opening", "ftimestep", "dt_min",
                 "dt_max", "dt_growth", "last_dt", "dim"):
        assert getattr(cfg, name) == getattr(cfg_ref, name), name
    cheap = S3.CheapSynchronousRK2Integrator(db)
    cheap.appendPhysicsPackage(gravity)
    with pytest.raises(NotImplementedError):
        cheap.config()


def test_node_generators_match_the_problem_setups():
    """GenerateNodeDistribution1d as Noh-planar-1d.py calls it and the 3-D lattice clipped to a
    sphere as CollisionlessSphereCollapse.py calls it give the problems' own node sets."""
    _, st = problems.noh_planar_1d()
    nodes = S1.makeFluidNodeList("nodes1", S1.GammaLawGasMKS(5.0/3.0, 1.0), nPerh=1.35)
    gen = S1.GenerateNodeDistribution1d(n=100, rho=1.0, xmin=0.0, xmax=1.0, nNodePerh=1.35)
    S1.distributeNodes1d((nodes, gen))
    assert np.array_equal(nodes.positions(), st["x"]) and np.array_equal(nodes.mass(), st["m"])
    assert np.array_equal(nodes.Hfield(), st["H"]) and np.array_equal(nodes.massDensity(), st["rho"])
    cfg, st3, _ = problems.collisionless_sphere_3d()
    r0 = 149597870700.0
    rho0 = 5.9722e24/(4.0/3.0*np.pi*r0*r0*r0)
    nodes3 = S3.makeFluidNodeList("nodes", S3.GammaLawGasMKS(5.0/3.0, 1.0))
    gen3 = S3.GenerateNodeDistribution3d(20, 20, 20, rho0, distributionType="lattice", xmin=(-r0, -r0, -r0),
                                         xmax=(r0, r0, r0), rmin=0.0, rmax=r0)
    S3.distributeNodes3d((nodes3, gen3))
    assert np.array_equal(nodes3.positions(), st3["x"])
    m = nodes3.mass()
    fmass = 5.9722e24/sum(m.tolist())        # the script's renormalisation (a sequential sum)
    assert np.array_equal(m*fmass, st3["m"])
