"""Reference-named Python surface of the SPH + TreeGravity path.

A Spheral script builds its problem from a handful of objects -- TableKernel,
makeFluidNodeList, DataBase, (Limited)MonaghanGingoldViscosity, SPH, TreeGravity, boundaries,
CheapSynchronousRK2Integrator, SpheralController -- and then calls
``control.advance(goalTime)``.  This module provides those names with the reference's
keyword arguments and defaults, and maps them onto one engine run (spx_run) on the GPU.
Use it through the per-dimension modules::

    from spheral_amd.Spheral3d import *

Reference signatures followed (paths relative to the reference's src/):
  TableKernel(kernel, numPoints=100, minNperh=0.25, maxNperh=64)    PYB11/Kernel/Kernel.py:260-275
  makeFluidNodeList(name, eos, ..., nPerh=2.01, kernelExtent=2.0)   NodeList/FluidNodeLists.py:15-35
  MonaghanGingoldViscosity(Clinear, Cquadratic, ...)                PYB11/ArtificialViscosity/MonaghanGingoldViscosity.py:25-30
  LimitedMonaghanGingoldViscosity(Clinear, Cquadratic, ...,
      etaCritFrac=1.0, etaFoldFrac=0.2)                             PYB11/ArtificialViscosity/LimitedMonaghanGingoldViscosity.py:24-32
  NBSplineKernel(order)                                             Kernel/NBSplineKernel.hh:20-30
  SPH(W, WPi=None, ..., XSPH=True, ...)                             SPH/SPHHydros.py:9-33
  TreeGravity(G, softeningLength, opening, ftimestep, ...)          PYB11/Gravity/TreeGravity.py:23-28
  SpheralController(integrator, kernel=None, ...).advance(goal)    SimulationControl/SpheralController.py:23-57,383-426

Options the GPU path does not implement raise NotImplementedError at construction
(or at the first step) instead of silently computing something else.  Several fluid
NodeLists with the same equation of state and node parameters are evolved as one run:
with sumMassDensityOverAllNodeLists=True the reference treats pairs across NodeLists
exactly as pairs within one.
"""
import numpy as np

from . import abi, connectivity, engine

# Enumerations used as keyword values in the reference's scripts.
RigorousSumDensity = "RigorousSumDensity"
IntegrateDensity = "IntegrateDensity"
SumDensity = "SumDensity"
IdealH = "IdealH"
IntegrateH = "IntegrateH"
GatherScatter = "GatherScatter"


def _tuple(v, dim):
    a = np.zeros(3)
    vv = np.atleast_1d(np.asarray(v, np.float64))
    a[:min(3, vv.size)] = vv[:3]
    return a


class Vector(tuple):
    """Geometric vector (GeomVector): Vector(x[, y[, z]])."""

    def __new__(cls, *c):
        return super().__new__(cls, tuple(float(x) for x in c))


class Plane:
    """Geometry/GeomPlane: a point on the plane and its normal."""

    def __init__(self, point, normal):
        self.point = np.asarray(point, np.float64)
        self.normal = np.asarray(normal, np.float64)

    def axis_side(self, dim):
        """(axis, side) for an axis-aligned plane: side 0 = low boundary (normal +e_a)."""
        n = self.normal[:dim]
        nz = np.flatnonzero(np.abs(n) > 0.0)
        if len(nz) != 1 or abs(abs(n[nz[0]]) - 1.0) > 1e-12:
            raise NotImplementedError("boundary planes must be axis-aligned with unit normals")
        a = int(nz[0])
        return a, (0 if n[a] > 0 else 1)


class PeriodicBoundary:
    """Boundary/PeriodicBoundary(plane1, plane2)."""

    def __init__(self, plane1, plane2):
        self.planes = (plane1, plane2)


class ReflectingBoundary:
    """Boundary/ReflectingBoundary(plane)."""

    def __init__(self, plane):
        self.plane = plane


class _BaseKernel:
    code = None
    extent = None

    def __init__(self, *args, **kwargs):
        if args or kwargs:
            raise NotImplementedError(f"{type(self).__name__} takes no parameters on the GPU path")


class BSplineKernel(_BaseKernel):
    code = abi.SPX_KERNEL_BSPLINE
    extent = 2.0


class WendlandC4Kernel(_BaseKernel):
    code = abi.SPX_KERNEL_WENDLANDC4
    extent = 1.0


class NBSplineKernel(_BaseKernel):
    """Kernel/NBSplineKernel(order): centred B-spline with order + 1 knots, extent (order + 1)/2
    in integer arithmetic (Kernel/NBSplineKernel.cc:99)."""
    code = abi.SPX_KERNEL_NBSPLINE

    def __init__(self, order):
        if not 2 <= int(order) <= 9:
            raise NotImplementedError("NBSplineKernel orders 2..9 are implemented on the GPU path")
        self.order = int(order)
        self.extent = abi.kernel_extent(abi.SPX_KERNEL_NBSPLINE, self.order)


class TableKernel:
    def __init__(self, kernel, numPoints=100, minNperh=0.25, maxNperh=64.0):
        if not isinstance(kernel, _BaseKernel):
            raise NotImplementedError("TableKernel on the GPU path supports BSplineKernel, WendlandC4Kernel and "
                                      "NBSplineKernel")
        self.base = kernel
        self.numPoints = int(numPoints)
        self.minNperh = float(minNperh)
        self.maxNperh = float(maxNperh)
        self.kernelExtent = kernel.extent


class GammaLawGas:
    """Material/GammaLawGas(gamma, mu, ...): P = (gamma - 1) rho eps, cs^2 = gamma (gamma - 1) eps."""

    def __init__(self, gamma, mu=2.0, constants=None, minimumPressure=-1e200, maximumPressure=1e200, **kwargs):
        if minimumPressure != -1e200 or maximumPressure != 1e200 or kwargs:
            raise NotImplementedError("pressure limits and external pressure are not implemented on the GPU path")
        self.gamma = float(gamma)
        self.mu = float(mu)


GammaLawGasMKS = GammaLawGas
GammaLawGasCGS = GammaLawGas


class FluidNodeList:
    """A FluidNodeList's fields as host arrays in the reference's record layout."""

    def __init__(self, name, eos, dim, numInternal, hmin, hmax, nPerh, rhoMin, rhoMax, kernelExtent):
        self.name = name
        self.eos = eos
        self.dim = dim
        self.hmin, self.hmax, self.nPerh = float(hmin), float(hmax), float(nPerh)
        self.rhoMin, self.rhoMax = float(rhoMin), float(rhoMax)
        self.kernelExtent = float(kernelExtent)
        self.numInternalNodes = 0
        self.numNodes = 0
        self._integrator = None
        self.resize(numInternal)

    def resize(self, n):
        d = self.dim
        self.numInternalNodes = self.numNodes = int(n)
        self._x = np.zeros((n, d))
        self._v = np.zeros((n, d))
        self._m = np.zeros(n)
        self._rho = np.zeros(n)
        self._eps = np.zeros(n)
        self._H = np.zeros((n, d, d))

    def _sync(self):
        if self._integrator is not None:
            self._integrator._download()

    def positions(self):
        self._sync()
        return self._x

    def velocity(self):
        self._sync()
        return self._v

    def mass(self):
        self._sync()
        return self._m

    def massDensity(self):
        self._sync()
        return self._rho

    def specificThermalEnergy(self):
        self._sync()
        return self._eps

    def Hfield(self):
        self._sync()
        return self._H


class GenerateNodeDistribution1d:
    """NodeGenerators/GenerateNodeDistribution1d.py: n equally spaced nodes on [xmin, xmax],
    x_i = xmin + (i + 1/2) dx, m = rho(x) dx, H = 1/(nNodePerh dx)."""

    def __init__(self, n, rho, xmin, xmax, nNodePerh=2.01, offset=0.0, rejecter=None):
        if rejecter is not None or offset != 0.0:
            raise NotImplementedError("node rejecters and offsets are not implemented")
        assert xmax > xmin and n > 0
        rhof = (lambda x: float(rho)) if isinstance(rho, float) else rho
        dx = (xmax - xmin)/n
        h0 = nNodePerh*dx
        self.x = [xmin + (i + 0.5)*dx for i in range(n)]
        self.m = [rhof(xi)*dx for xi in self.x]
        self.rho = [rhof(xi) for xi in self.x]
        self.H = [np.array([[1.0/h0]]) for _ in range(n)]
        self.dim = 1

    def positions(self):
        return np.array(self.x).reshape(-1, 1)


class GenerateNodeDistribution3d:
    """NodeGenerators/GenerateNodeDistribution3d.py, distributionType "lattice"
    (latticeDistribution, lines 573-624): cell centres of an n1 x n2 x n3 lattice on [xmin, xmax],
    kept when rmin^2 <= |x - origin|^2 <= rmax^2; m = dx dy dz rho(x); H = diag(1/(nNodePerh d))."""

    def __init__(self, n1, n2, n3, rho, distributionType="lattice", xmin=(0.0, 0.0, 0.0), xmax=(1.0, 1.0, 1.0),
                 rmin=None, rmax=None, origin=(0.0, 0.0, 0.0), nNodePerh=2.01, **kwargs):
        if distributionType != "lattice" or kwargs:
            raise NotImplementedError("GenerateNodeDistribution3d: only the lattice distribution is implemented")
        rhof = (lambda x: float(rho)) if isinstance(rho, float) else rho
        dx, dy, dz = ((xmax[a] - xmin[a])/n for a, n in zip(range(3), (n1, n2, n3)))
        H0 = np.diag([1.0/(nNodePerh*dx), 1.0/(nNodePerh*dy), 1.0/(nNodePerh*dz)])
        self.x, self.m, self.H, self.rho = [], [], [], []
        for ig in range(n1*n2*n3):
            i, j, k = ig % n1, (ig // n1) % n2, ig // (n1*n2)
            xx, yy, zz = xmin[0] + (i + 0.5)*dx, xmin[1] + (j + 0.5)*dy, xmin[2] + (k + 0.5)*dz
            r2 = (xx - origin[0])**2 + (yy - origin[1])**2 + (zz - origin[2])**2
            if (rmin is None or r2 >= rmin**2) and (rmax is None or r2 <= rmax**2):
                self.x.append((xx, yy, zz))
                self.m.append(dx*dy*dz*rhof((xx, yy, zz)))
                self.H.append(H0)
                self.rho.append(rhof((xx, yy, zz)))
        self.dim = 3

    def positions(self):
        return np.array(self.x).reshape(-1, 3)


def _distribute(*pairs):
    """SortAndDivideRedistributeNodes / PeanoHilbertDistributeNodes on one process: each NodeList
    takes its generator's nodes (positions, masses, mass density, H) in generator order."""
    for nodes, gen in pairs:
        n = len(gen.m)
        nodes.resize(n)
        nodes._x[...] = gen.positions()
        nodes._m[...] = np.asarray(gen.m)
        nodes._rho[...] = np.asarray(gen.rho)
        nodes._H[...] = np.stack(gen.H) if n else nodes._H


def distributeNodes1d(*pairs):
    _distribute(*pairs)


def distributeNodes3d(*pairs, **kwargs):
    _distribute(*pairs)


def makeFluidNodeList(name, eos, numInternal=0, numGhost=0, hmin=1.0e-20, hmax=1.0e20, hminratio=0.1, nPerh=2.01,
                      maxNumNeighbors=500, rhoMin=1.0e-10, rhoMax=1e10, NeighborType=None, searchType=GatherScatter,
                      kernelExtent=2.0, xmin=None, xmax=None, dim=3):
    if searchType != GatherScatter:
        raise NotImplementedError("only GatherScatter neighbour searches are implemented")
    return FluidNodeList(name, eos, dim, numInternal, hmin, hmax, nPerh, rhoMin, rhoMax, kernelExtent)


class DataBase:
    def __init__(self, dim=3):
        self.nDim = dim
        self.nodeLists = []

    def appendNodeList(self, nodes):
        if nodes.dim != self.nDim:
            raise ValueError("NodeList dimension does not match the DataBase")
        self.nodeLists.append(nodes)

    @property
    def numNodeLists(self):
        return len(self.nodeLists)

    numFluidNodeLists = numNodeLists

    @property
    def numSolidNodeLists(self):
        return 0

    @property
    def maxKernelExtent(self):
        return max(nl.kernelExtent for nl in self.nodeLists)


class MonaghanGingoldViscosity:
    code = abi.SPX_Q_MONAGHAN_GINGOLD

    def __init__(self, Clinear=1.0, Cquadratic=1.0, kernel=None, linearInExpansion=False, quadraticInExpansion=False):
        self.Cl = float(Clinear)
        self.Cq = float(Cquadratic)
        # ArtificialViscosityHandle defaults (ArtificialViscosityHandle.cc:43-50); checked when the
        # run is configured, so a script may set them as the reference's scripts do
        self.linearInExpansion = bool(linearInExpansion)
        self.quadraticInExpansion = bool(quadraticInExpansion)
        self.epsilon2 = 1.0e-2
        self.balsaraShearCorrection = False
        self.rigorousVelocityGradient = False

    def requireVelocityGradient(self):
        return False

    def _check(self):
        if self.linearInExpansion or self.quadraticInExpansion:
            raise NotImplementedError(f"{type(self).__name__} in-expansion terms are not implemented")
        if self.epsilon2 != 1.0e-2:
            raise NotImplementedError("a Q epsilon2 other than 1e-2 is not implemented")
        if self.balsaraShearCorrection:
            raise NotImplementedError("the Balsara shear correction is not implemented")
        if self.rigorousVelocityGradient:
            raise NotImplementedError("rigorousVelocityGradient is not implemented")


class LimitedMonaghanGingoldViscosity(MonaghanGingoldViscosity):
    """ArtificialViscosity/LimitedMonaghanGingoldViscosity: the SPH() factory's default Q."""
    code = abi.SPX_Q_LIMITED_MONAGHAN_GINGOLD

    def __init__(self, Clinear=1.0, Cquadratic=1.0, kernel=None, linearInExpansion=False, quadraticInExpansion=False,
                 etaCritFrac=1.0, etaFoldFrac=0.2):
        super().__init__(Clinear, Cquadratic, kernel, linearInExpansion, quadraticInExpansion)
        self.etaCritFrac = float(etaCritFrac)
        self.etaFoldFrac = float(etaFoldFrac)

    def requireVelocityGradient(self):
        return True


class _Package:
    def __init__(self):
        self.boundaries = []

    def appendBoundary(self, bc):
        self.boundaries.append(bc)

    def boundaryConditions(self):
        return list(self.boundaries)


class SPHPackage(_Package):
    """The SPH package with its SPHSmoothingScale and viscosity (SPH/SPHHydros.py:88-140)."""

    def __init__(self, W, dataBase, Q, **opts):
        super().__init__()
        self.W = W
        self.dataBase = dataBase
        self.Q = Q
        self.opts = opts
        self.compatibleEnergyEvolution = opts["compatibleEnergyEvolution"]
        self.cfl = opts["cfl"]

    def label(self):
        return "SPH"


def SPH(W, WPi=None, WGrad=None, dataBase=None, Q=None, filter=None, cfl=0.25, useVelocityMagnitudeForDt=False,
        compatibleEnergyEvolution=True, evolveTotalEnergy=False, gradhCorrection=True, XSPH=True,
        correctVelocityGradient=True, sumMassDensityOverAllNodeLists=True, densityUpdate=RigorousSumDensity,
        HUpdate=IdealH, epsTensile=0.0, nTensile=4.0, damageRelieveRubble=False, strengthInDamage=False,
        xmin=(-1e100, -1e100, -1e100), xmax=(1e100, 1e100, 1e100), etaMinAxis=0.1, ASPH=False,
        smoothingScaleMethod=None):
    """SPH factory with the reference's keywords and defaults (SPH/SPHHydros.py:9-33)."""
    def no(cond, what):
        if cond:
            raise NotImplementedError(f"SPH: {what} is not implemented on the GPU path")
    no(dataBase is None, "a DataBase is required;")
    no(dataBase.numSolidNodeLists > 0, "solid NodeLists")
    if Q is None:
        # SPH/SPHHydros.py:88-94 (Cl, Cq pass through "%g" there: exact for these values)
        ext = dataBase.maxKernelExtent
        Q = LimitedMonaghanGingoldViscosity(Clinear=2.0*(ext/2.0), Cquadratic=2.0*(ext/2.0)**2, kernel=W)
    no(not isinstance(Q, MonaghanGingoldViscosity), f"viscosity {type(Q).__name__}")
    no(WPi is not None and WPi is not W, "a separate WPi kernel")
    no(not compatibleEnergyEvolution, "compatibleEnergyEvolution=False")
    no(evolveTotalEnergy, "evolveTotalEnergy=True")
    no(XSPH, "XSPH=True (pass XSPH=False)")
    no(not sumMassDensityOverAllNodeLists, "sumMassDensityOverAllNodeLists=False")
    no(densityUpdate != RigorousSumDensity, f"densityUpdate={densityUpdate}")
    no(HUpdate != IdealH, f"HUpdate={HUpdate}")
    no(epsTensile != 0.0, "tensile correction (epsTensile != 0)")
    no(ASPH or smoothingScaleMethod is not None, "ASPH / custom smoothing-scale methods")
    return SPHPackage(W, dataBase, Q, cfl=float(cfl), useVelocityMagnitudeForDt=bool(useVelocityMagnitudeForDt),
                      compatibleEnergyEvolution=True, gradhCorrection=bool(gradhCorrection),
                      correctVelocityGradient=bool(correctVelocityGradient))


class TreeGravity(_Package):
    """Gravity/TreeGravity(G, softeningLength, opening, ftimestep, timeStepChoice)."""

    def __init__(self, G, softeningLength, opening=0.5, ftimestep=0.1, timeStepChoice=None):
        super().__init__()
        self.timeStepChoice = timeStepChoice
        self.G = float(G)
        self.softeningLength = float(softeningLength)
        self.opening = float(opening)
        self.ftimestep = float(ftimestep)

    def label(self):
        return "TreeGravity"

    def time_step_code(self):
        """GravityTimeStepType (Gravity/TreeGravity.hh) as the engine's SPX_GRAVITY_DT_* code."""
        if self.timeStepChoice in (None, "AccelerationRatio"):
            return abi.SPX_GRAVITY_DT_ACCELERATION_RATIO
        if self.timeStepChoice == "DynamicalTime":
            return abi.SPX_GRAVITY_DT_DYNAMICAL_TIME
        raise NotImplementedError(f"TreeGravity timeStepChoice {self.timeStepChoice!r}")


class CheapSynchronousRK2Integrator:
    """Integrator/CheapSynchronousRK2 driving the packages on one engine run."""
    code = abi.SPX_INTEGRATOR_CHEAP_SYNCHRONOUS_RK2

    def __init__(self, dataBase):
        self.dataBase = dataBase
        self.packages = []
        self.lastDt = 1.0e-5
        self.dtMin = 0.0
        self.dtMax = 1.7976931348623157e308
        self.dtGrowth = 2.0
        self.currentTime = 0.0
        self.currentCycle = 0
        self.rigorousBoundaries = False
        self.verbose = False
        self._run = None
        self._stale = False
        self._slices = []

    def appendPhysicsPackage(self, package):
        self.packages.append(package)

    def physicsPackages(self):
        return list(self.packages)

    def havePhysicsPackage(self, package):
        return package in self.packages

    # -- engine binding ------------------------------------------------------------
    def _boundaries(self):
        """Integrator::uniqueBoundaryConditions: each package's boundaries in append order."""
        bcs = []
        for p in self.packages:
            for bc in p.boundaries:
                if bc not in bcs:
                    bcs.append(bc)
        return bcs

    def config(self):
        """spx_config of the problem described by the DataBase and the packages."""
        db = self.dataBase
        dim = db.nDim
        if db.numNodeLists == 0:
            raise ValueError("the DataBase holds no NodeList")
        hydro = [p for p in self.packages if isinstance(p, SPHPackage)]
        grav = [p for p in self.packages if isinstance(p, TreeGravity)]
        other = [p for p in self.packages if not isinstance(p, (SPHPackage, TreeGravity))]
        if not hydro and len(grav) == 1 and not other:
            return self._nbody_config(grav[0])
        if len(hydro) != 1 or other or len(grav) > 1:
            raise NotImplementedError("the GPU path runs one SPH package plus an optional TreeGravity, or "
                                      "TreeGravity alone under SynchronousRK2")
        h = hydro[0]
        n0 = db.nodeLists[0]
        for nl in db.nodeLists[1:]:
            same = (nl.eos.gamma == n0.eos.gamma and nl.hmin == n0.hmin and nl.hmax == n0.hmax and
                    nl.nPerh == n0.nPerh and nl.rhoMin == n0.rhoMin and nl.rhoMax == n0.rhoMax and
                    nl.kernelExtent == n0.kernelExtent)
            if not same:
                raise NotImplementedError("NodeLists with different materials or node parameters")
        W = h.W
        h.Q._check()
        cfg = abi.default_config(dim, kernel=W.base.code, kernel_order=getattr(W.base, "order", 5), nperh=n0.nPerh)
        cfg.kernel_points = W.numPoints
        cfg.min_nperh, cfg.max_nperh = W.minNperh, W.maxNperh
        cfg.neighbor_extent = n0.kernelExtent
        cfg.hmin, cfg.hmax = n0.hmin, n0.hmax
        cfg.rho_min, cfg.rho_max = n0.rhoMin, n0.rhoMax
        cfg.gamma = n0.eos.gamma
        cfg.Cl, cfg.Cq = h.Q.Cl, h.Q.Cq
        cfg.viscosity = h.Q.code
        if isinstance(h.Q, LimitedMonaghanGingoldViscosity):
            cfg.eta_crit_frac, cfg.eta_fold_frac = h.Q.etaCritFrac, h.Q.etaFoldFrac
        cfg.cfl = h.opts["cfl"]
        cfg.use_velocity_magnitude_for_dt = int(h.opts["useVelocityMagnitudeForDt"])
        cfg.compatible_energy = 1
        cfg.gradh_correction = int(h.opts["gradhCorrection"])
        cfg.correct_velocity_gradient = int(h.opts["correctVelocityGradient"])
        cfg.dt_min, cfg.dt_max, cfg.dt_growth, cfg.last_dt = self.dtMin, self.dtMax, self.dtGrowth, self.lastDt
        cfg.integrator = self.code
        bcs = self._boundaries()
        for bc in bcs:
            if isinstance(bc, PeriodicBoundary):
                (a1, s1), (a2, s2) = (pl.axis_side(dim) for pl in bc.planes)
                if a1 != a2 or s1 == s2:
                    raise NotImplementedError("periodic boundaries need two opposite planes on one axis")
                lo, hi = (bc.planes[0], bc.planes[1]) if s1 == 0 else (bc.planes[1], bc.planes[0])
                cfg.bc_lo[a1] = cfg.bc_hi[a1] = abi.SPX_BC_PERIODIC
                cfg.xmin[a1], cfg.xmax[a1] = lo.point[a1], hi.point[a1]
            elif isinstance(bc, ReflectingBoundary):
                a, side = bc.plane.axis_side(dim)
                if side == 0:
                    cfg.bc_lo[a] = abi.SPX_BC_REFLECTING
                    cfg.xmin[a] = bc.plane.point[a]
                else:
                    cfg.bc_hi[a] = abi.SPX_BC_REFLECTING
                    cfg.xmax[a] = bc.plane.point[a]
            else:
                raise NotImplementedError(f"boundary {type(bc).__name__}")
        for a in range(dim):
            if cfg.bc_lo[a] == abi.SPX_BC_REFLECTING and cfg.bc_hi[a] == abi.SPX_BC_NONE:
                cfg.xmax[a] = 1e100
            if cfg.bc_hi[a] == abi.SPX_BC_REFLECTING and cfg.bc_lo[a] == abi.SPX_BC_NONE:
                cfg.xmin[a] = -1e100
        if grav:
            g = grav[0]
            cfg.gravity = 1
            cfg.G, cfg.softening, cfg.opening, cfg.ftimestep = g.G, g.softeningLength, g.opening, g.ftimestep
            cfg.gravity_time_step = g.time_step_code()
        return cfg

    def _nbody_config(self, g):
        """TreeGravity as the only package (an N-body script): SynchronousRK2, no SPH."""
        if self.code != abi.SPX_INTEGRATOR_SYNCHRONOUS_RK2:
            raise NotImplementedError("TreeGravity alone is implemented under SynchronousRK2Integrator")
        db = self.dataBase
        n0 = db.nodeLists[0]
        cfg = abi.default_config(db.nDim, nperh=n0.nPerh)
        cfg.neighbor_extent = n0.kernelExtent
        cfg.hmin, cfg.hmax = n0.hmin, n0.hmax
        cfg.rho_min, cfg.rho_max = n0.rhoMin, n0.rhoMax
        cfg.gamma = n0.eos.gamma
        cfg.gravity, cfg.gravity_only, cfg.integrator = 1, 1, abi.SPX_INTEGRATOR_SYNCHRONOUS_RK2
        cfg.G, cfg.softening, cfg.opening, cfg.ftimestep = g.G, g.softeningLength, g.opening, g.ftimestep
        cfg.gravity_time_step = g.time_step_code()
        cfg.dt_min, cfg.dt_max, cfg.dt_growth, cfg.last_dt = self.dtMin, self.dtMax, self.dtGrowth, self.lastDt
        return cfg

    def _gather_state(self):
        nls = self.dataBase.nodeLists
        self._slices = []
        k = 0
        for nl in nls:
            self._slices.append((nl, k, k + nl.numInternalNodes))
            k += nl.numInternalNodes
        cat = lambda attr: np.concatenate([getattr(nl, attr) for nl in nls], 0)
        return dict(x=cat("_x"), v=cat("_v"), m=cat("_m"), rho=cat("_rho"), eps=cat("_eps"), H=cat("_H"))

    def initializeProblemStartup(self):
        st = self._gather_state()
        self._run = engine.make_run(self.config(), st)
        self._run.initialize()
        for nl in self.dataBase.nodeLists:
            nl._integrator = self
        self._stale = False

    def _download(self):
        if not self._stale or self._run is None:
            return
        r = self._run
        fields = {"_x": r.field("position"), "_v": r.field("velocity"), "_rho": r.field("mass density"),
                  "_eps": r.field("specific thermal energy"), "_H": r.field("H")}
        for nl, a, b in self._slices:
            for attr, arr in fields.items():
                getattr(nl, attr)[...] = arr[a:b]
        self._stale = False

    def step(self, maxTime, dt=None):
        """Integrator::step(maxTime): one CheapSynchronousRK2 step (Integrator.cc:67-111)."""
        if self._run is None:
            self.initializeProblemStartup()
        dtv = self._run.step(maxTime)
        self.currentTime, self.currentCycle, self.lastDt = self._run.time()
        self._stale = True
        return dtv

    def conservation(self):
        """Mass, momentum (3), kinetic, thermal and gravitational energy of the state."""
        return self._run.conservation()

    def connectivityMap(self):
        """DataBase::connectivityMap of the current state in the reference's numbering: the
        connectivity is rebuilt from the current positions (as the next Integrator::setGhostNodes
        does) and its ghosts numbered as the appended boundaries create and cull them
        (spheral_amd.connectivity).  connectivityForNode(nodeList, i) takes the NodeList."""
        if self._run is None:
            self.initializeProblemStartup()
        dim = self.dataBase.nDim
        seq = []
        for bc in self._boundaries():
            if isinstance(bc, PeriodicBoundary):
                a, side = bc.planes[0].axis_side(dim)
                seq.append(("periodic", a, 1 if side == 0 else 2))
            elif isinstance(bc, ReflectingBoundary):
                a, side = bc.plane.axis_side(dim)
                seq.append(("reflecting", a, 1 if side == 0 else 2))
        r = self._run
        r.update_connectivity()
        off, nbr, img = r.neighbors()
        nls = [(nl.name, a, b - a) for nl, a, b in self._slices]
        return connectivity.connectivity_map(r.cfg, r.field("position"), r.field("H"), off, nbr, img, nls, seq)


class SynchronousRK2Integrator(CheapSynchronousRK2Integrator):
    """Integrator/SynchronousRK2: derivatives at the start and at the midpoint of each step, the
    trial half step with the full policies (SynchronousRK2.cc:40-135)."""
    code = abi.SPX_INTEGRATOR_SYNCHRONOUS_RK2


AccelerationRatio = "AccelerationRatio"
DynamicalTime = "DynamicalTime"


class SpheralController:
    """SimulationControl/SpheralController: advance(goalTime, maxSteps) and step(steps)."""

    def __init__(self, integrator, kernel=None, statsStep=1, printStep=1, iterateInitialH=True, initialTime=0.0,
                 restartStep=None, restoreCycle=None, vizBaseName=None, **kwargs):
        if restartStep is not None or restoreCycle is not None:
            raise NotImplementedError("restart files are not implemented on the GPU path")
        if vizBaseName is not None:
            raise NotImplementedError("visualisation dumps are not implemented on the GPU path")
        if initialTime != 0.0:
            raise NotImplementedError("a non-zero initial time is not implemented")
        self.integrator = integrator
        self.kernel = kernel
        self.totalSteps = 0
        integrator.initializeProblemStartup()
        if iterateInitialH:
            self.iterateIdealH()

    def iterateIdealH(self, maxIdealHIterations=50, idealHTolerance=1.0e-4):
        """SpheralController.iterateIdealH (SpheralController.py:909-931) -> spx_run_iterate_ideal_h."""
        k = self.integrator._run.iterate_ideal_h(maxIdealHIterations, idealHTolerance)
        self.integrator._stale = True
        self.integrator._download()
        return k

    def time(self):
        return self.integrator.currentTime

    def lastDt(self):
        return self.integrator.lastDt

    def advance(self, goalTime, maxSteps=None):
        """Step until goalTime or maxSteps (SpheralController.py:383-426)."""
        it = self.integrator
        k = 0
        while it.currentTime < goalTime*(1.0 - 1.0e-15) and (maxSteps is None or k < maxSteps):
            it.step(goalTime)
            k += 1
            self.totalSteps += 1
        it._download()

    def step(self, steps=1):
        for _ in range(steps):
            self.integrator.step(1.0e300)
            self.totalSteps += 1
        self.integrator._download()


def namespace(dim):
    """The names a ``from Spheral{dim}d import *`` script uses, bound to dimension dim."""
    sfx = f"{dim}d"

    def mk_nl(*args, **kwargs):
        kwargs.setdefault("dim", dim)
        return makeFluidNodeList(*args, **kwargs)

    def mk_db():
        return DataBase(dim)

    ns = {
        "Vector": Vector, "Plane": Plane, "PeriodicBoundary": PeriodicBoundary,
        "ReflectingBoundary": ReflectingBoundary, "BSplineKernel": BSplineKernel, "WendlandC4Kernel": WendlandC4Kernel,
        "NBSplineKernel": NBSplineKernel, "LimitedMonaghanGingoldViscosity": LimitedMonaghanGingoldViscosity,
        "TableKernel": TableKernel, "GammaLawGas": GammaLawGas, "GammaLawGasMKS": GammaLawGasMKS,
        "GammaLawGasCGS": GammaLawGasCGS, "makeFluidNodeList": mk_nl, "DataBase": mk_db,
        "MonaghanGingoldViscosity": MonaghanGingoldViscosity, "SPH": SPH, "TreeGravity": TreeGravity,
        "CheapSynchronousRK2Integrator": CheapSynchronousRK2Integrator, "SpheralController": SpheralController,
        "SynchronousRK2Integrator": SynchronousRK2Integrator, "AccelerationRatio": AccelerationRatio,
        "DynamicalTime": DynamicalTime,
        "RigorousSumDensity": RigorousSumDensity, "IntegrateDensity": IntegrateDensity, "SumDensity": SumDensity,
        "IdealH": IdealH, "IntegrateH": IntegrateH,
        "GenerateNodeDistribution1d": GenerateNodeDistribution1d,
        "GenerateNodeDistribution3d": GenerateNodeDistribution3d,
        "distributeNodes1d": distributeNodes1d, "distributeNodes3d": distributeNodes3d,
    }
    for base in ("Vector", "Plane", "PeriodicBoundary", "ReflectingBoundary", "BSplineKernel", "WendlandC4Kernel",
                 "NBSplineKernel", "LimitedMonaghanGingoldViscosity",
                 "TableKernel", "makeFluidNodeList", "DataBase", "MonaghanGingoldViscosity", "TreeGravity",
                 "CheapSynchronousRK2Integrator", "SynchronousRK2Integrator", "SPH"):
        ns[base + sfx] = ns[base]
    if dim == 3:
        ns["OctTreeGravity"] = TreeGravity
    elif dim == 2:
        ns["QuadTreeGravity"] = TreeGravity
    return ns
