"""ctypes mirror of ``include/spx_abi.h``: the configuration record and constants.

Field meaning and defaults follow the reference (paths relative to its src/):
TableKernel ctor defaults (Kernel/TableKernel.hh:35-37), makeFluidNodeList
defaults (NodeList/FluidNodeLists.py:15-31), the SPH() factory defaults
(SPH/SPHHydros.py:9-33), Integrator member defaults (Integrator/Integrator.cc:41-47)
and the TreeGravity ctor (Gravity/TreeGravity.hh:46-50).
"""
import ctypes as C

SPX_OK = 0
SPX_ERR_ARG = -1
SPX_ERR_HIP = -2
SPX_ERR_STATE = -3
SPX_ERR_UNSUPPORTED = -4
SPX_ERR_CAPACITY = -5
SPX_ERR_NUMERIC = -6

SPX_KERNEL_BSPLINE = 0
SPX_KERNEL_WENDLANDC4 = 1
SPX_KERNEL_NBSPLINE = 2

SPX_Q_MONAGHAN_GINGOLD = 0
SPX_Q_LIMITED_MONAGHAN_GINGOLD = 1

SPX_INTEGRATOR_CHEAP_SYNCHRONOUS_RK2 = 0
SPX_INTEGRATOR_SYNCHRONOUS_RK2 = 1

SPX_GRAVITY_DT_ACCELERATION_RATIO = 0
SPX_GRAVITY_DT_DYNAMICAL_TIME = 1

SPX_BC_NONE = 0
SPX_BC_PERIODIC = 1
SPX_BC_REFLECTING = 2

SPX_PREC_FP64 = 0
SPX_PREC_MIXED = 1

KERNEL_EXTENT = {SPX_KERNEL_BSPLINE: 2.0, SPX_KERNEL_WENDLANDC4: 1.0}


def kernel_extent(kind, order=5):
    """The base kernel's extent; NBSplineKernel sets (order + 1)/2 in integer arithmetic
    (Kernel/NBSplineKernel.cc:99)."""
    if kind == SPX_KERNEL_NBSPLINE:
        return float((int(order) + 1) // 2)
    return KERNEL_EXTENT[kind]


class SpxConfig(C.Structure):
    _fields_ = [
        ("dim", C.c_int32),
        ("kernel", C.c_int32),
        ("kernel_points", C.c_int32),
        ("precision", C.c_int32),
        ("min_nperh", C.c_double),
        ("max_nperh", C.c_double),
        ("nperh", C.c_double),
        ("neighbor_extent", C.c_double),
        ("hmin", C.c_double),
        ("hmax", C.c_double),
        ("rho_min", C.c_double),
        ("rho_max", C.c_double),
        ("gamma", C.c_double),
        ("Cl", C.c_double),
        ("Cq", C.c_double),
        ("cfl", C.c_double),
        ("use_velocity_magnitude_for_dt", C.c_int32),
        ("compatible_energy", C.c_int32),
        ("gradh_correction", C.c_int32),
        ("correct_velocity_gradient", C.c_int32),
        ("dt_min", C.c_double),
        ("dt_max", C.c_double),
        ("dt_growth", C.c_double),
        ("last_dt", C.c_double),
        ("bc_lo", C.c_int32 * 3),
        ("bc_hi", C.c_int32 * 3),
        ("xmin", C.c_double * 3),
        ("xmax", C.c_double * 3),
        ("gravity", C.c_int32),
        ("reorder_interval", C.c_int32),
        ("G", C.c_double),
        ("softening", C.c_double),
        ("opening", C.c_double),
        ("ftimestep", C.c_double),
        ("kernel_order", C.c_int32),
        ("viscosity", C.c_int32),
        ("eta_crit_frac", C.c_double),
        ("eta_fold_frac", C.c_double),
        ("integrator", C.c_int32),
        ("gravity_only", C.c_int32),
        ("gravity_time_step", C.c_int32),
    ]

    def copy(self):
        c = SpxConfig()
        C.pointer(c)[0] = self
        return c

    def as_dict(self):
        out = {}
        for name, _ in self._fields_:
            val = getattr(self, name)
            if hasattr(val, "__len__"):
                val = list(val)
            out[name] = val
        return out


def default_config(dim=3, kernel=SPX_KERNEL_BSPLINE, **overrides):
    """Reference defaults; ``neighbor_extent`` follows the kernel's extent as the
    reference's problem scripts set it (``kernelExtent = WT.kernelExtent``)."""
    c = SpxConfig()
    c.dim = dim
    c.kernel = kernel
    c.kernel_points = 1000
    c.precision = SPX_PREC_FP64
    c.min_nperh = 0.25
    c.max_nperh = 64.0
    c.nperh = 2.01
    c.neighbor_extent = kernel_extent(kernel)
    c.hmin = 1.0e-20
    c.hmax = 1.0e20
    c.rho_min = 1.0e-10
    c.rho_max = 1.0e10
    c.gamma = 5.0 / 3.0
    c.Cl = 1.0
    c.Cq = 1.0
    c.cfl = 0.25
    c.use_velocity_magnitude_for_dt = 0
    c.compatible_energy = 1
    c.gradh_correction = 1
    c.correct_velocity_gradient = 1
    c.dt_min = 0.0
    c.dt_max = 1.7976931348623157e308
    c.dt_growth = 2.0
    c.last_dt = 1.0e-5
    for a in range(3):
        c.bc_lo[a] = SPX_BC_NONE
        c.bc_hi[a] = SPX_BC_NONE
        c.xmin[a] = 0.0
        c.xmax[a] = 1.0
    c.gravity = 0
    c.reorder_interval = 1
    c.G = 1.0
    c.softening = 0.01
    c.opening = 0.5
    c.ftimestep = 0.1
    c.kernel_order = 5
    c.viscosity = SPX_Q_MONAGHAN_GINGOLD
    c.eta_crit_frac = 1.0
    c.eta_fold_frac = 0.2
    c.integrator = SPX_INTEGRATOR_CHEAP_SYNCHRONOUS_RK2
    c.gravity_only = 0
    c.gravity_time_step = SPX_GRAVITY_DT_ACCELERATION_RATIO
    for k, v in overrides.items():
        if k in ("bc_lo", "bc_hi", "xmin", "xmax"):
            arr = getattr(c, k)
            for a, x in enumerate(v):
                arr[a] = x
        elif k == "kernel":
            pass
        else:
            setattr(c, k, v)
    if "neighbor_extent" not in overrides:
        c.neighbor_extent = kernel_extent(c.kernel, c.kernel_order)
    return c


def set_periodic(cfg, xmin, xmax):
    for a in range(cfg.dim):
        cfg.bc_lo[a] = SPX_BC_PERIODIC
        cfg.bc_hi[a] = SPX_BC_PERIODIC
        cfg.xmin[a] = xmin[a]
        cfg.xmax[a] = xmax[a]
    return cfg
