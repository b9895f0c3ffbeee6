"""C ABI: the library loads on the host, exports every entry point of include/spx_abi.h,
and its host-side pieces (defaults, kernel tables) agree with the oracle."""
import ctypes as C
import os
import re

import numpy as np

from spheral_amd import abi, engine

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "spx_abi.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(spx_[a-z_0-9]+)\s*\(", text)))


def test_header_and_binding_agree():
    assert sorted(engine.EXPORTS) == declared_symbols()


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(engine.LIB_PATH)
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing


def test_config_layout_and_defaults():
    for dim in (1, 2, 3):
        c = abi.SpxConfig()
        assert engine.lib().spx_config_default(dim, C.byref(c)) == 0
        py = abi.default_config(dim)
        py.neighbor_extent = 2.0            # the C default names the BSpline extent
        for name, _ in abi.SpxConfig._fields_:
            a, b = getattr(c, name), getattr(py, name)
            if hasattr(a, "__len__"):
                assert list(a) == list(b), name
            else:
                assert a == b, name


def test_bad_arguments_report_errors():
    c = abi.SpxConfig()
    assert engine.lib().spx_config_default(7, C.byref(c)) == abi.SPX_ERR_ARG
    assert engine.lib().spx_last_error()


def test_engine_table_is_the_oracle_table(oracle_mod):
    """The engine builds its tables on the host in FP64; they must be the oracle's doubles."""
    for dim in (1, 2, 3):
        for kern, order in ((abi.SPX_KERNEL_BSPLINE, 5), (abi.SPX_KERNEL_WENDLANDC4, 5), (abi.SPX_KERNEL_NBSPLINE, 3),
                            (abi.SPX_KERNEL_NBSPLINE, 5), (abi.SPX_KERNEL_NBSPLINE, 7)):
            cfg = abi.default_config(dim, kernel=kern, kernel_order=order)
            cfg.kernel = kern
            wo, go, xo = oracle_mod.table_coefficients(cfg)
            n = C.c_int64()
            assert engine.lib().spx_table_coefficients(C.byref(cfg), C.byref(n), None, None, None) == 0
            we, ge = np.zeros(n.value), np.zeros(n.value)
            xe = C.c_double()
            dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
            assert engine.lib().spx_table_coefficients(C.byref(cfg), C.byref(n), dp(we), dp(ge), C.byref(xe)) == 0
            assert np.array_equal(we, wo) and np.array_equal(ge, go) and xe.value == xo
            eta = np.linspace(0.0, 1.05*abi.kernel_extent(kern, order), 997)
            W1, G1, N1 = oracle_mod.table_eval(cfg, eta)
            W2, G2, N2 = np.zeros_like(eta), np.zeros_like(eta), np.zeros_like(eta)
            assert engine.lib().spx_table_eval(C.byref(cfg), dp(eta), eta.size, dp(W2), dp(G2), dp(N2)) == 0
            assert np.array_equal(W1, W2) and np.array_equal(G1, G2) and np.array_equal(N1, N2)
