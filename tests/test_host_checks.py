"""Host-side argument checks of the Python engine wrapper (CPU only).

The library reads and writes caller buffers of the sizes its own state implies (N x tracers x
levels doubles for a state, N x n for a weight set), so the wrapper refuses arrays of any other
size before it calls the library.  These wrappers are built without a device handle: every
case raises before the first library call."""
import ctypes as C

import numpy as np
import pytest

from se_rbf_projects_amd import engine as E
from se_rbf_projects_amd.ctypes_api import config_deform


def bare_engine(N=12, n=5, cfg=None):
    e = E.Engine.__new__(E.Engine)
    e._e = C.c_void_p()
    e.N, e.n, e.cfg = N, n, cfg
    return e


def bare_closure():
    c = E.Closure.__new__(E.Closure)
    c._c = C.c_void_p()
    return c


def test_state_needs_a_configured_engine_and_its_size():
    e = bare_engine()
    with pytest.raises(ValueError):
        e.set_state(np.zeros((12, 10)))          # not configured
    cfg = config_deform(12, Nv=10, n=5, deg=1, nv=3)
    cfg.ntracers = 4
    e.cfg = cfg
    with pytest.raises(ValueError):
        e.set_state(np.zeros((12, 10)))          # one tracer given, four configured
    with pytest.raises(ValueError):
        e.norms(q_ref=np.zeros((12, 4, 9)))      # one level short
    with pytest.raises(ValueError):
        e.set_gamma(np.zeros(12 * 4 * 10 - 1))


def test_node_and_stencil_shapes():
    e = bare_engine()
    with pytest.raises(ValueError):
        e.set_nodes(np.zeros((12, 2)))
    with pytest.raises(ValueError):
        e.set_nodes(np.zeros((0, 3)))
    with pytest.raises(ValueError):
        e.set_nodes(np.zeros((12, 3)), lonlat=(np.zeros(12), np.zeros(11)))
    with pytest.raises(ValueError):
        e.set_stencils(np.zeros((11, 5), np.int32))
    with pytest.raises(ValueError):
        e.set_stencil_sizes(np.full(11, 5), np.full(12, 2))
    assert (e.N, e.n) == (12, 5)                 # a refused call changes nothing


def test_weight_sizes():
    e = bare_engine()
    w = np.zeros((12, 5))
    with pytest.raises(ValueError):
        e.set_weights(w, np.zeros((12, 4)))
    with pytest.raises(ValueError):
        e.set_weights(w, w, np.zeros((12, 6)))


def test_configure_takes_a_config_only():
    e = bare_engine()
    with pytest.raises(TypeError):
        e.configure({"Nv": 10})


def test_closure_parameter_and_data_sizes():
    c = bare_closure()
    with pytest.raises(ValueError):
        c.set_params(np.zeros(E.Closure.NPARAMS - 1), np.zeros(32))
    with pytest.raises(ValueError):
        c.set_params(np.zeros(E.Closure.NPARAMS), np.zeros(16))
    with pytest.raises(ValueError):
        c.set_data(np.zeros((8, 10)), np.zeros(8), np.zeros(8))
    with pytest.raises(ValueError):
        c.set_data(np.zeros((8, 16)), np.zeros(8), np.zeros(7))


def test_run_history_arguments():
    e = bare_engine()
    with pytest.raises(ValueError):
        e.run_history(0, 10, 1.0, every=0)
    with pytest.raises(ValueError):
        e.run_history(0, -1, 1.0, every=2)
