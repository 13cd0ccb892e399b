"""Plan compilation of a Bermudan under lifted Heston (CPU only: no device
work).  The device prices it with the standardized-monomial CPQR policy of
total degree 1 or 2 on (spot, variance), or the network policy judged by it;
the weighted-Laguerre policy, degree 3 and a Brownian-bridge ordering are
refused at compile time with the status the ABI documents."""
import ctypes as C

import pytest

from pricing_amd import abi, engine, workloads

STEP = 2.0 / 256


def _model():
    return workloads.lifted_heston_model(engine.lifted_from_rough(0.1, 8, 2.5))


@pytest.mark.parametrize("kw,status", [
    (dict(method=abi.PE_LSM_POLY_LAGUERRE), abi.PE_ERR_UNSUPPORTED),
    (dict(max_degree=3), abi.PE_ERR_UNSUPPORTED),
], ids=["laguerre", "degree3"])
def test_unsupported_policies_are_refused(kw, status):
    req = workloads.rough_bermudan(units=1 << 10, train_units=1 << 10, **kw)
    with pytest.raises(engine.EngineError) as err:
        engine.RoughPlan.compile(req, _model(), STEP)
    assert err.value.status == status


def test_brownian_bridge_is_refused():
    req = workloads.rough_bermudan(units=1 << 10, train_units=1 << 10)
    req.c.lsm.training_engine.brownian_bridge = 1
    with pytest.raises(engine.EngineError) as err:
        engine.RoughPlan.compile(req, _model(), STEP)
    assert err.value.status == abi.PE_ERR_INVALID_ARGUMENT


def test_degree_one_and_two_and_the_network_compile_with_grid_diagnostics():
    for degree, method in ((1, abi.PE_LSM_POLY_CPQR), (2, abi.PE_LSM_POLY_CPQR), (2, abi.PE_LSM_MLP)):
        req = workloads.rough_bermudan(units=1 << 10, train_units=1 << 10, scrambles=4, max_degree=degree,
                                       method=method)
        plan = engine.RoughPlan.compile(req, _model(), STEP)
        d = abi.pe_mc_diagnostics()
        assert engine.load().pe_plan_diagnostics(plan._handle, C.byref(d)) == abi.PE_OK
        assert d.has_checksums and d.scramble_count == 4
