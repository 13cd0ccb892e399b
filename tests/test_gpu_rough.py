"""GPU parity of the lifted Heston adapter: FP64 device mode against the CPU
oracle (per-plan reference grid and shared smile grid) within 1e-12, the
mixed mode within 0.1 standard errors of the FP64 mode on the whole smile
grid, a non-finite path failing both modes, and the device quadrature
weights against the oracle's."""
import pytest

from _support import oracle_rough
from pricing_amd import abi, engine, workloads

pytestmark = pytest.mark.gpu

REL = 1e-12
STEP = 2.0 / 256


@pytest.fixture(scope="module", autouse=True)
def _init():
    engine.init(0)


def _model():
    return workloads.lifted_heston_model(engine.lifted_from_rough(0.1, 20, 2.5))


def test_device_from_rough_matches_oracle():
    import ctypes as C
    from _support import oracle

    w, x = engine.lifted_from_rough(0.1, 20, 2.5)
    ow, ox = (C.c_double * 20)(), (C.c_double * 20)()
    assert oracle().oracle_lifted_from_rough(0.1, 20, 2.5, ow, ox) == 0
    assert w == list(ow) and x == list(ox)


@pytest.mark.parametrize("scrambles,antithetic", [(0, True), (8, False), (4, True)])
def test_single_plan_fp64_matches_oracle(scrambles, antithetic):
    model = _model()
    reqs = workloads.rough_smile(points=1000 if scrambles else 3000, scrambles=scrambles, strikes=3,
                                 maturities=(0.25, 1.0), antithetic=antithetic)
    for r in reqs:
        gpu = engine.RoughPlan.compile(r, model, STEP).evaluate("fp64")
        ref, _ = oracle_rough([r], model, STEP, workers=8)
        g, o = gpu.estimates[0], ref[0].estimates[0]
        assert abs(g.value - o.value) <= REL * abs(o.value) + 1e-15, (g.value, o.value)
        assert abs(g.standard_error - o.standard_error) <= 1e-10 * o.standard_error, (g.standard_error, o.standard_error)
        assert gpu.time_steps == ref[0].time_steps


def test_shared_smile_fp64_matches_oracle():
    model = _model()
    reqs = workloads.rough_smile(points=512, scrambles=4, strikes=5, maturities=(0.1, 0.5, 2.0))
    plans = [engine.RoughPlan.compile(r, model, STEP) for r in reqs]
    gpu = engine.evaluate_shared(plans, "fp64")
    ref, _ = oracle_rough(reqs, model, STEP, workers=8)
    for g, o in zip(gpu, ref):
        assert abs(g.estimates[0].value - o.estimates[0].value) <= REL * abs(o.estimates[0].value) + 1e-15
        assert abs(g.estimates[0].standard_error - o.estimates[0].standard_error) <= 1e-10 * o.estimates[0].standard_error


def test_smile_mixed_within_tenth_se_on_the_full_grid():
    """The whole 21-strike x 8-maturity smile on 2^14 points x 16 scrambles:
    every mixed price within 0.1 SE of the FP64 mode.  The mixed mode keeps
    the variance process in FP64 (the truncated scheme amplifies any FP32
    rounding in it), so the remaining gap is the FP32 asset increment."""
    model = _model()
    reqs = workloads.rough_smile(points=1 << 14, scrambles=16)
    plans = [engine.RoughPlan.compile(r, model, STEP) for r in reqs]
    f64 = engine.evaluate_shared(plans, "fp64")
    mixed = engine.evaluate_shared(plans, "mixed")
    gaps = [abs(a.estimates[0].value - b.estimates[0].value) / a.estimates[0].standard_error for a, b in zip(f64, mixed)]
    print("max mixed - fp64 in SE: %.4f" % max(gaps), flush=True)
    assert len(gaps) == 168
    assert max(gaps) < 0.1, gaps


def test_mixed_flags_a_non_finite_path():
    """A variance that overflows to infinity fails the evaluation in mixed
    mode too (the FP64 mode and the reference fail it)."""
    model = abi.lifted_heston(1e300, 0.0, 0.0, 1e300, -0.5, [1e300], [0.0])
    reqs = workloads.rough_smile(points=256, scrambles=0, strikes=2, maturities=(0.25,))
    for backend in ("fp64", "mixed"):
        plans = [engine.RoughPlan.compile(r, model, STEP) for r in reqs]
        with pytest.raises(engine.EngineError) as err:
            engine.evaluate_shared(plans, backend)
        assert err.value.status == abi.PE_ERR_NONFINITE_STATE
