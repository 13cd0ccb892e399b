"""Concurrent evaluation through the C ABI: distinct plans evaluated from
four host threads at once (each call on its own stream, first evaluations --
launch-cache builds and HIP-graph captures -- included) give bit-identical
results to the same plans evaluated one after another.  The reference's
PricingPlan::evaluate takes &self and is Send + Sync
(crates/pricing/src/engine/risk/pricing.rs:86-88)."""
import threading

import pytest

from pricing_amd import abi, engine, workloads

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _init():
    engine.init(0)


def _makers():
    model = workloads.lifted_heston_model(engine.lifted_from_rough(0.1, 20, 2.5))
    expiry = workloads.weekday_dates(252)[-1] / 365.0
    grid = workloads.lv_dupire_grid(expiry, build=engine.local_variance_from_ssvi)
    return [
        ("european", lambda: [engine.PricingPlan.compile(workloads.european_bs(units=1 << 15))],
         lambda ps, b: [ps[0].evaluate(b)]),
        ("graph asian", lambda: [engine.PricingPlan.compile(workloads.graph_asian(units=1 << 14))],
         lambda ps, b: [ps[0].evaluate(b)]),
        ("local vol shared", lambda: [engine.PricingPlan.compile(r)
                                      for r in workloads.lv_path_dependent(grid, points=512, scrambles=2)],
         lambda ps, b: engine.evaluate_shared(ps, b)),
        ("rough smile", lambda: [engine.RoughPlan.compile(r, model, 2.0 / 256)
                                 for r in workloads.rough_smile(points=1024, scrambles=2, strikes=3,
                                                                maturities=(0.5, 1.0))],
         lambda ps, b: engine.evaluate_shared(ps, b)),
        ("batch", lambda: [engine.PricingPlan.compile(r)
                           for r in workloads.batch_trades(24, seed=3, min_log2=9.0, max_log2=11.0)],
         lambda ps, b: engine.evaluate_batch(ps, b)),
        ("american", lambda: [engine.PricingPlan.compile(r)
                              for r in workloads.american_batch(4, train_log2=11, value_log2=11, dates=8)],
         lambda ps, b: engine.evaluate_lsm(ps, b)[0]),
    ]


def _key(results):
    return [tuple((r.estimates[c].value, r.estimates[c].standard_error) for c in range(abi.COMPONENTS))
            for r in results]


@pytest.mark.parametrize("backend", ["fp64", "mixed"])
def test_four_threads_match_serial(backend):
    makers = _makers()
    serial = {}
    for name, make, run in makers:
        serial[name] = _key(run(make(), backend))
    # fresh plans: the threads build their launch caches and graphs concurrently
    for rnd in range(2):
        plans = {name: make() for name, make, _ in makers}
        got, errors = {}, []

        def work(names):
            try:
                for _ in range(2):
                    for name, _, run in makers:
                        if name in names:
                            got[name] = _key(run(plans[name], backend))
            except Exception as e:  # surfaced below
                errors.append(repr(e))

        groups = [[m[0] for m in makers[i::4]] for i in range(4)]
        threads = [threading.Thread(target=work, args=(g,)) for g in groups]
        for t in threads:
            t.start()
        for t in threads:
            t.join()
        assert not errors, errors
        for name in serial:
            assert got[name] == serial[name], (rnd, name)
