"""Named workloads (BASELINE configurations) built as engine requests.

Each builder returns Request objects; sizes are parameters so tests can run
the same workload shape at reduced path counts.
"""
import datetime as _dt
import math

from . import abi, request

VALUATION = _dt.date(2026, 9, 4)


def _flat_curve(rate, horizon=1.0):
    return [0.0, horizon], [1.0, math.exp(-rate * horizon)]


def european_bs(units=1 << 16, seed=0x0123456789ABCDEF):
    """European call under Black-Scholes: S0=K=100, r=5%, q=0, sigma=20%, T=1,
    one exact lognormal step, pseudo-MC, pathwise Delta/Vega (FP64 method)."""
    req = request.Request()
    dt, df = _flat_curve(0.05)
    req.set_market(100.0, dt, df, [0.0, 1.0], [1.0, 1.0])
    req.set_black_scholes(0.2)
    req.set_product(abi.PE_PRODUCT_EUROPEAN, abi.PE_CALL, 365, 100.0)
    req.set_pseudo(seed, units)
    req.set_risk(delta=True, vega=True)
    return req


def weekday_dates(count, start=VALUATION):
    out = []
    d = start
    while len(out) < count:
        d += _dt.timedelta(days=1)
        if d.weekday() < 5:
            out.append((d - start).days)
    return out


ESSVI = {"rho": -0.6, "eta": 1.0, "gamma": 0.4, "theta_slope": 0.04}


def lv_dupire_grid(expiry, build=None, time_count=32, left=-1.5, right=1.0, nl=32, nr=31, shape=3.0,
                   floor=1e-8, cap=4.0):
    """The local-vol config's Dupire grid: StandardSsvi with theta(t) = 0.04 t
    (a two-knot theta PCHIP with slope 0.04), rho=-0.6, power-law phi
    (eta=1.0, gamma=0.4); time nodes T (i/31)^2, i = 0..31; 64 log-moneyness
    nodes piecewise-sinh(-1.5, 1.0, 32, 31, 3, 3).  `build` computes the
    values (the GPU builder or the oracle)."""
    times = [expiry * (i / (time_count - 1)) ** 2 for i in range(time_count)]
    times[-1] = expiry
    xs = []
    for i in range(nl + 1):
        frac = i / nl
        ratio = math.sinh(shape * (1.0 - frac)) / math.sinh(shape)
        xs.append(0.0 if i == nl else left * ratio)
    for i in range(1, nr + 1):
        frac = i / nr
        xs.append(right * math.sinh(shape * frac) / math.sinh(shape))
    theta_t = [expiry / 2.0, expiry]
    theta_v = [ESSVI["theta_slope"] * t for t in theta_t]
    values, repairs = build(theta_t, theta_v, ESSVI["theta_slope"], ESSVI["rho"], ESSVI["eta"], ESSVI["gamma"],
                            times, xs, floor, cap)
    return times, xs, values, floor, cap, repairs


def lv_path_dependent(values_grid, points=1 << 16, scrambles=16, seed=0xFEDCBA9876543210, steps=252,
                      barrier=130.0, strike=100.0, rate=0.05):
    """Local-vol Log-Euler path-dependent set: RQMC Sobol + Brownian bridge,
    daily monitoring over the first `steps` weekdays; an up-and-out call and
    an arithmetic Asian call on the same paths.  Returns the two requests."""
    times, xs, values, floor, cap, _ = values_grid
    dates = weekday_dates(steps)
    expiry = dates[-1]
    reqs = []
    for kind in ("barrier", "asian"):
        req = request.Request()
        horizon = expiry / 365.0
        req.set_market(100.0, [0.0, horizon], [1.0, math.exp(-rate * horizon)], [0.0, horizon], [1.0, 1.0])
        req.set_local_vol(times, xs, values, floor, cap)
        if kind == "barrier":
            req.set_product(abi.PE_PRODUCT_BARRIER, abi.PE_CALL, expiry, strike, barrier=barrier, dates=dates)
        else:
            req.set_product(abi.PE_PRODUCT_ASIAN, abi.PE_CALL, expiry, strike, dates=dates,
                            weights=[1.0 / steps] * steps)
        req.set_rqmc(seed, points, scrambles, antithetic=False, brownian_bridge=True)
        req.set_risk()
        reqs.append(req)
    return reqs
