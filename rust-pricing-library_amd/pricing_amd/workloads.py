"""Named workloads (the benchmark configurations: local-vol path-dependent set, trade batch, American batch, lifted-Heston smile, graph Asian) built as engine requests.

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


def graph_asian(units=1 << 20, fixings=52, as_graph=True, seed=0x6EA9_A51A):
    """Weekly arithmetic Asian call (K=100) over one year under Black-Scholes
    (S0=100, r=5%, sigma=20%), pseudo-MC, price only: as a product graph
    (the reference's Asian graph, compile.rs:118-160: a running weighted sum,
    strike, positive part, notional) or as the fixed Asian product."""
    req = request.Request()
    dt, df = _flat_curve(0.05)
    req.set_market(100.0, dt, df, [0.0, 1.0], [1.0, 1.0])
    req.set_black_scholes(0.2)
    days = [7 * (i + 1) for i in range(fixings)]
    w = 1.0 / fixings
    if as_graph:
        nodes = [(0, "literal", [], {"value": 0.0})]
        acc = 0
        for d in days:
            n = len(nodes)
            nodes.append((n, "terminal_spot", [], {"underlying_id": 1, "observation_day": d}))
            nodes.append((n + 1, "literal", [], {"value": w}))
            nodes.append((n + 2, "multiply", [n, n + 1], {}))
            nodes.append((n + 3, "add", [acc, n + 2], {}))
            acc = n + 3
        n = len(nodes)
        nodes += [(n, "literal", [], {"value": 100.0}), (n + 1, "subtract", [acc, n], {}),
                  (n + 2, "literal", [], {"value": 0.0}), (n + 3, "maximum", [n + 1, n + 2], {}),
                  (n + 4, "literal", [], {"value": 1.0}), (n + 5, "multiply", [n + 3, n + 4], {})]
        req.set_graph_product(nodes, [(n + 5, days[-1])], [1])
    else:
        req.set_product(abi.PE_PRODUCT_ASIAN, abi.PE_CALL, days[-1], 100.0, dates=days, weights=[w] * fixings)
    req.set_pseudo(seed, units)
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


# ------------------------------------------------------------ trade batch --
class SplitMix64:
    """SplitMix64 (Steele, Lea, Flood 2014): the batch generator's stream."""

    def __init__(self, seed):
        self.state = seed & 0xFFFFFFFFFFFFFFFF

    def next(self):
        self.state = (self.state + 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
        return z ^ (z >> 31)

    def uniform(self, lo=0.0, hi=1.0):
        return lo + (hi - lo) * ((self.next() >> 11) * (1.0 / (1 << 53)))

    def integer(self, lo, hi):
        """Uniform integer in [lo, hi]."""
        return lo + self.next() % (hi - lo + 1)


BATCH_RATE = 0.03
BATCH_HALF_WIDTH = 1.0


def batch_trades(count=4096, seed=0x00C0FFEE, min_log2=12.0, max_log2=18.0, risk=True):
    """Batched uneven BS/Black-76 trades with affine discrete dividends: trade
    id mod 5 is European / cash Digital / Barrier (discrete for odd ids,
    continuous for even ids) / arithmetic Asian / fixed Lookback; call and put
    alternate; K ~ U[70, 130], sigma ~ U[0.1, 0.5], T ~ U[0.1, 5] years
    (expiry = round(365 T) days); 12-365 equally spaced dates (Europeans and
    Digitals: one); pseudo-MC paths round(2^U[min_log2, max_log2]) with a seed
    per trade; 0-4 dividends of 1.0 cash + 1% proportional at U(0, T) ex-days;
    up-and-out call barrier max(S0, K) U[1.1, 1.4] or down-and-out put barrier
    min(S0, K) U[0.6, 0.9]; compact-C2 smoothing of half width 1.0; r = 3%,
    q = 0; BS for even ids, Black-76 for odd ids; price with Delta and Vega."""
    rng = SplitMix64(seed)
    s0 = 100.0
    out = []
    for tid in range(count):
        kind = tid % 5
        call = tid % 2 == 0
        strike = rng.uniform(70.0, 130.0)
        vol = rng.uniform(0.1, 0.5)
        t_years = rng.uniform(0.1, 5.0)
        expiry = max(1, int(round(365.0 * t_years)))
        steps = rng.integer(12, 365)
        paths = int(round(2.0 ** rng.uniform(min_log2, max_log2)))
        n_div = rng.integer(0, 4)
        ex_days = sorted({rng.integer(1, max(1, expiry - 1)) for _ in range(n_div)}) if expiry > 1 else []
        barrier_u = rng.uniform(0.0, 1.0)
        seed_t = rng.next()
        horizon = expiry / 365.0
        req = request.Request()
        req.set_market(s0, [0.0, horizon], [1.0, math.exp(-BATCH_RATE * horizon)], [0.0, horizon], [1.0, 1.0],
                       [(d / 365.0, 1.0, 0.01) for d in ex_days])
        req.set_black_scholes(vol, black76=tid % 2 == 1)
        side = abi.PE_CALL if call else abi.PE_PUT
        n = min(steps, expiry)
        dates = sorted({max(1, int(round(expiry * (k + 1) / n))) for k in range(n)})
        if kind == 0:
            req.set_product(abi.PE_PRODUCT_EUROPEAN, side, expiry, strike)
        elif kind == 1:
            req.set_product(abi.PE_PRODUCT_DIGITAL, side, expiry, strike, payout=1.0)
        elif kind == 2:
            if call:
                barrier = max(s0, strike) * (1.1 + 0.3 * barrier_u)
                down = 0
            else:
                barrier = min(s0, strike) * (0.6 + 0.3 * barrier_u)
                down = 1
            req.set_product(abi.PE_PRODUCT_BARRIER, side, expiry, strike, barrier=barrier, direction_down=down,
                            continuous=int(tid % 2 == 0), dates=dates)
        elif kind == 3:
            w = [1.0 / len(dates)] * len(dates)
            req.set_product(abi.PE_PRODUCT_ASIAN, side, expiry, strike, dates=dates, weights=w,
                            fixings=[None] * len(dates))
        else:
            req.set_product(abi.PE_PRODUCT_LOOKBACK, side, expiry, strike, dates=dates)
        req.set_pseudo(seed_t, paths)
        req.set_risk(delta=risk, vega=risk, smoothing_half_width=BATCH_HALF_WIDTH)
        req.trade_id = tid
        out.append(req)
    return out


# ----------------------------------------------------- lifted Heston smile --
ROUGH = {"hurst": 0.1, "v0": 0.02, "theta": 0.02, "kappa": 0.3, "nu": 0.3, "rho": -0.7, "factors": 20,
         "ratio": 2.5}
SMILE_MATURITIES = (0.1, 0.25, 0.5, 0.75, 1.0, 1.25, 1.5, 2.0)


def lifted_heston_model(weights_rates):
    """RoughHeston{H=0.1, v0=0.02, kappa=0.3, theta=0.02, nu=0.3, rho=-0.7}
    lifted with 20 geometric factors of ratio 2.5; weights_rates =
    (weights, rates) from LiftedHeston::from_rough."""
    w, x = weights_rates
    r = ROUGH
    return abi.lifted_heston(r["v0"], r["kappa"], r["theta"], r["nu"], r["rho"], w, x)


def rough_smile(points=1 << 16, scrambles=16, strikes=21, maturities=SMILE_MATURITIES, rate=0.0, seed=0x5EED5EED,
                antithetic=False):
    """European call carriers of the lifted Heston smile: 21 strikes
    0.8..1.2 S0 x 8 maturities (round(365 m) days), RQMC points x scrambles
    (pseudo-MC units when scrambles = 0)."""
    out = []
    for m in maturities:
        day = int(round(365 * m))
        for i in range(strikes):
            k = 100.0 * (0.8 + 0.4 * i / (strikes - 1))
            req = request.Request()
            req.set_market(100.0, [0.0, 1.0], [1.0, math.exp(-rate)], [0.0, 1.0], [1.0, 1.0])
            req.set_black_scholes(0.2)
            req.set_product(abi.PE_PRODUCT_EUROPEAN, abi.PE_CALL, day, k)
            if scrambles:
                req.set_rqmc(seed, points, scrambles, antithetic=antithetic)
            else:
                req.set_pseudo(seed, points, antithetic=antithetic)
            req.set_risk()
            out.append(req)
    return out


AMERICAN_RATE = 0.03
BERMUDAN_RATE = 0.05


def rough_bermudan(units=1 << 18, train_units=1 << 18, dates=12, strike=100.0, rate=BERMUDAN_RATE, expiry_days=365,
                   seed=0xBE5D_0001, train_seed=0xBE5D_0002, method=abi.PE_LSM_POLY_CPQR, max_degree=2,
                   scrambles=0, antithetic=False, mlp_epochs=4, mlp_batch=8192):
    """Bermudan put on the lifted-Heston smile model: S0 = 100, K = 100,
    `dates` exercise dates at round(expiry_days k / dates) days, r = 5%
    (an American put with r = 0 never exercises early), q = 0; pseudo-MC
    valuation paths (domain Valuation) and independent training paths
    (domain LsmTrain), or RQMC with `scrambles`.  The regression features are
    the spot and the variance; the basis is the standardized monomials of
    total degree <= max_degree (CPQR policy) or the network policy."""
    days = sorted({max(1, int(round(expiry_days * (k + 1) / dates))) for k in range(dates)})
    horizon = max(1.0, days[-1] / 365.0)
    req = request.Request()
    req.set_market(100.0, [0.0, horizon], [1.0, math.exp(-rate * horizon)], [0.0, horizon], [1.0, 1.0])
    req.set_black_scholes(0.2)
    req.set_product(abi.PE_PRODUCT_AMERICAN, abi.PE_PUT, days[-1], strike, dates=days)
    if scrambles:
        req.set_rqmc(seed, units, scrambles, antithetic=antithetic)
        training = (abi.PE_ENGINE_RQMC, train_seed, train_units, scrambles)
    else:
        req.set_pseudo(seed, units, antithetic=antithetic)
        training = (abi.PE_ENGINE_PSEUDO, train_seed, train_units, 0)
    req.set_risk()
    req.set_lsm(training, antithetic=antithetic, method=method, max_degree=max_degree, mlp_epochs=mlp_epochs,
                mlp_batch=mlp_batch)
    return req


def american_batch(count=256, seed=0x5EED_1A5E, train_log2=18, value_log2=18, dates=50,
                   method=abi.PE_LSM_POLY_LAGUERRE, antithetic=False, dividends=False, mlp_epochs=4,
                   mlp_batch=8192):
    """American puts priced by LSM: S0 = 100, K ~ U[80, 120], sigma ~ U[0.2,
    0.4], r = 3%, q = 0, T = 1 year, `dates` exercise dates at round(365 k /
    dates) days; per trade 2^train_log2 training paths (pseudo-MC, domain
    LsmTrain) and 2^value_log2 independent valuation paths, seeds per trade;
    the weighted-Laguerre basis of degree 3 (or the method given).  With
    `dividends`, one 1.0 cash + 1% dividend at half a year."""
    rng = SplitMix64(seed)
    out = []
    horizon = 1.0
    days = sorted({max(1, int(round(365.0 * (k + 1) / dates))) for k in range(dates)})
    for tid in range(count):
        strike = rng.uniform(80.0, 120.0)
        vol = rng.uniform(0.2, 0.4)
        seed_t, seed_v = rng.next(), rng.next()
        req = request.Request()
        divs = [(182 / 365.0, 1.0, 0.01)] if dividends else []
        req.set_market(100.0, [0.0, horizon], [1.0, math.exp(-AMERICAN_RATE * horizon)], [0.0, horizon], [1.0, 1.0],
                       divs)
        req.set_black_scholes(vol)
        req.set_product(abi.PE_PRODUCT_AMERICAN, abi.PE_PUT, days[-1], strike, dates=days)
        req.set_pseudo(seed_v, 1 << value_log2, antithetic=antithetic)
        req.set_risk()
        req.set_lsm((abi.PE_ENGINE_PSEUDO, seed_t, 1 << train_log2, 0), antithetic=antithetic, method=method,
                    mlp_epochs=mlp_epochs, mlp_batch=mlp_batch)
        req.trade_id = tid
        out.append(req)
    return out
