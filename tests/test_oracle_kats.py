"""Pins the oracle's sampling and reduction layers against the reference's own
known-answer tests (cited per test) and against mathematical identities."""
import ctypes as C
import hashlib
import math

import numpy as np
import pytest
from scipy.special import ndtri

from _support import darr, oracle

U32 = C.c_uint32


def philox(seed, counter):
    out = (U32 * 4)()
    oracle().oracle_philox(seed, (U32 * 4)(*counter), out)
    return list(out)


def test_direction_asset_sha256_matches_reference_digest():
    # crates/pricing/data/README.md:20-21 records the digest of the embedded asset.
    from pricing_amd import sobol_directions

    data = sobol_directions.asset_bytes()
    assert hashlib.sha256(data).hexdigest() == sobol_directions.SHA256
    assert len(data) == 20 + 21201 * 32 * 4


def test_philox_random123_zero_vector():
    # random.rs:290-295
    assert philox(0, [0, 0, 0, 0]) == [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]


def test_philox_coordinate_and_lane_layout():
    # random.rs:297-325: counter = [path lo, path hi, dim/4, domain], lane = dim%4.
    lib = oracle()
    block = philox(17, [29, 0, 0, 0])
    for dim in range(4):
        assert lib.oracle_philox_word(17, 29, dim, 0) == block[dim]
    seed, path = 0x0123456789ABCDEF, 0x1122334455667788
    block = philox(seed, [0x55667788, 0x11223344, 1, 1])
    for dim in range(4, 8):
        assert lib.oracle_philox_word(seed, path, dim, 1) == block[dim - 4]
    # random.rs:327-334 domain separation
    assert lib.oracle_philox_word(42, 5, 7, 0) != lib.oracle_philox_word(42, 5, 7, 1)


def test_philox_matches_independent_python_model():
    # Independent restatement of the Random123 round (random.rs:131-143).
    def model(seed, c):
        k0, k1 = seed & 0xFFFFFFFF, seed >> 32
        c = list(c)
        for r in range(10):
            p0 = 0xD2511F53 * c[0]
            p1 = 0xCD9E8D57 * c[2]
            c = [((p1 >> 32) ^ c[1] ^ k0) & 0xFFFFFFFF, p1 & 0xFFFFFFFF,
                 ((p0 >> 32) ^ c[3] ^ k1) & 0xFFFFFFFF, p0 & 0xFFFFFFFF]
            if r < 9:
                k0 = (k0 + 0x9E3779B9) & 0xFFFFFFFF
                k1 = (k1 + 0xBB67AE85) & 0xFFFFFFFF
        return c

    rng = np.random.default_rng(3)
    for _ in range(50):
        seed = int(rng.integers(0, 2**63)) * 2 + 1
        ctr = [int(x) for x in rng.integers(0, 2**32, 4)]
        assert philox(seed, ctr) == model(seed, ctr)


def test_midpoint_mapping_is_open():
    # random.rs:337-342
    lib = oracle()
    assert lib.oracle_open_unit_interval(0) == 2.0**-33
    assert lib.oracle_open_unit_interval(0xFFFFFFFF) == 1.0 - 2.0**-33


def as241(p):
    out = C.c_double()
    rc = oracle().oracle_as241(p, C.byref(out))
    return rc, out.value


def test_as241_reference_quantiles():
    # random.rs:345-356
    assert as241(0.5) == (0, 0.0)
    assert abs(as241(0.975)[1] - 1.959963984540054) < 2e-15
    assert abs(as241(1e-10)[1] - -6.361340902404056) < 2e-15
    assert abs(as241(0.975)[1] + as241(0.025)[1]) < 2e-15
    assert as241(0.0)[0] != 0 and as241(1.0)[0] != 0 and as241(math.nan)[0] != 0


def test_as241_against_scipy_ndtri_all_branches():
    # Wichura AS241 has relative error ~1e-16; a wrong coefficient or branch
    # bound shows up as a gross deviation from ndtri in one of the regions.
    # Domain reachable from a u32 midpoint: [2^-33, 1 - 2^-33].
    lo = 2.0**-33
    ps = np.concatenate([np.geomspace(lo, 0.075, 400), np.linspace(0.075, 0.925, 400),
                         1 - np.geomspace(lo, 0.075, 400)])
    for p in ps:
        rc, z = as241(float(p))
        assert rc == 0
        ref = ndtri(p)
        assert abs(z - ref) <= 1e-13 * max(1.0, abs(ref)), (p, z, ref)


def test_sobol_first_words_three_dimensions():
    # qmc.rs:377-401
    expected = [
        [0, 0, 0],
        [0x80000000, 0x80000000, 0x80000000],
        [0xC0000000, 0x40000000, 0x40000000],
        [0x40000000, 0xC0000000, 0xC0000000],
        [0x60000000, 0x60000000, 0xA0000000],
        [0xE0000000, 0xE0000000, 0x20000000],
        [0xA0000000, 0x20000000, 0xE0000000],
        [0x20000000, 0xA0000000, 0x60000000],
    ]
    lib = oracle()
    for point, row in enumerate(expected):
        for dim, word in enumerate(row):
            assert lib.oracle_sobol_word(dim, point) == word


def rqmc_words(seed, scrambles, dims, points, tables=False):
    words = (U32 * (scrambles * dims * points))()
    shifts = (U32 * (scrambles * dims))() if tables else None
    rows = (U32 * (scrambles * dims * 32))() if tables else None
    assert oracle().oracle_rqmc_words(seed, scrambles, dims, points, words, shifts, rows) == 0
    w = np.array(words, dtype=np.uint64).reshape(scrambles, points, dims)
    if tables:
        return w, np.array(shifts).reshape(scrambles, dims), np.array(rows).reshape(scrambles, dims, 32)
    return w


def test_scramble_abi_known_answer_words():
    # qmc.rs:467-482
    w = rqmc_words(0x0123456789ABCDEF, 3, 2, 8)
    assert w[0, 0, 0] == 0x46C5FD35
    assert w[0, 1, 0] == 0xE684A5D7
    assert w[0, 2, 1] == 0x7D21CC82
    assert w[1, 2, 0] == 0xAA93E083
    assert w[2, 7, 1] == 0xC2B40C80


def test_lms_rows_are_unit_lower_triangular_and_point_zero_is_shift():
    # qmc.rs:415-424 and :445-453
    w, shifts, rows = rqmc_words(123, 2, 8, 4, tables=True)
    for s in range(2):
        for d in range(8):
            assert rows[s, d, 0] == 0x80000000
            for r in range(32):
                diag = 1 << (31 - r)
                assert rows[s, d, r] & diag
                assert rows[s, d, r] & (diag - 1) == 0
            assert w[s, 0, d] == shifts[s, d]


def test_lms_linearity_prescrambled_directions():
    # The device engine XORs pre-scrambled direction words: L(a^b) = L(a)^L(b).
    w, shifts, rows = rqmc_words(99, 1, 3, 64, tables=True)
    lib = oracle()

    def lms(x, r):
        out = 0
        for i in range(32):
            out |= (bin(x & int(r[i])).count("1") & 1) << (31 - i)
        return out

    for d in range(3):
        for p in range(64):
            raw = lib.oracle_sobol_word(d, p)
            assert w[0, p, d] == lms(raw, rows[0, d]) ^ shifts[0, d]


def bridge(times, normals):
    n = len(times)
    inc = (C.c_double * (n - 1))()
    nodes = (U32 * (n - 1))()
    ranks = (U32 * (n - 1))()
    lw = (C.c_double * (n - 1))()
    rw = (C.c_double * (n - 1))()
    sd = (C.c_double * (n - 1))()
    assert oracle().oracle_bridge_apply(darr(times), n, darr(normals), inc, nodes, ranks, lw, rw, sd) == 0
    return list(inc), list(nodes), list(ranks), list(lw), list(rw), list(sd)


def test_bridge_order_and_coefficients():
    # bridge.rs:481-516
    _, nodes, ranks, _, _, _ = bridge([0.0, 0.25, 0.5, 0.75, 1.0], [0.0] * 4)
    assert nodes == [4, 2, 1, 3]
    assert ranks == [0, 1, 2, 3]
    _, nodes, _, lw, rw, sd = bridge([0.0, 0.1, 0.4, 1.0], [0.0] * 3)
    assert nodes[:2] == [3, 2]
    assert lw[1] == 0.6 and rw[1] == 0.4 and sd[1] == math.sqrt(0.24)


def test_bridge_forward_known_increments():
    # bridge.rs:518-535
    inc, *_ = bridge([0.0, 0.25, 0.5, 0.75, 1.0], [1.0, 2.0, 3.0, 4.0])
    r8 = math.sqrt(0.125)
    expected = [2 * (0.75 + 3 * r8), 2 * (0.75 - 3 * r8), 2 * (-0.25 + 4 * r8), 2 * (-0.25 - 4 * r8)]
    assert max(abs(a - b) for a, b in zip(inc, expected)) < 2e-15


def test_bridge_reproduces_brownian_covariance():
    # The bridge is a linear map z -> increments; increments must be i.i.d.
    # N(0,1) (identity covariance) for any grid: a wrong weight or variance fails.
    rng = np.random.default_rng(11)
    times = np.concatenate([[0.0], np.sort(rng.uniform(0.01, 2.0, 9))])
    n = len(times) - 1
    m = np.zeros((n, n))
    for k in range(n):
        e = [0.0] * n
        e[k] = 1.0
        m[:, k] = bridge(list(times), e)[0]
    assert np.allclose(m @ m.T, np.eye(n), atol=1e-12)


def test_bridge_reverse_is_transpose():
    # bridge.rs:559-580
    times = [0.0, 0.1, 0.4, 0.7, 1.0]
    z = [0.3, -1.1, 0.7, 2.0]
    adj = [-0.2, 0.9, 1.3, -0.4]
    inc = bridge(times, z)[0]
    na = (C.c_double * 4)()
    assert oracle().oracle_bridge_reverse(darr(times), 5, darr(adj), na) == 0
    assert abs(sum(a * b for a, b in zip(inc, adj)) - sum(a * b for a, b in zip(z, na))) < 2e-15


def test_bridge_rejects_invalid_grids():
    # bridge.rs:582-620
    lib = oracle()
    out = (C.c_double * 4)()
    assert lib.oracle_bridge_apply(darr([0.0]), 1, darr([0.0]), out, None, None, None, None, None) != 0
    assert lib.oracle_bridge_apply(darr([-0.0, 1.0]), 2, darr([0.0]), out, None, None, None, None, None) != 0
    assert lib.oracle_bridge_apply(darr([0.0, 0.5, 0.5]), 3, darr([0.0, 0.0]), out, None, None, None, None,
                                   None) != 0


def test_neumaier_and_two_pass_kats():
    # reduction.rs:477-500
    lib = oracle()
    v = [1e16, 1.0, -1e16, 2.0]
    assert lib.oracle_neumaier(darr(v), 4) == 3.0
    out = (C.c_double * 3)()
    lib.oracle_two_pass(darr(v), 4, out)
    assert out[0] == 4 and out[1] == 0.75 and out[2] > 0


def test_block_reduction_is_worker_count_invariant():
    # execution.rs:801-829: results depend on the block table only.
    rng = np.random.default_rng(5)
    values = rng.standard_normal(10007) * 1e3 + 1e6
    lib = oracle()
    outs = []
    for workers in (1, 2, 7):
        out = (C.c_double * 4)()
        lib.oracle_block_reduce(values.ctypes.data_as(C.POINTER(C.c_double)), len(values), 256, workers, out)
        outs.append(list(out))
    assert outs[0] == outs[1] == outs[2]
    assert outs[0][1] == len(values)
    assert abs(outs[0][0] - math.fsum(values)) <= 1e-9
    assert abs(outs[0][3] / (len(values) - 1) - np.var(values, ddof=1)) < 1e-6 * np.var(values)
