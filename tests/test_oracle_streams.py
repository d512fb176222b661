"""Pins the oracle's random-number layer (CPU only).

  * raw PCG64DXSM streams against NumPy's PCG64DXSM (a library routine), with
    the state/increment closed form of pcg64.rs:23-30, at every block size the
    steppers use and at path indices up to 2^20;
  * NumPy's PCG64DXSM against rand_pcg's Lcg128CmDxsm64::new(42, 54) known
    answer (the vector rand_pcg's own tests hold), so the chain
    reference -> rand_pcg -> NumPy -> oracle is closed;
  * the ziggurat tables against the reference's literals (fixture) and the
    integer fast-path thresholds against the floating-point test they replace;
  * the ziggurat's output distribution (moments, KS) and its draw rate
    against the rate implied by the tables.
"""
import json
import math
import os
import re
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path.insert(0, os.path.join(ROOT, "stocha_amd", "csrc"))

M = 0xDA942042E4DD58B5
MASK = (1 << 128) - 1
STREAM_OFFSET = 0xDA3E39CB94B95BDB7EF55DA5C0374F5A


def numpy_stream(seed, block, path):
    """NumPy PCG64DXSM positioned where stocha's path `path` starts."""
    stream = (seed + STREAM_OFFSET) & MASK
    inc = ((stream << 1) | 1) & MASK
    s0 = ((seed + inc) * M + inc) & MASK
    bg = np.random.PCG64DXSM()
    bg.state = {"bit_generator": "PCG64DXSM", "state": {"state": s0, "inc": inc},
                "has_uint32": 0, "uinteger": 0}
    bg.advance(path * block)
    return bg


def test_rand_pcg_known_answer():
    # Lcg128CmDxsm64::new(state=42, stream=54): inc = (54 << 1) | 1, state = (42 + inc)
    # stepped once -- the known-answer vector of rand_pcg's test suite.
    inc = (54 << 1) | 1
    s = ((42 + inc) * M + inc) & MASK
    bg = np.random.PCG64DXSM()
    bg.state = {"bit_generator": "PCG64DXSM", "state": {"state": s, "inc": inc},
                "has_uint32": 0, "uinteger": 0}
    assert bg.random_raw(6).tolist() == [17331114245835578256, 10267467544499227306,
                                         9726600296081716989, 10165951391103677450,
                                         12131334649314727261, 10134094537930450875]


@pytest.mark.parametrize("seed", [0, 1, 42, 12345, 2**64 - 1])
@pytest.mark.parametrize("block", [2552, 7656, 10208, 2148, 34304])
def test_oracle_streams_match_numpy(orc, seed, block):
    for path in (0, 1, 2, 255, 256, 99_999, 262_143, 1_048_575):
        ref = numpy_stream(seed, block, path).random_raw(40)
        np.testing.assert_array_equal(orc.raw_u64(seed, block, path, 40), ref)


def test_stocha_seed42_stream_start(orc):
    # the first draws of stocha's seed-42 stream (pcg64.rs:23-30)
    assert orc.raw_u64(42, 2552, 0, 4).tolist() == [4165707071428774973, 13625534677717622090,
                                                     8542257868084568114, 4313261270648151989]


def test_advance_is_repeated_stepping(orc):
    st = orc.stream_state(42, 2552, 0)
    a = st.copy()
    orc.lib().orc_advance(a.ctypes.data_as(orc._pu), 1000, 0)
    b = orc.stream_state(42, 1000, 1)  # path 1 with block 1000 = advance(1000)
    np.testing.assert_array_equal(a[:2], b[:2])
    # and against NumPy at a 2^70 jump (high word of delta)
    c = st.copy()
    orc.lib().orc_advance(c.ctypes.data_as(orc._pu), 7, 1 << 6)
    bg = numpy_stream(42, 1, 0)
    bg.advance((1 << 70) + 7)
    s = bg.state["state"]["state"]
    assert int(c[0]) == s & ((1 << 64) - 1) and int(c[1]) == s >> 64


# -------------------------------------------------------------- ziggurat --
def _header_values(path, name):
    text = open(path).read()
    body = text.split(f"#define {name} \\")[1].split("\n\n")[0]
    vals = re.findall(r"(-?0x[0-9a-fp.+-]+|\d+)ull|(-?0x[0-9a-fp.+-]+)", body)
    out = []
    for a, b in vals:
        out.append(int(a) if a and not a.startswith("0x") else float.fromhex(a or b))
    return out


@pytest.fixture(scope="module")
def tables():
    import gen_ziggurat_tables as g
    r, X, Y = g.build_tables()
    return r, X, Y, g.fast_thresholds(X)


def test_tables_equal_reference_literals(tables):
    r, X, Y, _ = tables
    ref = json.load(open(os.path.join(HERE, "golden", "zig_reference_tables.json")))
    assert r == float.fromhex(ref["R"])
    assert X == [float.fromhex(v) for v in ref["X"]]
    assert Y == [float.fromhex(v) for v in ref["Y"]]


@pytest.mark.parametrize("path", ["stocha_amd/csrc/zig_tables.h", "oracle/zig_tables.h"])
def test_committed_headers_are_generated(tables, path):
    r, X, Y, T = tables
    p = os.path.join(ROOT, path)
    assert _header_values(p, "ZIG_X_VALUES") == X
    assert _header_values(p, "ZIG_Y_VALUES") == Y
    assert _header_values(p, "ZIG_T_VALUES") == T


def test_thresholds_equal_float_test(tables):
    _, X, _, T = tables
    rng = np.random.default_rng(0)
    sc = 1.0 / float(1 << 53)
    for l in range(256):
        acc = lambda m: (float(m) * sc) * X[l] < X[l + 1]  # noqa: E731  normal.rs:23-27
        t = T[l]
        if t > 0:
            assert acc(t - 1)
        if t < (1 << 53):
            assert not acc(t)
        for m in rng.integers(0, 1 << 53, 50, dtype=np.int64):
            assert acc(int(m)) == (int(m) < t)
    assert T[255] == 0  # layer 255 never takes the fast path
    assert T[254] == 1


def test_layer255_wedge_always_rejects(tables):
    _, X, Y, _ = tables
    # y = Y[255] + u*(Y[256]-Y[255]) = 1 and exp(-x^2/2) rounds to 1 for x <= X[255]
    assert Y[255] == 1.0 and Y[256] == 1.0
    assert math.exp(-0.5 * X[255] * X[255]) == 1.0


def _reference_density(tables):
    """Output density of the reference's ziggurat (normal.rs:18-67), derived
    from its table geometry: layer l is the rectangle [0, X[l]] x [Y[l], Y[l+1]]
    (X[l](Y[l+1]-Y[l]) = V); layer 0 accepts x < X[1] and sends the rest to the
    tail.  Summing the fast and wedge acceptances over layers telescopes to
    f(x) - f(R) on [0, X[1]), nothing on [X[1], R), and the normal tail beyond R
    with weight V (X[0]-X[1])/X[0].  (The base strip [0,R] x [0,f(R)] of the
    textbook ziggurat is never sampled, so the variance is below 1.)"""
    from scipy import integrate
    r, X, Y, _ = tables
    f = lambda x: math.exp(-0.5 * x * x)  # noqa: E731
    tail = math.sqrt(math.pi / 2) * math.erfc(r / math.sqrt(2))
    v = r * f(r) + tail
    wt = v * (X[0] - X[1]) / X[0]
    m0 = integrate.quad(lambda x: f(x) - f(r), 0, X[1], epsabs=1e-15)[0] + wt
    m2 = integrate.quad(lambda x: x * x * (f(x) - f(r)), 0, X[1], epsabs=1e-15)[0] + \
        wt * integrate.quad(lambda x: x * x * f(x), r, np.inf)[0] / tail

    def cdf_abs(a, b):  # P(a <= |Z| < b)
        lo, hi = min(a, X[1]), min(b, X[1])
        p = 0.0
        if hi > lo:
            p += integrate.quad(lambda x: f(x) - f(r), lo, hi, epsabs=1e-15)[0]
        lo, hi = max(a, r), max(b, r)
        if hi > lo:
            p += wt * integrate.quad(f, lo, hi)[0] / tail
        return p / m0

    return m2 / m0, cdf_abs, m0 / (256 * v), v


def test_normals_match_derived_distribution(orc, tables):
    var, cdf_abs, _, _ = _reference_density(tables)
    assert abs(math.sqrt(var) - 0.99278) < 1e-4
    z, _ = orc.normals(42, 2552, 0, 2_000_000)
    assert abs(z.mean()) < 5 / math.sqrt(z.size)
    # sample variance vs the derived one (sd of s^2 ~ sqrt(2/n) * var)
    assert abs(z.var() - var) < 5 * math.sqrt(2.0 / z.size) * var
    # the reference's own tolerance (tests/test_rng.py:26-31)
    assert abs(z.std() - 1.0) < 0.01
    # binned |z| against the derived law, including the empty [X1, R) band
    edges = [0, 0.5, 1, 1.5, 2, 2.5, 3, 3.3, tables[1][1], tables[0], 4.0, 60.0]
    h, _ = np.histogram(np.abs(z), edges)
    for i in range(len(edges) - 1):
        p = cdf_abs(edges[i], edges[i + 1])
        sd = math.sqrt(max(p * (1 - p) * z.size, 1.0))
        assert abs(h[i] - p * z.size) < 5 * sd, (edges[i], h[i], p * z.size)
    assert h[8] == 0  # no output between X[1] and R


def test_normals_ks_reference_tolerance(orc):
    from scipy import stats
    z, _ = orc.normals(123, 2552, 0, 100_000)
    assert stats.kstest(z, "norm").pvalue > 0.01  # tests/test_rng.py:105-110


def test_draw_rate_matches_tables(orc, tables):
    """u64 draws per normal predicted from the table geometry."""
    _, X, _, T = tables
    _, _, acc, _ = _reference_density(tables)
    p_fast = sum(t / 2.0**53 for t in T) / 256
    p_tail = (1 / 256) * (1 - T[0] / 2.0**53)
    # a non-fast attempt in layers 1..255 draws one wedge uniform; a tail
    # excursion draws pairs (acceptance ~0.93 per pair at R = 3.654)
    per_attempt = 1.0 + (1.0 - p_fast - p_tail) + p_tail * 2 / 0.93
    draws = per_attempt / acc
    n = 400_000
    _, used = orc.normals(7, 2552, 3, n)
    assert abs(used / n - draws) < 0.002, (used / n, draws)


def test_tie_fixture_positions_are_ties(orc, tables):
    # tests/golden/zig_ties.json (tests/golden/find_zig_ties.c): each listed
    # draw is the first attempt of its normal and ties the fast test on the
    # high word, (u >> 32) == T[l] >> 21, with the exact test's outcome as
    # recorded; the GPU tie tests rely on these positions
    _, _, _, T = tables
    fix = json.load(open(os.path.join(HERE, "golden", "zig_ties.json")))
    seed, block = fix["seed"], fix["block"]
    assert {t["accepted"] for t in fix["ties"]} == {True, False}
    for t in fix["ties"]:
        u = int(orc.raw_u64(seed, block, t["path"], t["draw"] + 1)[t["draw"]])
        l = u & 0xFF
        assert (u >> 32) == (T[l] >> 21), t
        assert ((u >> 11) < T[l]) == t["accepted"], t
        _, before = orc.normals(seed, block, t["path"], t["normal"])
        assert before == t["draw"], t
