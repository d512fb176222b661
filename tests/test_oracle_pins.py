"""Pins the oracle's ziggurat bit layout, the Heston Euler recursion and the
Merton draw order/stream block against independently known values (CPU only).

The u64 draws come from NumPy's PCG64DXSM (pinned to rand_pcg by the
known-answer test in test_oracle_streams.py) and the ziggurat tables from the
reference's own literals (tests/golden/zig_reference_tables.json).  Each check
is a special case where the reference's expression collapses to something that
can be written down from those draws, chosen so that a plausible slip in the
oracle fails it:

  * ziggurat (normal.rs:18-67): layer = u & 0xFF, sign = bit 8 (bits 9-10
    unused), x = (u >> 11) * 2^-53 * X[layer]; the wedge uniform is the NEXT
    draw; a rejected attempt restarts on the draw after it; layer 255 always
    rejects; a layer-0 miss takes the tail pair (u1, u2) from the next two
    draws and returns sign * (-ln(u1)/R + R);
  * Heston Euler (heston.rs:44-87):
      - vanishing vol-of-vol with v0 != theta: the variance follows
        theta + (v0 - theta)(1 - kappa dt)^n and ln S_T is the sum of the
        per-step log increments with the PRE-update variance;
      - two steps at rho in {-1, 0, 1, -0.7}: v1 from the closed form of
        dw2 at those rho (dw2 = +-dw1 exactly at |rho| = 1, dz at rho = 0);
      - a variance that goes negative: the next log increment is exactly
        mu*dt (full truncation in the S step) and the variance after it is
        v1 + kappa*theta*dt (full truncation in the drift);
  * Merton (jump_diffusion.rs:48-105): the diffusion normal first, then
    always one uniform (also at lambda = 0), then the jump normal when
    u < lambda*dt; path p starts at position p*(steps+1024)*8 of the seed's
    stream.
"""
import json
import math
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
M = 0xDA942042E4DD58B5
MASK = (1 << 128) - 1
STREAM_OFFSET = 0xDA3E39CB94B95BDB7EF55DA5C0374F5A
TWO_M53 = 1.0 / float(1 << 53)


def numpy_base_stream(seed):
    """NumPy PCG64DXSM at position 0 of stocha's stream for `seed` (pcg64.rs:23-30)."""
    stream = (seed + STREAM_OFFSET) & MASK
    inc = ((stream << 1) | 1) & MASK
    s0 = ((seed + inc) * M + inc) & MASK
    bg = np.random.PCG64DXSM()
    bg.state = {"bit_generator": "PCG64DXSM", "state": {"state": s0, "inc": inc},
                "has_uint32": 0, "uinteger": 0}
    return bg


@pytest.fixture(scope="module")
def ref_tables():
    ref = json.load(open(os.path.join(HERE, "golden", "zig_reference_tables.json")))
    return (float.fromhex(ref["R"]), [float.fromhex(v) for v in ref["X"]],
            [float.fromhex(v) for v in ref["Y"]])


def f64(u):
    """next_f64 of a draw (pcg64.rs:37-39)."""
    return float(int(u) >> 11) * TWO_M53


def attempt(u, X):
    """(layer, sign, x, fast) of one ziggurat attempt on draw u (normal.rs:20-27)."""
    u = int(u)
    layer = u & 0xFF
    sign = -1.0 if (u >> 8) & 1 else 1.0
    x = float(u >> 11) * TWO_M53 * X[layer]
    return layer, sign, x, x < X[layer + 1]


# ------------------------------------------------------------- ziggurat --
@pytest.fixture(scope="module")
def stream42():
    return numpy_base_stream(42).random_raw(300_000)


def test_fast_path_bit_layout(orc, ref_tables, stream42):
    _, X, _ = ref_tables
    raw = stream42
    seen = set()
    n = 0
    for pos in range(0, 6000):
        layer, sign, x, fast = attempt(raw[pos], X)
        if not fast:
            continue
        z, used = orc.normals(42, 1, pos, 1)
        assert used == 1
        assert z[0] == sign * x, (pos, hex(int(raw[pos])))
        seen.add((int(raw[pos]) >> 8) & 7)
        n += 1
    # every combination of bits 8, 9, 10 occurred: the sign is bit 8 alone
    assert seen == set(range(8)) and n > 5000


def test_fast_path_explicit_values(orc, ref_tables, stream42):
    # the first fast-path draws with the sign bit clear and set, written out
    _, X, _ = ref_tables
    raw = stream42
    done = {0: False, 1: False}
    for pos in range(200):
        u = int(raw[pos])
        layer, _, x, fast = attempt(u, X)
        b8 = (u >> 8) & 1
        if not fast or done[b8]:
            continue
        expect = ((u >> 11) * 2.0**-53) * X[u & 255]
        expect = -expect if b8 else expect
        assert orc.normals(42, 1, pos, 1)[0][0] == expect
        done[b8] = True
    assert all(done.values())


def test_wedge_uses_next_draw(orc, ref_tables, stream42):
    _, X, Y = ref_tables
    raw = stream42
    acc = rej = 0
    for pos in range(0, 60_000):
        layer, sign, x, fast = attempt(raw[pos], X)
        if fast or layer == 0:
            continue
        y = Y[layer] + f64(raw[pos + 1]) * (Y[layer + 1] - Y[layer])
        z, used = orc.normals(42, 1, pos, 1)
        if y < math.exp(-0.5 * x * x):  # accepted in the wedge
            assert used == 2 and z[0] == sign * x, pos
            acc += 1
        else:  # rejected: the next attempt starts two draws on
            l2, s2, x2, f2 = attempt(raw[pos + 2], X)
            if f2:
                assert used == 3 and z[0] == s2 * x2, pos
                rej += 1
    assert acc > 50 and rej > 50


def test_layer255_always_rejects(orc, ref_tables, stream42):
    _, X, _ = ref_tables
    raw = stream42
    hits = 0
    for pos in np.nonzero((raw[:100_000] & np.uint64(0xFF)) == 255)[0][:200]:
        l2, s2, x2, f2 = attempt(raw[pos + 2], X)
        if not f2:
            continue
        z, used = orc.normals(42, 1, int(pos), 1)
        assert used == 3 and z[0] == s2 * x2
        hits += 1
    assert hits > 100


def test_tail_pair(orc, ref_tables, stream42):
    R, X, _ = ref_tables
    raw = stream42
    acc = 0
    for pos in range(0, len(raw) - 8):
        layer, sign, x, fast = attempt(raw[pos], X)
        if layer != 0 or fast:
            continue
        u1, u2 = f64(raw[pos + 1]), f64(raw[pos + 2])
        if u1 <= 0.0 or u2 <= 0.0:
            continue
        xt = -math.log(u1) / R
        z, used = orc.normals(42, 1, pos, 1)
        if -2.0 * math.log(u2) >= xt * xt:
            assert used == 3 and z[0] == sign * (xt + R), pos
            acc += 1
        else:
            assert used >= 5
    assert acc > 20


# --------------------------------------------------------------- Heston --
def test_heston_vanishing_volvol_variance_decay(orc):
    s0, v0, mu, kappa, theta, xi, rho, t, steps = 100.0, 0.09, 0.05, 3.0, 0.04, 1e-14, -0.7, 1.0, 40
    n = 6
    paths = orc.heston(s0, v0, mu, kappa, theta, xi, rho, t, steps, n, seed=11)
    dt = t / steps
    v = theta + (v0 - theta) * (1.0 - kappa * dt) ** np.arange(steps)  # v before step j
    for i in range(n):
        z, _ = orc.normals(11, orc.block_heston(steps), i, 2 * steps)
        dw1 = z[0::2]
        ls = math.log(s0) + np.cumsum((mu - 0.5 * v) * dt + np.sqrt(v * dt) * dw1)
        np.testing.assert_allclose(np.log(paths[i, 1:]), ls, rtol=0, atol=1e-11)


@pytest.mark.parametrize("rho", [-1.0, 0.0, 1.0, -0.7])
def test_heston_two_steps_rho_cases(orc, rho):
    s0, v0, mu, kappa, theta, xi, t, steps = 100.0, 0.04, 0.05, 2.0, 0.06, 0.3, 0.02, 2
    n = 40
    paths = orc.heston(s0, v0, mu, kappa, theta, xi, rho, t, steps, n, seed=5)
    dt = t / steps
    for i in range(n):
        z, _ = orc.normals(5, orc.block_heston(steps), i, 4)
        dw1_0, dz_0, dw1_1 = z[0], z[1], z[2]
        if rho == 1.0:
            dw2 = dw1_0
        elif rho == -1.0:
            dw2 = -dw1_0
        elif rho == 0.0:
            dw2 = dz_0
        else:
            dw2 = rho * dw1_0 + math.sqrt(1.0 - rho * rho) * dz_0
        v1 = v0 + kappa * (theta - v0) * dt + xi * math.sqrt(v0) * math.sqrt(dt) * dw2
        assert v1 > 0
        l1 = math.log(paths[i, 1] / s0)
        assert l1 == pytest.approx((mu - 0.5 * v0) * dt + math.sqrt(v0 * dt) * dw1_0,
                                   abs=1e-14)
        l2 = math.log(paths[i, 2] / paths[i, 1])
        assert l2 == pytest.approx((mu - 0.5 * v1) * dt + math.sqrt(v1 * dt) * dw1_1, abs=1e-13)


def test_heston_negative_variance_full_truncation(orc):
    s0, v0, mu, kappa, theta, xi, rho, t, steps = 100.0, 0.01, 0.05, 4.0, 0.09, 3.0, 0.3, 1.0, 4
    n = 400
    paths = orc.heston(s0, v0, mu, kappa, theta, xi, rho, t, steps, n, seed=3)
    dt = t / steps
    rc = math.sqrt(1.0 - rho * rho)
    checked_trunc = checked_next = 0
    for i in range(n):
        z, _ = orc.normals(3, orc.block_heston(steps), i, 6)
        dw2 = rho * z[0] + rc * z[1]
        v1 = v0 + kappa * (theta - v0) * dt + xi * math.sqrt(v0) * math.sqrt(dt) * dw2
        if v1 >= -1e-3:
            continue
        # v1 < 0: the second step's log increment is exactly mu*dt
        assert paths[i, 2] == paths[i, 1] * math.exp(mu * dt), i
        checked_trunc += 1
        # and v2 = v1 + kappa*theta*dt (the drift uses v+ = 0)
        v2 = v1 + kappa * theta * dt
        if v2 > 1e-3:
            l3 = math.log(paths[i, 3] / paths[i, 2])
            assert l3 == pytest.approx((mu - 0.5 * v2) * dt + math.sqrt(v2 * dt) * z[4],
                                       abs=1e-12), i
            checked_next += 1
    assert checked_trunc > 40 and checked_next > 20


# --------------------------------------------------------------- Merton --
@pytest.mark.parametrize("path", [0, 1, 7, 1000])
def test_merton_lambda0_draw_order_and_block(orc, path):
    s0, mu, sigma, lam, mu_j, sigma_j, t, steps, seed = 100.0, 0.05, 0.2, 0.0, -0.1, 0.2, 1.0, 2, 9
    p = orc.merton(s0, mu, sigma, lam, mu_j, sigma_j, t, steps, path + 1, seed=seed)[path]
    dt = t / steps
    drift = (mu - 0.5 * sigma * sigma - lam * (math.exp(mu_j + 0.5 * sigma_j * sigma_j) - 1.0)) * dt
    diff = sigma * math.sqrt(dt)
    base = path * (steps + 1024) * 8
    z0, c0 = orc.normals(seed, 1, base, 1)
    # one uniform at base + c0, drawn although lambda = 0, then the next normal
    z1, _ = orc.normals(seed, 1, base + c0 + 1, 1)
    assert p[1] == s0 * math.exp(drift + diff * z0[0])
    assert p[2] == p[1] * math.exp(drift + diff * z1[0])


def test_merton_jump_normal_follows_uniform(orc):
    s0, mu, sigma, lam, mu_j, sigma_j, t, steps, seed = 100.0, 0.05, 0.2, 60.0, -0.1, 0.2, 1.0, 200, 4
    n = 300
    paths = orc.merton(s0, mu, sigma, lam, mu_j, sigma_j, t, steps, n, seed=seed)
    dt = t / steps
    m_bar = math.exp(mu_j + 0.5 * sigma_j * sigma_j) - 1.0
    drift = (mu - 0.5 * sigma * sigma - lam * m_bar) * dt
    diff = sigma * math.sqrt(dt)
    block = (steps + 1024) * 8
    jumps = nojumps = 0
    for i in range(n):
        base = i * block
        z0, c0 = orc.normals(seed, 1, base, 1)
        g = numpy_base_stream(seed)
        g.advance(base + c0)
        u = f64(g.random_raw(1)[0])
        ret = drift + diff * z0[0]
        if u < lam * dt:
            zj, _ = orc.normals(seed, 1, base + c0 + 1, 1)
            ret = ret + (mu_j + sigma_j * zj[0])
            jumps += 1
        else:
            nojumps += 1
        assert paths[i, 1] == s0 * math.exp(ret), i
    assert jumps > 50 and nojumps > 50


@pytest.mark.parametrize("above", [False, True])
def test_merton_jump_test_is_strict(orc, above):
    """jump_diffusion.rs:80: a jump needs u < lambda*dt, so lambda*dt equal to
    the drawn uniform gives no jump and the next double above it gives one.
    With t = steps = 1, dt = 1 and lambda*dt is lambda itself."""
    s0, mu, sigma, mu_j, sigma_j, t, steps, seed = 100.0, 0.05, 0.2, -0.1, 0.2, 1.0, 1, 11
    z0, c0 = orc.normals(seed, 1, 0, 1)
    g = numpy_base_stream(seed)
    g.advance(int(c0))
    u = f64(g.random_raw(1)[0])
    assert 0.0 < u < 1.0
    lam = math.nextafter(u, 2.0) if above else u
    p = orc.merton(s0, mu, sigma, lam, mu_j, sigma_j, t, steps, 1, seed=seed)[0]
    m_bar = math.exp(mu_j + 0.5 * sigma_j * sigma_j) - 1.0
    ret = (mu - 0.5 * sigma * sigma - lam * m_bar) * 1.0 + sigma * math.sqrt(1.0) * z0[0]
    if above:
        zj, _ = orc.normals(seed, 1, int(c0) + 1, 1)
        ret = ret + (mu_j + sigma_j * zj[0])
    assert p[1] == s0 * math.exp(ret)


# ----------------------------------------------------------------- LSMC --
def test_lsmc_regression_needs_basis_plus_one_itm_paths(orc):
    """lsmc.rs:62: a date is regressed only with at least ncols + 1 paths in
    the money.  Small put books are searched for dates holding exactly ncols
    and exactly ncols + 1 such paths: the first must be skipped (no
    coefficients), the second solved."""
    deg, ncols = 3, 4
    seen = set()
    for seed in range(200):
        coeffs, itm, _, _ = orc.lsmc_trace(100.0, 92.0, 0.05, 0.3, 1.0, 20, 12, True, deg, seed)
        for step in range(1, 20):
            m = int(itm[step])
            if m == ncols:
                assert np.all(np.isnan(coeffs[step])), (seed, step)
                seen.add(m)
            elif m == ncols + 1:
                assert np.all(np.isfinite(coeffs[step])), (seed, step)
                seen.add(m)
        if seen == {ncols, ncols + 1}:
            return
    pytest.fail(f"no dates with {ncols} and {ncols + 1} paths in the money found")


@pytest.mark.parametrize("is_put,deg", [(True, 3), (False, 2)])
def test_lsmc_last_date_coefficients_by_numpy(orc, is_put, deg):
    """lsmc.rs:66-84 at the last regression date, where the cash flow is the
    discounted terminal payoff (no earlier exercise): the basis is powers
    0..deg of the spot standardised by the population std (numpy.std's
    default, ddof = 0) of the in-the-money spots, and the coefficients are
    the least-squares solution (numpy.linalg.lstsq)."""
    s0, k, r, sigma, t, steps, n = 100.0, (104.0 if is_put else 96.0), 0.05, 0.25, 1.0, 10, 4000
    coeffs, itm, _, _ = orc.lsmc_trace(s0, k, r, sigma, t, steps, n, is_put, deg, 42)
    paths = orc.gbm(s0, r, sigma, t, steps, n, seed=42)
    intr = (lambda s: np.maximum(k - s, 0.0)) if is_put else (lambda s: np.maximum(s - k, 0.0))
    spot = paths[:, steps - 1]
    sel = intr(spot) > 0
    assert int(itm[steps - 1]) == int(sel.sum())
    s = spot[sel]
    x = (s - s.mean()) / np.std(s)
    y = intr(paths[sel, steps]) * math.exp(-r * t / steps)
    want = np.linalg.lstsq(np.vander(x, deg + 1, increasing=True), y, rcond=None)[0]
    np.testing.assert_allclose(coeffs[steps - 1], want, rtol=1e-9, atol=1e-12)
