"""The RNG class (src/lib.rs:75-184) on the host: construction from a 128-bit
seed, the JSON state (pcg64.rs:57-77), argument checks.  No device work.

The initial state is pinned by NumPy's PCG64DXSM: Pcg64Dxsm::new(seed) sets
inc = ((seed + 0xda3e...4f5a) << 1) | 1 and takes one LCG step from
seed + inc (pcg64.rs:23-30); NumPy driven from that (state, inc) reproduces
the reference's first draws (for seed 42 the known stream 4165707071428774973,
13625534677717622090, ...), and the oracle's sequential RNG agrees.
"""
import json

import numpy as np
import pytest

import stocha_amd as S


def _numpy_draws(state, inc, n):
    bg = np.random.PCG64DXSM()
    bg.state = {"bit_generator": "PCG64DXSM", "state": {"state": state, "inc": inc},
                "has_uint32": 0, "uinteger": 0}
    return bg.random_raw(n)


@pytest.mark.parametrize("seed", [0, 42, 2**64 - 1, 2**64, 2**127 + 5, 2**128 - 1])
def test_initial_state_matches_numpy_and_oracle(orc, seed):
    r = S.RNG(seed)
    st, inc = r._state, r._inc
    o = orc.Rng(seed)
    assert (st, inc) == o.state()
    raw = _numpy_draws(st, inc, 4)
    if seed == 42:
        assert list(raw[:2]) == [4165707071428774973, 13625534677717622090]
    u = o.uniform(4)
    np.testing.assert_array_equal(u, (raw >> np.uint64(11)).astype(np.float64) * 2.0 ** -53)


def test_json_state_format_and_roundtrip():
    r = S.RNG(123)
    r._state = 987654321987654321987654321  # a mid-stream position
    s = r.save_state()
    obj = json.loads(s)
    assert list(obj) == ["version", "rng", "seed"] and obj["version"] == 2
    assert list(obj["rng"]) == ["state", "increment"]
    assert " " not in s
    back = S.RNG.from_state(s)
    assert (back._state, back._inc, back.seed) == (r._state, r._inc, 123)
    legacy = S.RNG.from_state('{"seed":42}')
    fresh = S.RNG(42)
    assert (legacy._state, legacy._inc, legacy.seed) == (fresh._state, fresh._inc, 42)
    assert repr(fresh) == "RNG(seed=42, algorithm='pcg64dxsm')"


@pytest.mark.parametrize("bad", ["not-valid-json", "[]", '{"seed": -1}', '{"seed": 1.5}',
                                 '{"seed": true}', '{"version": 2}',
                                 '{"seed": 340282366920938463463374607431768211456}'])
def test_from_state_rejects(bad):
    with pytest.raises(ValueError):
        S.RNG.from_state(bad)


def test_rng_arguments_like_pyo3():
    with pytest.raises(OverflowError):
        S.RNG(-1)
    with pytest.raises(OverflowError):
        S.RNG(2**128)
    with pytest.raises(TypeError):
        S.RNG(1.0)
    with pytest.raises(ValueError, match="Unknown algorithm: 'mt19937'"):
        S.RNG(0, algorithm="mt19937")
    assert S.RNG(0, algorithm="default").seed == 0
    with pytest.raises(ValueError, match="scale must be positive"):
        S.RNG(0).normal(size=10, scale=-1.0)
    with pytest.raises(OverflowError):
        S.RNG(0).uniform(size=-1)
    # empty draws need no device and leave the position alone
    r = S.RNG(3)
    st = r._state
    assert r.uniform(0).shape == (0,) and r.standard_normal(0).shape == (0,)
    assert r._state == st
