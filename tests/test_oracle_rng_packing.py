"""Pins of the oracle's integer and format code: the RNG (PM:56-65, PM:251-253, DS:332-334) and
the packed depth format (PK:3-33, depth-maps.ts:44-85)."""
import json
import os

import numpy as np
import pytest

import oracle as O
from lr3d import packing

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "pcg_vectors.json")


@pytest.fixture(scope="module")
def golden():
    return json.load(open(GOLDEN))


def test_pcg_matches_published_hash_vectors(golden):
    for x, y in zip(golden["inputs"], golden["pcg"]):
        assert O.pcg(x) == y, x


def test_pixel_streams_match_vectors(golden):
    for s in golden["streams"]:
        state = O.pixel_rng(s["seed"], s["iteration"], s["phase"], s["index"])
        draws = O.random_stream(state, len(s["draws_24bit"]))
        assert [round(d * 2**24) for d in draws] == s["draws_24bit"]
        assert all(0.0 <= d < 1.0 for d in draws)


def test_pcg_is_a_permutation_on_a_block():
    # the output function is a bijection of u32: no collisions on 2^16 consecutive inputs
    outs = {O.pcg(v) for v in range(1 << 16)}
    assert len(outs) == 1 << 16


def _js_imul(a, b):
    def to_i32(x):
        x &= 0xFFFFFFFF
        return x - (1 << 32) if x & 0x80000000 else x
    return to_i32(to_i32(a) * to_i32(b))


@pytest.mark.parametrize("ref", [0, 1, 7, 31, 199, 70000])
@pytest.mark.parametrize("it,phase,flags", [(0, 0, 0), (3, 1, 0), (7, 1, 1), (12, 0, 1)])
def test_host_seed_matches_javascript_semantics(ref, it, phase, flags):
    # DS:333-334: Math.imul(ref + 1, 0x9e3779b1) + it*97 + phase*13 + flags*7, then >>> 0
    js = (_js_imul(ref + 1, 0x9E3779B1) + it * 97 + phase * 13 + flags * 7) % (1 << 32)
    assert O.host_seed(ref, it, phase, flags) == js


def test_random_is_uniform_with_24_bit_resolution():
    draws = np.array(O.random_stream(12345, 20000))
    assert np.all((draws * 2**24) == np.round(draws * 2**24))
    hist, _ = np.histogram(draws, bins=20, range=(0, 1))
    expected = len(draws) / 20
    chi2 = ((hist - expected) ** 2 / expected).sum()
    assert chi2 < 45  # 19 dof, p < 1e-3


def test_f16_conversion_matches_ieee_round_to_nearest_even():
    rng = np.random.default_rng(0)
    vals = np.concatenate([
        rng.uniform(0.01, 10, 4000), rng.uniform(-70000, 70000, 2000), rng.uniform(-1e-5, 1e-5, 2000),
        np.array([0.0, -0.0, 65504.0, 65520.0, 1e9, 5.96e-8, 2.98e-8, 1.0 + 2**-11, 1.0 + 3 * 2**-11,
                  np.inf, -np.inf]),
    ]).astype(np.float32)
    # exact ties: midpoints between consecutive halves
    halves = rng.integers(0x0400, 0x7BFF, 1000).astype(np.uint16)
    lo = halves.view(np.float16).astype(np.float64)
    hi = (halves + 1).view(np.float16).astype(np.float64)
    vals = np.concatenate([vals, ((lo + hi) / 2).astype(np.float32)])
    with np.errstate(over="ignore"):  # values past 65504 round to infinity, as intended
        for v in vals:
            assert O.f32_to_f16(float(v)) == int(np.float16(v).view(np.uint16)), v


def test_f16_decode_matches_numpy_for_every_code():
    codes = np.arange(65536, dtype=np.uint32)
    ref = codes.astype(np.uint16).view(np.float16).astype(np.float32)
    got = np.array([O.f16_to_f32(int(c)) for c in codes], dtype=np.float32)
    nan = np.isnan(ref)
    assert np.array_equal(np.isnan(got), nan)
    assert np.array_equal(got[~nan].view(np.uint32), ref[~nan].view(np.uint32))
    # the host decode (depth-maps.ts halfToFloat) agrees
    assert np.array_equal(packing.half_to_float(codes)[~nan], ref[~nan])


def test_octahedral_normals_round_trip_every_code():
    worst = 0.0
    stable = 0
    for qx in range(64):
        for qy in range(64):
            code = (qx << 16) | (qy << 22) | 0x3C00  # depth 1.0
            n = O.unpack_normal(code)
            assert abs(np.linalg.norm(n) - 1) < 1e-6
            again = O.pack_sample(1.0, n, 0.0)
            stable += again == code
            m = O.unpack_normal(again)
            worst = max(worst, np.degrees(np.arccos(np.clip(np.dot(n, m), -1, 1))))
            # the float64 host decode agrees with the kernel decode
            assert np.allclose(packing.unpack_normal(code), n, atol=2e-6)
    assert stable >= 4096 * 0.96  # 3969 codes are fixed points; the rest sit on the fold edges
    assert worst < 3.0


def test_pack_quantisation_of_random_normals():
    rng = np.random.default_rng(3)
    v = rng.standard_normal((2000, 3))
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    errs = []
    for n in v.astype(np.float32):
        m = O.unpack_normal(O.pack_sample(1.5, n, 0.3))
        errs.append(np.degrees(np.arccos(np.clip(np.dot(n, m), -1, 1))))
    assert max(errs) < 4.0 and np.mean(errs) < 1.5  # 6-bit octahedral cells span ~3.5 deg at most


@pytest.mark.parametrize("cost,expected", [(0.0, 0), (0.0666, 0), (0.0667, 1), (0.5, 7), (0.9999, 14)])
def test_cost_quantisation(cost, expected):
    p = O.pack_sample(1.0, [0.0, 0.0, -1.0], cost)
    assert O.unpack_cost(p) == expected == min(15, int(np.float32(cost) * np.float32(15)))


@pytest.mark.parametrize("depth,cost", [(0.0, 0.1), (-1.0, 0.1), (float("nan"), 0.1), (1.0, 1.0), (1.0, 2.0)])
def test_invalid_samples_pack_to_zero(depth, cost):
    assert O.pack_sample(depth, [0.0, 0.0, -1.0], cost) == 0


def test_depth_survives_packing_to_half_precision():
    for d in (0.3, 0.5, 1.0, 1.234567, 2.999, 4.0):
        p = O.pack_sample(d, [0.1, -0.2, -0.97], 0.2)
        assert abs(O.unpack_depth(p) - d) <= d * 2**-11
