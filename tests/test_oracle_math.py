"""Pins the oracle's scalar functions to what mathematics fixes: library
transcendentals (error bounds), published hash inverses, closed forms
(Henyey-Greenstein normalisation, the fog integral, Hable's curve, slab
intersection) and lattice/continuity properties of the value noise."""
from __future__ import annotations

import math

import numpy as np
import pytest


@pytest.fixture(scope="module")
def O():
    from oracle import oracle

    return oracle.LIB


def ulp_err(got, ref):
    ref32 = np.float32(ref)
    return abs(float(got) - float(ref)) / float(np.spacing(np.abs(ref32)) if ref32 != 0 else np.float32(1e-45))


def test_sin_cos_accuracy(O):
    xs = np.concatenate([np.linspace(-math.pi, math.pi, 4001), np.linspace(-200.0, 200.0, 4001)]).astype(np.float32)
    for x in xs:
        s, c = O.vxo_sin(float(x)), O.vxo_cos(float(x))
        assert abs(s - math.sin(float(x))) <= 4e-7 * max(1.0, abs(float(x)) / math.pi)
        assert abs(c - math.cos(float(x))) <= 4e-7 * max(1.0, abs(float(x)) / math.pi)
    # WGSL bound: absolute error <= 2^-11 on [-pi, pi]; ours is far tighter
    assert O.vxo_sin(0.0) == 0.0 and O.vxo_cos(0.0) == 1.0


def test_exp_log_pow_accuracy(O):
    for x in np.linspace(-80.0, 80.0, 3001, dtype=np.float32):
        assert ulp_err(O.vxo_exp(float(x)), math.exp(float(x))) <= 3.0
    for x in np.geomspace(1e-30, 1e30, 3001).astype(np.float32):
        assert abs(O.vxo_log(float(x)) - math.log(float(x))) <= 2e-7 * max(1.0, abs(math.log(float(x))))
    rng = np.random.default_rng(1)
    for _ in range(3000):
        x = float(np.float32(rng.uniform(1e-4, 4.0)))
        y = float(np.float32(rng.uniform(0.05, 96.0)))
        ref = x ** y
        if 1e-30 < ref < 1e30:
            assert abs(O.vxo_pow(x, y) - ref) <= 1e-5 * ref + 1e-37
    assert O.vxo_pow(0.0, 2.0) == 0.0 and O.vxo_pow(0.7, 1.0) == np.float32(0.7) and O.vxo_pow(5.0, 0.0) == 1.0


def test_hash_u32_is_lowbias32(O):
    """hash_u32 (common.wgsl:597-606) is Wellons' lowbias32; its published
    inverse recovers the input, which pins both multipliers and all shifts."""

    def inverse(x):
        x ^= x >> 16
        x = (x * 0x43021123) & 0xFFFFFFFF
        x ^= (x >> 15) ^ (x >> 30)
        x = (x * 0x1D69E2A5) & 0xFFFFFFFF
        x ^= x >> 16
        return x

    rng = np.random.default_rng(7)
    for x in list(range(64)) + [int(v) for v in rng.integers(0, 2 ** 32, 2000, dtype=np.uint64)]:
        assert inverse(O.vxo_hash_u32(x)) == x
    assert O.vxo_hash_u32(0) == 0


def test_value_noise_lattice_and_range(O):
    rng = np.random.default_rng(3)
    for _ in range(500):
        ix, iy = (int(v) for v in rng.integers(-5000, 5000, 2))
        # at lattice points the bilinear blend returns the corner hash exactly
        assert O.vxo_value_noise(float(ix), float(iy)) == O.vxo_hash12(float(ix), float(iy))
        x, y = float(rng.uniform(-1e3, 1e3)), float(rng.uniform(-1e3, 1e3))
        v = O.vxo_value_noise(x, y)
        assert 0.0 <= v < 1.0
        # continuity across a cell edge (smoothstep weights match at u = 1)
        e = math.floor(x) + 1.0
        a, b = O.vxo_value_noise(float(np.nextafter(np.float32(e), np.float32(-1e9))), y), O.vxo_value_noise(e, y)
        assert abs(a - b) < 1e-5
        f = O.vxo_fbm(x, y)
        assert 0.0 <= f < 0.96875  # sum of octave amplitudes 1/2 + ... + 1/32


def test_phase_mie_is_normalised_hg(O):
    """phase_mie (phase.wgsl:5-10) is Henyey-Greenstein with g = 0.15: its
    integral over the sphere is 1."""
    c = np.linspace(-1.0, 1.0, 20001)
    vals = np.array([O.vxo_phase_mie(float(v)) for v in c])
    integral = 2.0 * math.pi * np.trapezoid(vals, c)
    assert abs(integral - 1.0) < 1e-4
    g = 0.15
    for v in (-1.0, 0.0, 0.5, 1.0):
        ref = (1 - g * g) / (4 * math.pi * (1 + g * g - 2 * g * v) ** 1.5)
        assert abs(O.vxo_phase_mie(v) - ref) < 1e-6


def test_fog_optical_depth_is_the_height_fog_integral(O):
    """fog_optical_depth_with_base (fog.wgsl:31-45) = integral of
    base * exp(-k (y0 + dy t)) dt over [0, t], k = 0.18."""
    import ctypes as C

    rng = np.random.default_rng(5)
    for _ in range(300):
        ro = np.array([0.0, rng.uniform(0.0, 30.0), 0.0], np.float32)
        rd = rng.normal(size=3)
        rd = (rd / np.linalg.norm(rd)).astype(np.float32)
        t = float(rng.uniform(0.1, 120.0))
        base = 0.002
        ts = np.linspace(0.0, t, 4001)
        ref = np.trapezoid(base * np.exp(-0.18 * (ro[1] + rd[1] * ts)), ts)
        got = O.vxo_fog_od(base, ro.ctypes.data, rd.ctypes.data, t)
        assert abs(got - ref) <= 2e-4 * ref + 1e-9


def test_hable_tonemap_closed_form(O):
    A, B, Cc, D, E, F = 0.22, 0.30, 0.10, 0.20, 0.01, 0.30

    def hable(x):
        return (x * (A * x + Cc * B) + D * E) / (x * (A * x + B) + D * F) - E / F

    for x in (0.0, 0.01, 0.18, 0.5, 1.0, 2.0, 6.0, 20.0):
        inp = np.array([x, x, x], np.float32)
        out = np.zeros(3, np.float32)
        O.vxo_tonemap(inp.ctypes.data, out.ctypes.data)
        ref = hable(x) / hable(6.0)
        assert abs(out[0] - ref) <= 2e-6 * max(1.0, abs(ref))
    # white point maps to 1
    inp = np.array([6.0, 6.0, 6.0], np.float32)
    out = np.zeros(3, np.float32)
    O.vxo_tonemap(inp.ctypes.data, out.ctypes.data)
    assert abs(out[0] - 1.0) < 1e-6


def test_intersect_aabb_against_sampling(O):
    rng = np.random.default_rng(11)
    bmin = np.array([1.0, 2.0, 3.0], np.float32)
    bmax = np.array([4.0, 5.5, 6.0], np.float32)
    out = np.zeros(2, np.float32)
    for _ in range(400):
        ro = rng.uniform(-5, 10, 3).astype(np.float32)
        rd = rng.normal(size=3)
        rd = (rd / np.linalg.norm(rd)).astype(np.float32)
        O.vxo_intersect_aabb(ro.ctypes.data, rd.ctypes.data, bmin.ctypes.data, bmax.ctypes.data, out.ctypes.data)
        ts = np.linspace(-40, 40, 80001)
        pts = ro[None, :] + ts[:, None] * rd[None, :]
        inside = np.all((pts >= bmin) & (pts <= bmax), axis=1)
        if inside.any():
            t_in, t_out = ts[inside].min(), ts[inside].max()
            assert out[0] <= out[1]
            assert abs(out[0] - t_in) < 2e-3 and abs(out[1] - t_out) < 2e-3
        else:
            assert out[0] > out[1] or out[1] < -39.99 or out[0] > 39.99
    # axis-parallel ray outside the slab: miss encoded as (1, 0)
    ro = np.array([0.0, 0.0, 0.0], np.float32)
    rd = np.array([1.0, 0.0, 0.0], np.float32)
    O.vxo_intersect_aabb(ro.ctypes.data, rd.ctypes.data, bmin.ctypes.data, bmax.ctypes.data, out.ctypes.data)
    assert out[0] == 1.0 and out[1] == 0.0
