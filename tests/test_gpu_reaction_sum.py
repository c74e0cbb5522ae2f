"""The wall reaction's ordered subtraction, evaluated by one wavefront.

A wall side's reaction is its starting value minus every term in key order,
one rounded subtraction at a time (contact.h:339,374,411-412,428 through
timeloop.h:753-769).  The reaction kernel evaluates a long chain in 64 chunks
from guessed start values and corrects the guesses until every chunk starts
where the one before it ends.  These chains are built to break guesses:
running sums that cross many binades, sums that hover around zero, terms
spanning 600 orders of magnitude, exact ties of round-to-nearest-even,
signed zeros, subnormals, infinities and NaNs.  The wavefront's result, one
device thread's plain loop and the host's loop must agree bit for bit.
"""
import ctypes
import math
import struct

import numpy as np
import pytest

from perigrain_amd import _ffi

pytestmark = pytest.mark.gpu


def _bits(x):
    return struct.unpack("<q", struct.pack("<d", x))[0]


def _same(a, b):
    if math.isnan(a) or math.isnan(b):
        return math.isnan(a) and math.isnan(b)
    return _bits(a) == _bits(b)


def _host(s0, ent, two):
    x = float(s0)
    for row in ent:
        if two:
            x = x - float(row[0])
            x = x - float(row[1])
        else:
            x = x - float(row[2])
    return x


def _run(chains):
    """chains: list of (s0, entries [m][4], two)."""
    off = np.zeros(len(chains) + 1, np.uint64)
    for c, (_, ent, _) in enumerate(chains):
        off[c + 1] = off[c] + len(ent)
    ent = np.concatenate([np.asarray(e, np.float64).reshape(-1, 4) for _, e, _ in chains])
    ent = np.ascontiguousarray(ent)
    two = np.array([1 if t else 0 for _, _, t in chains], np.int32)
    s0 = np.array([s for s, _, _ in chains], np.float64)
    wave = np.zeros(len(chains))
    seq = np.zeros(len(chains))
    dp = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    _ffi.check(_ffi.lib().pg_check_ordered_sum(
        0, len(chains), off.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)),
        two.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), dp(s0), dp(ent), dp(wave), dp(seq)))
    for c, (s, e, t) in enumerate(chains):
        h = _host(s, e, t)
        assert _same(seq[c], h), (c, len(e), t, seq[c], h)
        assert _same(wave[c], h), (c, len(e), t, wave[c], h)


def _ent(terms_a, terms_b=None):
    m = len(terms_a)
    e = np.zeros((m, 4))
    e[:, 0] = terms_a
    e[:, 1] = terms_b if terms_b is not None else terms_a[::-1]
    e[:, 2] = terms_a
    return e


def test_ordered_sum_random_and_growing_chains():
    rng = np.random.default_rng(0)
    chains = []
    for m in (1, 2, 31, 32, 33, 64, 65, 100, 512, 3000):
        for two in (True, False):
            chains.append((0.0, _ent(rng.normal(size=m), rng.normal(size=m)), two))
    # a wall pressed by many nodes: same-sign terms from zero, the running
    # sum crossing a binade every time it doubles
    for m in (512, 976, 5000):
        t = -rng.uniform(1e-5, 5e-5, m)
        chains.append((0.0, _ent(t, -rng.uniform(0, 1e-6, m)), True))
        chains.append((0.0, _ent(t), False))
    # a long chain whose chunks are hundreds of terms
    chains.append((1.0, _ent(rng.normal(size=20000) * 1e-3, rng.normal(size=20000)), True))
    _run(chains)


def test_ordered_sum_near_zero_and_wide_range():
    rng = np.random.default_rng(1)
    chains = []
    # running sums hovering around zero: sign changes and binade changes
    for m in (200, 700):
        t = rng.normal(size=m) * 1e-3
        t -= t.mean()
        chains.append((0.0, _ent(t, -t[::-1] * (1 + 1e-9)), True))
    # terms over 600 orders of magnitude, random signs
    for m in (300, 1024):
        t = rng.choice([-1.0, 1.0], m) * 10.0 ** rng.uniform(-300, 300, m)
        chains.append((0.0, _ent(t, t[::-1] * 0.5), True))
        chains.append((1e10, _ent(t), False))
    # subnormals
    t = rng.normal(size=400) * 1e-310
    chains.append((0.0, _ent(t), True))
    chains.append((5e-324, _ent(t), False))
    _run(chains)


def test_ordered_sum_ties_zeros_and_specials():
    rng = np.random.default_rng(2)
    chains = []
    # exact ties: from 2.0 (ulp 2^-51) terms of half an ulp round to even,
    # so each result depends on the last bit of the running sum
    q = [2.0 ** -52, -(2.0 ** -52), 2.0 ** -51, 3 * 2.0 ** -52, -(2.0 ** -51)]
    for m in (100, 600, 2000):
        t = np.array([q[k] for k in rng.integers(0, len(q), m)])
        u = np.array([q[k] for k in rng.integers(0, len(q), m)])
        chains.append((2.0, _ent(t, u), True))
        chains.append((2.0, _ent(t), False))
    # signed zeros: (-0) - (+0) = -0, (-0) - (-0) = +0
    for s0 in (0.0, -0.0):
        z = rng.choice([0.0, -0.0], 150)
        chains.append((s0, _ent(z, rng.choice([0.0, -0.0], 150)), True))
        chains.append((s0, _ent(np.full(150, 0.0)), False))
    # an infinity, then its opposite (NaN from then on), and a NaN term
    t = rng.normal(size=300)
    t_inf = t.copy()
    t_inf[100] = np.inf
    chains.append((0.0, _ent(t_inf), False))
    t_nan = t.copy()
    t_nan[150] = np.inf
    t_nan[200] = -np.inf
    chains.append((0.0, _ent(t_nan), False))
    t_n = t.copy()
    t_n[250] = np.nan
    chains.append((0.0, _ent(t_n), True))
    _run(chains)


def test_ordered_sum_rejects_bad_offsets():
    off = np.array([0, 4, 2], np.uint64)
    two = np.zeros(2, np.int32)
    z = np.zeros(16)
    dp = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    rc = _ffi.lib().pg_check_ordered_sum(
        0, 2, off.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)),
        two.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), dp(z), dp(z), dp(z), dp(z))
    assert rc != 0
