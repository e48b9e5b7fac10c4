"""Oracle semantic RNG against the reference's own known answers.

Pins: Philox4x32-10 KATs and the frozen address vector of
lib/CorePotts/test/test_rng_semantics.jl:15-37; transform domains (:59-113).
"""
import numpy as np
import pytest

from oracle import bridge


def test_philox_known_answers():
    # test_rng_semantics.jl:15-25
    assert bridge.philox((0, 0, 0, 0), (0, 0)) == (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)
    assert bridge.philox((0xffffffff,) * 4, (0xffffffff,) * 2) == (
        0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)
    assert bridge.philox((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344), (0xa4093822, 0x299f31d0)) == (
        0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1)


def test_address_packing_frozen_vector():
    # test_rng_semantics.jl:27-37: ProposalRecipientStream = 3, SiteEntity = 1
    seed = 0x0123456789abcdef
    c0, c1, key = bridge.counter_key(seed, 3, mcs=17, subround=2, operation=33, kind=1, entity=91,
                                     invocation=4, draw=7)
    assert c0 == 0x0402000000000011
    assert c1 == 0x01d030210000005b
    assert key == 0x5b145ecefa74362b
    assert bridge.rng_words(seed, 3, mcs=17, subround=2, operation=33, kind=1, entity=91,
                            invocation=4, draw=7) == (0x5b00c369, 0x92f3ce04, 0x12df44ce, 0x8daeafe9)


def test_generation_changes_key():
    seed = 0x0123456789abcdef
    a = bridge.rng_words(seed, 11, mcs=4, operation=2, kind=2, entity=8, generation=3)
    b = bridge.rng_words(seed, 11, mcs=4, operation=2, kind=2, entity=8, generation=2)
    assert a != b


def test_uniform_open01_rounding():
    lib = bridge.load()
    # (Float32(w >> 8) + 0.5f0) * 2^-24: exact below 2^23, ties-to-even above
    assert lib.orc_uniform_open01_f32(0) == np.float32(0.5) * np.float32(2.0 ** -24)
    top = lib.orc_uniform_open01_f32(0xffffffff)
    assert top == np.float32(1.0)  # 2^24 - 1 + 0.5 rounds to 2^24 in Float32
    for w in (0x12345678, 0x7fffff00, 0x80000000, 0xdeadbeef):
        expected = (np.float32(w >> 8) + np.float32(0.5)) * np.float32(2.0 ** -24)
        assert lib.orc_uniform_open01_f32(w) == expected


def test_bounded_uint_domain_and_retry():
    lib = bridge.load()
    seed = 77
    for bound in (1, 2, 3, 4, 7, 9, 26, 1000003):
        for entity in range(50):
            v = lib.orc_bounded_uint(seed, 5, 9, 0, 1, 1, entity, 0, bound)
            assert 0 <= v < bound
    # bounded_uint equals word 1 mod bound whenever word 1 clears the rejection threshold
    bound = 7
    thr = (2 ** 32 - bound) % bound
    for entity in range(200):
        w = bridge.rng_words(seed, 4, mcs=3, kind=1, entity=entity)[0]
        v = lib.orc_bounded_uint(seed, 4, 3, 0, 0, 1, entity, 0, bound)
        if w >= thr:
            assert v == w % bound


def test_bounded_uint_uniformity():
    lib = bridge.load()
    counts = np.zeros(6, int)
    for entity in range(30000):
        counts[lib.orc_bounded_uint(1234, 4, 1, 0, 0, 1, entity, 0, 6)] += 1
    expected = 30000 / 6
    chi2 = ((counts - expected) ** 2 / expected).sum()
    assert chi2 < 25  # 5 dof, p ~ 1e-4


def test_small_permutation_is_permutation():
    # checkerboard colour order: Fisher-Yates over colour ids with CheckerboardOrderStream
    import ctypes
    lib = bridge.load()
    seen = set()
    for mcs in range(1, 40):
        order = list(range(1, 5))
        for c in range(4, 1, -1):
            sel = lib.orc_bounded_uint(0xC0102, 8, mcs, 0, 0, 0, 0, 4 - c, c) + 1
            order[c - 1], order[sel - 1] = order[sel - 1], order[c - 1]
        assert sorted(order) == [1, 2, 3, 4]
        seen.add(tuple(order))
    assert len(seen) > 1
