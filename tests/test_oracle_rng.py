"""Random streams of the oracle: Philox4x32-10 against the Random123 known-answer
vectors, the MSVC rand() emulation against the values the reference's own run
log prints (outputs/output_release.txt:3-12), and the stream layout."""
import ctypes
import json
import os

import numpy as np

GOLD = os.path.join(os.path.dirname(__file__), "golden", "reference_run_log.json")


def test_philox_known_answers(oracle):
    # Random123 kat_vectors, philox4x32 with 10 rounds
    kat = [
        ((0, 0, 0, 0), (0, 0), (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
        ((0xffffffff,) * 4, (0xffffffff, 0xffffffff), (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
        ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344), (0xa4093822, 0x299f31d0),
         (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1)),
    ]
    for ctr, key, want in kat:
        assert tuple(oracle.philox(ctr, key)) == want


def test_msvc_rand_matches_reference_log(oracle):
    g = json.load(open(GOLD))
    out = np.zeros(10)
    oracle.lib().oracle_lcg(ctypes.c_uint32(g["msvc_seed"]), 10, oracle.dptr(out))
    # the log prints std::to_string (6 decimals)
    assert np.allclose(out, g["msvc_first_values"], atol=5e-7)


def test_stream_uniforms_open_interval_and_fp32_exact(oracle):
    u = oracle.stream(12345, 0, 7, 4096)
    assert u.min() > 0.0 and u.max() < 1.0
    # every value is an odd multiple of 2^-24, hence exact in FP32
    k = u * 2.0 ** 24
    assert np.all(k == np.round(k)) and np.all(np.round(k) % 2 == 1)
    assert np.all(u.astype(np.float32).astype(np.float64) == u)
    # mean / variance of U(0,1) within 5 sigma
    assert abs(u.mean() - 0.5) < 5 * np.sqrt(1 / 12 / u.size)


def test_streams_depend_on_lineage_primary_seed(oracle):
    a = oracle.stream(1, 0, 0, 64)
    assert not np.array_equal(a, oracle.stream(1, 0, 1, 64))
    assert not np.array_equal(a, oracle.stream(2, 0, 0, 64))
    child = oracle.lib().oracle_child_lineage(0, 0)
    assert child != 0
    assert not np.array_equal(a, oracle.stream(1, child, 0, 64))
    kids = {oracle.lib().oracle_child_lineage(0, k) for k in range(1000)}
    assert len(kids) == 1000


def test_stream_blocks_are_consecutive_philox_counters(oracle):
    # draw 4b+i is word i of Philox(counter=(b, lineage_lo, lineage_hi, primary))
    seed, lin, prim = 0x1234_5678_9abc_def0, 0x0fed_cba9_8765_4321, 77
    u = oracle.stream(seed, lin, prim, 12)
    for b in range(3):
        w = oracle.philox((b, lin & 0xffffffff, lin >> 32, prim), (seed & 0xffffffff, seed >> 32))
        want = [((x >> 9) + 0.5) * 2.0 ** -23 for x in w]
        assert np.allclose(u[4 * b: 4 * b + 4], want, rtol=0, atol=0)
