"""Bit-exact checks of the device's integer and index work through the C-ABI
probes: the Philox4x32-10 generator behind every random draw (Random123
known-answer vectors, then random counters and keys against the oracle's
independent implementation), the FP32 uniforms made of its words, the
histogram bin index of the escape tallies (Histogram.cu:73-79 with
Algorithm.cu:5-14, both modes) against the oracle's, and the Gaussian-beam
start points of a job's primaries (GaussianBeam.cu:52-65), which must follow
the Rayleigh law of the beam width and equal the formula evaluated on the
same stream words."""
import numpy as np
import pytest
from scipy import stats

pytestmark = pytest.mark.gpu

KEY = (0x6D6F6E73, 0x656C3235)   # the transport kernel's fixed Philox key
PRIMARY_BLOCK = 0xFFFFFFFF       # counter block of a primary's start point
M64 = (1 << 64) - 1


def _ctx(M):
    return M.Device.get(0).ctx_address()


def _splitmix64(z):
    z = (z + 0x9E3779B97F4A7C15) & M64
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
    return z ^ (z >> 31)


def test_philox_known_answers(M):
    from cudamonsel_amd import abi
    kat = [
        ((0, 0, 0, 0), (0, 0), (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
        ((0xffffffff,) * 4, (0xffffffff, 0xffffffff), (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
        ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344), (0xa4093822, 0x299f31d0),
         (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1)),
    ]
    out, _ = abi.philox(_ctx(M), [c for c, _, _ in kat], [k for _, k, _ in kat])
    for row, (_, _, want) in zip(out, kat):
        assert tuple(int(x) for x in row) == want


def test_philox_and_uniforms_match_oracle(oracle, M):
    from cudamonsel_amd import abi
    rng = np.random.default_rng(1)
    n = 4096
    ctr = rng.integers(0, 1 << 32, (n, 4), dtype=np.uint64).astype(np.uint32)
    key = rng.integers(0, 1 << 32, (n, 2), dtype=np.uint64).astype(np.uint32)
    ctr[:8, 0] = [0, 1, 2, 3, 0xFFFFFFFE, 0xFFFFFFFF, 0x80000000, 0x7FFFFFFF]
    out, u = abi.philox(_ctx(M), ctr, key)
    want = np.array([oracle.philox(tuple(int(x) for x in c), tuple(int(x) for x in k)) for c, k in zip(ctr, key)],
                    dtype=np.uint32)
    assert np.array_equal(out, want)
    # u = (x >> 9) 2^-23 + 2^-24: exact in FP32, open interval (0, 1)
    expect = ((want >> 9).astype(np.float64) * 2.0 ** -23 + 2.0 ** -24).astype(np.float32)
    assert np.array_equal(u, expect)
    assert u.min() > 0.0 and u.max() < 1.0


@pytest.mark.parametrize("compat", [False, True], ids=["fixed", "compat"])
@pytest.mark.parametrize("lo,hi,nbins", [(0.0, 500.0, 50), (0.0, 1000.0, 200), (0.0, np.pi, 180),
                                         (0.0, 2 * np.pi, 360), (0.0, 377.0, 37)])
def test_hist_bin_matches_oracle(oracle, M, compat, lo, hi, nbins):
    from cudamonsel_amd import abi
    rng = np.random.default_rng(nbins)
    d = (hi - lo) / nbins
    edges = lo + np.arange(nbins + 1) * d
    vals = np.concatenate([
        edges,                                    # exactly on the edges
        np.nextafter(edges, -np.inf), np.nextafter(edges, np.inf),
        rng.uniform(lo, hi, 5000),                # interior
        [lo - 1.0, -1e-300, hi * 1.5, hi + d, 1e300, -1e300, 0.0],  # under / over range
    ])
    g = abi.hist_bin(_ctx(M), vals, lo, hi, nbins, compat)
    o = np.array([oracle.hist_bin(float(v), lo, hi, nbins, compat) for v in vals])
    assert np.array_equal(g, o), vals[g != o][:5]


def test_primaries_follow_the_beam(oracle, M):
    """Start points of the device's primaries: the Rayleigh radius and uniform
    azimuth of the beam, pixel order, and the value of the createElectron
    formula on the same Philox words (FP32: within 1e-4 nm)."""
    from cudamonsel_amd import scenes
    from cudamonsel_amd.abi import probe_for
    sc = scenes.flat("PMMA", beam_eV=500.0)
    pr, cs = probe_for(sc)
    w = 0.5e-9
    centers = np.array([[x * 1e-9, 2e-9, sc.beam_z] for x in (-40.0, 0.0, 25.0, 63.0)])
    per = 50000
    seed = 2024
    n = per * centers.shape[0]
    pos, reg = pr.primaries(centers, per, 500.0, w, seed, n)
    assert np.all(reg == 0)
    cen_nm = np.repeat(centers * 1e9, per, 0)
    d = pos.astype(np.float64) - cen_nm
    assert np.max(np.abs(d[:, 2])) < 1e-4
    r = np.hypot(d[:, 0], d[:, 1]) / (w * 1e9)
    assert stats.kstest(r, stats.rayleigh.cdf).pvalue > 1e-3
    th = np.mod(np.arctan2(d[:, 1], d[:, 0]), 2 * np.pi) / (2 * np.pi)
    assert stats.kstest(th, "uniform").pvalue > 1e-3
    # the formula on the stream words: counter (block, lineage lo/hi, primary)
    lin = _splitmix64(seed ^ 0x5EED5EED5EED5EED)
    k = np.arange(0, n, 997)
    words = np.array([oracle.philox((PRIMARY_BLOCK, lin & 0xFFFFFFFF, lin >> 32, int(i)), KEY) for i in k],
                     dtype=np.uint64)
    u1 = (words[:, 0] >> 9) * 2.0 ** -23 + 2.0 ** -24
    u2 = (words[:, 1] >> 9) * 2.0 ** -23 + 2.0 ** -24
    rr = np.sqrt(-2.0 * np.log(u1)) * w * 1e9
    x = cen_nm[k, 0] + rr * np.cos(2 * np.pi * u2)
    y = cen_nm[k, 1] + rr * np.sin(2 * np.pi * u2)
    assert np.max(np.abs(pos[k, 0] - x)) < 1e-4 and np.max(np.abs(pos[k, 1] - y)) < 1e-4
    # a job split by primary_offset starts the same primaries at the same points
    p2, _ = pr.primaries(centers[1:], per, 500.0, w, seed, 1000, primary_offset=per)
    assert np.array_equal(p2, pos[per:per + 1000])
