"""Histogram binning (Histogram.cu:73-79 with Algorithm::binarySearch,
Algorithm.cu:5-14).  The reference's binarySearch returns the insertion point
where Java's returns -(insertion point)-1, so Histogram::bin files a value
between two edges one bin high.  Fixed mode follows the Java contract:
bin k holds [min + k d, min + (k+1) d), -1 below, nbins at or above max."""
import pytest


@pytest.mark.parametrize("val,fixed,compat", [
    (35.0, 3, 4),     # between edges 30 and 40
    (40.0, 4, 4),     # exactly on an edge: found by the search in both modes
    (0.0, 0, 0),
    (-1.0, -1, 0),    # below the range: underflow vs first bin
    (0.5, 0, 1),
    (499.9, 49, 50),  # last bin vs over-range slot
    (500.0, 50, 50),  # == max
    (600.0, 50, 50),  # above max: the reference writes past the array; clamped
])
def test_bin_semantics(oracle, val, fixed, compat):
    assert oracle.hist_bin(val, 0.0, 500.0, 50, False) == fixed
    assert oracle.hist_bin(val, 0.0, 500.0, 50, True) == compat


def test_se_window_of_the_line_scan_driver(oracle):
    """LinesOnLayers.cu:368-373 sums bins j < (int)(50 / (E0 / nbins)): with
    10 eV bins that is E < 50 eV in fixed mode but E <= 40 eV in compat mode."""
    maxSEbin = int(50.0 / (500.0 / 50))
    se = lambda e, c: 0 <= oracle.hist_bin(e, 0.0, 500.0, 50, c) < maxSEbin
    assert se(45.0, False) and not se(45.0, True)
    assert se(40.0, False) and se(40.0, True)
    assert se(39.99, True) and not se(40.01, True)
