"""Statistical comparison of two transport runs (GPU vs oracle).

Yields are not binomial (a primary can emit many SEs), so each run reports
per-pixel sums of n and n^2 over primaries and the standard error of a pixel
yield is sqrt((<n^2> - <n>^2) / N).  Histograms are compared with the
two-sample Kolmogorov-Smirnov statistic on their cumulative counts and a
chi-square homogeneity test."""
import numpy as np
from scipy import stats


def pixel_yields(back, se, se_sq, bse_sq, per):
    back = np.asarray(back, np.float64)
    se = np.asarray(se, np.float64)
    ys = se / per
    yb = (back - se) / per
    vs = np.maximum(np.asarray(se_sq) / per - ys ** 2, 1e-300) / per
    vb = np.maximum(np.asarray(bse_sq) / per - yb ** 2, 1e-300) / per
    return ys, yb, vs, vb


def from_gpu(r):
    return pixel_yields(r.back, r.se, r.seSq, r.bseSq, r.perPixel)


def from_oracle(o):
    return pixel_yields(o["back"], o["se"], o["seSq"], o["bseSq"], o["perPixel"])


def z_scores(a, b):
    ys1, yb1, vs1, vb1 = a
    ys2, yb2, vs2, vb2 = b
    zs = (ys1 - ys2) / np.sqrt(vs1 + vs2)
    zb = (yb1 - yb2) / np.sqrt(vb1 + vb2)
    return zs, zb


def total_z(a, b):
    """z of the all-pixel mean yields."""
    out = []
    for k in (0, 1):
        m1, m2 = a[k].mean(), b[k].mean()
        v1 = a[k + 2].sum() / len(a[k]) ** 2
        v2 = b[k + 2].sum() / len(b[k]) ** 2
        out.append((m1 - m2) / np.sqrt(v1 + v2))
    return out


def ks_binned(h1, h2):
    """Two-sample KS on binned counts; returns (D, p) with the asymptotic
    distribution for sample sizes n1, n2 (conservative for binned data)."""
    h1 = np.asarray(h1, np.float64)
    h2 = np.asarray(h2, np.float64)
    n1, n2 = h1.sum(), h2.sum()
    if n1 == 0 or n2 == 0:
        return 0.0, 1.0
    D = np.max(np.abs(np.cumsum(h1) / n1 - np.cumsum(h2) / n2))
    en = np.sqrt(n1 * n2 / (n1 + n2))
    p = stats.kstwobign.sf(D * en)
    return float(D), float(p)


def chi2_homogeneity(h1, h2, min_count=5):
    h1 = np.asarray(h1, np.float64)
    h2 = np.asarray(h2, np.float64)
    keep = (h1 + h2) >= min_count
    t = np.stack([h1[keep], h2[keep]])
    if t.shape[1] < 2:
        return 1.0
    return float(stats.chi2_contingency(t, correction=False)[1])


def chi2_pixels(z):
    """p-value of sum z^2 ~ chi2(len(z))."""
    z = np.asarray(z)
    return float(stats.chi2.sf(np.sum(z * z), len(z)))
