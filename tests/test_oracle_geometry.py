"""Geometry of the oracle: region location and first boundary crossings.

- RegionBase::findEndOfStep through the CSG walks of NormalIntersectionShape /
  NormalUnionShape is checked by brute force: dense point location
  (containingSubRegion) along the segment finds the first region change;
- NShapes::createLine (fixed mode) is checked against an independent
  construction of a trapezoid whose top corners are replaced by arcs
  (triangle-minus-disk), and the compat mode against the line without a left
  sidewall that NShapes.cu:140 produces;
- the reference tests' invariant t + t' = 1 for a segment and its reverse
  (CylindricalShapeTest.cu:106-417, SphereTest.cu:51-66)."""
import math

import numpy as np
import pytest

H, W, TH, R = 50.0, 20.0, 3.0, 5.0


@pytest.fixture(scope="module", params=[True, False], ids=["compat", "fixed"])
def line_scene(request, oracle):
    from cudamonsel_amd import scenes
    sc = scenes.line_on_substrate(h_nm=H, w_nm=W, theta_deg=TH, r_nm=R, compat=request.param)
    cs = sc.monte.compileScene()
    return request.param, sc, cs, oracle.OracleScene(cs)


def rounded_trapezoid(x, z, compat):
    """Inside test for the line cross-section (|y| < L/2), built from the
    geometry: sidewalls at angle TH, top corners rounded by arcs of radius R
    tangent to the top face and the sidewall (triangle-minus-disk).  With the
    reference's defect the left part is the right-hand half-space again, so
    (A u B) n (A u C) = A u (B n C): the right corner is cut at the chord
    between the tangent points and only the lens of the two corner disks is
    added back."""
    t = math.radians(TH)
    inside = (z >= -H) & (z <= 0)
    xc = W / 2 - R / math.cos(t) + (R - H) * math.tan(t)
    zc = -H + R

    def cross(a, b, px, pz):
        return (b[0] - a[0]) * (pz - a[1]) - (b[1] - a[1]) * (px - a[0])

    def parts(xx):
        side = (xx - W / 2) * math.cos(t) - z * math.sin(t) <= 0
        V = np.array([W / 2 - H * math.tan(t), -H])
        T1 = np.array([xc, -H])
        T2 = np.array([xc + R * math.cos(t), zc - R * math.sin(t)])
        c1, c2, c3 = cross(V, T1, xx, z), cross(T1, T2, xx, z), cross(T2, V, xx, z)
        in_tri = ((c1 >= 0) & (c2 >= 0) & (c3 >= 0)) | ((c1 <= 0) & (c2 <= 0) & (c3 <= 0))
        in_disk = (xx - xc) ** 2 + (z - zc) ** 2 <= R * R
        same_side_as_center = np.sign(cross(T1, T2, xx, z)) == np.sign(cross(T1, T2, xc, zc))
        return side, in_tri, in_disk, same_side_as_center

    side, in_tri, in_disk, below_chord = parts(x)
    if compat:
        lens = in_disk & ((x + xc) ** 2 + (z - zc) ** 2 <= R * R)
        return inside & ((side & below_chord) | lens)
    lside, lin_tri, lin_disk, _ = parts(-x)
    return inside & side & ~(in_tri & ~in_disk) & lside & ~(lin_tri & ~lin_disk)


def test_create_line_cross_section(line_scene):
    compat, sc, cs, osc = line_scene
    rng = np.random.default_rng(5)
    n = 60000
    x = rng.uniform(-20, 20, n)
    z = rng.uniform(-55, 5, n)
    pts = np.stack([x, np.zeros(n), z], axis=1)
    # keep away from the exact boundary (sampling margin)
    reg = osc.locate(pts)
    line_idx = cs.region_index(sc.regions["line0"])
    sub_idx = cs.region_index(sc.regions["substrate"])
    want = rounded_trapezoid(x, z, compat)
    got_line = reg == line_idx
    eps = 1e-6
    near = np.zeros(n, bool)
    for dx, dz in ((eps, 0), (-eps, 0), (0, eps), (0, -eps)):
        near |= rounded_trapezoid(x + dx, z + dz, compat) != want
    ok = (got_line == want) | near
    assert ok.all(), np.stack([x, z], 1)[~ok][:5]
    assert np.all(reg[(z > 0) & ~got_line] == sub_idx)


def brute_first_change(osc, p0, p1, n=4096):
    t = (np.arange(1, n + 1) / n)[:, None]
    pts = p0[None, :] + t * (p1 - p0)[None, :]
    reg = osc.locate(pts)
    r0 = osc.locate(p0[None, :])[0]
    ch = np.nonzero(reg != r0)[0]
    if ch.size == 0:
        return None, r0, r0
    return (ch[0] + 1) / n, r0, reg[ch[0]]


def test_find_end_of_step_matches_brute_force(line_scene):
    compat, sc, cs, osc = line_scene
    rng = np.random.default_rng(6 if compat else 7)
    checked = 0
    for _ in range(250):
        p0 = np.array([rng.uniform(-25, 25), rng.uniform(-30, 30), rng.uniform(-60, 10)])
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        p1 = p0 + d * rng.uniform(1, 40)
        ub, r0, rb = brute_first_change(osc, p0, p1)
        if r0 < 0:
            continue
        u, nxt, _ = osc.find_end_of_step(r0, p0[None, :], p1[None, :])
        if ub is None:
            assert u[0] > 1.0 or nxt[0] == r0
            continue
        assert u[0] <= 1.0
        assert abs(u[0] - ub) <= 1.0 / 4096 + 1e-9
        assert nxt[0] == rb
        checked += 1
    assert checked > 50


def test_segment_and_reverse_sum_to_one(line_scene):
    compat, sc, cs, osc = line_scene
    rng = np.random.default_rng(8)
    done = 0
    for _ in range(400):
        p0 = np.array([rng.uniform(-25, 25), rng.uniform(-30, 30), rng.uniform(-60, 10)])
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        p1 = p0 + d * rng.uniform(1, 30)
        r0 = osc.locate(p0[None, :])[0]
        r1 = osc.locate(p1[None, :])[0]
        if r0 < 0 or r1 < 0 or r0 == r1:
            continue
        ub, _, rb = brute_first_change(osc, p0, p1, 1024)
        ubr, _, rbr = brute_first_change(osc, p1, p0, 1024)
        if ub is None or ubr is None or rb != r1 or rbr != r0 or abs(ub + ubr - 1.0) > 3.0 / 1024:
            continue  # more than one crossing
        u, _, _ = osc.find_end_of_step(r0, p0[None, :], p1[None, :])
        ur, _, _ = osc.find_end_of_step(r1, p1[None, :], p0[None, :])
        assert u[0] + ur[0] == pytest.approx(1.0, abs=1e-8)
        done += 1
    assert done > 20


def test_chamber_sphere_reverse_sum(oracle):
    from cudamonsel_amd import scenes
    sc = scenes.flat("PMMA")
    cs = sc.monte.compileScene()
    osc = oracle.OracleScene(cs)
    rng = np.random.default_rng(9)
    for _ in range(20):
        d = rng.normal(size=3)
        d[2] = -abs(d[2])
        d /= np.linalg.norm(d)
        p0 = np.array([0.0, 0.0, -10.0])
        p1 = p0 + d * 2e8  # 0.2 m, past the 0.1 m chamber wall
        u, nxt, _ = osc.find_end_of_step(0, p0[None, :], p1[None, :])
        assert nxt[0] == -1
        hit = p0 + u[0] * (p1 - p0)
        assert np.linalg.norm(hit) == pytest.approx(1e8, rel=1e-9)
