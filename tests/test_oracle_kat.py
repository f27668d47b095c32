"""Pins the CPU oracle to the reference's own known-answer tests and invariants.

Each test cites the reference test it restates.  The oracle is the restatement of SerialBackend
that the GPU results are compared against, so these run on the CPU.
"""
import math

import numpy as np
import pytest

import _oracle

L1, L2, L3 = (0, 1, 0), (0, 2, 0), (0, 3, 0)
T1, T2, T3 = (0, 0, 1), (0, 0, 2), (0, 0, 3)
L2T1, L1T2 = (0, 2, 1), (0, 1, 2)
S2, S3, T2C, L1T2C = (1, 0, 0), (2, 0, 0), (3, 0, 0), (4, 0, 0)
FULL3 = (5, 3, 0)
EARTH_R = 6.371e6


def mean_of(x, u, edges, op):
    s, c, _ = _oracle.hist(x, u, edges, op=op)
    return s / np.where(c == 0, np.nan, c), s, c


def test_tiny_case_mean():
    """test/test_core_correctness.jl:41-59 and test/baseline_correctness.jl:6-23."""
    x = np.array([[0.0, 1.0], [0.0, 0.0]])
    u = np.array([[1.0, 2.0], [0.0, 0.0]])
    m, s, c = mean_of(x, u, [0.0, 2.0], L2)
    assert m[0] == 1.0 and c[0] == 1


def test_three_point_mean():
    """test/test_core_correctness.jl:61-80: sum 4, count 3, mean 4/3."""
    x = np.array([[0.0, 1.0, 0.0], [0.0, 0.0, 1.0]])
    u = x.copy()
    m, s, c = mean_of(x, u, [0.0, 2.0], L2)
    assert c[0] == 3 and abs(s[0] - 4.0) < 1e-15 and abs(m[0] - 4 / 3) < 1e-15


def test_pair_counts():
    """test/test_core_correctness.jl:23-39."""
    x3 = np.array([[0.0, 1.0, 2.0], [0.0, 0.0, 0.0]])
    _, _, c = mean_of(x3, np.zeros((2, 3)), [0.0, 3.0], S2)
    assert c.sum() == 3
    x4 = np.array([[0.0, 1.0, 2.0, 3.0], [0.0, 0.0, 0.0, 0.0]])
    _, _, c = mean_of(x4, np.zeros((2, 4)), [0.0, 3.0], S2)
    assert c.sum() == 6


def test_transverse_wiring_exact():
    """test/test_core_correctness.jl:83-140: du = (0,1) along rhat = (1,0)."""
    x = np.array([[0.0, 1.0], [0.0, 0.0]])
    u = np.array([[0.0, 0.0], [1.0, 2.0]])
    e = [0.0, 2.0]
    assert mean_of(x, u, e, L2)[0][0] == 0.0
    assert mean_of(x, u, e, T2)[0][0] == 1.0
    assert mean_of(x, u, e, L3)[0][0] == 0.0
    assert mean_of(x, u, e, T3)[0][0] == 1.0     # right-handed nhat = zhat x rhat
    assert mean_of(x, u, e, L1T2)[0][0] == 0.0


def pair_value(du, op, D):
    x = np.zeros((D, 2))
    x[0, 1] = 1.0
    u = np.zeros((D, 2))
    u[:, 1] = du
    return _oracle.pair(x, u, [0.0, 2.0], 0, 1, op=op)[2]


def test_3d_operator_values():
    """test/test_core_correctness.jl:142-156: du = (2,3,4), rhat = xhat."""
    du = [2.0, 3.0, 4.0]
    v = {k: pair_value(du, op, 3) for k, op in
         dict(L2=L2, T2=T2, T2C=T2C, L3=L3, L1T2=L1T2, L1T2C=L1T2C, S2=S2, S3=S3, F3=FULL3).items()}
    assert v["L2"] == pytest.approx(4.0) and v["T2"] == pytest.approx(25.0)
    assert v["T2C"] == pytest.approx(12.5) and v["L3"] == pytest.approx(8.0)
    assert v["L1T2"] == pytest.approx(50.0) and v["L1T2C"] == pytest.approx(25.0)
    assert v["S2"] == pytest.approx(v["L2"] + v["T2"])
    assert v["S3"] == pytest.approx(v["L3"] + v["L1T2"])
    assert v["S3"] != v["F3"] and v["F3"] == pytest.approx(math.sqrt(29.0) ** 3)


def rot_pair(theta, flip=False):
    R = np.array([[math.cos(theta), -math.sin(theta)], [math.sin(theta), math.cos(theta)]])
    if flip:
        R = np.array([[1.0, 0.0], [0.0, -1.0]])
    x = np.zeros((2, 2))
    x[:, 1] = R @ [1.0, 0.0]
    u = np.zeros((2, 2))
    u[:, 1] = R @ [2.0, 3.0]
    return x, u


def test_2d_orientation():
    """test/test_core_correctness.jl:158-177: L2T1 = 12, T3 = 27; rotation-invariant, odd under
    reflection; T2 even."""
    e = [0.0, 2.0]
    x, u = rot_pair(0.0)
    base = {op: _oracle.pair(x, u, e, 0, 1, op=op)[2] for op in (L2T1, T3, T2)}
    assert base[L2T1] == pytest.approx(12.0) and base[T3] == pytest.approx(27.0)
    x, u = rot_pair(0.7)
    for op in (L2T1, T3, T2):
        assert _oracle.pair(x, u, e, 0, 1, op=op)[2] == pytest.approx(base[op], rel=1e-12)
    x, u = rot_pair(0.0, flip=True)
    assert _oracle.pair(x, u, e, 0, 1, op=L2T1)[2] == pytest.approx(-base[L2T1])
    assert _oracle.pair(x, u, e, 0, 1, op=T3)[2] == pytest.approx(-base[T3])
    assert _oracle.pair(x, u, e, 0, 1, op=T2)[2] == pytest.approx(base[T2])


def test_3d_nhat_orientation():
    """test/test_helpers.jl:47-50: rhat = yhat -> nhat = normalize(zhat x yhat) = -xhat."""
    x = np.zeros((3, 2))
    x[1, 1] = 1.0
    u = np.zeros((3, 2))
    u[:, 1] = [1.0, 0.0, 0.0]
    assert _oracle.pair(x, u, [0.0, 2.0], 0, 1, op=T1)[2] == pytest.approx(-1.0)


def test_digitize_half_open():
    """test/test_helpers.jl:7-27: bins (a, b], edges [0, 1, 2, 5]."""
    edges = [0.0, 1.0, 2.0, 5.0]
    want = {0.5: 1, 1.5: 2, 4.0: 3, 0.0: 0, 5.0: 3, 6.0: 4}
    for v, b in want.items():
        x = np.array([[0.0, 1.0], [0.0, 0.0]])
        u = np.array([[0.0, v], [0.0, 0.0]])
        ok, d, val, db, vb = _oracle.pair(x, u, [0.0, 2.0], 0, 1, op=L1, vedges=edges)
        assert val == v and vb == b
    # the distance axis is the same half-open convention: r exactly on an edge goes below it
    x = np.array([[0.0, 1.0, 2.0], [0.0, 0.0, 0.0]])
    s, c, _ = _oracle.hist(x, np.zeros((2, 3)), [0.0, 1.0, 2.0], op=S2)
    assert c.tolist() == [2, 1]


def haversine(lo1, la1, lo2, la2, R):
    import mpmath as mp
    mp.mp.dps = 40
    f1, f2 = mp.radians(la1), mp.radians(la2)
    dl = mp.radians(lo2) - mp.radians(lo1)
    a = mp.sin((f2 - f1) / 2) ** 2 + mp.cos(f1) * mp.cos(f2) * mp.sin(dl / 2) ** 2
    return float(2 * R * mp.asin(mp.sqrt(a)))


def test_spherical_separation_matches_haversine():
    """test/test_spherical_geometry.jl:44-62: sigma*R == Haversine(R) at rtol 1e-12; degrees and
    radians agree."""
    for lo1, la1, lo2, la2 in ((10.0, 45.0, 13.0, 47.0), (-170.0, -33.0, 175.0, 12.0),
                               (0.0, 0.0, 0.0, 90.0), (100.0, 60.0, 100.0, 60.5)):
        x = np.array([[lo1, lo2], [la1, la2]])
        u = np.array([[0.3, -1.2], [0.7, 0.4]])
        okd, rd, vd, _, _ = _oracle.pair(x, u, [0.0, 1e8], 0, 1, op=L1, geometry="deg", radius=EARTH_R)
        okr, rr, vr, _, _ = _oracle.pair(np.radians(x), u, [0.0, 10.0], 0, 1, op=L1, geometry="rad")
        assert okd and okr
        assert rd == pytest.approx(haversine(lo1, la1, lo2, la2, EARTH_R), rel=1e-12)
        assert rd == pytest.approx(EARTH_R * rr, rel=1e-12)
        assert vd == pytest.approx(vr, abs=1e-12)


def rand_sphere(rng, N, lon_span=360.0, lat_span=120.0):
    lon = lon_span * rng.uniform(size=N) - lon_span / 2
    lat = lat_span * rng.uniform(size=N) - lat_span / 2
    return np.stack([lon, lat])


def test_solid_body_rotation():
    """test/test_spherical_geometry.jl:67-90: a rigid rotation has no longitudinal increment."""
    rng = np.random.default_rng(20260817)
    x = rand_sphere(rng, 250)
    om = 7.292e-5
    u = np.stack([om * EARTH_R * np.cos(np.radians(x[1])), np.zeros(250)])
    bins = np.linspace(0.0, 8.0e6, 21)
    sL, cL, _ = _oracle.hist(x, u, bins, op=L2, geometry="deg", radius=EARTH_R)
    sS, cS, _ = _oracle.hist(x, u, bins, op=S2, geometry="deg", radius=EARTH_R)
    occ = cL > 0
    assert occ.any() and sL[occ].sum() / sS[occ].sum() < 1e-24


def test_thin_shell_radial_not_in_longitudinal():
    """test/test_spherical_geometry.jl:92-114."""
    rng = np.random.default_rng(5)
    x = rand_sphere(rng, 200, 300.0, 100.0)
    ue = 7.292e-5 * EARTH_R * np.cos(np.radians(x[1]))
    u2 = np.stack([ue, np.zeros(200)])
    u3 = np.stack([ue, np.zeros(200), 3.0 * np.sin(np.radians(2 * x[1]))])
    bins = np.linspace(0.0, 8.0e6, 21)
    kw = dict(geometry="deg", radius=EARTH_R)
    s2, c2, _ = _oracle.hist(x, u2, bins, op=L2, **kw)
    s3, c3, _ = _oracle.hist(x, u3, bins, op=L2, **kw)
    assert np.array_equal(s2, s3) and np.array_equal(c2, c3)
    S2_2 = _oracle.hist(x, u2, bins, op=S2, **kw)[0]
    S2_3 = _oracle.hist(x, u3, bins, op=S2, **kw)[0]
    T2_3 = _oracle.hist(x, u3, bins, op=T2, **kw)[0]
    assert S2_3.sum() > S2_2.sum()
    np.testing.assert_allclose(S2_3, s3 + T2_3, rtol=1e-12)


def test_flat_limit_convergence():
    """test/test_spherical_geometry.jl:150-179: spherical -> Cartesian like r/R.

    Restated per pair: the reference compares binned L2 sums, whose difference here is dominated
    by pairs crossing bin edges (haversine vs tangent-plane distance) and so depends on the random
    draw; comparing each pair's du_L isolates the O(r/R) frame difference the test is about."""
    def deviation(hw):
        rng = np.random.default_rng(4242)
        N, lat0 = 60, 35.0
        dlon = hw * (2 * rng.uniform(size=N) - 1)
        dlat = hw * (2 * rng.uniform(size=N) - 1)
        uu = rng.normal(size=(2, N))
        xs = np.stack([dlon, lat0 + dlat])
        xf = np.stack([EARTH_R * np.radians(dlon) * math.cos(math.radians(lat0)), EARTH_R * np.radians(dlat)])
        dev, ref = 0.0, 0.0
        for i in range(N):
            for j in range(i + 1, N):
                vs = _oracle.pair(xs, uu, [0.0, 1e9], i, j, op=L1, geometry="deg", radius=EARTH_R)[2]
                vf = _oracle.pair(xf, uu, [0.0, 1e9], i, j, op=L1)[2]
                dev, ref = max(dev, abs(vs - vf)), max(ref, abs(vf))
        return dev / ref
    d_big, d_small = deviation(4.0), deviation(1.0)
    assert d_small < d_big and d_big / d_small > 2.5


@pytest.mark.parametrize("geometry,D", [("flat", 2), ("flat", 3), ("deg", 2)])
def test_joint_mass_conservation(geometry, D):
    """test/test_2d_binning.jl:32-108: with value bins covering every value, the joint histogram
    marginalises to the 1D one."""
    rng = np.random.default_rng(9)
    N = 300
    x = rng.uniform(0, 1, (D, N)) if geometry == "flat" else rand_sphere(rng, N)
    u = rng.normal(size=(D, N))
    edges = np.geomspace(1e-3, 2.0, 17) if geometry == "flat" else np.geomspace(1e-3, math.pi, 17)
    for op in (L1, L2, L3, T3):
        s1, c1, _ = _oracle.hist(x, u, edges, op=op, geometry=geometry)
        s2, c2, _ = _oracle.hist(x, u, edges, op=op, geometry=geometry, vedges=[-1e9, 0.0, 1e9])
        assert np.array_equal(c2.sum(axis=1), c1)
        np.testing.assert_allclose(s2.sum(axis=1), s1, rtol=1e-12, atol=1e-12)


def test_threaded_equals_serial():
    """ThreadedBackend == SerialBackend (test/test_threads.jl, test/test_parallel_equivalence.jl)."""
    rng = np.random.default_rng(2)
    x = rng.uniform(0, 1, (2, 500))
    u = rng.normal(size=(2, 500))
    e = np.geomspace(1e-3, 1.5, 20)
    s1, c1, _ = _oracle.hist(x, u, e, op=L3, nthreads=1)
    s8, c8, _ = _oracle.hist(x, u, e, op=L3, nthreads=8)
    assert np.array_equal(c1, c8)
    np.testing.assert_allclose(s1, s8, rtol=1e-13, atol=1e-13)


def test_compensated_matches_naive_closely():
    rng = np.random.default_rng(3)
    x = rng.uniform(0, 1, (2, 400))
    u = rng.normal(size=(2, 400))
    e = np.geomspace(1e-3, 1.5, 12)
    sc, cc, a = _oracle.hist(x, u, e, op=L2, compensated=True)
    sn, cn, _ = _oracle.hist(x, u, e, op=L2, compensated=False)
    assert np.array_equal(cc, cn)
    assert np.all(np.abs(sc - sn) <= 1e-12 * a)
