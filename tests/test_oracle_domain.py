"""Pins of the oracle at Domain level: setup closed forms, TimeIncrement's
branch table, closed-form first motion of the Sedov blast, the symmetries
and the size self-similarity of the discrete problem, the monotonic-Q and
gradient closed forms, and the documented LULESH 2.0 result at -s 30.
"""
import json
import os

import numpy as np
import pytest

from oracle_lib import OracleDomain

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def test_mesh_and_connectivity_closed_form():
    s = 4
    d = OracleDomain(s)
    n = s + 1
    idx = np.arange(n)
    K, J, I = np.meshgrid(idx, idx, idx, indexing="ij")
    np.testing.assert_array_equal(d.f64("x"), (1.125 * I / s).ravel())
    np.testing.assert_array_equal(d.f64("y"), (1.125 * J / s).ravel())
    np.testing.assert_array_equal(d.f64("z"), (1.125 * K / s).ravel())
    nl = d.i32("nodelist").reshape(-1, 8)
    e = np.arange(s ** 3)
    k, j, i = e // (s * s), (e // s) % s, e % s
    base = k * n * n + j * n + i
    expect = np.stack([base, base + 1, base + n + 1, base + n, base + n * n, base + n * n + 1,
                       base + n * n + n + 1, base + n * n + n], axis=1)
    np.testing.assert_array_equal(nl, expect)
    # face neighbours in the interior
    inner = (i > 0) & (i < s - 1) & (j > 0) & (j < s - 1) & (k > 0) & (k < s - 1)
    for name, off in [("lxim", -1), ("lxip", 1), ("letam", -s), ("letap", s),
                      ("lzetam", -s * s), ("lzetap", s * s)]:
        np.testing.assert_array_equal(d.i32(name)[inner], e[inner] + off)


def test_boundary_flags_and_symmetry_sets():
    s = 3
    d = OracleDomain(s)
    bc = d.i32("elemBC")
    e = np.arange(s ** 3)
    k, j, i = e // (s * s), (e // s) % s, e % s
    expect = ((i == 0) * 0x1 | (i == s - 1) * 0x10 | (j == 0) * 0x40 | (j == s - 1) * 0x400 |
              (k == 0) * 0x1000 | (k == s - 1) * 0x10000)
    np.testing.assert_array_equal(bc, expect)
    n = s + 1
    x, y, z = d.f64("x"), d.f64("y"), d.f64("z")
    assert sorted(d.i32("symmX")) == sorted(np.nonzero(x == 0)[0])
    assert sorted(d.i32("symmY")) == sorted(np.nonzero(y == 0)[0])
    assert sorted(d.i32("symmZ")) == sorted(np.nonzero(z == 0)[0])
    assert len(d.i32("symmX")) == n * n


def test_corner_lists_ascending_element_order():
    """Interior node (K,J,I) sees elements (K-1..K, J-1..J, I-1..I) in
    ascending index order with local corners 6,7,5,4,2,3,1,0."""
    s = 4
    d = OracleDomain(s)
    n = s + 1
    start, cl = d.i32("nodeElemStart"), d.i32("nodeElemCornerList")
    K, J, I = 2, 1, 3
    g = K * n * n + J * n + I
    corners = cl[start[g]:start[g + 1]]
    expect = []
    for (dk, dj, di, c) in [(-1, -1, -1, 6), (-1, -1, 0, 7), (-1, 0, -1, 5), (-1, 0, 0, 4),
                            (0, -1, -1, 2), (0, -1, 0, 3), (0, 0, -1, 1), (0, 0, 0, 0)]:
        el = (K + dk) * s * s + (J + dj) * s + (I + di)
        expect.append(8 * el + c)
    np.testing.assert_array_equal(corners, expect)
    assert start[-1] == 8 * s ** 3
    # every corner exactly once
    assert sorted(cl) == list(range(8 * s ** 3))


def _gather(start, cl, a, b, order):
    """Per-node sums of corner values a (stress) and b (hourglass) in a given
    order, as Python float additions (IEEE doubles, left to right)."""
    out = np.zeros(len(start) - 1)
    for g in range(len(start) - 1):
        corners = [int(c) for c in cl[start[g]:start[g + 1]]]
        acc = 0.0
        if order == "interleaved":        # lulesh.cc:849-857
            for c in corners:
                acc += float(a[c])
                acc += float(b[c])
        elif order == "stress_then_hourglass":
            for c in corners:
                acc += float(a[c])
            for c in corners:
                acc += float(b[c])
        elif order == "descending":
            for c in reversed(corners):
                acc += float(a[c])
                acc += float(b[c])
        out[g] = acc
    return out


def test_node_force_gather_order():
    """The oracle's node forces are the reference's gather
    (combineVolumeForcesTaskFunc, lulesh.cc:838-862): per node, over its
    corners in nodeElemCornerList order (ascending element index), stress
    then hourglass force of each corner, one rounded add at a time.  Rebuilt
    here from the oracle's own corner arrays (ora_calc_corner_forces, the
    reference's fx_elem layout [8 elem + corner]) and held to the oracle's
    fx/fy/fz bit for bit, on a state whose hourglass forces are non-zero;
    the two plausible slips -- all stress before all hourglass forces, or
    the corners in descending order -- change the bits, so the comparison
    pins the order."""
    s = 6
    d = OracleDomain(s)
    assert d.run(25) == 0
    start, cl = d.i32("nodeElemStart"), d.i32("nodeElemCornerList")
    rc, (sx, sy, sz, hx, hy, hz) = d.corner_forces()
    assert rc == 0
    assert np.max(np.abs(hx)) > 0 and np.max(np.abs(sx)) > 0
    assert d.nodal() == 0
    for f, a, b in (("fx", sx, hx), ("fy", sy, hy), ("fz", sz, hz)):
        got = d.f64(f)
        np.testing.assert_array_equal(got, _gather(start, cl, a, b, "interleaved"), err_msg=f)
        for wrong in ("stress_then_hourglass", "descending"):
            assert np.any(got != _gather(start, cl, a, b, wrong)), (f, wrong)


def test_volumes_and_masses():
    s = 6
    d = OracleDomain(s)
    h = 1.125 / s
    np.testing.assert_allclose(d.f64("volo"), h ** 3, rtol=1e-13)
    np.testing.assert_array_equal(d.f64("volo"), d.f64("elemMass"))
    # nodal mass = (number of incident elements) * h^3 / 8
    count = np.diff(d.i32("nodeElemStart"))
    np.testing.assert_allclose(d.f64("nodalMass"), count * h ** 3 / 8, rtol=1e-13)
    assert d.f64("nodalMass").sum() == pytest.approx(d.f64("volo").sum(), rel=1e-13)


def test_initial_energy_and_dt():
    for s, einit in [(30, 1.169998815e7), (45, 3.948746e7), (150, 1.462498519e9)]:
        if s > 45:
            continue  # the large domain is checked by the closed form below
        d = OracleDomain(s)
        e = d.f64("e")
        assert e[0] == pytest.approx(einit, rel=1e-9)
        assert np.count_nonzero(e) == 1
        vol0 = d.f64("volo")[0]
        assert d.scalar("deltatime") == pytest.approx(0.5 * np.cbrt(vol0) / np.sqrt(2 * e[0]),
                                                      rel=1e-15)
    assert 3.948746e7 * (150 / 45) ** 3 == pytest.approx(1.462498519e9, rel=1e-9)


def test_region_sizes_s30():
    """Region sizes at -s 30 -r 11 -b 1 -c 1 (the region map comes from glibc
    rand() after srand(0)); the list is an independent derivation of the
    reference's CreateRegionIndexSets (lulesh-init.cc:341-445), made apart
    from this oracle."""
    d = OracleDomain(30)
    sizes = [len(d.region(r)) for r in range(d.num_reg)]
    assert sizes == [123, 457, 493, 2733, 1004, 896, 3236, 2435, 1198, 5980, 8445]
    assert [d.rep(r) for r in range(11)] == [1, 1, 1, 1, 1, 2, 2, 2, 2, 2, 20]
    # regions partition the elements, each list ascending
    allel = np.concatenate([d.region(r) for r in range(11)])
    assert sorted(allel) == list(range(27000))
    for r in range(11):
        assert np.all(np.diff(d.region(r)) > 0)


def test_region_rep_rule():
    d16 = OracleDomain(2, nr=16, cost=2)
    # reg < 8 -> 1 ; reg < 16 - 31//20 = 15 -> 1+cost ; else 10*(1+cost)
    assert [d16.rep(r) for r in range(16)] == [1] * 8 + [3] * 7 + [30]
    d1 = OracleDomain(3, nr=1)
    # nr = 1: reg 0 is not below 1/2 but is below 1 - 16//20 -> 1 + cost
    assert len(d1.region(0)) == 27 and d1.rep(0) == 2


def test_region_balance_weights():
    """-b 2: region r+1 is drawn with weight (r+1)^2, so higher regions get
    more runs; every element is still assigned once."""
    d = OracleDomain(20, nr=5, balance=2)
    sizes = [len(d.region(r)) for r in range(5)]
    assert sum(sizes) == 8000
    assert sizes[4] > sizes[0]


def ts_case(**kw):
    d = OracleDomain(2)
    for k, v in kw.items():
        if k == "cycle":
            d.cycle = v
        else:
            d.set_scalar(k, v)
    d.time_increment()
    return d.scalar("deltatime"), d.scalar("time"), d.cycle


def test_time_increment_branches():
    base = dict(time=0.0, stoptime=1.0, deltatime=1e-3, dtmax=1e-2, cycle=5)
    # cycle 0 keeps dt0
    assert ts_case(**{**base, "cycle": 0, "dtcourant": 1e-9, "dthydro": 1e-9})[0] == 1e-3
    # shrinking: take the constraint (courant/2)
    assert ts_case(**base, dtcourant=1e-3, dthydro=1e20)[0] == 1e-3 / 2
    # hydro * 2/3
    assert ts_case(**base, dtcourant=1e20, dthydro=1.2e-3)[0] == 1.2e-3 * 2 / 3
    # 1 <= ratio < 1.1: hold
    assert ts_case(**base, dtcourant=2.1e-3, dthydro=1e20)[0] == 1e-3
    # ratio > 1.2: cap at 1.2 * olddt
    assert ts_case(**base, dtcourant=1e-1, dthydro=1e20)[0] == 1e-3 * 1.2
    # 1.1 <= ratio <= 1.2: take it
    assert ts_case(**base, dtcourant=2.3e-3, dthydro=1e20)[0] == 2.3e-3 / 2
    # dtmax
    assert ts_case(**{**base, "deltatime": 9.5e-3}, dtcourant=1.0, dthydro=1e20)[0] == 1e-2
    # stop-time approach: target in (dt, 4/3 dt) -> 2/3 dt
    dt, t, c = ts_case(**{**base, "time": 1.0 - 1.2e-3, "cycle": 0})
    assert dt == 2.0 * 1e-3 / 3.0 and c == 1
    # target below dt -> clipped to the stop time
    dt, t, c = ts_case(**{**base, "time": 1.0 - 4e-4, "cycle": 0})
    assert dt == pytest.approx(4e-4, rel=1e-12) and t == pytest.approx(1.0, rel=1e-15)


def test_first_motion_closed_form():
    """Cycle 1 only converts e(0) into p(0) = (2/3) e(0); cycle 2 pushes the
    eight nodes of element 0 with the outward quarter-face normals,
    f = p h^2/4 per axis, and nodes on the symmetry planes keep zero
    normal velocity.  Node (1,1,1) has mass h^3 and moves along (1,1,1)."""
    s = 5
    d = OracleDomain(s)
    h = 1.125 / s
    einit = d.f64("e")[0]
    dt0 = d.scalar("deltatime")
    d.time_increment()
    assert d.leapfrog() == 0
    assert np.all(d.f64("xd") == 0)
    assert d.f64("p")[0] == pytest.approx(2 / 3 * einit, rel=1e-15)
    assert d.scalar("dtcourant") == 1e20  # no element moves: vdov == 0 everywhere
    d.time_increment()
    dt1 = d.scalar("deltatime")
    assert dt1 == pytest.approx(1.2 * dt0, rel=1e-15)
    assert d.leapfrog() == 0
    n = s + 1
    g111 = 1 * n * n + 1 * n + 1
    v = (2 / 3 * einit) * h * h / 4 / h ** 3 * dt1
    for f in ("xd", "yd", "zd"):
        assert d.f64(f)[g111] == pytest.approx(v, rel=1e-13)
    g100 = 1  # node (0,0,1) in (K,J,I): on y=0 and z=0 planes
    assert d.f64("xd")[g100] > 0 and d.f64("yd")[g100] == 0 and d.f64("zd")[g100] == 0
    assert d.f64("xd")[0] == d.f64("yd")[0] == d.f64("zd")[0] == 0
    # untouched nodes
    far = n * n * n - 1
    assert d.f64("xd")[far] == 0


def test_symmetric_under_axis_permutation():
    """The Sedov set-up is symmetric under any permutation of x, y, z; the
    discrete solution keeps it to rounding (checked on e and positions)."""
    s = 8
    d = OracleDomain(s)
    assert d.run(60) == 0
    e = d.f64("e").reshape(s, s, s)  # [k][j][i]
    scale = np.max(np.abs(e))
    for perm in [(0, 2, 1), (1, 0, 2), (2, 1, 0), (1, 2, 0)]:
        assert np.max(np.abs(e - e.transpose(perm))) < 1e-10 * scale
    n = s + 1
    x = d.f64("x").reshape(n, n, n)
    y = d.f64("y").reshape(n, n, n)
    z = d.f64("z").reshape(n, n, n)
    assert np.max(np.abs(x - y.transpose(0, 2, 1))) < 1e-13
    assert np.max(np.abs(x - z.transpose(2, 1, 0))) < 1e-13
    # symmetry-plane nodes never leave their plane (lulesh.cc:566-588)
    assert np.all(x[:, :, 0] == 0) and np.all(y[:, 0, :] == 0) and np.all(z[0, :, :] == 0)
    rep = d.report()
    assert rep[1] < 1e-10 * scale


def test_size_self_similarity():
    """Lengths scale as 1/s and the deposited energy as s^3, so until the
    blast reaches the far boundary the discrete solution is the same in
    index space for every s: e*volo, and dt*s^2.5, agree across sizes."""
    a, b = OracleDomain(10), OracleDomain(24)
    assert a.run(60) == 0 and b.run(60) == 0
    ea = (a.f64("e") * a.f64("volo")).reshape(10, 10, 10)[:6, :6, :6]
    eb = (b.f64("e") * b.f64("volo")).reshape(24, 24, 24)[:6, :6, :6]
    assert np.max(np.abs(ea - eb)) < 1e-9 * np.max(np.abs(ea))
    assert a.scalar("deltatime") * 10 ** 2.5 == pytest.approx(b.scalar("deltatime") * 24 ** 2.5,
                                                              rel=1e-9)


def test_q_gradients_homologous_expansion():
    """v = alpha * x on the undeformed mesh, dt = 0: delx_* = h,
    delv_* = alpha, vdov = 3 alpha, arealg = h, vnew = 1 (closed forms of
    lulesh.cc:1349-1416 and 1180-1239)."""
    s = 3
    d = OracleDomain(s)
    h = 1.125 / s
    alpha = -0.7
    for c in ("x", "y", "z"):
        d.f64(c + "d")[:] = alpha * d.f64(c)
    d.set_scalar("deltatime", 0.0)
    assert d.elements() == 0
    for f in ("delx_xi", "delx_eta", "delx_zeta", "arealg"):
        np.testing.assert_allclose(d.f64(f), h, rtol=1e-13)
    for f in ("delv_xi", "delv_eta", "delv_zeta"):
        np.testing.assert_allclose(d.f64(f), alpha, rtol=1e-13)
    np.testing.assert_allclose(d.f64("vdov"), 3 * alpha, rtol=1e-13)
    np.testing.assert_allclose(d.f64("vnew"), 1.0, rtol=1e-14)


def test_monotonic_q_uniform_compression():
    """Uniform compression: every limiter ratio is 1, so phi = 1 and q = 0
    except across a free surface, where the outside gradient is 0, phi = 0
    and q takes its full linear and quadratic values (lulesh.cc:1457-1645)."""
    s = 3
    d = OracleDomain(s)
    h = 1.125 / s
    alpha = -0.5
    for c in ("x", "y", "z"):
        d.f64(c + "d")[:] = alpha * d.f64(c)
    d.set_scalar("deltatime", 0.0)
    assert d.elements() == 0
    assert d.material() == 0
    ql, qq = d.f64("ql"), d.f64("qq")
    e = np.arange(s ** 3)
    k, j, i = e // (s * s), (e // s) % s, e % s
    nfree = (i == s - 1).astype(int) + (j == s - 1) + (k == s - 1)
    dvx = alpha * h  # delv * delx, negative
    np.testing.assert_allclose(ql, -0.5 * 1.0 * dvx * nfree, rtol=1e-12, atol=1e-300)
    np.testing.assert_allclose(qq, 2 / 3 * dvx * dvx * nfree, rtol=1e-12, atol=1e-300)
    # expansion: no viscosity at all
    d2 = OracleDomain(s)
    for c in ("x", "y", "z"):
        d2.f64(c + "d")[:] = 0.5 * d2.f64(c)
    d2.set_scalar("deltatime", 0.0)
    d2.elements()
    d2.material()
    assert np.all(d2.f64("ql") == 0) and np.all(d2.f64("qq") == 0)


def test_constraints_uniform_state():
    s = 3
    d = OracleDomain(s)
    d.f64("ss")[:] = 2.0
    d.f64("arealg")[:] = 0.1
    d.f64("vdov")[:] = -0.5
    d.f64("vdov")[4] = 0.0  # skipped element
    d.constraints()
    assert d.scalar("dtcourant") == pytest.approx(0.1 / np.sqrt(4 + 256 * 0.01 * 0.25), rel=1e-15)
    assert d.scalar("dthydro") == pytest.approx(0.1 / 0.5, rel=1e-15)
    d.f64("vdov")[:] = 0.0
    d.constraints()
    assert d.scalar("dtcourant") == 1e20 and d.scalar("dthydro") == 1e20


def test_volume_error_paths():
    d = OracleDomain(3)
    d.f64("v")[4] = -1.0
    d.time_increment()
    assert d.leapfrog() == -1  # volo*v <= 0 in the hourglass task, lulesh.cc:906-910


def test_s30_published_result():
    """LULESH 2.0's documented run at -s 30 (to completion): 932 cycles and
    final origin energy 2.025075e+05.  The HPX path prints the same report
    (lulesh-util.cc:74-79), so the oracle must reproduce both."""
    with open(os.path.join(GOLDEN, "s30_published.json")) as f:
        gold = json.load(f)
    d = OracleDomain(30)
    assert d.run(10 ** 7) == 0
    assert d.cycle == gold["iterations"]
    rep = d.report()
    assert f"{rep[0]:.6e}" == gold["final_origin_energy_6"]
    # symmetry diffs stay at rounding level
    assert rep[1] < 1e-9 and rep[3] < 1e-11
