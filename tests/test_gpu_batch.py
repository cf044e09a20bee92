"""GPU checks of the batched ensemble (many Domains in one context) and of
the autotuner: every problem of a batch must evolve exactly as it does when
run alone, with its own time step, time and cycle count."""
import numpy as np
import pytest

from oracle_lib import OracleDomain

pytestmark = pytest.mark.gpu

FIELDS = ["x", "y", "z", "xd", "yd", "zd", "e", "p", "q", "v", "ss", "vdov", "arealg"]


def banner(msg):
    print(msg, flush=True)


def single_run(pyl, s, nr, factor, its, precision=0):
    dom = pyl.make_domain(s, nr, factor)
    dev = pyl.Device(dom, precision)
    ts = dom.time_state()
    assert dev.run(its, ts) == 0
    return dev, ts


def test_batch_params_deterministic(pyl):
    s1, r1, f1 = pyl.batch_params(0, 64)
    s2, r2, f2 = pyl.batch_params(0, 64)
    assert (s1 == s2).all() and (r1 == r2).all() and (f1 == f2).all()
    assert s1.min() >= 20 and s1.max() <= 80
    assert set(r1) <= {11, 16, 21} and f1.min() >= 0.5 and f1.max() < 2.0


def test_batch_equals_individual_runs(pyl):
    banner("batch of 5 vs individual runs")
    params = [(5, 11, 1.0), (9, 16, 0.6), (12, 21, 1.7), (7, 11, 1.3), (16, 16, 0.9)]
    its = 70
    doms = [pyl.make_domain(s, nr, f) for s, nr, f in params]
    batch = pyl.Batch(doms)
    rc, states = batch.run(its)
    assert rc == 0
    for k, (s, nr, f) in enumerate(params):
        dev, ts = single_run(pyl, s, nr, f, its)
        assert states[k].cycle == ts.cycle == its
        assert states[k].time == ts.time and states[k].deltatime == ts.deltatime
        for name in FIELDS:
            np.testing.assert_array_equal(batch.download(k, name), dev.download(name),
                                          err_msg=f"problem {k} {name}")


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_ragged_batch_with_tiny_problems(pyl, mode):
    """A batch mixing one- to eight-element problems (most regions empty)
    with larger ones, run until every problem reaches its stop time: each
    problem stops at its own cycle, and every problem's state equals the
    same problem run alone, bit for bit, in each material mode."""
    banner(f"ragged batch, material_mode={mode}")
    params = [(1, 11, 1.0), (14, 16, 0.8), (2, 11, 1.5), (1, 1, 0.7), (9, 21, 1.2),
              (3, 21, 1.0)]
    batch = pyl.Batch([pyl.make_domain(s, nr, f) for s, nr, f in params])
    t = batch.tuning()
    t.material_mode = mode
    batch.set_tuning(t)
    rc, states = batch.run(10 ** 7)
    assert rc == 0
    cycles = set()
    for k, (s, nr, f) in enumerate(params):
        dom = pyl.make_domain(s, nr, f)
        dev = pyl.Device(dom)
        dt = dev.tuning()
        dt.material_mode = mode
        dev.set_tuning(dt)
        ts = dom.time_state()
        assert dev.run(10 ** 7, ts) == 0
        assert states[k].cycle == ts.cycle and states[k].time == ts.time, k
        cycles.add(ts.cycle)
        for name in FIELDS:
            np.testing.assert_array_equal(batch.download(k, name), dev.download(name),
                                          err_msg=f"problem {k} {name}")
        dev.close()
    assert len(cycles) > 1  # the problems stop at different cycles


def test_batch_problem_matches_oracle(pyl):
    banner("batch member vs oracle (scaled energy)")
    doms = [pyl.make_domain(6, 11, 1.0), pyl.make_domain(10, 21, 1.8)]
    batch = pyl.Batch(doms)
    rc, states = batch.run(120)
    assert rc == 0
    ora = OracleDomain(10, 21)
    ora.scale_energy(1.8)
    assert ora.run(120) == 0
    assert states[1].cycle == ora.cycle
    e = batch.download(1, "e")
    assert np.max(np.abs(e - ora.f64("e"))) <= 1e-9 * np.max(np.abs(ora.f64("e")))


def test_batch_per_problem_stop_time(pyl):
    """Small problems reach the stop time at their own cycle counts while
    the batch continues (each with its own dt)."""
    banner("batch to completion s=4..7")
    params = [(4, 11, 0.5), (5, 11, 2.0), (6, 16, 1.0), (7, 21, 1.4)]
    doms = [pyl.make_domain(*p) for p in params]
    batch = pyl.Batch(doms)
    rc, states = batch.run(10 ** 6)
    assert rc == 0
    cycles = set()
    for k, p in enumerate(params):
        dev, ts = single_run(pyl, *p, 10 ** 6)
        assert states[k].cycle == ts.cycle
        assert abs(states[k].time - 1e-2) < 1e-12
        np.testing.assert_array_equal(batch.download(k, "e"), dev.download("e"))
        cycles.add(ts.cycle)
    assert len(cycles) > 1


@pytest.mark.parametrize("fma", [False, True])
def test_autotune_keeps_state(pyl, fma):
    """The autotuner restores the state it started from: a run after tuning
    equals a run without it, bit for bit (both arithmetic forms)."""
    banner(f"autotune fma={fma}")
    opts = pyl.OPT_FMA if fma else 0
    params = [(10, 11, 1.0), (14, 16, 1.2)]
    a = pyl.Batch([pyl.make_domain(*p) for p in params], opts=opts)
    rc, _ = a.run(20)
    assert rc == 0
    t, ms = a.autotune(6)
    assert t.zplanes >= 1 and t.block_elem in (64, 128, 256) and ms > 0
    rc, sa = a.run(50)
    b = pyl.Batch([pyl.make_domain(*p) for p in params], opts=opts)
    rc2, sb = b.run(50)
    assert rc == rc2 == 0
    for k in range(2):
        assert sa[k].cycle == sb[k].cycle == 50
        np.testing.assert_array_equal(a.download(k, "e"), b.download(k, "e"))
        np.testing.assert_array_equal(a.download(k, "x"), b.download(k, "x"))


@pytest.mark.parametrize("tune", [(3, 64, 128, 4, 1), (7, 128, 64, 6, 0), (40, 256, 256, 16, 1),
                                  (2, 256, 128, 8, 0)])
def test_tunings_do_not_change_results(pyl, tune):
    banner(f"tuning {tune}")
    dom = pyl.Domain(18)
    dev = pyl.Device(dom)
    t = pyl.Tuning(*tune)
    dev.set_tuning(t)
    ts = dom.time_state()
    assert dev.run(45, ts) == 0
    ref, _ = single_run(pyl, 18, 11, 1.0, 45)
    for name in FIELDS:
        np.testing.assert_array_equal(dev.download(name), ref.download(name), err_msg=name)


def test_mixed_batch(pyl):
    banner("mixed batch")
    params = [(8, 11, 1.0), (12, 16, 1.5)]
    batch = pyl.Batch([pyl.make_domain(*p) for p in params], precision=pyl.MIXED)
    rc, states = batch.run(100)
    assert rc == 0
    for k, p in enumerate(params):
        dev, ts = single_run(pyl, *p, 100)
        e64 = dev.download("e")[0]
        e32 = batch.download(k, "e")[0]
        assert abs(e32 - e64) <= 1e-6 * abs(e64)


@pytest.mark.parametrize("k", [2, 3])
def test_stream_groups_equal_one_stream(pyl, k):
    """A batch split over k streams (each driven by its own host thread)
    leaves every problem exactly as the one-stream batch does."""
    banner(f"batch over {k} streams vs one stream")
    params = [(5, 11, 1.0), (9, 16, 0.6), (12, 21, 1.7), (7, 11, 1.3), (16, 16, 0.9)]
    sizes = [p[0] for p in params]
    regions = [p[1] for p in params]
    factors = [p[2] for p in params]
    its = 60
    one = pyl.Batch.sedov(sizes, regions, factors)
    rc1, s1 = one.run(its)
    grp = pyl.Batch.sedov(sizes, regions, factors, opts=pyl.OPT_STREAMS(k))
    rc2, s2 = grp.run(its)
    assert rc1 == rc2 == 0
    for p in range(len(params)):
        assert s1[p].cycle == s2[p].cycle == its
        assert s1[p].time == s2[p].time and s1[p].deltatime == s2[p].deltatime
        for name in FIELDS:
            np.testing.assert_array_equal(grp.download(p, name), one.download(p, name),
                                          err_msg=f"problem {p} {name}")
    # a second run continues every lane from where it stopped (the smallest
    # problems reach the stop time on the way)
    rc1, s1 = one.run(its + 25)
    rc2, s2 = grp.run(its + 25)
    assert rc1 == rc2 == 0
    assert max(st.cycle for st in s1) == its + 25
    for p in range(len(params)):
        assert s1[p].cycle == s2[p].cycle and s1[p].time == s2[p].time
        np.testing.assert_array_equal(grp.download(p, "e"), one.download(p, "e"))
    buf = np.empty(8, np.float64)
    assert grp.L.lulesh_hip_batch_download(grp.h, len(params), pyl.FIELD_ID["e"],
                                           buf.ctypes.data, 8) == pyl.INVALID_ARGUMENT


def test_stream_group_host_domains_and_tuning(pyl):
    """The group built from host Domains, autotuned per lane and given an
    explicit tuning, still equals the individual runs."""
    banner("stream group from host Domains, autotune")
    params = [(6, 11, 1.0), (10, 21, 1.8), (8, 16, 0.7)]
    its = 40
    grp = pyl.Batch([pyl.make_domain(*p) for p in params], opts=pyl.OPT_STREAMS(2))
    t, ms = grp.autotune(4)
    assert ms > 0 and t.zplanes >= 1
    grp.set_tuning(pyl.Tuning(3, 64, 128, 4, 1))
    rc, states = grp.run(its)
    assert rc == 0
    for p, prm in enumerate(params):
        dev, ts = single_run(pyl, *prm, its)
        assert states[p].cycle == ts.cycle and states[p].time == ts.time
        for name in FIELDS:
            np.testing.assert_array_equal(grp.download(p, name), dev.download(name),
                                          err_msg=f"problem {p} {name}")


def test_benchmark_batch_of_64(pyl):
    """The benchmark ensemble itself: 64 problems drawn with seed 0 (sizes
    in [20, 80], regions in {11, 16, 21}, deposit x [0.5, 2)), 1000 cycles
    or each problem's stop time, in the benchmark's form (multiply-adds
    fused, two streams).  Every problem ends bit for bit where the same
    problem run alone ends; three of them (the smallest mesh, which has 21
    regions; the largest; a 21-region s=52 one) match the oracle's
    full-length runs (tests/golden/oracle_batch64_p*.json) at 1e-9."""
    from test_gpu_long import LIVE_ELEM, NODE, compare_with_fixture, load_fixture

    banner("benchmark batch: 64 problems, 1000 cycles")
    sizes, regions, factors = pyl.batch_params(0, 64)
    its = 1000
    b = pyl.Batch.sedov(sizes, regions, factors, opts=pyl.OPT_FMA | pyl.OPT_STREAMS(2))
    rc, states = b.run(its)
    assert rc == 0
    cyc = [st.cycle for st in states]
    print(f"[batch64] cycles {min(cyc)}..{max(cyc)}", flush=True)
    assert max(cyc) == its and min(cyc) < its  # the small problems stop early
    for k in range(64):
        one = pyl.Batch.sedov([sizes[k]], [regions[k]], [factors[k]], opts=pyl.OPT_FMA)
        rc1, st1 = one.run(its)
        assert rc1 == 0
        assert (st1[0].cycle, st1[0].time, st1[0].deltatime) == \
            (states[k].cycle, states[k].time, states[k].deltatime), k
        for name in FIELDS:
            np.testing.assert_array_equal(b.download(k, name), one.download(0, name),
                                          err_msg=f"problem {k} {name}")
        one.close()
    for k in (40, 17, 26):
        meta, arrays = load_fixture(f"batch64_p{k}")
        assert (meta["s"], meta["nr"]) == (int(sizes[k]), int(regions[k]))
        assert meta["factor"] == float(factors[k])
        print(f"[batch64] problem {k}: s={meta['s']} regions={meta['nr']}", flush=True)
        compare_with_fixture(meta, arrays, lambda f: b.download(k, f), states[k], 1e-9,
                             NODE + LIVE_ELEM)
    b.close()
