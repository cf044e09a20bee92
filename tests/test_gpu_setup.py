"""Set-up in HBM (lulesh_hip_create_sedov) against the host set-up
(lulesh_domain_create + lulesh_hip_create, itself bitwise equal to the
oracle's): every field and the time state bit for bit after creation, and
after cycles on both contexts; the report numbers from the device fields
equal the host report."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FIELDS_ALL = ["x", "y", "z", "xd", "yd", "zd", "xdd", "ydd", "zdd", "fx", "fy", "fz",
              "nodalMass", "e", "p", "q", "ql", "qq", "v", "volo", "delv", "vdov", "arealg",
              "ss", "elemMass", "vnew", "delv_xi", "delv_eta", "delv_zeta", "delx_xi",
              "delx_eta", "delx_zeta"]


def banner(msg):
    print(f"\n[setup] {msg}", flush=True)


def state_tuple(ts):
    return (ts.time, ts.deltatime, ts.dtcourant, ts.dthydro, ts.cycle)


@pytest.mark.parametrize("s,nr,b,c", [(1, 1, 1, 1), (5, 11, 1, 1), (17, 16, 2, 1),
                                      (30, 21, 1, 2), (45, 11, 1, 1)])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_device_setup_equals_host_setup(pyl, s, nr, b, c, precision):
    banner(f"s={s} nr={nr} b={b} c={c} {precision}")
    prec = pyl.FP64 if precision == "fp64" else pyl.MIXED
    dom = pyl.Domain(s, nr, b, c)
    host = pyl.Device(dom, prec)
    # (the host Domain's context keeps fx.. / xdd.. by default; this one asks)
    dev = pyl.Device.sedov(s, nr, b, c, prec, opts=pyl.OPT_NODE_FORCES)
    assert dev.structured()
    for f in FIELDS_ALL:
        np.testing.assert_array_equal(dev.download(f), host.download(f), err_msg=f)
    th = dom.time_state()
    td = dev.time_state()
    assert state_tuple(td) == state_tuple(th)
    assert host.run(40, th) == 0 and dev.run(40, td) == 0
    assert state_tuple(td) == state_tuple(th)
    for f in FIELDS_ALL:
        np.testing.assert_array_equal(dev.download(f), host.download(f), err_msg=f)
    host.download_view()
    np.testing.assert_array_equal(dev.report(), dom.report())
    dev.close()
    host.close()


def test_device_setup_s250(pyl):
    """The s=250, 21-region problem: fields equal to the host set-up."""
    import time
    banner("s=250 nr=21")
    t0 = time.perf_counter()
    dom = pyl.Domain(250, 21)
    host = pyl.Device(dom)
    t1 = time.perf_counter()
    dev = pyl.Device.sedov(250, 21)
    t2 = time.perf_counter()
    print(f"[setup] host set-up + upload {t1 - t0:.2f} s, set-up in HBM {t2 - t1:.2f} s",
          flush=True)
    for f in ["x", "y", "z", "nodalMass", "e", "v", "volo", "elemMass"]:
        np.testing.assert_array_equal(dev.download(f), host.download(f), err_msg=f)
    th, td = dom.time_state(), dev.time_state()
    assert state_tuple(td) == state_tuple(th)
    assert host.run(3, th) == 0 and dev.run(3, td) == 0
    for f in ["x", "xd", "e", "p", "q", "ss", "v"]:
        np.testing.assert_array_equal(dev.download(f), host.download(f), err_msg=f)
    dev.close()
    host.close()


def test_batch_setup_in_hbm_equals_host_batch(pyl):
    """The ensemble built in HBM equals the batch of host Domains, bit for
    bit, before and after 60 cycles (per-problem dt and stop)."""
    banner("batch of 12")
    sizes, regions, factors = pyl.batch_params(3, 12, 5, 30)
    doms = [pyl.make_domain(int(s), int(r), float(f)) for s, r, f in zip(sizes, regions, factors)]
    host = pyl.Batch(doms, opts=pyl.OPT_NODE_FORCES)
    dev = pyl.Batch.sedov(sizes, regions, factors, opts=pyl.OPT_NODE_FORCES)
    for k, s in enumerate(sizes):
        for f in FIELDS_ALL:
            np.testing.assert_array_equal(dev.download(k, f), host.download(k, f),
                                          err_msg=f"{k} {f}")
    rh, sh = host.run(60)
    rd, sd = dev.run(60)
    assert rh == rd == 0
    for k in range(len(sizes)):
        assert state_tuple(sh[k]) == state_tuple(sd[k])
        for f in ["x", "xd", "e", "p", "q", "v", "ss"]:
            np.testing.assert_array_equal(dev.download(k, f), host.download(k, f),
                                          err_msg=f"{k} {f}")
