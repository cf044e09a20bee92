"""GPU checks of the contracted arithmetic form (LULESH_OPT_FMA): the same
kernels with multiply-adds fused, as the reference's GCC -O3 -mfma build
fuses them.  It cannot be bitwise the separately rounded oracle, so it is
held to the oracle within the tolerance the reference's results are quoted
to: origin energy within 1e-9 relative after full runs, the documented s=30
result, symmetry diffs at the oracle's level, fields within 1e-11 of the
oracle's scale after a few cycles; and it must be deterministic (a batch
member equals the same problem run alone, bit for bit)."""
import os

import numpy as np
import pytest

from oracle_lib import OracleDomain, set_threads

pytestmark = pytest.mark.gpu

STATE = ["x", "y", "z", "xd", "yd", "zd", "e", "p", "q", "v", "ss", "delv", "vdov", "arealg"]


def banner(msg):
    print(f"\n[fma] {msg}", flush=True)


def threads():
    return max(1, min(16, os.cpu_count() or 1))


def test_fma_s30_to_completion(pyl):
    """s=30 to stoptime with contraction: 932 cycles and the documented
    origin energy 2.025075e+05, within 1e-9 of the oracle."""
    banner("s=30 to completion, contracted")
    ora = OracleDomain(30)
    assert ora.run(10 ** 7) == 0
    dom, dev, ts = pyl.run_device(30, 10 ** 7, opts=pyl.OPT_FMA)
    assert ts.cycle == ora.cycle == 932
    e0, oe0 = dom.field("e")[0], ora.f64("e")[0]
    assert f"{e0:.6e}" == "2.025075e+05"
    assert abs(e0 - oe0) <= 1e-9 * oe0
    rep, orep = dom.report(), ora.report()
    print(f"[fma] report {rep.tolist()} oracle {orep.tolist()}", flush=True)
    assert rep[1] <= max(10 * orep[1], 1e-9 * e0)
    assert rep[3] <= max(10 * orep[3], 1e-12)


@pytest.mark.parametrize("s,its", [(12, 25), (31, 12)])
def test_fma_fields_close_to_oracle(pyl, s, its):
    """Every state field after a few cycles within 1e-11 of the oracle,
    relative to the field's largest magnitude."""
    banner(f"s={s}: {its} cycles contracted vs oracle")
    ora = OracleDomain(s)
    assert ora.run(its) == 0
    dom, dev, ts = pyl.run_device(s, its, opts=pyl.OPT_FMA)
    assert ts.cycle == ora.cycle == its
    assert abs(ts.deltatime - ora.scalar("deltatime")) <= 1e-11 * ts.deltatime
    for f in STATE:
        a, b = dom.field(f), ora.f64(f)
        scale = max(float(np.max(np.abs(b))), 1e-300)
        err = float(np.max(np.abs(a - b))) / scale
        assert err <= 1e-11, (f, err)


def test_fma_s150_300_cycles_vs_oracle(pyl):
    banner("s=150: 300 cycles contracted vs oracle")
    set_threads(threads())
    ora = OracleDomain(150, 11)
    assert ora.run(300) == 0
    dom, dev, ts = pyl.run_device(150, 300, 11, opts=pyl.OPT_FMA)
    assert ts.cycle == ora.cycle == 300
    e0, oe0 = dom.field("e")[0], ora.f64("e")[0]
    print(f"[fma] e0 rel diff {abs(e0 - oe0) / abs(oe0):.3e}", flush=True)
    assert abs(e0 - oe0) <= 1e-9 * abs(oe0)
    rep, orep = dom.report(), ora.report()
    print(f"[fma] gpu report {rep.tolist()} oracle {orep.tolist()}", flush=True)
    assert rep[1] <= max(10 * orep[1], 1e-11 * abs(e0))  # MaxAbsDiff
    assert rep[3] <= max(10 * orep[3], 1e-12)            # MaxRelDiff
    dev.close()
    ora.close()


def test_fma_setup_in_hbm_equals_host_setup(pyl):
    """The set-up keeps separate rounding in the contracted form: the problem
    built in HBM equals the uploaded host Domain, bit for bit."""
    banner("contracted form: set-up in HBM vs host set-up")
    dev = pyl.Device.sedov(20, 16, opts=pyl.OPT_FMA)
    dom = pyl.Domain(20, 16)
    ref = pyl.Device(dom, opts=pyl.OPT_FMA)
    for f in ("x", "y", "z", "volo", "elemMass", "nodalMass", "v", "e"):
        np.testing.assert_array_equal(dev.download(f), ref.download(f), err_msg=f)


def test_fma_batch_member_equals_single_run(pyl):
    """The contracted form is deterministic: a problem of a batch (built in
    HBM, two streams) evolves exactly as the same problem alone."""
    banner("contracted batch vs single runs")
    sizes, regions, factors = [9, 14, 20], [11, 16, 21], [1.0, 1.6, 0.7]
    b = pyl.Batch.sedov(sizes, regions, factors, opts=pyl.OPT_FMA | pyl.OPT_STREAMS(2))
    rc, states = b.run(60)
    assert rc == 0
    for p in range(3):
        dom = pyl.make_domain(sizes[p], regions[p], factors[p])
        dev = pyl.Device(dom, opts=pyl.OPT_FMA)
        ts = dom.time_state()
        assert dev.run(60, ts) == 0
        assert states[p].cycle == ts.cycle and states[p].time == ts.time
        for f in ("x", "xd", "e", "p", "q", "v", "ss"):
            np.testing.assert_array_equal(b.download(p, f), dev.download(f), err_msg=f)


def test_fma_mixed_close_to_fp64(pyl):
    """Mixed precision with contraction: origin energy within 1e-6 of the
    separately rounded FP64 form."""
    banner("contracted mixed vs fp64")
    _, d64, _ = pyl.run_device(45, 300)
    _, dmx, _ = pyl.run_device(45, 300, precision=pyl.MIXED, opts=pyl.OPT_FMA)
    e64, emx = d64.download("e")[0], dmx.download("e")[0]
    assert abs(emx - e64) <= 1e-6 * abs(e64)


def _max_rel(a, b):
    scale = max(float(np.max(np.abs(b))), 1e-300)
    return float(np.max(np.abs(np.asarray(a) - np.asarray(b)))) / scale


@pytest.mark.parametrize("path", ["structured", "generic"])
@pytest.mark.parametrize("s", [8, 16, 31])
def test_fma_phases_from_perturbed_state(pyl, s, path):
    """One nodal, one kinematics and one material phase of the contracted
    form against the oracle from the same state: a developed state (30
    oracle cycles) whose velocities get seeded noise of 10% of their largest
    magnitude (every element's hourglass modes excited) and whose
    coordinates get noise of 1% of the mesh spacing (no element a
    parallelepiped).  The contracted build evaluates the hourglass force,
    the stress corner forces, the Jacobian, the Q gradients, the velocity
    divergence, the EOS and the constraints in a rearranged algebra (the
    structured nodal kernel) or with Newton reciprocals (the others); every
    result stays within 1e-12 of the oracle relative to the field's largest
    magnitude, and the time constraints within 1e-13."""
    banner(f"contracted phases from a perturbed state s={s} {path}")
    ora = OracleDomain(s)
    assert ora.run(30) == 0
    rng = np.random.default_rng(1000 + s)
    vmax = max(float(np.max(np.abs(ora.f64(f)))) for f in ("xd", "yd", "zd"))
    for f in ("xd", "yd", "zd"):
        v = ora.f64(f)
        v += rng.uniform(-0.1, 0.1, v.size) * vmax
    h = 1.125 / s
    for f in ("x", "y", "z"):
        c = ora.f64(f)
        c += rng.uniform(-0.01, 0.01, c.size) * h
    dom = pyl.Domain(s)
    opts = pyl.OPT_NODE_FORCES | pyl.OPT_FMA | (pyl.OPT_GENERIC if path == "generic" else 0)
    dev = pyl.Device(dom, opts=opts)
    assert dev.structured() == (path == "structured")
    for f in pyl.FIELDS:
        dev.upload(f, ora.f64(f))
    ora.time_increment()
    dt = ora.scalar("deltatime")
    assert dev.phase(pyl.PHASE_NODAL, dt) == 0
    assert ora.nodal() == 0
    worst = {}
    for f in ["fx", "fy", "fz", "xdd", "ydd", "zdd", "xd", "yd", "zd", "x", "y", "z"]:
        worst[f] = _max_rel(dev.download(f), ora.f64(f))
    for f in pyl.FIELDS:
        dev.upload(f, ora.f64(f))
    assert dev.phase(pyl.PHASE_ELEMENTS, dt) == 0
    assert ora.elements() == 0
    for f in ["vnew", "delv", "arealg", "vdov", "v", "delv_xi", "delv_eta", "delv_zeta",
              "delx_xi", "delx_eta", "delx_zeta"]:
        worst[f] = _max_rel(dev.download(f), ora.f64(f))
    for f in pyl.FIELDS:
        dev.upload(f, ora.f64(f))
    assert dev.phase(pyl.PHASE_MATERIAL, dt) == 0
    assert ora.material() == 0
    for f in ["e", "p", "q", "ql", "qq", "ss"]:
        worst[f] = _max_rel(dev.download(f), ora.f64(f))
    ora.constraints()
    dtc, dth = dev.constraints_result()
    assert abs(dtc - ora.scalar("dtcourant")) <= 1e-13 * ora.scalar("dtcourant")
    assert abs(dth - ora.scalar("dthydro")) <= 1e-13 * ora.scalar("dthydro")
    print(f"[fma] worst {max(worst, key=worst.get)} {max(worst.values()):.2e}; "
          + " ".join(f"{k}={v:.1e}" for k, v in worst.items()), flush=True)
    for f, w in worst.items():
        assert w <= 1e-12, (f, w)
    dev.close()
    ora.close()
