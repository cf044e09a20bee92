"""The fused elements + material pass (material_mode 2, k_elemmat_zmarch)
against the separate kernels (k_elems, then k_material_ordered: material
mode 0) on the same problem.

Per element both run the same operations on the same inputs; the fused pass
recomputes the gradients of the elements bordering each tile and of the
planes bordering each plane group, with the owner's arithmetic.  So in the
separately rounded form every state field, the time state and the cycle
are bit for bit those of mode 0, over tiles that end inside the mesh, tiles
larger than the mesh, one to many planes per workgroup, every region
option and a batch of differently sized problems (the mixed form too: its
FP32 gradients are separately rounded as well).  With multiply-adds fused
the compiler may contract the recomputed border gradients differently from
the owner's: the fields then agree to a few ulps of their scale after 60
cycles (checked at 1e-11).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

STATE = ["x", "y", "z", "xd", "yd", "zd", "e", "p", "q", "v", "ss", "delv", "vdov", "arealg",
         "vnew", "ql", "qq"]
FORMS = {"exact": (0, 0), "fused": (0, 0x400), "mixed": (1, 0)}


def banner(msg):
    print(f"\n[elemmat] {msg}", flush=True)


def run_mode(pyl, s, nr, b, c, prec, opts, mode, em_planes, cycles):
    # (ql and qq, dead between cycles, are stored only with LULESH_OPT_NODE_FORCES)
    dev = pyl.Device.sedov(s, nr, b, c, precision=prec, opts=opts | pyl.OPT_NODE_FORCES)
    t = dev.tuning()
    t.material_mode = mode
    t.em_planes = em_planes
    dev.set_tuning(t)
    assert dev.tuning().material_mode == mode
    ts = dev.time_state()
    assert dev.run(cycles, ts) == 0
    out = {f: dev.download(f) for f in STATE}
    tstate = (ts.time, ts.deltatime, ts.dtcourant, ts.dthydro, ts.cycle)
    dev.close()
    return out, tstate


def max_rel(a, b):
    scale = max(float(np.max(np.abs(b))), 1e-300)
    return float(np.max(np.abs(a - b))) / scale


# (s, nr, b, c, em_planes): s=3 and s=6 lie inside one tile; s=17 and s=31
# end tiles inside the mesh; plane counts 1 (every plane a group), 2, 3, 7
# and 0 (sized by the library)
CASES = [(3, 1, 1, 1, 0), (6, 11, 1, 1, 1), (17, 16, 2, 1, 2), (31, 21, 1, 2, 3),
         (45, 11, 1, 1, 7), (45, 11, 1, 1, 0), (64, 21, 3, 0, 5)]


@pytest.mark.parametrize("form", ["exact", "fused", "mixed"])
@pytest.mark.parametrize("s,nr,b,c,em", CASES)
def test_fused_pass_equals_separate_kernels(pyl, s, nr, b, c, em, form):
    prec, opts = FORMS[form]
    banner(f"s={s} nr={nr} b={b} c={c} em_planes={em} {form}")
    cycles = 60
    ref, tref = run_mode(pyl, s, nr, b, c, prec, opts, 0, 0, cycles)
    got, tgot = run_mode(pyl, s, nr, b, c, prec, opts, 2, em, cycles)
    # (s=3 reaches its stop time at cycle 23)
    assert tgot[4] == tref[4] == min(cycles, 23 if s == 3 else cycles)
    if form == "exact" or form == "mixed":
        # separately rounded: bit for bit
        assert tgot == tref
        for f in STATE:
            np.testing.assert_array_equal(got[f], ref[f], err_msg=f)
    else:
        worst = max(max_rel(got[f], ref[f]) for f in STATE)
        print(f"[elemmat] worst field difference {worst:.2e}", flush=True)
        assert abs(tgot[0] - tref[0]) <= 1e-12 * tref[0]
        for f in STATE:
            assert max_rel(got[f], ref[f]) <= 1e-11, f


def test_fused_pass_in_a_batch(pyl):
    """A batch of nine problems (s = 5..40, regions 1..21) advanced by the
    fused pass: every problem's state and time state bit for bit those of
    the separate kernels on the same batch."""
    banner("batch of 9")
    sizes, regions, factors = pyl.batch_params(7, 9, 5, 40)
    res = []
    for mode in (0, 2):
        bt = pyl.Batch.sedov(sizes, regions, factors, opts=pyl.OPT_NODE_FORCES)
        t = bt.tuning()
        t.material_mode = mode
        t.em_planes = 3 if mode == 2 else 0
        bt.set_tuning(t)
        assert bt.tuning().material_mode == mode
        rc, states = bt.run(80)
        assert rc == 0
        res.append(([(st.time, st.deltatime, st.cycle) for st in states],
                    [{f: bt.download(k, f) for f in STATE} for k in range(len(sizes))]))
        bt.close()
    assert res[0][0] == res[1][0]
    for k in range(len(sizes)):
        for f in STATE:
            np.testing.assert_array_equal(res[1][1][k][f], res[0][1][k][f],
                                          err_msg=f"problem {k} {f}")


def test_gradients_after_a_fused_cycle(pyl):
    """The fused pass does not store the Q gradients (cycle temporaries in
    the reference, lulesh.cc:2117-2136): after a fused cycle their download is
    refused and download_view leaves them untouched; an elements phase stores
    them again."""
    banner("gradient downloads")
    dom = pyl.Domain(8)
    dev = pyl.Device(dom)
    t = dev.tuning()
    t.material_mode = 2
    dev.set_tuning(t)
    assert dev.tuning().material_mode == 2
    before = dev.download("delv_xi")
    ts = dom.time_state()
    assert dev.run(5, ts) == 0
    with pytest.raises(pyl.LuleshError):
        dev.download("delv_xi")
    host = dom.field("delx_eta").copy()
    dev.download_view()
    np.testing.assert_array_equal(dom.field("delx_eta"), host)
    assert dev.phase(pyl.PHASE_NODAL, ts.deltatime) == 0
    assert dev.phase(pyl.PHASE_ELEMENTS, ts.deltatime) == 0
    after = dev.download("delv_xi")
    assert after.shape == before.shape and np.any(after != 0)
    dev.close()
