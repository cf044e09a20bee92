"""Benchmark of the MI355X LULESH Lagrange leapfrog cycle.

Workload: plain FP64 Sedov blast at s=150 (3,375,000 elements, 151^3 nodes,
11 regions, -b 1 -c 1), the configuration the headline metric is quoted on.
A step is one cycle (device TimeIncrement + LagrangeLeapFrog) over the whole
mesh; the fields are resident in HBM before the timed region starts.

Every line runs in one arithmetic form (--arith): "fma" (default:
multiply-adds fused, as the reference's own GCC -O3 -mfma build) or "exact"
(every multiply and add rounded on its own: bitwise the oracle); the s=150
FP64 line of the other form is measured beside the headline, with its
origin-energy difference from it.

Also reported on the same GPU: the FP32-mixed form at s=150 (with its
origin-energy error against the exact FP64 form after the same cycles), the
s=250 / 21-region problem (FP64 and mixed), the 64-problem batch, per-kernel
HIP-event timings and the dominant kernel's roofline, a device-copy
bandwidth reference, and the CPU oracle (OpenMP port of the reference
arithmetic) timed on this host's cores.

Prints one JSON line.
"""
import argparse
import ctypes
import importlib.util
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "lulesh-task-based-hpx_amd")
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6  # MI355X FP64 (vector and matrix)


def load_pylulesh():
    spec = importlib.util.spec_from_file_location("pylulesh", os.path.join(PKG, "pylulesh.py"))
    mod = importlib.util.module_from_spec(spec)
    sys.modules["pylulesh"] = mod
    spec.loader.exec_module(mod)
    return mod


def model_bytes_per_zone_cycle(s):
    """Compulsory HBM bytes per zone-cycle of the minimal 4-pass schedule:
    B(s) = 816 + 204 (s+1)^3 / s^3."""
    return 816.0 + 204.0 * (s + 1) ** 3 / s ** 3


def kernel_bytes(name, E, N, mixed=False, material_mode=0, keep_dead=False):
    """Algorithmic HBM bytes of one launch of each kernel over the mesh
    (structured path): every array element counted once per launch, node
    arrays once per node.  keep_dead: the context also stores the fields
    that are dead between cycles (fx..zdd by the nodal kernel, ql qq by the
    material kernels; LULESH_OPT_NODE_FORCES)."""
    gr = 4 if mixed else 8
    dead_e = 16 if keep_dead else 0
    if name == "nodal":
        # z-march: node x..zd + nodalMass in, x..zd out; p q volo v ss elemMass in
        return E * 48 + N * (48 + 8 + 48 + (48 if keep_dead else 0))
    if name == "elements":
        # node x..zd in; volo v in; v vnew delv arealg vdov out; 6 gradients out
        return E * (16 + 40 + 6 * gr) + N * 48
    if name == "material":
        if material_mode == 0:
            # element order: repetition count (u16), elemBC, 6 gradients, vdov
            # elemMass volo vnew e delv p q arealg in; p e q ss out
            return E * (2 + 4 + 6 * gr + 72 + 32 + dead_e)
        # region lists: list, elemBC, 6 gradients + 6 neighbour gradients
        # (gathered), vdov elemMass volo vnew e delv p q arealg in; p e q ss out
        return E * (4 + 4 + 12 * gr + 72 + 32 + dead_e)
    if name == "elements_material":
        # the fused pass: node x..zd in; volo v in; v vnew delv arealg vdov out;
        # repetition count (u16), elemBC, elemMass e p q in; p e q ss out
        # (the Q gradients stay on chip)
        return E * (16 + 40 + 2 + 4 + 32 + 32 + dead_e) + N * 48
    if name == "time_increment":
        return 128
    return 0


# committed rocprofv3 summaries of the s=150 FP64 leg, per arithmetic form
TRAFFIC_FILES = {"fma": "profiles/hbm_traffic_sedov_s150_fp64.md",
                 "exact": "profiles/hbm_traffic_sedov_s150_fp64_exact.md"}
PROF_CYCLES = 64  # cycles of the per-kernel event timing pass


def measured(key, kernel, arith="fma"):
    """A per-launch figure of `kernel` at s=150 FP64 from the committed
    rocprofv3 summary of that arithmetic form: key "traffic" = HBM bytes
    (FETCH_SIZE + WRITE_SIZE, copy-calibrated), "valu_busy" = fraction of
    time the VALUs issue, "fp64_flop" = FP64 lane operations (FMA = 2).
    None if absent."""
    path = os.path.join(ROOT, TRAFFIC_FILES[arith])
    if not os.path.exists(path):
        return None
    for line in open(path):
        f = line.split()
        if len(f) == 3 and f[0] == key and f[1] == kernel:
            return float(f[2])
    return None


def cpu_baseline(s, nr, seconds, threads):
    """OpenMP oracle (the reference arithmetic) on the host: zone-cycles/s."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_lib import OracleDomain, set_threads

    set_threads(threads)
    d = OracleDomain(s, nr)
    d.run(1)  # cycle 1 outside the timing (first-touch of every array)
    cycles, t0 = 0, time.perf_counter()
    while time.perf_counter() - t0 < seconds:
        target = d.cycle + 2
        d.run(target)
        cycles += 2
    el = time.perf_counter() - t0
    d.close()
    return s ** 3 * cycles / el, cycles, el


def timed_run(pyl, dev, dom, warmup, steps, tune_cycles, material_mode=None):
    """Autotune (from a snapshot that is restored), warm up, then time
    `steps` cycles with the fields resident in HBM.  material_mode: the
    material kernel to use (library default if None)."""
    if material_mode is not None:
        t = dev.tuning()
        t.material_mode = material_mode
        dev.set_tuning(t)
        assert dev.tuning().material_mode == material_mode
    tuning = None
    if tune_cycles:
        t, _ = dev.autotune(tune_cycles)
        tuning = t.as_dict()
    ts = dom.time_state()
    if warmup:
        rc = dev.run(warmup, ts)
        assert rc == 0, rc
    t0 = time.perf_counter()
    rc = dev.run(ts.cycle + steps, ts)
    t1 = time.perf_counter()
    assert rc == 0, rc
    return t1 - t0, ts, tuning


def batch_bench(pyl, nprob, its, seed, tune_cycles, precision, streams=1, arith_opts=0):
    """The ensemble: nprob Sedov problems with seeded sizes in [20, 80],
    region counts in {11, 16, 21} and energy factors in [0.5, 2), each run
    for `its` cycles (or to its stop time) in one context, split over
    `streams` streams; the zone-cycles after one untimed graph replay are
    timed."""
    sizes, regions, factors = pyl.batch_params(seed, nprob)
    # built in HBM: bitwise the batch of host Domains (tests/test_gpu_setup.py)
    b = pyl.Batch.sedov(sizes, regions, factors, precision=precision,
                        opts=arith_opts | (pyl.OPT_STREAMS(streams) if streams > 1 else 0))
    tuning = None
    if tune_cycles:
        t, _ = b.autotune(tune_cycles)
        tuning = t.as_dict()
    # untimed warm-up: one replay of the cycle graph (its capture included)
    warm = b.tuning().cycles_per_graph
    rc, states0 = b.run(warm)
    assert rc == 0, rc
    t0 = time.perf_counter()
    rc, states = b.run(its)
    el = time.perf_counter() - t0
    assert rc == 0, rc
    zc = sum(int(s) ** 3 * (st.cycle - s0.cycle) for s, st, s0 in zip(sizes, states, states0))
    out = {"value": zc / el, "unit": "zone-cycles/s", "problems": nprob,
           "elements": int(sum(int(s) ** 3 for s in sizes)), "zone_cycles": zc,
           "elapsed_s": el, "device_ms": b.last_run_ms(),
           "cycles_min": min(st.cycle for st in states),
           "cycles_max": max(st.cycle for st in states), "warmup_cycles": warm, "seed": seed,
           "tuning": tuning,
           "streams": streams,
           "dtype": "f64" if precision == pyl.FP64 else "mixed"}
    b.close()
    return out


def copy_calibration(nbytes, reps):
    """A plain device-to-device copy of `nbytes` (torch's copy kernel), run
    `reps` times: the reference point that scales rocprofv3's FETCH_SIZE and
    WRITE_SIZE counters to bytes moved (each copy reads and writes nbytes)."""
    import torch

    src = torch.ones(nbytes // 4, dtype=torch.int32, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    for _ in range(reps):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes": nbytes, "reps": reps,
                      "note": "per copy dispatch: nbytes read, nbytes written"}))


def copy_bandwidth(pyl, nbytes=1 << 31, reps=30):
    """Bytes moved per second (read + write) by the library's hand-written
    streaming copy of `nbytes` (16-byte non-temporal loads and stores, the
    best of five launch shapes, HIP events): the practical HBM ceiling of a
    kernel that streams reads and writes, reported beside the 8 TB/s peak."""
    gbs = ctypes.c_double()
    rc = pyl.lib().lulesh_hip_stream_bandwidth(ctypes.c_int64(nbytes), reps, ctypes.byref(gbs))
    return gbs.value if rc == 0 and gbs.value > 0 else None


def fma_rate(pyl, reps=20):
    """FP64 TFLOP/s of the library's hand-written multiply-add issue kernel
    (independent v_fma_f64 chains, about 0.1 s): the practical FP64 vector
    ceiling, reported beside the 78.6 TFLOP/s peak."""
    tf = ctypes.c_double()
    rc = pyl.lib().lulesh_hip_fp64_fma_rate(reps, ctypes.byref(tf))
    return tf.value if rc == 0 and tf.value > 0 else None


def published_check(pyl, opts):
    """LULESH 2.0's documented run, -s 30 to its stop time (932 cycles,
    final origin energy 2.025075e+05), built in HBM and run on the GPU in the
    headline's arithmetic form: a check of the library on this box (about
    0.05 s).  It also runs every cycle kernel of the process once before the
    headline (the first cycles a process runs are a few % slower)."""
    d = pyl.Device.sedov(30, 11, opts=opts)
    ts = d.time_state()
    rc = d.run(10 ** 7, ts)
    e0 = float(d.report()[0])
    d.close()
    return {"rc": rc, "cycles": ts.cycle, "origin_energy": e0, "published_cycles": 932,
            "published_origin_energy": 2.025075e+05,
            "matches": rc == 0 and ts.cycle == 932 and f"{e0:.6e}" == "2.025075e+05"}


def measured_traffic_per_cycle(arith, names):
    """HBM bytes one s=150 FP64 cycle moves by the counters of the committed
    rocprofv3 summary (FETCH_SIZE + WRITE_SIZE of each kernel in `names`,
    calibrated), or None."""
    tot = 0.0
    for k in names:
        t = measured("traffic", k, arith)
        if t is None:
            return None
        tot += t
    return tot


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--s", type=int, default=150)
    ap.add_argument("--regions", type=int, default=11)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-large", action="store_true", help="skip the s=250 line")
    ap.add_argument("--arith", choices=("fma", "exact"), default="fma",
                    help="arithmetic form of every line: multiply-adds fused (as the "
                         "reference's GCC -O3 -mfma build) or every operation rounded on its "
                         "own (bitwise the oracle); the s=150 FP64 line of the other form "
                         "is reported beside the headline")
    ap.add_argument("--no-other-form", action="store_true",
                    help="skip the s=150 FP64 line of the other arithmetic form")
    ap.add_argument("--batch", type=int, default=64)
    ap.add_argument("--batch-its", type=int, default=1000)
    ap.add_argument("--batch-streams", type=int, default=2,
                    help="streams the batch is split over (1..15)")
    ap.add_argument("--material-mode", type=int, default=None, choices=(0, 1, 2),
                    help="material kernel of the s=150 legs: 0 element order (the default), "
                         "1 region lists, 2 fused with the elements phase")
    ap.add_argument("--tune-cycles", type=int, default=0,
                    help="autotune the s=150 legs first (0: the library's default launch shapes)")
    ap.add_argument("--batch-tune-cycles", type=int, default=6,
                    help="autotune the batch leg first (0: default launch shapes)")
    ap.add_argument("--no-copy-reference", action="store_true",
                    help="skip the 2 GiB device-copy bandwidth reference")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a 1 GiB device copy (counter calibration)")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration(1 << 30, 5)
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the multi-GPU "
              "decomposition is a design sketch, not built.", file=sys.stderr)
        sys.exit(2)

    pyl = load_pylulesh()
    s, nr = args.s, args.regions
    E = s ** 3
    N = (s + 1) ** 3

    # ---- FP64 headline -------------------------------------------------
    print(f"[bench] building Domain s={s}", file=sys.stderr, flush=True)
    dom = pyl.Domain(s, nr)
    t_up = time.perf_counter()
    AO = pyl.OPT_FMA if args.arith == "fma" else 0
    # allocation, upload of every field, layout checks; fx.. / xdd.. are not
    # stored (dead between cycles, nothing reads them)
    NF = pyl.OPT_NO_NODE_FORCES
    dev = pyl.Device(dom, pyl.FP64, AO | NF)
    t_up = time.perf_counter() - t_up
    # Before the headline's warm-up and timed cycles: the s=30 published-
    # result check (the first cycles a process runs are a few % slower than
    # later ones, whatever the problem), then the HBM and FP64 ceiling
    # references (about 0.1 s of streaming copies, then 0.1 s of
    # multiply-adds, the last straight before the warm-up): they bring the
    # GPU out of the idle clock state the host-side set-up left it in.  From
    # idle (even 50 ms of it) the first 20 cycles take about 10% longer than
    # the steady state, every kernel slower, and the next 20 about 3%; a
    # short window measured from idle times the clock ramp, not the cycle.
    s30 = None if args.no_copy_reference else published_check(pyl, AO)
    copy_gbs = None if args.no_copy_reference else copy_bandwidth(pyl)
    fma_tf = None if args.no_copy_reference else fma_rate(pyl)
    elapsed, ts, tuning = timed_run(pyl, dev, dom, args.warmup, args.steps, args.tune_cycles,
                                    args.material_mode)
    dev_ms = dev.last_run_ms()  # HIP events around the replayed graphs of the timed region
    cycles = args.steps
    zcps = E * cycles / elapsed
    print(f"[bench] fp64: {cycles} cycles in {elapsed:.3f}s -> {zcps:.4e} zc/s",
          file=sys.stderr, flush=True)

    # Per-kernel HIP-event timings: straight after the timed region (the GPU
    # still in its steady state), the next PROF_CYCLES cycles run with an
    # event pair around every kernel launch on the context's stream (the
    # kernels and tuning of the timed region, launched one by one, back to
    # back; the events are read when the stream next synchronises).
    dev.set_timing(True)
    rc = dev.run(ts.cycle + PROF_CYCLES, ts)
    assert rc == 0, rc
    kt = dev.kernel_times()
    dev.set_timing(False)
    prof_cycles = PROF_CYCLES
    t_down = time.perf_counter()
    dev.download_view()
    t_down = time.perf_counter() - t_down
    e0_fp64 = float(dom.field("e")[0])
    rep64 = dom.report()
    mode = dev.tuning().material_mode
    kernels = {}
    for name, (avg_ms, n) in kt.items():
        if n == 0:
            continue
        b = kernel_bytes(name, E, N, material_mode=mode)
        per_cycle = n / prof_cycles
        kernels[name] = {"avg_ms": round(avg_ms, 5), "launches_per_cycle": per_cycle,
                         "bytes_per_launch": b / max(per_cycle, 1),
                         "GBs": round((b / max(per_cycle, 1)) / (avg_ms * 1e-3) / 1e9, 1)}
    dom_k = max(kernels, key=lambda k: kernels[k]["avg_ms"] * kernels[k]["launches_per_cycle"])
    dk = kernels[dom_k]
    achieved = dk["bytes_per_launch"] / (dk["avg_ms"] * 1e-3) / 1e9
    prof = (s, nr) == (150, 11)
    traffic = measured("traffic", dom_k, args.arith) if prof else None
    valu_busy = measured("valu_busy", dom_k, args.arith) if prof else None
    fp64_flop = measured("fp64_flop", dom_k, args.arith) if prof else None
    fp64 = None
    if fp64_flop:
        tf = fp64_flop / (dk["avg_ms"] * 1e-3) / 1e12
        fp64 = {"flop_per_launch": int(fp64_flop), "achieved": round(tf, 2),
                "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": round(tf / FP64_PEAK_TFLOPS, 4),
                "measured_fma_ceiling": round(fma_tf, 2) if fma_tf else None,
                "frac_of_measured_ceiling": round(tf / fma_tf, 4) if fma_tf else None}
    ms_cycle_events = sum(k["avg_ms"] * k["launches_per_cycle"] for k in kernels.values())
    dev.close()

    # ---- the same FP64 line in the other arithmetic form ------------------
    other = None
    e0_exact = e0_fp64 if args.arith == "exact" else None
    if not args.no_other_form:
        oname = "exact" if args.arith == "fma" else "fma"
        domo = pyl.Domain(s, nr)
        devo = pyl.Device(domo, pyl.FP64, (pyl.OPT_FMA if oname == "fma" else 0) | NF)
        if not args.no_copy_reference:
            fma_rate(pyl)  # out of the idle clocks of the host set-up, as for the headline
        el_o, tso, _ = timed_run(pyl, devo, domo, args.warmup, args.steps, args.tune_cycles,
                                 args.material_mode)
        dev_ms_o = devo.last_run_ms()
        devo.set_timing(True)
        rc = devo.run(tso.cycle + PROF_CYCLES, tso)  # to the cycle the headline state is read at
        assert rc == 0, rc
        kto = devo.kernel_times()
        devo.set_timing(False)
        devo.download_view()
        e0o = float(domo.field("e")[0])
        repo = domo.report()
        if oname == "exact":
            e0_exact = e0o
        other = {"arith": oname, "value": E * args.steps / el_o,
                 "ms_per_step": 1e3 * el_o / args.steps,
                 "device_ms_per_step": dev_ms_o / args.steps,
                 "kernels_avg_ms": {k: round(a, 5) for k, (a, n) in kto.items() if n},
                 "origin_energy": e0o,
                 "origin_energy_rel_diff_vs_headline": abs(e0o - e0_fp64) / abs(e0_fp64),
                 "symmetry_max_rel_diff": float(repo[3]), "cycles": tso.cycle}
        devo.close()
        domo.close()
        print(f"[bench] fp64 {oname}: {other['value']:.4e} zc/s", file=sys.stderr, flush=True)

    # ---- FP32-mixed form -------------------------------------------------
    mixed = None
    if not args.no_mixed:
        domm = pyl.Domain(s, nr)
        devm = pyl.Device(domm, pyl.MIXED, AO | NF)
        if not args.no_copy_reference:
            fma_rate(pyl)  # out of the idle clocks of the host set-up, as for the headline
        el_m, tsm, _ = timed_run(pyl, devm, domm, args.warmup, args.steps, args.tune_cycles,
                                 args.material_mode)
        rc = devm.run(tsm.cycle + PROF_CYCLES, tsm)  # to the cycle the FP64 state is read at
        assert rc == 0, rc
        devm.download_view()
        e0m = float(domm.field("e")[0])
        repm = domm.report()
        e0ref = e0_exact if e0_exact is not None else e0_fp64
        mixed = {"value": E * args.steps / el_m, "ms_per_step": 1e3 * el_m / args.steps,
                 "arith": args.arith,
                 "origin_energy_rel_err_vs_fp64": abs(e0m - e0ref) / abs(e0ref),
                 "origin_energy_reference": "FP64 every operation rounded on its own "
                                            "(bitwise the oracle), same cycle"
                                            if e0_exact is not None else "FP64 headline form",
                 "symmetry_max_abs_diff": float(repm[1]), "symmetry_max_rel_diff": float(repm[3]),
                 "cycles": tsm.cycle}
        devm.close()
        domm.close()

    # ---- the large single problem: s=250, 21 regions, 500 cycles ----------
    large = None
    if not args.no_large:
        large = {"config": "s=250, 21 regions, -b 1 -c 1, 500 cycles from the start, "
                           "built in HBM (bitwise the host set-up)"}
        e0 = {}
        for name, prec in (("fp64", pyl.FP64), ("mixed", pyl.MIXED)):
            dl = pyl.Device.sedov(250, 21, precision=prec, opts=AO)
            tl = dl.time_state()
            rc = dl.run(10, tl)
            assert rc == 0, rc
            t0 = time.perf_counter()
            rc = dl.run(500, tl)
            el = time.perf_counter() - t0
            assert rc == 0 and tl.cycle == 500, rc
            rep = dl.report()
            e0[name] = rep[0]
            large[name] = {"value": 250 ** 3 * 490 / el, "ms_per_step": 1e3 * el / 490,
                           "symmetry_max_rel_diff": float(rep[3])}
            dl.close()
        large["mixed"]["origin_energy_rel_err_vs_fp64"] = abs(e0["mixed"] - e0["fp64"]) / abs(
            e0["fp64"])
        print(f"[bench] s=250: fp64 {large['fp64']['value']:.4e} zc/s, "
              f"mixed {large['mixed']['value']:.4e} zc/s", file=sys.stderr, flush=True)

    # ---- 64-problem batch -------------------------------------------------
    batch = None
    if not args.no_batch:
        print(f"[bench] batch of {args.batch}", file=sys.stderr, flush=True)
        batch = batch_bench(pyl, args.batch, args.batch_its, 0, args.batch_tune_cycles, pyl.FP64,
                            args.batch_streams, AO)
        print(f"[bench] batch: {batch['value']:.4e} zc/s", file=sys.stderr, flush=True)
        # the single s=150 problem over as many cycles as the batch runs, in
        # the same form, for the batch-vs-single comparison
        dl = pyl.Device.sedov(s, nr, precision=pyl.FP64, opts=AO)
        tl = dl.time_state()
        assert dl.run(args.warmup, tl) == 0
        c0 = tl.cycle
        t0 = time.perf_counter()
        assert dl.run(c0 + args.batch_its, tl) == 0
        el = time.perf_counter() - t0
        single = {"value": E * (tl.cycle - c0) / el, "cycles": tl.cycle - c0,
                  "device_ms_per_step": dl.last_run_ms() / (tl.cycle - c0)}
        dl.close()
        batch["single_s150_same_cycles"] = single
        batch["vs_single_s150"] = batch["value"] / single["value"]
        print(f"[bench] single s={s} over {single['cycles']} cycles: {single['value']:.4e} zc/s",
              file=sys.stderr, flush=True)

    # ---- CPU baseline ----------------------------------------------------
    cpu = None
    if not args.no_cpu:
        # every host core this process may run on, unless the environment
        # sets an OpenMP thread budget below that (the GPU box allots 16
        # threads per GPU); both figures go into the line
        avail = len(os.sched_getaffinity(0))
        budget = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
        threads = max(1, min(avail, budget) if budget > 0 else avail)
        cs = min(s, 150)
        v, ncyc, el = cpu_baseline(cs, nr, args.cpu_seconds, threads)
        cpu = {"value": v, "unit": "zone-cycles/s", "cores": threads, "kind": "port",
               "host_cpus_available": avail, "host_cpus_total": os.cpu_count(),
               "omp_num_threads_env": budget or None,
               "sample": f"OpenMP oracle, s={cs}, {nr} regions, {ncyc} cycles after cycle 1, "
                         f"{el:.1f}s wall, gcc -O3 -ffp-contract=off"}
        # the same port on one core, a shorter sample
        v1, ncyc1, el1 = cpu_baseline(cs, nr, max(2.0, args.cpu_seconds / 4), 1)
        cpu["serial"] = {"value": v1, "cores": 1,
                         "sample": f"serial oracle, s={cs}, {nr} regions, {ncyc1} cycles after "
                                   f"cycle 1, {el1:.1f}s wall"}

    bsz = model_bytes_per_zone_cycle(s)
    big = [k for k in kernels if k not in ("time_increment", "minima")]
    traffic_cycle = measured_traffic_per_cycle(args.arith, big) if prof else None
    out = {
        "metric": "zone-cycles/sec FP64 s=150 vs FP32-mixed & 64-problem batch; % HBM peak",
        "value": zcps,
        "unit": "zone-cycles/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * elapsed / cycles,
        "device_ms_per_step": dev_ms / cycles,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (deterministic Sedov initial conditions)",
        "config": {"workload": f"LULESH Sedov s={s} ({E} elements, {N} nodes), {nr} regions, "
                               "-b 1 -c 1, FP64, cycles from the start of the run",
                   "s": s, "regions": nr, "elements": E,
                   "arith": ("FP64 with multiply-adds fused (as the reference's GCC -O3 -mfma "
                             "build)" if args.arith == "fma" else
                             "FP64, every multiply and add rounded on its own (bitwise the "
                             "oracle)")},
        "roofline": {"bound": "hbm", "kernel": dom_k, "achieved": round(achieved, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": int(traffic) if traffic else None,
                     "valu_busy": valu_busy,
                     "fp64_vector": fp64,
                     "source_of_traffic_and_valu_busy": TRAFFIC_FILES[args.arith]
                     if traffic else None},
        "hbm_stream_copy_GBs": round(copy_gbs, 1) if copy_gbs else None,
        "fp64_fma_rate_TFLOPs": round(fma_tf, 2) if fma_tf else None,
        "before_warmup": ("the s=30 published-result check, then the HBM stream-copy and FP64 "
                          "multiply-add references (they bring the GPU out of idle clocks)"
                          if copy_gbs else "nothing (GPU idle during the host set-up)"),
        "s30_published_check": s30,
        "model_bytes_per_zone_cycle": round(bsz, 1),
        "model_hbm_frac": round(bsz * zcps / (HBM_PEAK_GBS * 1e9), 4),
        # bytes the kernels actually move per cycle (rocprofv3 counters of the
        # committed summary) over the measured time per cycle
        "measured_bytes_per_cycle": int(traffic_cycle) if traffic_cycle else None,
        "measured_hbm_frac": round(traffic_cycle / (elapsed / cycles) / (HBM_PEAK_GBS * 1e9), 4)
        if traffic_cycle else None,
        "kernels": kernels,
        "ms_per_cycle_kernel_sum": round(ms_cycle_events, 4),
        "value_incl_host_transfers": E * cycles / (elapsed + t_up + t_down),
        "host_transfer_s": {"create_and_upload": round(t_up, 4), "download": round(t_down, 4)},
        "origin_energy": e0_fp64,
        "fp64_other_form": other,
        "symmetry_max_abs_diff": float(rep64[1]),
        "symmetry_max_rel_diff": float(rep64[3]),
        "tuning": tuning,
        "mixed": mixed,
        "batch": batch,
        "large": large,
        "cpu_baseline": cpu,
        "speedup_vs_cpu": (zcps / cpu["value"]) if cpu else None,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
