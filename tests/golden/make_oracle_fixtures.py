"""Writes the oracle fixtures of the long runs (tests/golden/oracle_*.json +
.npz): the CPU oracle (oracle/liblulesh_oracle.so, OpenMP; its results do
not depend on the thread count) run over the benchmark configurations at
their full iteration counts.

  s150_r11_i2000   s=150, 11 regions, -b 1 -c 1, 2000 cycles
  s150_r11_stop    the same problem run to its stop time (t = 1e-2, 5382 cycles)
  s250_r21_i500    s=250, 21 regions, -b 1 -c 1, 500 cycles
  batch64_p<k>     problem k of the seeded 64-problem ensemble (seed 0:
                   sizes in [20, 80], regions in {11, 16, 21}, deposit
                   scaled by a factor in [0.5, 2)), 1000 cycles or its stop
                   time, for k = 40 (the smallest mesh, 21 regions),
                   17 (the largest mesh) and 26 (21 regions, s = 52)

Each fixture holds the final time state, the report numbers
(VerifyAndWriteFinalOutput, lulesh-util.cc:60-115) with 17 significant
digits, and per state field: max |f|, sum f, sum |f|, sum f^2 and the values
at a fixed set of indices (every element / node of the cube next to the
origin, where the blast is, and a seeded random sample of the rest).

Usage (in the build container, a few tens of minutes on 8 cores; s150_r11_stop
about 15 minutes alone):
    python tests/golden/make_oracle_fixtures.py [name ...]
"""
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(HERE))

NODE_FIELDS = ["x", "y", "z", "xd", "yd", "zd"]
ELEM_FIELDS = ["e", "p", "q", "ql", "qq", "v", "ss", "delv", "vdov", "arealg", "vnew"]

# seeded 64-problem ensemble (seed 0): (k, s, regions, factor) of the members
# kept as fixtures; the test checks these against the library's own draw
BATCH_MEMBERS = (40, 17, 26)


def configs():
    out = {
        "s150_r11_i2000": dict(s=150, nr=11, balance=1, cost=1, its=2000, factor=1.0),
        "s150_r11_stop": dict(s=150, nr=11, balance=1, cost=1, its=10 ** 7, factor=1.0),
        "s250_r21_i500": dict(s=250, nr=21, balance=1, cost=1, its=500, factor=1.0),
    }
    sys.path.insert(0, os.path.join(ROOT, "lulesh-task-based-hpx_amd"))
    import pylulesh  # host-side helper only: the seeded ensemble parameters

    sizes, regions, factors = pylulesh.batch_params(0, 64)
    for k in BATCH_MEMBERS:
        out[f"batch64_p{k}"] = dict(s=int(sizes[k]), nr=int(regions[k]), balance=1, cost=1,
                                    its=1000, factor=float(factors[k]), member=k, seed=0)
    return out


def sample_indices(s, node):
    """Every element (node) of the m^3 cube at the origin, m = min(s, 8)
    (m + 1 for nodes), and 512 seeded random others."""
    n = s + 1 if node else s
    m = min(s, 8) + (1 if node else 0)
    k, j, i = np.meshgrid(np.arange(m), np.arange(m), np.arange(m), indexing="ij")
    corner = ((k * n + j) * n + i).ravel()
    total = n ** 3
    rng = np.random.default_rng(20240 + s + (1 if node else 0))
    rest = np.setdiff1d(np.arange(total), corner)
    pick = rng.choice(rest, size=min(512, rest.size), replace=False)
    return np.sort(np.concatenate([corner, pick])).astype(np.int64)


def field_stats(a):
    return {"max_abs": float(np.max(np.abs(a))), "sum": float(np.sum(a)),
            "sum_abs": float(np.sum(np.abs(a))), "sum_sq": float(np.sum(a * a))}


def make(name, cfg, threads):
    from oracle_lib import OracleDomain, set_threads

    set_threads(threads)
    t0 = time.time()
    d = OracleDomain(cfg["s"], cfg["nr"], cfg["balance"], cfg["cost"])
    if cfg["factor"] != 1.0:
        d.scale_energy(cfg["factor"])
    step = 100
    target = 0
    while target < cfg["its"]:
        target = min(target + step, cfg["its"])
        rc = d.run(target)
        assert rc == 0, rc
        print(f"[{name}] cycle {d.cycle} t={d.scalar('time'):.6e} "
              f"({time.time() - t0:.0f}s)", flush=True)
        if d.cycle < target:  # stop time reached
            break
    rep = d.report()
    nidx = sample_indices(cfg["s"], True)
    eidx = sample_indices(cfg["s"], False)
    meta = dict(cfg)
    meta.update({
        "cycle": int(d.cycle), "time": d.scalar("time"), "deltatime": d.scalar("deltatime"),
        "dtcourant": d.scalar("dtcourant"), "dthydro": d.scalar("dthydro"),
        "report": {"e0": float(rep[0]), "MaxAbsDiff": float(rep[1]),
                   "TotalAbsDiff": float(rep[2]), "MaxRelDiff": float(rep[3])},
        "stats": {f: field_stats(d.f64(f)) for f in NODE_FIELDS + ELEM_FIELDS},
        "oracle": "oracle/lulesh_oracle.c, gcc -O3 -ffp-contract=off, OpenMP",
        "oracle_seconds": round(time.time() - t0, 1),
    })
    arrays = {"node_index": nidx, "elem_index": eidx}
    for f in NODE_FIELDS:
        arrays[f] = d.f64(f)[nidx].copy()
    for f in ELEM_FIELDS:
        arrays[f] = d.f64(f)[eidx].copy()
    d.close()
    with open(os.path.join(HERE, f"oracle_{name}.json"), "w") as fh:
        json.dump(meta, fh, indent=1)
    np.savez_compressed(os.path.join(HERE, f"oracle_{name}.npz"), **arrays)
    print(f"[{name}] done: cycle {meta['cycle']}, e0 {meta['report']['e0']!r}", flush=True)


def main():
    cfgs = configs()
    names = sys.argv[1:] or list(cfgs)
    threads = len(os.sched_getaffinity(0))
    for name in names:
        make(name, cfgs[name], threads)


if __name__ == "__main__":
    main()
