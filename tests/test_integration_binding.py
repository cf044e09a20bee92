"""The C++ binding shown in the repository's integration guide (make_view:
a lulesh_domain_view built from a Domain's accessors) compiles and builds
the view the library expects.  The snippet is taken from the guide as
written and compiled
with g++ against this repository's host Domain, which keeps the reference
Domain's accessor names (lulesh.h:248-448); the program then compares the
view field by field with the view the library builds itself
(Domain::view)."""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

MAIN = r'''
#include <cstdio>
#include <cstring>
#include "domain.hpp"
using namespace lulesh;
%s
int main() {
    Domain d(10, 11, 1, 1);
    ViewTables t;
    lulesh_domain_view a = make_view(d, t), b = d.view();
    int bad = 0;
#define SAME(f) if (a.f != b.f) { std::printf("field %%s differs\n", #f); ++bad; }
    SAME(sizeX) SAME(numElem) SAME(numNode) SAME(numReg) SAME(cost)
    SAME(x) SAME(y) SAME(z) SAME(xd) SAME(yd) SAME(zd) SAME(xdd) SAME(ydd) SAME(zdd)
    SAME(fx) SAME(fy) SAME(fz) SAME(nodalMass) SAME(symmX) SAME(symmY) SAME(symmZ)
    SAME(nodelist) SAME(lxim) SAME(lxip) SAME(letam) SAME(letap) SAME(lzetam) SAME(lzetap)
    SAME(elemBC) SAME(e) SAME(p) SAME(q) SAME(ql) SAME(qq) SAME(v) SAME(volo) SAME(delv)
    SAME(vdov) SAME(arealg) SAME(ss) SAME(elemMass) SAME(vnew) SAME(regNumList)
    SAME(regElemSize) SAME(nodeElemCornerList)
    SAME(e_cut) SAME(p_cut) SAME(q_cut) SAME(v_cut) SAME(u_cut) SAME(hgcoef) SAME(ss4o3)
    SAME(qstop) SAME(monoq_max_slope) SAME(monoq_limiter_mult) SAME(qlc_monoq) SAME(qqc_monoq)
    SAME(qqc) SAME(eosvmax) SAME(eosvmin) SAME(pmin) SAME(emin) SAME(dvovmax) SAME(refdens)
    SAME(dtcourant) SAME(dthydro) SAME(dtfixed) SAME(time) SAME(deltatime)
    SAME(deltatimemultlb) SAME(deltatimemultub) SAME(dtmax) SAME(stoptime) SAME(cycle)
    for (int r = 0; r < a.numReg; ++r)
        if (a.regElemlist[r] != b.regElemlist[r]) { std::printf("regElemlist %%d\n", r); ++bad; }
    if (std::memcmp(a.nodeElemStart, b.nodeElemStart, sizeof(Index_t) * (a.numNode + 1)) != 0) {
        std::printf("nodeElemStart differs\n"); ++bad;
    }
    std::printf(bad ? "FAIL\n" : "OK\n");
    return bad;
}
'''


def test_integration_binding_compiles_and_matches(tmp_path):
    text = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    code = next(b for b in re.findall(r"```cpp\n(.*?)```", text, re.S) if "make_view" in b)
    # the declarations up to the end of make_view (the rest is the cycle loop)
    end = code.index("    return v;\n}\n") + len("    return v;\n}\n")
    decl = code[:end].replace('#include "lulesh_hip.h"', "")
    src = tmp_path / "binding.cc"
    src.write_text(MAIN % decl)
    exe = tmp_path / "binding"
    pkg = os.path.join(ROOT, "lulesh-task-based-hpx_amd")
    r = subprocess.run(["g++", "-std=c++17", "-O0", "-Wall", "-I", os.path.join(ROOT, "include"),
                        "-I", os.path.join(pkg, "host"), str(src),
                        os.path.join(pkg, "host", "domain.cc"), "-o", str(exe)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    out = subprocess.run([str(exe)], capture_output=True, text=True)
    assert out.returncode == 0 and out.stdout.strip().endswith("OK"), out.stdout
