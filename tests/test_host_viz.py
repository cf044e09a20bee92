"""The -v dump (DumpToVisit, lulesh-viz.cc:27-197) as VTK legacy files,
read back here with an independent parser: mesh, connectivity, regions,
zone fields e/p/v/q, node fields speed/xd/yd/zd, time and cycle, whole and
split into slabs (-f).  Host-only: no device is needed."""
import os
import re

import numpy as np
import pytest


def read_vtk(path):
    """Minimal reader of the binary legacy unstructured grid this build
    writes: returns dict of arrays and the header fields."""
    raw = open(path, "rb").read()
    pos = 0
    out = {"fields": {}, "cell": {}, "point": {}}

    def line():
        nonlocal pos
        end = raw.index(b"\n", pos)
        text = raw[pos:end].decode()
        pos = end + 1
        return text

    def take(dtype, count):
        nonlocal pos
        dt = np.dtype(dtype).newbyteorder(">")
        a = np.frombuffer(raw, dt, count, pos).astype(np.dtype(dtype))
        pos += dt.itemsize * count
        assert raw[pos:pos + 1] == b"\n"
        pos += 1
        return a

    assert line().startswith("# vtk DataFile Version")
    out["title"] = line()
    assert line() == "BINARY"
    assert line() == "DATASET UNSTRUCTURED_GRID"
    section = None
    types = {"float": "f4", "double": "f8", "int": "i4"}
    while pos < len(raw):
        t = line()
        if not t:
            continue
        w = t.split()
        if w[0] == "FIELD":
            for _ in range(int(w[2])):
                name, comps, n, ty = line().split()
                out["fields"][name] = take(types[ty], int(comps) * int(n))
        elif w[0] == "POINTS":
            out["points"] = take(types[w[2]], 3 * int(w[1])).reshape(-1, 3)
        elif w[0] == "CELLS":
            out["cells"] = take("i4", int(w[2])).reshape(int(w[1]), 9)
        elif w[0] == "CELL_TYPES":
            out["types"] = take("i4", int(w[1]))
        elif w[0] in ("CELL_DATA", "POINT_DATA"):
            section = "cell" if w[0] == "CELL_DATA" else "point"
            out[section + "_count"] = int(w[1])
        elif w[0] == "SCALARS":
            assert line() == "LOOKUP_TABLE default"
            n = out[section + "_count"]
            out[section][w[1]] = take(types[w[2]], n)
        else:
            raise AssertionError(f"unexpected line {t!r}")
    return out


def filled_domain(pyl, s, seed=3):
    dom = pyl.Domain(s, 11)
    rng = np.random.default_rng(seed)
    for f in ("x", "y", "z", "xd", "yd", "zd", "e", "p", "v", "q"):
        a = dom.field(f)
        a[:] = rng.standard_normal(a.size) * 10.0 ** rng.integers(-3, 4)
    ts = dom.time_state()
    ts.cycle, ts.time = 37, 1.25e-4
    dom.set_time_state(ts)
    return dom


def check_piece(pyl, dom, piece, k0, k1):
    s = dom.s
    n = s + 1
    e0, e1 = k0 * s * s, k1 * s * s
    g0, g1 = k0 * n * n, (k1 + 1) * n * n
    assert piece["fields"]["CYCLE"][0] == 37
    assert piece["fields"]["TIME"][0] == 1.25e-4
    xyz = np.stack([dom.field(f)[g0:g1] for f in ("x", "y", "z")], axis=1)
    np.testing.assert_array_equal(piece["points"], xyz.astype(np.float32))
    assert np.all(piece["cells"][:, 0] == 8) and np.all(piece["types"] == 12)
    conn = dom.field("nodelist").reshape(-1, 8)[e0:e1] - g0
    np.testing.assert_array_equal(piece["cells"][:, 1:], conn)
    np.testing.assert_array_equal(piece["cell"]["regions"], dom.field("regNumList")[e0:e1])
    for f in ("e", "p", "v", "q"):
        np.testing.assert_array_equal(piece["cell"][f], dom.field(f)[e0:e1].astype(np.float32))
    vel = [dom.field(f)[g0:g1].astype(np.float32) for f in ("xd", "yd", "zd")]
    for f, a in zip(("xd", "yd", "zd"), vel):
        np.testing.assert_array_equal(piece["point"][f], a)
    speed = np.sqrt(vel[0] * vel[0] + vel[1] * vel[1] + vel[2] * vel[2])
    np.testing.assert_allclose(piece["point"]["speed"], speed, rtol=1e-7)


def test_viz_single_file(pyl, tmp_path):
    dom = filled_domain(pyl, 3)
    dom.write_viz(str(tmp_path), 1)
    assert sorted(os.listdir(tmp_path)) == ["lulesh_plot_c37.vtk"]
    piece = read_vtk(tmp_path / "lulesh_plot_c37.vtk")
    assert piece["points"].shape == (64, 3) and piece["cells"].shape == (27, 9)
    check_piece(pyl, dom, piece, 0, 3)


@pytest.mark.parametrize("s,pieces", [(3, 2), (5, 3), (4, 9)])
def test_viz_split_into_slabs(pyl, tmp_path, s, pieces):
    dom = filled_domain(pyl, s, seed=s)
    dom.write_viz(str(tmp_path), pieces)
    used = min(pieces, s)
    listing = (tmp_path / "lulesh_plot_c37.visit").read_text().split("\n")
    assert listing[0] == f"!NBLOCKS {used}"
    names = [x for x in listing[1:] if x]
    assert names == [f"lulesh_plot_c37.{k:03d}.vtk" for k in range(used)]
    total = 0
    for k, name in enumerate(names):
        k0, k1 = s * k // used, s * (k + 1) // used
        piece = read_vtk(tmp_path / name)
        check_piece(pyl, dom, piece, k0, k1)
        total += piece["cells"].shape[0]
        assert re.search(r"cycle 37 ", piece["title"])
    assert total == s ** 3


def test_viz_bad_arguments(pyl, tmp_path):
    dom = pyl.Domain(2)
    L = pyl.lib()
    assert L.lulesh_domain_write_viz(None, str(tmp_path).encode(), 1) == pyl.INVALID_ARGUMENT
    assert L.lulesh_domain_write_viz(dom.h, str(tmp_path).encode(), 0) == pyl.INVALID_ARGUMENT
    assert L.lulesh_domain_write_viz(dom.h, str(tmp_path / "missing" / "dir").encode(), 1) == -4
