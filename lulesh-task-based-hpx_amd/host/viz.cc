// Visualization dump of the final state, the counterpart of the reference's
// DumpToVisit (lulesh-viz.cc:27-197, called after the cycle loop when -v is
// given, lulesh.cc:2397-2400).  The reference writes a Silo file; Silo is not
// available to this build, so the same objects go into VTK legacy
// unstructured-grid files, which VisIt (and ParaView) read directly:
//
//   mesh     node coordinates as float, hexahedra in the Domain's local node
//            order (the VTK hexahedron order: bottom face counter-clockwise,
//            then the top face), time and cycle as field data
//   regions  the region number of each element (regNumList, 1-based)
//   e p v q  zone-centred, float
//   speed xd yd zd   node-centred, float (speed = |velocity|, lulesh-viz.cc:165)
//
// File name: lulesh_plot_c<cycle>.vtk as the reference's basename
// (lulesh-viz.cc:35).  With pieces > 1 (the -f flag, "number of files to
// split the viz dump into") the element planes are split into that many
// slabs, lulesh_plot_c<cycle>.<k>.vtk, tied together by
// lulesh_plot_c<cycle>.visit (VisIt's multi-block list), as the reference's
// multi-block objects tie its per-rank blocks together (lulesh-viz.cc:193-290).
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "domain.hpp"

namespace lulesh {

namespace {

// VTK legacy binary data is big-endian.
template <typename T>
void put_be(std::vector<char> &out, T v) {
    char b[sizeof(T)];
    std::memcpy(b, &v, sizeof(T));
    for (size_t k = 0; k < sizeof(T); ++k) out.push_back(b[sizeof(T) - 1 - k]);
}

void put_text(std::vector<char> &out, const std::string &s) {
    out.insert(out.end(), s.begin(), s.end());
}

// One slab of element planes [k0, k1) of the structured cube.
int write_piece(const VizFields &f, const std::string &path, Index_t k0, Index_t k1) {
    const int64_t s = f.nx, n = s + 1, E = s * s * (k1 - k0), N = n * n * (k1 - k0 + 1);
    const int64_t e0 = int64_t(k0) * s * s, g0 = int64_t(k0) * n * n;
    std::vector<char> out;
    out.reserve(size_t(E) * 60 + size_t(N) * 32 + 4096);
    char title[256];
    std::snprintf(title, sizeof title, "LULESH cycle %d time %.17e elements %lld-%lld", f.cycle,
                  f.time, (long long)e0, (long long)(e0 + E - 1));
    put_text(out, "# vtk DataFile Version 3.0\n");
    put_text(out, std::string(title) + "\nBINARY\nDATASET UNSTRUCTURED_GRID\n");
    put_text(out, "FIELD FieldData 2\nTIME 1 1 double\n");
    put_be<double>(out, f.time);
    put_text(out, "\nCYCLE 1 1 int\n");
    put_be<int32_t>(out, f.cycle);
    put_text(out, "\nPOINTS " + std::to_string(N) + " float\n");
    for (int64_t g = g0; g < g0 + N; ++g) {
        put_be<float>(out, float(f.x[g]));
        put_be<float>(out, float(f.y[g]));
        put_be<float>(out, float(f.z[g]));
    }
    put_text(out, "\nCELLS " + std::to_string(E) + " " + std::to_string(9 * E) + "\n");
    for (int64_t e = e0; e < e0 + E; ++e) {
        put_be<int32_t>(out, 8);
        for (int a = 0; a < 8; ++a) {
            int64_t g;
            if (f.nodelist) {
                g = f.nodelist[8 * e + a];
            } else { // the reference's structured connectivity, lulesh-init.cc:236-244
                const int64_t k = e / (s * s), j = (e / s) % s, i = e % s;
                const int64_t b = (k * n + j) * n + i;
                const int64_t nd[8] = {b, b + 1, b + n + 1, b + n, b + n * n, b + n * n + 1,
                                       b + n * n + n + 1, b + n * n + n};
                g = nd[a];
            }
            if (g < g0 || g >= g0 + N) return -3; // not the structured cube
            put_be<int32_t>(out, int32_t(g - g0));
        }
    }
    put_text(out, "\nCELL_TYPES " + std::to_string(E) + "\n");
    for (int64_t e = 0; e < E; ++e) put_be<int32_t>(out, 12); // VTK_HEXAHEDRON
    put_text(out, "\nCELL_DATA " + std::to_string(E) + "\n");
    put_text(out, "SCALARS regions int 1\nLOOKUP_TABLE default\n");
    for (int64_t e = e0; e < e0 + E; ++e) put_be<int32_t>(out, f.regNumList ? f.regNumList[e] : 1);
    const Real_t *zone[4] = {f.e, f.p, f.v, f.q};
    const char *zname[4] = {"e", "p", "v", "q"};
    for (int k = 0; k < 4; ++k) {
        put_text(out, std::string("\nSCALARS ") + zname[k] + " float 1\nLOOKUP_TABLE default\n");
        for (int64_t e = e0; e < e0 + E; ++e) put_be<float>(out, float(zone[k][e]));
    }
    put_text(out, "\nPOINT_DATA " + std::to_string(N) + "\n");
    put_text(out, "SCALARS speed float 1\nLOOKUP_TABLE default\n");
    for (int64_t g = g0; g < g0 + N; ++g) {
        const float xd = float(f.xd[g]), yd = float(f.yd[g]), zd = float(f.zd[g]);
        put_be<float>(out, std::sqrt((xd * xd) + (yd * yd) + (zd * zd)));
    }
    const Real_t *node[3] = {f.xd, f.yd, f.zd};
    const char *nname[3] = {"xd", "yd", "zd"};
    for (int k = 0; k < 3; ++k) {
        put_text(out, std::string("\nSCALARS ") + nname[k] + " float 1\nLOOKUP_TABLE default\n");
        for (int64_t g = g0; g < g0 + N; ++g) put_be<float>(out, float(node[k][g]));
    }
    put_text(out, "\n");
    FILE *fp = std::fopen(path.c_str(), "wb");
    if (!fp) return -4;
    const size_t w = std::fwrite(out.data(), 1, out.size(), fp);
    const int rc = std::fclose(fp);
    return (w == out.size() && rc == 0) ? 0 : -4;
}

} // namespace

VizFields VizOf(Domain &d) {
    VizFields f;
    f.nx = d.sizeX();
    f.cycle = d.cycle();
    f.time = d.time();
    f.x = &d.x(0);
    f.y = &d.y(0);
    f.z = &d.z(0);
    f.xd = &d.xd(0);
    f.yd = &d.yd(0);
    f.zd = &d.zd(0);
    f.e = &d.e(0);
    f.p = &d.p(0);
    f.v = &d.v(0);
    f.q = &d.q(0);
    f.regNumList = &d.regNumList(0);
    f.nodelist = d.nodelist(0);
    return f;
}

int WriteViz(const VizFields &f, const char *dir, Int_t pieces) {
    if (f.nx < 1 || pieces < 1 || !f.x || !f.y || !f.z || !f.xd || !f.yd || !f.zd || !f.e ||
        !f.p || !f.v || !f.q)
        return -3;
    if (pieces > f.nx) pieces = f.nx; // at least one element plane per slab
    std::string base = dir && *dir ? std::string(dir) + "/" : std::string();
    base += "lulesh_plot_c" + std::to_string(f.cycle);
    if (pieces == 1) return write_piece(f, base + ".vtk", 0, f.nx);
    std::string list = "!NBLOCKS " + std::to_string(pieces) + "\n";
    for (Int_t k = 0; k < pieces; ++k) {
        const Index_t k0 = Index_t(int64_t(f.nx) * k / pieces);
        const Index_t k1 = Index_t(int64_t(f.nx) * (k + 1) / pieces);
        char suffix[32];
        std::snprintf(suffix, sizeof suffix, ".%03d.vtk", int(k));
        const std::string path = base + suffix;
        if (int rc = write_piece(f, path, k0, k1)) return rc;
        const size_t slash = path.find_last_of('/');
        list += (slash == std::string::npos ? path : path.substr(slash + 1)) + "\n";
    }
    FILE *fp = std::fopen((base + ".visit").c_str(), "w");
    if (!fp) return -4;
    const size_t w = std::fwrite(list.data(), 1, list.size(), fp);
    const int rc = std::fclose(fp);
    return (w == list.size() && rc == 0) ? 0 : -4;
}

} // namespace lulesh
