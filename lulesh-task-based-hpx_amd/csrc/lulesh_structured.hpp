// Structured-mesh kernels for the Sedov cube built by the reference set-up
// (lulesh-init.cc:202-252): node (K,J,I) = (K*n + J)*n + I with n = s+1,
// element (k,j,i) = (k*s + j)*s + i with local nodes in the reference order.
// These kernels replace the nodelist / corner-CSR indirections by index
// arithmetic; the host only selects them after checking that the Domain's
// connectivity, corner lists and symmetry sets are exactly the structured
// ones.
//
// k_nodal_zmarch fuses the whole nodal phase (lulesh.cc:2017-2110) into one
// pass without storing corner forces in HBM.  A workgroup owns a tile of
// 15x15 node columns over a range of node planes and marches upward in z:
// for each element plane k its 16x16 threads compute the stress and
// hourglass corner forces of the 16x16 element patch covering the tile
// (one element of halo on the low x/y side), stage them in LDS, and the
// owning threads add them to their nodes in the reference's corner order
// (ascending element index, stress then hourglass per corner,
// lulesh.cc:849-857):
//   node plane k+1 starts with corners 6,7,5,4 of element plane k,
//   node plane k   ends   with corners 2,3,1,0 of element plane k,
// so the partial sums of a node plane live in registers for one step.  A
// finished node plane is advanced (acceleration, symmetry BCs, velocity,
// position) and written to the other half of a double buffer, since the
// neighbouring tiles still read the old positions.
#pragma once

#include "lulesh_elem.hpp"

namespace lulesh_dev {

constexpr int kPatch = 16;            // element patch edge (threads per tile edge)
constexpr int kTile = kPatch - 1;     // node tile edge owned by a workgroup
constexpr int kZmarchThreads = kPatch * kPatch;

struct NodeIn {
    const double *x, *y, *z, *xd, *yd, *zd;
};
struct NodeOut {
    double *x, *y, *z, *xd, *yd, *zd;
};
struct NodeForceOut { // optional copies of fx.. / xdd.. (Domain fidelity)
    double *fx, *fy, *fz, *xdd, *ydd, *zdd;
};

// One problem of a context (a context holds one Domain or a batch of
// independent Domains stored back to back).
struct ProbDesc {
    int32_t s, E, N;
    int32_t elemOff, nodeOff; // offsets into the concatenated element / node arrays
    int32_t zplanes, tiles;   // z-march: node planes per workgroup, node tiles per edge
    int32_t zBlock0, eBlock0; // first workgroup of this problem in the z-march / element grids
};

template <typename HG, bool STORE_F>
__global__ __launch_bounds__(kZmarchThreads) void k_nodal_zmarch(
    const ProbDesc *__restrict__ pds, const int32_t *__restrict__ blockProb, NodeBufs nb,
    NodeForceOut fo0, const double *__restrict__ p0,
    const double *__restrict__ q0, const double *__restrict__ volo0,
    const double *__restrict__ v0, const double *__restrict__ ss0,
    const double *__restrict__ elemMass0, const double *__restrict__ nodalMass0, double hgcoef,
    double u_cut, DevState *states) {
    const int prob = blockProb[blockIdx.x];
    DevState *st = states + prob;
    if (!st->active) return;
    const ProbDesc pd = pds[prob];
    const int s = pd.s, planes_per_block = pd.zplanes;
    const int eo = pd.elemOff, no = pd.nodeOff;
    const int cur = st->cur;
    double *const *bi = nb.b[cur];
    double *const *bo = nb.b[cur ^ 1];
    const NodeIn in{bi[0] + no, bi[1] + no, bi[2] + no, bi[3] + no, bi[4] + no, bi[5] + no};
    const NodeOut out{bo[0] + no, bo[1] + no, bo[2] + no, bo[3] + no, bo[4] + no, bo[5] + no};
    NodeForceOut fo = fo0;
    if (STORE_F) fo = {fo0.fx + no, fo0.fy + no, fo0.fz + no, fo0.xdd + no, fo0.ydd + no,
                       fo0.zdd + no};
    const double *__restrict__ p = p0 + eo;
    const double *__restrict__ q = q0 + eo;
    const double *__restrict__ volo = volo0 + eo;
    const double *__restrict__ v = v0 + eo;
    const double *__restrict__ ss = ss0 + eo;
    const double *__restrict__ elemMass = elemMass0 + eo;
    const double *__restrict__ nodalMass = nodalMass0 + no;
    const int lb = blockIdx.x - pd.zBlock0;
    // LDS: 4 corners x (3 stress + 3 hourglass) values per element of the patch
    __shared__ double lds[24][kZmarchThreads];

    const int n = s + 1;
    const int tid = threadIdx.x;
    const int jj = tid / kPatch, ii = tid % kPatch;
    const int tilesX = pd.tiles;
    const int tile = lb % (tilesX * tilesX);
    const int J0 = (tile / tilesX) * kTile, I0 = (tile % tilesX) * kTile;
    const int K0 = (lb / (tilesX * tilesX)) * planes_per_block;
    const int K1 = min(K0 + planes_per_block, n); // owned node planes [K0, K1)
    const int ej = J0 - 1 + jj, ei = I0 - 1 + ii; // this thread's element column
    const bool elemCol = (ej >= 0) && (ej < s) && (ei >= 0) && (ei < s);
    const int J = J0 + jj, I = I0 + ii; // this thread's node column
    const bool ownsNode = (jj < kTile) && (ii < kTile) && (J < n) && (I < n);
    // validity of the four elements around node (J,I) in a plane:
    // (J-1,I-1) (J-1,I) (J,I-1) (J,I)  <->  patch (jj,ii) (jj,ii+1) (jj+1,ii) (jj+1,ii+1)
    const bool v00 = (J >= 1) && (I >= 1), v01 = (J >= 1) && (I < s), v10 = (J < s) && (I >= 1),
               v11 = (J < s) && (I < s);
    const double dt = st->deltatime;

    double px = 0.0, py = 0.0, pz = 0.0; // partial sums of the node in the current plane

    auto finish_node = [&](int K) {
        const int g = (K * n + J) * n + I;
        const double mass = nodalMass[g];
        double ax = px / mass, ay = py / mass, az = pz / mass;
        if (STORE_F) {
            fo.fx[g] = px;
            fo.fy[g] = py;
            fo.fz[g] = pz;
        }
        if (I == 0) ax = 0.0; // symmX: the x = 0 plane
        if (J == 0) ay = 0.0; // symmY
        if (K == 0) az = 0.0; // symmZ
        if (STORE_F) {
            fo.xdd[g] = ax;
            fo.ydd[g] = ay;
            fo.zdd[g] = az;
        }
        double vx = in.xd[g] + ax * dt;
        if (fabs_(vx) < u_cut) vx = 0.0;
        double vy = in.yd[g] + ay * dt;
        if (fabs_(vy) < u_cut) vy = 0.0;
        double vz = in.zd[g] + az * dt;
        if (fabs_(vz) < u_cut) vz = 0.0;
        out.xd[g] = vx;
        out.yd[g] = vy;
        out.zd[g] = vz;
        out.x[g] = in.x[g] + vx * dt;
        out.y[g] = in.y[g] + vy * dt;
        out.z[g] = in.z[g] + vz * dt;
    };
    // add stress then hourglass of one corner staged at lds[4c.. ] of patch thread t
    auto add = [&](int c, int t) {
        px += lds[6 * c + 0][t];
        px += lds[6 * c + 3][t];
        py += lds[6 * c + 1][t];
        py += lds[6 * c + 4][t];
        pz += lds[6 * c + 2][t];
        pz += lds[6 * c + 5][t];
    };

    const int kbeg = max(K0 - 1, 0);
    const int kend = min(K1, s); // element planes [kbeg, kend)
    for (int k = kbeg; k < kend; ++k) {
        // ---- element (k, ej, ei): stress + hourglass corner forces ----------
        double up[24]; // corners 4..7: (sx, sy, sz, hx, hy, hz) each
        if (elemCol) {
            const int e = (k * s + ej) * s + ei;
            const int b = (k * n + ej) * n + ei;
            const int nd[8] = {b, b + 1, b + n + 1, b + n, b + n * n, b + n * n + 1,
                               b + n * n + n + 1, b + n * n + n};
            double x[8], y[8], z[8];
#pragma unroll
            for (int a = 0; a < 8; ++a) {
                x[a] = in.x[nd[a]];
                y[a] = in.y[nd[a]];
                z[a] = in.z[nd[a]];
            }
            double sc[3][8];
            {
                const double det = shape_derivatives<false>(x, y, z, nullptr);
                if (det == 0) st->error = kVolumeError;
                node_normals(sc, x, y, z);
                const double sig = -p[e] - q[e];
#pragma unroll
                for (int a = 0; a < 8; ++a) {
                    sc[0][a] = -(sig * sc[0][a]);
                    sc[1][a] = -(sig * sc[1][a]);
                    sc[2][a] = -(sig * sc[2][a]);
                }
            }
            const double determ = volo[e] * v[e];
            if (determ <= 0.0) st->error = kVolumeError;
            double xd[8], yd[8], zd[8];
#pragma unroll
            for (int a = 0; a < 8; ++a) {
                xd[a] = in.xd[nd[a]];
                yd[a] = in.yd[nd[a]];
                zd[a] = in.zd[nd[a]];
            }
            double hg[3][8];
            if constexpr (sizeof(HG) == 8) {
                const double coefficient = -hgcoef * 0.01 * ss[e] * elemMass[e] / cbrt_(determ);
                hourglass_force<double>(x, y, z, xd, yd, zd, determ, coefficient, hg);
            } else {
                float xf[8], yf[8], zf[8], xdf[8], ydf[8], zdf[8], hf[3][8];
#pragma unroll
                for (int a = 0; a < 8; ++a) {
                    xf[a] = float(x[a] - x[0]);
                    yf[a] = float(y[a] - y[0]);
                    zf[a] = float(z[a] - z[0]);
                    xdf[a] = float(xd[a]);
                    ydf[a] = float(yd[a]);
                    zdf[a] = float(zd[a]);
                }
                const float df = float(determ);
                const float coefficient =
                    float(-hgcoef * 0.01) * float(ss[e]) * float(elemMass[e]) / cbrt_(df);
                hourglass_force<float>(xf, yf, zf, xdf, ydf, zdf, df, coefficient, hf);
#pragma unroll
                for (int a = 0; a < 8; ++a) {
                    hg[0][a] = double(hf[0][a]);
                    hg[1][a] = double(hf[1][a]);
                    hg[2][a] = double(hf[2][a]);
                }
            }
#pragma unroll
            for (int c = 0; c < 4; ++c) {
#pragma unroll
                for (int d = 0; d < 3; ++d) {
                    lds[6 * c + d][tid] = sc[d][c];
                    lds[6 * c + 3 + d][tid] = hg[d][c];
                    up[6 * c + d] = sc[d][c + 4];
                    up[6 * c + 3 + d] = hg[d][c + 4];
                }
            }
        } else {
#pragma unroll
            for (int c = 0; c < 24; ++c) up[c] = 0.0;
        }
        __syncthreads();
        // ---- node plane k ends with corners 2,3,1,0 ----------------------------
        if (ownsNode && k >= K0) {
            if (v00) add(2, tid);
            if (v01) add(3, tid + 1);
            if (v10) add(1, tid + kPatch);
            if (v11) add(0, tid + kPatch + 1);
            finish_node(k);
        }
        __syncthreads();
        if (elemCol) {
#pragma unroll
            for (int c = 0; c < 24; ++c) lds[c][tid] = up[c];
        }
        __syncthreads();
        // ---- node plane k+1 starts with corners 6,7,5,4 -------------------------
        // lds now holds corners 4..7 at slots 0..3
        px = py = pz = 0.0;
        if (ownsNode && k + 1 < K1) {
            if (v00) add(2, tid);             // corner 6 of element (J-1, I-1)
            if (v01) add(3, tid + 1);         // corner 7 of element (J-1, I)
            if (v10) add(1, tid + kPatch);    // corner 5 of element (J, I-1)
            if (v11) add(0, tid + kPatch + 1); // corner 4 of element (J, I)
        }
        __syncthreads();
    }
    // The top node plane (K = s) has no element plane above it.
    if (ownsNode && K1 == n && kend == s && s >= K0) finish_node(s);
}

} // namespace lulesh_dev
