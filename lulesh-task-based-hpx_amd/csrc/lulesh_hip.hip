// MI355X (gfx950) implementation of the LULESH Lagrange leapfrog cycle and
// its C-ABI (include/lulesh_hip.h).
//
// One cycle of the reference (TimeIncrement, lulesh.cc:182-229, then
// LagrangeLeapFrogWithTasks, lulesh.cc:1977-2240) becomes a short sequence of
// stream-ordered kernels over device-resident SoA fields:
//
//   k_time_increment  one thread: new dt from the previous cycle's minima
//   k_force           per element: stress (IntegrateStress) and FB hourglass
//                     corner forces, written corner-major (coalesced)
//   k_nodes           per node: corner->node gather in the reference's
//                     ascending-element order (no atomics), acceleration,
//                     symmetry BCs, velocity and position
//   k_elems           per element: kinematics, volume update, Q gradients
//   k_material        per element of a region list: monotonic Q, EOS x rep,
//                     sound speed, and the dt minima (wave shuffles + one
//                     64-bit atomic min per block on the bit pattern of the
//                     positive double, which is exact and order-free)
//
// Cycles are captured in a HIP graph and replayed; every kernel reads an
// `active` flag written by k_time_increment, so cycles past the stop time or
// cycle limit are no-ops and the host only polls the time state between
// graph launches.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "lulesh_elem.hpp"
#include "lulesh_hip.h"
#include "lulesh_structured.hpp"

using namespace lulesh_dev;

namespace {

constexpr int kBlock = 256;

__device__ __forceinline__ unsigned long long dbits(double v) {
    return (unsigned long long)__double_as_longlong(v);
}
__host__ __device__ inline double bits_to_double(unsigned long long b) {
    double d;
    memcpy(&d, &b, sizeof d);
    return d;
}
inline unsigned long long host_dbits(double v) {
    unsigned long long b;
    memcpy(&b, &v, sizeof b);
    return b;
}

// ---------------------------------------------------------------------------
// TimeIncrement on the device, lulesh.cc:182-229.  The loop condition of
// lulesh.cc:2374-2375 decides whether this cycle runs at all.
// ---------------------------------------------------------------------------
__global__ void k_time_increment(DevState *st) {
    DevState s = *st;
    const bool run = !s.error && (s.time < s.stoptime) && (s.cycle < s.its);
    st->active = run ? 1 : 0;
    if (!run) return;
    double targetdt = s.stoptime - s.time;
    double deltatime = s.deltatime;
    if ((s.dtfixed <= 0.0) && (s.cycle != 0)) {
        const double olddt = deltatime;
        const double dtcourant = bits_to_double(s.dtc_bits);
        const double dthydro = bits_to_double(s.dth_bits);
        double gnewdt = 1.0e+20;
        if (dtcourant < gnewdt) gnewdt = dtcourant / 2.0;
        if (dthydro < gnewdt) gnewdt = dthydro * 2.0 / 3.0;
        double newdt = gnewdt;
        const double ratio = newdt / olddt;
        if (ratio >= 1.0) {
            if (ratio < s.deltatimemultlb)
                newdt = olddt;
            else if (ratio > s.deltatimemultub)
                newdt = olddt * s.deltatimemultub;
        }
        if (newdt > s.dtmax) newdt = s.dtmax;
        deltatime = newdt;
    }
    if ((targetdt > deltatime) && (targetdt < (4.0 * deltatime / 3.0)))
        targetdt = 2.0 * deltatime / 3.0;
    if (targetdt < deltatime) deltatime = targetdt;
    st->deltatime = deltatime;
    st->time = s.time + deltatime;
    st->cycle = s.cycle + 1;
    // dtcourant = dthydro = 1e20 before the constraints (lulesh.cc:2211-2212)
    st->dtc_bits = dbits(1.0e+20);
    st->dth_bits = dbits(1.0e+20);
}

// A cycle driven by a host-side TimeIncrement: take its dt, reset minima.
__global__ void k_begin_cycle(DevState *st, double deltatime, int reset_minima) {
    st->deltatime = deltatime;
    st->active = 1;
    st->error = 0;
    if (reset_minima) {
        st->dtc_bits = dbits(1.0e+20);
        st->dth_bits = dbits(1.0e+20);
    }
}

__device__ __forceinline__ void load8(const double *__restrict__ f, const int n[8], double o[8]) {
#pragma unroll
    for (int k = 0; k < 8; ++k) o[k] = f[n[k]];
}

__device__ __forceinline__ void load_nodelist(const int32_t *__restrict__ nodelist, int i,
                                              int n[8]) {
    const int4 a = reinterpret_cast<const int4 *>(nodelist)[2 * i];
    const int4 b = reinterpret_cast<const int4 *>(nodelist)[2 * i + 1];
    n[0] = a.x; n[1] = a.y; n[2] = a.z; n[3] = a.w;
    n[4] = b.x; n[5] = b.y; n[6] = b.z; n[7] = b.w;
}

struct NodeFields {
    double *x, *y, *z, *xd, *yd, *zd;
};

// ---------------------------------------------------------------------------
// Element corner forces: InitIntegrateStressForElemsTask (lulesh.cc:787-836)
// and CalcHourglassForElemsTask (lulesh.cc:864-1113).  Stress corners are
// FP64; hourglass corners are FP64 or, in the mixed form, FP32 computed from
// node-0-relative coordinates.
// ---------------------------------------------------------------------------
template <typename HG>
__global__ __launch_bounds__(kBlock) void k_force(
    int32_t E, const int32_t *__restrict__ nodelist, NodeFields nf,
    const double *__restrict__ p, const double *__restrict__ q, const double *__restrict__ volo,
    const double *__restrict__ v, const double *__restrict__ ss,
    const double *__restrict__ elemMass, double hgcoef, double *__restrict__ csx,
    double *__restrict__ csy, double *__restrict__ csz, HG *__restrict__ chx,
    HG *__restrict__ chy, HG *__restrict__ chz, DevState *st) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= E || !st->active) return;
    int n[8];
    load_nodelist(nodelist, i, n);
    double x[8], y[8], z[8];
    load8(nf.x, n, x);
    load8(nf.y, n, y);
    load8(nf.z, n, z);
    // stress part
    {
        const double det = shape_derivatives<false>(x, y, z, nullptr);
        if (det == 0) st->error = LULESH_VOLUME_ERROR; // lulesh.cc:819-822
        double B[3][8];
        node_normals(B, x, y, z);
        const double sig = -p[i] - q[i];
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const size_t o = size_t(k) * E + i;
            csx[o] = -(sig * B[0][k]);
            csy[o] = -(sig * B[1][k]);
            csz[o] = -(sig * B[2][k]);
        }
    }
    // hourglass part
    const double determ = volo[i] * v[i];
    if (determ <= 0.0) st->error = LULESH_VOLUME_ERROR; // lulesh.cc:906-910
    double xd[8], yd[8], zd[8];
    load8(nf.xd, n, xd);
    load8(nf.yd, n, yd);
    load8(nf.zd, n, zd);
    HG hgf[3][8];
    if constexpr (sizeof(HG) == 8) {
        const double coefficient = -hgcoef * 0.01 * ss[i] * elemMass[i] / cbrt_(determ);
        hourglass_force<double>(x, y, z, xd, yd, zd, determ, coefficient, hgf);
    } else {
        float xf[8], yf[8], zf[8], xdf[8], ydf[8], zdf[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            xf[k] = float(x[k] - x[0]);
            yf[k] = float(y[k] - y[0]);
            zf[k] = float(z[k] - z[0]);
            xdf[k] = float(xd[k]);
            ydf[k] = float(yd[k]);
            zdf[k] = float(zd[k]);
        }
        const float df = float(determ);
        const float coefficient = float(-hgcoef * 0.01) * float(ss[i]) * float(elemMass[i]) /
                                  cbrt_(df);
        hourglass_force<float>(xf, yf, zf, xdf, ydf, zdf, df, coefficient, hgf);
    }
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const size_t o = size_t(k) * E + i;
        chx[o] = hgf[0][k];
        chy[o] = hgf[1][k];
        chz[o] = hgf[2][k];
    }
}

// ---------------------------------------------------------------------------
// Nodes: combineVolumeForcesTaskFunc (lulesh.cc:838-862) in the reference's
// corner order with stress and hourglass terms interleaved,
// CalcAccelerationForNodesTask (1115-1123), the symmetry-plane BCs
// (566-588, as a per-node mask), CalcVelocityAndPositionForNodesTask
// (1125-1158).
// ---------------------------------------------------------------------------
template <typename HG>
__global__ __launch_bounds__(kBlock) void k_nodes(
    int32_t N, const int32_t *__restrict__ start, const int32_t *__restrict__ cornerMap,
    const double *__restrict__ csx, const double *__restrict__ csy,
    const double *__restrict__ csz, const HG *__restrict__ chx, const HG *__restrict__ chy,
    const HG *__restrict__ chz, const double *__restrict__ nodalMass,
    const uint8_t *__restrict__ bcMask, NodeFields nf, double *__restrict__ fx,
    double *__restrict__ fy, double *__restrict__ fz, double *__restrict__ xdd,
    double *__restrict__ ydd, double *__restrict__ zdd, double u_cut, DevState *st) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= N || !st->active) return;
    const double dt = st->deltatime;
    const int b = start[g], e = start[g + 1];
    double sx = 0.0, sy = 0.0, sz = 0.0;
    for (int c = b; c < e; ++c) {
        const int cm = cornerMap[c];
        sx += csx[cm];
        sx += double(chx[cm]);
        sy += csy[cm];
        sy += double(chy[cm]);
        sz += csz[cm];
        sz += double(chz[cm]);
    }
    fx[g] = sx;
    fy[g] = sy;
    fz[g] = sz;
    const double mass = nodalMass[g];
    double ax = sx / mass, ay = sy / mass, az = sz / mass;
    const uint8_t m = bcMask[g];
    if (m & 1) ax = 0.0;
    if (m & 2) ay = 0.0;
    if (m & 4) az = 0.0;
    xdd[g] = ax;
    ydd[g] = ay;
    zdd[g] = az;
    double vx = nf.xd[g] + ax * dt;
    if (fabs_(vx) < u_cut) vx = 0.0;
    double vy = nf.yd[g] + ay * dt;
    if (fabs_(vy) < u_cut) vy = 0.0;
    double vz = nf.zd[g] + az * dt;
    if (fabs_(vz) < u_cut) vz = 0.0;
    nf.xd[g] = vx;
    nf.yd[g] = vy;
    nf.zd[g] = vz;
    nf.x[g] = nf.x[g] + vx * dt;
    nf.y[g] = nf.y[g] + vy * dt;
    nf.z[g] = nf.z[g] + vz * dt;
}

// ---------------------------------------------------------------------------
// Elements: CalcKinematicsForElemsTask (lulesh.cc:1169-1275, including the
// volume check and UpdateVolumes) and CalcMonotonicQGradientsForElemsTask
// (lulesh.cc:1277-1418).  Gradients are FP64, or FP32 in the mixed form.
// ---------------------------------------------------------------------------
template <typename QG, bool STRUCT>
__global__ __launch_bounds__(kBlock) void k_elems(
    int32_t E, int32_t s, const int32_t *__restrict__ nodelist, NodeFields nf,
    const double *__restrict__ volo, double *__restrict__ v, double *__restrict__ vnew,
    double *__restrict__ delv, double *__restrict__ arealg, double *__restrict__ vdov,
    QG *__restrict__ delv_xi, QG *__restrict__ delv_eta, QG *__restrict__ delv_zeta,
    QG *__restrict__ delx_xi, QG *__restrict__ delx_eta, QG *__restrict__ delx_zeta,
    double v_cut, double eosvmin, double eosvmax, DevState *st) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= E || !st->active) return;
    const double deltaTime = st->deltatime;
    int n[8];
    if (STRUCT) {
        // element (k,j,i) -> its eight nodes, lulesh-init.cc:236-244
        const int nn = s + 1;
        const int k = i / (s * s), r = i - k * s * s, j = r / s, ci = r - j * s;
        const int b = (k * nn + j) * nn + ci;
        n[0] = b; n[1] = b + 1; n[2] = b + nn + 1; n[3] = b + nn;
        n[4] = b + nn * nn; n[5] = b + nn * nn + 1; n[6] = b + nn * nn + nn + 1;
        n[7] = b + nn * nn + nn;
    } else {
        load_nodelist(nodelist, i, n);
    }
    double x[8], y[8], z[8], xd[8], yd[8], zd[8];
    load8(nf.x, n, x);
    load8(nf.y, n, y);
    load8(nf.z, n, z);
    const double volume = elem_volume<double>(x, y, z);
    const double voloi = volo[i];
    const double rel = volume / voloi;
    const double vold = v[i];
    vnew[i] = rel;
    delv[i] = rel - vold;
    arealg[i] = characteristic_length(x, y, z, volume);
    load8(nf.xd, n, xd);
    load8(nf.yd, n, yd);
    load8(nf.zd, n, zd);
    // Q gradients on the end-of-step geometry (lulesh.cc:1349)
    {
        QG g[6];
        if constexpr (sizeof(QG) == 8) {
            q_gradients<double>(x, y, z, xd, yd, zd, voloi * rel, g);
        } else {
            float xf[8], yf[8], zf[8], xv[8], yv[8], zv[8];
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                xf[k] = float(x[k] - x[0]);
                yf[k] = float(y[k] - y[0]);
                zf[k] = float(z[k] - z[0]);
                xv[k] = float(xd[k]);
                yv[k] = float(yd[k]);
                zv[k] = float(zd[k]);
            }
            q_gradients<float>(xf, yf, zf, xv, yv, zv, float(voloi * rel), g);
        }
        delx_xi[i] = g[0];
        delx_eta[i] = g[1];
        delx_zeta[i] = g[2];
        delv_xi[i] = g[3];
        delv_eta[i] = g[4];
        delv_zeta[i] = g[5];
    }
    // Velocity divergence on the half-step geometry (lulesh.cc:1218-1239)
    const double dt2 = 0.5 * deltaTime;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        x[k] -= dt2 * xd[k];
        y[k] -= dt2 * yd[k];
        z[k] -= dt2 * zd[k];
    }
    double B[3][8];
    const double detJ = shape_derivatives<true>(x, y, z, B);
    vdov[i] = velocity_divergence(xd, yd, zd, B, detJ);
    // CheckAndUpdateVolumeForElems, lulesh.cc:1255-1270
    double vc = vold;
    if (eosvmin != 0.0 && vc < eosvmin) vc = eosvmin;
    if (eosvmax != 0.0 && vc > eosvmax) vc = eosvmax;
    if (vc < 0.0) st->error = LULESH_VOLUME_ERROR;
    v[i] = (fabs_(rel - 1.0) < v_cut) ? 1.0 : rel;
}

struct ElemFields {
    const int32_t *elemBC, *lxim, *lxip, *letam, *letap, *lzetam, *lzetap;
    const double *vdov, *elemMass, *volo, *vnew, *delv, *arealg;
    double *e, *p, *q, *ql, *qq, *ss;
};

// Minimum over a block of two doubles, then one atomic per block.
__device__ __forceinline__ void block_min_atomic(double c, double h, DevState *st) {
    __shared__ double sc[kBlock / 64], sh[kBlock / 64];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const double oc = __shfl_xor(c, off, 64);
        const double oh = __shfl_xor(h, off, 64);
        c = (oc < c) ? oc : c;
        h = (oh < h) ? oh : h;
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) {
        sc[wave] = c;
        sh[wave] = h;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < int(blockDim.x >> 6); ++w) {
            c = (sc[w] < c) ? sc[w] : c;
            h = (sh[w] < h) ? sh[w] : h;
        }
        atomicMin(&st->dtc_bits, dbits(c));
        atomicMin(&st->dth_bits, dbits(h));
    }
}

// ---------------------------------------------------------------------------
// Material: CalcMonotonicQRegionForElemsAndApplyInitTask (lulesh.cc:1420-1669),
// EvalEOSAllInOneTask repeated `rep` times (lulesh.cc:2190-2194; every
// repetition re-reads the same inputs, so it recomputes the same values),
// CalcSoundSpeedForElemsAndSaveTask (1873-1924), then the element's Courant
// and hydro constraints (1938-1974) folded into the block minimum.
// `list` holds the elements of regions sharing one repetition count.
// ---------------------------------------------------------------------------
template <typename QG>
__global__ __launch_bounds__(kBlock) void k_material(
    int32_t count, const int32_t *__restrict__ list, int32_t rep, ElemFields f,
    const QG *__restrict__ delv_xi, const QG *__restrict__ delv_eta,
    const QG *__restrict__ delv_zeta, const QG *__restrict__ delx_xi,
    const QG *__restrict__ delx_eta, const QG *__restrict__ delx_zeta, Params P,
    DevState *st) {
    if (!st->active) return; // uniform across the grid
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    double dtc = 1.7976931348623157e+308, dth = 1.7976931348623157e+308;
    if (t < count) {
        const int i = list[t];
        const int32_t bc = f.elemBC[i];
        const double sxi = double(delv_xi[i]), seta = double(delv_eta[i]),
                     szeta = double(delv_zeta[i]);
        auto nbr = [&](const QG *a, int32_t bits, int32_t comm, int32_t symm, const int32_t *l,
                       double self) {
            return (bits == comm || bits == 0) ? double(a[l[i]]) : (bits == symm ? self : 0.0);
        };
        const double phixi = mono_phi(
            sxi, nbr(delv_xi, bc & XI_M, XI_M_COMM, XI_M_SYMM, f.lxim, sxi),
            nbr(delv_xi, bc & XI_P, XI_P_COMM, XI_P_SYMM, f.lxip, sxi), P);
        const double phieta = mono_phi(
            seta, nbr(delv_eta, bc & ETA_M, ETA_M_COMM, ETA_M_SYMM, f.letam, seta),
            nbr(delv_eta, bc & ETA_P, ETA_P_COMM, ETA_P_SYMM, f.letap, seta), P);
        const double phizeta = mono_phi(
            szeta, nbr(delv_zeta, bc & ZETA_M, ZETA_M_COMM, ZETA_M_SYMM, f.lzetam, szeta),
            nbr(delv_zeta, bc & ZETA_P, ZETA_P_COMM, ZETA_P_SYMM, f.lzetap, szeta), P);
        const double dv[3] = {sxi, seta, szeta};
        const double dx[3] = {double(delx_xi[i]), double(delx_eta[i]), double(delx_zeta[i])};
        const double vdov = f.vdov[i], vnew = f.vnew[i];
        double qlin, qquad;
        q_terms(phixi, phieta, phizeta, vdov, dv, dx, f.elemMass[i], f.volo[i], vnew, P, qlin,
                qquad);
        f.qq[i] = qquad;
        f.ql[i] = qlin;
        // vnewc clamp, lulesh.cc:1654-1666
        double vnewc = vnew;
        if (P.eosvmin != 0.0 && vnewc < P.eosvmin) vnewc = P.eosvmin;
        if (P.eosvmax != 0.0 && vnewc > P.eosvmax) vnewc = P.eosvmax;
        const double e0 = f.e[i], dv0 = f.delv[i], p0 = f.p[i], q0 = f.q[i];
        double en = 0, pn = 0, qn = 0, ssn = 0;
        for (int r = 0; r < rep; ++r) {
            double ei = e0;
            opaque(ei);
            eos(ei, dv0, p0, q0, qquad, qlin, vnewc, P, en, pn, qn, ssn);
        }
        f.ss[i] = ssn;
        f.p[i] = pn;
        f.e[i] = en;
        f.q[i] = qn;
        constraints(ssn, vdov, f.arealg[i], P, dtc, dth);
    }
    block_min_atomic(dtc, dth, st);
}

// Stand-alone constraint pass over all elements (CalcConstraintForElemsTask,
// lulesh.cc:1938-1974): the same minima as the fused path, for phase-level
// checks on arbitrary inputs.
__global__ __launch_bounds__(kBlock) void k_constraints(int32_t E, const double *__restrict__ ss,
                                                        const double *__restrict__ vdov,
                                                        const double *__restrict__ arealg,
                                                        Params P, DevState *st) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double dtc = 1.7976931348623157e+308, dth = 1.7976931348623157e+308;
    if (i < E) constraints(ss[i], vdov[i], arealg[i], P, dtc, dth);
    block_min_atomic(dtc, dth, st);
}

template <typename T>
__global__ void k_widen(int32_t n, const T *__restrict__ src, double *__restrict__ dst) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = double(src[i]);
}

template <typename T>
__global__ void k_narrow(int32_t n, const double *__restrict__ src, T *__restrict__ dst) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = T(src[i]);
}

inline unsigned grid_for(int64_t n, int block) { return unsigned((n + block - 1) / block); }

} // namespace

// ---------------------------------------------------------------------------
// Context
// ---------------------------------------------------------------------------
struct KernelTimer {
    const char *name;
    double total_ms = 0;
    int64_t launches = 0;
};

struct lulesh_ctx {
    int precision = 0;
    int32_t s = 0, E = 0, N = 0, numReg = 0, cost = 0;
    Params P{};
    double hgcoef = 0;
    hipStream_t stream = nullptr;
    std::vector<void *> allocs;
    int64_t bytes = 0;
    double *f[LULESH_F_COUNT] = {};
    // mixed-form FP32 gradient storage
    float *g32[6] = {};
    int32_t *nodelist = nullptr, *lxim = nullptr, *lxip = nullptr, *letam = nullptr,
            *letap = nullptr, *lzetam = nullptr, *lzetap = nullptr, *elemBC = nullptr;
    int32_t *nodeStart = nullptr, *cornerMap = nullptr, *matList = nullptr;
    uint8_t *bcMask = nullptr;
    double *cs[3] = {};
    void *ch[3] = {};
    struct Segment {
        int32_t offset, count, rep;
    };
    std::vector<Segment> segs;
    // structured Sedov cube: fused z-march nodal kernel over a double-buffered
    // node state (x, y, z, xd, yd, zd); f[] points at the current half
    bool structured = false;
    bool store_forces = false; // also write fx.. and xdd.. (Domain fidelity)
    int zplanes = 12;          // node planes per z-march workgroup
    double *nodeBuf[2][6] = {};
    int parity = 0;
    DevState *st = nullptr;
    DevState *hst = nullptr; // pinned mirror
    DevState init{};
    hipGraphExec_t graph = nullptr;
    int graph_cycles = 0;
    int graph_parity = 0;
    bool timing = false;
    std::vector<KernelTimer> timers;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    std::string err;
};

namespace {

int hip_fail(lulesh_ctx *c, hipError_t e, const char *what) {
    if (c) c->err = std::string(what) + ": " + hipGetErrorString(e);
    return LULESH_HIP_ERROR - int(e);
}

#define HIP_TRY(ctx, call)                                                                   \
    do {                                                                                     \
        hipError_t e_ = (call);                                                              \
        if (e_ != hipSuccess) return hip_fail((ctx), e_, #call);                             \
    } while (0)

template <typename T>
int dalloc(lulesh_ctx *c, T **p, size_t n) {
    void *q = nullptr;
    const size_t bytes = std::max<size_t>(n, 1) * sizeof(T);
    HIP_TRY(c, hipMalloc(&q, bytes));
    c->allocs.push_back(q);
    c->bytes += int64_t(bytes);
    *p = static_cast<T *>(q);
    return 0;
}

template <typename T>
int upload(lulesh_ctx *c, T *dst, const T *src, size_t n) {
    if (!src || n == 0) return 0;
    HIP_TRY(c, hipMemcpy(dst, src, n * sizeof(T), hipMemcpyHostToDevice));
    return 0;
}

bool is_node_field(int id) { return id <= LULESH_F_NODALMASS; }
bool is_grad_field(int id) { return id >= LULESH_F_DELV_XI && id <= LULESH_F_DELX_ZETA; }

double *view_field(const lulesh_domain_view *v, int id) {
    double *const tab[LULESH_F_COUNT] = {
        v->x, v->y, v->z, v->xd, v->yd, v->zd, v->xdd, v->ydd, v->zdd, v->fx, v->fy, v->fz,
        v->nodalMass, v->e, v->p, v->q, v->ql, v->qq, v->v, v->volo, v->delv, v->vdov, v->arealg,
        v->ss, v->elemMass, v->vnew, v->delv_xi, v->delv_eta, v->delv_zeta, v->delx_xi,
        v->delx_eta, v->delx_zeta};
    return tab[id];
}

// Index of the FP32 gradient buffer for a field id (mixed form).
int grad_slot(int id) {
    switch (id) {
    case LULESH_F_DELV_XI: return 0;
    case LULESH_F_DELV_ETA: return 1;
    case LULESH_F_DELV_ZETA: return 2;
    case LULESH_F_DELX_XI: return 3;
    case LULESH_F_DELX_ETA: return 4;
    default: return 5;
    }
}

template <typename T> struct KernelSet;

// Launch one kernel, optionally bracketed by events for per-kernel timing.
template <typename F>
int launch(lulesh_ctx *c, int timer_id, F &&fn) {
    if (c->timing) HIP_TRY(c, hipEventRecord(c->ev0, c->stream));
    fn();
    HIP_TRY(c, hipGetLastError());
    if (c->timing) {
        HIP_TRY(c, hipEventRecord(c->ev1, c->stream));
        HIP_TRY(c, hipEventSynchronize(c->ev1));
        float ms = 0;
        HIP_TRY(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
        c->timers[size_t(timer_id)].total_ms += ms;
        c->timers[size_t(timer_id)].launches += 1;
    }
    return 0;
}

enum { T_TIME = 0, T_FORCE, T_NODES, T_ELEMS, T_MATERIAL, T_COUNT };

void set_node_parity(lulesh_ctx *c, int parity) {
    c->parity = parity;
    for (int k = 0; k < 6; ++k) c->f[LULESH_F_X + k] = c->nodeBuf[parity][k];
}

template <typename R>
int enqueue_nodal(lulesh_ctx *c) {
    if (c->structured) {
        const int cur = c->parity, nxt = cur ^ 1;
        double *const *a = c->nodeBuf[cur];
        double *const *b = c->nodeBuf[nxt];
        const NodeIn in{a[0], a[1], a[2], a[3], a[4], a[5]};
        const NodeOut out{b[0], b[1], b[2], b[3], b[4], b[5]};
        const NodeForceOut fo{c->f[LULESH_F_FX], c->f[LULESH_F_FY], c->f[LULESH_F_FZ],
                              c->f[LULESH_F_XDD], c->f[LULESH_F_YDD], c->f[LULESH_F_ZDD]};
        const int n = c->s + 1;
        const int tiles = (n + kTile - 1) / kTile;
        const dim3 grid(unsigned(tiles * tiles), unsigned((n + c->zplanes - 1) / c->zplanes));
        int rc = launch(c, T_FORCE, [&] {
            auto kern = c->store_forces ? k_nodal_zmarch<R, true> : k_nodal_zmarch<R, false>;
            kern<<<grid, kZmarchThreads, 0, c->stream>>>(
                c->s, c->zplanes, in, out, fo, c->f[LULESH_F_P], c->f[LULESH_F_Q],
                c->f[LULESH_F_VOLO], c->f[LULESH_F_V], c->f[LULESH_F_SS], c->f[LULESH_F_ELEMMASS],
                c->f[LULESH_F_NODALMASS], c->hgcoef, c->P.u_cut, c->st);
        });
        set_node_parity(c, nxt);
        return rc;
    }
    const NodeFields nf{c->f[LULESH_F_X], c->f[LULESH_F_Y], c->f[LULESH_F_Z],
                        c->f[LULESH_F_XD], c->f[LULESH_F_YD], c->f[LULESH_F_ZD]};
    int rc = launch(c, T_FORCE, [&] {
        k_force<R><<<grid_for(c->E, kBlock), kBlock, 0, c->stream>>>(
            c->E, c->nodelist, nf, c->f[LULESH_F_P], c->f[LULESH_F_Q], c->f[LULESH_F_VOLO],
            c->f[LULESH_F_V], c->f[LULESH_F_SS], c->f[LULESH_F_ELEMMASS], c->hgcoef, c->cs[0],
            c->cs[1], c->cs[2], static_cast<R *>(c->ch[0]), static_cast<R *>(c->ch[1]),
            static_cast<R *>(c->ch[2]), c->st);
    });
    if (rc) return rc;
    return launch(c, T_NODES, [&] {
        k_nodes<R><<<grid_for(c->N, kBlock), kBlock, 0, c->stream>>>(
            c->N, c->nodeStart, c->cornerMap, c->cs[0], c->cs[1], c->cs[2],
            static_cast<const R *>(c->ch[0]), static_cast<const R *>(c->ch[1]),
            static_cast<const R *>(c->ch[2]), c->f[LULESH_F_NODALMASS], c->bcMask, nf,
            c->f[LULESH_F_FX], c->f[LULESH_F_FY], c->f[LULESH_F_FZ], c->f[LULESH_F_XDD],
            c->f[LULESH_F_YDD], c->f[LULESH_F_ZDD], c->P.u_cut, c->st);
    });
}

template <typename R>
int enqueue_elements(lulesh_ctx *c) {
    const NodeFields nf{c->f[LULESH_F_X], c->f[LULESH_F_Y], c->f[LULESH_F_Z],
                        c->f[LULESH_F_XD], c->f[LULESH_F_YD], c->f[LULESH_F_ZD]};
    R *g[6];
    for (int k = 0; k < 6; ++k)
        g[k] = sizeof(R) == 8 ? reinterpret_cast<R *>(c->f[LULESH_F_DELV_XI + k])
                              : reinterpret_cast<R *>(c->g32[k]);
    return launch(c, T_ELEMS, [&] {
        auto kern = c->structured ? k_elems<R, true> : k_elems<R, false>;
        kern<<<grid_for(c->E, kBlock), kBlock, 0, c->stream>>>(
            c->E, c->s, c->nodelist, nf, c->f[LULESH_F_VOLO], c->f[LULESH_F_V], c->f[LULESH_F_VNEW],
            c->f[LULESH_F_DELV], c->f[LULESH_F_AREALG], c->f[LULESH_F_VDOV], g[0], g[1], g[2],
            g[3], g[4], g[5], c->P.v_cut, c->P.eosvmin, c->P.eosvmax, c->st);
    });
}

template <typename R>
int enqueue_material(lulesh_ctx *c) {
    const ElemFields ef{c->elemBC, c->lxim, c->lxip, c->letam, c->letap, c->lzetam, c->lzetap,
                        c->f[LULESH_F_VDOV], c->f[LULESH_F_ELEMMASS], c->f[LULESH_F_VOLO],
                        c->f[LULESH_F_VNEW], c->f[LULESH_F_DELV], c->f[LULESH_F_AREALG],
                        c->f[LULESH_F_E], c->f[LULESH_F_P], c->f[LULESH_F_Q], c->f[LULESH_F_QL],
                        c->f[LULESH_F_QQ], c->f[LULESH_F_SS]};
    const R *g[6];
    for (int k = 0; k < 6; ++k)
        g[k] = sizeof(R) == 8 ? reinterpret_cast<const R *>(c->f[LULESH_F_DELV_XI + k])
                              : reinterpret_cast<const R *>(c->g32[k]);
    for (const auto &sg : c->segs) {
        if (sg.count == 0) continue;
        int rc = launch(c, T_MATERIAL, [&] {
            k_material<R><<<grid_for(sg.count, kBlock), kBlock, 0, c->stream>>>(
                sg.count, c->matList + sg.offset, sg.rep, ef, g[0], g[1], g[2], g[3], g[4], g[5],
                c->P, c->st);
        });
        if (rc) return rc;
    }
    return 0;
}

int enqueue_cycle_body(lulesh_ctx *c) {
    int rc;
    if (c->precision == LULESH_PRECISION_MIXED) {
        if ((rc = enqueue_nodal<float>(c))) return rc;
        if ((rc = enqueue_elements<float>(c))) return rc;
        return enqueue_material<float>(c);
    }
    if ((rc = enqueue_nodal<double>(c))) return rc;
    if ((rc = enqueue_elements<double>(c))) return rc;
    return enqueue_material<double>(c);
}

int enqueue_cycle(lulesh_ctx *c) {
    int rc = launch(c, T_TIME, [&] { k_time_increment<<<1, 1, 0, c->stream>>>(c->st); });
    if (rc) return rc;
    return enqueue_cycle_body(c);
}

int build_graph(lulesh_ctx *c, int cycles) {
    // an even number of cycles returns the node double buffer to the parity
    // the graph was captured with
    cycles += cycles & 1;
    if (c->graph && c->graph_cycles == cycles && c->graph_parity == c->parity) return 0;
    if (c->graph) {
        (void)hipGraphExecDestroy(c->graph);
        c->graph = nullptr;
    }
    hipGraph_t g = nullptr;
    HIP_TRY(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    int rc = 0;
    for (int k = 0; k < cycles && rc == 0; ++k) rc = enqueue_cycle(c);
    hipError_t e = hipStreamEndCapture(c->stream, &g);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess) return hip_fail(c, e, "hipStreamEndCapture");
    e = hipGraphInstantiate(&c->graph, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) return hip_fail(c, e, "hipGraphInstantiate");
    c->graph_cycles = cycles;
    c->graph_parity = c->parity;
    return 0;
}

int read_state(lulesh_ctx *c) {
    HIP_TRY(c, hipMemcpyAsync(c->hst, c->st, sizeof(DevState), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return 0;
}

} // namespace

namespace {

// True when the Domain is the structured Sedov cube of the reference set-up
// (lulesh-init.cc:202-310, 448-468): canonical connectivity, corner lists in
// ascending element order, and the three symmetry sets are the x = 0, y = 0
// and z = 0 node planes.  Only then may the index-arithmetic kernels run.
bool is_structured_cube(const lulesh_domain_view *v) {
    const int64_t s = v->sizeX, n = s + 1, E = v->numElem, N = v->numNode;
    std::vector<int32_t> count(size_t(N) + 1, 0);
    for (int64_t e = 0; e < E; ++e) {
        const int64_t k = e / (s * s), j = (e / s) % s, i = e % s;
        const int64_t b = (k * n + j) * n + i;
        const int64_t nd[8] = {b, b + 1, b + n + 1, b + n, b + n * n, b + n * n + 1,
                               b + n * n + n + 1, b + n * n + n};
        for (int a = 0; a < 8; ++a) {
            if (v->nodelist[8 * e + a] != nd[a]) return false;
            ++count[size_t(nd[a]) + 1];
        }
    }
    for (int64_t g = 0; g < N; ++g) {
        count[size_t(g) + 1] += count[size_t(g)];
        if (v->nodeElemStart[g] != count[size_t(g)]) return false;
    }
    if (v->nodeElemStart[N] != 8 * E) return false;
    std::vector<int32_t> fill(count.begin(), count.end() - 1);
    for (int64_t c = 0; c < 8 * E; ++c) {
        const int32_t g = v->nodelist[c];
        if (v->nodeElemCornerList[fill[size_t(g)]++] != int32_t(c)) return false;
    }
    if (!v->symmX || !v->symmY || !v->symmZ) return false;
    const int64_t P = n * n;
    for (int axis = 0; axis < 3; ++axis) {
        const int32_t *set = axis == 0 ? v->symmX : axis == 1 ? v->symmY : v->symmZ;
        std::vector<int32_t> got(set, set + P), want;
        want.reserve(size_t(P));
        for (int64_t a = 0; a < n; ++a)
            for (int64_t b = 0; b < n; ++b)
                want.push_back(int32_t(axis == 0 ? (a * n + b) * n        // I == 0
                                      : axis == 1 ? (a * n) * n + b       // J == 0
                                                  : a * n + b));          // K == 0
        std::sort(got.begin(), got.end());
        std::sort(want.begin(), want.end());
        if (got != want) return false;
    }
    return true;
}

} // namespace

extern "C" {

int lulesh_hip_create(const lulesh_domain_view *v, int precision_opts, lulesh_ctx **out) {
    const int precision = precision_opts & 0xff;
    if (precision_opts & ~(0xff | LULESH_OPT_NODE_FORCES | LULESH_OPT_GENERIC))
        return LULESH_INVALID_ARGUMENT;
    if (!v || !out || v->sizeX < 1 || v->numElem != v->sizeX * v->sizeX * v->sizeX ||
        v->numNode != (v->sizeX + 1) * (v->sizeX + 1) * (v->sizeX + 1) || v->numReg < 1 ||
        (precision != LULESH_PRECISION_FP64 && precision != LULESH_PRECISION_MIXED) ||
        !v->nodelist || !v->nodeElemStart || !v->nodeElemCornerList || !v->regElemSize ||
        !v->regElemlist || !v->elemBC || !v->x || !v->e || !v->nodalMass || !v->volo ||
        !v->elemMass)
        return LULESH_INVALID_ARGUMENT;
    if (int64_t(v->numElem) * 8 >= int64_t(1) << 31) return LULESH_INVALID_ARGUMENT;
    *out = nullptr;
    auto *c = new lulesh_ctx();
    c->precision = precision;
    c->store_forces = (precision_opts & LULESH_OPT_NODE_FORCES) != 0;
    c->structured = !(precision_opts & LULESH_OPT_GENERIC) && is_structured_cube(v);
    c->s = v->sizeX;
    c->E = v->numElem;
    c->N = v->numNode;
    c->numReg = v->numReg;
    c->cost = v->cost;
    c->hgcoef = v->hgcoef;
    c->P = Params{v->e_cut, v->p_cut, v->q_cut, v->v_cut, v->u_cut, v->hgcoef,
                  v->monoq_max_slope, v->monoq_limiter_mult, v->qlc_monoq, v->qqc_monoq, v->qqc,
                  v->eosvmax, v->eosvmin, v->pmin, v->emin, v->dvovmax, v->refdens};
    c->timers = {{"time_increment"}, {"force"}, {"nodes"}, {"elements"}, {"material"}};
    const size_t E = size_t(c->E), N = size_t(c->N);
    auto fail = [&](int rc) {
        lulesh_hip_destroy(c);
        return rc;
    };
    int rc;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess)
        return fail(LULESH_HIP_ERROR);
    if (hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess)
        return fail(LULESH_HIP_ERROR);
    for (int id = 0; id < LULESH_F_COUNT; ++id) {
        const size_t n = is_node_field(id) ? N : E;
        if (precision == LULESH_PRECISION_MIXED && is_grad_field(id)) {
            if ((rc = dalloc(c, &c->g32[grad_slot(id)], n))) return fail(rc);
            continue;
        }
        if ((rc = dalloc(c, &c->f[id], n))) return fail(rc);
        if (hipMemset(c->f[id], 0, n * sizeof(double)) != hipSuccess) return fail(LULESH_HIP_ERROR);
        if ((rc = upload(c, c->f[id], (const double *)view_field(v, id), n))) return fail(rc);
    }
    for (int k = 0; k < 6; ++k) c->nodeBuf[0][k] = c->f[LULESH_F_X + k];
    if (c->structured) {
        for (int k = 0; k < 6; ++k)
            if ((rc = dalloc(c, &c->nodeBuf[1][k], N))) return fail(rc);
    } else if ((rc = dalloc(c, &c->nodelist, 8 * E)) ||
               (rc = upload(c, c->nodelist, v->nodelist, 8 * E))) {
        return fail(rc);
    }
    int32_t **conn[7] = {&c->lxim, &c->lxip, &c->letam, &c->letap, &c->lzetam, &c->lzetap,
                         &c->elemBC};
    const int32_t *hconn[7] = {v->lxim, v->lxip, v->letam, v->letap, v->lzetam, v->lzetap,
                               v->elemBC};
    for (int k = 0; k < 7; ++k) {
        if (!hconn[k]) return fail(LULESH_INVALID_ARGUMENT);
        if ((rc = dalloc(c, conn[k], E)) || (rc = upload(c, *conn[k], hconn[k], E)))
            return fail(rc);
    }
    // Corner CSR with corner ids remapped to the corner-major layout of the
    // force arrays: corner 8*e+k lives at k*E+e (indexed path only).
    if (!c->structured) {
        std::vector<int32_t> cmap(8 * E);
        for (size_t i = 0; i < 8 * E; ++i) {
            const int32_t cl = v->nodeElemCornerList[i];
            cmap[i] = (cl & 7) * c->E + (cl >> 3);
        }
        if ((rc = dalloc(c, &c->nodeStart, N + 1)) ||
            (rc = upload(c, c->nodeStart, v->nodeElemStart, N + 1)) ||
            (rc = dalloc(c, &c->cornerMap, 8 * E)) ||
            (rc = upload(c, c->cornerMap, cmap.data(), 8 * E)))
            return fail(rc);
    }
    // Symmetry-plane node sets as a per-node mask (bit0 x, bit1 y, bit2 z).
    if (!c->structured) {
        std::vector<uint8_t> mask(N, 0);
        const size_t P = size_t(c->s + 1) * (c->s + 1);
        const int32_t *sets[3] = {v->symmX, v->symmY, v->symmZ};
        for (int a = 0; a < 3; ++a)
            if (sets[a])
                for (size_t k = 0; k < P; ++k) mask[size_t(sets[a][k])] |= uint8_t(1u << a);
        if ((rc = dalloc(c, &c->bcMask, N)) || (rc = upload(c, c->bcMask, mask.data(), N)))
            return fail(rc);
    }
    // Region lists grouped by EOS repetition count (lulesh.cc:2159-2169);
    // within a group, elements ascend.
    {
        std::vector<int32_t> reps(size_t(c->numReg));
        for (int r = 0; r < c->numReg; ++r) reps[size_t(r)] = region_rep(r, c->numReg, c->cost);
        std::vector<int32_t> uniq(reps);
        std::sort(uniq.begin(), uniq.end());
        uniq.erase(std::unique(uniq.begin(), uniq.end()), uniq.end());
        std::vector<int32_t> all;
        all.reserve(E);
        for (int32_t rv : uniq) {
            const int32_t off = int32_t(all.size());
            for (int r = 0; r < c->numReg; ++r)
                if (reps[size_t(r)] == rv)
                    all.insert(all.end(), v->regElemlist[r], v->regElemlist[r] + v->regElemSize[r]);
            std::sort(all.begin() + off, all.end());
            c->segs.push_back({off, int32_t(all.size()) - off, rv});
        }
        if (all.size() != E) return fail(LULESH_INVALID_ARGUMENT);
        if ((rc = dalloc(c, &c->matList, E)) || (rc = upload(c, c->matList, all.data(), E)))
            return fail(rc);
    }
    // Corner force arrays (allocated once; the reference mallocs them every
    // cycle, lulesh.cc:2017-2022).
    for (int k = 0; k < 3 && !c->structured; ++k) {
        if ((rc = dalloc(c, &c->cs[k], 8 * E))) return fail(rc);
        if (precision == LULESH_PRECISION_MIXED) {
            float *p = nullptr;
            if ((rc = dalloc(c, &p, 8 * E))) return fail(rc);
            c->ch[k] = p;
        } else {
            double *p = nullptr;
            if ((rc = dalloc(c, &p, 8 * E))) return fail(rc);
            c->ch[k] = p;
        }
    }
    // Mixed form: narrow any provided gradients into FP32 storage.
    if (precision == LULESH_PRECISION_MIXED) {
        for (int id = LULESH_F_DELV_XI; id <= LULESH_F_DELX_ZETA; ++id) {
            const double *src = view_field(v, id);
            if (!src) continue;
            double *tmp = nullptr;
            if (hipMalloc(&tmp, E * sizeof(double)) != hipSuccess) return fail(LULESH_HIP_ERROR);
            (void)hipMemcpy(tmp, src, E * sizeof(double), hipMemcpyHostToDevice);
            k_narrow<float><<<grid_for(c->E, 256), 256, 0, c->stream>>>(c->E, tmp,
                                                                        c->g32[grad_slot(id)]);
            (void)hipStreamSynchronize(c->stream);
            (void)hipFree(tmp);
        }
    }
    if (c->structured) {
        // z extent of a z-march workgroup: enough workgroups to fill the chip
        // (about 4 per CU) while keeping the one-plane halo small.
        const int n = c->s + 1, tiles = (n + kTile - 1) / kTile;
        int zp = n;
        while (zp > 4 && int64_t(tiles) * tiles * ((n + zp - 1) / zp) < 1024) zp = (zp + 1) / 2;
        c->zplanes = std::max(zp, 1);
    }
    if ((rc = dalloc(c, &c->st, 1))) return fail(rc);
    if (hipHostMalloc(reinterpret_cast<void **>(&c->hst), sizeof(DevState)) != hipSuccess)
        return fail(LULESH_HIP_ERROR);
    DevState s{};
    s.time = v->time;
    s.deltatime = v->deltatime;
    s.stoptime = v->stoptime;
    s.dtfixed = v->dtfixed;
    s.deltatimemultlb = v->deltatimemultlb;
    s.deltatimemultub = v->deltatimemultub;
    s.dtmax = v->dtmax;
    s.dtc_bits = host_dbits(v->dtcourant);
    s.dth_bits = host_dbits(v->dthydro);
    s.cycle = v->cycle;
    s.its = v->cycle;
    s.active = 0;
    s.error = 0;
    c->init = s;
    if (hipMemcpy(c->st, &s, sizeof s, hipMemcpyHostToDevice) != hipSuccess)
        return fail(LULESH_HIP_ERROR);
    if (hipDeviceSynchronize() != hipSuccess) return fail(LULESH_HIP_ERROR);
    *out = c;
    return 0;
}

int lulesh_hip_leapfrog(lulesh_ctx *c, double deltatime, double *dtcourant, double *dthydro) {
    if (!c) return LULESH_INVALID_ARGUMENT;
    int rc = launch(c, T_TIME, [&] { k_begin_cycle<<<1, 1, 0, c->stream>>>(c->st, deltatime, 1); });
    if (rc) return rc;
    if ((rc = enqueue_cycle_body(c))) return rc;
    if ((rc = read_state(c))) return rc;
    if (c->hst->error) {
        c->err = "VolumeError";
        return c->hst->error;
    }
    if (dtcourant) *dtcourant = bits_to_double(c->hst->dtc_bits);
    if (dthydro) *dthydro = bits_to_double(c->hst->dth_bits);
    return 0;
}

int lulesh_hip_phase(lulesh_ctx *c, int phase, double deltatime) {
    if (!c || phase < LULESH_PHASE_NODAL || phase > LULESH_PHASE_CONSTRAINTS)
        return LULESH_INVALID_ARGUMENT;
    int rc = launch(c, T_TIME, [&] {
        k_begin_cycle<<<1, 1, 0, c->stream>>>(c->st, deltatime, phase >= LULESH_PHASE_MATERIAL);
    });
    if (rc) return rc;
    const bool mixed = c->precision == LULESH_PRECISION_MIXED;
    switch (phase) {
    case LULESH_PHASE_NODAL: rc = mixed ? enqueue_nodal<float>(c) : enqueue_nodal<double>(c); break;
    case LULESH_PHASE_ELEMENTS:
        rc = mixed ? enqueue_elements<float>(c) : enqueue_elements<double>(c);
        break;
    case LULESH_PHASE_MATERIAL:
        rc = mixed ? enqueue_material<float>(c) : enqueue_material<double>(c);
        break;
    default:
        k_constraints<<<grid_for(c->E, kBlock), kBlock, 0, c->stream>>>(
            c->E, c->f[LULESH_F_SS], c->f[LULESH_F_VDOV], c->f[LULESH_F_AREALG], c->P, c->st);
        HIP_TRY(c, hipGetLastError());
    }
    if (rc) return rc;
    if ((rc = read_state(c))) return rc;
    return c->hst->error;
}

int lulesh_hip_constraints_result(lulesh_ctx *c, double *dtcourant, double *dthydro) {
    if (!c) return LULESH_INVALID_ARGUMENT;
    int rc = read_state(c);
    if (rc) return rc;
    if (dtcourant) *dtcourant = bits_to_double(c->hst->dtc_bits);
    if (dthydro) *dthydro = bits_to_double(c->hst->dth_bits);
    return 0;
}

int lulesh_hip_run(lulesh_ctx *c, int32_t max_cycles, lulesh_time_state *io) {
    if (!c || !io) return LULESH_INVALID_ARGUMENT;
    DevState s = c->init;
    s.time = io->time;
    s.deltatime = io->deltatime;
    s.dtc_bits = host_dbits(io->dtcourant);
    s.dth_bits = host_dbits(io->dthydro);
    s.cycle = io->cycle;
    s.its = max_cycles;
    s.active = 1;
    s.error = 0;
    *c->hst = s;
    HIP_TRY(c, hipMemcpyAsync(c->st, c->hst, sizeof s, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    // Cycles per graph: enough to amortise launch and polling, few enough
    // that the no-op tail after the last cycle stays short.
    const int per = c->E >= 1000000 ? 8 : 32;
    int rc = 0;
    if (!c->timing && (rc = build_graph(c, per))) return rc;
    for (;;) {
        if (c->timing) {
            for (int k = 0; k < per && rc == 0; ++k) rc = enqueue_cycle(c);
            if (rc) return rc;
        } else {
            HIP_TRY(c, hipGraphLaunch(c->graph, c->stream));
        }
        if ((rc = read_state(c))) return rc;
        const DevState &h = *c->hst;
        if (h.error || !(h.time < h.stoptime && h.cycle < h.its)) break;
    }
    const DevState &h = *c->hst;
    io->time = h.time;
    io->deltatime = h.deltatime;
    io->dtcourant = bits_to_double(h.dtc_bits);
    io->dthydro = bits_to_double(h.dth_bits);
    io->cycle = h.cycle;
    io->status = h.error;
    if (h.error) c->err = "VolumeError";
    return h.error;
}

int lulesh_hip_download(lulesh_ctx *c, int id, void *dst, size_t count) {
    if (!c || !dst || id < 0 || id >= LULESH_F_COUNT) return LULESH_INVALID_ARGUMENT;
    const size_t n = is_node_field(id) ? size_t(c->N) : size_t(c->E);
    if (count > n) return LULESH_INVALID_ARGUMENT;
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    if (c->precision == LULESH_PRECISION_MIXED && is_grad_field(id)) {
        double *tmp = nullptr;
        HIP_TRY(c, hipMalloc(&tmp, n * sizeof(double)));
        k_widen<float><<<grid_for(int64_t(n), 256), 256, 0, c->stream>>>(int32_t(n),
                                                                         c->g32[grad_slot(id)], tmp);
        hipError_t e = hipMemcpyAsync(dst, tmp, count * sizeof(double), hipMemcpyDeviceToHost,
                                      c->stream);
        (void)hipStreamSynchronize(c->stream);
        (void)hipFree(tmp);
        if (e != hipSuccess) return hip_fail(c, e, "download");
        return 0;
    }
    HIP_TRY(c, hipMemcpyAsync(dst, c->f[id], count * sizeof(double), hipMemcpyDeviceToHost,
                              c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return 0;
}

int lulesh_hip_upload(lulesh_ctx *c, int id, const void *src, size_t count) {
    if (!c || !src || id < 0 || id >= LULESH_F_COUNT) return LULESH_INVALID_ARGUMENT;
    const size_t n = is_node_field(id) ? size_t(c->N) : size_t(c->E);
    if (count > n) return LULESH_INVALID_ARGUMENT;
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    if (c->precision == LULESH_PRECISION_MIXED && is_grad_field(id)) {
        double *tmp = nullptr;
        HIP_TRY(c, hipMalloc(&tmp, n * sizeof(double)));
        hipError_t e = hipMemcpy(tmp, src, count * sizeof(double), hipMemcpyHostToDevice);
        if (e == hipSuccess) {
            k_narrow<float><<<grid_for(int64_t(count), 256), 256, 0, c->stream>>>(
                int32_t(count), tmp, c->g32[grad_slot(id)]);
            e = hipStreamSynchronize(c->stream);
        }
        (void)hipFree(tmp);
        if (e != hipSuccess) return hip_fail(c, e, "upload");
        return 0;
    }
    HIP_TRY(c, hipMemcpy(c->f[id], src, count * sizeof(double), hipMemcpyHostToDevice));
    return 0;
}

int lulesh_hip_download_view(lulesh_ctx *c, lulesh_domain_view *dst) {
    if (!c || !dst) return LULESH_INVALID_ARGUMENT;
    for (int id = 0; id < LULESH_F_COUNT; ++id) {
        double *p = view_field(dst, id);
        if (!p) continue;
        int rc = lulesh_hip_download(c, id, p, is_node_field(id) ? size_t(c->N) : size_t(c->E));
        if (rc) return rc;
    }
    int rc = read_state(c);
    if (rc) return rc;
    dst->time = c->hst->time;
    dst->deltatime = c->hst->deltatime;
    dst->dtcourant = bits_to_double(c->hst->dtc_bits);
    dst->dthydro = bits_to_double(c->hst->dth_bits);
    dst->cycle = c->hst->cycle;
    return 0;
}

int lulesh_hip_set_timing(lulesh_ctx *c, int enable) {
    if (!c) return LULESH_INVALID_ARGUMENT;
    c->timing = enable != 0;
    for (auto &t : c->timers) {
        t.total_ms = 0;
        t.launches = 0;
    }
    return 0;
}

int lulesh_hip_kernel_times(lulesh_ctx *c, const char **names, double *avg_ms,
                            int64_t *launches, int cap) {
    if (!c) return LULESH_INVALID_ARGUMENT;
    const int n = int(c->timers.size());
    for (int k = 0; k < n && k < cap; ++k) {
        const auto &t = c->timers[size_t(k)];
        if (names) names[k] = t.name;
        if (avg_ms) avg_ms[k] = t.launches ? t.total_ms / double(t.launches) : 0.0;
        if (launches) launches[k] = t.launches;
    }
    return n;
}

int64_t lulesh_hip_device_bytes(const lulesh_ctx *c) { return c ? c->bytes : 0; }

const char *lulesh_hip_last_error(const lulesh_ctx *c) { return c ? c->err.c_str() : "null context"; }

void lulesh_hip_destroy(lulesh_ctx *c) {
    if (!c) return;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    if (c->graph) (void)hipGraphExecDestroy(c->graph);
    for (void *p : c->allocs) (void)hipFree(p);
    if (c->hst) (void)hipHostFree(c->hst);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

} // extern "C"
