// MI355X (gfx950) implementation of the LULESH Lagrange leapfrog cycle and
// its C-ABI (include/lulesh_hip.h).
//
// One cycle of the reference (TimeIncrement, lulesh.cc:182-229, then
// LagrangeLeapFrogWithTasks, lulesh.cc:1977-2240) is four stream-ordered
// kernel launches over device-resident SoA fields:
//
//   k_time_increment  one thread per problem: new dt from the previous
//                     cycle's minima, loop condition, minima reset
//   k_nodal_zmarch    the whole nodal phase fused (structured Sedov cube):
//                     stress + FB hourglass corner forces staged in LDS and
//                     gathered to nodes in the reference's corner order, then
//                     acceleration, symmetry BCs, velocity and position
//                     (lulesh_structured.hpp).  Meshes with any other
//                     connectivity use k_force + k_nodes (indexed gathers
//                     over nodelist and the corner CSR, corner forces in HBM).
//   k_elems           per element: kinematics, volume update, Q gradients
//   k_material        per element of the region lists, grouped by EOS
//                     repetition count: monotonic Q, EOS x rep, sound speed,
//                     and the element's dt constraints folded into a block
//                     minimum and one 64-bit atomic min per block on the bit
//                     pattern of the (positive) double -- exact, order-free.
//
// A context holds one problem or a batch of independent problems stored back
// to back; every launch covers all problems through block -> problem tables,
// and each problem keeps its own time state (DevState).  Cycles are captured
// in a HIP graph and replayed; kernels of a problem whose loop condition is
// false exit at once, so the host only polls the time states between graph
// launches.
//
// This file is compiled twice, into two arithmetic forms of the same code:
// LULESH_IMPL_NS = lulesh_impl_exact with -ffp-contract=off (every multiply
// and add rounded on its own, in the reference's operand order: bitwise the
// oracle) and LULESH_IMPL_NS = lulesh_impl_fma with
// -ffp-contract=fast-honor-pragmas (multiply-adds fused, as the reference's
// own GCC -O3 -mfma build does; the set-up keeps separate rounding).
// capi.cc implements the C-ABI and forwards each call to the form its
// context was created with.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <climits>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>
#include <type_traits>
#include <vector>

#include "domain.hpp"
#include "lulesh_elem.hpp"
#include "lulesh_hip.h"
#include "lulesh_structured.hpp"

#ifndef LULESH_IMPL_NS
#define LULESH_IMPL_NS lulesh_impl_exact
#endif
// 1 in the build whose multiply-adds are fused (LULESH_OPT_FMA contexts)
#ifndef LULESH_FUSED_FORM
#define LULESH_FUSED_FORM 0
#endif

namespace LULESH_IMPL_NS {

struct lulesh_ctx;
#include "lulesh_api.inc"

using namespace lulesh_dev;

namespace {

constexpr int kBlock = 256; // default and maximum block of the element kernels
constexpr bool kFusedForm = LULESH_FUSED_FORM != 0;

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

// The Courant / hydro minima of a cycle (lulesh.cc:1938-1974, 2204-2239) are
// reduced in two levels.  Each workgroup takes the minimum of its elements
// and merges it with an atomic min into one of kMinSlots slots of its problem
// (one 128-byte line each, so the atomics of different workgroups go to
// different addresses instead of queueing on one), and k_reduce_minima then
// folds the slots into the problem's DevState and re-arms them.  Minima are
// exact, so the grouping does not change the result.  The values are
// positive doubles, whose bit patterns order like the values.
constexpr int kMinSlots = 256;
constexpr int kSlotU64 = 16;
constexpr unsigned long long kMinInit = 0x4415AF1D78B58C40ull; // 1.0e+20

// Minimum over the kMinSlots slots of one problem, by a workgroup of
// kMinSlots threads; the slots are re-armed to 1e20.  The result is valid in
// thread 0.
__device__ __forceinline__ void fold_slots(unsigned long long *prob_slots,
                                           unsigned long long &c, unsigned long long &h) {
    unsigned long long *sl = prob_slots + size_t(threadIdx.x) * kSlotU64;
    c = sl[0];
    h = sl[1];
    sl[0] = kMinInit;
    sl[1] = kMinInit;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const unsigned long long oc = __shfl_xor(c, off, 64);
        const unsigned long long oh = __shfl_xor(h, off, 64);
        c = oc < c ? oc : c;
        h = oh < h ? oh : h;
    }
    __shared__ unsigned long long rc[kMinSlots / 64], rh[kMinSlots / 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) {
        rc[wave] = c;
        rh[wave] = h;
    }
    __syncthreads();
    if (threadIdx.x == 0)
        for (int w = 1; w < kMinSlots / 64; ++w) {
            c = rc[w] < c ? rc[w] : c;
            h = rh[w] < h ? rh[w] : h;
        }
}

// One workgroup per problem: if a cycle's minima are pending, dtcourant /
// dthydro = minimum over the problem's slots (used after the last cycle of a
// run and after host-driven cycles and phases).
__global__ __launch_bounds__(kMinSlots) void k_reduce_minima(DevState *states,
                                                             unsigned long long *slots) {
    DevState *st = states + blockIdx.x;
    if (!st->pending) return; // uniform across the workgroup
    unsigned long long c, h;
    fold_slots(slots + size_t(blockIdx.x) * kMinSlots * kSlotU64, c, h);
    if (threadIdx.x == 0) {
        st->dtc_bits = c;
        st->dth_bits = h;
        st->pending = 0;
    }
}

// ---------------------------------------------------------------------------
// TimeIncrement on the device, lulesh.cc:182-229, one workgroup per problem:
// the workgroup first folds the previous cycle's constraint minima out of
// the slots, then thread 0 steps the time.  The loop condition of
// lulesh.cc:2374-2375 decides whether the cycle runs.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kMinSlots) void k_time_increment(DevState *states,
                                                              unsigned long long *slots) {
    DevState *st = states + blockIdx.x;
    const bool fold = st->pending != 0; // uniform across the workgroup
    unsigned long long c = 0, h = 0;
    if (fold) fold_slots(slots + size_t(blockIdx.x) * kMinSlots * kSlotU64, c, h);
    if (threadIdx.x != 0) return;
    DevState s = *st;
    if (fold) {
        s.dtc_bits = st->dtc_bits = c;
        s.dth_bits = st->dth_bits = h;
    }
    if (s.fresh) st->cur = s.cur ^ 1; // adopt the node state the last cycle wrote
    const bool run = !s.error && (s.time < s.stoptime) && (s.cycle < s.its);
    st->active = run ? 1 : 0;
    st->pending = run ? 1 : 0;
    st->fresh = run ? 1 : 0; // this cycle's nodal phase writes the other half
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

// A cycle (or a nodal phase) driven by a host-side TimeIncrement: a nodal
// phase (`adopt`) first adopts the node state the previous nodal phase
// wrote; take the host's dt, reset minima.
__global__ void k_begin_cycle(DevState *st, double deltatime, int reset_minima, int adopt) {
    if (adopt) {
        if (st->fresh) st->cur ^= 1;
        st->fresh = 1;
    }
    st->deltatime = deltatime;
    st->active = 1;
    st->pending = reset_minima ? 1 : 0; // phases that produce minima
    st->error = 0;
    if (reset_minima) {
        st->dtc_bits = dbits(1.0e+20);
        st->dth_bits = dbits(1.0e+20);
    }
}

// The time state and cycle limit of lulesh_hip_run, set on the device in
// stream order (no host round trip); the node-buffer bookkeeping (cur,
// fresh, active) stays as the device has it.  The caller's minima replace
// any left in the slots.
__global__ void k_set_time_state(DevState *st, double time, double deltatime,
                                 unsigned long long dtc_bits, unsigned long long dth_bits,
                                 int32_t cycle, int32_t its) {
    st->time = time;
    st->deltatime = deltatime;
    st->dtc_bits = dtc_bits;
    st->dth_bits = dth_bits;
    st->pending = 0;
    st->cycle = cycle;
    st->its = its;
    st->error = 0;
}

// The cycle limit of every problem of a batch.
__global__ void k_set_its(DevState *st, int32_t nprob, int32_t its) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < nprob) st[p].its = its;
}

__device__ __forceinline__ void load8(const double *__restrict__ f, const int n[8], double o[8]) {
#pragma unroll
    for (int k = 0; k < 8; ++k) o[k] = f[n[k]];
}

// The eight node values of a structured element in four 16-byte loads: its
// nodes pair up along x (n1 = n0 + 1, n2 = n3 + 1, n5 = n4 + 1, n6 = n7 + 1).
// A pair is 8-byte aligned; gfx950 reads it with one dwordx4 load.
typedef double node_pair __attribute__((ext_vector_type(2)));
__device__ __forceinline__ void load8_pairs(const double *__restrict__ f, const int n[8],
                                            double o[8]) {
    node_pair p[4];
    __builtin_memcpy(&p[0], f + n[0], sizeof(node_pair));
    __builtin_memcpy(&p[1], f + n[3], sizeof(node_pair));
    __builtin_memcpy(&p[2], f + n[4], sizeof(node_pair));
    __builtin_memcpy(&p[3], f + n[7], sizeof(node_pair));
    o[0] = p[0].x; o[1] = p[0].y; o[3] = p[1].x; o[2] = p[1].y;
    o[4] = p[2].x; o[5] = p[2].y; o[7] = p[3].x; o[6] = p[3].y;
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
// Indexed path (any connectivity, one problem).  Element corner forces:
// InitIntegrateStressForElemsTask (lulesh.cc:787-836) and
// CalcHourglassForElemsTask (lulesh.cc:864-1113), written corner-major.
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
    {
        // (the scaled forms give the reference's values bit for bit)
        if (jacobian_det_x64(x, y, z) == 0) st->error = kVolumeError; // lulesh.cc:819-822
        double B[3][8];
        node_normals_x16(B, x, y, z);
        const double sig16 = (-p[i] - q[i]) * 0.0625;
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const size_t o = size_t(k) * E + i;
            csx[o] = -(sig16 * B[0][k]);
            csy[o] = -(sig16 * B[1][k]);
            csz[o] = -(sig16 * B[2][k]);
        }
    }
    const double determ = volo[i] * v[i];
    if (determ <= 0.0) st->error = kVolumeError; // lulesh.cc:906-910
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

// Indexed path, nodes: combineVolumeForcesTaskFunc (lulesh.cc:838-862) in the
// reference's corner order with stress and hourglass interleaved,
// CalcAccelerationForNodesTask (1115-1123), symmetry BCs (566-588, as a
// per-node mask), CalcVelocityAndPositionForNodesTask (1125-1158).
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
// STRUCT: node indices from the element's (k,j,i); else from nodelist.
// The end-of-step node state is the newest half (cur ^ fresh): the one the
// nodal phase of this cycle wrote, or the start state when no nodal phase
// ran since (a phase called on its own).
// NT: the results are stored with the non-temporal hint (stream_put).
// ---------------------------------------------------------------------------
// Q gradients of one element in the form's gradient type: FP64, or FP32 on
// coordinates taken relative to node 0 (their differences keep their
// significant bits) in the mixed form.  vol = volo * (V / volo), the
// reference's volo[i] * vnew[i] (lulesh.cc:1349).
template <typename QG>
__device__ __forceinline__ void element_q_gradients(const double x[8], const double y[8],
                                                    const double z[8], const double xd[8],
                                                    const double yd[8], const double zd[8],
                                                    double vol, QG g[6]) {
    if constexpr (sizeof(QG) == 8) {
        q_gradients<double>(x, y, z, xd, yd, zd, vol, g);
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
        q_gradients<float>(xf, yf, zf, xv, yv, zv, float(vol), g);
    }
}

struct ElemKin {
    const double *volo;
    double *v, *vnew, *delv, *arealg, *vdov;
};

template <typename QG, bool STRUCT, bool NT>
__global__ __launch_bounds__(kBlock) void k_elems(
    const ProbDesc *__restrict__ pds, const int32_t *__restrict__ blockProb,
    const int32_t *__restrict__ nodelist, NodeBufs nb, ElemKin f0, QG *__restrict__ dvxi,
    QG *__restrict__ dveta, QG *__restrict__ dvzeta, QG *__restrict__ dxxi,
    QG *__restrict__ dxeta, QG *__restrict__ dxzeta, double v_cut, double eosvmin,
    double eosvmax, DevState *states) {
    const int blk = xcd_block(blockIdx.x, gridDim.x);
    const int prob = blockProb[blk];
    DevState *st = states + prob;
    const ProbDesc pd = pds[prob];
    const int li = (blk - pd.eBlock0) * blockDim.x + threadIdx.x;
    if (li >= pd.E || !st->active) return;
    const int i = pd.elemOff + li; // element index in the concatenated arrays
    const double deltaTime = st->deltatime;
    double *const *nf = nb.b[st->cur ^ st->fresh];
    int n[8];
    if (STRUCT) {
        // element (k,j,i) -> its eight nodes, lulesh-init.cc:236-244
        const int s = pd.s, nn = s + 1;
        const int k = int((double(li) + 0.5) * pd.inv_ss), r = li - k * s * s;
        const int j = int((double(r) + 0.5) * pd.inv_s), ci = r - j * s;
        const int b = pd.nodeOff + (k * nn + j) * nn + ci;
        n[0] = b; n[1] = b + 1; n[2] = b + nn + 1; n[3] = b + nn;
        n[4] = b + nn * nn; n[5] = b + nn * nn + 1; n[6] = b + nn * nn + nn + 1;
        n[7] = b + nn * nn + nn;
    } else {
        load_nodelist(nodelist, li, n);
    }
    auto load8_node = [&](const double *f, const int nd[8], double o[8]) {
        if constexpr (STRUCT)
            load8_pairs(f, nd, o);
        else
            load8(f, nd, o);
    };
    double x[8], y[8], z[8], xd[8], yd[8], zd[8];
    // every load is issued before the first store (the stores could alias
    // the node arrays as far as the compiler knows)
    load8_node(nf[0], n, x);
    load8_node(nf[1], n, y);
    load8_node(nf[2], n, z);
    load8_node(nf[3], n, xd);
    load8_node(nf[4], n, yd);
    load8_node(nf[5], n, zd);
    const double voloi = f0.volo[i];
    const double vold = f0.v[i];
    const double dt2 = 0.5 * deltaTime;
    auto put_gradients = [&](const QG g[6]) {
        stream_put<NT>(&dxxi[i], g[0]);
        stream_put<NT>(&dxeta[i], g[1]);
        stream_put<NT>(&dxzeta[i], g[2]);
        stream_put<NT>(&dvxi[i], g[3]);
        stream_put<NT>(&dveta[i], g[4]);
        stream_put<NT>(&dvzeta[i], g[5]);
    };
    double rel;
    if constexpr (kFusedForm) {
        // multiply-add and mixed forms: the Q gradients and the velocity
        // divergence from three sums per coordinate and velocity component
        // (the velocities reduced to theirs at once, the positions after
        // the volume and the characteristic length): 744 -> 602 VALU
        // instructions and 150 -> 126 VGPRs (FP64), s=150 kinematics
        // 0.138 -> 0.135 ms
        double V[3][3], X[3][3];
        lin_sums<double>(xd, V[0]);
        lin_sums<double>(yd, V[1]);
        lin_sums<double>(zd, V[2]);
        const double volume = elem_volume<double>(x, y, z);
        rel = volume / voloi;
        stream_put<NT>(&f0.vnew[i], rel);
        stream_put<NT>(&f0.delv[i], rel - vold);
        stream_put<NT>(&f0.arealg[i], characteristic_length(x, y, z, volume));
        lin_sums<double>(x, X[0]);
        lin_sums<double>(y, X[1]);
        lin_sums<double>(z, X[2]);
        QG g[6];
        if constexpr (sizeof(QG) == 8) {
            q_gradients_sums<double>(X, V, voloi * rel, g);
        } else { // mixed form: the sums rounded to FP32 for the gradients
            float Xf[3][3], Vf[3][3];
#pragma unroll
            for (int c = 0; c < 3; ++c)
#pragma unroll
                for (int d = 0; d < 3; ++d) {
                    Xf[c][d] = float(X[c][d]);
                    Vf[c][d] = float(V[c][d]);
                }
            q_gradients_sums<float>(Xf, Vf, float(voloi * rel), g);
        }
        put_gradients(g);
        stream_put<NT>(&f0.vdov[i], vdov_sums(X, V, dt2));
    } else {
        const double volume = elem_volume<double>(x, y, z);
        rel = volume / voloi;
        stream_put<NT>(&f0.vnew[i], rel);
        stream_put<NT>(&f0.delv[i], rel - vold);
        stream_put<NT>(&f0.arealg[i], characteristic_length(x, y, z, volume));
        // Q gradients on the end-of-step geometry (lulesh.cc:1349)
        {
            QG g[6];
            element_q_gradients<QG>(x, y, z, xd, yd, zd, voloi * rel, g);
            put_gradients(g);
        }
        // Velocity divergence on the half-step geometry (lulesh.cc:1218-1239)
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            x[k] -= dt2 * xd[k];
            y[k] -= dt2 * yd[k];
            z[k] -= dt2 * zd[k];
        }
        double B[3][8];
        const double detJ = shape_derivatives<true>(x, y, z, B);
        stream_put<NT>(&f0.vdov[i], velocity_divergence(xd, yd, zd, B, detJ));
    }
    // CheckAndUpdateVolumeForElems, lulesh.cc:1255-1270
    double vc = vold;
    if (eosvmin != 0.0 && vc < eosvmin) vc = eosvmin;
    if (eosvmax != 0.0 && vc > eosvmax) vc = eosvmax;
    if (vc < 0.0) st->error = kVolumeError;
    stream_put<NT>(&f0.v[i], (fabs_(rel - 1.0) < v_cut) ? 1.0 : rel);
}

struct ElemFields {
    const int32_t *elemBC, *lxim, *lxip, *letam, *letap, *lzetam, *lzetap;
    const double *vdov, *elemMass, *volo, *vnew, *delv, *arealg;
    double *e, *p, *q, *ql, *qq, *ss;
};

// One segment of the material launch: the elements (concatenated indices,
// ascending) of one problem's regions that share a repetition count.
struct MatSeg {
    int32_t prob, rep, offset, count, block0;
};

__device__ __forceinline__ void block_min_atomic(double c, double h,
                                                 unsigned long long *prob_slots) {
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
        unsigned long long *sl = prob_slots + size_t(blockIdx.x % kMinSlots) * kSlotU64;
        atomicMin(sl, dbits(c));
        atomicMin(sl + 1, dbits(h));
    }
}


// ---------------------------------------------------------------------------
// Material: CalcMonotonicQRegionForElemsAndApplyInitTask (lulesh.cc:1420-1669),
// EvalEOSAllInOneTask repeated `rep` times (lulesh.cc:2190-2194; every
// repetition re-reads its inputs from the Domain, so all of them are treated
// as new values each time), CalcSoundSpeedForElemsAndSaveTask (1873-1924),
// then the element's Courant and hydro constraints (1938-1974) folded into
// the block minimum.
// ---------------------------------------------------------------------------
template <typename QG, bool RHO1, bool STRUCT>
__global__ __launch_bounds__(kBlock) void k_material(
    const ProbDesc *__restrict__ pds, const MatSeg *__restrict__ segs,
    const int32_t *__restrict__ blockSeg,
    const int32_t *__restrict__ list, ElemFields f, const QG *__restrict__ delv_xi,
    const QG *__restrict__ delv_eta, const QG *__restrict__ delv_zeta,
    const QG *__restrict__ delx_xi, const QG *__restrict__ delx_eta,
    const QG *__restrict__ delx_zeta, Params P, DevState *states,
    unsigned long long *__restrict__ minslots) {
    // (no XCD remap here: the costliest segments come first and must stay
    // spread over all XCDs)
    const int blk = blockIdx.x;
    const MatSeg sg = segs[blockSeg[blk]];
    DevState *st = states + sg.prob;
    if (!st->active) return; // uniform across the block
    const int t = (blk - sg.block0) * blockDim.x + threadIdx.x;
    double dtc = 1.7976931348623157e+308, dth = 1.7976931348623157e+308;
    if (t < sg.count) {
        const int i = list[sg.offset + t];
        const int32_t bc = f.elemBC[i];
        const double sxi = double(delv_xi[i]), seta = double(delv_eta[i]),
                     szeta = double(delv_zeta[i]);
        // On the structured cube the neighbour the switch reads is the
        // adjacent element (checked at create), so no index array is loaded.
        int ps = 0;
        if (STRUCT) ps = pds[sg.prob].s;
        auto nbr = [&](const QG *a, int32_t bits, int32_t comm, int32_t symm, const int32_t *l,
                       int step, double self) {
            if (bits == comm || bits == 0) return double(a[STRUCT ? i + step : l[i]]);
            return bits == symm ? self : 0.0;
        };
        const double phixi = mono_phi(
            sxi, nbr(delv_xi, bc & XI_M, XI_M_COMM, XI_M_SYMM, f.lxim, -1, sxi),
            nbr(delv_xi, bc & XI_P, XI_P_COMM, XI_P_SYMM, f.lxip, 1, sxi), P);
        const double phieta = mono_phi(
            seta, nbr(delv_eta, bc & ETA_M, ETA_M_COMM, ETA_M_SYMM, f.letam, -ps, seta),
            nbr(delv_eta, bc & ETA_P, ETA_P_COMM, ETA_P_SYMM, f.letap, ps, seta), P);
        const double phizeta = mono_phi(
            szeta, nbr(delv_zeta, bc & ZETA_M, ZETA_M_COMM, ZETA_M_SYMM, f.lzetam, -ps * ps, szeta),
            nbr(delv_zeta, bc & ZETA_P, ZETA_P_COMM, ZETA_P_SYMM, f.lzetap, ps * ps, szeta), P);
        const double dv[3] = {sxi, seta, szeta};
        const double dx[3] = {double(delx_xi[i]), double(delx_eta[i]), double(delx_zeta[i])};
        const double vdov = f.vdov[i], vnew = f.vnew[i];
        double qlin, qquad;
        q_terms(phixi, phieta, phizeta, vdov, dv, dx, f.elemMass[i], f.volo[i], vnew, P, qlin,
                qquad);
        if (f.qq) { // (null: not kept, dead between cycles)
            f.qq[i] = qquad;
            f.ql[i] = qlin;
        }
        // vnewc clamp, lulesh.cc:1654-1666
        double vnewc = vnew;
        if (P.eosvmin != 0.0 && vnewc < P.eosvmin) vnewc = P.eosvmin;
        if (P.eosvmax != 0.0 && vnewc > P.eosvmax) vnewc = P.eosvmax;
        const double e0 = f.e[i], dv0 = f.delv[i], p0 = f.p[i], q0 = f.q[i];
        double en = 0, pn = 0, qn = 0, ssn = 0;
        for (int r = 0; r < sg.rep; ++r) {
            double ei = e0, di = dv0, pi = p0, qi = q0, qqi = qquad, qli = qlin, vi = vnewc;
            opaque(ei);
            opaque(di);
            opaque(pi);
            opaque(qi);
            opaque(qqi);
            opaque(qli);
            opaque(vi);
            eos<RHO1>(ei, di, pi, qi, qqi, qli, vi, P, en, pn, qn, ssn);
        }
        f.ss[i] = ssn;
        f.p[i] = pn;
        f.e[i] = en;
        f.q[i] = qn;
        constraints(ssn, vdov, f.arealg[i], P, dtc, dth);
    }
    block_min_atomic(dtc, dth, minslots + size_t(sg.prob) * kMinSlots * kSlotU64);
}

constexpr int kMapWords = 16; // rounds of 64 evaluations per start map
constexpr int kMatSlots = 11; // LDS doubles per element: 7 EOS inputs, 4 outputs

// EvalEOSAllInOneTask x rep (lulesh.cc:2190-2194) of the 64 elements of a
// wave, dealt out over its lanes.  mat_lds holds kMatSlots doubles per
// thread of the workgroup: the 7 EOS inputs of its element at [0..6] on
// entry, the outputs e, p, q, ss of the element's last repetition at
// [7..10] on return.  smap: kMapWords start-map words per wave; sincl: one
// int per thread.  Every lane of the wave calls it (rep = 0 for lanes
// without an element).  The k-th start of the map belongs to the k-th lane
// with rep > 0: lane k itself when such lanes come first in the wave
// (COMPACT false), else looked up in lanes (one int per thread, COMPACT).
template <bool RHO1, bool COMPACT>
__device__ __forceinline__ void eos_wave_balanced(double *mat_lds, unsigned long long *smap,
                                                  int32_t *sincl, int32_t *lanes, int tid,
                                                  int rep, const Params &P) {
    const int lane = tid & 63, wbase = tid & ~63;
    if constexpr (COMPACT) {
        // lanes[wbase + k] = the lane of the wave's k-th element with rep > 0
        const unsigned long long act = __ballot(rep > 0);
        const int rank = __builtin_amdgcn_mbcnt_hi(uint32_t(act >> 32),
                                                   __builtin_amdgcn_mbcnt_lo(uint32_t(act), 0));
        if (rep > 0) lanes[wbase + rank] = lane;
    }
    // exclusive prefix sum of the repetition counts over the wave
    int incl = rep;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int v = __shfl_up(incl, o, 64);
        if (lane >= o) incl += v;
    }
    const int total = __builtin_amdgcn_readlane(incl, 63);
    const int excl = incl - rep;
    sincl[tid] = incl;
    // Evaluation w of the wave belongs to the element whose range
    // [excl, excl + rep) holds it.  Bit b of word k of the wave's start map
    // marks an element whose first evaluation is k*64 + b, so the owner of
    // evaluation r*64 + L is (starts before word r) + (starts in word r at
    // bits <= L) - 1.  The map covers kMapWords rounds at a time.
    unsigned long long *wmap = smap + (tid >> 6) * kMapWords;
    int started = 0; // elements whose first evaluation precedes the current round
    for (int c0 = 0; c0 < total; c0 += 64 * kMapWords) {
        if (lane < kMapWords) wmap[lane] = 0;
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
        if (rep > 0 && excl >= c0 && excl < c0 + 64 * kMapWords)
            atomicOr(&wmap[(excl - c0) >> 6], 1ull << (excl & 63));
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
        const int nev = min(total - c0, 64 * kMapWords);
        for (int r = 0; r * 64 < nev; ++r) {
            const unsigned long long m = wmap[r];
            const uint32_t mlo = __builtin_amdgcn_readfirstlane(uint32_t(m));
            const uint32_t mhi = __builtin_amdgcn_readfirstlane(uint32_t(m >> 32));
            const int below = __builtin_amdgcn_mbcnt_hi(mhi, __builtin_amdgcn_mbcnt_lo(mlo, 0));
            const int self = int((lane < 32 ? mlo >> lane : mhi >> (lane - 32)) & 1u);
            if (r * 64 + lane < nev) {
                const int k = started + below + self - 1; // the owner's rank
                const int j = wbase + (COMPACT ? lanes[wbase + k] : k);
                double en, pn, qn, ssn;
                double *el = mat_lds + kMatSlots * j;
                eos<RHO1>(el[0], el[1], el[2], el[3], el[4], el[5], el[6], P, en, pn, qn, ssn);
                if (c0 + r * 64 + lane == sincl[j] - 1) { // the element's last repetition
                    el[7] = en;
                    el[8] = pn;
                    el[9] = qn;
                    el[10] = ssn;
                }
            }
            started += __builtin_popcount(mlo) + __builtin_popcount(mhi);
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

// ---------------------------------------------------------------------------
// Material in element order.  Same per-element work as k_material, but the
// `rep` evaluations of EvalEOSAllInOneTask of the 64 elements of a wave are
// dealt out over its 64 lanes: the wave's sum of repetitions is split into
// rounds of 64 evaluations, lane L of round r evaluates repetition
// r*64 + L, which belongs to the element whose prefix range holds it.  Every
// evaluation reads its inputs from LDS (so none can be folded into another)
// and the element's last repetition writes its outputs.  A wave needs
// ceil(sum rep / 64) rounds instead of max(rep), and the element fields are
// read in order (coalesced) with one round trip to memory.  NT: the
// results are stored with the non-temporal hint (stream_put).
// ---------------------------------------------------------------------------
template <typename QG, bool RHO1, bool STRUCT, bool NT>
__global__ __launch_bounds__(kBlock) void k_material_ordered(
    const ProbDesc *__restrict__ pds, const int32_t *__restrict__ blockProb,
    const uint16_t *__restrict__ erep, ElemFields f, const QG *__restrict__ delv_xi,
    const QG *__restrict__ delv_eta, const QG *__restrict__ delv_zeta,
    const QG *__restrict__ delx_xi, const QG *__restrict__ delx_eta,
    const QG *__restrict__ delx_zeta, Params P, DevState *states,
    unsigned long long *__restrict__ minslots) {
    // dynamic LDS, sized to the block: per element 7 EOS inputs and 4
    // outputs (kMatSlots doubles, one address and immediate offsets per
    // element), then the start maps and the inclusive repetition sums
    extern __shared__ double mat_lds[];
    const int nt = blockDim.x;
    unsigned long long *smap =
        reinterpret_cast<unsigned long long *>(mat_lds + kMatSlots * nt);
    int32_t *sincl = reinterpret_cast<int32_t *>(smap + (nt >> 6) * kMapWords);
    const int blk = xcd_block(blockIdx.x, gridDim.x);
    const int prob = blockProb[blk];
    DevState *st = states + prob;
    if (!st->active) return; // uniform across the block
    const ProbDesc pd = pds[prob];
    const int tid = threadIdx.x;
    const int li = (blk - pd.mBlock0) * blockDim.x + tid;
    const bool valid = li < pd.E;
    const int i = pd.elemOff + (valid ? li : pd.E - 1);
    double dtc = 1.7976931348623157e+308, dth = 1.7976931348623157e+308;
    // all loads of the element and its six face neighbours are independent
    const int32_t bc = f.elemBC[i];
    const double sxi = double(delv_xi[i]), seta = double(delv_eta[i]),
                 szeta = double(delv_zeta[i]);
    double nb[6];
    if (STRUCT) {
        const int s = pd.s, lo = pd.elemOff, hi = pd.elemOff + pd.E - 1;
        const int off[6] = {-1, 1, -s, s, -s * s, s * s};
        const QG *arr[6] = {delv_xi, delv_xi, delv_eta, delv_eta, delv_zeta, delv_zeta};
#pragma unroll
        for (int d = 0; d < 6; ++d) nb[d] = double(arr[d][min(max(i + off[d], lo), hi)]);
    } else {
        nb[0] = double(delv_xi[f.lxim[i]]);
        nb[1] = double(delv_xi[f.lxip[i]]);
        nb[2] = double(delv_eta[f.letam[i]]);
        nb[3] = double(delv_eta[f.letap[i]]);
        nb[4] = double(delv_zeta[f.lzetam[i]]);
        nb[5] = double(delv_zeta[f.lzetap[i]]);
    }
    const double dx[3] = {double(delx_xi[i]), double(delx_eta[i]), double(delx_zeta[i])};
    const double vdov = f.vdov[i], vnew = f.vnew[i], emass = f.elemMass[i], volo = f.volo[i];
    const double e0 = f.e[i], dv0 = f.delv[i], p0 = f.p[i], q0 = f.q[i], arealg = f.arealg[i];
    const int rep = valid ? int(erep[i]) : 0;
    // BC switch of lulesh.cc:1460-1597 on the loaded neighbours
    auto sel = [&](int32_t bits, int32_t comm, int32_t symm, double nbv, double self) {
        return (bits == comm || bits == 0) ? nbv : (bits == symm ? self : 0.0);
    };
    const double phixi = mono_phi(sxi, sel(bc & XI_M, XI_M_COMM, XI_M_SYMM, nb[0], sxi),
                                  sel(bc & XI_P, XI_P_COMM, XI_P_SYMM, nb[1], sxi), P);
    const double phieta = mono_phi(seta, sel(bc & ETA_M, ETA_M_COMM, ETA_M_SYMM, nb[2], seta),
                                   sel(bc & ETA_P, ETA_P_COMM, ETA_P_SYMM, nb[3], seta), P);
    const double phizeta =
        mono_phi(szeta, sel(bc & ZETA_M, ZETA_M_COMM, ZETA_M_SYMM, nb[4], szeta),
                 sel(bc & ZETA_P, ZETA_P_COMM, ZETA_P_SYMM, nb[5], szeta), P);
    const double dv[3] = {sxi, seta, szeta};
    double qlin, qquad;
    q_terms(phixi, phieta, phizeta, vdov, dv, dx, emass, volo, vnew, P, qlin, qquad);
    double vnewc = vnew;
    if (P.eosvmin != 0.0 && vnewc < P.eosvmin) vnewc = P.eosvmin;
    if (P.eosvmax != 0.0 && vnewc > P.eosvmax) vnewc = P.eosvmax;
    if (valid && f.qq) { // (null: not kept, dead between cycles)
        stream_put<NT>(&f.qq[i], qquad);
        stream_put<NT>(&f.ql[i], qlin);
    }
    {
        double *me = mat_lds + kMatSlots * tid;
        me[0] = e0;
        me[1] = dv0;
        me[2] = p0;
        me[3] = q0;
        me[4] = qquad;
        me[5] = qlin;
        me[6] = vnewc;
    }
    eos_wave_balanced<RHO1, false>(mat_lds, smap, sincl, nullptr, tid, rep, P);
    if (valid) {
        const double *me = mat_lds + kMatSlots * tid;
        const double en = me[7], pn = me[8], qn = me[9], ssn = me[10];
        stream_put<NT>(&f.ss[i], ssn);
        stream_put<NT>(&f.p[i], pn);
        stream_put<NT>(&f.e[i], en);
        stream_put<NT>(&f.q[i], qn);
        constraints(ssn, vdov, arealg, P, dtc, dth);
    }
    block_min_atomic(dtc, dth, minslots + size_t(prob) * kMinSlots * kSlotU64);
}

// ---------------------------------------------------------------------------
// Elements and material in one pass over the structured cube, without
// storing the Q gradients: CalcKinematicsForElemsTask, UpdateVolumes and
// CalcMonotonicQGradientsForElemsTask (lulesh.cc:1169-1418), then
// CalcMonotonicQRegionForElemsAndApplyInitTask, EvalEOSAllInOneTask x rep,
// CalcSoundSpeedForElemsAndSaveTask and the element constraints
// (lulesh.cc:1420-1974), per element the same arithmetic as k_elems and
// k_material_ordered.
//
// A workgroup owns a tile of pw x ph element columns (one thread each) over
// a range [k0, k1) of element planes and marches upward in z.  Step t
// computes the kinematics and gradients of plane t, then the material of
// plane t-1, whose monotonic limiter reads the gradients of its six face
// neighbours:
//   zeta (planes t-2 and t)  the same thread's registers;
//   xi, eta (in the plane)   the tile's gradient grids in LDS, whose border
//                            holds the elements just outside the tile,
//                            recomputed by the first wave (2 (pw + ph) <= 64
//                            of them per plane);
//   below k0, above k1 - 1   delv_zeta of planes k0-1 and k1 recomputed by
//                            every thread for its column.
// A recomputed element repeats its owner's operations on the same inputs,
// so it has the owner's gradients (bitwise, in the separately rounded
// form).  The grids rotate over three slots, so one barrier per step
// orders the writes of plane t after the reads of plane t-2.  The EOS
// repetitions are dealt out over each wave as in k_material_ordered (idle
// lanes of a partial tile sit between the elements of a wave: COMPACT).
// NT: the results are stored with the non-temporal hint (stream_put).
//
// Measured at s = 150 (FP64, multiply-adds fused): 0.42 ms per cycle against
// 0.145 + 0.133 ms for k_elems + k_material_ordered, although it moves about
// 100 fewer bytes per element.  254 VGPRs hold it to two waves per SIMD
// (three spill 97 VGPRs: 0.84 ms); its kinematics steps alone take 0.20 ms
// and its material steps alone 0.167 ms (each slower than its own kernel,
// which runs at three and seven waves per SIMD), the recomputed border
// 0.03 ms.  material_mode 2 selects it; the default stays 0.
// ---------------------------------------------------------------------------
constexpr int kEmThreads = 256;
constexpr int kEmGrid = 18 * 18; // (pw + 2)(ph + 2) for pw + ph <= 32

template <typename QG, bool RHO1, bool NT>
__global__ __launch_bounds__(kEmThreads) void k_elemmat_zmarch(
    const ProbDesc *__restrict__ pds, const int32_t *__restrict__ blockProb, NodeBufs nb,
    ElemKin f0, ElemFields f, const uint16_t *__restrict__ erep, Params P, DevState *states,
    unsigned long long *__restrict__ minslots) {
    __shared__ double mat_lds[kMatSlots * kEmThreads];
    __shared__ unsigned long long smap[(kEmThreads / 64) * kMapWords];
    __shared__ int32_t sincl[kEmThreads], elanes[kEmThreads];
    __shared__ QG gxi[3][kEmGrid], geta[3][kEmGrid];
    const int blk = xcd_block(blockIdx.x, gridDim.x);
    const int prob = blockProb[blk];
    DevState *st = states + prob;
    if (!st->active) return; // uniform across the block
    const ProbDesc pd = pds[prob];
    const int s = pd.s, nn = s + 1, pw = pd.emPW, ph = pd.emPH, gw = pw + 2;
    const int eo = pd.elemOff, no = pd.nodeOff;
    const int lb = blk - pd.emBlock0, tilesXY = pd.emTX * pd.emTY;
    const int tile = lb % tilesXY, zg = lb / tilesXY;
    const int E0x = (tile % pd.emTX) * pw, E0y = (tile / pd.emTX) * ph;
    const int k0 = zg * pd.emPlanes, k1 = min(k0 + pd.emPlanes, s); // planes [k0, k1)
    const int tid = threadIdx.x;
    const int jj = tid / pw, ii = tid - jj * pw;
    const bool inTile = jj < ph;
    const int ei = E0x + ii, ej = E0y + jj; // the thread's element column
    const bool elemCol = inTile && ei < s && ej < s;
    // grid position of the thread's element (threads past the tile: a
    // harmless interior slot; their results are not stored)
    const int own = inTile ? (jj + 1) * gw + ii + 1 : gw + 1;
    // the border element of lane tid < 2 (pw + ph) of the first wave
    // (sides xi-, xi+, eta-, eta+), the gradient it contributes and where
    int hx = -1, hy = -1, hpos = 0;
    bool hxi = false;
    {
        int h = tid;
        if (h < ph) {
            hx = E0x - 1, hy = E0y + h, hpos = (h + 1) * gw, hxi = true;
        } else if ((h -= ph) < ph) {
            hx = E0x + pw, hy = E0y + h, hpos = (h + 1) * gw + pw + 1, hxi = true;
        } else if ((h -= ph) < pw) {
            hx = E0x + h, hy = E0y - 1, hpos = h + 1;
        } else if ((h -= pw) < pw) {
            hx = E0x + h, hy = E0y + ph, hpos = (ph + 1) * gw + h + 1;
        }
    }
    const bool border = hx >= 0 && hx < s && hy >= 0 && hy < s;
    const double deltaTime = st->deltatime;
    double *const *nf = nb.b[st->cur ^ st->fresh];
    // element (k, j, i) -> its eight nodes, lulesh-init.cc:236-244
    auto nodes_of = [&](int k, int j, int i, int n[8]) {
        const int b = no + (k * nn + j) * nn + i;
        n[0] = b; n[1] = b + 1; n[2] = b + nn + 1; n[3] = b + nn;
        n[4] = b + nn * nn; n[5] = b + nn * nn + 1; n[6] = b + nn * nn + nn + 1;
        n[7] = b + nn * nn + nn;
    };
    // the gradients alone (volume, then the Q gradients on it)
    auto gradients_of = [&](int k, int j, int i, QG g[6]) {
        int n[8];
        nodes_of(k, j, i, n);
        double x[8], y[8], z[8], xd[8], yd[8], zd[8];
        load8_pairs(nf[0], n, x);
        load8_pairs(nf[1], n, y);
        load8_pairs(nf[2], n, z);
        load8_pairs(nf[3], n, xd);
        load8_pairs(nf[4], n, yd);
        load8_pairs(nf[5], n, zd);
        const double voloi = f0.volo[eo + (k * s + j) * s + i];
        const double rel = elem_volume<double>(x, y, z) / voloi;
        element_q_gradients<QG>(x, y, z, xd, yd, zd, voloi * rel, g);
    };
    // material inputs of plane t-1 carried from its kinematics
    double m_vdov = 0.0, m_vnew = 0.0, m_delv = 0.0, m_arealg = 0.0, m_volo = 0.0;
    QG m_dx[3] = {QG(0), QG(0), QG(0)};
    QG z_lo = QG(0), z_mid = QG(0); // delv_zeta of planes t-2, t-1
    double dtc = 1.7976931348623157e+308, dth = 1.7976931348623157e+308;
    const int tlo = k0 > 0 ? k0 - 1 : 0;
    for (int t = tlo; t <= k1; ++t) {
        const int sl = t % 3;
        QG z_hi = QG(0);
        double n_vdov = 0.0, n_vnew = 0.0, n_delv = 0.0, n_arealg = 0.0, n_volo = 0.0;
        QG n_dx[3] = {QG(0), QG(0), QG(0)};
        if (t >= k0 && t < k1) {
            // ---- kinematics of plane t (k_elems) -------------------------
            if (elemCol) {
                const int i = eo + (t * s + ej) * s + ei;
                int n[8];
                nodes_of(t, ej, ei, n);
                double x[8], y[8], z[8], xd[8], yd[8], zd[8];
                load8_pairs(nf[0], n, x);
                load8_pairs(nf[1], n, y);
                load8_pairs(nf[2], n, z);
                load8_pairs(nf[3], n, xd);
                load8_pairs(nf[4], n, yd);
                load8_pairs(nf[5], n, zd);
                const double voloi = f0.volo[i];
                const double vold = f0.v[i];
                const double volume = elem_volume<double>(x, y, z);
                const double rel = volume / voloi;
                const double arealg = characteristic_length(x, y, z, volume);
                QG g[6];
                element_q_gradients<QG>(x, y, z, xd, yd, zd, voloi * rel, g);
                const double dt2 = 0.5 * deltaTime;
#pragma unroll
                for (int k = 0; k < 8; ++k) {
                    x[k] -= dt2 * xd[k];
                    y[k] -= dt2 * yd[k];
                    z[k] -= dt2 * zd[k];
                }
                double B[3][8];
                const double detJ = shape_derivatives<true>(x, y, z, B);
                const double vdov = velocity_divergence(xd, yd, zd, B, detJ);
                double vc = vold;
                if (P.eosvmin != 0.0 && vc < P.eosvmin) vc = P.eosvmin;
                if (P.eosvmax != 0.0 && vc > P.eosvmax) vc = P.eosvmax;
                if (vc < 0.0) st->error = kVolumeError;
                stream_put<NT>(&f0.vnew[i], rel);
                stream_put<NT>(&f0.delv[i], rel - vold);
                stream_put<NT>(&f0.arealg[i], arealg);
                stream_put<NT>(&f0.vdov[i], vdov);
                stream_put<NT>(&f0.v[i], (fabs_(rel - 1.0) < P.v_cut) ? 1.0 : rel);
                gxi[sl][own] = g[3];
                geta[sl][own] = g[4];
                z_hi = g[5];
                n_dx[0] = g[0];
                n_dx[1] = g[1];
                n_dx[2] = g[2];
                n_vdov = vdov;
                n_vnew = rel;
                n_delv = rel - vold;
                n_arealg = arealg;
                n_volo = voloi;
            }
            if (border) { // the tile's border element of this lane
                QG g[6];
                gradients_of(t, hy, hx, g);
                (hxi ? gxi[sl] : geta[sl])[hpos] = hxi ? g[3] : g[4];
            }
        } else if (t < s && elemCol) { // plane k0 - 1 or k1: delv_zeta alone
            QG g[6];
            gradients_of(t, ej, ei, g);
            z_hi = g[5];
        }
        __syncthreads();
        if (t > k0) {
            // ---- material of plane m = t - 1 (k_material_ordered) ----------
            const int m = t - 1, ms = m % 3;
            const int i = eo + (elemCol ? (m * s + ej) * s + ei : 0);
            const int32_t bc = f.elemBC[i];
            const double emass = f.elemMass[i];
            const double e0 = f.e[i], p0 = f.p[i], q0 = f.q[i];
            const int rep = elemCol ? int(erep[i]) : 0;
            const double sxi = double(gxi[ms][own]), seta = double(geta[ms][own]),
                         szeta = double(z_mid);
            // BC switch of lulesh.cc:1460-1597 on the neighbours' gradients
            auto sel = [&](int32_t bits, int32_t comm, int32_t symm, double nbv, double self) {
                return (bits == comm || bits == 0) ? nbv : (bits == symm ? self : 0.0);
            };
            const double phixi =
                mono_phi(sxi, sel(bc & XI_M, XI_M_COMM, XI_M_SYMM, double(gxi[ms][own - 1]), sxi),
                         sel(bc & XI_P, XI_P_COMM, XI_P_SYMM, double(gxi[ms][own + 1]), sxi), P);
            const double phieta = mono_phi(
                seta, sel(bc & ETA_M, ETA_M_COMM, ETA_M_SYMM, double(geta[ms][own - gw]), seta),
                sel(bc & ETA_P, ETA_P_COMM, ETA_P_SYMM, double(geta[ms][own + gw]), seta), P);
            const double phizeta =
                mono_phi(szeta, sel(bc & ZETA_M, ZETA_M_COMM, ZETA_M_SYMM, double(z_lo), szeta),
                         sel(bc & ZETA_P, ZETA_P_COMM, ZETA_P_SYMM, double(z_hi), szeta), P);
            const double dv[3] = {sxi, seta, szeta};
            const double dx[3] = {double(m_dx[0]), double(m_dx[1]), double(m_dx[2])};
            double qlin, qquad;
            q_terms(phixi, phieta, phizeta, m_vdov, dv, dx, emass, m_volo, m_vnew, P, qlin,
                    qquad);
            double vnewc = m_vnew;
            if (P.eosvmin != 0.0 && vnewc < P.eosvmin) vnewc = P.eosvmin;
            if (P.eosvmax != 0.0 && vnewc > P.eosvmax) vnewc = P.eosvmax;
            if (elemCol && f.qq) { // (null: not kept, dead between cycles)
                stream_put<NT>(&f.qq[i], qquad);
                stream_put<NT>(&f.ql[i], qlin);
            }
            double *me = mat_lds + kMatSlots * tid;
            me[0] = e0;
            me[1] = m_delv;
            me[2] = p0;
            me[3] = q0;
            me[4] = qquad;
            me[5] = qlin;
            me[6] = vnewc;
            eos_wave_balanced<RHO1, true>(mat_lds, smap, sincl, elanes, tid, rep, P);
            if (elemCol) {
                const double en = me[7], pn = me[8], qn = me[9], ssn = me[10];
                stream_put<NT>(&f.ss[i], ssn);
                stream_put<NT>(&f.p[i], pn);
                stream_put<NT>(&f.e[i], en);
                stream_put<NT>(&f.q[i], qn);
                double ec, eh; // this element's constraints, folded into the thread's minima
                constraints(ssn, m_vdov, m_arealg, P, ec, eh);
                dtc = (ec < dtc) ? ec : dtc;
                dth = (eh < dth) ? eh : dth;
            }
        }
        z_lo = z_mid;
        z_mid = z_hi;
        m_vdov = n_vdov;
        m_vnew = n_vnew;
        m_delv = n_delv;
        m_arealg = n_arealg;
        m_volo = n_volo;
        m_dx[0] = n_dx[0];
        m_dx[1] = n_dx[1];
        m_dx[2] = n_dx[2];
    }
    block_min_atomic(dtc, dth, minslots + size_t(prob) * kMinSlots * kSlotU64);
}

// Stand-alone constraint pass over one problem's elements
// (CalcConstraintForElemsTask, lulesh.cc:1938-1974): the same minima as the
// fused path, for phase-level checks on arbitrary inputs.
__global__ __launch_bounds__(kBlock) void k_constraints(int32_t E, const double *__restrict__ ss,
                                                        const double *__restrict__ vdov,
                                                        const double *__restrict__ arealg,
                                                        Params P,
                                                        unsigned long long *minslots) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double dtc = 1.7976931348623157e+308, dth = 1.7976931348623157e+308;
    if (i < E) constraints(ss[i], vdov[i], arealg[i], P, dtc, dth);
    block_min_atomic(dtc, dth, minslots);
}

#if !LULESH_FUSED_FORM
// Arm every constraint-minima slot of one problem at 1e20 (a fill kernel,
// so that the clearing can sit in a captured graph).
__global__ __launch_bounds__(kMinSlots) void k_arm_slots(unsigned long long *slots) {
    unsigned long long *sl = slots + size_t(threadIdx.x) * kSlotU64;
    sl[0] = kMinInit;
    sl[1] = kMinInit;
}

// Fold one problem's slots into out[0] = dtcourant, out[1] = dthydro (and
// re-arm them); one workgroup of kMinSlots threads.
__global__ __launch_bounds__(kMinSlots) void k_fold_minima_to(unsigned long long *slots,
                                                              double *out) {
    unsigned long long c, h;
    fold_slots(slots, c, h);
    if (threadIdx.x == 0) {
        out[0] = bits_to_double(c);
        out[1] = bits_to_double(h);
    }
}
#endif

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

// Streaming copy for the HBM ceiling measurement: every lane moves UNROLL
// 16-byte vectors per pass, loads first, with the non-temporal hint on both
// sides; GRID_STRIDE: a fixed grid walks the buffer, else one pass per lane.
typedef double vec2d __attribute__((ext_vector_type(2)));

template <int UNROLL, bool GRID_STRIDE>
__global__ __launch_bounds__(256) void k_stream_copy(const vec2d *__restrict__ src,
                                                     vec2d *__restrict__ dst, int64_t n) {
    const int64_t lanes = int64_t(gridDim.x) * blockDim.x;
    int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    do {
        vec2d v[UNROLL];
#pragma unroll
        for (int u = 0; u < UNROLL; ++u) {
            const int64_t k = i + int64_t(u) * lanes;
            if (k < n) v[u] = __builtin_nontemporal_load(src + k);
        }
#pragma unroll
        for (int u = 0; u < UNROLL; ++u) {
            const int64_t k = i + int64_t(u) * lanes;
            if (k < n) __builtin_nontemporal_store(v[u], dst + k);
        }
        i += int64_t(UNROLL) * lanes;
    } while (GRID_STRIDE && i < n);
}

} // namespace

// ---------------------------------------------------------------------------
// Context
// ---------------------------------------------------------------------------
struct KernelTimer {
    const char *name;
    double total_ms = 0;
    int64_t launches = 0;
};

struct ProbHost {
    int32_t s = 0, E = 0, N = 0, numReg = 0, cost = 0;
    int32_t elemOff = 0, nodeOff = 0;
    std::vector<MatSeg> segs; // offsets into matList; block0 set by configure()
};

struct lulesh_ctx {
    // A batch split over several streams is a group: `lanes` are complete
    // contexts over disjoint subsets of the problems (problem p is problem
    // local_of[p] of lane lane_of[p]) and the group holds no device data.
    std::vector<lulesh_ctx *> lanes;
    std::vector<int32_t> lane_of, local_of;
    int precision = 0;
    bool structured = false;
    bool store_forces = false; // also write fx.. and xdd.. (Domain fidelity)
    int nprob = 0;
    int cus = 256;         // compute units of the device
    bool zSquare16 = true; // every problem's z-march patch is 16 x 16
    // The kinematics and material kernels store their results with the
    // non-temporal hint: contexts of one problem.  Measured at s = 60-250
    // (FP64 and mixed): 1-3% less time per cycle, the next kernel's reads
    // gaining more than the stores cost; the 64-problem batch on two
    // streams ran 0.3-0.9% slower with it.
    bool streamStores = false;
    std::vector<ProbHost> probs;
    int64_t totE = 0, totN = 0;
    Params P{};
    double hgcoef = 0;
    hipStream_t stream = nullptr;
    std::vector<void *> allocs;
    int64_t bytes = 0;
    // field base pointers (concatenated over problems); the node fields
    // x..zd live in nodeBuf (two halves on the structured path)
    double *f[LULESH_F_COUNT] = {};
    float *g32[6] = {}; // mixed form: FP32 gradient storage
    NodeBufs nb{};
    int32_t *lxim = nullptr, *lxip = nullptr, *letam = nullptr, *letap = nullptr,
            *lzetam = nullptr, *lzetap = nullptr, *elemBC = nullptr, *matList = nullptr;
    // indexed path (one problem)
    int32_t *nodelist = nullptr, *nodeStart = nullptr, *cornerMap = nullptr;
    uint8_t *bcMask = nullptr;
    double *cs[3] = {};
    void *ch[3] = {};
    // launch tables (rebuilt by configure())
    lulesh_tuning tune{};
    std::vector<ProbDesc> hpd;
    std::vector<MatSeg> hsegs;
    int zBlocks = 0, eBlocks = 0, mBlocks = 0;
    ProbDesc *d_pd = nullptr;
    MatSeg *d_segs = nullptr;
    int32_t *d_zBlockProb = nullptr, *d_eBlockProb = nullptr, *d_mBlockSeg = nullptr;
    int32_t *d_zMap = nullptr; // z-march: hardware block -> logical block (-1: idle)
    int zGrid = 0;
    int32_t *d_oBlockProb = nullptr; // element-order material grid
    int oBlocks = 0;
    int32_t *d_emBlockProb = nullptr; // fused elements + material grid
    int emBlocks = 0;
    // the Q gradients in HBM are those of the last elements phase; a cycle
    // of the fused elements + material pass does not store them
    bool grads_stale = false;
    uint16_t *erep = nullptr;        // EOS repetitions per element
    unsigned long long *minslots = nullptr; // constraint minima slots, kMinSlots per problem
    // time states
    DevState *st = nullptr;
    DevState *hst = nullptr; // pinned mirror of all problems' states
    std::vector<DevState> init;
    // captured cycle graphs: graphs[k] replays graph_sizes[k] cycles; the
    // sizes are cycles_per_graph and the powers of two below it, so any
    // count of cycles up to cycles_per_graph is a sum of distinct graphs
    std::vector<hipGraphExec_t> graphs;
    std::vector<int> graph_sizes;
    bool timing = false;
    std::vector<KernelTimer> timers;
    hipEvent_t run0 = nullptr, run1 = nullptr;
    // timing mode: an event pair per launch, read when the stream next
    // synchronises, so that timed kernels still run back to back
    std::vector<hipEvent_t> tev;
    std::vector<int> tev_timer;
    size_t tev_used = 0;
    double last_run_ms = 0;
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
    HIP_TRY(c, hipMemset(q, 0, bytes));
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
bool is_state_node_field(int id) { return id >= LULESH_F_X && id <= LULESH_F_ZD; }
bool is_grad_field(int id) { return id >= LULESH_F_DELV_XI && id <= LULESH_F_DELX_ZETA; }
bool is_force_field(int id) { return id >= LULESH_F_XDD && id <= LULESH_F_FZ; }
// The fields that are dead between cycles: fx.., xdd.. and ql, qq.  A
// context writes them only when it keeps them (store_forces).
bool is_dead_field(int id) {
    return is_force_field(id) || id == LULESH_F_QL || id == LULESH_F_QQ;
}

double *view_field(const lulesh_domain_view *v, int id) {
    double *const tab[LULESH_F_COUNT] = {
        v->x, v->y, v->z, v->xd, v->yd, v->zd, v->xdd, v->ydd, v->zdd, v->fx, v->fy, v->fz,
        v->nodalMass, v->e, v->p, v->q, v->ql, v->qq, v->v, v->volo, v->delv, v->vdov, v->arealg,
        v->ss, v->elemMass, v->vnew, v->delv_xi, v->delv_eta, v->delv_zeta, v->delx_xi,
        v->delx_eta, v->delx_zeta};
    return tab[id];
}

int grad_slot(int id) { return id - LULESH_F_DELV_XI; }

// The node state x, y, z, xd, yd, zd: one allocation per half of the double
// buffer (the structured path has two halves, the indexed path one), the six
// fields totN doubles apart.  f[X..ZD] are the first half.
int alloc_node_state(lulesh_ctx *c, bool two_halves) {
    const size_t N = size_t(c->totN);
    for (int h = 0; h < 2; ++h) {
        if (h == 1 && !two_halves) {
            for (int k = 0; k < 6; ++k) c->nb.b[1][k] = c->nb.b[0][k];
            break;
        }
        double *block = nullptr;
        if (int rc = dalloc(c, &block, 6 * N)) return rc;
        for (int k = 0; k < 6; ++k) c->nb.b[h][k] = block + size_t(k) * N;
    }
    for (int k = 0; k < 6; ++k) c->f[LULESH_F_X + k] = c->nb.b[0][k];
    return 0;
}

template <typename F>
int launch(lulesh_ctx *c, int timer_id, F &&fn) {
    if (c->timing) {
        const size_t k = c->tev_used;
        while (c->tev.size() < 2 * (k + 1)) {
            // timing-only events: no system-scope cache writeback at record
            hipEvent_t e = nullptr;
            HIP_TRY(c, hipEventCreateWithFlags(&e, hipEventDisableSystemFence));
            c->tev.push_back(e);
        }
        c->tev_timer.resize(k + 1);
        c->tev_timer[k] = timer_id;
        HIP_TRY(c, hipEventRecord(c->tev[2 * k], c->stream));
        fn();
        HIP_TRY(c, hipGetLastError());
        HIP_TRY(c, hipEventRecord(c->tev[2 * k + 1], c->stream));
        c->tev_used = k + 1;
        return 0;
    }
    fn();
    HIP_TRY(c, hipGetLastError());
    return 0;
}

// Add the recorded launch durations to the timers (the caller has
// synchronised the stream).
int flush_timers(lulesh_ctx *c) {
    for (size_t k = 0; k < c->tev_used; ++k) {
        float ms = 0;
        HIP_TRY(c, hipEventElapsedTime(&ms, c->tev[2 * k], c->tev[2 * k + 1]));
        auto &t = c->timers[size_t(c->tev_timer[k])];
        t.total_ms += ms;
        t.launches += 1;
    }
    c->tev_used = 0;
    return 0;
}

enum { T_TIME = 0, T_NODAL, T_NODES, T_ELEMS, T_MATERIAL, T_MINIMA, T_ELEMMAT, T_COUNT };

// Node planes per z-march workgroup when the tuning leaves it open (one
// problem: the cost model below; a batch: the fewest plane groups that give
// every CU a workgroup).
int auto_zplanes(const lulesh_ctx *c) {
    if (c->nprob == 1) {
        // One problem: the count that minimises (rounds of workgroups) x
        // (element planes per workgroup), a round being the two z-march
        // workgroups per CU that fit at once, among the counts that fill at
        // least 90% of a round (all counts compete when none does).  This
        // model's pick was the fastest of every group count measured at
        // s = 150 and s = 250.
        const int n = c->probs[0].s + 1;
        int pw, ph;
        zmarch_shape(n, pw, ph);
        const int64_t t2 = int64_t(zmarch_tiles(n, pw)) * zmarch_tiles(n, ph);
        const int64_t slots = 2 * int64_t(std::max(c->cus, 1));
        bool any_full = false;
        for (int z = 1; z <= n && !any_full; ++z) any_full = 10 * t2 * zmarch_groups(n, z) >= 9 * slots;
        int best = 1;
        int64_t best_cost = INT64_MAX;
        for (int z = n; z >= 1; --z) { // ties: the smaller z (more workgroups)
            const int64_t wgs = t2 * zmarch_groups(n, z);
            if (any_full && 10 * wgs < 9 * slots) continue;
            const int64_t cost = ((wgs + slots - 1) / slots) * zmarch_steps(n, z);
            if (cost <= best_cost) {
                best_cost = cost;
                best = z;
            }
        }
        return best;
    }
    int nmax = 1;
    for (const auto &p : c->probs) nmax = std::max(nmax, p.s + 1);
    auto blocks = [&](int z) {
        int64_t b = 0;
        for (const auto &p : c->probs) {
            const int n = p.s + 1;
            int pw, ph;
            zmarch_shape(n, pw, ph);
            b += int64_t(zmarch_tiles(n, pw)) * zmarch_tiles(n, ph) * zmarch_groups(n, z);
        }
        return b;
    };
    // A batch: the fewest plane groups along z (for its largest problem)
    // that still give every CU a workgroup; the z-march's workgroups of a
    // batch walk few planes, so each further group adds its recomputed
    // plane and its prologue to little work.  Measured on the 64-problem
    // batch (two streams, s = 21..80): one group per problem 8.08e9
    // zone-cycles/s, two 7.97e9, the previous rule (>= 2048 workgroups, 12
    // planes) 7.60e9.
    for (int g = 1; g <= 40; ++g) {
        const int z = (nmax + 2 * g - 2) / g; // the smallest z with at most g plane groups
        if (z >= 2 && blocks(z) >= std::max(c->cus, 1)) return z;
    }
    return std::min(4, nmax);
}

// Element planes per workgroup of the fused elements + material pass when
// the tuning leaves it open.  One problem: the count that minimises (rounds
// of workgroups) x (planes walked, the two recomputed delv_zeta planes
// counted as one), a round being two workgroups per CU.  A batch: the
// largest count that still gives 2048 workgroups.
int auto_em_planes(const lulesh_ctx *c) {
    auto tiles = [](int s) {
        int pw, ph;
        elemmat_shape(s, pw, ph);
        return int64_t((s + pw - 1) / pw) * ((s + ph - 1) / ph);
    };
    if (c->nprob == 1) {
        const int s = c->probs[0].s;
        const int64_t t2 = tiles(s), slots = 2 * int64_t(std::max(c->cus, 1));
        int best = s;
        int64_t best_cost = INT64_MAX;
        for (int z = s; z >= 1; --z) { // ties: the larger z (fewer recomputed planes)
            const int64_t wgs = t2 * ((s + z - 1) / z);
            const int64_t cost = ((wgs + slots - 1) / slots) * (z + 1);
            if (cost < best_cost) {
                best_cost = cost;
                best = z;
            }
        }
        return best;
    }
    int smax = 1;
    for (const auto &p : c->probs) smax = std::max(smax, p.s);
    auto blocks = [&](int z) {
        int64_t b = 0;
        for (const auto &p : c->probs) b += tiles(p.s) * ((p.s + z - 1) / z);
        return b;
    };
    for (int z : {32, 24, 16, 12, 10, 8, 6, 5, 4})
        if (z <= smax && blocks(z) >= 2048) return z;
    return std::min(4, smax);
}

void destroy_graphs(lulesh_ctx *c) {
    for (hipGraphExec_t g : c->graphs)
        if (g) (void)hipGraphExecDestroy(g);
    c->graphs.clear();
    c->graph_sizes.clear();
}

// The order of xcd_block as a map: every XCD a contiguous range of blocks.
std::vector<int32_t> contiguous_block_map(int nblocks) {
    std::vector<int32_t> map(static_cast<size_t>(nblocks));
    const int per = nblocks / kXcds, rem = nblocks % kXcds;
    for (int b = 0; b < nblocks; ++b) {
        const int xcd = b % kXcds, idx = b / kXcds;
        map[size_t(b)] = xcd * per + std::min(xcd, rem) + idx;
    }
    return map;
}

// Hardware block -> logical block of the z-march launch (-1: an idle
// block).  Workgroups are dispatched round-robin over the XCDs (hardware
// block b to XCD b % 8) and each XCD's L2 serves its own workgroups.  One
// problem: the order of xcd_block, every XCD a contiguous range of tiles
// (its workgroups all walk about the same planes).  A batch, whose
// workgroups walk different numbers of planes (small problems fewer): runs
// of up to 8 adjacent tiles of one plane group (which share halo nodes in
// L2), the longest walks first, each to the XCD with the least work so far
// (planes walked + 1 per workgroup), so that every XCD gets the same work
// and starts its longest workgroups first.  XCD x's k-th block is hardware
// block 8k + x; XCDs with fewer blocks get idle ones at the end.
// Measured on the 64-problem batch (s = 21..80): nodal 65.2 -> 58.1 ps per
// element.  (The same dealing for the element-order material, weighted by
// EOS rounds, was slower than contiguous ranges for the single problem and
// the batch alike: 133 -> 134 and 39.9 -> 45.1 ps per element.)
std::vector<int32_t> zmarch_block_map(const std::vector<ProbDesc> &pds, int nblocks) {
    if (pds.size() <= 1) return contiguous_block_map(nblocks);
    struct Run {
        int64_t steps;
        int32_t first, count;
    };
    std::vector<Run> runs;
    for (const ProbDesc &d : pds) {
        const int n = d.s + 1, tiles = d.tilesX * d.tilesY;
        for (int g = 0, G = zmarch_groups(n, d.zplanes); g < G; ++g) {
            const int K0 = zmarch_group_start(g, d.zplanes);
            const int K1 = std::min(zmarch_group_start(g + 1, d.zplanes), n);
            const int64_t steps = std::min(K1, n - 1) - std::max(K0 - 1, 0) + 1;
            for (int t = 0; t < tiles; t += 8)
                runs.push_back({steps, d.zBlock0 + g * tiles + t, std::min(8, tiles - t)});
        }
    }
    std::stable_sort(runs.begin(), runs.end(),
                     [](const Run &a, const Run &b) { return a.steps > b.steps; });
    std::vector<std::vector<int32_t>> lists(kXcds);
    std::vector<int64_t> load(kXcds, 0);
    for (const Run &r : runs) {
        const int x = int(std::min_element(load.begin(), load.end()) - load.begin());
        for (int k = 0; k < r.count; ++k) lists[size_t(x)].push_back(r.first + k);
        load[size_t(x)] += r.steps * r.count;
    }
    size_t L = 0;
    for (const auto &l : lists) L = std::max(L, l.size());
    std::vector<int32_t> map(L * kXcds, -1);
    for (int x = 0; x < kXcds; ++x)
        for (size_t k = 0; k < lists[size_t(x)].size(); ++k)
            map[k * kXcds + size_t(x)] = lists[size_t(x)][k];
    return map;
}

// Rebuild the block -> problem / segment tables for the current tuning.
int configure(lulesh_ctx *c) {
    lulesh_tuning &t = c->tune;
    // kinematics workgroups: 256 threads for one problem (the autotuner's pick
    // at s = 150: 137 against 139.5 us for 128), 128 for a batch
    if (t.block_elem != 64 && t.block_elem != 128 && t.block_elem != 256)
        t.block_elem = c->nprob == 1 ? 256 : 128;
    // material workgroups: 64 threads for one problem (s = 150: 0.3-1.1% less
    // time per cycle than 128 in each arithmetic form), 128 for a batch
    if (t.block_material != 64 && t.block_material != 128 && t.block_material != 256)
        t.block_material = c->nprob == 1 ? 64 : 128;
    if (t.zplanes <= 0) t.zplanes = auto_zplanes(c);
    if (t.cycles_per_graph <= 0) t.cycles_per_graph = c->totE >= 1000000 ? 16 : 32;
    // material: 0 (element order) where the per-element repetition counts
    // exist, else 1; 2 (fused with the elements, structured cube) only when
    // asked for: at s = 150 it takes 0.42 ms per cycle against 0.28 ms for
    // the two kernels of mode 0 (k_elemmat_zmarch)
    const bool can_fuse = c->structured && c->erep;
    if (t.material_mode < 0 || t.material_mode > 2) t.material_mode = 0;
    if (t.material_mode == 2 && !can_fuse) t.material_mode = 0;
    if (t.material_mode == 0 && !c->erep) t.material_mode = 1;
    if (t.material_mode == 1 && !c->matList) t.material_mode = 0; // no region lists
    if (t.em_planes <= 0) t.em_planes = auto_em_planes(c);
    std::vector<int32_t> zb, eb, mb, ob, emb;
    c->hpd.clear();
    c->zSquare16 = true;
    c->streamStores = c->nprob == 1;
    c->hsegs.clear();
    for (int pi = 0; pi < c->nprob; ++pi) {
        const ProbHost &ph = c->probs[size_t(pi)];
        ProbDesc d{};
        d.s = ph.s;
        d.E = ph.E;
        d.N = ph.N;
        d.elemOff = ph.elemOff;
        d.nodeOff = ph.nodeOff;
        const int n = ph.s + 1;
        zmarch_shape(n, d.pw, d.ph);
        c->zSquare16 = c->zSquare16 && d.pw == 16 && d.ph == 16;
        d.tilesX = zmarch_tiles(n, d.pw);
        d.tilesY = zmarch_tiles(n, d.ph);
        d.inv_s = 1.0 / double(ph.s);
        d.inv_ss = 1.0 / (double(ph.s) * double(ph.s));
        d.zplanes = std::min(t.zplanes, n);
        d.zBlock0 = int32_t(zb.size());
        const int64_t nz = int64_t(d.tilesX) * d.tilesY * zmarch_groups(n, d.zplanes);
        zb.insert(zb.end(), size_t(nz), pi);
        d.eBlock0 = int32_t(eb.size());
        eb.insert(eb.end(), size_t(grid_for(ph.E, t.block_elem)), pi);
        d.mBlock0 = int32_t(ob.size());
        ob.insert(ob.end(), size_t(grid_for(ph.E, t.block_material)), pi);
        elemmat_shape(ph.s, d.emPW, d.emPH);
        d.emTX = (ph.s + d.emPW - 1) / d.emPW;
        d.emTY = (ph.s + d.emPH - 1) / d.emPH;
        d.emPlanes = std::max(1, std::min(t.em_planes, ph.s));
        d.emBlock0 = int32_t(emb.size());
        emb.insert(emb.end(),
                   size_t(d.emTX) * size_t(d.emTY) * size_t((ph.s + d.emPlanes - 1) / d.emPlanes),
                   pi);
        c->hpd.push_back(d);
    }
    // material segments, the costliest (largest rep) first so they start early
    std::vector<MatSeg> all;
    for (const auto &ph : c->probs) all.insert(all.end(), ph.segs.begin(), ph.segs.end());
    std::stable_sort(all.begin(), all.end(),
                     [](const MatSeg &a, const MatSeg &b) { return a.rep > b.rep; });
    for (auto &sg : all) {
        if (sg.count == 0) continue;
        sg.block0 = int32_t(mb.size());
        mb.insert(mb.end(), size_t(grid_for(sg.count, t.block_material)),
                  int32_t(c->hsegs.size()));
        c->hsegs.push_back(sg);
    }
    c->zBlocks = int(zb.size());
    const std::vector<int32_t> zmap = zmarch_block_map(c->hpd, c->zBlocks);
    c->zGrid = int(zmap.size());
    c->eBlocks = int(eb.size());
    c->mBlocks = int(mb.size());
    c->oBlocks = int(ob.size());
    c->emBlocks = int(emb.size());
    auto put = [&](auto **dptr, const auto &vec) -> int {
        using T = std::remove_reference_t<decltype(vec[0])>;
        if (*dptr) (void)hipFree(*dptr);
        *dptr = nullptr;
        HIP_TRY(c, hipMalloc(reinterpret_cast<void **>(dptr),
                             std::max<size_t>(vec.size(), 1) * sizeof(T)));
        if (!vec.empty())
            HIP_TRY(c, hipMemcpy(*dptr, vec.data(), vec.size() * sizeof(T),
                                 hipMemcpyHostToDevice));
        return 0;
    };
    int rc;
    if ((rc = put(&c->d_pd, c->hpd)) || (rc = put(&c->d_segs, c->hsegs)) ||
        (rc = put(&c->d_zBlockProb, zb)) || (rc = put(&c->d_zMap, zmap)) ||
        (rc = put(&c->d_eBlockProb, eb)) ||
        (rc = put(&c->d_mBlockSeg, mb)) || (rc = put(&c->d_oBlockProb, ob)) ||
        (rc = put(&c->d_emBlockProb, emb)))
        return rc;
    destroy_graphs(c);
    return 0;
}

template <typename R>
int enqueue_nodal(lulesh_ctx *c) {
    if (c->structured) {
        const NodeForceOut fo{c->f[LULESH_F_FX], c->f[LULESH_F_FY], c->f[LULESH_F_FZ],
                              c->f[LULESH_F_XDD], c->f[LULESH_F_YDD], c->f[LULESH_F_ZDD]};
        return launch(c, T_NODAL, [&] {
            constexpr int M = sizeof(R) == 4 ? MODE_MIXED : kFusedForm ? MODE_FUSED : MODE_EXACT;
            constexpr bool D = M != MODE_EXACT; // both corner sets staged at once
            auto kern = c->zSquare16
                            ? (c->store_forces ? k_nodal_zmarch<M, true, true, D>
                                               : k_nodal_zmarch<M, false, true, D>)
                            : (c->store_forces ? k_nodal_zmarch<M, true, false, D>
                                               : k_nodal_zmarch<M, false, false, D>);
            kern<<<c->zGrid, kZmarchThreads, 0, c->stream>>>(
                c->d_pd, c->d_zMap, c->d_zBlockProb, c->nb, fo, c->f[LULESH_F_P], c->f[LULESH_F_Q],
                c->f[LULESH_F_VOLO], c->f[LULESH_F_V], c->f[LULESH_F_SS], c->f[LULESH_F_ELEMMASS],
                c->f[LULESH_F_NODALMASS], c->hgcoef, c->P.u_cut, c->st);
        });
    }
    double *const *b = c->nb.b[0];
    const NodeFields nf{b[0], b[1], b[2], b[3], b[4], b[5]};
    const int32_t E = int32_t(c->totE), N = int32_t(c->totN);
    int rc = launch(c, T_NODAL, [&] {
        k_force<R><<<grid_for(E, kBlock), kBlock, 0, c->stream>>>(
            E, c->nodelist, nf, c->f[LULESH_F_P], c->f[LULESH_F_Q], c->f[LULESH_F_VOLO],
            c->f[LULESH_F_V], c->f[LULESH_F_SS], c->f[LULESH_F_ELEMMASS], c->hgcoef, c->cs[0],
            c->cs[1], c->cs[2], static_cast<R *>(c->ch[0]), static_cast<R *>(c->ch[1]),
            static_cast<R *>(c->ch[2]), c->st);
    });
    if (rc) return rc;
    return launch(c, T_NODES, [&] {
        k_nodes<R><<<grid_for(N, kBlock), kBlock, 0, c->stream>>>(
            N, c->nodeStart, c->cornerMap, c->cs[0], c->cs[1], c->cs[2],
            static_cast<const R *>(c->ch[0]), static_cast<const R *>(c->ch[1]),
            static_cast<const R *>(c->ch[2]), c->f[LULESH_F_NODALMASS], c->bcMask, nf,
            c->f[LULESH_F_FX], c->f[LULESH_F_FY], c->f[LULESH_F_FZ], c->f[LULESH_F_XDD],
            c->f[LULESH_F_YDD], c->f[LULESH_F_ZDD], c->P.u_cut, c->st);
    });
}

template <typename R>
R *grad_ptr(lulesh_ctx *c, int k) {
    if constexpr (sizeof(R) == 8)
        return c->f[LULESH_F_DELV_XI + k];
    else
        return c->g32[k];
}

template <typename R>
int enqueue_elements(lulesh_ctx *c) {
    const ElemKin ek{c->f[LULESH_F_VOLO], c->f[LULESH_F_V], c->f[LULESH_F_VNEW],
                     c->f[LULESH_F_DELV], c->f[LULESH_F_AREALG], c->f[LULESH_F_VDOV]};
    return launch(c, T_ELEMS, [&] {
        auto kern = c->streamStores
                        ? (c->structured ? k_elems<R, true, true> : k_elems<R, false, true>)
                        : (c->structured ? k_elems<R, true, false> : k_elems<R, false, false>);
        kern<<<c->eBlocks, c->tune.block_elem, 0, c->stream>>>(
            c->d_pd, c->d_eBlockProb, c->nodelist, c->nb, ek, grad_ptr<R>(c, 0),
            grad_ptr<R>(c, 1), grad_ptr<R>(c, 2), grad_ptr<R>(c, 3), grad_ptr<R>(c, 4),
            grad_ptr<R>(c, 5), c->P.v_cut, c->P.eosvmin, c->P.eosvmax, c->st);
    });
}

template <typename R, bool NT>
auto material_ordered_kernel(bool rho1, bool structured) {
    return structured ? (rho1 ? k_material_ordered<R, true, true, NT>
                              : k_material_ordered<R, false, true, NT>)
                      : (rho1 ? k_material_ordered<R, true, false, NT>
                              : k_material_ordered<R, false, false, NT>);
}

// ql / qq (CalcMonotonicQRegionForElems, lulesh.cc:1647-1648) are read back
// only by the same phase's EOS (lulesh.cc:1726-1727): stored when the
// context keeps the fields that are dead between cycles, else null (16
// bytes per element fewer: s=150 material 0.134 -> 0.128 ms).
inline double *qterm_ptr(lulesh_ctx *c, int id) { return c->store_forces ? c->f[id] : nullptr; }

template <typename R>
int enqueue_material(lulesh_ctx *c) {
    if (c->tune.material_mode != 1) { // (mode 2 outside the fused cycle: element order)
        const ElemFields ef{c->elemBC, c->lxim, c->lxip, c->letam, c->letap, c->lzetam,
                            c->lzetap, c->f[LULESH_F_VDOV], c->f[LULESH_F_ELEMMASS],
                            c->f[LULESH_F_VOLO], c->f[LULESH_F_VNEW], c->f[LULESH_F_DELV],
                            c->f[LULESH_F_AREALG], c->f[LULESH_F_E], c->f[LULESH_F_P],
                            c->f[LULESH_F_Q], qterm_ptr(c, LULESH_F_QL), qterm_ptr(c, LULESH_F_QQ),
                            c->f[LULESH_F_SS]};
        return launch(c, T_MATERIAL, [&] {
            const bool rho1 = c->P.refdens == 1.0;
            auto kern = c->streamStores ? material_ordered_kernel<R, true>(rho1, c->structured)
                                        : material_ordered_kernel<R, false>(rho1, c->structured);
            const size_t lds = size_t(c->tune.block_material) * (kMatSlots * sizeof(double) + 4) +
                               size_t(c->tune.block_material / 64) * kMapWords * 8;
            kern<<<c->oBlocks, c->tune.block_material, lds, c->stream>>>(
                c->d_pd, c->d_oBlockProb, c->erep, ef, grad_ptr<R>(c, 0), grad_ptr<R>(c, 1),
                grad_ptr<R>(c, 2), grad_ptr<R>(c, 3), grad_ptr<R>(c, 4), grad_ptr<R>(c, 5), c->P,
                c->st, c->minslots);
        });
    }
    if (c->mBlocks == 0) return 0;
    const ElemFields ef{c->elemBC, c->lxim, c->lxip, c->letam, c->letap, c->lzetam, c->lzetap,
                        c->f[LULESH_F_VDOV], c->f[LULESH_F_ELEMMASS], c->f[LULESH_F_VOLO],
                        c->f[LULESH_F_VNEW], c->f[LULESH_F_DELV], c->f[LULESH_F_AREALG],
                        c->f[LULESH_F_E], c->f[LULESH_F_P], c->f[LULESH_F_Q],
                        qterm_ptr(c, LULESH_F_QL), qterm_ptr(c, LULESH_F_QQ), c->f[LULESH_F_SS]};
    return launch(c, T_MATERIAL, [&] {
        const bool rho1 = c->P.refdens == 1.0;
        auto kern = c->structured ? (rho1 ? k_material<R, true, true> : k_material<R, false, true>)
                                  : (rho1 ? k_material<R, true, false> : k_material<R, false, false>);
        kern<<<c->mBlocks, c->tune.block_material, 0, c->stream>>>(
            c->d_pd, c->d_segs, c->d_mBlockSeg, c->matList, ef, grad_ptr<R>(c, 0), grad_ptr<R>(c, 1),
            grad_ptr<R>(c, 2), grad_ptr<R>(c, 3), grad_ptr<R>(c, 4), grad_ptr<R>(c, 5), c->P,
            c->st, c->minslots);
    });
}

// The fused elements + material pass (material_mode 2).
template <typename R>
int enqueue_elements_material(lulesh_ctx *c) {
    const ElemKin ek{c->f[LULESH_F_VOLO], c->f[LULESH_F_V], c->f[LULESH_F_VNEW],
                     c->f[LULESH_F_DELV], c->f[LULESH_F_AREALG], c->f[LULESH_F_VDOV]};
    const ElemFields ef{c->elemBC, c->lxim, c->lxip, c->letam, c->letap, c->lzetam, c->lzetap,
                        c->f[LULESH_F_VDOV], c->f[LULESH_F_ELEMMASS], c->f[LULESH_F_VOLO],
                        c->f[LULESH_F_VNEW], c->f[LULESH_F_DELV], c->f[LULESH_F_AREALG],
                        c->f[LULESH_F_E], c->f[LULESH_F_P], c->f[LULESH_F_Q],
                        qterm_ptr(c, LULESH_F_QL), qterm_ptr(c, LULESH_F_QQ), c->f[LULESH_F_SS]};
    c->grads_stale = true;
    return launch(c, T_ELEMMAT, [&] {
        const bool rho1 = c->P.refdens == 1.0;
        auto kern = c->streamStores
                        ? (rho1 ? k_elemmat_zmarch<R, true, true> : k_elemmat_zmarch<R, false, true>)
                        : (rho1 ? k_elemmat_zmarch<R, true, false>
                                : k_elemmat_zmarch<R, false, false>);
        kern<<<c->emBlocks, kEmThreads, 0, c->stream>>>(c->d_pd, c->d_emBlockProb, c->nb, ek, ef,
                                                        c->erep, c->P, c->st, c->minslots);
    });
}

// Fold the per-slot constraint minima into each problem's DevState.
int enqueue_minima(lulesh_ctx *c) {
    return launch(c, T_MINIMA, [&] {
        k_reduce_minima<<<c->nprob, kMinSlots, 0, c->stream>>>(c->st, c->minslots);
    });
}

template <typename R>
int enqueue_material_and_minima(lulesh_ctx *c) {
    const int rc = enqueue_material<R>(c);
    return rc ? rc : enqueue_minima(c);
}

// The three phases of a cycle; the constraint minima stay in the slots
// (pending) until the next k_time_increment or k_reduce_minima folds them.
int enqueue_cycle_body(lulesh_ctx *c) {
    int rc;
    const bool fused = c->tune.material_mode == 2;
    if (c->precision == LULESH_PRECISION_MIXED) {
        if ((rc = enqueue_nodal<float>(c))) return rc;
        if (fused) return enqueue_elements_material<float>(c);
        if ((rc = enqueue_elements<float>(c))) return rc;
        return enqueue_material<float>(c);
    }
    if ((rc = enqueue_nodal<double>(c))) return rc;
    if (fused) return enqueue_elements_material<double>(c);
    if ((rc = enqueue_elements<double>(c))) return rc;
    return enqueue_material<double>(c);
}

int enqueue_cycle(lulesh_ctx *c) {
    int rc = launch(c, T_TIME, [&] {
        k_time_increment<<<c->nprob, kMinSlots, 0, c->stream>>>(c->st, c->minslots);
    });
    if (rc) return rc;
    return enqueue_cycle_body(c);
}

// Capture `cycles` cycles into one executable graph, uploaded to the device
// so that its first replay costs no more than the later ones.
int capture_cycles(lulesh_ctx *c, int cycles, hipGraphExec_t *out) {
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
    e = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) return hip_fail(c, e, "hipGraphInstantiate");
    HIP_TRY(c, hipGraphUpload(*out, c->stream));
    return 0;
}

// The graph set of the current tuning: cycles_per_graph cycles, and every
// power of two below it (built together, before any timed replay).
int build_graphs(lulesh_ctx *c) {
    const int per = std::max(c->tune.cycles_per_graph, 1);
    if (!c->graph_sizes.empty() && c->graph_sizes[0] == per) return 0;
    destroy_graphs(c);
    std::vector<int> sizes{per};
    for (int p = 1; p < per; p <<= 1) sizes.push_back(p);
    std::sort(sizes.begin() + 1, sizes.end(), std::greater<int>());
    for (int n : sizes) {
        hipGraphExec_t g = nullptr;
        if (int rc = capture_cycles(c, n, &g)) {
            destroy_graphs(c);
            return rc;
        }
        c->graphs.push_back(g);
        c->graph_sizes.push_back(n);
    }
    return 0;
}

// Enqueue exactly m <= cycles_per_graph cycles: the full graph, or a sum of
// distinct power-of-two graphs (timing mode: the launches themselves).
int enqueue_cycles(lulesh_ctx *c, int m) {
    if (c->timing) {
        for (int k = 0; k < m; ++k)
            if (int rc = enqueue_cycle(c)) return rc;
        return 0;
    }
    for (size_t k = 0; k < c->graphs.size() && m > 0; ++k)
        if (c->graph_sizes[k] <= m) {
            HIP_TRY(c, hipGraphLaunch(c->graphs[k], c->stream));
            m -= c->graph_sizes[k];
        }
    return 0;
}

int read_states(lulesh_ctx *c) {
    HIP_TRY(c, hipMemcpyAsync(c->hst, c->st, sizeof(DevState) * size_t(c->nprob),
                              hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return flush_timers(c);
}

int write_states(lulesh_ctx *c) {
    HIP_TRY(c, hipMemcpyAsync(c->st, c->hst, sizeof(DevState) * size_t(c->nprob),
                              hipMemcpyHostToDevice, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return 0;
}

bool still_running(const DevState &h) {
    return !h.error && h.time < h.stoptime && h.cycle < h.its;
}

// Cycles (up to cap) that a problem in host state h certainly still runs:
// the loop condition of lulesh.cc:2374-2375 holds before each of them.
// TimeIncrement (lulesh.cc:182-229) never grows the time step by more than
// deltatimemultub per cycle, so the time before the j-th next cycle is at
// most time + dt (u + u^2 + ... + u^j); the bound is padded against
// rounding.  It only decides how far the host may queue cycles ahead of its
// next look at the device; the device evaluates the condition itself.
int64_t sure_cycles(const DevState &h, int64_t cap) {
    if (!still_running(h)) return 0;
    const double u = std::max(h.deltatimemultub, 1.0) * (1.0 + 1e-12);
    double t = h.time, dt = h.deltatime;
    int64_t j = 1;
    while (j < cap && int64_t(h.cycle) + j < int64_t(h.its)) {
        dt *= u;
        t += dt * (1.0 + 1e-12);
        if (!(t < h.stoptime)) break;
        ++j;
    }
    return j;
}

// Advance every problem until none runs.  The host queues cycles while it
// can tell from the last state it read that the first queued cycle will run
// for some problem (a problem's cycle limit is exact, its stop time is
// bounded by sure_cycles), at most cycles_per_graph at a time and never past
// the largest cycle limit; only when that is uncertain does it wait for the
// device and read the states again.  A run of k cycles towards a cycle
// limit therefore replays exactly k cycles with no wait in between; near the
// stop time it reads the state after every replay.
int run_loop(lulesh_ctx *c) {
    int rc = 0;
    if (!c->timing && (rc = build_graphs(c))) return rc;
    const int per = std::max(c->tune.cycles_per_graph, 1);
    HIP_TRY(c, hipEventRecord(c->run0, c->stream));
    int64_t queued = 0; // cycles enqueued since the host states were read
    for (;;) {
        int64_t sure = 0, maybe = 0;
        for (int pi = 0; pi < c->nprob; ++pi) {
            const DevState &h = c->hst[pi];
            if (!still_running(h)) continue;
            sure = std::max(sure, sure_cycles(h, queued + per));
            maybe = std::max(maybe, int64_t(h.its) - int64_t(h.cycle));
        }
        if (sure > queued) {
            const int m = int(std::min<int64_t>(per, maybe - queued));
            if ((rc = enqueue_cycles(c, m))) return rc;
            queued += m;
            if (!c->timing) continue;
        } else if (queued == 0) {
            break; // the states are current and no problem runs
        }
        if ((rc = read_states(c))) return rc;
        queued = 0;
    }
    // the last cycle's minima, for the time state handed back to the caller
    if ((rc = enqueue_minima(c))) return rc;
    HIP_TRY(c, hipEventRecord(c->run1, c->stream));
    HIP_TRY(c, hipEventSynchronize(c->run1));
    float ms = 0;
    HIP_TRY(c, hipEventElapsedTime(&ms, c->run0, c->run1));
    c->last_run_ms = ms;
    if ((rc = read_states(c))) return rc;
    for (int pi = 0; pi < c->nprob; ++pi)
        if (c->hst[pi].error) {
            c->err = "VolumeError in problem " + std::to_string(pi);
            return c->hst[pi].error;
        }
    return 0;
}

void to_time_state(const DevState &h, lulesh_time_state *ts) {
    ts->time = h.time;
    ts->deltatime = h.deltatime;
    ts->dtcourant = bits_to_double(h.dtc_bits);
    ts->dthydro = bits_to_double(h.dth_bits);
    ts->cycle = h.cycle;
    ts->status = h.error;
}

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
    // Face neighbours: wherever the BC switch of lulesh.cc:1460-1597 reads a
    // neighbour (bits 0 or *_COMM), it must be the adjacent element.
    const int32_t *nb[6] = {v->lxim, v->lxip, v->letam, v->letap, v->lzetam, v->lzetap};
    const int32_t masks[6] = {XI_M, XI_P, ETA_M, ETA_P, ZETA_M, ZETA_P};
    const int32_t comms[6] = {XI_M_COMM, XI_P_COMM, ETA_M_COMM, ETA_P_COMM, ZETA_M_COMM,
                              ZETA_P_COMM};
    const int64_t step[6] = {-1, 1, -s, s, -s * s, s * s};
    for (int64_t e = 0; e < E; ++e)
        for (int d = 0; d < 6; ++d) {
            const int32_t bits = v->elemBC[e] & masks[d];
                        if ((bits == 0 || bits == comms[d]) && nb[d][e] != e + step[d]) return false;
        }
    if (!v->symmX || !v->symmY || !v->symmZ) return false;
    const int64_t P = n * n;
    for (int axis = 0; axis < 3; ++axis) {
        const int32_t *set = axis == 0 ? v->symmX : axis == 1 ? v->symmY : v->symmZ;
        std::vector<int32_t> got(set, set + P), want;
        want.reserve(size_t(P));
        for (int64_t a = 0; a < n; ++a)
            for (int64_t b = 0; b < n; ++b)
                want.push_back(int32_t(axis == 0 ? (a * n + b) * n   // I == 0
                                      : axis == 1 ? (a * n) * n + b  // J == 0
                                                  : a * n + b));     // K == 0
        std::sort(got.begin(), got.end());
        std::sort(want.begin(), want.end());
        if (got != want) return false;
    }
    return true;
}

bool view_ok(const lulesh_domain_view *v) {
    return v && v->sizeX >= 1 && v->numElem == v->sizeX * v->sizeX * v->sizeX &&
           v->numNode == (v->sizeX + 1) * (v->sizeX + 1) * (v->sizeX + 1) && v->numReg >= 1 &&
           v->nodelist && v->nodeElemStart && v->nodeElemCornerList && v->regElemSize &&
           v->regElemlist && v->elemBC && v->lxim && v->lxip && v->letam && v->letap &&
           v->lzetam && v->lzetap && v->x && v->y && v->z && v->xd && v->yd && v->zd && v->e &&
           v->p && v->q && v->v && v->ss && v->nodalMass && v->volo && v->elemMass &&
           int64_t(v->numElem) * 8 < (int64_t(1) << 31);
}

// Every element in exactly one region list (lulesh-init.cc:341-445 builds
// a partition): the material kernels take each element's repetition count
// from its region, and the element-order kernel maps the EOS evaluations of
// a wave onto its elements by those counts.
bool regions_ok(const lulesh_domain_view *v) {
    std::vector<uint8_t> hit(size_t(v->numElem), 0);
    int64_t seen = 0;
    for (int r = 0; r < v->numReg; ++r) {
        const int32_t n = v->regElemSize[r];
        if (n < 0 || (n > 0 && !v->regElemlist[r])) return false;
        for (int32_t k = 0; k < n; ++k) {
            const int32_t el = v->regElemlist[r][k];
            if (el < 0 || el >= v->numElem || hit[size_t(el)]) return false;
            hit[size_t(el)] = 1;
        }
        seen += n;
    }
    return seen == v->numElem;
}

Params params_of(const lulesh_domain_view *v) {
    return Params{v->e_cut, v->p_cut, v->q_cut, v->v_cut, v->u_cut, v->hgcoef,
                  v->monoq_max_slope, v->monoq_limiter_mult, v->qlc_monoq, v->qqc_monoq, v->qqc,
                  v->eosvmax, v->eosvmin, v->pmin, v->emin, v->dvovmax, v->refdens};
}

DevState state_of(const lulesh_domain_view *v) {
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
    return s;
}

// Views of host Domains into a context.  single_view: lulesh_hip_create,
// whose context keeps fx.. and xdd.. unless LULESH_OPT_NO_NODE_FORCES.
int create_impl(const lulesh_domain_view *const *views, int nprob, int precision_opts,
                bool single_view, lulesh_ctx **out) {
    if (!out) return LULESH_INVALID_ARGUMENT;
    *out = nullptr;
    const int precision = precision_opts & 0xff;
    const int forces_opts = precision_opts & (LULESH_OPT_NODE_FORCES | LULESH_OPT_NO_NODE_FORCES);
    if ((precision_opts & ~(0xff | LULESH_OPT_NODE_FORCES | LULESH_OPT_NO_NODE_FORCES |
                            LULESH_OPT_GENERIC)) ||
        forces_opts == (LULESH_OPT_NODE_FORCES | LULESH_OPT_NO_NODE_FORCES) ||
        (!single_view && (forces_opts & LULESH_OPT_NO_NODE_FORCES)) ||
        (precision != LULESH_PRECISION_FP64 && precision != LULESH_PRECISION_MIXED) ||
        nprob < 1 || !views)
        return LULESH_INVALID_ARGUMENT;
    int64_t totE = 0, totN = 0;
    for (int pi = 0; pi < nprob; ++pi) {
        if (!view_ok(views[pi]) || !regions_ok(views[pi])) return LULESH_INVALID_ARGUMENT;
        totE += views[pi]->numElem;
        totN += views[pi]->numNode;
        if (std::memcmp(&views[pi]->e_cut, &views[0]->e_cut, sizeof(double) * 19) != 0)
            return LULESH_INVALID_ARGUMENT; // run constants are shared by the kernels
    }
    if (totE * 8 >= (int64_t(1) << 31) || totN >= (int64_t(1) << 31))
        return LULESH_INVALID_ARGUMENT;
    bool structured = !(precision_opts & LULESH_OPT_GENERIC);
    for (int pi = 0; pi < nprob && structured; ++pi) structured = is_structured_cube(views[pi]);
    if (!structured && nprob > 1) return LULESH_INVALID_ARGUMENT; // batches need cubes

    auto *c = new lulesh_ctx();
    auto fail = [&](int rc) {
        lulesh_hip_destroy(c);
        return rc;
    };
    c->precision = precision;
    // (the indexed kernels always write them)
    c->store_forces = !structured || (precision_opts & LULESH_OPT_NODE_FORCES) != 0 ||
                      (single_view && !(precision_opts & LULESH_OPT_NO_NODE_FORCES));
    c->structured = structured;
    c->nprob = nprob;
    c->totE = totE;
    c->totN = totN;
    c->P = params_of(views[0]);
    c->hgcoef = views[0]->hgcoef;
    c->timers = {{"time_increment"}, {"nodal"}, {"nodes"}, {"elements"}, {"material"},
                 {"minima"}, {"elements_material"}};
    int rc;
    int dev = 0;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&c->run0) != hipSuccess || hipEventCreate(&c->run1) != hipSuccess ||
        hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&c->cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        return fail(LULESH_HIP_ERROR);
    {
        int32_t eo = 0, no = 0;
        for (int pi = 0; pi < nprob; ++pi) {
            ProbHost ph;
            ph.s = views[pi]->sizeX;
            ph.E = views[pi]->numElem;
            ph.N = views[pi]->numNode;
            ph.numReg = views[pi]->numReg;
            ph.cost = views[pi]->cost;
            ph.elemOff = eo;
            ph.nodeOff = no;
            eo += ph.E;
            no += ph.N;
            c->probs.push_back(ph);
        }
    }
    const size_t E = size_t(totE), N = size_t(totN);
    if ((rc = alloc_node_state(c, structured))) return fail(rc);
    // Fields, concatenated over problems.
    for (int id = 0; id < LULESH_F_COUNT; ++id) {
        const size_t n = is_node_field(id) ? N : E;
        if (precision == LULESH_PRECISION_MIXED && is_grad_field(id)) {
            if ((rc = dalloc(c, &c->g32[grad_slot(id)], n))) return fail(rc);
            continue;
        }
        if (!is_state_node_field(id) && (rc = dalloc(c, &c->f[id], n))) return fail(rc);
        for (int pi = 0; pi < nprob; ++pi) {
            const ProbHost &ph = c->probs[size_t(pi)];
            const size_t off = is_node_field(id) ? size_t(ph.nodeOff) : size_t(ph.elemOff);
            const size_t cnt = is_node_field(id) ? size_t(ph.N) : size_t(ph.E);
            if ((rc = upload(c, c->f[id] + off, (const double *)view_field(views[pi], id), cnt)))
                return fail(rc);
        }
    }
    // Face neighbours and BC flags, shifted to concatenated element indices.
    {
        int32_t **dst[7] = {&c->lxim, &c->lxip, &c->letam, &c->letap, &c->lzetam, &c->lzetap,
                            &c->elemBC};
        std::vector<int32_t> buf(E);
        for (int k = 0; k < 7; ++k) {
            for (int pi = 0; pi < nprob; ++pi) {
                const lulesh_domain_view *v = views[pi];
                const int32_t *src[7] = {v->lxim, v->lxip, v->letam, v->letap,
                                         v->lzetam, v->lzetap, v->elemBC};
                const int32_t off = c->probs[size_t(pi)].elemOff;
                for (int32_t i = 0; i < v->numElem; ++i)
                    buf[size_t(off + i)] = k < 6 ? src[k][i] + off : src[k][i];
            }
            if ((rc = dalloc(c, dst[k], E)) || (rc = upload(c, *dst[k], buf.data(), E)))
                return fail(rc);
        }
    }
    // Region lists grouped by EOS repetition count (lulesh.cc:2159-2169),
    // ascending within a group, as concatenated element indices.
    {
        std::vector<int32_t> all;
        all.reserve(E);
        for (int pi = 0; pi < nprob; ++pi) {
            const lulesh_domain_view *v = views[pi];
            ProbHost &ph = c->probs[size_t(pi)];
            std::vector<int32_t> reps(size_t(v->numReg));
            for (int r = 0; r < v->numReg; ++r) reps[size_t(r)] = region_rep(r, v->numReg, v->cost);
            std::vector<int32_t> uniq(reps);
            std::sort(uniq.begin(), uniq.end());
            uniq.erase(std::unique(uniq.begin(), uniq.end()), uniq.end());
            int64_t seen = 0;
            for (int32_t rv : uniq) {
                const int32_t off = int32_t(all.size());
                for (int r = 0; r < v->numReg; ++r)
                    if (reps[size_t(r)] == rv)
                        for (int32_t k = 0; k < v->regElemSize[r]; ++k)
                            all.push_back(v->regElemlist[r][k] + ph.elemOff);
                std::sort(all.begin() + off, all.end());
                ph.segs.push_back({pi, rv, off, int32_t(all.size()) - off, 0});
                seen += int32_t(all.size()) - off;
            }
            if (seen != ph.E) return fail(LULESH_INVALID_ARGUMENT);
        }
        if ((rc = dalloc(c, &c->matList, E)) || (rc = upload(c, c->matList, all.data(), E)))
            return fail(rc);
        // repetitions per element for the element-order material kernel
        std::vector<uint16_t> rep(E, 0);
        bool fits = true;
        for (int pi = 0; pi < nprob; ++pi) {
            const lulesh_domain_view *v = views[pi];
            for (int r = 0; r < v->numReg; ++r) {
                const int32_t rv = region_rep(r, v->numReg, v->cost);
                fits = fits && rv >= 1 && rv <= 65535;
                for (int32_t k = 0; k < v->regElemSize[r]; ++k)
                    rep[size_t(v->regElemlist[r][k] + c->probs[size_t(pi)].elemOff)] =
                        uint16_t(rv);
            }
        }
        if (fits && ((rc = dalloc(c, &c->erep, E)) || (rc = upload(c, c->erep, rep.data(), E))))
            return fail(rc);
    }
    {
        const size_t ns = size_t(nprob) * kMinSlots * kSlotU64;
        std::vector<unsigned long long> init(ns, kMinInit);
        if ((rc = dalloc(c, &c->minslots, ns)) || (rc = upload(c, c->minslots, init.data(), ns)))
            return fail(rc);
    }
    if (!structured) {
        const lulesh_domain_view *v = views[0];
        if ((rc = dalloc(c, &c->nodelist, 8 * E)) ||
            (rc = upload(c, c->nodelist, v->nodelist, 8 * E)))
            return fail(rc);
        // Corner CSR with corner ids remapped to the corner-major layout of
        // the force arrays: corner 8*e+k lives at k*E+e.
        std::vector<int32_t> cmap(8 * E);
        for (size_t i = 0; i < 8 * E; ++i) {
            const int32_t cl = v->nodeElemCornerList[i];
            cmap[i] = (cl & 7) * int32_t(E) + (cl >> 3);
        }
        if ((rc = dalloc(c, &c->nodeStart, N + 1)) ||
            (rc = upload(c, c->nodeStart, v->nodeElemStart, N + 1)) ||
            (rc = dalloc(c, &c->cornerMap, 8 * E)) ||
            (rc = upload(c, c->cornerMap, cmap.data(), 8 * E)))
            return fail(rc);
        // Symmetry-plane node sets as a per-node mask (bit0 x, bit1 y, bit2 z).
        std::vector<uint8_t> mask(N, 0);
        const size_t P = size_t(v->sizeX + 1) * (v->sizeX + 1);
        const int32_t *sets[3] = {v->symmX, v->symmY, v->symmZ};
        for (int a = 0; a < 3; ++a)
            if (sets[a])
                for (size_t k = 0; k < P; ++k) mask[size_t(sets[a][k])] |= uint8_t(1u << a);
        if ((rc = dalloc(c, &c->bcMask, N)) || (rc = upload(c, c->bcMask, mask.data(), N)))
            return fail(rc);
        for (int k = 0; k < 3; ++k) {
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
    }
    // Mixed form: narrow any provided gradients into FP32 storage.
    if (precision == LULESH_PRECISION_MIXED) {
        double *tmp = nullptr;
        if (hipMalloc(&tmp, E * sizeof(double)) != hipSuccess) return fail(LULESH_HIP_ERROR);
        for (int id = LULESH_F_DELV_XI; id <= LULESH_F_DELX_ZETA; ++id) {
            (void)hipMemset(tmp, 0, E * sizeof(double));
            for (int pi = 0; pi < nprob; ++pi)
                if (const double *src = view_field(views[pi], id))
                    (void)hipMemcpy(tmp + c->probs[size_t(pi)].elemOff, src,
                                    size_t(c->probs[size_t(pi)].E) * sizeof(double),
                                    hipMemcpyHostToDevice);
            k_narrow<float><<<grid_for(int64_t(E), 256), 256, 0, c->stream>>>(
                int32_t(E), tmp, c->g32[grad_slot(id)]);
        }
        (void)hipStreamSynchronize(c->stream);
        (void)hipFree(tmp);
    }
    if ((rc = dalloc(c, &c->st, size_t(nprob)))) return fail(rc);
    if (hipHostMalloc(reinterpret_cast<void **>(&c->hst), sizeof(DevState) * size_t(nprob)) !=
        hipSuccess)
        return fail(LULESH_HIP_ERROR);
    for (int pi = 0; pi < nprob; ++pi) c->init.push_back(state_of(views[pi]));
    if (hipMemcpy(c->st, c->init.data(), sizeof(DevState) * size_t(nprob),
                  hipMemcpyHostToDevice) != hipSuccess)
        return fail(LULESH_HIP_ERROR);
    if ((rc = configure(c))) return fail(rc);
    if (hipDeviceSynchronize() != hipSuccess) return fail(LULESH_HIP_ERROR);
    *out = c;
    return 0;
}

// Host <-> device copy of one problem's field.  Node state fields live in
// the half of the double buffer named by the problem's DevState.
int field_copy(lulesh_ctx *c, int prob, int id, void *host, size_t count, bool down) {
    if (!c || !host || id < 0 || id >= LULESH_F_COUNT || prob < 0 || prob >= c->nprob ||
        !c->lanes.empty())
        return LULESH_INVALID_ARGUMENT;
    const ProbHost &ph = c->probs[size_t(prob)];
    const size_t n = is_node_field(id) ? size_t(ph.N) : size_t(ph.E);
    const size_t off = is_node_field(id) ? size_t(ph.nodeOff) : size_t(ph.elemOff);
    if (count > n) return LULESH_INVALID_ARGUMENT;
    if (is_dead_field(id) && !c->store_forces) {
        c->err = "fx/fy/fz, xdd/ydd/zdd and ql/qq are not kept by this context "
                 "(LULESH_OPT_NODE_FORCES was not given)";
        return LULESH_INVALID_ARGUMENT;
    }
    if (down && is_grad_field(id) && c->grads_stale) {
        c->err = "the Q gradients were not stored by the last cycle (fused elements + "
                 "material pass); an elements phase stores them";
        return LULESH_INVALID_ARGUMENT;
    }
    int rc = read_states(c);
    if (rc) return rc;
    if (c->precision == LULESH_PRECISION_MIXED && is_grad_field(id)) {
        double *tmp = nullptr;
        HIP_TRY(c, hipMalloc(&tmp, std::max<size_t>(count, 1) * sizeof(double)));
        hipError_t e;
        float *g = c->g32[grad_slot(id)] + off;
        if (down) {
            k_widen<float><<<grid_for(int64_t(count), 256), 256, 0, c->stream>>>(int32_t(count),
                                                                                 g, tmp);
            e = hipMemcpyAsync(host, tmp, count * sizeof(double), hipMemcpyDeviceToHost,
                               c->stream);
        } else {
            e = hipMemcpyAsync(tmp, host, count * sizeof(double), hipMemcpyHostToDevice,
                               c->stream);
            k_narrow<float><<<grid_for(int64_t(count), 256), 256, 0, c->stream>>>(int32_t(count),
                                                                                  tmp, g);
        }
        const hipError_t e2 = hipStreamSynchronize(c->stream);
        (void)hipFree(tmp);
        if (e != hipSuccess) return hip_fail(c, e, "field copy");
        if (e2 != hipSuccess) return hip_fail(c, e2, "field copy");
        return 0;
    }
    double *base = c->f[id];
    if (is_state_node_field(id)) {
        const DevState &h = c->hst[prob];
        base = c->nb.b[h.cur ^ h.fresh][id - LULESH_F_X];
    }
    double *dev = base + off;
    HIP_TRY(c, hipMemcpyAsync(down ? host : static_cast<void *>(dev),
                              down ? static_cast<const void *>(dev) : host,
                              count * sizeof(double),
                              down ? hipMemcpyDeviceToHost : hipMemcpyHostToDevice, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return 0;
}

// Copies of every allocation that holds state, for the autotuner's restore.
struct Snapshot {
    std::vector<std::pair<void *, size_t>> regions;
    std::vector<void *> copies;
};

} // namespace

// ---------------------------------------------------------------------------
// Sedov set-up in HBM (the Domain constructor, lulesh-init.cc:10-183, single
// rank, structured cube): the same values the host set-up produces, bit for
// bit, without the O(E) host loops and the upload.  Separate rounding in
// both arithmetic forms (contract(off) here and in elem_volume_rounded).
// ---------------------------------------------------------------------------
// Node coordinates from their lattice indices, lulesh-init.cc:207-227.
__global__ void k_sedov_nodes(int32_t s, double *__restrict__ x, double *__restrict__ y,
                              double *__restrict__ z) {
#pragma clang fp contract(off)
    const int n = s + 1;
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n * n * n) return;
    const int col = g % n, row = (g / n) % n, plane = g / (n * n);
    x[g] = 1.125 * double(col) / double(s);
    y[g] = 1.125 * double(row) / double(s);
    z[g] = 1.125 * double(plane) / double(s);
}

// Element volumes = reference volumes = masses (lulesh-init.cc:148-160),
// v = 1, and the single-rank boundary flags (lulesh-init.cc:548-607).
__global__ void k_sedov_elems(int32_t s, const double *__restrict__ x,
                              const double *__restrict__ y, const double *__restrict__ z,
                              double *__restrict__ volo, double *__restrict__ elemMass,
                              double *__restrict__ v, int32_t *__restrict__ elemBC) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= s * s * s) return;
    const int n = s + 1;
    const int i = e % s, j = (e / s) % s, k = e / (s * s);
    const int b = (k * n + j) * n + i;
    const int nd[8] = {b, b + 1, b + n + 1, b + n, b + n * n, b + n * n + 1,
                       b + n * n + n + 1, b + n * n + n};
    double xl[8], yl[8], zl[8];
#pragma unroll
    for (int a = 0; a < 8; ++a) {
        xl[a] = x[nd[a]];
        yl[a] = y[nd[a]];
        zl[a] = z[nd[a]];
    }
    const double vol = elem_volume_rounded<double>(xl, yl, zl);
    volo[e] = vol;
    elemMass[e] = vol;
    v[e] = 1.0;
    int32_t bc = 0;
    if (k == 0) bc |= ZETA_M_SYMM;
    if (k == s - 1) bc |= ZETA_P_FREE;
    if (j == 0) bc |= ETA_M_SYMM;
    if (j == s - 1) bc |= ETA_P_FREE;
    if (i == 0) bc |= XI_M_SYMM;
    if (i == s - 1) bc |= XI_P_FREE;
    elemBC[e] = bc;
}

// Nodal mass = sum of volume/8 over the node's elements in ascending
// element order, starting from 0 (the constructor's element-by-element
// accumulation, lulesh-init.cc:161-166).
__global__ void k_sedov_nodal_mass(int32_t s, const double *__restrict__ volo,
                                   double *__restrict__ nodalMass) {
#pragma clang fp contract(off)
    const int n = s + 1;
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n * n * n) return;
    const int I = g % n, J = (g / n) % n, K = g / (n * n);
    double m = 0.0;
    for (int k = K - 1; k <= K; ++k)
        for (int j = J - 1; j <= J; ++j)
            for (int i = I - 1; i <= I; ++i)
                if (k >= 0 && k < s && j >= 0 && j < s && i >= 0 && i < s)
                    m += volo[(k * s + j) * s + i] / 8.0;
    nodalMass[g] = m;
}

// EOS repetitions per element from the region runs: run r covers elements
// [start[r], start[r+1]) of region region[r] (1-based), lulesh.cc:2159-2169.
__global__ void k_sedov_reps(int32_t E, int32_t nruns, const int32_t *__restrict__ start,
                             const int32_t *__restrict__ region, int32_t nr, int32_t cost,
                             uint16_t *__restrict__ rep) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= E) return;
    int lo = 0, hi = nruns - 1; // last run with start <= e
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (start[mid] <= e) lo = mid;
        else hi = mid - 1;
    }
    rep[e] = uint16_t(region_rep(region[lo] - 1, nr, cost));
}

// Set-up in HBM of one or more Sedov problems (sizes[p], regions[p], the
// deposit of element 0 scaled by factors[p] as lulesh_domain_scale_energy
// does), concatenated like a batch of host Domains.
int create_sedov_impl(const int32_t *sizes, const int32_t *regions, const double *factors,
                      int32_t nprob, int32_t balance, int32_t cost, int precision_opts,
                      lulesh_ctx **out) {
    if (!out) return LULESH_INVALID_ARGUMENT;
    *out = nullptr;
    const int precision = precision_opts & 0xff;
    if ((precision_opts & ~(0xff | LULESH_OPT_NODE_FORCES)) ||
        (precision != LULESH_PRECISION_FP64 && precision != LULESH_PRECISION_MIXED) ||
        nprob < 1 || !sizes || !regions || !factors || balance < 0 || cost < 0)
        return LULESH_INVALID_ARGUMENT;
    int64_t totE = 0, totN = 0;
    for (int pi = 0; pi < nprob; ++pi) {
        const int32_t s = sizes[pi], nr = regions[pi];
        if (s < 1 || nr < 1 || !(factors[pi] > 0)) return LULESH_INVALID_ARGUMENT;
        for (int32_t r = 0; r < nr; ++r) {
            const int32_t rv = region_rep(r, nr, cost);
            if (rv < 1 || rv > 65535) return LULESH_INVALID_ARGUMENT;
        }
        totE += int64_t(s) * s * s;
        totN += int64_t(s + 1) * (s + 1) * (s + 1);
        if (totE * 8 >= (int64_t(1) << 31) || totN >= (int64_t(1) << 31))
            return LULESH_INVALID_ARGUMENT;
    }
    auto *c = new lulesh_ctx();
    auto fail = [&](int rc) {
        lulesh_hip_destroy(c);
        return rc;
    };
    using D = lulesh::Domain;
    c->precision = precision;
    c->store_forces = (precision_opts & LULESH_OPT_NODE_FORCES) != 0;
    c->structured = true;
    c->nprob = nprob;
    c->totE = totE;
    c->totN = totN;
    c->P = Params{D::m_e_cut, D::m_p_cut, D::m_q_cut, D::m_v_cut, D::m_u_cut, D::m_hgcoef,
                  D::m_monoq_max_slope, D::m_monoq_limiter_mult, D::m_qlc_monoq,
                  D::m_qqc_monoq, D::m_qqc, D::m_eosvmax, D::m_eosvmin, D::m_pmin, D::m_emin,
                  D::m_dvovmax, D::m_refdens};
    c->hgcoef = D::m_hgcoef;
    c->timers = {{"time_increment"}, {"nodal"}, {"nodes"}, {"elements"}, {"material"},
                 {"minima"}, {"elements_material"}};
    int dev = 0;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&c->run0) != hipSuccess || hipEventCreate(&c->run1) != hipSuccess ||
        hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&c->cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        return fail(LULESH_HIP_ERROR);
    {
        int32_t eo = 0, no = 0;
        for (int pi = 0; pi < nprob; ++pi) {
            ProbHost ph;
            ph.s = sizes[pi];
            ph.E = ph.s * ph.s * ph.s;
            ph.N = (ph.s + 1) * (ph.s + 1) * (ph.s + 1);
            ph.numReg = regions[pi];
            ph.cost = cost;
            ph.elemOff = eo;
            ph.nodeOff = no;
            eo += ph.E;
            no += ph.N;
            c->probs.push_back(ph);
        }
    }
    const size_t E = size_t(totE), N = size_t(totN);
    int rc;
    // every field zero-filled (the constructor's allocations), then set
    if ((rc = alloc_node_state(c, true))) return fail(rc);
    for (int id = 0; id < LULESH_F_COUNT; ++id) {
        const size_t n = is_node_field(id) ? N : E;
        if (precision == LULESH_PRECISION_MIXED && is_grad_field(id)) {
            if ((rc = dalloc(c, &c->g32[grad_slot(id)], n))) return fail(rc);
            continue;
        }
        if (!is_state_node_field(id) && (rc = dalloc(c, &c->f[id], n))) return fail(rc);
    }
    if ((rc = dalloc(c, &c->elemBC, E)) || (rc = dalloc(c, &c->erep, E))) return fail(rc);
    {
        const size_t ns = size_t(nprob) * kMinSlots * kSlotU64;
        std::vector<unsigned long long> init(ns, kMinInit);
        if ((rc = dalloc(c, &c->minslots, ns)) || (rc = upload(c, c->minslots, init.data(), ns)))
            return fail(rc);
    }
    // region runs of every problem from the reference's rand() stream
    // (O(runs) on the host), as concatenated element indices
    std::vector<int32_t> rstart, rregion, rfirst{0};
    {
        std::vector<lulesh::RegionRun> runs;
        for (int pi = 0; pi < nprob; ++pi) {
            const ProbHost &ph = c->probs[size_t(pi)];
            lulesh::SedovRegionRuns(ph.E, ph.numReg, balance, runs);
            for (const auto &r : runs) {
                rstart.push_back(r.start);
                rregion.push_back(r.region);
            }
            rfirst.push_back(int32_t(rstart.size()));
        }
    }
    int32_t *d_rs = nullptr, *d_rr = nullptr;
    if ((rc = dalloc(c, &d_rs, rstart.size())) ||
        (rc = upload(c, d_rs, rstart.data(), rstart.size())) ||
        (rc = dalloc(c, &d_rr, rregion.size())) ||
        (rc = upload(c, d_rr, rregion.data(), rregion.size())))
        return fail(rc);
    // the zero fills and uploads above are complete before the set-up kernels
    if (hipDeviceSynchronize() != hipSuccess) return fail(LULESH_HIP_ERROR);
    for (int pi = 0; pi < nprob; ++pi) {
        const ProbHost &ph = c->probs[size_t(pi)];
        const size_t eo = size_t(ph.elemOff), no = size_t(ph.nodeOff);
        const int gN = grid_for(ph.N, 256), gE = grid_for(ph.E, 256);
        k_sedov_nodes<<<gN, 256, 0, c->stream>>>(ph.s, c->f[LULESH_F_X] + no,
                                                   c->f[LULESH_F_Y] + no, c->f[LULESH_F_Z] + no);
        k_sedov_elems<<<gE, 256, 0, c->stream>>>(
            ph.s, c->f[LULESH_F_X] + no, c->f[LULESH_F_Y] + no, c->f[LULESH_F_Z] + no,
            c->f[LULESH_F_VOLO] + eo, c->f[LULESH_F_ELEMMASS] + eo, c->f[LULESH_F_V] + eo,
            c->elemBC + eo);
        k_sedov_nodal_mass<<<gN, 256, 0, c->stream>>>(ph.s, c->f[LULESH_F_VOLO] + eo,
                                                        c->f[LULESH_F_NODALMASS] + no);
        k_sedov_reps<<<gE, 256, 0, c->stream>>>(
            ph.E, rfirst[size_t(pi) + 1] - rfirst[size_t(pi)], d_rs + rfirst[size_t(pi)],
            d_rr + rfirst[size_t(pi)], ph.numReg, cost, c->erep + eo);
        HIP_TRY(c, hipGetLastError());
    }
    if ((rc = dalloc(c, &c->st, size_t(nprob)))) return fail(rc);
    if (hipHostMalloc(reinterpret_cast<void **>(&c->hst), sizeof(DevState) * size_t(nprob)) !=
        hipSuccess)
        return fail(LULESH_HIP_ERROR);
    for (int pi = 0; pi < nprob; ++pi) {
        const ProbHost &ph = c->probs[size_t(pi)];
        // the deposit of element 0 and dt0, lulesh-init.cc:172-181 (scaled as
        // lulesh_domain_scale_energy scales a host Domain)
        const double einit = lulesh::SedovInitialEnergy(ph.s) * factors[pi];
        if ((rc = upload(c, c->f[LULESH_F_E] + ph.elemOff, &einit, 1))) return fail(rc);
        // run defaults, lulesh-init.cc:135-145
        DevState st{};
        st.time = 0.0;
        st.deltatime = lulesh::SedovInitialDeltaTime(ph.s, einit);
        st.stoptime = 1.0e-2;
        st.dtfixed = -1.0e-6;
        st.deltatimemultlb = 1.1;
        st.deltatimemultub = 1.2;
        st.dtmax = 1.0e-2;
        st.dtc_bits = host_dbits(1.0e+20);
        st.dth_bits = host_dbits(1.0e+20);
        c->init.push_back(st);
    }
    if (hipMemcpy(c->st, c->init.data(), sizeof(DevState) * size_t(nprob),
                  hipMemcpyHostToDevice) != hipSuccess)
        return fail(LULESH_HIP_ERROR);
    if ((rc = configure(c))) return fail(rc);
    if (hipStreamSynchronize(c->stream) != hipSuccess) return fail(LULESH_HIP_ERROR);
    *out = c;
    return 0;
}

// Streams requested by the options (LULESH_OPT_STREAMS), at most one per
// problem, 1 if none.
int requested_streams(int precision_opts, int nprob) {
    const int k = std::min((precision_opts >> 12) & 0xf, nprob);
    return k < 1 ? 1 : k;
}

// Deal nprob problems of E[p] elements out over k lanes: largest first, each
// to the lane with the fewest elements so far (ties: the lowest lane); each
// lane keeps its problems in ascending order.
std::vector<std::vector<int32_t>> deal_problems(const std::vector<int64_t> &E, int k) {
    std::vector<int32_t> order(E.size());
    for (size_t p = 0; p < E.size(); ++p) order[p] = int32_t(p);
    std::stable_sort(order.begin(), order.end(),
                     [&](int32_t a, int32_t b) { return E[size_t(a)] > E[size_t(b)]; });
    std::vector<std::vector<int32_t>> lanes(static_cast<size_t>(k));
    std::vector<int64_t> load(static_cast<size_t>(k), 0);
    for (int32_t p : order) {
        const size_t l = size_t(std::min_element(load.begin(), load.end()) - load.begin());
        lanes[l].push_back(p);
        load[l] += E[size_t(p)];
    }
    for (auto &l : lanes) std::sort(l.begin(), l.end());
    return lanes;
}

// A group over `parts` (lane -> global problem indices) whose lanes are made
// by `make(lane, out)`.
template <typename Make>
int make_group(const std::vector<std::vector<int32_t>> &parts, int nprob, Make &&make,
               lulesh_ctx **out) {
    *out = nullptr;
    auto *g = new lulesh_ctx();
    g->nprob = nprob;
    g->lane_of.assign(size_t(nprob), 0);
    g->local_of.assign(size_t(nprob), 0);
    for (size_t l = 0; l < parts.size(); ++l) {
        lulesh_ctx *lane = nullptr;
        const int rc = make(l, &lane);
        if (rc) {
            lulesh_hip_destroy(g);
            return rc;
        }
        g->lanes.push_back(lane);
        g->bytes += lane->bytes;
        for (size_t k = 0; k < parts[l].size(); ++k) {
            g->lane_of[size_t(parts[l][k])] = int32_t(l);
            g->local_of[size_t(parts[l][k])] = int32_t(k);
        }
    }
    g->precision = g->lanes[0]->precision;
    g->structured = g->lanes[0]->structured;
    g->tune = g->lanes[0]->tune;
    *out = g;
    return 0;
}

// ---------------------------------------------------------------------------
// Entry points (declared by lulesh_api.inc; capi.cc forwards the C-ABI here)
// ---------------------------------------------------------------------------
int lulesh_hip_create(const lulesh_domain_view *v, int precision_opts, lulesh_ctx **out) {
    const lulesh_domain_view *views[1] = {v};
    return create_impl(views, 1, precision_opts, true, out);
}

int lulesh_hip_batch_create(const lulesh_domain_view *const *views, int32_t nprob,
                            int precision_opts, lulesh_ctx **out) {
    const int k = requested_streams(precision_opts, nprob);
    const int opts = precision_opts & ~LULESH_OPT_STREAMS(0xf);
    if (k == 1 || !views || !out) return create_impl(views, nprob, opts, false, out);
    std::vector<int64_t> E(static_cast<size_t>(nprob));
    for (int p = 0; p < nprob; ++p) {
        if (!views[p]) return LULESH_INVALID_ARGUMENT;
        E[size_t(p)] = views[p]->numElem;
    }
    const auto parts = deal_problems(E, k);
    return make_group(parts, nprob, [&](size_t l, lulesh_ctx **lane) {
        std::vector<const lulesh_domain_view *> sub;
        for (int32_t p : parts[l]) sub.push_back(views[p]);
        return create_impl(sub.data(), int(sub.size()), opts, false, lane);
    }, out);
}

int lulesh_hip_create_sedov(int32_t s, int32_t nr, int32_t balance, int32_t cost,
                            int precision_opts, lulesh_ctx **out) {
    const double one = 1.0;
    return create_sedov_impl(&s, &nr, &one, 1, balance, cost, precision_opts, out);
}

int lulesh_hip_batch_create_sedov(const int32_t *sizes, const int32_t *regions,
                                  const double *factors, int32_t nprob, int32_t balance,
                                  int32_t cost, int precision_opts, lulesh_ctx **out) {
    const int k = requested_streams(precision_opts, nprob);
    const int opts = precision_opts & ~LULESH_OPT_STREAMS(0xf);
    if (k == 1 || !sizes || !regions || !factors || !out)
        return create_sedov_impl(sizes, regions, factors, nprob, balance, cost, opts, out);
    std::vector<int64_t> E(static_cast<size_t>(nprob));
    for (int p = 0; p < nprob; ++p) E[size_t(p)] = int64_t(sizes[p]) * sizes[p] * sizes[p];
    const auto parts = deal_problems(E, k);
    return make_group(parts, nprob, [&](size_t l, lulesh_ctx **lane) {
        std::vector<int32_t> ss, rr;
        std::vector<double> ff;
        for (int32_t p : parts[l]) {
            ss.push_back(sizes[p]);
            rr.push_back(regions[p]);
            ff.push_back(factors[p]);
        }
        return create_sedov_impl(ss.data(), rr.data(), ff.data(), int32_t(ss.size()), balance,
                                 cost, opts, lane);
    }, out);
}

int lulesh_hip_get_time_state(lulesh_ctx *c, lulesh_time_state *ts) {
    if (!c || !ts || c->nprob != 1) return LULESH_INVALID_ARGUMENT;
    if (int rc = read_states(c)) return rc;
    to_time_state(c->hst[0], ts);
    return 0;
}

int lulesh_hip_report(lulesh_ctx *c, double out[4]) {
    if (!c || !out || c->nprob != 1) return LULESH_INVALID_ARGUMENT;
    const int32_t nx = c->probs[0].s;
    std::vector<double> e(size_t(nx) * nx);
    if (int rc = field_copy(c, 0, LULESH_F_E, e.data(), e.size(), true)) return rc;
    lulesh::ReportNumbers(e.data(), nx, out);
    return 0;
}

int32_t lulesh_hip_num_problems(const lulesh_ctx *c) { return c ? c->nprob : 0; }

int lulesh_hip_leapfrog(lulesh_ctx *c, double deltatime, double *dtcourant, double *dthydro) {
    if (!c || c->nprob != 1) return LULESH_INVALID_ARGUMENT;
    int rc = launch(c, T_TIME, [&] {
        k_begin_cycle<<<1, 1, 0, c->stream>>>(c->st, deltatime, 1, 1);
    });
    if (rc) return rc;
    if ((rc = enqueue_cycle_body(c)) || (rc = enqueue_minima(c))) return rc;
    if ((rc = read_states(c))) return rc;
    if (c->hst->error) {
        c->err = "VolumeError";
        return c->hst->error;
    }
    if (dtcourant) *dtcourant = bits_to_double(c->hst->dtc_bits);
    if (dthydro) *dthydro = bits_to_double(c->hst->dth_bits);
    return 0;
}

int lulesh_hip_phase(lulesh_ctx *c, int phase, double deltatime) {
    if (!c || c->nprob != 1 || phase < LULESH_PHASE_NODAL || phase > LULESH_PHASE_CONSTRAINTS)
        return LULESH_INVALID_ARGUMENT;
    // the nodal phase starts a new cycle (and adopts the previous node
    // state); the later phases belong to the same cycle
    int rc = launch(c, T_TIME, [&] {
        k_begin_cycle<<<1, 1, 0, c->stream>>>(c->st, deltatime, phase >= LULESH_PHASE_MATERIAL,
                                              phase == LULESH_PHASE_NODAL);
    });
    if (rc) return rc;
    const bool mixed = c->precision == LULESH_PRECISION_MIXED;
    switch (phase) {
    case LULESH_PHASE_NODAL: rc = mixed ? enqueue_nodal<float>(c) : enqueue_nodal<double>(c); break;
    case LULESH_PHASE_ELEMENTS:
        rc = mixed ? enqueue_elements<float>(c) : enqueue_elements<double>(c);
        if (rc == 0) c->grads_stale = false;
        break;
    case LULESH_PHASE_MATERIAL:
        rc = mixed ? enqueue_material_and_minima<float>(c)
                   : enqueue_material_and_minima<double>(c);
        break;
    default:
        k_constraints<<<grid_for(c->totE, kBlock), kBlock, 0, c->stream>>>(
            int32_t(c->totE), c->f[LULESH_F_SS], c->f[LULESH_F_VDOV], c->f[LULESH_F_AREALG], c->P,
            c->minslots);
        HIP_TRY(c, hipGetLastError());
        rc = enqueue_minima(c);
    }
    if (rc) return rc;
    if ((rc = read_states(c))) return rc;
    return c->hst->error;
}

#if !LULESH_FUSED_FORM
// ---------------------------------------------------------------------------
// Constraints without allocation (lulesh_hip_constraints_noalloc): the
// kernels of the constraints phase (k_constraints into per-slot atomic
// minima, then the slot fold) with the slots in this file-scope array
// instead of a context's allocation.  Each device has its own copy; its
// address is looked up once per device.  Calls on one device share it:
// calls that overlap on one device (different streams, or host threads)
// race.
// ---------------------------------------------------------------------------
static __device__ __align__(256) unsigned long long scratch[kMinSlots * kSlotU64];
// largest element count of lulesh_hip_constraints_noalloc (int32 indices and
// a grid of 256-thread workgroups)
constexpr int32_t kNoAllocMaxElems = INT32_MAX - 256;

namespace {
constexpr int kMaxDevices = 64;
unsigned long long *g_scratch_ptr[kMaxDevices] = {};
std::mutex g_scratch_mu;

// The device address of `scratch` on the current device (looked up once).
hipError_t scratch_on_device(unsigned long long **out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch_ptr[dev] = static_cast<unsigned long long *>(p);
    }
    *out = g_scratch_ptr[dev];
    return hipSuccess;
}
} // namespace

int lulesh_hip_constraints_noalloc(const double *ss, const double *vdov, const double *arealg,
                                   int32_t numElem, double qqc, double dvovmax, double *out,
                                   void *stream) {
    if (!out || numElem < 0 || numElem > kNoAllocMaxElems ||
        (numElem > 0 && (!ss || !vdov || !arealg)))
        return LULESH_INVALID_ARGUMENT;
    unsigned long long *slots = nullptr;
    hipError_t e = scratch_on_device(&slots);
    if (e != hipSuccess) return LULESH_HIP_ERROR - int(e);
    hipStream_t st = static_cast<hipStream_t>(stream);
    Params P{};
    P.qqc = qqc;
    P.dvovmax = dvovmax;
    k_arm_slots<<<1, kMinSlots, 0, st>>>(slots);
    if (numElem > 0)
        k_constraints<<<grid_for(numElem, kBlock), kBlock, 0, st>>>(numElem, ss, vdov, arealg, P,
                                                                    slots);
    k_fold_minima_to<<<1, kMinSlots, 0, st>>>(slots, out);
    e = hipGetLastError();
    return e == hipSuccess ? 0 : LULESH_HIP_ERROR - int(e);
}
#endif

int lulesh_hip_constraints_result(lulesh_ctx *c, double *dtcourant, double *dthydro) {
    if (!c) return LULESH_INVALID_ARGUMENT;
    int rc = read_states(c);
    if (rc) return rc;
    if (dtcourant) *dtcourant = bits_to_double(c->hst->dtc_bits);
    if (dthydro) *dthydro = bits_to_double(c->hst->dth_bits);
    return 0;
}

int lulesh_hip_run(lulesh_ctx *c, int32_t max_cycles, lulesh_time_state *io) {
    if (!c || !io || c->nprob != 1) return LULESH_INVALID_ARGUMENT;
    // the host mirror the run loop plans with: the run constants of the
    // problem and the caller's time state
    DevState &s = c->hst[0];
    s = c->init[0];
    s.time = io->time;
    s.deltatime = io->deltatime;
    s.dtc_bits = host_dbits(io->dtcourant);
    s.dth_bits = host_dbits(io->dthydro);
    s.pending = 0;
    s.cycle = io->cycle;
    s.its = max_cycles;
    s.error = 0;
    k_set_time_state<<<1, 1, 0, c->stream>>>(c->st, s.time, s.deltatime, s.dtc_bits,
                                             s.dth_bits, s.cycle, s.its);
    HIP_TRY(c, hipGetLastError());
    const int rc = run_loop(c);
    to_time_state(c->hst[0], io);
    return rc;
}

int lulesh_hip_batch_run(lulesh_ctx *c, int32_t max_cycles, lulesh_time_state *states) {
    if (!c) return LULESH_INVALID_ARGUMENT;
    if (!c->lanes.empty()) {
        // every lane on its own stream, driven by its own host thread
        const size_t L = c->lanes.size();
        std::vector<std::vector<lulesh_time_state>> st(L);
        std::vector<int> rcs(L, 0);
        for (size_t l = 0; l < L; ++l) st[l].resize(size_t(c->lanes[l]->nprob));
        std::vector<std::thread> th;
        for (size_t l = 1; l < L; ++l)
            th.emplace_back([&, l] { rcs[l] = lulesh_hip_batch_run(c->lanes[l], max_cycles,
                                                                   st[l].data()); });
        rcs[0] = lulesh_hip_batch_run(c->lanes[0], max_cycles, st[0].data());
        for (auto &t : th) t.join();
        // the lanes start together: the slowest lane's device time
        c->last_run_ms = 0;
        for (lulesh_ctx *l : c->lanes) c->last_run_ms = std::max(c->last_run_ms, l->last_run_ms);
        if (states)
            for (int p = 0; p < c->nprob; ++p)
                states[p] = st[size_t(c->lane_of[size_t(p)])][size_t(c->local_of[size_t(p)])];
        for (size_t l = 0; l < L; ++l)
            if (rcs[l]) {
                c->err = c->lanes[l]->err;
                return rcs[l];
            }
        return 0;
    }
    int rc = read_states(c);
    if (rc) return rc;
    for (int pi = 0; pi < c->nprob; ++pi) c->hst[pi].its = max_cycles;
    k_set_its<<<grid_for(c->nprob, 64), 64, 0, c->stream>>>(c->st, c->nprob, max_cycles);
    HIP_TRY(c, hipGetLastError());
    rc = run_loop(c);
    if (states)
        for (int pi = 0; pi < c->nprob; ++pi) to_time_state(c->hst[pi], &states[pi]);
    return rc;
}

double lulesh_hip_last_run_ms(const lulesh_ctx *c) { return c ? c->last_run_ms : 0.0; }

int lulesh_hip_stream_bandwidth(int64_t bytes, int32_t reps, double *GBs) {
    if (bytes < (int64_t(1) << 20) || reps < 1 || !GBs) return LULESH_INVALID_ARGUMENT;
    *GBs = 0;
    const int64_t n = bytes / 16;
    void *a = nullptr, *b = nullptr;
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = 0;
    auto fail = [&](hipError_t e) {
        if (rc == 0 && e != hipSuccess) rc = LULESH_HIP_ERROR - int(e);
        return rc != 0;
    };
    int dev = 0, cus = 256;
    if (!fail(hipGetDevice(&dev)) &&
        !fail(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev)) &&
        !fail(hipMalloc(&a, size_t(n) * 16)) && !fail(hipMalloc(&b, size_t(n) * 16)) &&
        !fail(hipMemset(a, 1, size_t(n) * 16)) &&
        !fail(hipStreamCreateWithFlags(&st, hipStreamNonBlocking)) &&
        !fail(hipEventCreate(&e0)) && !fail(hipEventCreate(&e1))) {
        const vec2d *src = static_cast<const vec2d *>(a);
        vec2d *dst = static_cast<vec2d *>(b);
        // launch shapes: one pass per lane (1, 2 or 4 vectors), and a grid of
        // 8 or 16 workgroups per CU walking the buffer with 4 vectors
        for (int shape = 0; shape < 5 && rc == 0; ++shape) {
            auto go = [&] {
                switch (shape) {
                case 0: k_stream_copy<1, false><<<grid_for(n, 256), 256, 0, st>>>(src, dst, n); break;
                case 1: k_stream_copy<2, false><<<grid_for(n, 512), 256, 0, st>>>(src, dst, n); break;
                case 2: k_stream_copy<4, false><<<grid_for(n, 1024), 256, 0, st>>>(src, dst, n); break;
                case 3: k_stream_copy<4, true><<<8 * cus, 256, 0, st>>>(src, dst, n); break;
                default: k_stream_copy<4, true><<<16 * cus, 256, 0, st>>>(src, dst, n); break;
                }
            };
            go(); // warm-up
            if (fail(hipEventRecord(e0, st))) break;
            for (int r = 0; r < reps; ++r) go();
            if (fail(hipGetLastError()) || fail(hipEventRecord(e1, st)) ||
                fail(hipEventSynchronize(e1)))
                break;
            float ms = 0;
            if (fail(hipEventElapsedTime(&ms, e0, e1))) break;
            if (ms > 0) *GBs = std::max(*GBs, 2.0 * double(n) * 16 * reps / (ms * 1e-3) / 1e9);
        }
    }
    if (st) (void)hipStreamSynchronize(st);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (st) (void)hipStreamDestroy(st);
    if (a) (void)hipFree(a);
    if (b) (void)hipFree(b);
    return rc;
}

// FP64 multiply-add issue rate: CHAINS independent fused multiply-add chains
// per lane (x <- x a + b, bounded for 0 < a < 1), so that every SIMD always
// has an independent v_fma_f64 to issue.  The final store never happens (no
// sum equals the sentinel) but keeps the chains live.
template <int CHAINS>
__global__ __launch_bounds__(256) void k_fma_rate(double *out, int iters, double a, double b) {
    double x[CHAINS];
#pragma unroll
    for (int k = 0; k < CHAINS; ++k) x[k] = double(threadIdx.x + k) * 1e-3;
#pragma unroll 4
    for (int i = 0; i < iters; ++i)
#pragma unroll
        for (int k = 0; k < CHAINS; ++k) x[k] = __builtin_fma(x[k], a, b);
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < CHAINS; ++k) s += x[k];
    if (s == -1.0) out[blockIdx.x] = s;
}

int lulesh_hip_fp64_fma_rate(int32_t reps, double *TFLOPs) {
    if (reps < 1 || !TFLOPs) return LULESH_INVALID_ARGUMENT;
    *TFLOPs = 0;
    constexpr int kChains = 8, kIters = 16384;
    double *out = nullptr;
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = 0;
    auto fail = [&](hipError_t e) {
        if (rc == 0 && e != hipSuccess) rc = LULESH_HIP_ERROR - int(e);
        return rc != 0;
    };
    int dev = 0, cus = 256;
    if (!fail(hipGetDevice(&dev)) &&
        !fail(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev))) {
        const int blocks = 16 * cus; // four waves per SIMD, four rounds
        if (!fail(hipMalloc(&out, size_t(blocks) * sizeof(double))) &&
            !fail(hipStreamCreateWithFlags(&st, hipStreamNonBlocking)) &&
            !fail(hipEventCreate(&e0)) && !fail(hipEventCreate(&e1))) {
            const double a = 0.999999, b = 1e-6;
            k_fma_rate<kChains><<<blocks, 256, 0, st>>>(out, kIters, a, b); // warm-up
            if (!fail(hipEventRecord(e0, st))) {
                for (int r = 0; r < reps; ++r)
                    k_fma_rate<kChains><<<blocks, 256, 0, st>>>(out, kIters, a, b);
                float ms = 0;
                if (!fail(hipGetLastError()) && !fail(hipEventRecord(e1, st)) &&
                    !fail(hipEventSynchronize(e1)) && !fail(hipEventElapsedTime(&ms, e0, e1)) &&
                    ms > 0)
                    *TFLOPs = 2.0 * kChains * double(kIters) * 256.0 * blocks * reps /
                              (ms * 1e-3) / 1e12;
            }
        }
    }
    if (st) (void)hipStreamSynchronize(st);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (st) (void)hipStreamDestroy(st);
    if (out) (void)hipFree(out);
    return rc;
}

int lulesh_hip_download(lulesh_ctx *c, int id, void *dst, size_t count) {
    return field_copy(c, 0, id, dst, count, true);
}

int lulesh_hip_upload(lulesh_ctx *c, int id, const void *src, size_t count) {
    return field_copy(c, 0, id, const_cast<void *>(src), count, false);
}

int lulesh_hip_batch_download(lulesh_ctx *c, int32_t prob, int id, void *dst, size_t count) {
    if (c && !c->lanes.empty()) {
        if (prob < 0 || prob >= c->nprob) return LULESH_INVALID_ARGUMENT;
        return field_copy(c->lanes[size_t(c->lane_of[size_t(prob)])],
                          c->local_of[size_t(prob)], id, dst, count, true);
    }
    return field_copy(c, prob, id, dst, count, true);
}

int lulesh_hip_download_view(lulesh_ctx *c, lulesh_domain_view *dst) {
    if (!c || !dst || c->nprob != 1) return LULESH_INVALID_ARGUMENT;
    for (int id = 0; id < LULESH_F_COUNT; ++id) {
        double *p = view_field(dst, id);
        if (!p || (is_dead_field(id) && !c->store_forces) ||
            (is_grad_field(id) && c->grads_stale))
            continue;
        int rc = lulesh_hip_download(c, id, p, is_node_field(id) ? size_t(c->totN)
                                                                 : size_t(c->totE));
        if (rc) return rc;
    }
    int rc = read_states(c);
    if (rc) return rc;
    dst->time = c->hst->time;
    dst->deltatime = c->hst->deltatime;
    dst->dtcourant = bits_to_double(c->hst->dtc_bits);
    dst->dthydro = bits_to_double(c->hst->dth_bits);
    dst->cycle = c->hst->cycle;
    return 0;
}

int lulesh_hip_set_tuning(lulesh_ctx *c, const lulesh_tuning *t) {
    if (!c || !t) return LULESH_INVALID_ARGUMENT;
    if (!c->lanes.empty()) {
        for (lulesh_ctx *l : c->lanes)
            if (int rc = lulesh_hip_set_tuning(l, t)) return rc;
        c->tune = c->lanes[0]->tune;
        return 0;
    }
    c->tune = *t;
    return configure(c);
}

int lulesh_hip_get_tuning(const lulesh_ctx *c, lulesh_tuning *t) {
    if (!c || !t) return LULESH_INVALID_ARGUMENT;
    *t = c->tune;
    return 0;
}

int lulesh_hip_autotune(lulesh_ctx *c, int32_t cycles, lulesh_tuning *best,
                        double *best_ms_per_cycle) {
    if (!c || cycles < 2) return LULESH_INVALID_ARGUMENT;
    // The contexts timed together: the context itself, or every lane of a
    // stream group (one tuning for all lanes, the lanes running at once as
    // lulesh_hip_batch_run runs them).
    std::vector<lulesh_ctx *> cs;
    if (c->lanes.empty())
        cs.push_back(c);
    else
        cs = c->lanes;
    const size_t L = cs.size();
    // Snapshot every state buffer of every context; each candidate runs
    // `cycles` cycles from the snapshot and the state is restored afterwards.
    std::vector<Snapshot> snaps(L);
    std::vector<std::vector<DevState>> st0(L);
    std::vector<char> timing0(L);
    auto free_snaps = [&]() {
        for (auto &sn : snaps)
            for (void *q : sn.copies) (void)hipFree(q);
    };
    int rc = 0;
    for (size_t l = 0; l < L && rc == 0; ++l) {
        lulesh_ctx *x = cs[l];
        Snapshot &snap = snaps[l];
        for (int id = 0; id < LULESH_F_COUNT; ++id)
            if (x->f[id] && !is_state_node_field(id))
                snap.regions.push_back(
                    {x->f[id], size_t(is_node_field(id) ? x->totN : x->totE) * 8});
        for (int k = 0; k < 6; ++k) {
            if (x->g32[k]) snap.regions.push_back({x->g32[k], size_t(x->totE) * 4});
            snap.regions.push_back({x->nb.b[0][k], size_t(x->totN) * 8});
            if (x->nb.b[1][k] != x->nb.b[0][k])
                snap.regions.push_back({x->nb.b[1][k], size_t(x->totN) * 8});
        }
        snap.regions.push_back({x->st, sizeof(DevState) * size_t(x->nprob)});
        for (auto &r : snap.regions) {
            void *q = nullptr;
            if (hipMalloc(&q, r.second) != hipSuccess) {
                free_snaps();
                return hip_fail(x, hipErrorOutOfMemory, "autotune snapshot");
            }
            snap.copies.push_back(q);
            (void)hipMemcpyAsync(q, r.first, r.second, hipMemcpyDeviceToDevice, x->stream);
        }
        timing0[l] = x->timing;
        x->timing = false;
        rc = read_states(x);
        st0[l].assign(x->hst, x->hst + x->nprob);
    }
    if (rc) {
        for (size_t l = 0; l < L; ++l) cs[l]->timing = timing0[l] != 0;
        free_snaps();
        return rc;
    }
    auto restore = [&](size_t l) -> int {
        lulesh_ctx *x = cs[l];
        for (size_t k = 0; k < snaps[l].regions.size(); ++k)
            HIP_TRY(x, hipMemcpyAsync(snaps[l].regions[k].first, snaps[l].copies[k],
                                      snaps[l].regions[k].second, hipMemcpyDeviceToDevice,
                                      x->stream));
        return read_states(x);
    };
    // restore the snapshot and arm `cycles` cycles
    auto rearm = [&](size_t l) -> int {
        lulesh_ctx *x = cs[l];
        int r = restore(l);
        if (r) return r;
        for (int pi = 0; pi < x->nprob; ++pi) x->hst[pi].its = st0[l][size_t(pi)].cycle + cycles;
        return write_states(x);
    };
    // run_loop on every context at once (lane 0 on this thread)
    auto run_all = [&]() -> int {
        std::vector<int> rcs(L, 0);
        std::vector<std::thread> th;
        for (size_t l = 1; l < L; ++l) th.emplace_back([&, l] { rcs[l] = run_loop(cs[l]); });
        rcs[0] = run_loop(cs[0]);
        for (auto &t : th) t.join();
        for (int r : rcs)
            if (r) return r;
        return 0;
    };
    auto trial = [&](const lulesh_tuning &t, double *ms) -> int {
        int r = 0;
        for (size_t l = 0; l < L && r == 0; ++l) {
            cs[l]->tune = t;
            if (!(r = configure(cs[l])) && !(r = rearm(l)))
                r = build_graphs(cs[l]); // graph build outside the timing
        }
        // a first run warms the new graphs and launch tables up; the second,
        // from the same snapshot, is the one timed (device time of the
        // slowest context)
        if (r || (r = run_all())) return r;
        for (size_t l = 0; l < L && r == 0; ++l) r = rearm(l);
        if (r || (r = run_all())) return r;
        double slowest = 0;
        for (lulesh_ctx *x : cs) slowest = std::max(slowest, x->last_run_ms);
        *ms = slowest / cycles;
        return 0;
    };
    lulesh_tuning cur = cs[0]->tune;
    cur.zplanes = auto_zplanes(cs[0]);
    double cur_ms = 1e300;
    rc = trial(cur, &cur_ms);
    if (rc == 0) cur = cs[0]->tune;
    // coordinate descent over the launch parameters; the plane counts are
    // the smallest ones giving 1..40 workgroups along z for the largest
    // problem (each of them splits its node planes as evenly as a count can)
    int nmax = 1;
    for (lulesh_ctx *x : cs)
        for (const auto &p : x->probs) nmax = std::max(nmax, p.s + 1);
    std::vector<int> zcand;
    for (int g = 40; g >= 1; --g) {
        const int z = (nmax + 2 * g - 2) / g; // the smallest z with at most g plane groups
        if (z >= 2 && (zcand.empty() || zcand.back() != z)) zcand.push_back(z);
    }
    const int bcand[] = {64, 128, 256};
    const int gcand[] = {4, 8, 16, 32};
    const int mcand[] = {0, 1, 2};
    // element planes per workgroup of the fused pass: 2..40 groups along z
    std::vector<int> ecand;
    for (int g = 40; g >= 2; --g) {
        const int z = (nmax - 1 + g - 1) / g;
        if (z >= 2 && (ecand.empty() || ecand.back() != z)) ecand.push_back(z);
    }
    for (int pass = 0; pass < 6 && rc == 0; ++pass) {
        const int *cand = pass == 0   ? zcand.data()
                          : pass == 3 ? gcand
                          : pass == 4 ? mcand
                          : pass == 5 ? ecand.data()
                                      : bcand;
        const int ncand = pass == 0   ? int(zcand.size())
                          : pass == 3 ? 4
                          : pass == 4 ? 3
                          : pass == 5 ? int(ecand.size())
                                      : 3;
        if (pass == 5 && cur.material_mode != 2) break;
        for (int k = 0; k < ncand && rc == 0; ++k) {
            lulesh_tuning t = cur;
            int32_t *field = pass == 0   ? &t.zplanes
                             : pass == 1 ? &t.block_elem
                             : pass == 2 ? &t.block_material
                             : pass == 3 ? &t.cycles_per_graph
                             : pass == 4 ? &t.material_mode
                                         : &t.em_planes;
            if (*field == cand[k]) continue;
            *field = cand[k];
            double ms = 0;
            rc = trial(t, &ms);
            if (rc == 0 && ms < cur_ms) {
                cur = cs[0]->tune;
                cur_ms = ms;
            }
        }
    }
    int rc2 = 0, rc3 = 0;
    for (size_t l = 0; l < L; ++l) {
        cs[l]->tune = cur;
        if (int r = configure(cs[l])) rc2 = rc2 ? rc2 : r;
        if (int r = restore(l)) rc3 = rc3 ? rc3 : r;
        cs[l]->timing = timing0[l] != 0;
    }
    free_snaps();
    c->tune = cur;
    if (rc) return rc;
    if (rc2) return rc2;
    if (rc3) return rc3;
    if (best) *best = cur;
    if (best_ms_per_cycle) *best_ms_per_cycle = cur_ms;
    return 0;
}

int lulesh_hip_set_timing(lulesh_ctx *c, int enable) {
    if (!c) return LULESH_INVALID_ARGUMENT;
    if (!c->lanes.empty()) {
        for (lulesh_ctx *l : c->lanes)
            if (int rc = lulesh_hip_set_timing(l, enable)) return rc;
        return 0;
    }
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    c->tev_used = 0;
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
    if (!c->lanes.empty()) {
        // launches summed over the lanes, averages weighted by launches
        std::vector<double> tot(size_t(std::max(cap, 0)), 0.0);
        std::vector<int64_t> cnt(size_t(std::max(cap, 0)), 0);
        int n = 0;
        for (lulesh_ctx *l : c->lanes) {
            std::vector<double> a(tot.size());
            std::vector<int64_t> k(tot.size());
            n = lulesh_hip_kernel_times(l, names, a.data(), k.data(), cap);
            if (n < 0) return n;
            for (int i = 0; i < n && i < cap; ++i) {
                tot[size_t(i)] += a[size_t(i)] * double(k[size_t(i)]);
                cnt[size_t(i)] += k[size_t(i)];
            }
        }
        for (int i = 0; i < n && i < cap; ++i) {
            if (avg_ms) avg_ms[i] = cnt[size_t(i)] ? tot[size_t(i)] / double(cnt[size_t(i)]) : 0.0;
            if (launches) launches[i] = cnt[size_t(i)];
        }
        return n;
    }
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    if (int rc = flush_timers(c)) return rc;
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

int32_t lulesh_hip_is_structured(const lulesh_ctx *c) { return c && c->structured ? 1 : 0; }

const char *lulesh_hip_last_error(const lulesh_ctx *c) { return c ? c->err.c_str() : "null context"; }

void lulesh_hip_destroy(lulesh_ctx *c) {
    if (!c) return;
    for (lulesh_ctx *l : c->lanes) lulesh_hip_destroy(l);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    destroy_graphs(c);
    for (void *p : c->allocs) (void)hipFree(p);
    for (void *p : {(void *)c->d_pd, (void *)c->d_segs, (void *)c->d_zBlockProb, (void *)c->d_zMap,

                    (void *)c->d_eBlockProb, (void *)c->d_mBlockSeg, (void *)c->d_oBlockProb,
                    (void *)c->d_emBlockProb})
        if (p) (void)hipFree(p);
    if (c->hst) (void)hipHostFree(c->hst);
    for (hipEvent_t e : {c->run0, c->run1})
        if (e) (void)hipEventDestroy(e);
    for (hipEvent_t e : c->tev) (void)hipEventDestroy(e);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

} // namespace LULESH_IMPL_NS
