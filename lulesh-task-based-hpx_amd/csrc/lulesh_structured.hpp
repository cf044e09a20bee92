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
// node columns over a range of node planes and marches upward in z: for
// each element plane k its threads (one per element) compute the stress and
// hourglass corner forces of the pw x ph element patch covering the tile
// (16 x 16 unless another shape needs fewer tiles: zmarch_shape), stage them
// in LDS, and the owning threads add them to their nodes.  Along x the first
// tile owns node columns 0..pw-1 of patch element columns 0..pw-1; tile
// t >= 1 owns the pw-1 node columns (pw-1)t+1 .. (pw-1)t+pw-1 and its patch
// starts one element lower ((pw-1)t), the halo its first node column needs;
// likewise along y with ph.  The node forces are summed in the reference's
// corner order
// (ascending element index, stress then hourglass per corner,
// lulesh.cc:849-857):
//   node plane k+1 starts with corners 6,7,5,4 of element plane k,
//   node plane k   ends   with corners 2,3,1,0 of element plane k,
// so the partial sums of a node plane live in registers for one step.  A
// finished node plane is advanced (acceleration, symmetry BCs, velocity,
// position) and written to the other half of a double buffer, since the
// neighbouring tiles still read the old positions.
#pragma once

#include <algorithm>

#include "lulesh_elem.hpp"

namespace lulesh_dev {

constexpr int kZmarchThreads = 256;  // threads per z-march workgroup
constexpr int kPatchNodes = 17 * 17; // LDS node slots per plane: (pw + 1) (ph + 1) at most

// Tiles along an edge of n nodes for patches of P elements: P node columns
// for the first tile, P - 1 for each further one.
__host__ __device__ inline int zmarch_tiles(int n, int P) {
    return n <= P ? 1 : 1 + (n - P + P - 2) / (P - 1);
}

// Plane groups of the z-march (z node planes per workgroup): group 0 owns
// node planes [0, z), every later group z - 1 planes, because it also
// recomputes the element plane below its first node plane; so every
// workgroup walks at most z element planes.  z = 1: one plane per group.
__host__ __device__ inline int zmarch_group_start(int g, int z) {
    return z <= 1 ? g : (g == 0 ? 0 : z + (g - 1) * (z - 1));
}
__host__ __device__ inline int zmarch_groups(int n, int z) {
    if (z <= 1) return n;
    return n <= z ? 1 : 1 + (n - z + z - 2) / (z - 1);
}
// Element planes the longest plane group of an edge of n nodes walks.
inline int zmarch_steps(int n, int z) {
    int m = 0;
    for (int g = 0, G = zmarch_groups(n, z); g < G; ++g) {
        const int K0 = zmarch_group_start(g, z), K1 = std::min(zmarch_group_start(g + 1, z), n);
        m = std::max(m, std::min(K1, n - 1) - std::max(K0 - 1, 0));
    }
    return m;
}

// Patch shape (pw x ph elements) of one problem.  16 x 16 unless another
// shape that fits the workgroup (pw ph <= 256 threads) and the LDS node
// ring ((pw + 1)(ph + 1) <= 289) needs at least 3% fewer tiles per plane
// group; among those the fewest tiles, then more threads, then the wider
// patch (longer coalesced rows).  s = 150 keeps 16 x 16 (10 x 10 tiles);
// small meshes gain most (s = 20: two 23 x 11 tiles instead of four).
inline void zmarch_shape(int n, int &pw, int &ph) {
    const int t16 = zmarch_tiles(n, 16) * zmarch_tiles(n, 16);
    int best = 0, bestThreads = 0, bw = 16, bh = 16;
    for (int w = 8; w <= 32; ++w)
        for (int h = 8; h <= 32; ++h) {
            if (w * h > kZmarchThreads || (w + 1) * (h + 1) > kPatchNodes) continue;
            const int t = zmarch_tiles(n, w) * zmarch_tiles(n, h);
            if (best == 0 || t < best || (t == best && w * h > bestThreads) ||
                (t == best && w * h == bestThreads && w > bw)) {
                best = t;
                bestThreads = w * h;
                bw = w;
                bh = h;
            }
        }
    const bool take = 100 * best <= 97 * t16;
    pw = take ? bw : 16;
    ph = take ? bh : 16;
}

// Tile of the fused elements + material pass (k_elemmat_zmarch): pw x ph
// element columns, one thread each, with pw ph <= 256 and pw + ph <= 32 (the
// 2 (pw + ph) elements bordering the tile are recomputed by one wave).  The
// fewest tiles over an s x s plane, then the most threads, then the wider
// tile.  s = 150: 17 x 15 (9 x 10 tiles instead of 10 x 10 at 16 x 16).
inline void elemmat_shape(int s, int &pw, int &ph) {
    int best = 0, bestThreads = 0;
    pw = ph = 16;
    for (int w = 4; w <= 28; ++w)
        for (int h = 4; h <= 28; ++h) {
            if (w * h > 256 || w + h > 32) continue;
            const int t = ((s + w - 1) / w) * ((s + h - 1) / h);
            if (best == 0 || t < best || (t == best && w * h > bestThreads) ||
                (t == best && w * h == bestThreads && w > pw)) {
                best = t;
                bestThreads = w * h;
                pw = w;
                ph = h;
            }
        }
}

// Element i of an array whose base is the same for the whole workgroup, as
// a 32-bit byte offset from that base: the access then keeps the base in
// scalar registers and only the offset in a vector register (the saddr form
// of global_load / global_store) instead of a 64-bit address per access.
// The arrays of one problem stay below 4 GiB (at most 646^3 values of 8
// bytes).  In the nodal z-march: 1349 -> 1277 VALU instructions per element
// step with the power-of-two-free stress integration, 0.174 -> 0.171 ms.
template <typename T>
__device__ __forceinline__ T &at32(T *base, int i) {
    return *reinterpret_cast<T *>(reinterpret_cast<char *>(base) +
                                  uint32_t(i) * uint32_t(sizeof(T)));
}
template <typename T>
__device__ __forceinline__ const T &at32(const T *base, int i) {
    return *reinterpret_cast<const T *>(reinterpret_cast<const char *>(base) +
                                        uint32_t(i) * uint32_t(sizeof(T)));
}

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
    int32_t zplanes;          // z-march: node planes of the first plane group (zmarch_groups)
    int32_t pw, ph;           // z-march patch: elements per row, rows
    int32_t tilesX, tilesY;   // z-march tiles along x and y
    int32_t zBlock0, eBlock0; // first workgroup of this problem in the z-march / element grids
    int32_t mBlock0;          // first workgroup in the element-order material grid
    // fused elements + material pass (k_elemmat_zmarch): tile of emPW x emPH
    // element columns, emTX x emTY tiles, emPlanes element planes per
    // workgroup, first workgroup emBlock0
    int32_t emPW, emPH, emTX, emTY, emPlanes, emBlock0;
    // 1/s and 1/s^2 rounded to double: element (k, j, i) from its index li
    // as floor((li + 0.5) * (1/d)), exact because the quotient stays at
    // least 0.5/s^2 (> 2^-20 for s <= 645) away from an integer while the
    // two roundings move it by less than 2^-40
    double inv_s, inv_ss;
};

// Node-plane ring in LDS: two slots, each holding x, y, z, xd, yd, zd of the
// (pw+1) x (ph+1) patch nodes of one node plane (field f of patch node t at
// slot[f * kPatchNodes + t]).
//
// Forms of the kernel:
//  MODE_EXACT (FP64, every multiply and add rounded on its own): each
//    corner's stress and hourglass force are staged separately and the node
//    adds them one after the other, corner by corner, in the reference's
//    interleaved order (lulesh.cc:849-857): bitwise the reference
//    arithmetic.  24 staged values per element (77 KB of LDS); 212 VGPRs
//    with the two-pass hourglass and the carried bottom face: two
//    workgroups per CU.
//  MODE_FUSED (FP64, multiply-adds fused): the element adds its stress and
//    hourglass force per corner before staging (one more rounding per
//    corner, no change in what is summed), 12 staged values per element,
//    and forms the hourglass force in two register-lean passes: 228 VGPRs
//    without spills.  Two workgroups per CU.  At s=150 the nodal kernel took
//    0.185 ms against 0.211 ms staged as MODE_EXACT (0.168 ms with the
//    bottom-face coordinates carried in registers, CARRY; now 0.152 ms, 226
//    VGPRs, with fewer operations where the results need not be the
//    reference's bit for bit: the hourglass force factored, the stress
//    forces from the face diagonals, one transform for the Jacobian
//    columns and the hourglass modes, a reciprocal cube root for the
//    hourglass coefficient and volinv, a Newton reciprocal for 1/mass;
//    855 VALU instructions per element step against 982).
//  MODE_MIXED (FP32 hourglass): staged as MODE_FUSED; with its lower
//    corners' sums in registers (LOREG) and the next plane fetched before
//    the element work (EARLY), 252 VGPRs, 0.149 ms at s=150.
//  DBL (the fused and mixed forms): the upper corners get staging rows of
//    their own (77 KB of LDS with the ring), so both corner sets are staged
//    before one barrier, and each thread reads its node's old state from the
//    ring during the element work: two barriers per plane instead of four.
//    Two workgroups per CU (the VGPRs of the fused form allow no more; the
//    mixed form had three at 52 KB).  s=150: fused 0.183 -> 0.180 ms, mixed
//    0.177 -> 0.169 ms.
// Tried and dropped: FP64 staged combined at three workgroups per CU (needs
// more than the 168 VGPRs of three waves per SIMD and spills: 0.48-0.59
// ms); the next plane copied into LDS by global_load_lds during the element
// work (the copies' waits and spills cost more than the latency they hide;
// fetching the next plane into registers before the element work instead of
// after it gains nothing in the fused form, which the mixed form does).
// SQUARE16: every problem of the launch has 16 x 16 patches (the shape is
// then a compile-time constant; runtime shapes cost 2-6% at s=150).
enum ZmarchMode { MODE_EXACT = 0, MODE_FUSED = 1, MODE_MIXED = 2 };

template <int MODE, bool STORE_F, bool SQUARE16, bool DBL>
__global__ __launch_bounds__(kZmarchThreads)
__attribute__((amdgpu_waves_per_eu(2))) void k_nodal_zmarch(
    const ProbDesc *__restrict__ pds, const int32_t *__restrict__ blockMap,
    const int32_t *__restrict__ blockProb, NodeBufs nb, NodeForceOut fo0, const double *__restrict__ p0,
    const double *__restrict__ q0, const double *__restrict__ volo0,
    const double *__restrict__ v0, const double *__restrict__ ss0,
    const double *__restrict__ elemMass0, const double *__restrict__ nodalMass0, double hgcoef,
    double u_cut, DevState *states) {
    constexpr bool COMBINE = MODE != MODE_EXACT; // corner sums staged (12 values)
    // the logical block of this workgroup (its problem, tile and plane
    // group), or -1 for an idle padding block: the host's map orders the
    // blocks per XCD
    const int blk = blockMap[blockIdx.x];
    if (blk < 0) return;
    const int prob = blockProb[blk];
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
    const int lb = blk - pd.zBlock0;
    // LDS: the staged corner values of the element plane (4 corners x 3
    // stress + 3 hourglass values, or 3 sums), and the node-plane ring.
    constexpr int kStage = COMBINE ? 12 : 24; // staged values per element
    // DBL: the upper corners get their own staging rows (lds[kStage..]), so
    // both corner sets are staged before one barrier and read after it
    static_assert(!DBL || COMBINE, "two staging areas only for the combined corner sums");
    // LOREG: the lower corners' sums are kept in registers too and staged with
    // the upper ones (the mixed form: 160 against 168 us at s=150; the fused
    // form's VGPRs would reach 252 and it ran 3% slower)
    constexpr bool LOREG = DBL && MODE == MODE_MIXED;
    __shared__ double lds[DBL ? 2 * kStage : kStage][kZmarchThreads];
    __shared__ double ring[2][6 * kPatchNodes];
    auto slot_of = [](int K) { return K & 1; };
    auto pnv = [&](int slot, int f, int t) -> double & { return ring[slot][f * kPatchNodes + t]; };

    const int n = s + 1;
    const int tid = threadIdx.x;
    const int pw = SQUARE16 ? 16 : pd.pw, ph = SQUARE16 ? 16 : pd.ph; // patch: pw x ph elements
    const int jj = tid / pw, ii = tid - jj * pw; // (threads past pw * ph idle: jj >= ph)
    const int tilesX = pd.tilesX, tilesXY = pd.tilesX * pd.tilesY;
    const int tile = lb % tilesXY;
    const int ty = tile / tilesX, tx = tile % tilesX;
    const int dy = ty > 0, dx = tx > 0;                 // 1: the patch has a low-side halo
    const int E0y = ty * (ph - 1), E0x = tx * (pw - 1); // first element row / column
    const int zg = lb / tilesXY; // plane group
    const int K0 = zmarch_group_start(zg, planes_per_block);
    const int K1 = min(zmarch_group_start(zg + 1, planes_per_block), n); // owned planes [K0, K1)
    const int ej = E0y + jj, ei = E0x + ii;        // this thread's element column
    const bool elemCol = (jj < ph) && (ej < s) && (ei < s);
    const int J = E0y + dy + jj, I = E0x + dx + ii; // this thread's node column
    const bool ownsNode = (jj < ph - dy) && (ii < pw - dx) && (J < n) && (I < n);
    // validity of the four elements around node (J,I) in a plane:
    // (J-1,I-1) (J-1,I) (J,I-1) (J,I)  <->  patch threads tb, tb+1, tb+pw, tb+pw+1
    // (tb < 0 only where J = 0 or I = 0, whose elements are not read)
    const bool v00 = (J >= 1) && (I >= 1), v01 = (J >= 1) && (I < s), v10 = (J < s) && (I >= 1),
               v11 = (J < s) && (I < s);
    const int tb = tid + (dy - 1) * pw + (dx - 1);
    const double dt = st->deltatime;

    double px = 0.0, py = 0.0, pz = 0.0; // partial sums of the node in the current plane

    const int own = (jj + dy) * (pw + 1) + (ii + dx); // node (J, I) in the patch node array
    // the node's start-of-cycle state (x, y, z, xd, yd, zd) from the ring;
    // DBL reads it during the element work, before the barrier after which
    // the ring slot may be refilled, the other forms when they use it
    auto own_state = [&](int K, double ox[6]) {
#pragma unroll
        for (int f = 0; f < 6; ++f) ox[f] = pnv(slot_of(K), f, own);
    };
    // old(f): field f of the node's start-of-cycle state
    auto finish_node = [&](int K, double mass, auto old) {
        const int g = (K * n + J) * n + I;
        double ax, ay, az;
        if constexpr (MODE == MODE_FUSED) {
            // a reciprocal and three products (not bitwise the quotients)
            const double rm = recip_pos(mass);
            ax = px * rm;
            ay = py * rm;
            az = pz * rm;
        } else {
            ax = px / mass;
            ay = py / mass;
            az = pz / mass;
        }
        if (STORE_F) {
            at32(fo.fx, g) = px;
            at32(fo.fy, g) = py;
            at32(fo.fz, g) = pz;
        }
        if (I == 0) ax = 0.0; // symmX: the x = 0 plane
        if (J == 0) ay = 0.0; // symmY
        if (K == 0) az = 0.0; // symmZ
        if (STORE_F) {
            at32(fo.xdd, g) = ax;
            at32(fo.ydd, g) = ay;
            at32(fo.zdd, g) = az;
        }
        double vx = old(3) + ax * dt;
        if (fabs_(vx) < u_cut) vx = 0.0;
        double vy = old(4) + ay * dt;
        if (fabs_(vy) < u_cut) vy = 0.0;
        double vz = old(5) + az * dt;
        if (fabs_(vz) < u_cut) vz = 0.0;
        at32(out.xd, g) = vx;
        at32(out.yd, g) = vy;
        at32(out.zd, g) = vz;
        at32(out.x, g) = old(0) + vx * dt;
        at32(out.y, g) = old(1) + vy * dt;
        at32(out.z, g) = old(2) + vz * dt;
    };
    // Ring fill: node plane K goes into its slot in two steps: fetch_plane
    // issues the global loads into registers (coalesced rows of pw+1 nodes;
    // nodes outside the mesh are not touched -- no element reads them) and
    // store_plane writes them to LDS one phase later, so the load latency is
    // hidden behind the node work of the current plane.
    // Patch node t = tid + 256 m, m = 0, 1.
    int pg[2];
    bool pin[2];
#pragma unroll
    for (int m = 0; m < 2; ++m) {
        const int t = tid + m * kZmarchThreads;
        const int a = t / (pw + 1), b = t - a * (pw + 1);
        const int Jn = E0y + a, In = E0x + b;
        pin[m] = t < (pw + 1) * (ph + 1) && Jn < n && In < n;
        pg[m] = Jn * n + In;
    }
    double rn[2][6];
    // (every path assigns the registers, so their old values die at the
    // store and do not stay live through the element work)
    auto fetch_plane = [&](int K, bool go) {
#pragma unroll
        for (int m = 0; m < 2; ++m) {
            if (go && pin[m]) {
                const int g = K * n * n + pg[m];
                rn[m][0] = at32(in.x, g);
                rn[m][1] = at32(in.y, g);
                rn[m][2] = at32(in.z, g);
                rn[m][3] = at32(in.xd, g);
                rn[m][4] = at32(in.yd, g);
                rn[m][5] = at32(in.zd, g);
            } else {
#pragma unroll
                for (int f = 0; f < 6; ++f) rn[m][f] = 0.0;
            }
        }
    };
    auto store_plane = [&](int K) {
        const int slot = slot_of(K);
#pragma unroll
        for (int m = 0; m < 2; ++m)
            if (pin[m]) {
                const int t = tid + m * kZmarchThreads;
#pragma unroll
                for (int f = 0; f < 6; ++f) pnv(slot, f, t) = rn[m][f];
            }
    };
    // The element scalars needed first are fetched early as well:
    // p, q, volo, v of element (k, ej, ei).
    double ev[4];
    auto fetch_elem = [&](int k, bool go) {
        if (go && elemCol) {
            const int e = (k * s + ej) * s + ei;
            ev[0] = at32(p, e);
            ev[1] = at32(q, e);
            ev[2] = at32(volo, e);
            ev[3] = at32(v, e);
        } else {
#pragma unroll
            for (int f = 0; f < 4; ++f) ev[f] = 0.0;
        }
    };
    // add the corner staged at slot c of patch thread t: stress then
    // hourglass, or their sum (COMBINE)
    auto add = [&](int c, int t, int base = 0) { // base: kStage for the upper rows (DBL)
        if constexpr (COMBINE) {
            px += lds[base + 3 * c + 0][t];
            py += lds[base + 3 * c + 1][t];
            pz += lds[base + 3 * c + 2][t];
        } else {
            px += lds[6 * c + 0][t];
            px += lds[6 * c + 3][t];
            py += lds[6 * c + 1][t];
            py += lds[6 * c + 4][t];
            pz += lds[6 * c + 2][t];
            pz += lds[6 * c + 5][t];
        }
    };

    const int kbeg = max(K0 - 1, 0);
    const int kend = min(K1, s); // element planes [kbeg, kend)
    // nodal mass of this thread's node in the plane finished by the current
    // step (fetched one step ahead, like the node planes)
    auto fetch_mass = [&](int K) {
        return ownsNode && K <= s ? at32(nodalMass, (K * n + J) * n + I) : 0.0;
    };
    fetch_plane(kbeg, true);
    store_plane(kbeg);
    fetch_plane(kbeg + 1, kbeg < kend);
    fetch_elem(kbeg, kbeg < kend);
    double mcur = fetch_mass(kbeg), mnext = 0.0;
    double up[kStage]; // corners 4..7: (sx, sy, sz, hx, hy, hz) each, or their sums
    // CARRY (fused and separately rounded forms): the coordinates of the
    // element's bottom face are the previous step's top face, kept in
    // registers (read from the ring once before the first step, after one
    // more barrier): 12 fewer LDS reads per step at the start of the element
    // work, where it waits on them (s=150: nodal 0.170 -> 0.168 ms fused,
    // 0.221 -> 0.218 ms separately rounded).  The mixed form has no
    // registers left for it; read inside the loop only for steps after the
    // first, the compiler peeled the first step and the kernel was slower.
    constexpr bool CARRY = MODE != MODE_MIXED;
    double cxyz[3][4];
    if constexpr (CARRY) {
        __syncthreads(); // plane kbeg is in its ring slot
        const int q0 = jj * (pw + 1) + ii;
        const int nq[4] = {q0, q0 + 1, q0 + pw + 2, q0 + pw + 1};
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int c = 0; c < 3; ++c) cxyz[c][a] = elemCol ? pnv(slot_of(kbeg), c, nq[a]) : 0.0;
    }
    // ---- element (k, ej, ei): stress + hourglass corner forces ----------
    // (lower corners to LDS or lo, upper corners to up; DBL: both staged)
    auto element_work = [&](int k) {
        double lo[kStage]; // (LOREG) corners 0..3 likewise, staged with the upper ones
        // the running sum of lower-corner value c: in registers (LOREG) or in LDS
        auto lower = [&](int c) -> double & {
            if constexpr (LOREG)
                return lo[c];
            else
                return lds[c][tid];
        };
        // Wave priorities (s_setprio) for the waves sharing each SIMD: the
        // element work issues ahead of the node phase of the other
        // workgroup (its LDS-bound corner sums and node updates fill
        // the issue slots the FP64 chains leave), the hourglass part
        // highest.  Measured at s=150: nodal 0.301 -> 0.261 ms (separate
        // rounding), 0.243 -> 0.219 ms (fused); no priority order tried was
        // faster.
        __builtin_amdgcn_s_setprio(2);
        if (elemCol) {
            // sigxx = sigyy = sigzz = -p - q (lulesh.cc:629-633) and the
            // hourglass determinant volo*v (lulesh.cc:1098)
            const int e = (k * s + ej) * s + ei;
            // (times 1/16 for the scaled node normals, so that the products
            // are the reference's bit for bit; DIAG, the multiply-add form:
            // the stress forces from the face diagonals, nodal 0.162 ->
            // 0.158 ms at s=150; the mixed form ran 6% slower with them)
            constexpr bool DIAG = MODE == MODE_FUSED;
            const double sig16 = (-ev[0] - ev[1]) * 0.0625;
            const double determ = ev[2] * ev[3];
            const double sse = at32(ss, e);
            const double emass = at32(elemMass, e);
            // the element's nodes in the patch node arrays (reference local
            // order, lulesh-init.cc:237-244): bottom plane k, top plane k+1
            const int q0 = jj * (pw + 1) + ii;
            const int nq[4] = {q0, q0 + 1, q0 + pw + 2, q0 + pw + 1};
            const int sb = slot_of(k), st1 = slot_of(k + 1);
            auto node = [&](int f, int a) { return pnv(a < 4 ? sb : st1, f, nq[a & 3]); };
            double x[8], y[8], z[8];
#pragma unroll
            for (int a = 0; a < 8; ++a) {
                if (CARRY && a < 4) { // the previous step's top face
                    x[a] = cxyz[0][a];
                    y[a] = cxyz[1][a];
                    z[a] = cxyz[2][a];
                } else {
                    x[a] = node(0, a);
                    y[a] = node(1, a);
                    z[a] = node(2, a);
                }
            }
            if constexpr (CARRY) {
#pragma unroll
                for (int a = 0; a < 4; ++a) {
                    cxyz[0][a] = x[a + 4];
                    cxyz[1][a] = y[a + 4];
                    cxyz[2][a] = z[a + 4];
                }
            }
            // (DIAG) the coordinates' hourglass modes, from the transform
            // that also gives the Jacobian columns (with the stress forces'
            // shared face pairs and the reciprocal of the nodal mass: 897
            // -> 855 instructions, nodal 0.155 -> 0.152 ms)
            double hmc[3][4];
            {
                double sc[3][8]; // node normals times 16, or (DIAG) the corner stress forces
                if constexpr (DIAG) {
                    double jx[3], jy[3], jz[3];
                    hex_transform(x, jx[0], jx[1], jx[2], hmc[0]);
                    hex_transform(y, jy[0], jy[1], jy[2], hmc[1]);
                    hex_transform(z, jz[0], jz[1], jz[2], hmc[2]);
                    if (jacobian_det_cols(jx[0], jx[1], jx[2], jy[0], jy[1], jy[2], jz[0], jz[1],
                                          jz[2]) == 0)
                        st->error = kVolumeError;
                } else {
                    if (jacobian_det_x64(x, y, z) == 0) st->error = kVolumeError;
                }
                if constexpr (DIAG)
                    stress_forces_x8(sc, (ev[0] + ev[1]) * 0.125, x, y, z);
                else
                    node_normals_x16(sc, x, y, z);
                // corners 0..3 go to LDS now (its last readers passed the
                // previous barrier); corners 4..7 wait in registers
                constexpr int W = COMBINE ? 3 : 6; // staged values per corner
#pragma unroll
                for (int a = 0; a < 4; ++a)
#pragma unroll
                    for (int d = 0; d < 3; ++d) {
                        const double lv = DIAG ? sc[d][a] : -(sig16 * sc[d][a]);
                        const double uv = DIAG ? sc[d][a + 4] : -(sig16 * sc[d][a + 4]);
                        if constexpr (LOREG)
                            lo[W * a + d] = lv;
                        else
                            lds[W * a + d][tid] = lv;
                        up[W * a + d] = uv;
                    }
            }
            if (determ <= 0.0) st->error = kVolumeError;
            __builtin_amdgcn_s_setprio(3);
            if constexpr (MODE == MODE_MIXED) {
                // FP32 hourglass force added to each corner's stress in place
                // (lower corners in LDS, upper in registers), in two
                // register-lean passes, on coordinates relative to node 0
                // (their differences keep their significant bits)
                float dvdx[8], dvdy[8], dvdz[8], hmx[4], hmy[4], hmz[4];
                {
                    float xf[8], yf[8], zf[8];
#pragma unroll
                    for (int a = 0; a < 8; ++a) {
                        xf[a] = float(x[a] - x[0]);
                        yf[a] = float(y[a] - y[0]);
                        zf[a] = float(z[a] - z[0]);
                    }
                    volume_derivative<float>(dvdx, dvdy, dvdz, xf, yf, zf);
                    hourglass_modes<float>(xf, yf, zf, hmx, hmy, hmz);
                }
                const float df = float(determ);
                const float coefficient =
                    float(-hgcoef * 0.01) * float(sse) * float(emass) / cbrt_(df);
                fb_hourglass_two_pass<float>(
                    dvdx, dvdy, dvdz, hmx, hmy, hmz, 1.0f / df, coefficient,
                    [&](int c, int a) { return float(node(3 + c, a)); },
                    [&](int c, int a, float f) {
                        if (a < 4)
                            lower(3 * a + c) += double(f);
                        else
                            up[3 * (a - 4) + c] += double(f);
                    });
            } else if constexpr (MODE == MODE_FUSED) {
                // FP64 hourglass force added to each corner's stress in place
                // (velocities read from LDS when used), with hourgam factored
                // out and the factor 1/12 of the volume derivatives moved
                // into volinv (982 -> 910 VALU instructions per element step,
                // nodal 0.168 -> 0.161 ms at s=150)
                double dvdx[8], dvdy[8], dvdz[8];
                volume_derivative<double, true>(dvdx, dvdy, dvdz, x, y, z);
                // determ^(-1/3) without a division: the coefficient's divisor
                // and, cubed, volinv (with the diagonal node normals: 897
                // instructions, nodal 0.162 -> 0.155 ms)
                const double rc = rcbrt_pos(determ);
                const double coefficient = -hgcoef * 0.01 * sse * emass * rc;
                fb_hourglass_factored<double>(
                    dvdx, dvdy, dvdz, hmc[0], hmc[1], hmc[2], (rc * rc) * (rc * (1.0 / 12.0)),
                    coefficient,
                    [&](int c, int a) { return node(3 + c, a); },
                    [&](int c, int a, double f) {
                        if (a < 4)
                            lower(3 * a + c) += f;
                        else
                            up[3 * (a - 4) + c] += f;
                    });
            } else {
                // FP64 hourglass force of each corner staged beside its
                // stress, in the two-pass form: the same products summed in
                // the same order as hourglass_force (bit for bit), with fewer
                // values live (214 against 256 VGPRs and spills: s=150 nodal
                // 0.233 -> 0.221 ms)
                double dvdx[8], dvdy[8], dvdz[8], hmx[4], hmy[4], hmz[4];
                volume_derivative<double>(dvdx, dvdy, dvdz, x, y, z);
                hourglass_modes<double>(x, y, z, hmx, hmy, hmz);
                const double coefficient = -hgcoef * 0.01 * sse * emass / cbrt_(determ);
                fb_hourglass_two_pass<double>(
                    dvdx, dvdy, dvdz, hmx, hmy, hmz, 1.0 / determ, coefficient,
                    [&](int c, int a) { return node(3 + c, a); },
                    [&](int c, int a, double f) {
                        if (a < 4)
                            lds[6 * a + 3 + c][tid] = f;
                        else
                            up[6 * (a - 4) + 3 + c] = f;
                    });
            }
        } else {
#pragma unroll
            for (int c = 0; c < kStage; ++c) up[c] = 0.0;
        }
        if constexpr (DBL) { // the upper corners (LOREG: both sets) to their rows now
            if (elemCol) {
#pragma unroll
                for (int c = 0; c < kStage; ++c) {
                    if constexpr (LOREG) lds[c][tid] = lo[c];
                    lds[kStage + c][tid] = up[c];
                }
            }
        }
        __builtin_amdgcn_sched_barrier(0); // keep the prefetches out of the element work
    };
    // ---- node plane k ends with corners 2,3,1,0 ----------------------------
    auto node_finish = [&](int k, const double (&ox)[6]) {
        if (ownsNode && k >= K0) {
            if (v00) add(2, tb);
            if (v01) add(3, tb + 1);
            if (v10) add(1, tb + pw);
            if (v11) add(0, tb + pw + 1);
            if constexpr (DBL)
                finish_node(k, mcur, [&](int f) { return ox[f]; });
            else
                finish_node(k, mcur, [&](int f) { return pnv(slot_of(k), f, own); });
        }
    };
    // ---- node plane k+1 starts with corners 6,7,5,4 -------------------------
    // the upper corners' rows (lds[kStage..] with DBL, else lds[0..]) hold
    // corners 4..7 at slots 0..3
    auto node_start = [&](int k) {
        constexpr int kUp = DBL ? kStage : 0;
        px = py = pz = 0.0;
        if (ownsNode && k + 1 < K1) {
            if (v00) add(2, tb, kUp);          // corner 6 of element (J-1, I-1)
            if (v01) add(3, tb + 1, kUp);      // corner 7 of element (J-1, I)
            if (v10) add(1, tb + pw, kUp);     // corner 5 of element (J, I-1)
            if (v11) add(0, tb + pw + 1, kUp); // corner 4 of element (J, I)
        }
    };
    for (int k = kbeg; k < kend; ++k) {
        // Plane k is in its slot; plane k+1 (fetched during the previous
        // step) goes over plane k-1, whose last readers passed the
        // previous step's second barrier (DBL: the previous element work
        // and own_state read it; else: finish_node, before the third).
        // The previous step's last phase only read the corner staging
        // area, which is written again after this barrier.
        store_plane(k + 1);
        __syncthreads();
        double ox[6]; // (DBL) this thread's node of plane k, read while its slot is stable
        if constexpr (DBL) {
            if (ownsNode && k >= K0) own_state(k, ox);
        }
        // The mixed form fetches the next step's plane before the element
        // work, so that the loads have the whole step to arrive (252 VGPRs;
        // s=150: nodal 0.180 -> 0.158 ms); the fused FP64 form, whose element
        // work is longer, gains nothing from it, and the separately rounded
        // form would spill.
        constexpr bool EARLY = MODE == MODE_MIXED;
        if constexpr (EARLY) {
            fetch_plane(k + 2, k + 1 < kend);
            __builtin_amdgcn_sched_barrier(0);
        }
        element_work(k);
        // the next step's plane, element scalars and nodal mass
        if constexpr (!EARLY) fetch_plane(k + 2, k + 1 < kend);
        fetch_elem(k + 1, k + 1 < kend);
        mnext = fetch_mass(k + 1);
        __syncthreads();
        __builtin_amdgcn_s_setprio(0);
        node_finish(k, ox);
        if constexpr (!DBL) { // the upper corners over the lower ones, between barriers
            __syncthreads();
            if (elemCol) {
#pragma unroll
                for (int c = 0; c < kStage; ++c) lds[c][tid] = up[c];
            }
            __syncthreads();
        }
        node_start(k);
        mcur = mnext;
    }
    // The top node plane (K = s) has no element plane above it.
    if (ownsNode && K1 == n && kend == s && s >= K0)
        finish_node(s, mcur, [&](int f) { return pnv(slot_of(s), f, own); });
}

} // namespace lulesh_dev
