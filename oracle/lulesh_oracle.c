/*
 * LULESH oracle: serial / OpenMP CPU restatement of the reference HPX
 * Lagrange leapfrog path (esa-tu-darmstadt/lulesh-task-based-hpx).
 *
 * TEST INFRASTRUCTURE ONLY -- the checker, never the product.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.
 *
 * Why a serial restatement is faithful: every floating-point accumulation
 * of the HPX path happens inside one element, one node or one corner loop
 * (lulesh.cc:842-861, 1238-1247, 1692-1871) and the only cross-task
 * reductions are exact min()s (lulesh.cc:1931-1936, 1942-1971, 2233), so the
 * result does not depend on task size, thread count or loop order across
 * elements.  The OpenMP loops below only split those independent loops.
 *
 * Build: -O2 -ffp-contract=off (no FMA contraction), see oracle/Makefile.
 *
 * Pinning (how each function is checked independently of this file) is
 * listed in tests/test_oracle_*.py; a function without such a check would be
 * marked "parity unpinned" here -- at present only the libm cbrt() used by
 * the hourglass coefficient is unpinned at the ulp level (glibc cbrt is not
 * correctly rounded).
 */
#include "lulesh_oracle.h"

#include <float.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef double Real_t;
typedef int32_t Index_t;
typedef int32_t Int_t;

/* Boundary-condition bits, lulesh.h:47-75 */
#define XI_M 0x00007
#define XI_M_SYMM 0x00001
#define XI_M_FREE 0x00002
#define XI_M_COMM 0x00004
#define XI_P 0x00038
#define XI_P_SYMM 0x00008
#define XI_P_FREE 0x00010
#define XI_P_COMM 0x00020
#define ETA_M 0x001c0
#define ETA_M_SYMM 0x00040
#define ETA_M_FREE 0x00080
#define ETA_M_COMM 0x00100
#define ETA_P 0x00e00
#define ETA_P_SYMM 0x00200
#define ETA_P_FREE 0x00400
#define ETA_P_COMM 0x00800
#define ZETA_M 0x07000
#define ZETA_M_SYMM 0x01000
#define ZETA_M_FREE 0x02000
#define ZETA_M_COMM 0x04000
#define ZETA_P 0x38000
#define ZETA_P_SYMM 0x08000
#define ZETA_P_FREE 0x10000
#define ZETA_P_COMM 0x20000

/* Domain fields, lulesh.h:464-593 */
struct ora_domain {
    Index_t sizeX, numElem, numNode, numReg, cost;
    /* node-centred */
    Real_t *x, *y, *z, *xd, *yd, *zd, *xdd, *ydd, *zdd, *fx, *fy, *fz, *nodalMass;
    Index_t *symmX, *symmY, *symmZ;
    /* element-centred */
    Index_t *nodelist, *lxim, *lxip, *letam, *letap, *lzetam, *lzetap, *elemBC;
    Real_t *e, *p, *q, *ql, *qq, *v, *volo, *delv, *vdov, *arealg, *ss, *elemMass, *vnew;
    Real_t *delv_xi, *delv_eta, *delv_zeta, *delx_xi, *delx_eta, *delx_zeta;
    /* regions */
    Index_t *regNumList, *regElemSize, **regElemlist;
    /* corner CSR */
    Index_t *nodeElemStart, *nodeElemCornerList;
    /* constants, lulesh-init.cc:18-36 */
    Real_t e_cut, p_cut, q_cut, v_cut, u_cut, hgcoef, ss4o3, qstop, monoq_max_slope,
        monoq_limiter_mult, qlc_monoq, qqc_monoq, qqc, eosvmax, eosvmin, pmin, emin, dvovmax,
        refdens;
    /* time state, lulesh-init.cc:135-145 */
    Real_t dtcourant, dthydro, dtfixed, time, deltatime, deltatimemultlb, deltatimemultub, dtmax,
        stoptime;
    Int_t cycle;
    /* corner force scratch (lulesh.cc:2017-2022), kept for the life of the domain */
    Real_t *fxs, *fys, *fzs, *fxh, *fyh, *fzh;
};

static int g_threads = 0; /* 0 = OpenMP default */

void ora_set_threads(int32_t n) {
    g_threads = n;
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#endif
}

int32_t ora_max_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

static void *xalloc(size_t n) {
    void *p = calloc(n ? n : 1, 1);
    if (!p) {
        fprintf(stderr, "oracle: out of memory (%zu bytes)\n", n);
        abort();
    }
    return p;
}
#define ALLOC_R(n) ((Real_t *)xalloc(sizeof(Real_t) * (size_t)(n)))
#define ALLOC_I(n) ((Index_t *)xalloc(sizeof(Index_t) * (size_t)(n)))

/* ------------------------------------------------------------------------ */
/* Element geometry helpers                                                  */
/* ------------------------------------------------------------------------ */

/* CollectDomainNodesToElemNodes, lulesh.cc:233-273 */
static inline void collect_nodes(const Real_t *gx, const Real_t *gy, const Real_t *gz,
                                 const Index_t *n, Real_t x[8], Real_t y[8], Real_t z[8]) {
    for (int k = 0; k < 8; ++k) {
        x[k] = gx[n[k]];
        y[k] = gy[n[k]];
        z[k] = gz[n[k]];
    }
}

/* CalcElemShapeFunctionDerivatives, lulesh.cc:277-374 */
static inline void shape_function_derivatives(const Real_t x[8], const Real_t y[8],
                                              const Real_t z[8], Real_t b[3][8],
                                              Real_t *volume) {
    const Real_t x0 = x[0], x1 = x[1], x2 = x[2], x3 = x[3], x4 = x[4], x5 = x[5], x6 = x[6],
                 x7 = x[7];
    const Real_t y0 = y[0], y1 = y[1], y2 = y[2], y3 = y[3], y4 = y[4], y5 = y[5], y6 = y[6],
                 y7 = y[7];
    const Real_t z0 = z[0], z1 = z[1], z2 = z[2], z3 = z[3], z4 = z[4], z5 = z[5], z6 = z[6],
                 z7 = z[7];
    Real_t fjxxi, fjxet, fjxze, fjyxi, fjyet, fjyze, fjzxi, fjzet, fjzze;
    Real_t cjxxi, cjxet, cjxze, cjyxi, cjyet, cjyze, cjzxi, cjzet, cjzze;

    fjxxi = 0.125 * ((x6 - x0) + (x5 - x3) - (x7 - x1) - (x4 - x2));
    fjxet = 0.125 * ((x6 - x0) - (x5 - x3) + (x7 - x1) - (x4 - x2));
    fjxze = 0.125 * ((x6 - x0) + (x5 - x3) + (x7 - x1) + (x4 - x2));
    fjyxi = 0.125 * ((y6 - y0) + (y5 - y3) - (y7 - y1) - (y4 - y2));
    fjyet = 0.125 * ((y6 - y0) - (y5 - y3) + (y7 - y1) - (y4 - y2));
    fjyze = 0.125 * ((y6 - y0) + (y5 - y3) + (y7 - y1) + (y4 - y2));
    fjzxi = 0.125 * ((z6 - z0) + (z5 - z3) - (z7 - z1) - (z4 - z2));
    fjzet = 0.125 * ((z6 - z0) - (z5 - z3) + (z7 - z1) - (z4 - z2));
    fjzze = 0.125 * ((z6 - z0) + (z5 - z3) + (z7 - z1) + (z4 - z2));

    cjxxi = (fjyet * fjzze) - (fjzet * fjyze);
    cjxet = -(fjyxi * fjzze) + (fjzxi * fjyze);
    cjxze = (fjyxi * fjzet) - (fjzxi * fjyet);
    cjyxi = -(fjxet * fjzze) + (fjzet * fjxze);
    cjyet = (fjxxi * fjzze) - (fjzxi * fjxze);
    cjyze = -(fjxxi * fjzet) + (fjzxi * fjxet);
    cjzxi = (fjxet * fjyze) - (fjyet * fjxze);
    cjzet = -(fjxxi * fjyze) + (fjyxi * fjxze);
    cjzze = (fjxxi * fjyet) - (fjyxi * fjxet);

    b[0][0] = -cjxxi - cjxet - cjxze;
    b[0][1] = cjxxi - cjxet - cjxze;
    b[0][2] = cjxxi + cjxet - cjxze;
    b[0][3] = -cjxxi + cjxet - cjxze;
    b[0][4] = -b[0][2];
    b[0][5] = -b[0][3];
    b[0][6] = -b[0][0];
    b[0][7] = -b[0][1];
    b[1][0] = -cjyxi - cjyet - cjyze;
    b[1][1] = cjyxi - cjyet - cjyze;
    b[1][2] = cjyxi + cjyet - cjyze;
    b[1][3] = -cjyxi + cjyet - cjyze;
    b[1][4] = -b[1][2];
    b[1][5] = -b[1][3];
    b[1][6] = -b[1][0];
    b[1][7] = -b[1][1];
    b[2][0] = -cjzxi - cjzet - cjzze;
    b[2][1] = cjzxi - cjzet - cjzze;
    b[2][2] = cjzxi + cjzet - cjzze;
    b[2][3] = -cjzxi + cjzet - cjzze;
    b[2][4] = -b[2][2];
    b[2][5] = -b[2][3];
    b[2][6] = -b[2][0];
    b[2][7] = -b[2][1];

    *volume = 8. * (fjxet * cjxet + fjyet * cjyet + fjzet * cjzet);
}

/* SumElemFaceNormal, lulesh.cc:378-409 */
static inline void sum_face_normal(Real_t pf[3][8], int n0, int n1, int n2, int n3,
                                   const Real_t x[8], const Real_t y[8], const Real_t z[8]) {
    Real_t bisectX0 = 0.5 * (x[n3] + x[n2] - x[n1] - x[n0]);
    Real_t bisectY0 = 0.5 * (y[n3] + y[n2] - y[n1] - y[n0]);
    Real_t bisectZ0 = 0.5 * (z[n3] + z[n2] - z[n1] - z[n0]);
    Real_t bisectX1 = 0.5 * (x[n2] + x[n1] - x[n3] - x[n0]);
    Real_t bisectY1 = 0.5 * (y[n2] + y[n1] - y[n3] - y[n0]);
    Real_t bisectZ1 = 0.5 * (z[n2] + z[n1] - z[n3] - z[n0]);
    Real_t areaX = 0.25 * (bisectY0 * bisectZ1 - bisectZ0 * bisectY1);
    Real_t areaY = 0.25 * (bisectZ0 * bisectX1 - bisectX0 * bisectZ1);
    Real_t areaZ = 0.25 * (bisectX0 * bisectY1 - bisectY0 * bisectX1);
    pf[0][n0] += areaX;
    pf[0][n1] += areaX;
    pf[0][n2] += areaX;
    pf[0][n3] += areaX;
    pf[1][n0] += areaY;
    pf[1][n1] += areaY;
    pf[1][n2] += areaY;
    pf[1][n3] += areaY;
    pf[2][n0] += areaZ;
    pf[2][n1] += areaZ;
    pf[2][n2] += areaZ;
    pf[2][n3] += areaZ;
}

/* CalcElemNodeNormals, lulesh.cc:413-451 (face order 0123, 0451, 1562, 2673, 3740, 4765) */
static inline void node_normals(Real_t pf[3][8], const Real_t x[8], const Real_t y[8],
                                const Real_t z[8]) {
    for (int i = 0; i < 8; ++i) pf[0][i] = pf[1][i] = pf[2][i] = 0.0;
    sum_face_normal(pf, 0, 1, 2, 3, x, y, z);
    sum_face_normal(pf, 0, 4, 5, 1, x, y, z);
    sum_face_normal(pf, 1, 5, 6, 2, x, y, z);
    sum_face_normal(pf, 2, 6, 7, 3, x, y, z);
    sum_face_normal(pf, 3, 7, 4, 0, x, y, z);
    sum_face_normal(pf, 4, 7, 6, 5, x, y, z);
}

/* VoluDer, lulesh.cc:468-491 */
static inline void volu_der(Real_t x0, Real_t x1, Real_t x2, Real_t x3, Real_t x4, Real_t x5,
                            Real_t y0, Real_t y1, Real_t y2, Real_t y3, Real_t y4, Real_t y5,
                            Real_t z0, Real_t z1, Real_t z2, Real_t z3, Real_t z4, Real_t z5,
                            Real_t *dvdx, Real_t *dvdy, Real_t *dvdz) {
    const Real_t twelfth = 1.0 / 12.0;
    *dvdx = (y1 + y2) * (z0 + z1) - (y0 + y1) * (z1 + z2) + (y0 + y4) * (z3 + z4) -
            (y3 + y4) * (z0 + z4) - (y2 + y5) * (z3 + z5) + (y3 + y5) * (z2 + z5);
    *dvdy = -(x1 + x2) * (z0 + z1) + (x0 + x1) * (z1 + z2) - (x0 + x4) * (z3 + z4) +
            (x3 + x4) * (z0 + z4) + (x2 + x5) * (z3 + z5) - (x3 + x5) * (z2 + z5);
    *dvdz = -(y1 + y2) * (x0 + x1) + (y0 + y1) * (x1 + x2) - (y0 + y4) * (x3 + x4) +
            (y3 + y4) * (x0 + x4) + (y2 + y5) * (x3 + x5) - (y3 + y5) * (x2 + x5);
    *dvdx *= twelfth;
    *dvdy *= twelfth;
    *dvdz *= twelfth;
}

/* CalcElemVolumeDerivative, lulesh.cc:495-523 */
static inline void volume_derivative(Real_t dvdx[8], Real_t dvdy[8], Real_t dvdz[8],
                                     const Real_t x[8], const Real_t y[8], const Real_t z[8]) {
#define VD(a, b, c, d, e, f, o)                                                              \
    volu_der(x[a], x[b], x[c], x[d], x[e], x[f], y[a], y[b], y[c], y[d], y[e], y[f], z[a],  \
             z[b], z[c], z[d], z[e], z[f], &dvdx[o], &dvdy[o], &dvdz[o])
    VD(1, 2, 3, 4, 5, 7, 0);
    VD(0, 1, 2, 7, 4, 6, 3);
    VD(3, 0, 1, 6, 7, 5, 2);
    VD(2, 3, 0, 5, 6, 4, 1);
    VD(7, 6, 5, 0, 3, 1, 4);
    VD(4, 7, 6, 1, 0, 2, 5);
    VD(5, 4, 7, 2, 1, 3, 6);
    VD(6, 5, 4, 3, 2, 0, 7);
#undef VD
}

/* CalcElemVolume, lulesh.cc:592-674 */
static inline Real_t elem_volume(const Real_t x[8], const Real_t y[8], const Real_t z[8]) {
    const Real_t twelveth = 1.0 / 12.0;
    const Real_t x0 = x[0], x1 = x[1], x2 = x[2], x3 = x[3], x4 = x[4], x5 = x[5], x6 = x[6],
                 x7 = x[7];
    const Real_t y0 = y[0], y1 = y[1], y2 = y[2], y3 = y[3], y4 = y[4], y5 = y[5], y6 = y[6],
                 y7 = y[7];
    const Real_t z0 = z[0], z1 = z[1], z2 = z[2], z3 = z[3], z4 = z[4], z5 = z[5], z6 = z[6],
                 z7 = z[7];
    Real_t dx61 = x6 - x1, dy61 = y6 - y1, dz61 = z6 - z1;
    Real_t dx70 = x7 - x0, dy70 = y7 - y0, dz70 = z7 - z0;
    Real_t dx63 = x6 - x3, dy63 = y6 - y3, dz63 = z6 - z3;
    Real_t dx20 = x2 - x0, dy20 = y2 - y0, dz20 = z2 - z0;
    Real_t dx50 = x5 - x0, dy50 = y5 - y0, dz50 = z5 - z0;
    Real_t dx64 = x6 - x4, dy64 = y6 - y4, dz64 = z6 - z4;
    Real_t dx31 = x3 - x1, dy31 = y3 - y1, dz31 = z3 - z1;
    Real_t dx72 = x7 - x2, dy72 = y7 - y2, dz72 = z7 - z2;
    Real_t dx43 = x4 - x3, dy43 = y4 - y3, dz43 = z4 - z3;
    Real_t dx57 = x5 - x7, dy57 = y5 - y7, dz57 = z5 - z7;
    Real_t dx14 = x1 - x4, dy14 = y1 - y4, dz14 = z1 - z4;
    Real_t dx25 = x2 - x5, dy25 = y2 - y5, dz25 = z2 - z5;
#define TRIPLE_PRODUCT(x1, y1, z1, x2, y2, z2, x3, y3, z3)                                    \
    ((x1) * ((y2) * (z3) - (z2) * (y3)) + (x2) * ((z1) * (y3) - (y1) * (z3)) +                \
     (x3) * ((y1) * (z2) - (z1) * (y2)))
    Real_t volume =
        TRIPLE_PRODUCT(dx31 + dx72, dx63, dx20, dy31 + dy72, dy63, dy20, dz31 + dz72, dz63,
                       dz20) +
        TRIPLE_PRODUCT(dx43 + dx57, dx64, dx70, dy43 + dy57, dy64, dy70, dz43 + dz57, dz64,
                       dz70) +
        TRIPLE_PRODUCT(dx14 + dx25, dx61, dx50, dy14 + dy25, dy61, dy50, dz14 + dz25, dz61,
                       dz50);
#undef TRIPLE_PRODUCT
    volume *= twelveth;
    return volume;
}

/* AreaFace, lulesh.cc:678-692 */
static inline Real_t area_face(Real_t x0, Real_t x1, Real_t x2, Real_t x3, Real_t y0, Real_t y1,
                               Real_t y2, Real_t y3, Real_t z0, Real_t z1, Real_t z2,
                               Real_t z3) {
    Real_t fx = (x2 - x0) - (x3 - x1);
    Real_t fy = (y2 - y0) - (y3 - y1);
    Real_t fz = (z2 - z0) - (z3 - z1);
    Real_t gx = (x2 - x0) + (x3 - x1);
    Real_t gy = (y2 - y0) + (y3 - y1);
    Real_t gz = (z2 - z0) + (z3 - z1);
    Real_t area = (fx * fx + fy * fy + fz * fz) * (gx * gx + gy * gy + gz * gz) -
                  (fx * gx + fy * gy + fz * gz) * (fx * gx + fy * gy + fz * gz);
    return area;
}

/* std::max(a, b) as used at lulesh.cc:704-724: (a < b) ? b : a */
static inline Real_t std_max(Real_t a, Real_t b) { return (a < b) ? b : a; }

/* CalcElemCharacteristicLength, lulesh.cc:696-729 */
static inline Real_t characteristic_length(const Real_t x[8], const Real_t y[8],
                                           const Real_t z[8], Real_t volume) {
    Real_t a, charLength = 0.0;
#define AF(a0, a1, a2, a3)                                                                   \
    area_face(x[a0], x[a1], x[a2], x[a3], y[a0], y[a1], y[a2], y[a3], z[a0], z[a1], z[a2],   \
              z[a3])
    a = AF(0, 1, 2, 3);
    charLength = std_max(a, charLength);
    a = AF(4, 5, 6, 7);
    charLength = std_max(a, charLength);
    a = AF(0, 1, 5, 4);
    charLength = std_max(a, charLength);
    a = AF(1, 2, 6, 5);
    charLength = std_max(a, charLength);
    a = AF(2, 3, 7, 6);
    charLength = std_max(a, charLength);
    a = AF(3, 0, 4, 7);
    charLength = std_max(a, charLength);
#undef AF
    charLength = 4.0 * volume / sqrt(charLength);
    return charLength;
}

/* CalcElemVelocityGradient, lulesh.cc:733-781 */
static inline void velocity_gradient(const Real_t xvel[8], const Real_t yvel[8],
                                     const Real_t zvel[8], const Real_t b[3][8], Real_t detJ,
                                     Real_t d[6]) {
    const Real_t inv_detJ = 1.0 / detJ;
    Real_t dyddx, dxddy, dzddx, dxddz, dzddy, dyddz;
    const Real_t *pfx = b[0], *pfy = b[1], *pfz = b[2];
#define GRAD(pf, v)                                                                          \
    (inv_detJ * (pf[0] * (v[0] - v[6]) + pf[1] * (v[1] - v[7]) + pf[2] * (v[2] - v[4]) +     \
                 pf[3] * (v[3] - v[5])))
    d[0] = GRAD(pfx, xvel);
    d[1] = GRAD(pfy, yvel);
    d[2] = GRAD(pfz, zvel);
    dyddx = GRAD(pfx, yvel);
    dxddy = GRAD(pfy, xvel);
    dzddx = GRAD(pfx, zvel);
    dxddz = GRAD(pfz, xvel);
    dzddy = GRAD(pfy, zvel);
    dyddz = GRAD(pfz, yvel);
#undef GRAD
    d[5] = .5 * (dxddy + dyddx);
    d[4] = .5 * (dxddz + dzddx);
    d[3] = .5 * (dzddy + dyddz);
}

/* Hourglass gamma table, lulesh.cc:921-954 */
static const Real_t GAMMA[4][8] = {
    {1., 1., -1., -1., -1., -1., 1., 1.},
    {1., -1., -1., 1., -1., 1., 1., -1.},
    {1., -1., 1., -1., 1., -1., 1., -1.},
    {-1., 1., -1., 1., 1., -1., 1., -1.},
};

/* CalcElemFBHourglassForce, lulesh.cc:527-562 */
static inline void fb_hourglass_force(const Real_t xd[8], const Real_t yd[8], const Real_t zd[8],
                                      Real_t hourgam[8][4], Real_t coefficient, Real_t hgfx[8],
                                      Real_t hgfy[8], Real_t hgfz[8]) {
    Real_t hxx[4];
    const Real_t *vel[3] = {xd, yd, zd};
    Real_t *out[3] = {hgfx, hgfy, hgfz};
    for (int c = 0; c < 3; ++c) {
        const Real_t *v = vel[c];
        for (int i = 0; i < 4; i++) {
            hxx[i] = hourgam[0][i] * v[0] + hourgam[1][i] * v[1] + hourgam[2][i] * v[2] +
                     hourgam[3][i] * v[3] + hourgam[4][i] * v[4] + hourgam[5][i] * v[5] +
                     hourgam[6][i] * v[6] + hourgam[7][i] * v[7];
        }
        for (int i = 0; i < 8; i++) {
            out[c][i] = coefficient * (hourgam[i][0] * hxx[0] + hourgam[i][1] * hxx[1] +
                                       hourgam[i][2] * hxx[2] + hourgam[i][3] * hxx[3]);
        }
    }
}

/* Per-element hourglass body of CalcHourglassForElemsTask, lulesh.cc:880-1071 */
static inline void elem_hourglass(const Real_t x1[8], const Real_t y1[8], const Real_t z1[8],
                                  const Real_t xd1[8], const Real_t yd1[8], const Real_t zd1[8],
                                  Real_t determ, Real_t ss1, Real_t mass1, Real_t hgcoef,
                                  Real_t hgfx[8], Real_t hgfy[8], Real_t hgfz[8]) {
    Real_t dvdx[8], dvdy[8], dvdz[8];
    Real_t hourgam[8][4];
    volume_derivative(dvdx, dvdy, dvdz, x1, y1, z1);
    Real_t volinv = 1.0 / determ;
    for (int i1 = 0; i1 < 4; ++i1) {
        const Real_t *g = GAMMA[i1];
        Real_t hourmodx = x1[0] * g[0] + x1[1] * g[1] + x1[2] * g[2] + x1[3] * g[3] +
                          x1[4] * g[4] + x1[5] * g[5] + x1[6] * g[6] + x1[7] * g[7];
        Real_t hourmody = y1[0] * g[0] + y1[1] * g[1] + y1[2] * g[2] + y1[3] * g[3] +
                          y1[4] * g[4] + y1[5] * g[5] + y1[6] * g[6] + y1[7] * g[7];
        Real_t hourmodz = z1[0] * g[0] + z1[1] * g[1] + z1[2] * g[2] + z1[3] * g[3] +
                          z1[4] * g[4] + z1[5] * g[5] + z1[6] * g[6] + z1[7] * g[7];
        for (int j = 0; j < 8; ++j) {
            hourgam[j][i1] =
                g[j] - volinv * (dvdx[j] * hourmodx + dvdy[j] * hourmody + dvdz[j] * hourmodz);
        }
    }
    Real_t volume13 = cbrt(determ);
    Real_t coefficient = -hgcoef * 0.01 * ss1 * mass1 / volume13;
    fb_hourglass_force(xd1, yd1, zd1, hourgam, coefficient, hgfx, hgfy, hgfz);
}

/* ------------------------------------------------------------------------ */
/* Domain construction, lulesh-init.cc                                       */
/* ------------------------------------------------------------------------ */

/* BuildMesh, lulesh-init.cc:202-252 (tp = 1, colLoc = rowLoc = planeLoc = 0) */
static void build_mesh(ora_domain *d, Int_t nx) {
    Index_t edgeElems = nx, edgeNodes = nx + 1, meshEdgeElems = nx;
    Index_t nidx = 0;
    Real_t tz = 1.125 * (Real_t)(0 * nx) / (Real_t)meshEdgeElems;
    for (Index_t plane = 0; plane < edgeNodes; ++plane) {
        Real_t ty = 1.125 * (Real_t)(0 * nx) / (Real_t)meshEdgeElems;
        for (Index_t row = 0; row < edgeNodes; ++row) {
            Real_t tx = 1.125 * (Real_t)(0 * nx) / (Real_t)meshEdgeElems;
            for (Index_t col = 0; col < edgeNodes; ++col) {
                d->x[nidx] = tx;
                d->y[nidx] = ty;
                d->z[nidx] = tz;
                ++nidx;
                tx = 1.125 * (Real_t)(col + 1) / (Real_t)meshEdgeElems;
            }
            ty = 1.125 * (Real_t)(row + 1) / (Real_t)meshEdgeElems;
        }
        tz = 1.125 * (Real_t)(plane + 1) / (Real_t)meshEdgeElems;
    }
    Index_t zidx = 0;
    nidx = 0;
    for (Index_t plane = 0; plane < edgeElems; ++plane) {
        for (Index_t row = 0; row < edgeElems; ++row) {
            for (Index_t col = 0; col < edgeElems; ++col) {
                Index_t *ln = &d->nodelist[8 * zidx];
                ln[0] = nidx;
                ln[1] = nidx + 1;
                ln[2] = nidx + edgeNodes + 1;
                ln[3] = nidx + edgeNodes;
                ln[4] = nidx + edgeNodes * edgeNodes;
                ln[5] = nidx + edgeNodes * edgeNodes + 1;
                ln[6] = nidx + edgeNodes * edgeNodes + edgeNodes + 1;
                ln[7] = nidx + edgeNodes * edgeNodes + edgeNodes;
                ++zidx;
                ++nidx;
            }
            ++nidx;
        }
        nidx += edgeNodes;
    }
}

/* SetupThreadSupportStructures, lulesh-init.cc:256-310 */
static void setup_corner_csr(ora_domain *d) {
    Index_t numNode = d->numNode, numElem = d->numElem;
    Index_t *count = ALLOC_I(numNode);
    for (Index_t i = 0; i < numElem; ++i)
        for (int j = 0; j < 8; ++j) ++count[d->nodelist[8 * i + j]];
    d->nodeElemStart = ALLOC_I(numNode + 1);
    d->nodeElemStart[0] = 0;
    for (Index_t i = 1; i <= numNode; ++i)
        d->nodeElemStart[i] = d->nodeElemStart[i - 1] + count[i - 1];
    d->nodeElemCornerList = ALLOC_I(d->nodeElemStart[numNode]);
    memset(count, 0, sizeof(Index_t) * (size_t)numNode);
    for (Index_t i = 0; i < numElem; ++i) {
        for (int j = 0; j < 8; ++j) {
            Index_t m = d->nodelist[8 * i + j];
            d->nodeElemCornerList[d->nodeElemStart[m] + count[m]] = i * 8 + j;
            ++count[m];
        }
    }
    free(count);
}

/* CreateRegionIndexSets, lulesh-init.cc:341-445 (libc srand/rand, myRank = 0) */
static void create_region_index_sets(ora_domain *d, Int_t nr, Int_t balance) {
    srand(0);
    Index_t numElem = d->numElem;
    d->numReg = nr;
    d->regElemSize = ALLOC_I(nr);
    d->regElemlist = (Index_t **)xalloc(sizeof(Index_t *) * (size_t)nr);
    Index_t nextIndex = 0;
    if (nr == 1) {
        while (nextIndex < numElem) {
            d->regNumList[nextIndex] = 1;
            nextIndex++;
        }
        d->regElemSize[0] = 0;
    } else {
        Int_t regionNum, regionVar, lastReg = -1, binSize;
        Index_t elements, runto = 0;
        Int_t costDenominator = 0;
        Int_t *regBinEnd = ALLOC_I(nr);
        for (Index_t i = 0; i < nr; ++i) {
            d->regElemSize[i] = 0;
            costDenominator = (Int_t)((double)costDenominator + pow((double)(i + 1), balance));
            regBinEnd[i] = costDenominator;
        }
        while (nextIndex < numElem) {
            regionVar = rand() % costDenominator;
            Index_t i = 0;
            while (regionVar >= regBinEnd[i]) i++;
            regionNum = ((i + 0) % nr) + 1;
            while (regionNum == lastReg) {
                regionVar = rand() % costDenominator;
                i = 0;
                while (regionVar >= regBinEnd[i]) i++;
                regionNum = ((i + 0) % nr) + 1;
            }
            binSize = rand() % 1000;
            if (binSize < 773)
                elements = rand() % 15 + 1;
            else if (binSize < 937)
                elements = rand() % 16 + 16;
            else if (binSize < 970)
                elements = rand() % 32 + 32;
            else if (binSize < 974)
                elements = rand() % 64 + 64;
            else if (binSize < 978)
                elements = rand() % 128 + 128;
            else if (binSize < 981)
                elements = rand() % 256 + 256;
            else
                elements = rand() % 1537 + 512;
            runto = elements + nextIndex;
            while (nextIndex < runto && nextIndex < numElem) {
                d->regNumList[nextIndex] = regionNum;
                nextIndex++;
            }
            lastReg = regionNum;
        }
        free(regBinEnd);
    }
    for (Index_t i = 0; i < numElem; ++i) d->regElemSize[d->regNumList[i] - 1]++;
    for (Index_t i = 0; i < nr; ++i) {
        d->regElemlist[i] = ALLOC_I(d->regElemSize[i]);
        d->regElemSize[i] = 0;
    }
    for (Index_t i = 0; i < numElem; ++i) {
        Index_t r = d->regNumList[i] - 1;
        Index_t regndx = d->regElemSize[r]++;
        d->regElemlist[r][regndx] = i;
    }
}

/* SetupSymmetryPlanes, lulesh-init.cc:448-468 */
static void setup_symmetry_planes(ora_domain *d, Int_t edgeNodes) {
    Index_t nidx = 0;
    for (Index_t i = 0; i < edgeNodes; ++i) {
        Index_t planeInc = i * edgeNodes * edgeNodes;
        Index_t rowInc = i * edgeNodes;
        for (Index_t j = 0; j < edgeNodes; ++j) {
            d->symmZ[nidx] = rowInc + j;
            d->symmY[nidx] = planeInc + j;
            d->symmX[nidx] = planeInc + j * edgeNodes;
            ++nidx;
        }
    }
}

/* SetupElementConnectivities, lulesh-init.cc:473-500 */
static void setup_element_connectivities(ora_domain *d, Int_t edgeElems) {
    Index_t numElem = d->numElem;
    d->lxim[0] = 0;
    for (Index_t i = 1; i < numElem; ++i) {
        d->lxim[i] = i - 1;
        d->lxip[i - 1] = i;
    }
    d->lxip[numElem - 1] = numElem - 1;
    for (Index_t i = 0; i < edgeElems; ++i) {
        d->letam[i] = i;
        d->letap[numElem - edgeElems + i] = numElem - edgeElems + i;
    }
    for (Index_t i = edgeElems; i < numElem; ++i) {
        d->letam[i] = i - edgeElems;
        d->letap[i - edgeElems] = i;
    }
    for (Index_t i = 0; i < edgeElems * edgeElems; ++i) {
        d->lzetam[i] = i;
        d->lzetap[numElem - edgeElems * edgeElems + i] = numElem - edgeElems * edgeElems + i;
    }
    for (Index_t i = edgeElems * edgeElems; i < numElem; ++i) {
        d->lzetam[i] = i - edgeElems * edgeElems;
        d->lzetap[i - edgeElems * edgeElems] = i;
    }
}

/* SetupBoundaryConditions, lulesh-init.cc:503-608 for the single rank: every
 * *_COMM branch is dead (planeLoc = rowLoc = colLoc = 0 = tp - 1). */
static void setup_boundary_conditions(ora_domain *d, Int_t edgeElems) {
    Index_t numElem = d->numElem;
    for (Index_t i = 0; i < numElem; ++i) d->elemBC[i] = 0;
    for (Index_t i = 0; i < edgeElems; ++i) {
        Index_t planeInc = i * edgeElems * edgeElems;
        Index_t rowInc = i * edgeElems;
        for (Index_t j = 0; j < edgeElems; ++j) {
            d->elemBC[rowInc + j] |= ZETA_M_SYMM;
            d->elemBC[rowInc + j + numElem - edgeElems * edgeElems] |= ZETA_P_FREE;
            d->elemBC[planeInc + j] |= ETA_M_SYMM;
            d->elemBC[planeInc + j + edgeElems * edgeElems - edgeElems] |= ETA_P_FREE;
            d->elemBC[planeInc + j * edgeElems] |= XI_M_SYMM;
            d->elemBC[planeInc + j * edgeElems + edgeElems - 1] |= XI_P_FREE;
        }
    }
}

ora_domain *ora_create(int32_t nx, int32_t nr, int32_t balance, int32_t cost) {
    if (nx < 1 || nr < 1) return NULL;
    ora_domain *d = (ora_domain *)xalloc(sizeof(ora_domain));
    Index_t edgeElems = nx, edgeNodes = nx + 1;
    d->e_cut = 1.0e-7;
    d->p_cut = 1.0e-7;
    d->q_cut = 1.0e-7;
    d->v_cut = 1.0e-10;
    d->u_cut = 1.0e-7;
    d->hgcoef = 3.0;
    d->ss4o3 = 4.0 / 3.0;
    d->qstop = 1.0e+12;
    d->monoq_max_slope = 1.0;
    d->monoq_limiter_mult = 2.0;
    d->qlc_monoq = 0.5;
    d->qqc_monoq = 2.0 / 3.0;
    d->qqc = 2.0;
    d->eosvmax = 1.0e+9;
    d->eosvmin = 1.0e-9;
    d->pmin = 0.;
    d->emin = -1.0e+15;
    d->dvovmax = 0.1;
    d->refdens = 1.0;
    d->cost = cost;
    d->sizeX = edgeElems;
    d->numElem = edgeElems * edgeElems * edgeElems;
    d->numNode = edgeNodes * edgeNodes * edgeNodes;
    Index_t E = d->numElem, N = d->numNode;
    d->regNumList = ALLOC_I(E);
    /* AllocateElemPersistent / AllocateNodePersistent, lulesh.h:146-201 */
    d->nodelist = ALLOC_I(8 * (size_t)E);
    d->lxim = ALLOC_I(E);
    d->lxip = ALLOC_I(E);
    d->letam = ALLOC_I(E);
    d->letap = ALLOC_I(E);
    d->lzetam = ALLOC_I(E);
    d->lzetap = ALLOC_I(E);
    d->elemBC = ALLOC_I(E);
    d->e = ALLOC_R(E);
    d->p = ALLOC_R(E);
    d->q = ALLOC_R(E);
    d->ql = ALLOC_R(E);
    d->qq = ALLOC_R(E);
    d->v = ALLOC_R(E);
    d->volo = ALLOC_R(E);
    d->delv = ALLOC_R(E);
    d->vdov = ALLOC_R(E);
    d->arealg = ALLOC_R(E);
    d->ss = ALLOC_R(E);
    d->elemMass = ALLOC_R(E);
    d->vnew = ALLOC_R(E);
    d->delv_xi = ALLOC_R(E);
    d->delv_eta = ALLOC_R(E);
    d->delv_zeta = ALLOC_R(E);
    d->delx_xi = ALLOC_R(E);
    d->delx_eta = ALLOC_R(E);
    d->delx_zeta = ALLOC_R(E);
    d->x = ALLOC_R(N);
    d->y = ALLOC_R(N);
    d->z = ALLOC_R(N);
    d->xd = ALLOC_R(N);
    d->yd = ALLOC_R(N);
    d->zd = ALLOC_R(N);
    d->xdd = ALLOC_R(N);
    d->ydd = ALLOC_R(N);
    d->zdd = ALLOC_R(N);
    d->fx = ALLOC_R(N);
    d->fy = ALLOC_R(N);
    d->fz = ALLOC_R(N);
    d->nodalMass = ALLOC_R(N);
    /* SetupCommBuffers, lulesh-init.cc:331-336 */
    d->symmX = ALLOC_I((size_t)edgeNodes * edgeNodes);
    d->symmY = ALLOC_I((size_t)edgeNodes * edgeNodes);
    d->symmZ = ALLOC_I((size_t)edgeNodes * edgeNodes);
    d->fxs = ALLOC_R(8 * (size_t)E);
    d->fys = ALLOC_R(8 * (size_t)E);
    d->fzs = ALLOC_R(8 * (size_t)E);
    d->fxh = ALLOC_R(8 * (size_t)E);
    d->fyh = ALLOC_R(8 * (size_t)E);
    d->fzh = ALLOC_R(8 * (size_t)E);

    /* lulesh-init.cc:81-107 */
    for (Index_t i = 0; i < E; ++i) d->v[i] = 1.0;

    build_mesh(d, nx);
    setup_corner_csr(d);
    create_region_index_sets(d, nr, balance);
    setup_symmetry_planes(d, edgeNodes);
    setup_element_connectivities(d, edgeElems);
    setup_boundary_conditions(d, edgeElems);

    /* lulesh-init.cc:135-145 */
    d->dtfixed = -1.0e-6;
    d->stoptime = 1.0e-2;
    d->deltatimemultlb = 1.1;
    d->deltatimemultub = 1.2;
    d->dtcourant = 1.0e+20;
    d->dthydro = 1.0e+20;
    d->dtmax = 1.0e-2;
    d->time = 0.;
    d->cycle = 0;

    /* lulesh-init.cc:148-167 */
    for (Index_t i = 0; i < E; ++i) {
        Real_t xl[8], yl[8], zl[8];
        const Index_t *n = &d->nodelist[8 * (size_t)i];
        collect_nodes(d->x, d->y, d->z, n, xl, yl, zl);
        Real_t volume = elem_volume(xl, yl, zl);
        d->volo[i] = volume;
        d->elemMass[i] = volume;
        for (int j = 0; j < 8; ++j) d->nodalMass[n[j]] += volume / 8.0;
    }

    /* lulesh-init.cc:172-181 */
    const Real_t ebase = 3.948746e+7;
    Real_t scale = (nx * 1) / 45.0;
    Real_t einit = ebase * scale * scale * scale;
    d->e[0] = einit;
    d->deltatime = (.5 * cbrt(d->volo[0])) / sqrt(2.0 * einit);
    return d;
}

void ora_scale_energy(ora_domain *d, double factor) {
    Real_t einit = d->e[0] * factor;
    d->e[0] = einit;
    d->deltatime = (.5 * cbrt(d->volo[0])) / sqrt(2.0 * einit);
}

void ora_destroy(ora_domain *d) {
    if (!d) return;
    Real_t *rf[] = {d->x,    d->y,     d->z,       d->xd,       d->yd,      d->zd,
                    d->xdd,  d->ydd,   d->zdd,     d->fx,       d->fy,      d->fz,
                    d->nodalMass, d->e, d->p,      d->q,        d->ql,      d->qq,
                    d->v,    d->volo,  d->delv,    d->vdov,     d->arealg,  d->ss,
                    d->elemMass, d->vnew, d->delv_xi, d->delv_eta, d->delv_zeta, d->delx_xi,
                    d->delx_eta, d->delx_zeta, d->fxs, d->fys, d->fzs, d->fxh, d->fyh, d->fzh};
    for (size_t i = 0; i < sizeof(rf) / sizeof(rf[0]); ++i) free(rf[i]);
    Index_t *ifl[] = {d->symmX, d->symmY, d->symmZ, d->nodelist, d->lxim, d->lxip,
                      d->letam, d->letap, d->lzetam, d->lzetap, d->elemBC, d->regNumList,
                      d->regElemSize, d->nodeElemStart, d->nodeElemCornerList};
    for (size_t i = 0; i < sizeof(ifl) / sizeof(ifl[0]); ++i) free(ifl[i]);
    if (d->regElemlist) {
        for (Index_t r = 0; r < d->numReg; ++r) free(d->regElemlist[r]);
        free(d->regElemlist);
    }
    free(d);
}

/* ------------------------------------------------------------------------ */
/* TimeIncrement, lulesh.cc:182-229                                          */
/* ------------------------------------------------------------------------ */
void ora_time_increment(ora_domain *d) {
    Real_t targetdt = d->stoptime - d->time;
    if ((d->dtfixed <= 0.0) && (d->cycle != 0)) {
        Real_t ratio;
        Real_t olddt = d->deltatime;
        Real_t gnewdt = 1.0e+20;
        Real_t newdt;
        if (d->dtcourant < gnewdt) gnewdt = d->dtcourant / 2.0;
        if (d->dthydro < gnewdt) gnewdt = d->dthydro * 2.0 / 3.0;
        newdt = gnewdt;
        ratio = newdt / olddt;
        if (ratio >= 1.0) {
            if (ratio < d->deltatimemultlb)
                newdt = olddt;
            else if (ratio > d->deltatimemultub)
                newdt = olddt * d->deltatimemultub;
        }
        if (newdt > d->dtmax) newdt = d->dtmax;
        d->deltatime = newdt;
    }
    if ((targetdt > d->deltatime) && (targetdt < (4.0 * d->deltatime / 3.0)))
        targetdt = 2.0 * d->deltatime / 3.0;
    if (targetdt < d->deltatime) d->deltatime = targetdt;
    d->time += d->deltatime;
    ++d->cycle;
}

/* ------------------------------------------------------------------------ */
/* Phase 1: corner forces                                                    */
/* ------------------------------------------------------------------------ */

/* InitIntegrateStressForElemsTask (lulesh.cc:787-836) and
 * CalcHourglassForElemsTask (lulesh.cc:864-1113) for all elements. */
int ora_calc_corner_forces(ora_domain *d, double *sx, double *sy, double *sz, double *hx,
                           double *hy, double *hz) {
    const Index_t E = d->numElem;
    int err = 0;
#pragma omp parallel for schedule(static) reduction(| : err)
    for (Index_t i = 0; i < E; ++i) {
        const Index_t *n = &d->nodelist[8 * (size_t)i];
        Real_t xl[8], yl[8], zl[8];
        Real_t B[3][8];
        Real_t determ;
        /* stress: sig = -p - q (lulesh.cc:798) */
        Real_t sig = -d->p[i] - d->q[i];
        collect_nodes(d->x, d->y, d->z, n, xl, yl, zl);
        shape_function_derivatives(xl, yl, zl, B, &determ);
        if (determ == 0) err |= 1; /* lulesh.cc:819-822 */
        node_normals(B, xl, yl, zl);
        for (int k = 0; k < 8; ++k) {
            d->fxs[8 * (size_t)i + k] = -(sig * B[0][k]);
            d->fys[8 * (size_t)i + k] = -(sig * B[1][k]);
            d->fzs[8 * (size_t)i + k] = -(sig * B[2][k]);
        }
    }
#pragma omp parallel for schedule(static) reduction(| : err)
    for (Index_t i = 0; i < E; ++i) {
        Real_t determ = d->volo[i] * d->v[i]; /* lulesh.cc:903 */
        if (determ <= 0.0) err |= 1;          /* lulesh.cc:906-910 */
    }
    if (err) return ORA_VOLUME_ERROR;
#pragma omp parallel for schedule(static)
    for (Index_t i = 0; i < E; ++i) {
        const Index_t *n = &d->nodelist[8 * (size_t)i];
        Real_t xl[8], yl[8], zl[8], xdl[8], ydl[8], zdl[8];
        Real_t hgfx[8], hgfy[8], hgfz[8];
        collect_nodes(d->x, d->y, d->z, n, xl, yl, zl);
        collect_nodes(d->xd, d->yd, d->zd, n, xdl, ydl, zdl);
        Real_t determ = d->volo[i] * d->v[i];
        elem_hourglass(xl, yl, zl, xdl, ydl, zdl, determ, d->ss[i], d->elemMass[i], d->hgcoef,
                       hgfx, hgfy, hgfz);
        for (int k = 0; k < 8; ++k) {
            d->fxh[8 * (size_t)i + k] = hgfx[k];
            d->fyh[8 * (size_t)i + k] = hgfy[k];
            d->fzh[8 * (size_t)i + k] = hgfz[k];
        }
    }
    if (sx) {
        size_t nb = sizeof(Real_t) * 8 * (size_t)E;
        memcpy(sx, d->fxs, nb);
        memcpy(sy, d->fys, nb);
        memcpy(sz, d->fzs, nb);
        memcpy(hx, d->fxh, nb);
        memcpy(hy, d->fyh, nb);
        memcpy(hz, d->fzh, nb);
    }
    return 0;
}

/* Phases 1-3, lulesh.cc:2017-2110 */
int ora_lagrange_nodal(ora_domain *d) {
    int rc = ora_calc_corner_forces(d, NULL, NULL, NULL, NULL, NULL, NULL);
    if (rc) return rc;
    const Index_t N = d->numNode;
    const Real_t dt = d->deltatime, u_cut = d->u_cut;
    /* combineVolumeForcesTaskFunc (838-862) + CalcAccelerationForNodesTask (1115-1123) */
#pragma omp parallel for schedule(static)
    for (Index_t g = 0; g < N; ++g) {
        Index_t count = d->nodeElemStart[g + 1] - d->nodeElemStart[g];
        const Index_t *cl = &d->nodeElemCornerList[d->nodeElemStart[g]];
        Real_t fx = 0.0, fy = 0.0, fz = 0.0;
        for (Index_t ii = 0; ii < count; ++ii) {
            Index_t c = cl[ii];
            fx += d->fxs[c];
            fx += d->fxh[c];
            fy += d->fys[c];
            fy += d->fyh[c];
            fz += d->fzs[c];
            fz += d->fzh[c];
        }
        d->fx[g] = fx;
        d->fy[g] = fy;
        d->fz[g] = fz;
        Real_t mass = d->nodalMass[g];
        d->xdd[g] = fx / mass;
        d->ydd[g] = fy / mass;
        d->zdd[g] = fz / mass;
    }
    /* ApplyAccelerationBoundaryConditionsForNodes, lulesh.cc:566-588 */
    Index_t numNodeBC = (d->sizeX + 1) * (d->sizeX + 1);
    for (Index_t i = 0; i < numNodeBC; ++i) d->xdd[d->symmX[i]] = 0.0;
    for (Index_t i = 0; i < numNodeBC; ++i) d->ydd[d->symmY[i]] = 0.0;
    for (Index_t i = 0; i < numNodeBC; ++i) d->zdd[d->symmZ[i]] = 0.0;
    /* CalcVelocityAndPositionForNodesTask, lulesh.cc:1125-1158 */
#pragma omp parallel for schedule(static)
    for (Index_t i = 0; i < N; ++i) {
        Real_t xdnew = d->xd[i] + d->xdd[i] * dt;
        if (fabs(xdnew) < u_cut) xdnew = 0.0;
        d->xd[i] = xdnew;
        Real_t ydnew = d->yd[i] + d->ydd[i] * dt;
        if (fabs(ydnew) < u_cut) ydnew = 0.0;
        d->yd[i] = ydnew;
        Real_t zdnew = d->zd[i] + d->zdd[i] * dt;
        if (fabs(zdnew) < u_cut) zdnew = 0.0;
        d->zd[i] = zdnew;
        d->x[i] = d->x[i] + d->xd[i] * dt;
        d->y[i] = d->y[i] + d->yd[i] * dt;
        d->z[i] = d->z[i] + d->zd[i] * dt;
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Phase 4: kinematics + Q gradients, lulesh.cc:1169-1418                    */
/* ------------------------------------------------------------------------ */
int ora_lagrange_elements(ora_domain *d) {
    const Index_t E = d->numElem;
    const Real_t deltaTime = d->deltatime;
    const Real_t v_cut = d->v_cut, eosvmin = d->eosvmin, eosvmax = d->eosvmax;
    int err = 0;
#pragma omp parallel for schedule(static) reduction(| : err)
    for (Index_t i = 0; i < E; ++i) {
        const Index_t *n = &d->nodelist[8 * (size_t)i];
        Real_t B[3][8], D[6];
        Real_t xl[8], yl[8], zl[8], xdl[8], ydl[8], zdl[8];
        Real_t detJ = 0.0;
        collect_nodes(d->x, d->y, d->z, n, xl, yl, zl);
        /* CalcKinematicsForElemsTask, lulesh.cc:1180-1233 */
        Real_t volume = elem_volume(xl, yl, zl);
        Real_t relativeVolume = volume / d->volo[i];
        Real_t vold = d->v[i];
        d->vnew[i] = relativeVolume;
        d->delv[i] = relativeVolume - vold;
        d->arealg[i] = characteristic_length(xl, yl, zl, volume);
        collect_nodes(d->xd, d->yd, d->zd, n, xdl, ydl, zdl);
        Real_t dt2 = 0.5 * deltaTime;
        for (int j = 0; j < 8; ++j) {
            xl[j] -= dt2 * xdl[j];
            yl[j] -= dt2 * ydl[j];
            zl[j] -= dt2 * zdl[j];
        }
        shape_function_derivatives(xl, yl, zl, B, &detJ);
        velocity_gradient(xdl, ydl, zdl, B, detJ, D);
        /* CalcLagrangeElements, lulesh.cc:1238-1247 (deviatoric parts are dead) */
        Real_t vdov_tmp = D[0] + D[1] + D[2];
        d->vdov[i] = vdov_tmp;
        /* CheckAndUpdateVolumeForElems, lulesh.cc:1255-1270 */
        Real_t vc = vold;
        if (eosvmin != 0.0 && vc < eosvmin) vc = eosvmin;
        if (eosvmax != 0.0 && vc > eosvmax) vc = eosvmax;
        if (vc < 0.0) err |= 1;
        Real_t vnew_tmp = relativeVolume;
        if (fabs(vnew_tmp - 1.0) < v_cut) vnew_tmp = 1.0;
        d->v[i] = vnew_tmp;
    }
    if (err) return ORA_VOLUME_ERROR;
    /* CalcMonotonicQGradientsForElemsTask, lulesh.cc:1277-1418 */
#pragma omp parallel for schedule(static)
    for (Index_t i = 0; i < E; ++i) {
        const Real_t ptiny = 1.e-36;
        const Index_t *n = &d->nodelist[8 * (size_t)i];
        Real_t ax, ay, az, dxv, dyv, dzv;
        Real_t x0 = d->x[n[0]], x1 = d->x[n[1]], x2 = d->x[n[2]], x3 = d->x[n[3]],
               x4 = d->x[n[4]], x5 = d->x[n[5]], x6 = d->x[n[6]], x7 = d->x[n[7]];
        Real_t y0 = d->y[n[0]], y1 = d->y[n[1]], y2 = d->y[n[2]], y3 = d->y[n[3]],
               y4 = d->y[n[4]], y5 = d->y[n[5]], y6 = d->y[n[6]], y7 = d->y[n[7]];
        Real_t z0 = d->z[n[0]], z1 = d->z[n[1]], z2 = d->z[n[2]], z3 = d->z[n[3]],
               z4 = d->z[n[4]], z5 = d->z[n[5]], z6 = d->z[n[6]], z7 = d->z[n[7]];
        Real_t xv0 = d->xd[n[0]], xv1 = d->xd[n[1]], xv2 = d->xd[n[2]], xv3 = d->xd[n[3]],
               xv4 = d->xd[n[4]], xv5 = d->xd[n[5]], xv6 = d->xd[n[6]], xv7 = d->xd[n[7]];
        Real_t yv0 = d->yd[n[0]], yv1 = d->yd[n[1]], yv2 = d->yd[n[2]], yv3 = d->yd[n[3]],
               yv4 = d->yd[n[4]], yv5 = d->yd[n[5]], yv6 = d->yd[n[6]], yv7 = d->yd[n[7]];
        Real_t zv0 = d->zd[n[0]], zv1 = d->zd[n[1]], zv2 = d->zd[n[2]], zv3 = d->zd[n[3]],
               zv4 = d->zd[n[4]], zv5 = d->zd[n[5]], zv6 = d->zd[n[6]], zv7 = d->zd[n[7]];
        Real_t vol = d->volo[i] * d->vnew[i];
        Real_t norm = 1.0 / (vol + ptiny);
        Real_t dxj = -0.25 * ((x0 + x1 + x5 + x4) - (x3 + x2 + x6 + x7));
        Real_t dyj = -0.25 * ((y0 + y1 + y5 + y4) - (y3 + y2 + y6 + y7));
        Real_t dzj = -0.25 * ((z0 + z1 + z5 + z4) - (z3 + z2 + z6 + z7));
        Real_t dxi = 0.25 * ((x1 + x2 + x6 + x5) - (x0 + x3 + x7 + x4));
        Real_t dyi = 0.25 * ((y1 + y2 + y6 + y5) - (y0 + y3 + y7 + y4));
        Real_t dzi = 0.25 * ((z1 + z2 + z6 + z5) - (z0 + z3 + z7 + z4));
        Real_t dxk = 0.25 * ((x4 + x5 + x6 + x7) - (x0 + x1 + x2 + x3));
        Real_t dyk = 0.25 * ((y4 + y5 + y6 + y7) - (y0 + y1 + y2 + y3));
        Real_t dzk = 0.25 * ((z4 + z5 + z6 + z7) - (z0 + z1 + z2 + z3));

        ax = dyi * dzj - dzi * dyj;
        ay = dzi * dxj - dxi * dzj;
        az = dxi * dyj - dyi * dxj;
        d->delx_zeta[i] = vol / sqrt(ax * ax + ay * ay + az * az + ptiny);
        ax *= norm;
        ay *= norm;
        az *= norm;
        dxv = 0.25 * ((xv4 + xv5 + xv6 + xv7) - (xv0 + xv1 + xv2 + xv3));
        dyv = 0.25 * ((yv4 + yv5 + yv6 + yv7) - (yv0 + yv1 + yv2 + yv3));
        dzv = 0.25 * ((zv4 + zv5 + zv6 + zv7) - (zv0 + zv1 + zv2 + zv3));
        d->delv_zeta[i] = ax * dxv + ay * dyv + az * dzv;

        ax = dyj * dzk - dzj * dyk;
        ay = dzj * dxk - dxj * dzk;
        az = dxj * dyk - dyj * dxk;
        d->delx_xi[i] = vol / sqrt(ax * ax + ay * ay + az * az + ptiny);
        ax *= norm;
        ay *= norm;
        az *= norm;
        dxv = 0.25 * ((xv1 + xv2 + xv6 + xv5) - (xv0 + xv3 + xv7 + xv4));
        dyv = 0.25 * ((yv1 + yv2 + yv6 + yv5) - (yv0 + yv3 + yv7 + yv4));
        dzv = 0.25 * ((zv1 + zv2 + zv6 + zv5) - (zv0 + zv3 + zv7 + zv4));
        d->delv_xi[i] = ax * dxv + ay * dyv + az * dzv;

        ax = dyk * dzi - dzk * dyi;
        ay = dzk * dxi - dxk * dzi;
        az = dxk * dyi - dyk * dxi;
        d->delx_eta[i] = vol / sqrt(ax * ax + ay * ay + az * az + ptiny);
        ax *= norm;
        ay *= norm;
        az *= norm;
        dxv = -0.25 * ((xv0 + xv1 + xv5 + xv4) - (xv3 + xv2 + xv6 + xv7));
        dyv = -0.25 * ((yv0 + yv1 + yv5 + yv4) - (yv3 + yv2 + yv6 + yv7));
        dzv = -0.25 * ((zv0 + zv1 + zv5 + zv4) - (zv3 + zv2 + zv6 + zv7));
        d->delv_eta[i] = ax * dxv + ay * dyv + az * dzv;
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Phase 5: monotonic Q per region + EOS + sound speed                       */
/* ------------------------------------------------------------------------ */

/* One direction of the monotonic limiter, lulesh.cc:1457-1508.  The switch
 * on (bcMask & *_M): *_COMM and 0 read the neighbour, *_SYMM the element
 * itself, *_FREE and any other bit pattern (the default branch) give 0. */
static inline Real_t bc_delv(const Real_t *delv, Index_t ielem, Int_t bits, Int_t comm,
                             Int_t symm, Int_t freeb, Index_t nbr) {
    if (bits == comm || bits == 0) return delv[nbr];
    if (bits == symm) return delv[ielem];
    (void)freeb;
    return 0.0;
}

static inline Real_t mono_phi(const Real_t *delv, Index_t ielem, Real_t delvm, Real_t delvp,
                              Real_t ptiny, Real_t limiter_mult, Real_t max_slope) {
    Real_t norm = 1. / (delv[ielem] + ptiny);
    delvm = delvm * norm;
    delvp = delvp * norm;
    Real_t phi = .5 * (delvm + delvp);
    delvm *= limiter_mult;
    delvp *= limiter_mult;
    if (delvm < phi) phi = delvm;
    if (delvp < phi) phi = delvp;
    if (phi < 0.) phi = 0.;
    if (phi > max_slope) phi = max_slope;
    return phi;
}

/* CalcMonotonicQRegionForElemsAndApplyInitTask body, lulesh.cc:1450-1651:
 * writes qq, ql of the element. */
static inline void monotonic_q_elem(ora_domain *d, Index_t ielem, Real_t ptiny) {
    Int_t bcMask = d->elemBC[ielem];
    const Real_t lm = d->monoq_limiter_mult, ms = d->monoq_max_slope;
    Real_t phixi = mono_phi(
        d->delv_xi, ielem,
        bc_delv(d->delv_xi, ielem, bcMask & XI_M, XI_M_COMM, XI_M_SYMM, XI_M_FREE, d->lxim[ielem]),
        bc_delv(d->delv_xi, ielem, bcMask & XI_P, XI_P_COMM, XI_P_SYMM, XI_P_FREE, d->lxip[ielem]),
        ptiny, lm, ms);
    Real_t phieta = mono_phi(
        d->delv_eta, ielem,
        bc_delv(d->delv_eta, ielem, bcMask & ETA_M, ETA_M_COMM, ETA_M_SYMM, ETA_M_FREE,
                d->letam[ielem]),
        bc_delv(d->delv_eta, ielem, bcMask & ETA_P, ETA_P_COMM, ETA_P_SYMM, ETA_P_FREE,
                d->letap[ielem]),
        ptiny, lm, ms);
    Real_t phizeta = mono_phi(
        d->delv_zeta, ielem,
        bc_delv(d->delv_zeta, ielem, bcMask & ZETA_M, ZETA_M_COMM, ZETA_M_SYMM, ZETA_M_FREE,
                d->lzetam[ielem]),
        bc_delv(d->delv_zeta, ielem, bcMask & ZETA_P, ZETA_P_COMM, ZETA_P_SYMM, ZETA_P_FREE,
                d->lzetap[ielem]),
        ptiny, lm, ms);
    Real_t qlin, qquad;
    if (d->vdov[ielem] > 0.) {
        qlin = 0.;
        qquad = 0.;
    } else {
        Real_t delvxxi = d->delv_xi[ielem] * d->delx_xi[ielem];
        Real_t delvxeta = d->delv_eta[ielem] * d->delx_eta[ielem];
        Real_t delvxzeta = d->delv_zeta[ielem] * d->delx_zeta[ielem];
        if (delvxxi > 0.) delvxxi = 0.;
        if (delvxeta > 0.) delvxeta = 0.;
        if (delvxzeta > 0.) delvxzeta = 0.;
        Real_t rho = d->elemMass[ielem] / (d->volo[ielem] * d->vnew[ielem]);
        qlin = -d->qlc_monoq * rho *
               (delvxxi * (1. - phixi) + delvxeta * (1. - phieta) + delvxzeta * (1. - phizeta));
        qquad = d->qqc_monoq * rho *
                (delvxxi * delvxxi * (1. - phixi * phixi) +
                 delvxeta * delvxeta * (1. - phieta * phieta) +
                 delvxzeta * delvxzeta * (1. - phizeta * phizeta));
    }
    d->qq[ielem] = qquad;
    d->ql[ielem] = qlin;
}

/* CalcPressureForElemsTask for one element, lulesh.cc:1671-1690 */
static inline Real_t pressure_elem(Real_t *bvc, Real_t *pbvc, Real_t e_old, Real_t compression,
                                   Real_t vnewc, Real_t pmin, Real_t p_cut, Real_t eosvmax) {
    const Real_t cls = 2.0 / 3.0;
    *bvc = cls * (compression + 1.0);
    *pbvc = cls;
    Real_t newval = *bvc * e_old;
    if (fabs(newval) < p_cut || vnewc >= eosvmax) newval = 0.0;
    if (newval < pmin) newval = pmin;
    return newval;
}

/* EvalEOSAllInOneTask for one element (lulesh.cc:1692-1871) followed by the
 * sound speed of CalcSoundSpeedForElemsAndSaveTask (1885-1896).
 * in: e_old, delvc, p_old, q_old, qq_old, ql_old, vnewc (already clamped). */
static inline void eos_elem(Real_t e_old, Real_t delvc, Real_t p_old, Real_t q_old,
                            Real_t qq_old, Real_t ql_old, Real_t vnewc, Real_t emin, Real_t pmin,
                            Real_t p_cut, Real_t rho0, Real_t e_cut, Real_t q_cut,
                            Real_t eosvmin, Real_t eosvmax, Real_t *pe, Real_t *pp, Real_t *pq,
                            Real_t *pss) {
    Real_t compression, compHalfStep, work, e_new, p_new, q_new, bvc, pbvc, pHalfStep;
    /* EvalEOSInit, lulesh.cc:1720-1749 */
    compression = 1. / vnewc - 1.;
    Real_t vchalf = vnewc - delvc * .5;
    compHalfStep = 1. / vchalf - 1.;
    work = 0;
    if (eosvmin != 0.) {
        if (vnewc <= eosvmin) compHalfStep = compression;
    }
    if (eosvmax != 0.0) {
        if (vnewc >= eosvmax) {
            p_old = 0.;
            compression = 0.;
            compHalfStep = 0.;
        }
    }
    /* CalcEnergyForElemsInit, lulesh.cc:1754-1761 */
    e_new = e_old - 0.5 * delvc * (p_old + q_old) + 0.5 * work;
    if (e_new < emin) e_new = emin;
    pHalfStep = pressure_elem(&bvc, &pbvc, e_new, compHalfStep, vnewc, pmin, p_cut, eosvmax);
    /* lulesh.cc:1771-1791 */
    {
        Real_t vhalf = 1. / (1. + compHalfStep);
        if (delvc > 0.) {
            q_new = 0.;
        } else {
            Real_t ssc = (pbvc * e_new + vhalf * vhalf * bvc * pHalfStep) / rho0;
            if (ssc <= .1111111e-36)
                ssc = .3333333e-18;
            else
                ssc = sqrt(ssc);
            q_new = (ssc * ql_old + qq_old);
        }
        e_new = e_new + 0.5 * delvc * (3.0 * (p_old + q_old) - 4.0 * (pHalfStep + q_new));
    }
    /* lulesh.cc:1793-1803 */
    {
        Real_t newval = e_new + 0.5 * work;
        if (fabs(newval) < e_cut) newval = 0.0;
        if (newval < emin) newval = emin;
        e_new = newval;
    }
    p_new = pressure_elem(&bvc, &pbvc, e_new, compression, vnewc, pmin, p_cut, eosvmax);
    /* lulesh.cc:1812-1842 */
    {
        const Real_t sixth = 1.0 / 6.0;
        Real_t q_tilde;
        if (delvc > 0.) {
            q_tilde = 0.;
        } else {
            Real_t ssc = (pbvc * e_new + vnewc * vnewc * bvc * p_new) / rho0;
            if (ssc <= .1111111e-36)
                ssc = .3333333e-18;
            else
                ssc = sqrt(ssc);
            q_tilde = (ssc * ql_old + qq_old);
        }
        e_new = e_new - (7.0 * (p_old + q_old) - 8.0 * (pHalfStep + q_new) + (p_new + q_tilde)) *
                            delvc * sixth;
        if (fabs(e_new) < e_cut) e_new = 0.;
        if (e_new < emin) e_new = emin;
    }
    p_new = pressure_elem(&bvc, &pbvc, e_new, compression, vnewc, pmin, p_cut, eosvmax);
    /* lulesh.cc:1851-1869 */
    if (delvc <= 0.) {
        Real_t ssc = (pbvc * e_new + vnewc * vnewc * bvc * p_new) / rho0;
        if (ssc <= .1111111e-36)
            ssc = .3333333e-18;
        else
            ssc = sqrt(ssc);
        q_new = (ssc * ql_old + qq_old);
        if (fabs(q_new) < q_cut) q_new = 0.;
    }
    /* CalcSoundSpeedForElemsAndSaveTask, lulesh.cc:1885-1896 */
    Real_t ssTmp = (pbvc * e_new + vnewc * vnewc * bvc * p_new) / rho0;
    if (ssTmp <= .1111111e-36)
        ssTmp = .3333333e-18;
    else
        ssTmp = sqrt(ssTmp);
    *pe = e_new;
    *pp = p_new;
    *pq = q_new;
    *pss = ssTmp;
}

/* Region repetition count, lulesh.cc:2159-2169 */
static inline Int_t region_rep(Int_t reg, Int_t numReg, Int_t cost) {
    if (reg < numReg / 2) return 1;
    if (reg < (numReg - (numReg + 15) / 20)) return 1 + cost;
    return 10 * (1 + cost);
}

int32_t ora_reg_rep(ora_domain *d, int32_t r) { return region_rep(r, d->numReg, d->cost); }

/* ApplyMaterialPropertiesForElems phase, lulesh.cc:2138-2202.  The rep loop
 * re-reads the Domain each time (lulesh.cc:1720-1727) and the Domain is only
 * written after the last repetition, so every repetition computes the same
 * values; the oracle evaluates it rep times all the same, as the reference. */
int ora_apply_material_properties(ora_domain *d) {
    const Real_t ptiny = 1.e-36;
    const Real_t eosvmin = d->eosvmin, eosvmax = d->eosvmax;
    for (Int_t reg = 0; reg < d->numReg; ++reg) {
        const Index_t n = d->regElemSize[reg];
        const Index_t *list = d->regElemlist[reg];
        const Int_t rep = region_rep(reg, d->numReg, d->cost);
#pragma omp parallel for schedule(static)
        for (Index_t i = 0; i < n; ++i) {
            Index_t ielem = list[i];
            monotonic_q_elem(d, ielem, ptiny);
            /* vnewc clamp, lulesh.cc:1654-1666 */
            Real_t vnewc = d->vnew[ielem];
            if (eosvmin != 0.0) {
                if (vnewc < eosvmin) vnewc = eosvmin;
            }
            if (eosvmax != 0.0) {
                if (vnewc > eosvmax) vnewc = eosvmax;
            }
            Real_t e_new = 0, p_new = 0, q_new = 0, ss_new = 0;
            for (Int_t r = 0; r < rep; ++r) {
                volatile Real_t e_in = d->e[ielem]; /* re-read per repetition */
                eos_elem(e_in, d->delv[ielem], d->p[ielem], d->q[ielem], d->qq[ielem],
                         d->ql[ielem], vnewc, d->emin, d->pmin, d->p_cut, d->refdens, d->e_cut,
                         d->q_cut, eosvmin, eosvmax, &e_new, &p_new, &q_new, &ss_new);
            }
            d->ss[ielem] = ss_new;
            d->p[ielem] = p_new;
            d->e[ielem] = e_new;
            d->q[ielem] = q_new;
        }
    }
    return 0;
}

void ora_eval_eos_elem(const double in[7], double out[4]) {
    /* constants of lulesh-init.cc:18-36 */
    eos_elem(in[0], in[1], in[2], in[3], in[4], in[5], in[6], -1.0e+15, 0., 1.0e-7, 1.0, 1.0e-7,
             1.0e-7, 1.0e-9, 1.0e+9, &out[0], &out[1], &out[2], &out[3]);
}

/* ------------------------------------------------------------------------ */
/* Phase 6: time constraints, lulesh.cc:1938-1974, 2204-2239                 */
/* ------------------------------------------------------------------------ */
void ora_calc_time_constraints(ora_domain *d) {
    Real_t dtcourant = 1.0e+20, dthydro = 1.0e+20;
    const Real_t qqc2 = 64.0 * d->qqc * d->qqc;
    const Real_t dvovmax = d->dvovmax;
    const Index_t E = d->numElem;
    /* Region lists partition the elements and min() is exact, so the
     * minimum over regions equals the minimum over all elements. */
#pragma omp parallel for schedule(static) reduction(min : dtcourant, dthydro)
    for (Index_t indx = 0; indx < E; ++indx) {
        Real_t vd = d->vdov[indx];
        Real_t c, h;
        if (vd == 0.0) {
            c = DBL_MAX;
            h = DBL_MAX;
        } else {
            Real_t dtf = d->ss[indx] * d->ss[indx];
            if (vd < 0.0) dtf += qqc2 * d->arealg[indx] * d->arealg[indx] * vd * vd;
            dtf = sqrt(dtf);
            dtf = d->arealg[indx] / dtf;
            c = dtf;
            h = dvovmax / (fabs(vd) + 1.e-20);
        }
        if (c < dtcourant) dtcourant = c;
        if (h < dthydro) dthydro = h;
    }
    d->dtcourant = dtcourant;
    d->dthydro = dthydro;
}

int ora_lagrange_leapfrog(ora_domain *d) {
    int rc = ora_lagrange_nodal(d);
    if (rc) return rc;
    rc = ora_lagrange_elements(d);
    if (rc) return rc;
    rc = ora_apply_material_properties(d);
    if (rc) return rc;
    ora_calc_time_constraints(d);
    return 0;
}

int ora_run(ora_domain *d, int32_t its) {
    while ((d->time < d->stoptime) && (d->cycle < its)) {
        ora_time_increment(d);
        int rc = ora_lagrange_leapfrog(d);
        if (rc) return rc;
    }
    return 0;
}

/* VerifyAndWriteFinalOutput, lulesh-util.cc:60-97 */
void ora_report(ora_domain *d, double out[4]) {
    Index_t nx = d->sizeX;
    Real_t MaxAbsDiff = 0.0, TotalAbsDiff = 0.0, MaxRelDiff = 0.0;
    for (Index_t j = 0; j < nx; ++j) {
        for (Index_t k = j + 1; k < nx; ++k) {
            Real_t AbsDiff = fabs(d->e[j * nx + k] - d->e[k * nx + j]);
            TotalAbsDiff += AbsDiff;
            if (MaxAbsDiff < AbsDiff) MaxAbsDiff = AbsDiff;
            Real_t RelDiff = fabs(d->e[k * nx + j]) > 1e-8 ? AbsDiff / d->e[k * nx + j] : 0.0;
            if (MaxRelDiff < RelDiff) MaxRelDiff = RelDiff;
        }
    }
    out[0] = d->e[0];
    out[1] = MaxAbsDiff;
    out[2] = TotalAbsDiff;
    out[3] = MaxRelDiff;
}

/* ------------------------------------------------------------------------ */
/* Accessors                                                                 */
/* ------------------------------------------------------------------------ */
double *ora_field_f64(ora_domain *d, const char *name, int64_t *len) {
    const int64_t E = d->numElem, N = d->numNode;
    struct {
        const char *n;
        Real_t *p;
        int64_t len;
    } tab[] = {{"x", d->x, N},           {"y", d->y, N},
               {"z", d->z, N},           {"xd", d->xd, N},
               {"yd", d->yd, N},         {"zd", d->zd, N},
               {"xdd", d->xdd, N},       {"ydd", d->ydd, N},
               {"zdd", d->zdd, N},       {"fx", d->fx, N},
               {"fy", d->fy, N},         {"fz", d->fz, N},
               {"nodalMass", d->nodalMass, N}, {"e", d->e, E},
               {"p", d->p, E},           {"q", d->q, E},
               {"ql", d->ql, E},         {"qq", d->qq, E},
               {"v", d->v, E},           {"volo", d->volo, E},
               {"delv", d->delv, E},     {"vdov", d->vdov, E},
               {"arealg", d->arealg, E}, {"ss", d->ss, E},
               {"elemMass", d->elemMass, E}, {"vnew", d->vnew, E},
               {"delv_xi", d->delv_xi, E}, {"delv_eta", d->delv_eta, E},
               {"delv_zeta", d->delv_zeta, E}, {"delx_xi", d->delx_xi, E},
               {"delx_eta", d->delx_eta, E}, {"delx_zeta", d->delx_zeta, E}};
    for (size_t i = 0; i < sizeof(tab) / sizeof(tab[0]); ++i) {
        if (strcmp(tab[i].n, name) == 0) {
            if (len) *len = tab[i].len;
            return tab[i].p;
        }
    }
    return NULL;
}

int32_t *ora_field_i32(ora_domain *d, const char *name, int64_t *len) {
    const int64_t E = d->numElem, N = d->numNode;
    const int64_t P = (int64_t)(d->sizeX + 1) * (d->sizeX + 1);
    struct {
        const char *n;
        Index_t *p;
        int64_t len;
    } tab[] = {{"nodelist", d->nodelist, 8 * E},
               {"lxim", d->lxim, E},
               {"lxip", d->lxip, E},
               {"letam", d->letam, E},
               {"letap", d->letap, E},
               {"lzetam", d->lzetam, E},
               {"lzetap", d->lzetap, E},
               {"elemBC", d->elemBC, E},
               {"regNumList", d->regNumList, E},
               {"symmX", d->symmX, P},
               {"symmY", d->symmY, P},
               {"symmZ", d->symmZ, P},
               {"nodeElemStart", d->nodeElemStart, N + 1},
               {"nodeElemCornerList", d->nodeElemCornerList, 8 * E}};
    for (size_t i = 0; i < sizeof(tab) / sizeof(tab[0]); ++i) {
        if (strcmp(tab[i].n, name) == 0) {
            if (len) *len = tab[i].len;
            return tab[i].p;
        }
    }
    return NULL;
}

static Real_t *scalar_ptr(ora_domain *d, const char *name) {
    struct {
        const char *n;
        Real_t *p;
    } tab[] = {{"time", &d->time},
               {"deltatime", &d->deltatime},
               {"dtcourant", &d->dtcourant},
               {"dthydro", &d->dthydro},
               {"stoptime", &d->stoptime},
               {"dtmax", &d->dtmax},
               {"dtfixed", &d->dtfixed},
               {"deltatimemultlb", &d->deltatimemultlb},
               {"deltatimemultub", &d->deltatimemultub},
               {"hgcoef", &d->hgcoef},
               {"u_cut", &d->u_cut},
               {"e_cut", &d->e_cut},
               {"p_cut", &d->p_cut},
               {"q_cut", &d->q_cut},
               {"v_cut", &d->v_cut},
               {"qqc", &d->qqc},
               {"dvovmax", &d->dvovmax},
               {"eosvmin", &d->eosvmin},
               {"eosvmax", &d->eosvmax},
               {"emin", &d->emin},
               {"pmin", &d->pmin},
               {"refdens", &d->refdens},
               {"monoq_limiter_mult", &d->monoq_limiter_mult},
               {"monoq_max_slope", &d->monoq_max_slope},
               {"qlc_monoq", &d->qlc_monoq},
               {"qqc_monoq", &d->qqc_monoq},
               {"ss4o3", &d->ss4o3},
               {"qstop", &d->qstop}};
    for (size_t i = 0; i < sizeof(tab) / sizeof(tab[0]); ++i)
        if (strcmp(tab[i].n, name) == 0) return tab[i].p;
    return NULL;
}

double ora_get_scalar(ora_domain *d, const char *name) {
    Real_t *p = scalar_ptr(d, name);
    return p ? *p : NAN;
}

void ora_set_scalar(ora_domain *d, const char *name, double value) {
    Real_t *p = scalar_ptr(d, name);
    if (p) *p = value;
}

int32_t ora_cycle(ora_domain *d) { return d->cycle; }
void ora_set_cycle(ora_domain *d, int32_t c) { d->cycle = c; }
int32_t ora_num_reg(ora_domain *d) { return d->numReg; }
int32_t ora_reg_elem_size(ora_domain *d, int32_t r) { return d->regElemSize[r]; }
int32_t *ora_reg_elem_list(ora_domain *d, int32_t r) { return d->regElemlist[r]; }

/* Exported element helpers (unit tests) */
double ora_calc_elem_volume(const double x[8], const double y[8], const double z[8]) {
    return elem_volume(x, y, z);
}
void ora_calc_elem_shape_function_derivatives(const double x[8], const double y[8],
                                              const double z[8], double b[24], double *volume) {
    shape_function_derivatives(x, y, z, (Real_t(*)[8])b, volume);
}
void ora_calc_elem_node_normals(double pf[24], const double x[8], const double y[8],
                                const double z[8]) {
    node_normals((Real_t(*)[8])pf, x, y, z);
}
void ora_calc_elem_volume_derivative(double dv[24], const double x[8], const double y[8],
                                     const double z[8]) {
    volume_derivative(dv, dv + 8, dv + 16, x, y, z);
}
double ora_calc_elem_characteristic_length(const double x[8], const double y[8],
                                           const double z[8], double volume) {
    return characteristic_length(x, y, z, volume);
}
void ora_calc_elem_velocity_gradient(const double xv[8], const double yv[8], const double zv[8],
                                     const double b[24], double detJ, double d[6]) {
    velocity_gradient(xv, yv, zv, (const Real_t(*)[8])b, detJ, d);
}
void ora_calc_elem_fb_hourglass(const double x[8], const double y[8], const double z[8],
                                const double xd[8], const double yd[8], const double zd[8],
                                double determ, double ss, double mass, double hgcoef,
                                double hg[24]) {
    elem_hourglass(x, y, z, xd, yd, zd, determ, ss, mass, hgcoef, hg, hg + 8, hg + 16);
}
