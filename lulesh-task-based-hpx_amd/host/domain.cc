// Sedov problem setup on the host: the Domain constructor.
//
// Follows the reference constructor (lulesh-init.cc:10-183) for a single
// rank.  Every value must match the reference bit for bit, so the formulas
// that produce floating-point data (mesh coordinates, volumes, nodal masses,
// the initial energy and time step) keep the reference's expression shapes,
// and the region assignment replays the same libc rand() stream.
#include "domain.hpp"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <stdexcept>

namespace lulesh {

Real_t CalcElemVolume(const Real_t x[8], const Real_t y[8], const Real_t z[8]) {
    // Sum of three triple products of edge/diagonal differences,
    // lulesh.cc:592-674.
    const Real_t twelveth = Real_t(1.0) / Real_t(12.0);
    auto d = [](Real_t a, Real_t b) { return a - b; };
    const Real_t dx61 = d(x[6], x[1]), dy61 = d(y[6], y[1]), dz61 = d(z[6], z[1]);
    const Real_t dx70 = d(x[7], x[0]), dy70 = d(y[7], y[0]), dz70 = d(z[7], z[0]);
    const Real_t dx63 = d(x[6], x[3]), dy63 = d(y[6], y[3]), dz63 = d(z[6], z[3]);
    const Real_t dx20 = d(x[2], x[0]), dy20 = d(y[2], y[0]), dz20 = d(z[2], z[0]);
    const Real_t dx50 = d(x[5], x[0]), dy50 = d(y[5], y[0]), dz50 = d(z[5], z[0]);
    const Real_t dx64 = d(x[6], x[4]), dy64 = d(y[6], y[4]), dz64 = d(z[6], z[4]);
    const Real_t dx31 = d(x[3], x[1]), dy31 = d(y[3], y[1]), dz31 = d(z[3], z[1]);
    const Real_t dx72 = d(x[7], x[2]), dy72 = d(y[7], y[2]), dz72 = d(z[7], z[2]);
    const Real_t dx43 = d(x[4], x[3]), dy43 = d(y[4], y[3]), dz43 = d(z[4], z[3]);
    const Real_t dx57 = d(x[5], x[7]), dy57 = d(y[5], y[7]), dz57 = d(z[5], z[7]);
    const Real_t dx14 = d(x[1], x[4]), dy14 = d(y[1], y[4]), dz14 = d(z[1], z[4]);
    const Real_t dx25 = d(x[2], x[5]), dy25 = d(y[2], y[5]), dz25 = d(z[2], z[5]);
    // Triple product in the argument order and grouping of the reference's
    // TRIPLE_PRODUCT macro (lulesh.cc:649-651); the grouping fixes rounding.
    auto triple = [](Real_t x1, Real_t y1, Real_t z1, Real_t x2, Real_t y2, Real_t z2,
                     Real_t x3, Real_t y3, Real_t z3) {
        return x1 * (y2 * z3 - z2 * y3) + x2 * (z1 * y3 - y1 * z3) + x3 * (y1 * z2 - z1 * y2);
    };
    Real_t volume = triple(dx31 + dx72, dx63, dx20, dy31 + dy72, dy63, dy20, dz31 + dz72, dz63,
                           dz20) +
                    triple(dx43 + dx57, dx64, dx70, dy43 + dy57, dy64, dy70, dz43 + dz57, dz64,
                           dz70) +
                    triple(dx14 + dx25, dx61, dx50, dy14 + dy25, dy61, dy50, dz14 + dz25, dz61,
                           dz50);
    volume *= twelveth;
    return volume;
}

Domain::Domain(Index_t nx, Int_t nr, Int_t balance, Int_t cost) {
    if (nx < 1 || nr < 1) throw std::invalid_argument("problem size and regions must be >= 1");
    const Index_t edgeElems = nx, edgeNodes = nx + 1;
    m_cost = cost;
    m_sizeX = edgeElems;
    m_numElem = edgeElems * edgeElems * edgeElems;
    m_numNode = edgeNodes * edgeNodes * edgeNodes;
    const size_t E = size_t(m_numElem), N = size_t(m_numNode);

    // Persistent allocations (lulesh.h:146-201) zero-filled; v starts at 1
    // (lulesh-init.cc:81-107).
    for (auto *f : {&m_x, &m_y, &m_z, &m_xd, &m_yd, &m_zd, &m_xdd, &m_ydd, &m_zdd, &m_fx, &m_fy,
                    &m_fz, &m_nodalMass})
        f->assign(N, 0.0);
    for (auto *f : {&m_e, &m_p, &m_q, &m_ql, &m_qq, &m_volo, &m_vnew, &m_delv, &m_vdov,
                    &m_arealg, &m_ss, &m_elemMass, &m_delv_xi, &m_delv_eta, &m_delv_zeta,
                    &m_delx_xi, &m_delx_eta, &m_delx_zeta})
        f->assign(E, 0.0);
    m_v.assign(E, 1.0);
    m_nodelist.assign(8 * E, 0);
    for (auto *f : {&m_lxim, &m_lxip, &m_letam, &m_letap, &m_lzetam, &m_lzetap, &m_elemBC,
                    &m_regNumList})
        f->assign(E, 0);
    const size_t P = size_t(edgeNodes) * edgeNodes;
    m_symmX.assign(P, 0);
    m_symmY.assign(P, 0);
    m_symmZ.assign(P, 0);

    BuildMesh(nx);
    SetupThreadSupportStructures();
    CreateRegionIndexSets(nr, balance);
    SetupSymmetryPlanes(edgeNodes);
    SetupElementConnectivities(edgeElems);
    SetupBoundaryConditions(edgeElems);

    // Run defaults, lulesh-init.cc:135-145
    m_dtfixed = -1.0e-6;
    m_stoptime = 1.0e-2;
    m_deltatimemultlb = 1.1;
    m_deltatimemultub = 1.2;
    m_dtcourant = 1.0e+20;
    m_dthydro = 1.0e+20;
    m_dtmax = 1.0e-2;
    m_time = 0.;
    m_cycle = 0;

    // Reference volumes and masses, lulesh-init.cc:148-167: the nodal mass is
    // accumulated element by element, corner by corner (this order fixes its
    // rounding).
    for (Index_t i = 0; i < m_numElem; ++i) {
        Real_t xl[8], yl[8], zl[8];
        const Index_t *n = nodelist(i);
        for (int k = 0; k < 8; ++k) {
            xl[k] = m_x[n[k]];
            yl[k] = m_y[n[k]];
            zl[k] = m_z[n[k]];
        }
        const Real_t volume = CalcElemVolume(xl, yl, zl);
        m_volo[i] = volume;
        m_elemMass[i] = volume;
        for (int k = 0; k < 8; ++k) m_nodalMass[n[k]] += volume / Real_t(8.0);
    }

    // Energy deposit scaled by (nx/45)^3 into element 0, lulesh-init.cc:172-179
    const Real_t einit = SedovInitialEnergy(nx);
    m_e[0] = einit;
    // Initial time step from the CFL estimate, lulesh-init.cc:181
    m_deltatime = (Real_t(.5) * std::cbrt(m_volo[0])) / std::sqrt(Real_t(2.0) * einit);
}

Real_t SedovInitialEnergy(Index_t nx) {
    const Real_t ebase = 3.948746e+7;
    const Real_t scale = nx / Real_t(45.0);
    return ebase * scale * scale * scale;
}

Real_t SedovInitialDeltaTime(Index_t nx, Real_t einit) {
    // volo of element 0 from its lattice nodes, exactly as the constructor
    // computes it
    const Index_t n[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0},
                             {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
    Real_t x[8], y[8], z[8];
    for (int k = 0; k < 8; ++k) {
        x[k] = SedovCoord(n[k][0], nx);
        y[k] = SedovCoord(n[k][1], nx);
        z[k] = SedovCoord(n[k][2], nx);
    }
    const Real_t volo0 = CalcElemVolume(x, y, z);
    return (Real_t(.5) * std::cbrt(volo0)) / std::sqrt(Real_t(2.0) * einit);
}

void Domain::BuildMesh(Int_t nx) {
    // Lattice over [0, 1.125]^3; every coordinate is computed from its index
    // (lulesh-init.cc:207-227), never accumulated.
    const Index_t edgeNodes = nx + 1;
    auto coord = [nx](Index_t i) { return SedovCoord(i, nx); };
    Index_t nidx = 0;
    for (Index_t plane = 0; plane < edgeNodes; ++plane)
        for (Index_t row = 0; row < edgeNodes; ++row)
            for (Index_t col = 0; col < edgeNodes; ++col, ++nidx) {
                m_x[nidx] = coord(col);
                m_y[nidx] = coord(row);
                m_z[nidx] = coord(plane);
            }
    // Hex connectivity in the reference's local node order
    // (lulesh-init.cc:236-244): bottom face 0-1-2-3 counter-clockwise in
    // (col,row), top face 4-7 above it.
    const Index_t planeStride = edgeNodes * edgeNodes;
    Index_t zidx = 0;
    for (Index_t plane = 0; plane < nx; ++plane)
        for (Index_t row = 0; row < nx; ++row)
            for (Index_t col = 0; col < nx; ++col, ++zidx) {
                const Index_t base = plane * planeStride + row * edgeNodes + col;
                Index_t *ln = nodelist(zidx);
                ln[0] = base;
                ln[1] = base + 1;
                ln[2] = base + edgeNodes + 1;
                ln[3] = base + edgeNodes;
                ln[4] = base + planeStride;
                ln[5] = base + planeStride + 1;
                ln[6] = base + planeStride + edgeNodes + 1;
                ln[7] = base + planeStride + edgeNodes;
            }
}

void Domain::SetupThreadSupportStructures() {
    // Node -> corner CSR (lulesh-init.cc:256-310): corners listed in ascending
    // element order, which fixes the summation order of the node force gather.
    const Index_t N = m_numNode, E = m_numElem;
    m_nodeElemStart.assign(size_t(N) + 1, 0);
    for (Index_t i = 0; i < 8 * E; ++i) ++m_nodeElemStart[m_nodelist[i] + 1];
    for (Index_t n = 0; n < N; ++n) m_nodeElemStart[n + 1] += m_nodeElemStart[n];
    m_nodeElemCornerList.assign(size_t(m_nodeElemStart[N]), 0);
    std::vector<Index_t> fill(m_nodeElemStart.begin(), m_nodeElemStart.end() - 1);
    for (Index_t c = 0; c < 8 * E; ++c) m_nodeElemCornerList[fill[m_nodelist[c]]++] = c;
    for (Index_t c : m_nodeElemCornerList)
        if (c < 0 || c > 8 * E) {
            std::fprintf(stderr, "nodeElemCornerList entry out of range\n");
            std::exit(-1);
        }
}

void SedovRegionRuns(Index_t E, Int_t nr, Int_t balance, std::vector<RegionRun> &runs) {
    // Random region runs (lulesh-init.cc:341-445).  The reference draws from
    // libc rand() after srand(0); the same calls in the same order reproduce
    // its region map exactly.
    runs.clear();
    std::srand(0);
    if (nr == 1) {
        runs.push_back({0, 1});
        return;
    }
    // Cumulative weights (i+1)^balance; the running sum is kept in an int,
    // as in the reference.
    std::vector<Int_t> binEnd(static_cast<size_t>(nr));
    Int_t denom = 0;
    for (Int_t i = 0; i < nr; ++i) {
        denom = Int_t(double(denom) + std::pow(double(i + 1), double(balance)));
        binEnd[size_t(i)] = denom;
    }
    auto pick = [&]() {
        const Int_t r = std::rand() % denom;
        Int_t i = 0;
        while (r >= binEnd[size_t(i)]) ++i;
        return (i % nr) + 1;
    };
    Int_t lastReg = -1;
    Index_t next = 0;
    while (next < E) {
        Int_t regionNum = pick();
        while (regionNum == lastReg) regionNum = pick();
        // Run length: a skewed distribution of bin sizes.
        const Int_t bin = std::rand() % 1000;
        Index_t run;
        if (bin < 773) run = std::rand() % 15 + 1;
        else if (bin < 937) run = std::rand() % 16 + 16;
        else if (bin < 970) run = std::rand() % 32 + 32;
        else if (bin < 974) run = std::rand() % 64 + 64;
        else if (bin < 978) run = std::rand() % 128 + 128;
        else if (bin < 981) run = std::rand() % 256 + 256;
        else run = std::rand() % 1537 + 512;
        runs.push_back({next, regionNum});
        next += run;
        lastReg = regionNum;
    }
}

void Domain::CreateRegionIndexSets(Int_t nr, Int_t balance) {
    m_numReg = nr;
    m_regElemSize.assign(size_t(nr), 0);
    const Index_t E = m_numElem;
    std::vector<RegionRun> runs;
    SedovRegionRuns(E, nr, balance, runs);
    for (size_t r = 0; r < runs.size(); ++r) {
        const Index_t end = r + 1 < runs.size() ? runs[r + 1].start : E;
        for (Index_t i = runs[r].start; i < end && i < E; ++i) m_regNumList[i] = runs[r].region;
    }
    // Region index sets in ascending element order.
    for (Index_t i = 0; i < E; ++i) ++m_regElemSize[size_t(m_regNumList[i] - 1)];
    m_regElemlist.assign(size_t(nr), {});
    for (Int_t r = 0; r < nr; ++r) m_regElemlist[size_t(r)].reserve(size_t(m_regElemSize[r]));
    for (Index_t i = 0; i < E; ++i) m_regElemlist[size_t(m_regNumList[i] - 1)].push_back(i);
    m_regElemlistPtr.resize(size_t(nr));
    for (Int_t r = 0; r < nr; ++r) m_regElemlistPtr[size_t(r)] = m_regElemlist[size_t(r)].data();
}

void Domain::SetupSymmetryPlanes(Int_t edgeNodes) {
    // Nodes on the x = 0, y = 0 and z = 0 planes (lulesh-init.cc:448-468).
    Index_t k = 0;
    for (Index_t a = 0; a < edgeNodes; ++a)
        for (Index_t b = 0; b < edgeNodes; ++b, ++k) {
            m_symmZ[k] = a * edgeNodes + b;
            m_symmY[k] = a * edgeNodes * edgeNodes + b;
            m_symmX[k] = a * edgeNodes * edgeNodes + b * edgeNodes;
        }
}

void Domain::SetupElementConnectivities(Int_t edgeElems) {
    // Face neighbours (lulesh-init.cc:473-500).  At the mesh boundary the
    // reference points them at the element itself (low faces) or at the
    // wrapped index (xi faces inside a plane); the BC flags override them.
    const Index_t E = m_numElem, row = edgeElems, plane = edgeElems * edgeElems;
    for (Index_t i = 0; i < E; ++i) {
        m_lxim[i] = (i == 0) ? 0 : i - 1;
        m_lxip[i] = (i == E - 1) ? E - 1 : i + 1;
        m_letam[i] = (i < row) ? i : i - row;
        m_letap[i] = (i >= E - row) ? i : i + row;
        m_lzetam[i] = (i < plane) ? i : i - plane;
        m_lzetap[i] = (i >= E - plane) ? i : i + plane;
    }
}

void Domain::SetupBoundaryConditions(Int_t edgeElems) {
    // Single rank: symmetry on the three low faces, free surface on the three
    // high faces (lulesh-init.cc:548-607; the *_COMM branches need a
    // neighbouring rank).
    const Index_t n = edgeElems;
    for (Index_t k = 0; k < n; ++k)
        for (Index_t j = 0; j < n; ++j)
            for (Index_t i = 0; i < n; ++i) {
                Int_t bc = 0;
                if (k == 0) bc |= ZETA_M_SYMM;
                if (k == n - 1) bc |= ZETA_P_FREE;
                if (j == 0) bc |= ETA_M_SYMM;
                if (j == n - 1) bc |= ETA_P_FREE;
                if (i == 0) bc |= XI_M_SYMM;
                if (i == n - 1) bc |= XI_P_FREE;
                m_elemBC[(k * n + j) * n + i] = bc;
            }
}

lulesh_domain_view Domain::view() {
    lulesh_domain_view v{};
    v.sizeX = m_sizeX;
    v.numElem = m_numElem;
    v.numNode = m_numNode;
    v.numReg = m_numReg;
    v.cost = m_cost;
    v.x = m_x.data(); v.y = m_y.data(); v.z = m_z.data();
    v.xd = m_xd.data(); v.yd = m_yd.data(); v.zd = m_zd.data();
    v.xdd = m_xdd.data(); v.ydd = m_ydd.data(); v.zdd = m_zdd.data();
    v.fx = m_fx.data(); v.fy = m_fy.data(); v.fz = m_fz.data();
    v.nodalMass = m_nodalMass.data();
    v.symmX = m_symmX.data(); v.symmY = m_symmY.data(); v.symmZ = m_symmZ.data();
    v.nodelist = m_nodelist.data();
    v.lxim = m_lxim.data(); v.lxip = m_lxip.data();
    v.letam = m_letam.data(); v.letap = m_letap.data();
    v.lzetam = m_lzetam.data(); v.lzetap = m_lzetap.data();
    v.elemBC = m_elemBC.data();
    v.e = m_e.data(); v.p = m_p.data(); v.q = m_q.data(); v.ql = m_ql.data(); v.qq = m_qq.data();
    v.v = m_v.data(); v.volo = m_volo.data(); v.delv = m_delv.data(); v.vdov = m_vdov.data();
    v.arealg = m_arealg.data(); v.ss = m_ss.data(); v.elemMass = m_elemMass.data();
    v.vnew = m_vnew.data();
    v.delv_xi = m_delv_xi.data(); v.delv_eta = m_delv_eta.data(); v.delv_zeta = m_delv_zeta.data();
    v.delx_xi = m_delx_xi.data(); v.delx_eta = m_delx_eta.data(); v.delx_zeta = m_delx_zeta.data();
    v.regNumList = m_regNumList.data();
    v.regElemSize = m_regElemSize.data();
    v.regElemlist = m_regElemlistPtr.data();
    v.nodeElemStart = m_nodeElemStart.data();
    v.nodeElemCornerList = m_nodeElemCornerList.data();
    v.e_cut = m_e_cut; v.p_cut = m_p_cut; v.q_cut = m_q_cut; v.v_cut = m_v_cut; v.u_cut = m_u_cut;
    v.hgcoef = m_hgcoef; v.ss4o3 = m_ss4o3; v.qstop = m_qstop;
    v.monoq_max_slope = m_monoq_max_slope; v.monoq_limiter_mult = m_monoq_limiter_mult;
    v.qlc_monoq = m_qlc_monoq; v.qqc_monoq = m_qqc_monoq; v.qqc = m_qqc;
    v.eosvmax = m_eosvmax; v.eosvmin = m_eosvmin; v.pmin = m_pmin; v.emin = m_emin;
    v.dvovmax = m_dvovmax; v.refdens = m_refdens;
    v.dtcourant = m_dtcourant; v.dthydro = m_dthydro; v.dtfixed = m_dtfixed; v.time = m_time;
    v.deltatime = m_deltatime; v.deltatimemultlb = m_deltatimemultlb;
    v.deltatimemultub = m_deltatimemultub; v.dtmax = m_dtmax; v.stoptime = m_stoptime;
    v.cycle = m_cycle;
    return v;
}

} // namespace lulesh
