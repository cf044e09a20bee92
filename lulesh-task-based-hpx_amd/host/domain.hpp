// Host-side Domain of the MI355X LULESH path.
//
// Same data model and accessor names as the reference Domain
// (lulesh.h:130-606): SoA std::vectors for node and element fields, the
// corner CSR, region index sets and the run constants.  The hot path itself
// runs on the GPU (lulesh_hip.h); this class builds the Sedov problem on the
// host, hands plain pointers to the device context and receives the fields
// back for the final report.
#pragma once

#include <cstdint>
#include <vector>

#include "lulesh_hip.h"

namespace lulesh {

using Real_t = double;   // lulesh.h:27
using Index_t = int32_t; // lulesh.h:26
using Int_t = int32_t;   // lulesh.h:28

enum { VolumeError = -1, QStopError = -2 }; // lulesh.h:30

// Boundary-condition bits, lulesh.h:47-75: two bits per face side plus COMM.
enum : Int_t {
    XI_M = 0x00007, XI_M_SYMM = 0x00001, XI_M_FREE = 0x00002, XI_M_COMM = 0x00004,
    XI_P = 0x00038, XI_P_SYMM = 0x00008, XI_P_FREE = 0x00010, XI_P_COMM = 0x00020,
    ETA_M = 0x001c0, ETA_M_SYMM = 0x00040, ETA_M_FREE = 0x00080, ETA_M_COMM = 0x00100,
    ETA_P = 0x00e00, ETA_P_SYMM = 0x00200, ETA_P_FREE = 0x00400, ETA_P_COMM = 0x00800,
    ZETA_M = 0x07000, ZETA_M_SYMM = 0x01000, ZETA_M_FREE = 0x02000, ZETA_M_COMM = 0x04000,
    ZETA_P = 0x38000, ZETA_P_SYMM = 0x08000, ZETA_P_FREE = 0x10000, ZETA_P_COMM = 0x20000,
};

class Domain {
  public:
    // Domain(numRanks, colLoc, rowLoc, planeLoc, nx, tp, nr, balance, cost),
    // lulesh.h:135-137.  This build supports the single-rank decomposition
    // the reference always uses (numRanks = 1, lulesh.cc:2250).
    Domain(Index_t nx, Int_t nr, Int_t balance, Int_t cost);

    // Node-centred accessors, lulesh.h:248-276
    Real_t &x(Index_t i) { return m_x[i]; }
    Real_t &y(Index_t i) { return m_y[i]; }
    Real_t &z(Index_t i) { return m_z[i]; }
    Real_t &xd(Index_t i) { return m_xd[i]; }
    Real_t &yd(Index_t i) { return m_yd[i]; }
    Real_t &zd(Index_t i) { return m_zd[i]; }
    Real_t &xdd(Index_t i) { return m_xdd[i]; }
    Real_t &ydd(Index_t i) { return m_ydd[i]; }
    Real_t &zdd(Index_t i) { return m_zdd[i]; }
    Real_t &fx(Index_t i) { return m_fx[i]; }
    Real_t &fy(Index_t i) { return m_fy[i]; }
    Real_t &fz(Index_t i) { return m_fz[i]; }
    Real_t &nodalMass(Index_t i) { return m_nodalMass[i]; }
    Index_t &symmX(Index_t i) { return m_symmX[i]; }
    Index_t &symmY(Index_t i) { return m_symmY[i]; }
    Index_t &symmZ(Index_t i) { return m_symmZ[i]; }
    bool symmXempty() { return m_symmX.empty(); }
    bool symmYempty() { return m_symmY.empty(); }
    bool symmZempty() { return m_symmZ.empty(); }

    // Element-centred accessors, lulesh.h:281-355
    Index_t &regElemSize(Index_t i) { return m_regElemSize[i]; }
    Index_t &regNumList(Index_t i) { return m_regNumList[i]; }
    Index_t *regNumList() { return m_regNumList.data(); }
    Index_t *regElemlist(Int_t r) { return m_regElemlist[r].data(); }
    Index_t &regElemlist(Int_t r, Index_t i) { return m_regElemlist[r][i]; }
    Index_t *nodelist(Index_t i) { return &m_nodelist[Index_t(8) * i]; }
    Index_t &lxim(Index_t i) { return m_lxim[i]; }
    Index_t &lxip(Index_t i) { return m_lxip[i]; }
    Index_t &letam(Index_t i) { return m_letam[i]; }
    Index_t &letap(Index_t i) { return m_letap[i]; }
    Index_t &lzetam(Index_t i) { return m_lzetam[i]; }
    Index_t &lzetap(Index_t i) { return m_lzetap[i]; }
    Int_t &elemBC(Index_t i) { return m_elemBC[i]; }
    Real_t &vnew(Index_t i) { return m_vnew[i]; }
    Real_t &delv_xi(Index_t i) { return m_delv_xi[i]; }
    Real_t &delv_eta(Index_t i) { return m_delv_eta[i]; }
    Real_t &delv_zeta(Index_t i) { return m_delv_zeta[i]; }
    Real_t &delx_xi(Index_t i) { return m_delx_xi[i]; }
    Real_t &delx_eta(Index_t i) { return m_delx_eta[i]; }
    Real_t &delx_zeta(Index_t i) { return m_delx_zeta[i]; }
    Real_t &e(Index_t i) { return m_e[i]; }
    Real_t &p(Index_t i) { return m_p[i]; }
    Real_t &q(Index_t i) { return m_q[i]; }
    Real_t &ql(Index_t i) { return m_ql[i]; }
    Real_t &qq(Index_t i) { return m_qq[i]; }
    Real_t &v(Index_t i) { return m_v[i]; }
    Real_t &delv(Index_t i) { return m_delv[i]; }
    Real_t &volo(Index_t i) { return m_volo[i]; }
    Real_t &vdov(Index_t i) { return m_vdov[i]; }
    Real_t &arealg(Index_t i) { return m_arealg[i]; }
    Real_t &ss(Index_t i) { return m_ss[i]; }
    Real_t &elemMass(Index_t i) { return m_elemMass[i]; }
    Index_t nodeElemCount(Index_t i) { return m_nodeElemStart[i + 1] - m_nodeElemStart[i]; }
    Index_t *nodeElemCornerList(Index_t i) { return &m_nodeElemCornerList[m_nodeElemStart[i]]; }

    // Parameters, lulesh.h:397-418
    Real_t u_cut() const { return m_u_cut; }
    Real_t e_cut() const { return m_e_cut; }
    Real_t p_cut() const { return m_p_cut; }
    Real_t q_cut() const { return m_q_cut; }
    Real_t v_cut() const { return m_v_cut; }
    Real_t hgcoef() const { return m_hgcoef; }
    Real_t qstop() const { return m_qstop; }
    Real_t monoq_max_slope() const { return m_monoq_max_slope; }
    Real_t monoq_limiter_mult() const { return m_monoq_limiter_mult; }
    Real_t ss4o3() const { return m_ss4o3; }
    Real_t qlc_monoq() const { return m_qlc_monoq; }
    Real_t qqc_monoq() const { return m_qqc_monoq; }
    Real_t qqc() const { return m_qqc; }
    Real_t eosvmax() const { return m_eosvmax; }
    Real_t eosvmin() const { return m_eosvmin; }
    Real_t pmin() const { return m_pmin; }
    Real_t emin() const { return m_emin; }
    Real_t dvovmax() const { return m_dvovmax; }
    Real_t refdens() const { return m_refdens; }

    // Time step controls, lulesh.h:421-445
    Real_t &time() { return m_time; }
    Real_t &deltatime() { return m_deltatime; }
    Real_t &deltatimemultlb() { return m_deltatimemultlb; }
    Real_t &deltatimemultub() { return m_deltatimemultub; }
    Real_t &stoptime() { return m_stoptime; }
    Real_t &dtcourant() { return m_dtcourant; }
    Real_t &dthydro() { return m_dthydro; }
    Real_t &dtmax() { return m_dtmax; }
    Real_t &dtfixed() { return m_dtfixed; }
    Int_t &cycle() { return m_cycle; }
    Index_t &sizeX() { return m_sizeX; }
    Index_t &sizeY() { return m_sizeX; }
    Index_t &sizeZ() { return m_sizeX; }
    Index_t &numReg() { return m_numReg; }
    Int_t &cost() { return m_cost; }
    Index_t &numElem() { return m_numElem; }
    Index_t &numNode() { return m_numNode; }

    // Raw array pointers, lulesh.h:357-392
    Real_t *x_begin() { return m_x.data(); }
    Real_t *y_begin() { return m_y.data(); }
    Real_t *z_begin() { return m_z.data(); }
    Real_t *xd_begin() { return m_xd.data(); }
    Real_t *yd_begin() { return m_yd.data(); }
    Real_t *zd_begin() { return m_zd.data(); }
    Real_t *xdd_begin() { return m_xdd.data(); }
    Real_t *ydd_begin() { return m_ydd.data(); }
    Real_t *zdd_begin() { return m_zdd.data(); }
    Real_t *fx_begin() { return m_fx.data(); }
    Real_t *fy_begin() { return m_fy.data(); }
    Real_t *fz_begin() { return m_fz.data(); }
    Index_t *symmX_begin() { return m_symmX.data(); }
    Index_t *symmY_begin() { return m_symmY.data(); }
    Index_t *symmZ_begin() { return m_symmZ.data(); }
    Real_t *p_begin() { return m_p.data(); }
    Real_t *q_begin() { return m_q.data(); }
    Real_t *v_begin() { return m_v.data(); }
    Real_t *vnew_begin() { return m_vnew.data(); }
    Real_t *nodalMass_begin() { return m_nodalMass.data(); }
    Real_t *vdov_begin() { return m_vdov.data(); }
    Real_t *ss_begin() { return m_ss.data(); }
    Real_t *elemMass_begin() { return m_elemMass.data(); }

    // A plain-pointer view of every array for the C-ABI (lulesh_hip.h).
    lulesh_domain_view view();

  private:
    void BuildMesh(Int_t nx);
    void SetupThreadSupportStructures();
    void CreateRegionIndexSets(Int_t nreg, Int_t balance);
    void SetupSymmetryPlanes(Int_t edgeNodes);
    void SetupElementConnectivities(Int_t edgeElems);
    void SetupBoundaryConditions(Int_t edgeElems);

    std::vector<Real_t> m_x, m_y, m_z, m_xd, m_yd, m_zd, m_xdd, m_ydd, m_zdd, m_fx, m_fy, m_fz,
        m_nodalMass;
    std::vector<Index_t> m_symmX, m_symmY, m_symmZ;
    Int_t m_numReg = 0, m_cost = 0;
    std::vector<Index_t> m_regElemSize, m_regNumList;
    std::vector<std::vector<Index_t>> m_regElemlist;
    std::vector<Index_t *> m_regElemlistPtr;
    std::vector<Index_t> m_nodelist, m_lxim, m_lxip, m_letam, m_letap, m_lzetam, m_lzetap;
    std::vector<Int_t> m_elemBC;
    std::vector<Real_t> m_delv_xi, m_delv_eta, m_delv_zeta, m_delx_xi, m_delx_eta, m_delx_zeta;
    std::vector<Real_t> m_e, m_p, m_q, m_ql, m_qq, m_v, m_volo, m_vnew, m_delv, m_vdov,
        m_arealg, m_ss, m_elemMass;
    std::vector<Index_t> m_nodeElemStart, m_nodeElemCornerList;

  public:
    // Constants, lulesh-init.cc:18-36 (the same for every Domain)
    static constexpr Real_t m_e_cut = 1.0e-7, m_p_cut = 1.0e-7, m_q_cut = 1.0e-7,
                            m_v_cut = 1.0e-10, m_u_cut = 1.0e-7;
    static constexpr Real_t m_hgcoef = 3.0, m_ss4o3 = 4.0 / 3.0, m_qstop = 1.0e+12;
    static constexpr Real_t m_monoq_max_slope = 1.0, m_monoq_limiter_mult = 2.0,
                            m_qlc_monoq = 0.5, m_qqc_monoq = 2.0 / 3.0, m_qqc = 2.0;
    static constexpr Real_t m_eosvmax = 1.0e+9, m_eosvmin = 1.0e-9, m_pmin = 0.,
                            m_emin = -1.0e+15, m_dvovmax = 0.1, m_refdens = 1.0;

  private:

    Real_t m_dtcourant = 0, m_dthydro = 0, m_dtfixed = 0, m_time = 0, m_deltatime = 0,
           m_deltatimemultlb = 0, m_deltatimemultub = 0, m_dtmax = 0, m_stoptime = 0;
    Int_t m_cycle = 0;
    Index_t m_sizeX = 0, m_numElem = 0, m_numNode = 0;
};

// Element volume (the reference's CalcElemVolume, lulesh.cc:592-674), used by
// the Domain constructor for volo / elemMass / nodalMass.
Real_t CalcElemVolume(const Real_t x[8], const Real_t y[8], const Real_t z[8]);

// Lattice coordinate of node index i along an edge of nx elements,
// lulesh-init.cc:207-227.
inline Real_t SedovCoord(Index_t i, Index_t nx) { return Real_t(1.125) * Real_t(i) / Real_t(nx); }

// The energy deposited in element 0 and the initial time step of the Sedov
// problem of nx elements per edge (lulesh-init.cc:172-181).
Real_t SedovInitialEnergy(Index_t nx);
// dt0 for a deposit einit in element 0 (the constructor's, lulesh-init.cc:181,
// for einit = SedovInitialEnergy(nx))
Real_t SedovInitialDeltaTime(Index_t nx, Real_t einit);

// One run of the region map: elements [start, next run's start) belong to
// region `region` (1-based).  SedovRegionRuns replays the reference's libc
// rand() stream (lulesh-init.cc:341-445) for E elements, nr regions and the
// given balance; the last run ends at E.
struct RegionRun {
    Index_t start;
    Int_t region;
};
void SedovRegionRuns(Index_t E, Int_t nr, Int_t balance, std::vector<RegionRun> &runs);

// Report numbers of VerifyAndWriteFinalOutput (lulesh-util.cc:60-115) from
// the first nx*nx entries of the energy field: e(0), MaxAbsDiff,
// TotalAbsDiff, MaxRelDiff of the plane-0 x <-> y symmetry check.
void ReportNumbers(const Real_t *e, Index_t nx, Real_t out[4]);

// TimeIncrement, lulesh.cc:182-229.
void TimeIncrement(Domain &domain);

// The cycle body through the device context (lulesh.cc:2378).
int LagrangeLeapFrog(Domain &domain, lulesh_ctx *ctx);

// lulesh-util.cc:60-97: {e(0), MaxAbsDiff, TotalAbsDiff, MaxRelDiff}.
void VerifyNumbers(Domain &domain, Real_t out[4]);

Int_t RegionRep(Int_t reg, Int_t numReg, Int_t cost);

// The arrays of the visualization dump (lulesh-viz.cc:56-197) of one
// structured Sedov cube of nx elements per edge.  nodelist may be NULL (the
// structured connectivity of lulesh-init.cc:236-244), regNumList too (every
// element in region 1).
struct VizFields {
    Index_t nx = 0;
    Int_t cycle = 0;
    Real_t time = 0;
    const Real_t *x = nullptr, *y = nullptr, *z = nullptr, *xd = nullptr, *yd = nullptr,
                 *zd = nullptr;
    const Real_t *e = nullptr, *p = nullptr, *v = nullptr, *q = nullptr;
    const Index_t *regNumList = nullptr, *nodelist = nullptr;
};
// Write lulesh_plot_c<cycle>.vtk into directory `dir` (NULL or "" = the
// working directory), or `pieces` slabs and a .visit list.  0 on success,
// -3 bad arguments, -4 a file could not be written.
int WriteViz(const VizFields &f, const char *dir, Int_t pieces);
// The dump arrays of a host Domain (its current fields and time state).
VizFields VizOf(Domain &d);

} // namespace lulesh
