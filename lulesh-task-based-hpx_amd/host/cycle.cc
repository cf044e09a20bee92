// Host pieces of the cycle: TimeIncrement, the LagrangeLeapFrog boundary
// call into the device context, the verification numbers, and the C-ABI of
// the host Domain (lulesh_domain.h).
#include <cmath>
#include <new>
#include <random>

#include "domain.hpp"
#include "lulesh_domain.h"

namespace lulesh {

void TimeIncrement(Domain &domain) {
    // lulesh.cc:182-229: Courant/hydro limited step with lb/ub ratio clamps,
    // dtmax cap and the stoptime approach rule; cycle 0 keeps dt0.
    Real_t targetdt = domain.stoptime() - domain.time();
    if ((domain.dtfixed() <= Real_t(0.0)) && (domain.cycle() != Int_t(0))) {
        const Real_t olddt = domain.deltatime();
        Real_t gnewdt = Real_t(1.0e+20);
        if (domain.dtcourant() < gnewdt) gnewdt = domain.dtcourant() / Real_t(2.0);
        if (domain.dthydro() < gnewdt) gnewdt = domain.dthydro() * Real_t(2.0) / Real_t(3.0);
        Real_t newdt = gnewdt;
        const Real_t ratio = newdt / olddt;
        if (ratio >= Real_t(1.0)) {
            if (ratio < domain.deltatimemultlb())
                newdt = olddt;
            else if (ratio > domain.deltatimemultub())
                newdt = olddt * domain.deltatimemultub();
        }
        if (newdt > domain.dtmax()) newdt = domain.dtmax();
        domain.deltatime() = newdt;
    }
    if ((targetdt > domain.deltatime()) &&
        (targetdt < (Real_t(4.0) * domain.deltatime() / Real_t(3.0))))
        targetdt = Real_t(2.0) * domain.deltatime() / Real_t(3.0);
    if (targetdt < domain.deltatime()) domain.deltatime() = targetdt;
    domain.time() += domain.deltatime();
    ++domain.cycle();
}

int LagrangeLeapFrog(Domain &domain, lulesh_ctx *ctx) {
    // The cycle body of lulesh.cc:2378; constraint minima land in the Domain
    // as at lulesh.cc:2234-2235.
    double dtc = 0, dth = 0;
    const int rc = lulesh_hip_leapfrog(ctx, domain.deltatime(), &dtc, &dth);
    if (rc == 0) {
        domain.dtcourant() = dtc;
        domain.dthydro() = dth;
    }
    return rc;
}

void ReportNumbers(const Real_t *e, Index_t nx, Real_t out[4]) {
    // Plane-0 symmetry check of the energy field (lulesh-util.cc:81-97): the
    // Sedov solution is symmetric under x <-> y, so e(j*nx+k) must equal
    // e(k*nx+j).  RelDiff divides by the signed e(k*nx+j).
    Real_t maxAbs = 0.0, totalAbs = 0.0, maxRel = 0.0;
    for (Index_t j = 0; j < nx; ++j)
        for (Index_t k = j + 1; k < nx; ++k) {
            const Real_t a = e[j * nx + k], b = e[k * nx + j];
            const Real_t absDiff = std::fabs(a - b);
            totalAbs += absDiff;
            if (maxAbs < absDiff) maxAbs = absDiff;
            const Real_t rel = std::fabs(b) > 1e-8 ? absDiff / b : 0.0;
            if (maxRel < rel) maxRel = rel;
        }
    out[0] = e[0];
    out[1] = maxAbs;
    out[2] = totalAbs;
    out[3] = maxRel;
}

void VerifyNumbers(Domain &domain, Real_t out[4]) {
    ReportNumbers(&domain.e(0), domain.sizeX(), out);
}

Int_t RegionRep(Int_t reg, Int_t numReg, Int_t cost) {
    // EOS repetitions per region, lulesh.cc:2159-2169: the lower half of the
    // regions once, the upper half 1+cost times, and the last
    // (numReg+15)/20 regions 10*(1+cost) times.
    if (reg < numReg / 2) return 1;
    if (reg < numReg - (numReg + 15) / 20) return 1 + cost;
    return 10 * (1 + cost);
}

} // namespace lulesh

using lulesh::Domain;

struct lulesh_domain {
    Domain dom;
    lulesh_domain(int32_t nx, int32_t nr, int32_t b, int32_t c) : dom(nx, nr, b, c) {}
};

extern "C" {

lulesh_domain *lulesh_domain_create(int32_t nx, int32_t nr, int32_t balance, int32_t cost) {
    if (nx < 1 || nr < 1 || nx > 1600) return nullptr;
    try {
        return new lulesh_domain(nx, nr, balance, cost);
    } catch (...) {
        return nullptr;
    }
}

void lulesh_domain_destroy(lulesh_domain *d) { delete d; }

int lulesh_domain_get_view(lulesh_domain *d, lulesh_domain_view *out) {
    if (!d || !out) return LULESH_INVALID_ARGUMENT;
    *out = d->dom.view();
    return 0;
}

int lulesh_domain_set_time_state(lulesh_domain *d, const lulesh_time_state *ts) {
    if (!d || !ts) return LULESH_INVALID_ARGUMENT;
    d->dom.time() = ts->time;
    d->dom.deltatime() = ts->deltatime;
    d->dom.dtcourant() = ts->dtcourant;
    d->dom.dthydro() = ts->dthydro;
    d->dom.cycle() = ts->cycle;
    return 0;
}

int lulesh_domain_get_time_state(lulesh_domain *d, lulesh_time_state *ts) {
    if (!d || !ts) return LULESH_INVALID_ARGUMENT;
    ts->time = d->dom.time();
    ts->deltatime = d->dom.deltatime();
    ts->dtcourant = d->dom.dtcourant();
    ts->dthydro = d->dom.dthydro();
    ts->cycle = d->dom.cycle();
    ts->status = 0;
    return 0;
}

void lulesh_time_increment(lulesh_domain *d) {
    if (d) lulesh::TimeIncrement(d->dom);
}

int lulesh_lagrange_leapfrog(lulesh_domain *d, lulesh_ctx *ctx) {
    if (!d || !ctx) return LULESH_INVALID_ARGUMENT;
    return lulesh::LagrangeLeapFrog(d->dom, ctx);
}

void lulesh_verify_report(lulesh_domain *d, double out[4]) {
    if (d && out) lulesh::VerifyNumbers(d->dom, out);
}

int32_t lulesh_region_rep(int32_t reg, int32_t numReg, int32_t cost) {
    return lulesh::RegionRep(reg, numReg, cost);
}

int lulesh_domain_scale_energy(lulesh_domain *d, double factor) {
    if (!d || !(factor > 0)) return LULESH_INVALID_ARGUMENT;
    Domain &dom = d->dom;
    const double einit = dom.e(0) * factor;
    dom.e(0) = einit;
    dom.deltatime() = (0.5 * std::cbrt(dom.volo(0))) / std::sqrt(2.0 * einit);
    return 0;
}

int lulesh_domain_write_viz(lulesh_domain *d, const char *dir, int32_t pieces) {
    if (!d || pieces < 1) return LULESH_INVALID_ARGUMENT;
    return lulesh::WriteViz(lulesh::VizOf(d->dom), dir, pieces);
}

int lulesh_batch_params(uint64_t seed, int32_t nprob, int32_t smin, int32_t smax,
                        int32_t *sizes, int32_t *regions, double *factors) {
    if (nprob < 0 || smin < 1 || smax < smin || !sizes || !regions || !factors)
        return LULESH_INVALID_ARGUMENT;
    // Raw engine outputs mapped by hand so the draws do not depend on the
    // standard library's distribution algorithms.
    std::mt19937_64 gen(seed);
    const uint64_t span = uint64_t(smax - smin + 1);
    static const int32_t kRegions[3] = {11, 16, 21};
    for (int32_t p = 0; p < nprob; ++p) {
        sizes[p] = smin + int32_t(gen() % span);
        regions[p] = kRegions[gen() % 3];
        factors[p] = 0.5 + 1.5 * double(gen() >> 11) * 0x1.0p-53;
    }
    return 0;
}

} // extern "C"
