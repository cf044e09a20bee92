// lulesh-amd: command-line driver of the MI355X LULESH path.
//
// Same flags, defaults, cycle loop and report as the reference program
// (hpx_main, lulesh.cc:2244-2413; ParseCommandLineOptions and
// VerifyAndWriteFinalOutput, lulesh-util.cc:36-115).  The cycle body runs on
// the GPU through the C-ABI of lulesh_hip.h.
//
//   --s <n>   edge elements (default 30)        --i <n>  cycle limit (9999999)
//   --r <n>   regions (11)                       --b <n>  region balance (1)
//   --c <n>   extra cost of expensive regions (1) --q      quiet: CSV line only
//   --p       progress every 100 cycles
//   --precision fp64|mixed   arithmetic form (fp64)
//   --fma     fuse multiply-adds (LULESH_OPT_FMA, as the reference's own
//             GCC -O3 -mfma build); default: every multiply and add
//             rounded on its own
//   --host-loop   TimeIncrement on the host and one lulesh_hip_leapfrog call
//                 per cycle (the reference's loop shape); default is the
//                 device-side loop (lulesh_hip_run)
//   --device-setup  build the problem in HBM (lulesh_hip_create_sedov)
//                 instead of on the host; the report comes from the device
//                 fields (not with --host-loop)
//   --v       write the final state for VisIt (lulesh.cc:2397-2400): VTK
//             legacy files lulesh_plot_c<cycle>.vtk in the working directory
//             (the reference's Silo dump, lulesh-viz.cc, has the same objects)
//   --f <n>   split that dump into n slabs of element planes, listed in
//             lulesh_plot_c<cycle>.visit (default 1)
// Single-dash spellings (-s 30) are accepted too.
#include <sys/time.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iomanip>
#include <iostream>
#include <string>
#include <vector>

#include "domain.hpp"

using namespace lulesh;

namespace {

struct Opts {
    Int_t its = 9999999, nx = 30, numReg = 11, showProg = 0, quiet = 0, balance = 1, cost = 1;
    Int_t numFiles = 1, viz = 0;
    int precision = LULESH_PRECISION_FP64;
    bool hostLoop = false, deviceSetup = false, fma = false;
};

[[noreturn]] void usage(const char *msg) {
    std::fprintf(stderr, "%s\n", msg);
    std::fprintf(stderr,
                 "usage: lulesh-amd [--s n] [--i n] [--r n] [--b n] [--c n] [--q] [--p] [--v]\n"
                 "                  [--f n]\n"
                 "                  [--precision fp64|mixed] [--fma] [--host-loop] [--device-setup]\n");
    std::exit(-1);
}

Opts parse(int argc, char **argv) {
    Opts o;
    for (int a = 1; a < argc; ++a) {
        std::string k = argv[a];
        while (!k.empty() && k[0] == '-') k.erase(0, 1);
        auto next = [&]() -> std::string {
            if (a + 1 >= argc) usage(("missing value for --" + k).c_str());
            return argv[++a];
        };
        auto num = [&]() {
            const std::string v = next();
            char *end = nullptr;
            const long r = std::strtol(v.c_str(), &end, 10);
            if (end == v.c_str() || *end != '\0') usage(("bad integer for --" + k).c_str());
            return Int_t(r);
        };
        if (k == "s") o.nx = num();
        else if (k == "i") o.its = num();
        else if (k == "r") o.numReg = num();
        else if (k == "b") o.balance = num();
        else if (k == "c") o.cost = num();
        else if (k == "f") o.numFiles = num();
        else if (k == "q") o.quiet = 1;
        else if (k == "p") o.showProg = 1;
        else if (k == "v") o.viz = 1;
        else if (k == "host-loop") o.hostLoop = true;
        else if (k == "device-setup") o.deviceSetup = true;
        else if (k == "fma") o.fma = true;
        else if (k == "precision") {
            const std::string v = next();
            if (v == "fp64") o.precision = LULESH_PRECISION_FP64;
            else if (v == "mixed") o.precision = LULESH_PRECISION_MIXED;
            else usage("--precision must be fp64 or mixed");
        } else if (k == "h" || k == "help") usage("");
        else usage(("unknown option " + std::string(argv[a])).c_str());
    }
    if (o.nx < 1 || o.numReg < 1) usage("--s and --r must be >= 1");
    if (o.numFiles < 1) usage("--f must be >= 1");
    if (o.hostLoop && o.deviceSetup) usage("--host-loop needs the host set-up");
    return o;
}

double now_s() {
    timeval t;
    gettimeofday(&t, nullptr);
    return double(t.tv_sec) + double(t.tv_usec) / 1e6;
}

void check(int rc, lulesh_ctx *ctx) {
    if (rc == 0) return;
    if (rc == VolumeError || rc == QStopError) std::exit(rc); // lulesh.cc:821, 908, 1262
    std::fprintf(stderr, "lulesh-amd: device error %d: %s\n", rc, lulesh_hip_last_error(ctx));
    std::exit(-1);
}

// The viz dump of a problem built in HBM: its fields come back from the
// device, the region numbers from the region runs of the set-up.
void write_viz_from_device(const Opts &o, lulesh_ctx *ctx, const lulesh_time_state &ts) {
    const int64_t E = int64_t(o.nx) * o.nx * o.nx, N = int64_t(o.nx + 1) * (o.nx + 1) * (o.nx + 1);
    std::vector<Real_t> nf[6], ef[4];
    const int nid[6] = {LULESH_F_X, LULESH_F_Y, LULESH_F_Z, LULESH_F_XD, LULESH_F_YD, LULESH_F_ZD};
    const int eid[4] = {LULESH_F_E, LULESH_F_P, LULESH_F_V, LULESH_F_Q};
    for (int k = 0; k < 6; ++k) {
        nf[k].resize(size_t(N));
        check(lulesh_hip_download(ctx, nid[k], nf[k].data(), size_t(N)), ctx);
    }
    for (int k = 0; k < 4; ++k) {
        ef[k].resize(size_t(E));
        check(lulesh_hip_download(ctx, eid[k], ef[k].data(), size_t(E)), ctx);
    }
    std::vector<RegionRun> runs;
    SedovRegionRuns(Index_t(E), o.numReg, o.balance, runs);
    std::vector<Index_t> reg(static_cast<size_t>(E));
    for (size_t r = 0; r < runs.size(); ++r) {
        const int64_t end = r + 1 < runs.size() ? runs[r + 1].start : E;
        for (int64_t e = runs[r].start; e < end; ++e) reg[size_t(e)] = runs[r].region;
    }
    VizFields f;
    f.nx = o.nx;
    f.cycle = ts.cycle;
    f.time = ts.time;
    f.x = nf[0].data();
    f.y = nf[1].data();
    f.z = nf[2].data();
    f.xd = nf[3].data();
    f.yd = nf[4].data();
    f.zd = nf[5].data();
    f.e = ef[0].data();
    f.p = ef[1].data();
    f.v = ef[2].data();
    f.q = ef[3].data();
    f.regNumList = reg.data();
    if (WriteViz(f, nullptr, o.numFiles) != 0) std::printf("Error writing out viz file - rank 0\n");
}

// VerifyAndWriteFinalOutput, lulesh-util.cc:60-115; v = e(0), MaxAbsDiff,
// TotalAbsDiff, MaxRelDiff
void report(const Real_t v[4], Int_t cycles, double elapsed, Int_t nx) {
    const int64_t nx8 = nx;
    const double grind = ((elapsed * 1e6) / cycles) / double(nx8 * nx8 * nx8);
    std::cout << "Run completed:\n";
    std::cout << "   Problem size        =  " << nx << "\n";
    std::cout << "   Iteration count     =  " << cycles << "\n";
    std::cout << "   Final Origin Energy =  " << std::scientific << std::setprecision(6)
              << std::setw(12) << v[0] << "\n";
    std::cout << "   Testing Plane 0 of Energy Array on rank 0:\n";
    std::cout << "        MaxAbsDiff   = " << std::setw(12) << v[1] << "\n";
    std::cout << "        TotalAbsDiff = " << std::setw(12) << v[2] << "\n";
    std::cout << "        MaxRelDiff   = " << std::setw(12) << v[3] << "\n";
    std::cout << "   Origin Energy (17 digits) = " << std::setprecision(16) << v[0] << "\n";
    std::cout.unsetf(std::ios_base::floatfield);
    std::cout << std::fixed << std::setprecision(2);
    std::cout << "\nElapsed time         = " << std::setw(10) << elapsed << " (s)\n";
    std::cout << std::setprecision(8);
    std::cout << "Grind time (us/z/c)  = " << std::setw(10) << grind << " (per dom)  ("
              << std::setw(10) << elapsed << " overall)\n";
    // The reference labels this "(z/s)" but prints zone-cycles per millisecond.
    std::cout << "FOM                  = " << std::setw(10) << 1000.0 / grind << " (z/s)\n";
    std::cout << "Zone-cycles/s        = " << std::scientific << std::setprecision(4)
              << 1e6 / grind << "\n\n";
}

} // namespace

int main(int argc, char **argv) {
    const Opts o = parse(argc, argv);
    if (!o.quiet) {
        std::cout << "Running problem size " << o.nx << "^3 per domain until completion\n";
        std::cout << "Num processors: 1\n";
        std::cout << "Device: one MI355X (gfx950), " << (o.precision ? "mixed" : "fp64")
                  << (o.fma ? " fused multiply-add" : "") << " form, "
                  << (o.hostLoop ? "host" : "device") << " time loop\n";
        std::cout << "Total number of elements: " << int64_t(o.nx) * o.nx * o.nx << " \n\n";
    }
    if (o.deviceSetup) {
        lulesh_ctx *ctx = nullptr;
        check(lulesh_hip_create_sedov(o.nx, o.numReg, o.balance, o.cost,
                                      o.precision | (o.fma ? LULESH_OPT_FMA : 0), &ctx),
              ctx);
        lulesh_time_state ts{};
        check(lulesh_hip_get_time_state(ctx, &ts), ctx);
        const double stoptime = 1.0e-2; // lulesh-init.cc:136
        const double t0 = now_s();
        while (ts.time < stoptime && ts.cycle < o.its) {
            Int_t limit = o.its;
            if (o.showProg && !o.quiet) limit = std::min(o.its, (ts.cycle / 100 + 1) * 100);
            check(lulesh_hip_run(ctx, limit, &ts), ctx);
            if (o.showProg && !o.quiet && ts.cycle % 100 == 0)
                std::cout << "cycle = " << ts.cycle << ", " << std::scientific
                          << "time = " << ts.time << ", dt=" << ts.deltatime << "\n"
                          << std::defaultfloat;
        }
        const double elapsed = now_s() - t0;
        Real_t v[4];
        check(lulesh_hip_report(ctx, v), ctx);
        if (o.viz) write_viz_from_device(o, ctx, ts);
        lulesh_hip_destroy(ctx);
        if (!o.quiet)
            report(v, ts.cycle, elapsed, o.nx);
        else
            std::cout << o.nx << "," << o.numReg << "," << ts.cycle << ",1," << elapsed << ","
                      << std::scientific << std::setprecision(6) << std::setw(12) << v[0]
                      << std::endl;
        return 0;
    }
    Domain dom(o.nx, o.numReg, o.balance, o.cost);
    lulesh_domain_view view = dom.view();
    lulesh_ctx *ctx = nullptr;
    // the report needs e only: the dead node forces and accelerations are not stored
    check(lulesh_hip_create(&view,
                            o.precision | (o.fma ? LULESH_OPT_FMA : 0) | LULESH_OPT_NO_NODE_FORCES,
                            &ctx),
          ctx);

    // Timed window as in the reference (lulesh.cc:2368-2395): the cycle loop
    // only, with the fields already resident on the device.
    const double t0 = now_s();
    if (o.hostLoop) {
        while (dom.time() < dom.stoptime() && dom.cycle() < o.its) {
            TimeIncrement(dom);
            check(LagrangeLeapFrog(dom, ctx), ctx);
            if (o.showProg && !o.quiet && dom.cycle() % 100 == 0)
                std::cout << "cycle = " << dom.cycle() << ", " << std::scientific
                          << "time = " << dom.time() << ", dt=" << dom.deltatime() << "\n"
                          << std::defaultfloat;
        }
    } else {
        lulesh_time_state ts{dom.time(), dom.deltatime(), dom.dtcourant(), dom.dthydro(),
                             dom.cycle(), 0};
        while (ts.time < dom.stoptime() && ts.cycle < o.its) {
            Int_t limit = o.its;
            if (o.showProg && !o.quiet) limit = std::min(o.its, (ts.cycle / 100 + 1) * 100);
            check(lulesh_hip_run(ctx, limit, &ts), ctx);
            if (o.showProg && !o.quiet && ts.cycle % 100 == 0)
                std::cout << "cycle = " << ts.cycle << ", " << std::scientific
                          << "time = " << ts.time << ", dt=" << ts.deltatime << "\n"
                          << std::defaultfloat;
        }
        dom.time() = ts.time;
        dom.deltatime() = ts.deltatime;
        dom.dtcourant() = ts.dtcourant;
        dom.dthydro() = ts.dthydro;
        dom.cycle() = ts.cycle;
    }
    const double elapsed = now_s() - t0;

    // Bring the fields back for the report (outside the timed window).
    lulesh_domain_view back = dom.view();
    const Real_t keepTime = dom.time();
    const Int_t keepCycle = dom.cycle();
    check(lulesh_hip_download_view(ctx, &back), ctx);
    dom.time() = keepTime;
    dom.cycle() = keepCycle;
    lulesh_hip_destroy(ctx);

    // Write out the final viz file (lulesh.cc:2397-2400)
    if (o.viz && WriteViz(VizOf(dom), nullptr, o.numFiles) != 0)
        std::printf("Error writing out viz file - rank 0\n");

    if (!o.quiet) {
        Real_t v[4];
        VerifyNumbers(dom, v);
        report(v, dom.cycle(), elapsed, o.nx);
    } else {
        std::cout << o.nx << "," << o.numReg << "," << dom.cycle() << ",1," << elapsed << ","
                  << std::scientific << std::setprecision(6) << std::setw(12) << dom.e(0)
                  << std::endl;
    }
    return 0;
}
