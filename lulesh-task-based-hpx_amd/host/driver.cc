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
// Single-dash spellings (-s 30) are accepted too.  --f and --v (VisIt
// output) are not supported.
#include <sys/time.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iomanip>
#include <iostream>
#include <string>

#include "domain.hpp"

using namespace lulesh;

namespace {

struct Opts {
    Int_t its = 9999999, nx = 30, numReg = 11, showProg = 0, quiet = 0, balance = 1, cost = 1;
    int precision = LULESH_PRECISION_FP64;
    bool hostLoop = false, deviceSetup = false, fma = false;
};

[[noreturn]] void usage(const char *msg) {
    std::fprintf(stderr, "%s\n", msg);
    std::fprintf(stderr,
                 "usage: lulesh-amd [--s n] [--i n] [--r n] [--b n] [--c n] [--q] [--p]\n"
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
        else if (k == "f") (void)num(); // number of VisIt files: accepted, unused
        else if (k == "q") o.quiet = 1;
        else if (k == "p") o.showProg = 1;
        else if (k == "v") usage("Use of -v requires compiling with -DVIZ_MESH");
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
    check(lulesh_hip_create(&view, o.precision | (o.fma ? LULESH_OPT_FMA : 0), &ctx), ctx);

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
