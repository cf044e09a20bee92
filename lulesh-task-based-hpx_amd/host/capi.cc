// The C-ABI of include/lulesh_hip.h.
//
// The device code and its context management are compiled twice
// (csrc/lulesh_hip.hip): lulesh_impl_exact rounds every multiply and add on
// its own, lulesh_impl_fma contracts multiply-adds into fused operations.
// A context made with LULESH_OPT_FMA belongs to the second form; this layer
// remembers the form of each context and forwards every call to it.
#include <cstddef>
#include <cstdint>

#include "lulesh_hip.h"

namespace lulesh_impl_exact {
struct lulesh_ctx;
#include "lulesh_api.inc"
} // namespace lulesh_impl_exact

namespace lulesh_impl_fma {
struct lulesh_ctx;
#include "lulesh_api.inc"
} // namespace lulesh_impl_fma

struct lulesh_ctx {
    bool fma;
    void *impl;
};

namespace {

namespace ex = lulesh_impl_exact;
namespace fm = lulesh_impl_fma;

ex::lulesh_ctx *E(const lulesh_ctx *c) { return static_cast<ex::lulesh_ctx *>(c->impl); }
fm::lulesh_ctx *F(const lulesh_ctx *c) { return static_cast<fm::lulesh_ctx *>(c->impl); }

// Run the create call of the form LULESH_OPT_FMA selects (make_exact or
// make_fma, given the options without that bit and a pointer of the form's
// own context type) and wrap the context it produced.
template <typename MakeExact, typename MakeFma>
int create_in_form(int precision_opts, lulesh_ctx **out, MakeExact &&make_exact,
                   MakeFma &&make_fma) {
    if (!out) return LULESH_INVALID_ARGUMENT;
    *out = nullptr;
    const bool fma = (precision_opts & LULESH_OPT_FMA) != 0;
    const int opts = precision_opts & ~LULESH_OPT_FMA;
    void *impl = nullptr;
    int rc;
    if (fma) {
        fm::lulesh_ctx *p = nullptr;
        rc = make_fma(opts, &p);
        impl = p;
    } else {
        ex::lulesh_ctx *p = nullptr;
        rc = make_exact(opts, &p);
        impl = p;
    }
    if (rc != 0 || !impl) return rc;
    *out = new lulesh_ctx{fma, impl};
    return 0;
}

} // namespace

// A null context goes to the exact form, which rejects it as usual.
#define LULESH_FWD(c, fn, ...)                                                              \
    (!(c) ? ex::fn(nullptr, __VA_ARGS__)                                                    \
          : (c)->fma ? fm::fn(F(c), __VA_ARGS__) : ex::fn(E(c), __VA_ARGS__))
#define LULESH_FWD0(c, fn) (!(c) ? ex::fn(nullptr) : (c)->fma ? fm::fn(F(c)) : ex::fn(E(c)))

extern "C" {

int lulesh_hip_create(const lulesh_domain_view *host, int precision, lulesh_ctx **out) {
    return create_in_form(
        precision, out,
        [&](int o, ex::lulesh_ctx **p) { return ex::lulesh_hip_create(host, o, p); },
        [&](int o, fm::lulesh_ctx **p) { return fm::lulesh_hip_create(host, o, p); });
}

int lulesh_hip_batch_create(const lulesh_domain_view *const *views, int32_t nprob,
                            int precision, lulesh_ctx **out) {
    return create_in_form(
        precision, out,
        [&](int o, ex::lulesh_ctx **p) { return ex::lulesh_hip_batch_create(views, nprob, o, p); },
        [&](int o, fm::lulesh_ctx **p) { return fm::lulesh_hip_batch_create(views, nprob, o, p); });
}

int lulesh_hip_create_sedov(int32_t s, int32_t nr, int32_t balance, int32_t cost,
                            int precision_opts, lulesh_ctx **out) {
    return create_in_form(
        precision_opts, out,
        [&](int o, ex::lulesh_ctx **p) {
            return ex::lulesh_hip_create_sedov(s, nr, balance, cost, o, p);
        },
        [&](int o, fm::lulesh_ctx **p) {
            return fm::lulesh_hip_create_sedov(s, nr, balance, cost, o, p);
        });
}

int lulesh_hip_batch_create_sedov(const int32_t *sizes, const int32_t *regions,
                                  const double *factors, int32_t nprob, int32_t balance,
                                  int32_t cost, int precision_opts, lulesh_ctx **out) {
    return create_in_form(
        precision_opts, out,
        [&](int o, ex::lulesh_ctx **p) {
            return ex::lulesh_hip_batch_create_sedov(sizes, regions, factors, nprob, balance,
                                                     cost, o, p);
        },
        [&](int o, fm::lulesh_ctx **p) {
            return fm::lulesh_hip_batch_create_sedov(sizes, regions, factors, nprob, balance,
                                                     cost, o, p);
        });
}

int lulesh_hip_leapfrog(lulesh_ctx *c, double deltatime, double *dtcourant, double *dthydro) {
    return LULESH_FWD(c, lulesh_hip_leapfrog, deltatime, dtcourant, dthydro);
}
int lulesh_hip_run(lulesh_ctx *c, int32_t max_cycles, lulesh_time_state *inout) {
    return LULESH_FWD(c, lulesh_hip_run, max_cycles, inout);
}
int lulesh_hip_phase(lulesh_ctx *c, int phase, double deltatime) {
    return LULESH_FWD(c, lulesh_hip_phase, phase, deltatime);
}
int lulesh_hip_constraints_result(lulesh_ctx *c, double *dtcourant, double *dthydro) {
    return LULESH_FWD(c, lulesh_hip_constraints_result, dtcourant, dthydro);
}
int lulesh_hip_download(lulesh_ctx *c, int field_id, void *dst, size_t count) {
    return LULESH_FWD(c, lulesh_hip_download, field_id, dst, count);
}
int lulesh_hip_upload(lulesh_ctx *c, int field_id, const void *src, size_t count) {
    return LULESH_FWD(c, lulesh_hip_upload, field_id, src, count);
}
int lulesh_hip_download_view(lulesh_ctx *c, lulesh_domain_view *dst) {
    return LULESH_FWD(c, lulesh_hip_download_view, dst);
}
int lulesh_hip_set_timing(lulesh_ctx *c, int enable) {
    return LULESH_FWD(c, lulesh_hip_set_timing, enable);
}
int lulesh_hip_kernel_times(lulesh_ctx *c, const char **names, double *avg_ms,
                            int64_t *launches, int cap) {
    return LULESH_FWD(c, lulesh_hip_kernel_times, names, avg_ms, launches, cap);
}
int64_t lulesh_hip_device_bytes(const lulesh_ctx *c) {
    return LULESH_FWD0(c, lulesh_hip_device_bytes);
}
int32_t lulesh_hip_is_structured(const lulesh_ctx *c) {
    return LULESH_FWD0(c, lulesh_hip_is_structured);
}
double lulesh_hip_last_run_ms(const lulesh_ctx *c) {
    return LULESH_FWD0(c, lulesh_hip_last_run_ms);
}
int lulesh_hip_stream_bandwidth(int64_t bytes, int32_t reps, double *GBs) {
    return ex::lulesh_hip_stream_bandwidth(bytes, reps, GBs);
}
int lulesh_hip_fp64_fma_rate(int32_t reps, double *TFLOPs) {
    return ex::lulesh_hip_fp64_fma_rate(reps, TFLOPs);
}
int lulesh_hip_constraints_noalloc(const double *ss, const double *vdov, const double *arealg,
                                   int32_t numElem, double qqc, double dvovmax, double *out,
                                   void *stream) {
    return ex::lulesh_hip_constraints_noalloc(ss, vdov, arealg, numElem, qqc, dvovmax, out,
                                              stream);
}
int lulesh_hip_set_tuning(lulesh_ctx *c, const lulesh_tuning *tuning) {
    return LULESH_FWD(c, lulesh_hip_set_tuning, tuning);
}
int lulesh_hip_get_tuning(const lulesh_ctx *c, lulesh_tuning *tuning) {
    return LULESH_FWD(c, lulesh_hip_get_tuning, tuning);
}
int lulesh_hip_autotune(lulesh_ctx *c, int32_t cycles, lulesh_tuning *best,
                        double *best_ms_per_cycle) {
    return LULESH_FWD(c, lulesh_hip_autotune, cycles, best, best_ms_per_cycle);
}
int lulesh_hip_get_time_state(lulesh_ctx *c, lulesh_time_state *ts) {
    return LULESH_FWD(c, lulesh_hip_get_time_state, ts);
}
int lulesh_hip_report(lulesh_ctx *c, double out[4]) {
    return LULESH_FWD(c, lulesh_hip_report, out);
}
int lulesh_hip_batch_run(lulesh_ctx *c, int32_t max_cycles, lulesh_time_state *states) {
    return LULESH_FWD(c, lulesh_hip_batch_run, max_cycles, states);
}
int lulesh_hip_batch_download(lulesh_ctx *c, int32_t prob, int field_id, void *dst,
                              size_t count) {
    return LULESH_FWD(c, lulesh_hip_batch_download, prob, field_id, dst, count);
}
int32_t lulesh_hip_num_problems(const lulesh_ctx *c) {
    return LULESH_FWD0(c, lulesh_hip_num_problems);
}
const char *lulesh_hip_last_error(const lulesh_ctx *c) {
    return LULESH_FWD0(c, lulesh_hip_last_error);
}
void lulesh_hip_destroy(lulesh_ctx *c) {
    if (!c) return;
    if (c->fma)
        fm::lulesh_hip_destroy(F(c));
    else
        ex::lulesh_hip_destroy(E(c));
    delete c;
}

} // extern "C"
