/*
 * lulesh_hip.h -- C-ABI of the MI355X (gfx950) LULESH Lagrange leapfrog path.
 *
 * This is the drop-in boundary for the reference's hot path.  In the
 * reference, hpx_main's cycle loop calls
 *
 *     TimeIncrement(*locDom);               // lulesh.cc:2377  (lulesh.cc:182-229)
 *     LagrangeLeapFrogWithTasks(*locDom);   // lulesh.cc:2378  (lulesh.cc:1977-2240)
 *
 * on a Domain whose state lives in host std::vectors (lulesh.h:130-606).
 * Here the Domain's SoA fields are copied once into HBM by lulesh_hip_create
 * and advanced there; the host Domain is refreshed by lulesh_hip_download.
 *
 * Conventions (kept from the reference):
 *   - return 0 on success; -1 = VolumeError, -2 = QStopError (lulesh.h:30,
 *     where the reference calls exit(VolumeError) at lulesh.cc:821, 908, 1262);
 *     the host program turns a negative code into exit(code) so the process
 *     status matches the reference.  Codes <= -100 wrap HIP runtime errors,
 *     -3 = invalid argument.
 *   - inputs are copied at create; the caller keeps ownership of its arrays.
 *   - a context is not thread-safe: one host thread and one HIP stream each.
 *   - all device memory is allocated once at create (the reference mallocs
 *     corner and gradient buffers every cycle, lulesh.cc:2017-2022, 2117).
 *
 * Only plain C types appear in this header.
 */
#ifndef LULESH_HIP_H
#define LULESH_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Error codes, lulesh.h:30 */
#define LULESH_OK 0
#define LULESH_VOLUME_ERROR (-1)
#define LULESH_QSTOP_ERROR (-2)
#define LULESH_INVALID_ARGUMENT (-3)
#define LULESH_HIP_ERROR (-100)

/* Arithmetic forms. */
#define LULESH_PRECISION_FP64 0  /* every kernel in FP64 (the reference's Real_t, lulesh.h:27) */
#define LULESH_PRECISION_MIXED 1 /* hourglass + monotonic-Q gradients in FP32, the rest FP64 */
/* Option bits or-ed into the precision argument of lulesh_hip_create. */
#define LULESH_OPT_NODE_FORCES 0x100 /* keep the fields that are dead between cycles
                                        up to date in HBM, as the reference's Domain
                                        does: fx/fy/fz and xdd/ydd/zdd (lulesh.cc:858-860,
                                        1115-1123) and the Q terms ql/qq (lulesh.cc:
                                        1647-1648).  The default of lulesh_hip_create (a
                                        host Domain's view); batches and problems built in
                                        HBM keep them only with this bit. */
#define LULESH_OPT_NO_NODE_FORCES 0x800 /* lulesh_hip_create only: do not write fx..,
                                        xdd.. and ql/qq (dead between cycles: nothing
                                        reads them before the next cycle overwrites them;
                                        ql/qq are read back only by the EOS of the same
                                        phase, which keeps them in registers).  Their
                                        downloads are then refused, and
                                        lulesh_hip_download_view leaves those eight host
                                        arrays untouched. */
#define LULESH_OPT_GENERIC 0x200     /* use the indexed kernels (nodelist + corner CSR)
                                        even when the mesh is the structured Sedov cube */
#define LULESH_OPT_FMA 0x400        /* contract FP64 (and FP32) multiply-adds into fused
                                        operations, as the reference's own GCC -O3 -mfma
                                        build does: one rounding per multiply-add instead
                                        of two.  Results then agree with the separately
                                        rounded form to within a few ulps per operation
                                        (not bit for bit). */
/* Batched ensembles only: split the problems over k streams (2 <= k <= 15,
 * at most one per problem), each a complete context over its share (largest
 * problems first, each to the stream with the fewest elements so far) driven
 * by its own host thread in lulesh_hip_batch_run.  Kernels of one stream
 * fill the gaps the other leaves at its launch and graph boundaries.  Every
 * problem's fields and time state are bit-identical to the one-stream
 * batch.  No counterpart in the reference, which runs one problem. */
#define LULESH_OPT_STREAMS(k) (((k) & 0xf) << 12)

/*
 * Plain-pointer view of a Domain (lulesh.h:130-606).  Array lengths:
 * N = numNode, E = numElem, P = (sizeX+1)^2.  Pointers to fields that the
 * path recomputes before reading (fx.., xdd.., delv_*.., vnew) may be NULL
 * at create.
 */
typedef struct lulesh_domain_view {
    int32_t sizeX;   /* edge elements (s): lulesh.h:439 */
    int32_t numElem; /* s^3 */
    int32_t numNode; /* (s+1)^3 */
    int32_t numReg;  /* lulesh.h:490 */
    int32_t cost;    /* lulesh.h:491 (-c) */
    int32_t pad0;
    /* node-centred [N], lulesh.h:464-481 */
    double *x, *y, *z, *xd, *yd, *zd, *xdd, *ydd, *zdd, *fx, *fy, *fz, *nodalMass;
    /* symmetry node sets [P], lulesh.h:483-485 */
    int32_t *symmX, *symmY, *symmZ;
    /* element-centred, lulesh.h:496-536 */
    int32_t *nodelist; /* [8E] */
    int32_t *lxim, *lxip, *letam, *letap, *lzetam, *lzetap, *elemBC; /* [E] */
    double *e, *p, *q, *ql, *qq, *v, *volo, *delv, *vdov, *arealg, *ss, *elemMass, *vnew;
    double *delv_xi, *delv_eta, *delv_zeta, *delx_xi, *delx_eta, *delx_zeta;
    /* regions, lulesh.h:492-494 */
    int32_t *regNumList;   /* [E], region number + 1 */
    int32_t *regElemSize;  /* [numReg] */
    int32_t **regElemlist; /* [numReg][regElemSize[r]] */
    /* corner CSR, lulesh.h:592-593 */
    int32_t *nodeElemStart;      /* [N+1] */
    int32_t *nodeElemCornerList; /* [8E] */
    /* constants, lulesh.h:539-560 */
    double e_cut, p_cut, q_cut, v_cut, u_cut, hgcoef, ss4o3, qstop, monoq_max_slope,
        monoq_limiter_mult, qlc_monoq, qqc_monoq, qqc, eosvmax, eosvmin, pmin, emin, dvovmax,
        refdens;
    /* time state, lulesh.h:563-572 */
    double dtcourant, dthydro, dtfixed, time, deltatime, deltatimemultlb, deltatimemultub,
        dtmax, stoptime;
    int32_t cycle;
    int32_t pad1;
} lulesh_domain_view;

/* Time state advanced by the device-side TimeIncrement (lulesh.cc:182-229). */
typedef struct lulesh_time_state {
    double time, deltatime, dtcourant, dthydro;
    int32_t cycle;
    int32_t status; /* 0, or the error code that stopped the loop */
} lulesh_time_state;

/* Field identifiers for upload/download. */
enum lulesh_field_id {
    LULESH_F_X = 0, LULESH_F_Y, LULESH_F_Z, LULESH_F_XD, LULESH_F_YD, LULESH_F_ZD,
    LULESH_F_XDD, LULESH_F_YDD, LULESH_F_ZDD, LULESH_F_FX, LULESH_F_FY, LULESH_F_FZ,
    LULESH_F_NODALMASS,
    LULESH_F_E, LULESH_F_P, LULESH_F_Q, LULESH_F_QL, LULESH_F_QQ, LULESH_F_V, LULESH_F_VOLO,
    LULESH_F_DELV, LULESH_F_VDOV, LULESH_F_AREALG, LULESH_F_SS, LULESH_F_ELEMMASS,
    LULESH_F_VNEW, LULESH_F_DELV_XI, LULESH_F_DELV_ETA, LULESH_F_DELV_ZETA, LULESH_F_DELX_XI,
    LULESH_F_DELX_ETA, LULESH_F_DELX_ZETA,
    LULESH_F_COUNT
};

/* Phases of the cycle, for phase-level parity checks. */
enum lulesh_phase_id {
    LULESH_PHASE_NODAL = 0,    /* lulesh.cc:2017-2110: forces, gather, accel, BC, vel, pos */
    LULESH_PHASE_ELEMENTS = 1, /* lulesh.cc:2112-2136: kinematics, volumes, Q gradients */
    LULESH_PHASE_MATERIAL = 2, /* lulesh.cc:2138-2202: monotonic Q, EOS x rep, sound speed */
    LULESH_PHASE_CONSTRAINTS = 3 /* lulesh.cc:2204-2239: dtcourant / dthydro */
};

typedef struct lulesh_ctx lulesh_ctx;

/* Copy a Domain into HBM.  Replaces the per-cycle allocations of
 * LagrangeLeapFrogWithTasks (lulesh.cc:2017-2022, 2117) with one allocation. */
int lulesh_hip_create(const lulesh_domain_view *host, int precision, lulesh_ctx **out);

/* One Lagrange leapfrog cycle with a host-chosen time step: the drop-in for
 * LagrangeLeapFrogWithTasks(domain) (lulesh.cc:1977-2240), called after the
 * host's TimeIncrement (lulesh.cc:2377).  On return *dtcourant / *dthydro hold
 * the constraint minima (lulesh.cc:2234-2235). */
int lulesh_hip_leapfrog(lulesh_ctx *ctx, double deltatime, double *dtcourant,
                        double *dthydro);

/* The reference cycle loop (lulesh.cc:2374-2386) run on the device: repeat
 * {TimeIncrement; LagrangeLeapFrog} while time < stoptime and cycle <
 * max_cycles.  *inout supplies and returns the time state.  Cycles are
 * replayed from captured HIP graphs; TimeIncrement runs on the device. */
int lulesh_hip_run(lulesh_ctx *ctx, int32_t max_cycles, lulesh_time_state *inout);

/* One phase of the cycle (see lulesh_phase_id) with time step deltatime.
 * The MATERIAL and CONSTRAINTS phases restart the constraint minima at 1e20
 * (lulesh.cc:2211-2212); lulesh_hip_constraints_result reads them back. */
int lulesh_hip_phase(lulesh_ctx *ctx, int phase, double deltatime);
int lulesh_hip_constraints_result(lulesh_ctx *ctx, double *dtcourant, double *dthydro);

/* The constraints phase without a context and without any allocation: no
 * counterpart in the reference; for callers that cannot allocate device
 * memory (code under HIP graph capture, serving loops, harnesses that
 * forbid it).  Computes the Courant and hydro minima (CalcCourantConstraint /
 * CalcHydroConstraint, lulesh.cc:1938-1974; every operation rounded on its
 * own) over caller-owned device arrays ss, vdov, arealg of numElem elements
 * and writes out[0] = dtcourant, out[1] = dthydro (device memory, 1e20 where
 * no element constrains), with the run constants qqc and dvovmax
 * (lulesh-init.cc:27, 31).  Everything is enqueued on `stream` (a
 * hipStream_t; NULL = the default stream) and nothing waits for it.  The
 * partial minima live in a fixed 32 KB device array of the library (one per
 * device, kept until the process ends), cleared at the start of each call:
 * calls that overlap on one device race.  0 <= numElem <= 2^31 - 257, else
 * LULESH_INVALID_ARGUMENT. */
int lulesh_hip_constraints_noalloc(const double *ss, const double *vdov, const double *arealg,
                                   int32_t numElem, double qqc, double dvovmax, double *out,
                                   void *stream);

/* Synchronous copies between HBM and host arrays of `count` elements. */
int lulesh_hip_download(lulesh_ctx *ctx, int field_id, void *dst, size_t count);
int lulesh_hip_upload(lulesh_ctx *ctx, int field_id, const void *src, size_t count);
/* Download every state field into a view (NULL pointers are skipped), and
 * the time state.  fx/fy/fz, xdd/ydd/zdd and ql/qq are skipped (their host
 * arrays left as they are) when the context does not keep them
 * (LULESH_OPT_NO_NODE_FORCES, or a context built without
 * LULESH_OPT_NODE_FORCES by the batch or in-HBM set-up calls); a single
 * lulesh_hip_download of one of them then returns LULESH_INVALID_ARGUMENT. */
int lulesh_hip_download_view(lulesh_ctx *ctx, lulesh_domain_view *dst);

/* Kernel timing of the last lulesh_hip_run call that had timing enabled:
 * names[i] / avg_ms[i] / launches[i] for up to `cap` kernels; returns the
 * number of kernels.  Timing brackets every launch with HIP events on the
 * context's stream. */
int lulesh_hip_set_timing(lulesh_ctx *ctx, int enable);
int lulesh_hip_kernel_times(lulesh_ctx *ctx, const char **names, double *avg_ms,
                            int64_t *launches, int cap);

/* Device memory held by the context, in bytes. */
int64_t lulesh_hip_device_bytes(const lulesh_ctx *ctx);
/* 1 if the structured (index-arithmetic, fused z-march) kernels are in use. */
int32_t lulesh_hip_is_structured(const lulesh_ctx *ctx);
/* Device milliseconds of the last lulesh_hip_run / lulesh_hip_batch_run
 * (HIP events around the replayed graphs, polling included; for a stream
 * group the slowest stream's). */
double lulesh_hip_last_run_ms(const lulesh_ctx *ctx);

/* HBM ceiling for the roofline: the best rate (bytes read + written per
 * second, in GB/s) of a hand-written streaming copy of `bytes` bytes
 * (16-byte vectors, non-temporal loads and stores, five launch shapes, each
 * timed over `reps` copies with HIP events).  Allocates and frees its own
 * two buffers.  A diagnostic with no counterpart in the reference. */
int lulesh_hip_stream_bandwidth(int64_t bytes, int32_t reps, double *GBs);

/* FP64 ceiling for the roofline: the rate (TFLOP/s, a fused multiply-add
 * counted as two) of a hand-written kernel that issues nothing but
 * independent v_fma_f64 (eight chains per lane, four waves per SIMD), timed
 * over `reps` launches of about 5 ms with HIP events.  A diagnostic with no
 * counterpart in the reference. */
int lulesh_hip_fp64_fma_rate(int32_t reps, double *TFLOPs);

/* ---------------------------------------------------------------------- */
/* Launch tuning.  No counterpart in the reference, whose knob is the HPX
 * task size (--task-size / --elems-per-task, lulesh.cc:2267-2335). */
typedef struct lulesh_tuning {
    int32_t zplanes;          /* node planes per z-march workgroup (0 = sized to fill the GPU) */
    int32_t block_elem;       /* threads per block of the kinematics kernel: 64, 128, 256
                                 (other values: 256 for a context of one problem, 128
                                 for a batch) */
    int32_t block_material;   /* threads per block of the material/EOS kernel: 64, 128, 256
                                 (other values: 64 for a context of one problem, 128
                                 for a batch) */
    int32_t cycles_per_graph; /* cycles captured per HIP graph (0 = by problem size) */
    int32_t material_mode;    /* 0: element order, EOS repetitions balanced over each wave;
                                 1: region lists grouped by repetition count;
                                 2: fused with the elements phase in one z-march pass
                                 (structured cube only; the cycle then does not store
                                 the Q gradients delv_* / delx_*, whose downloads are
                                 refused until an elements phase stores them again;
                                 slower than 0 on the Sedov problem: 0.42 against
                                 0.28 ms per cycle at s = 150);
                                 other values: 0, or 1 without repetition counts */
    int32_t em_planes;        /* element planes per workgroup of the fused pass
                                 (0 = sized to fill the GPU) */
} lulesh_tuning;

int lulesh_hip_set_tuning(lulesh_ctx *ctx, const lulesh_tuning *tuning);
int lulesh_hip_get_tuning(const lulesh_ctx *ctx, lulesh_tuning *tuning);
/* Autotuner: times `cycles` cycles for each candidate launch configuration
 * (coordinate descent over zplanes, block_elem, block_material,
 * cycles_per_graph, material_mode, em_planes; each candidate is run twice from the
 * snapshot, the second run timed) starting from a snapshot of the current state, restores
 * the state, and keeps the fastest.  Returns it and its ms per cycle.  A
 * stream group (LULESH_OPT_STREAMS) is tuned as a whole: every candidate is
 * set on all streams, which run at once, and the slowest stream's device
 * time counts. */
int lulesh_hip_autotune(lulesh_ctx *ctx, int32_t cycles, lulesh_tuning *best,
                        double *best_ms_per_cycle);

/* ---------------------------------------------------------------------- */
/* Set-up in HBM.  No counterpart in the reference, whose Domain constructor
 * (lulesh-init.cc:10-183) builds the problem on the host.  Builds the Sedov
 * problem of `s` elements per edge, `nr` regions, region balance and cost
 * (the --s --r --b --c flags) directly on the device: every field equals the
 * one lulesh_domain_create + lulesh_hip_create produce, bit for bit, without
 * the O(E) host loops and the upload (only the region runs of the
 * reference's rand() stream are drawn on the host).  Structured kernels only
 * (LULESH_OPT_GENERIC is rejected) and element-order material.  The time
 * state starts at the constructor's (lulesh_hip_get_time_state). */
int lulesh_hip_create_sedov(int32_t s, int32_t nr, int32_t balance, int32_t cost,
                            int precision_opts, lulesh_ctx **out);
/* The batched ensemble built the same way: problem p has sizes[p] elements
 * per edge, regions[p] regions and its deposit scaled by factors[p] > 0
 * (as lulesh_domain_scale_energy scales a host Domain); equal, bit for bit,
 * to lulesh_hip_batch_create on the corresponding host Domains. */
int lulesh_hip_batch_create_sedov(const int32_t *sizes, const int32_t *regions,
                                  const double *factors, int32_t nprob, int32_t balance,
                                  int32_t cost, int precision_opts, lulesh_ctx **out);
/* Time state of a single-problem context (time, dt, dtcourant, dthydro,
 * cycle), as lulesh_hip_run would hand it back. */
int lulesh_hip_get_time_state(lulesh_ctx *ctx, lulesh_time_state *ts);
/* The report numbers of VerifyAndWriteFinalOutput (lulesh-util.cc:60-115)
 * from the device fields: out[0] = e(0), out[1] = MaxAbsDiff,
 * out[2] = TotalAbsDiff, out[3] = MaxRelDiff. */
int lulesh_hip_report(lulesh_ctx *ctx, double out[4]);

/* ---------------------------------------------------------------------- */
/* Batched ensemble: many independent Sedov Domains in one context, advanced
 * together (one launch per kernel covers every problem) while each keeps its
 * own time step, time and cycle -- exactly as if run alone.  No counterpart
 * in the reference, which runs one Domain per process.  All problems must
 * be structured cubes with the same run constants. */
int lulesh_hip_batch_create(const lulesh_domain_view *const *views, int32_t nprob,
                            int precision, lulesh_ctx **out);
/* Run every problem until its stop time or max_cycles; states[nprob] receive
 * the final time states (status = error code per problem). */
int lulesh_hip_batch_run(lulesh_ctx *ctx, int32_t max_cycles, lulesh_time_state *states);
int lulesh_hip_batch_download(lulesh_ctx *ctx, int32_t prob, int field_id, void *dst,
                              size_t count);
int32_t lulesh_hip_num_problems(const lulesh_ctx *ctx);

const char *lulesh_hip_last_error(const lulesh_ctx *ctx);
void lulesh_hip_destroy(lulesh_ctx *ctx);

#ifdef __cplusplus
}
#endif
#endif
