/*
 * scft_api.h -- C-ABI of the MI355X (gfx950) SCFT string-method library (libscft_hip.so).
 *
 * Drop-in boundary for SCFT_string's solve_field hot path. The reference binds this path
 * through the C++ class `density` (density.h:225-548) and the host solvers of scft.cu;
 * each entry point below names the reference interface it replaces.
 *
 * Conventions
 *  - Every call returns SCFT_OK (0) or an error code; scft_last_error() gives the message.
 *    Nothing in the library calls exit() (the reference's HANDLE_ERROR does,
 *    GPU_kernels.h:16-48).
 *  - Host arrays at the boundary are FP64 whatever the internal precision.
 *  - Field layout per replica (scft.cu:1037-1051): W row = [W- | W+] (2M doubles),
 *    densities = [phiA | phiB | phih | phicB] (4M doubles); grid index
 *    r = z + y*m2 + x*m2*m1 (scft.cu:462-464).
 *  - One context per host thread; all work is queued on the context's HIP stream and
 *    every call that returns host data synchronises that stream.
 */
#ifndef SCFT_API_H
#define SCFT_API_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum scft_status {
    SCFT_OK = 0,
    SCFT_ERR_INVALID = 1,     /* bad argument or parameter */
    SCFT_ERR_UNSUPPORTED = 2, /* grid size / precision combination without a kernel */
    SCFT_ERR_HIP = 3,         /* HIP runtime error */
    SCFT_ERR_NOMEM = 4,       /* device allocation failed */
    SCFT_ERR_NAN = 5,         /* a field error norm became NaN (scft.cu:539, 770 assert) */
    SCFT_ERR_NOCONV = 6       /* an iteration limit was reached before its tolerance */
} scft_status;

typedef enum scft_precision {
    SCFT_FP64 = 0, /* propagator and FFTs in FP64 */
    SCFT_FP32 = 1  /* propagator and FFTs in FP32; Q, norms, mixing, spline stay FP64 */
} scft_precision;

/* Run parameters: the keys of the reference's input.dat (scft.cu:120-251) plus the
 * compile-time replica count strn (scft.cu:35), made a runtime value here. */
/* Propagator schedule: which kernels run the contour loop and how they are tiled. The
 * reference has no counterpart (cuFFT picks its own plans, density.h:41-75). A zero field
 * keeps the default: the grid's measured choice, or the SCFT_* environment override. A
 * non-zero field the grid cannot run makes scft_ctx_create return SCFT_ERR_INVALID. */
typedef enum {
    SCFT_SCHED_DEFAULT = 0,
    SCFT_SCHED_TWO_STAGE = 1, /* plane/column kernels on register FFTs (grids compiled for them) */
    SCFT_SCHED_STOCKHAM = 2,  /* plane/column kernels on LDS Stockham FFTs */
    SCFT_SCHED_PENCIL = 3,    /* one line pass per axis, any grid whose axes factor into 2, 3, 5 */
    SCFT_SCHED_TWO_PASS = 4,  /* one plane pass and one column pass per contour step, the kinetic
                                 factor split by axis (grids with the two-stage kernels) */
    SCFT_SCHED_PENCIL_SPLIT = 5 /* line passes with the kinetic factor split by axis: x, y, z, y
                                   per contour step, any grid the line kernels take */
} scft_sched_family;

typedef struct scft_schedule {
    int family;   /* scft_sched_family */
    int coltile;  /* two-stage, two-pass: (y, z) columns per workgroup of the x pass (32 or 64) */
    int plane_nt; /* two-stage: threads per workgroup of the plane kernels (the grid's default
                     or half of it, for 512- and 768-thread planes) */
    int streams;  /* replica groups on their own streams, 1 .. 4 */
} scft_schedule;

typedef struct scft_params {
    int m[3];        /* grid points, x slowest */
    double D[3];     /* box lengths in R0 */
    int N;           /* contour steps of the copolymer */
    double f;        /* A fraction as read; rounded to round(N f)/N (scft.cu:1008) */
    double alpha;    /* homopolymer length / N (scft.cu:38) */
    int ens;         /* 1 canonical, 2 grand canonical */
    double phidb;    /* copolymer volume fraction (ens 1) or fugacity (ens 2) */
    double chi;      /* chi N */
    int nrep;        /* replicas on the string (strn) */
    int precision;   /* scft_precision */
    int fix[2];      /* fix first / last replica */
    int dostring;    /* 1: string method, 0: independent replicas */
    int justFE;      /* 1: only the free energy (main skips solve_field) */
    int springsteps; /* iterations with the spring term before redistribution starts */
    int device;      /* HIP device ordinal */
    int fe_fp64;     /* FP32 contexts: 1 = the props that scft_free_energy evaluates F from
                        runs in FP64 (on the FP32-relaxed fields), 0 = in FP32 */
    double phihA;    /* A homopolymer, which the reference declares but holds at zA = 0
                        (density.h:352, 390-397, README.md:3). ens 1: its volume fraction
                        (B homopolymer: 1 - phidb - phihA); ens 2: its activity zA. Length
                        Nh, propagated by q1, so it needs Nh <= n1. 0 = the reference. */
    int anderson;            /* dostring = 0 only: Anderson mixing over this many histories (at
                                most 40) in scft_string_iterate, one props per iteration. The
                                reference keeps only its arrays (scft.cu:30-31, 1029-1031).
                                0 = the reference's mixing. */
    double anderson_lambda;  /* its mixing fraction (<= 0: 0.2) */
    double anderson_start;   /* simple mixing (lambda 0.05) until errs < this (<= 0: 1e-2;
                                started much earlier, Anderson can land on another saddle
                                point, e.g. the disordered one) */
    int autotune;            /* 1: scft_ctx_create first runs scft_autotune for these
                                parameters (a lamellar trial field) and creates the context
                                with the fastest schedule; sched is ignored. 0 = sched. */
    int ckpt;                /* 0: keep the propagators of every contour index (2(N+1) fields
                                per replica, as density.h:293). K > 0: keep every K-th index
                                (K ~ sqrt(N)) and recompute the rest in windows of K for the
                                densities and the box gradient: N/K + 2K + 5 fields per
                                replica, about twice the propagation work. Results agree with
                                K = 0 to rounding. */
    scft_schedule sched;     /* propagator schedule (all zero: the defaults) */
} scft_params;

typedef struct scft_ctx scft_ctx;

/* Replaces `new density(N, D)` + setupSplineBuffers (scft.cu:1010-1011, density.h:272-315,
 * 462-476) and the cuFFT plans of step::step (density.h:41-75). Allocates all device
 * memory for nrep replicas. */
scft_status scft_ctx_create(const scft_params *p, scft_ctx **out);
scft_status scft_ctx_destroy(scft_ctx *ctx);
/* Message of the last failed call on ctx; with ctx NULL, of the calling thread's last failed
 * call that takes no context (scft_spline_derivs_noalloc). */
const char *scft_last_error(const scft_ctx *ctx);
scft_status scft_sync(scft_ctx *ctx);
/* FFT kernels the propagator runs on: "two-stage" or "stockham" (plane/column kernels
 * compiled for the grid) or "pencil" (one line pass per axis, any 2-3-5 grid). The
 * reference has no counterpart: cuFFT picks its own plan (density.h:41-75). */
const char *scft_fft_path(const scft_ctx *ctx);
/* The propagator configuration in words ("two-stage, 2 replica-group streams, direct
 * launches"), followed after an autotune by every candidate's contour-loop time in ms. */
const char *scft_engine_info(scft_ctx *ctx);

/* Autotuner (no counterpart in the reference, whose cuFFT plans are fixed per grid): times
 * the contour loop of one props over p->nrep rows for every schedule the grid can run (kernel
 * family; for the two-stage kernels, column tile x plane workgroup size; replica-group
 * streams 1 .. min(3, nrep)), each in a context of its own that is created, timed and
 * destroyed in turn, so no context is ever reconfigured. The fields are W [nrep][2M] when W
 * is non-NULL, else a lamellar trial field. Each candidate: one untimed props, then the
 * best of `reps` (>= 1) timed ones; candidates within 3% of the fastest are timed again in a
 * fresh context (reps + 1 runs) and keep their best time. The rows timed are those a string
 * iteration propagates (nrep less the fixed ends when dostring = 1 and justFE = 0). *best
 * receives the fastest schedule; report (may be NULL) every candidate with its ms per
 * contour step and the choice, NUL-terminated within report_len. p->sched and p->autotune
 * are ignored. */
scft_status scft_autotune(const scft_params *p, const double *W, int reps, scft_schedule *best,
                          char *report, size_t report_len);

/* Replaces density::props (density.h:344-459), batched over `nrep` independent rows:
 * W [nrep][2M] in; phi [nrep][4M], lnQ [nrep], phis [nrep][5] out (phi, phis may be NULL).
 * Uses scratch state only; the context's field state (scft_set_fields) is untouched. */
scft_status scft_props_batch(scft_ctx *ctx, const double *W, double *phi, double *lnQ,
                             double *phis, int nrep);

/* Same as scft_props_batch with device pointers for W (row stride 2M) and phi
 * (row stride 4M) on the context's device; lnQ and phis are host arrays. W is read on the
 * context's own (non-blocking) stream: the caller completes the work that writes W before
 * the call (e.g. synchronises its producer stream). phi is complete when the call returns. */
scft_status scft_props_batch_device(scft_ctx *ctx, const double *dW, double *dphi,
                                    double *lnQ, double *phis, int nrep);

/* Context field state: W rows [nrep][2M] (the W-, W+ part of scft.cu's W[ii]). */
scft_status scft_set_fields(scft_ctx *ctx, const double *W);
scft_status scft_get_fields(scft_ctx *ctx, double *W);
/* Densities of the last props per replica, [nrep][4M] (W[ii][2M:6M] in scft.cu). */
scft_status scft_get_densities(scft_ctx *ctx, double *phi);
/* Tangent dW-/dalpha of the last projection, [nrep][M] (W[ii][6M:7M] in scft.cu). */
scft_status scft_get_tangent(scft_ctx *ctx, double *dWda);

/* Replaces solve_field0 (scft.cu:516-554) for every replica with which[ii] != 0 (NULL:
 * all), batched, each replica with its own mixing parameter and early exit.
 * k_out[ii] receives the value solve_field0 returns (may be NULL). */
scft_status scft_solve_w_plus(scft_ctx *ctx, int maxIter, double errTol, const int *which,
                              int *k_out);

/* Per-iteration report of one solve_field body (scft.cu:592-757). Arrays hold nrep
 * entries; the caller provides them (any may be NULL). */
typedef struct scft_iter_report {
    double *errs, *errs3, *x, *xref, *dotWs, *dotWs2, *normDER, *normDEV;
    int *steps;
    double fitup;
    double stepav;
} scft_iter_report;

/* One iteration k of solve_field (scft.cu:592-757) for the single-process layout:
 * solve_field0(20, errTol) + props for every active replica, DEV, errs, perpendicular
 * projection, fixed-lambda mixing, arc length, spline redistribution. Fixed ends
 * (params.fix) are excluded as in scft.cu:587-588. alf [nrep] receives the arc-length
 * coordinate (scft.cu:737); may be NULL. */
scft_status scft_string_iterate(scft_ctx *ctx, int k, double errTol, double *alf,
                                scft_iter_report *rep);

/* Replaces FreeE (scft.cu:812-853), single process: solve_field0(100, 1e-4) and props for
 * every active replica plus both ends, then FEs0 = -lnQ, FEs1, FEs, phctot, alf.
 * Any output may be NULL. */
scft_status scft_free_energy(scft_ctx *ctx, double *FE, double *FEs, double *FEs0,
                             double *FEs1, double *phctot, double *alf);

/* Box derivative of lnQ (density.h:344-399) at fixed W for nrow rows W[nrow][2M]:
 * lnQ[nrow] (may be NULL) and dlnQ_dD[nrow][3] out. The exact derivative of the discrete
 * Richardson propagator: the stored q1 and q2 serve as each other's adjoints. Uses
 * scratch state only. The reference has no counterpart (it computes no gradients); this
 * is the backward pass for stress-free box relaxation. */
scft_status scft_dlnq_dbox(scft_ctx *ctx, const double *W, int nrow, double *lnQ,
                           double *dlnQ_dD);

/* Free-energy derivatives at the current field state, per replica (props on every row):
 * dFdD[nrep][3] = -d lnQ/dD_i, the box stress, and dFdchi[nrep] = <phiA phiB>, the chiN
 * derivative of FEs (scft.cu:823-828). Both are the total derivatives of FreeE's F at a
 * saddle point, where F is stationary in W. Either output may be NULL. No counterpart in
 * the reference. */
scft_status scft_box_stress(scft_ctx *ctx, double *dFdD, double *dFdchi);

/* Backward pass of the replicas' free energies F_r (FEs of FreeE, scft.cu:812-853) at the
 * current saddle fields, as an autograd backward takes it: from dL_dF[nrep] it returns
 * dL_dD[nrep][3] = dL_dF[r] dF_r/dD_i and dL_dchi[nrep] = dL_dF[r] dF_r/dchiN (sum over r
 * for the shared box and chiN). The derivatives are those of scft_box_stress. dL_dF must
 * not be NULL; either output may be. No counterpart in the reference. */
scft_status scft_backward(scft_ctx *ctx, const double *dL_dF, double *dL_dD, double *dL_dchi);

/* New box lengths D[3] for every later call: rebuilds the kinetic tables (setup_k_sq,
 * density.h:79-95, which the reference runs once at start-up). */
scft_status scft_set_box(scft_ctx *ctx, const double D[3]);

/* Independent solves of uneven size run together: one scft_string_iterate (iteration k[i],
 * tolerance errTol) on each of the n contexts at once, each from its own host thread. Every
 * context owns its streams, so the GPU runs their kernels concurrently, and small solves
 * fill the CUs that one of them alone leaves idle. Results are the same as n calls in
 * sequence. alf[i] and reps[i] are the per-context outputs of scft_string_iterate (alf and
 * reps may be NULL). The contexts must be distinct. Returns SCFT_OK or the status of the
 * first context (in index order) that failed; its message is in scft_last_error(ctxs[i]).
 * No counterpart in the reference, which runs one replica per MPI rank. */
scft_status scft_iterate_concurrent(scft_ctx *const *ctxs, int n, const int *k, double errTol,
                                    double *const *alf, scft_iter_report *reps);

/* Stress-free box relaxation of a single-replica context with dostring = 0. Starting
 * from box D (in/out), alternately (1) iterates the fields with the solve_field body
 * (scft_string_iterate, solve_w_plus tolerance 1e-3 as in scft.cu:1123) until errs <= errTol
 * or maxFieldIter iterations, and (2) moves each box length selected in `axes` (bit i for
 * D_i) by a secant step on dF/dD_i (first step 1% of D_i, every step at most 10%), until
 * |dF/dD_i| <= stressTol on every selected axis. dFdD[3] (may be NULL) gets the last
 * stress, *boxes (may be NULL) the number of boxes visited. Returns SCFT_ERR_NOCONV when
 * maxBoxes boxes pass without reaching stressTol. No counterpart in the reference. */
scft_status scft_relax_box(scft_ctx *ctx, int axes, double stressTol, double errTol,
                           int maxFieldIter, int maxBoxes, double D[3], double *dFdD,
                           int *boxes);

/* The string's spline on host rows, chunk by chunk, as class density exposes it
 * (density.h:462-528, called from scft.cu:643-650, 746-751, 1011). ww[ii] are the nrep
 * replica rows (host; W- of replica ii from index 0), a chunk is grid points r0 .. r0+n-1.
 *
 * Replaces setupSplineBuffers (density.h:462-476): device and pinned staging for chunks of
 * up to chunkM points of every replica (min(chunkM, M) is kept). Needed once before the
 * chunk calls; calling it again resizes. */
scft_status scft_spline_setup(scft_ctx *ctx, int chunkM);
/* Replace setSplineX / setSplineXRef (density.h:478-488): nodes x[nrep] (strictly
 * increasing; the reference divides by zero otherwise) and evaluation points xref[nrep] of
 * the following chunk calls (copied). */
scft_status scft_spline_set_x(scft_ctx *ctx, const double *x);
scft_status scft_spline_set_xref(scft_ctx *ctx, const double *xref);
/* Replaces splineDerivsChunked (density.h:491-509): out[ii][r] = derivative at x[ii] of the
 * natural cubic spline through (x[j], ww[j][r]), for every ii with doiis[ii] == 1 and r in
 * the chunk. Other rows of out are not touched. */
scft_status scft_spline_derivs(scft_ctx *ctx, const double *const *ww, int r0, int n, const int *doiis,
                               double *const *out);
/* Replaces splineRedistChunked (density.h:511-528): ww[ii][r] += lambda (s(xref[ii]) - ww[ii][r])
 * with s the spline through (x[j], ww[j][r]), for interior ii (0 < ii < nrep - 1) with
 * doiis[ii] == 1 and r in the chunk; the spline is fitted to the values before the update. */
scft_status scft_spline_redist(scft_ctx *ctx, double *const *ww, int r0, int n, const int *doiis, double lambda);

/* No counterpart in the reference. For callers that cannot allocate device memory (code
 * under HIP graph capture, serving loops, harnesses that forbid hipMalloc): the spline
 * tangent of scft_spline_derivs (splineDerivsChunked, density.h:491-509) on device rows and
 * without a context. dout[ii][r] = derivative at x[ii] of the natural cubic spline through
 * (x[j], dW[j][r]), for every ii with doiis[ii] == 1 and r < M; other rows of dout are not
 * touched. dW: nrep device rows of stride wstride >= M (W- first); dout: device rows of
 * stride ostride >= M; x (strictly increasing) and doiis: host arrays of nrep.
 * The work is queued on `stream` (a hipStream_t; NULL: the null stream) and not waited for.
 * Besides dW and dout it uses only a static device array of the library (33 KB per GPU,
 * held until the process ends): nrep above 64 is refused with SCFT_ERR_UNSUPPORTED. Every
 * call on one GPU shares that array and rewrites all of it, so calls that overlap in time on
 * one GPU (on other streams or threads) race; calls on one stream are ordered. It allocates
 * nothing and makes no host-device copy. scft_last_error(NULL) gives the calling thread's
 * last message. */
scft_status scft_spline_derivs_noalloc(int nrep, long long M, const double *x, const int *doiis,
                                       const double *dW, long long wstride, double *dout,
                                       long long ostride, void *stream);

/* Weights of the natural cubic spline through (x[i], a[i]), i < nrep (cubefit_chunk_g,
 * GPU_kernels.h:232-284), at the points xe[erows[j]], j < ne: G[j][i] (row-major, ne x nrep)
 * is the derivative (deriv = 1) or the value (deriv = 0) there per unit a[i]; a point lies
 * on the interval of the last node strictly below it (GPU_kernels.h:291-297). x strictly
 * increasing, 2 <= nrep <= 64. Host only, no device work: the string kernels apply these
 * weights to every grid column. No counterpart in the reference, which refits the spline
 * at every grid point. */
scft_status scft_spline_weights(int nrep, const double *x, const double *xe, const int *erows, int ne,
                                int deriv, double *G);

/* Diagnostics for tests: after a props call on slot `slot` of the scratch batch,
 * qq[s] = <q1'[s] q2'[N-s]> for s = 0..N (the A4 symmetry invariant, with the
 * mean-field factoring left in; all entries equal QC' in exact arithmetic). */
scft_status scft_debug_qq_invariant(scft_ctx *ctx, int slot, double *qq);

/* Diagnostics for the benchmark: runs props on the first nrep rows of the context field
 * state `reps` times and returns the mean time per contour step (ms), measured with HIP
 * events on the context stream around the contour loop (begin + N steps). */
scft_status scft_debug_time_props(scft_ctx *ctx, int nrep, int reps, double *ms_per_step);

#ifdef __cplusplus
}
#endif
#endif
