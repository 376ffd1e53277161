/*
 * scft_oracle.h -- FP64 CPU restatement of SCFT_string's solve_field path.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker that the HIP product
 * path is compared against. Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it. The product (libscft_hip.so, the
 * C++ host driver) never links or calls it.
 *
 * Each function cites the reference file:line it restates. Files:
 *   density.h      class step (propagator), class density (props, spline wrappers)
 *   GPU_kernels.h  elementwise / reduction / spline kernels
 *   scft.cu        solve_field0, solve_field, FreeE, EDist
 *
 * Parity pins (what checks each function; see tests/test_oracle_*.py):
 *   orc_rfft3/orc_irfft3    numpy.fft.rfftn/irfftn fixtures (tests/golden)
 *   orc_ksq                 closed form + numpy.fft.fftfreq
 *   orc_step                uniform-field and single-mode closed forms,
 *                           operator symmetry (<q1(s) q2(N-s)> independent of s)
 *   orc_props               W=0 Simpson closed forms (odd/even n1, Nh truncation),
 *                           uniform-field Q closed forms, canonical normalisation
 *   orc_spline_*            scipy.interpolate.CubicSpline(bc_type='natural') fixtures
 *   orc_solve_field0        one-step map against hand-derived uniform-field update
 *   orc_string_iter         straight-line string fixed point, projection invariant
 *                           (parity of the control flow itself: transcription only)
 *   orc_free_energy         disordered-phase closed form
 *   orc_dlnq_dbox           central differences of orc_props in D; zero for a uniform field
 */
#ifndef SCFT_ORACLE_H
#define SCFT_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_grid orc_grid;

/* grid: m[3] points (x slowest, z fastest; r = z + y*m2 + x*m2*m1), box D[3] in R0 */
orc_grid *orc_grid_create(const int m[3], const double D[3]);
void orc_grid_destroy(orc_grid *g);
int orc_grid_M(const orc_grid *g);
int orc_grid_Mk(const orc_grid *g);
void orc_set_threads(int n);

/* density.h:79-95 step::setup_k_sq -> ksq[Mk] */
void orc_ksq(const orc_grid *g, double *ksq);

/* cufftExecD2Z / Z2D equivalents (density.h:67-74): unnormalised 3-D real<->half-complex.
 * out/in are interleaved complex, Mk = m0*m1*(m2/2+1) entries, kz fastest. */
void orc_rfft3(orc_grid *g, const double *in, double *out);
void orc_irfft3(orc_grid *g, const double *in, double *out);

/* density.h:129-220 step::do_step for ONE chain with the factors of its species:
 *   e4 = exp(-w ds/4) (GPU_kernels.h:220), e2 = e4^2 (GPU_kernels.h:221),
 *   k12 = exp(-k^2 ds/12)/M, k6 = k12^2*M (GPU_kernels.h:211-212). */
void orc_step(orc_grid *g, const double *e4, const double *e2, const double *k12,
              const double *k6, const double *in, double *out);

/* density.h:344-459 density::props.
 *   w[0:2M] = (W-, W+) in;  w[2M:6M] = (phiA, phiB, phih, phicB) out.
 *   Q[3] = (Q1, Q2, QC) out (may be NULL).
 *   qstore: optional [N+1][2][M] copy of (q1[s], q2[s]) (density.h:296-301). */
void orc_props(orc_grid *g, double *w, double *lnQ, double *phis, int Ns, double alpha,
               double phic, double f, int ens, double *Q, double *qstore);

/* scft.cu:516-554 solve_field0 on one replica row Wrow (7M doubles, scft.cu:1037-1051),
 * DEV = that replica's DEV[ii][0] (2M). Returns the final k. */
int orc_solve_field0(orc_grid *g, double *Wrow, double *DEV, double chi, double f, int N,
                     double alpha, double phic, int ens, int maxIter, double errTol);

/* GPU_kernels.h:232-284 cubefit_chunk_g for one grid point: natural cubic spline through
 * (x[i], a[i]), i < n. Outputs c, b, d (a is the input). */
void orc_spline_coeffs(int n, const double *x, const double *a, double *c, double *b,
                       double *d);
/* GPU_kernels.h:286-306: derivative at every node x[j] (interval = last i with x[i] < x[j]) */
void orc_spline_deriv_nodes(int n, const double *x, const double *a, double *out);
/* GPU_kernels.h:308-329: value at xr[j] (interval = last i with x[i] < xr[j]) */
void orc_spline_value_at(int n, const double *x, const double *a, const double *xr,
                         double *out);

/* scft.cu:351-355 EDist over W- of two rows */
double orc_edist(int M, const double *Wa, const double *Wb);

/* Run parameters shared by the string routines (scft.cu globals + main locals). */
typedef struct {
    double chi, f, alpha, phidb;
    int N, ens, nrep, dostring;
    int fix[2];
    int springsteps;
} orc_run;

/* Per-iteration report of one solve_field body (scft.cu:592-806). */
typedef struct {
    double errs[256], errs3[256], x[256], xref[256], dotWs[256], dotWs2[256];
    double normDER[256], normDEV[256];
    int steps[256];
    double fitup;
} orc_iter_report;

/* One iteration k of scft.cu:592-757 (before the periodic output / FreeE block).
 * W: nrep rows of 7M, DEV: nrep rows of 2M, doiis: nrep flags (already masked by fix,
 * scft.cu:587-588), alf: nrep out. errTol is solve_field's errTol (1e-3 in main). */
void orc_string_iter(orc_grid *g, const orc_run *run, double **W, double **DEV, int *doiis,
                     int k, double errTol, double *alf, orc_iter_report *rep);

/* scft.cu:812-853 FreeE for a single process (procid 0 == numprocs-1).
 * Mutates W+ through solve_field0(100, 1e-4). Returns FE; per-replica arrays out. */
double orc_free_energy(orc_grid *g, const orc_run *run, double **W, double **DEV,
                       const int *doiis, double *alf, double *FEs0, double *FEs1, double *FEs,
                       double *phctot);

/* A homopolymer for every later orc_props / orc_dlnq_dbox on this grid (default 0, the
 * reference's zA = 0, density.h:352, 390-397). ens 1: volume fraction (B homopolymer
 * 1 - phic - phihA); ens 2: activity zA. It uses q1 up to Nh (needs Nh <= n1). Pinned by
 * the A <-> B mirror symmetry: negating W- and swapping f <-> 1 - f maps it onto the
 * reference's B homopolymer. */
void orc_grid_set_phiha(orc_grid *g, double phihA);

/* Exact derivative of the discrete lnQ of density::props (density.h:344-399) with respect
 * to the box lengths D[0..2] at fixed W, from the stored q1, q2 as adjoints. w as in
 * orc_props (phi out). dlnQ[3] out. Pinned by central differences of orc_props. */
void orc_dlnq_dbox(orc_grid *g, double *w, int Ns, double alpha, double phic, double f, int ens,
                   double *lnQ, double dlnQ[3]);

/* F3: opt-in Anderson mixing for dostring = 0 (the reference keeps only its arrays, DIM /
 * DDEV / WIN, scft.cu:30-31, 1029-1031). State of one replica with H histories. */
typedef struct orc_anderson orc_anderson;
orc_anderson *orc_anderson_create(int M, int H);
void orc_anderson_destroy(orc_anderson *a);
/* Gaussian elimination with partial pivoting; 0 when singular (pivot < 1e-14 max|U|) */
int orc_small_solve(int n, double *U, double *V, double *c);
/* One iteration on one replica row W (7M): props, DEV, errs; the reference's simple mixing
 * (lambda 0.05) while errs >= start, Anderson with mixing fraction lambda after. Returns
 * errs. Pinned by: H = 0 equals simple mixing with lambda; the fixed point equals simple
 * mixing's; a linear residual map is solved exactly after enough histories. */
double orc_anderson_iter(orc_grid *g, double *W, orc_anderson *a, double chi, double f, int N,
                         double alpha, double phic, int ens, double lambda, double start);

#ifdef __cplusplus
}
#endif
#endif
