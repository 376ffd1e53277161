/*
 * scft_oracle.c -- FP64 CPU restatement of SCFT_string's solve_field path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, never the product). Loaded by tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg; nothing in
 * scft_string_amd/ links or calls it.
 *
 * The reference (CUDA + cuFFT + MPI) cannot be built here (no nvcc, no cuFFT), so this
 * is a line-by-line restatement of its arithmetic in plain C99 + OpenMP, with a
 * project-written mixed-radix FFT in place of cuFFT. The R2C/C2R transforms follow the
 * textbook decomposition (real z-lines -> half spectrum -> complex y, x transforms), which
 * is deliberately different from the GPU path (which packs q and q-dagger into one complex
 * field), so the two implementations share no FFT code or layout.
 *
 * Reference files: density.h, GPU_kernels.h, scft.cu (cited per function).
 */
#define _GNU_SOURCE
#include "scft_oracle.h"

#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef double complex cplx;

/* Mutation switch for the oracle's own tests (tests/test_oracle_mutations.py): 0 builds
 * the restatement; a value k > 0 plants one known mistake in the string iteration or
 * FreeE (see the ORC_MUTANT sites) so the tests can show their closed forms catch it.
 * liborc.so is always built with 0. */
#ifndef ORC_MUTANT
#define ORC_MUTANT 0
#endif

/* ------------------------------------------------------------------------------------ */
/* 1-D mixed-radix Stockham FFT (unnormalised; forward uses exp(-2*pi*i*jk/n), like cuFFT) */
/* ------------------------------------------------------------------------------------ */
typedef struct {
    int n, nf, fac[40];
    cplx *tw; /* tw[t] = exp(-2 pi i t / n), computed in long double, rounded once */
} fft1d;

static void fft1d_init(fft1d *p, int n)
{
    p->n = n;
    p->nf = 0;
    int r = n;
    while (r % 4 == 0) { p->fac[p->nf++] = 4; r /= 4; }
    while (r % 2 == 0) { p->fac[p->nf++] = 2; r /= 2; }
    for (int q = 3; r > 1; q += 2)
        while (r % q == 0) { p->fac[p->nf++] = q; r /= q; }
    p->tw = (cplx *)malloc(sizeof(cplx) * (size_t)n);
    const long double pi = acosl(-1.0L);
    for (int t = 0; t < n; t++) {
        long double a = -2.0L * pi * (long double)t / (long double)n;
        p->tw[t] = (double)cosl(a) + I * (double)sinl(a);
    }
}

static void fft1d_free(fft1d *p) { free(p->tw); p->tw = NULL; }

/* small DFT of size R on v (in place), sign from `inv`; generic O(R^2) for R not in {2,3,4,5} */
static void dft_small(cplx *v, int R, int inv, const fft1d *p)
{
    cplx o[64];
    if (R == 2) {
        cplx a = v[0], b = v[1];
        v[0] = a + b; v[1] = a - b;
        return;
    }
    if (R == 4) {
        cplx a = v[0] + v[2], b = v[0] - v[2], c = v[1] + v[3], d = v[1] - v[3];
        /* forward: multiply d by -i; inverse: by +i */
        cplx dj = inv ? (-cimag(d) + I * creal(d)) : (cimag(d) - I * creal(d));
        v[0] = a + c; v[2] = a - c; v[1] = b + dj; v[3] = b - dj;
        return;
    }
    const int step = p->n / R;
    for (int u = 0; u < R; u++) {
        cplx s = 0;
        for (int r = 0; r < R; r++) {
            cplx w = p->tw[((u * r) % R) * step];
            if (inv) w = conj(w);
            s += v[r] * w;
        }
        o[u] = s;
    }
    for (int u = 0; u < R; u++) v[u] = o[u];
}

/* x (length n, stride 1) transformed in place; work has n entries */
static void fft1d_exec(const fft1d *p, cplx *x, cplx *work, int inv)
{
    const int n = p->n;
    cplx *a = x, *b = work;
    int Ns = 1;
    for (int f = 0; f < p->nf; f++) {
        const int R = p->fac[f], nb = n / R, tws = n / (Ns * R);
        for (int j = 0; j < nb; j++) {
            const int k = j % Ns;
            cplx v[64];
            for (int r = 0; r < R; r++) {
                cplx w = p->tw[r * k * tws];
                if (inv) w = conj(w);
                v[r] = a[j + r * nb] * w;
            }
            dft_small(v, R, inv, p);
            const int idx = (j / Ns) * Ns * R + k;
            for (int u = 0; u < R; u++) b[idx + u * Ns] = v[u];
        }
        cplx *t = a; a = b; b = t;
        Ns *= R;
    }
    if (a != x) memcpy(x, a, sizeof(cplx) * (size_t)n);
}

/* ------------------------------------------------------------------------------------ */
/* grid                                                                                  */
/* ------------------------------------------------------------------------------------ */
struct orc_grid {
    int m[3], M, Mk, mh; /* mh = m2/2+1 */
    double D[3];
    fft1d p[3];
    double *ksq;            /* [Mk] */
    cplx *h;                /* [Mk] half-spectrum work */
    double *k12, *k6;       /* [Mk] */
    double *e4A, *e2A, *e4B, *e2B; /* [M] */
    double *qs0, *qs1;      /* [M] */
    double *qq;             /* [2(N+1)M] propagator store */
    size_t qq_cap;
    double phihA;           /* A homopolymer (0 = the reference, zA = 0) */
};

void orc_grid_set_phiha(orc_grid *g, double phihA) { g->phihA = phihA; }

void orc_set_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

/* density.h:79-95 step::setup_k_sq */
static void setup_ksq(orc_grid *g)
{
    const int *m = g->m;
    const double *DA = g->D;
    const double pi = 4 * atan(1.0); /* scft.cu:38 */
    for (int k0 = -(m[0] - 1) / 2; k0 <= m[0] / 2; k0++) {
        const int K0 = (k0 < 0) ? (k0 + m[0]) : k0;
        const double A0 = k0 * k0 / (DA[0] * DA[0]);
        for (int k1 = -(m[1] - 1) / 2; k1 <= m[1] / 2; k1++) {
            const int K1 = (k1 < 0) ? (k1 + m[1]) : k1;
            const double A1 = A0 + k1 * k1 / (DA[1] * DA[1]);
            for (int k2 = 0; k2 <= m[2] / 2; k2++) {
                const double A2 = A1 + k2 * k2 / (DA[2] * DA[2]);
                const int k = k2 + (m[2] / 2 + 1) * (K1 + m[1] * K0);
                g->ksq[k] = 4 * pi * pi * A2;
            }
        }
    }
}

orc_grid *orc_grid_create(const int m[3], const double D[3])
{
    orc_grid *g = (orc_grid *)calloc(1, sizeof(orc_grid));
    for (int i = 0; i < 3; i++) { g->m[i] = m[i]; g->D[i] = D[i]; fft1d_init(&g->p[i], m[i]); }
    g->M = m[0] * m[1] * m[2];
    g->mh = m[2] / 2 + 1;
    g->Mk = m[0] * m[1] * g->mh; /* scft.cu:1006 */
    g->ksq = (double *)malloc(sizeof(double) * g->Mk);
    g->k12 = (double *)malloc(sizeof(double) * g->Mk);
    g->k6 = (double *)malloc(sizeof(double) * g->Mk);
    g->h = (cplx *)malloc(sizeof(cplx) * g->Mk);
    g->e4A = (double *)malloc(sizeof(double) * g->M);
    g->e2A = (double *)malloc(sizeof(double) * g->M);
    g->e4B = (double *)malloc(sizeof(double) * g->M);
    g->e2B = (double *)malloc(sizeof(double) * g->M);
    g->qs0 = (double *)malloc(sizeof(double) * g->M);
    g->qs1 = (double *)malloc(sizeof(double) * g->M);
    setup_ksq(g);
    return g;
}

void orc_grid_destroy(orc_grid *g)
{
    if (!g) return;
    for (int i = 0; i < 3; i++) fft1d_free(&g->p[i]);
    free(g->ksq); free(g->k12); free(g->k6); free(g->h);
    free(g->e4A); free(g->e2A); free(g->e4B); free(g->e2B);
    free(g->qs0); free(g->qs1); free(g->qq);
    free(g);
}

int orc_grid_M(const orc_grid *g) { return g->M; }
int orc_grid_Mk(const orc_grid *g) { return g->Mk; }
void orc_ksq(const orc_grid *g, double *ksq) { memcpy(ksq, g->ksq, sizeof(double) * g->Mk); }

/* ------------------------------------------------------------------------------------ */
/* 3-D real <-> half-complex transforms (cufftExecD2Z / cufftExecZ2D, density.h:67-74)   */
/* ------------------------------------------------------------------------------------ */
static void rfft3_core(orc_grid *g, const double *in, cplx *H)
{
    const int m0 = g->m[0], m1 = g->m[1], m2 = g->m[2], mh = g->mh;
    /* z: real lines -> half spectrum */
#pragma omp parallel
    {
        cplx *line = (cplx *)malloc(sizeof(cplx) * m2), *work = (cplx *)malloc(sizeof(cplx) * m2);
#pragma omp for schedule(static)
        for (int xy = 0; xy < m0 * m1; xy++) {
            for (int z = 0; z < m2; z++) line[z] = in[(size_t)xy * m2 + z];
            fft1d_exec(&g->p[2], line, work, 0);
            for (int k = 0; k < mh; k++) H[(size_t)xy * mh + k] = line[k];
        }
        free(line); free(work);
    }
    /* y: complex lines of stride mh */
#pragma omp parallel
    {
        cplx *line = (cplx *)malloc(sizeof(cplx) * m1), *work = (cplx *)malloc(sizeof(cplx) * m1);
#pragma omp for schedule(static)
        for (int xk = 0; xk < m0 * mh; xk++) {
            const int x = xk / mh, k = xk % mh;
            cplx *base = H + (size_t)x * m1 * mh + k;
            for (int y = 0; y < m1; y++) line[y] = base[(size_t)y * mh];
            fft1d_exec(&g->p[1], line, work, 0);
            for (int y = 0; y < m1; y++) base[(size_t)y * mh] = line[y];
        }
        free(line); free(work);
    }
    /* x: complex lines of stride m1*mh */
#pragma omp parallel
    {
        cplx *line = (cplx *)malloc(sizeof(cplx) * m0), *work = (cplx *)malloc(sizeof(cplx) * m0);
#pragma omp for schedule(static)
        for (int yk = 0; yk < m1 * mh; yk++) {
            cplx *base = H + yk;
            for (int x = 0; x < m0; x++) line[x] = base[(size_t)x * m1 * mh];
            fft1d_exec(&g->p[0], line, work, 0);
            for (int x = 0; x < m0; x++) base[(size_t)x * m1 * mh] = line[x];
        }
        free(line); free(work);
    }
}

/* H is consumed (overwritten) */
static void irfft3_core(orc_grid *g, cplx *H, double *out)
{
    const int m0 = g->m[0], m1 = g->m[1], m2 = g->m[2], mh = g->mh;
#pragma omp parallel
    {
        cplx *line = (cplx *)malloc(sizeof(cplx) * m0), *work = (cplx *)malloc(sizeof(cplx) * m0);
#pragma omp for schedule(static)
        for (int yk = 0; yk < m1 * mh; yk++) {
            cplx *base = H + yk;
            for (int x = 0; x < m0; x++) line[x] = base[(size_t)x * m1 * mh];
            fft1d_exec(&g->p[0], line, work, 1);
            for (int x = 0; x < m0; x++) base[(size_t)x * m1 * mh] = line[x];
        }
        free(line); free(work);
    }
#pragma omp parallel
    {
        cplx *line = (cplx *)malloc(sizeof(cplx) * m1), *work = (cplx *)malloc(sizeof(cplx) * m1);
#pragma omp for schedule(static)
        for (int xk = 0; xk < m0 * mh; xk++) {
            const int x = xk / mh, k = xk % mh;
            cplx *base = H + (size_t)x * m1 * mh + k;
            for (int y = 0; y < m1; y++) line[y] = base[(size_t)y * mh];
            fft1d_exec(&g->p[1], line, work, 1);
            for (int y = 0; y < m1; y++) base[(size_t)y * mh] = line[y];
        }
        free(line); free(work);
    }
    /* z: Hermitian extension of each line, inverse, real part */
#pragma omp parallel
    {
        cplx *line = (cplx *)malloc(sizeof(cplx) * m2), *work = (cplx *)malloc(sizeof(cplx) * m2);
#pragma omp for schedule(static)
        for (int xy = 0; xy < m0 * m1; xy++) {
            const cplx *src = H + (size_t)xy * mh;
            for (int k = 0; k < mh; k++) line[k] = src[k];
            for (int k = mh; k < m2; k++) line[k] = conj(src[m2 - k]);
            fft1d_exec(&g->p[2], line, work, 1);
            for (int z = 0; z < m2; z++) out[(size_t)xy * m2 + z] = creal(line[z]);
        }
        free(line); free(work);
    }
}

void orc_rfft3(orc_grid *g, const double *in, double *out)
{
    rfft3_core(g, in, (cplx *)out);
}

void orc_irfft3(orc_grid *g, const double *in, double *out)
{
    cplx *H = (cplx *)malloc(sizeof(cplx) * g->Mk);
    memcpy(H, in, sizeof(cplx) * g->Mk);
    irfft3_core(g, H, out);
    free(H);
}

/* ------------------------------------------------------------------------------------ */
/* A4: one Richardson-extrapolated step for one chain (density.h:129-220)               */
/* ------------------------------------------------------------------------------------ */
static void sandwich(orc_grid *g, double *q, const double *kf)
{
    /* D2Z -> Mult_self(expKsq*) -> Z2D (density.h:149-156, GPU_kernels.h:175-181) */
    rfft3_core(g, q, g->h);
    const int Mk = g->Mk;
#pragma omp parallel for schedule(static)
    for (int k = 0; k < Mk; k++) g->h[k] = creal(g->h[k]) * kf[k] + I * (cimag(g->h[k]) * kf[k]);
    irfft3_core(g, g->h, q);
}

void orc_step(orc_grid *g, const double *e4, const double *e2, const double *k12,
              const double *k6, const double *in, double *out)
{
    const int M = g->M;
    double *qs0 = g->qs0, *qs1 = g->qs1;
    /* (1) qs0 = in * expWds2  (Mult, density.h:137-145) */
#pragma omp parallel for schedule(static)
    for (int r = 0; r < M; r++) qs0[r] = in[r] * e4[r];
    sandwich(g, qs0, k12);
    /* qs0 *= expWds  (Mult_self, density.h:160-170) */
#pragma omp parallel for schedule(static)
    for (int r = 0; r < M; r++) qs0[r] *= e2[r];
    sandwich(g, qs0, k12);
    /* full step: qs1 = in * expWds (density.h:185-195) */
#pragma omp parallel for schedule(static)
    for (int r = 0; r < M; r++) qs1[r] = in[r] * e2[r];
    sandwich(g, qs1, k6);
    /* Richardson (GPU_kernels.h:199-203) */
#pragma omp parallel for schedule(static)
    for (int r = 0; r < M; r++) out[r] = (qs0[r] * e4[r] * 4 - qs1[r] * e2[r]) / 3;
}

/* ------------------------------------------------------------------------------------ */
/* A5-A9: props (density.h:344-459)                                                      */
/* ------------------------------------------------------------------------------------ */
static double sum_all(const double *a, int M)
{
    double s = 0;
    for (int r = 0; r < M; r++) s += a[r];
    return s;
}

void orc_props(orc_grid *g, double *w, double *lnQ, double *phis, int Ns, double alpha,
               double phic, double f, int ens, double *Qout, double *qstore)
{
    const int M = g->M, Mk = g->Mk;
    const int n1 = (int)round(Ns * f), n2 = Ns - n1;
    const int Nh = (int)(alpha * Ns); /* density.h:346: truncation */
    const double ds = 1.0 / Ns;
    double Q1 = 0, Q2 = 0, QC = 0, zA = 0, zB = 0, zC = 0;

    const size_t need = (size_t)2 * (Ns + 1) * M;
    if (g->qq_cap < need) {
        free(g->qq);
        g->qq = (double *)malloc(sizeof(double) * need);
        g->qq_cap = need;
    }
    /* q1[s] = qq + 2 s M, q2[s] = qq + (2 s + 1) M  (density.h:296-301) */
#define Q1P(s) (g->qq + (size_t)2 * (s) * M)
#define Q2P(s) (g->qq + ((size_t)2 * (s) + 1) * M)

    /* prepareFactors (density.h:118-125; GPU_kernels.h:207-225) */
#pragma omp parallel for schedule(static)
    for (int k = 0; k < Mk; k++) {
        g->k12[k] = exp(-g->ksq[k] * ds / 12) / M;
        g->k6[k] = g->k12[k] * g->k12[k] * M;
    }
    const double *Wm = w, *Wp = w + M;
#pragma omp parallel for schedule(static)
    for (int r = 0; r < M; r++) {
        g->e4A[r] = exp(-(Wm[r] + Wp[r]) * ds / 4);
        g->e2A[r] = g->e4A[r] * g->e4A[r];
        g->e4B[r] = exp(-(-Wm[r] + Wp[r]) * ds / 4);
        g->e2B[r] = g->e4B[r] * g->e4B[r];
    }
    for (int r = 0; r < M; r++) { Q1P(0)[r] = 1.0; Q2P(0)[r] = 1.0; }

    /* contour loop (density.h:374); species by the step's start index (density.h:136-146) */
    for (int s = 0; s < Ns; s++) {
        if (s < n1) orc_step(g, g->e4A, g->e2A, g->k12, g->k6, Q1P(s), Q1P(s + 1));
        else        orc_step(g, g->e4B, g->e2B, g->k12, g->k6, Q1P(s), Q1P(s + 1));
        if (s < n2) orc_step(g, g->e4B, g->e2B, g->k12, g->k6, Q2P(s), Q2P(s + 1));
        else        orc_step(g, g->e4A, g->e2A, g->k12, g->k6, Q2P(s), Q2P(s + 1));
    }

    /* Q (density.h:377-386); Q1 of the A homopolymer over q1 up to Nh */
    Q2 = sum_all(Q2P(Nh), M);
    QC = sum_all(Q1P(Ns), M);
    Q1 = sum_all(Q1P(Nh), M);
    Q1 /= M; Q2 /= M; QC /= M;

    /* ensemble (density.h:388-399). With g->phihA = 0 this is the reference (zA = 0);
     * otherwise ens 1 gives the A homopolymer the volume fraction phihA and the B
     * homopolymer 1 - phic - phihA, and ens 2 gives it the activity zA = phihA. */
    if (ens == 1) {
        const double phihB = 1.0 - phic - g->phihA;
        *lnQ = phic * log(QC) + (phihB / alpha) * log(Q2);
        if (g->phihA != 0.0) *lnQ += (g->phihA / alpha) * log(Q1);
        zA = g->phihA / (Q1 * alpha);
        zB = phihB / (Q2 * alpha);
        zC = phic / QC;
    } else if (ens == 2) {
        zA = g->phihA;
        zB = 1.0;
        zC = phic;
        *lnQ = zA * Q1 + zB * Q2 + zC * QC;
    } else {
        *lnQ = NAN;
        zB = zC = NAN;
    }
    if (Qout) { Qout[0] = Q1; Qout[1] = Q2; Qout[2] = QC; }

    double *rA = w + 2 * M, *rB = w + 3 * M, *rh = w + 4 * M, *rCB = w + 5 * M;
    double *rhA = (double *)malloc(sizeof(double) * M), *rhB = (double *)malloc(sizeof(double) * M);
    double *rCA = (double *)malloc(sizeof(double) * M), *rC = (double *)malloc(sizeof(double) * M);
#pragma omp parallel for schedule(static)
    for (int r = 0; r < M; r++) {
        /* rA_rB_init_Ns_homo (GPU_kernels.h:92-101) */
        /* The reference starts rhA at 0 (GPU_kernels.h:96, "no A homopolymers"). The A
         * homopolymer here gets the same end terms as the B one, 2 q1[Nh]. Both are
         * multiplied by zA, which is 0 unless phihA is set. */
        double hA = 2.0 * Q1P(Nh)[r];
        double hB = 2.0 * Q2P(Nh)[r];
        double cA = Q2P(Ns)[r] + Q1P(n1)[r] * Q2P(n2)[r];
        double cB = Q1P(Ns)[r] + Q1P(n1)[r] * Q2P(n2)[r];
        int s, wt;
        /* homopolymer A (density.h:405-407): zA = 0 in the reference (A.9) */
        for (s = 1, wt = 4; s < Nh; s++, wt = 6 - wt) hA += wt * Q1P(s)[r] * Q1P(Nh - s)[r];
        for (s = 1, wt = 4; s < Nh; s++, wt = 6 - wt) hB += wt * Q2P(s)[r] * Q2P(Nh - s)[r];
        for (s = 1, wt = 4; s < n1; s++, wt = 6 - wt) cA += wt * Q1P(s)[r] * Q2P(Ns - s)[r];
        for (s = n1 + 1, wt = 4; s < Ns; s++, wt = 6 - wt) cB += wt * Q1P(s)[r] * Q2P(Ns - s)[r];
        /* scale_concentrations (GPU_kernels.h:129-139) */
        hA = hA * zA * ds / 3.0;
        hB = hB * zB * ds / 3.0;
        cA = cA * zC * ds / 3.0;
        cB = cB * zC * ds / 3.0;
        /* sumconcs2 (GPU_kernels.h:114-127) */
        rhA[r] = hA; rhB[r] = hB; rCA[r] = cA; rCB[r] = cB;
        rA[r] = hA + cA;
        rB[r] = hB + cB;
        rh[r] = hA + hB;
        rC[r] = cA + cB;
    }
    /* phis (density.h:433-454) */
    phis[1] = sum_all(rA, M) / M;
    phis[2] = sum_all(rB, M) / M;
    phis[0] = sum_all(rC, M) / M;
    phis[3] = sum_all(rhA, M) / M;
    phis[4] = sum_all(rhB, M) / M;
    if (qstore) memcpy(qstore, g->qq, sizeof(double) * need);
    free(rhA); free(rhB); free(rCA); free(rC);
#undef Q1P
#undef Q2P
}

/* ------------------------------------------------------------------------------------ */
/* A10: solve_field0 (scft.cu:516-554)                                                   */
/* ------------------------------------------------------------------------------------ */
int orc_solve_field0(orc_grid *g, double *W, double *DEV, double chi, double f, int N,
                     double alpha, double phic, int ens, int maxIter, double errTol)
{
    const int M = g->M;
    double lambda = 0.05, err = 1.0, lnQ, S1, phis[6];
    double errt = 1, dlam1 = 1.09, dlam2 = 1.10;
    int k;
    const double *phiA = W + 2 * M, *phiB = W + 3 * M;
    for (k = 1; k < maxIter && err > errTol; k++) {
        orc_props(g, W, &lnQ, phis, N, alpha, phic, f, ens, NULL, NULL);
        for (int r = 0; r < M; r++) {
            DEV[r] = 0;
            DEV[r + M] = chi * (phiB[r] + phiA[r] - 1.0);
        }
        S1 = 0.0;
        for (int r = 0; r < 2 * M; r++) S1 += DEV[r] * DEV[r];
        err = pow(S1 / (M), 0.5);
        for (int r = 0; r < 2 * M; r++) W[r] = W[r] + lambda * DEV[r];
        if (err < errt) lambda = fmin(0.1, lambda * dlam1);
        else lambda = fmax(0.005, lambda / dlam2);
        errt = err;
    }
    return k;
}

/* ------------------------------------------------------------------------------------ */
/* A13: natural cubic spline (GPU_kernels.h:232-329)                                     */
/* ------------------------------------------------------------------------------------ */
void orc_spline_coeffs(int strn, const double *x, const double *a, double *c, double *b,
                       double *d)
{
    const int n = strn - 1;
    double h[256], A[256], l[256], u[256], z[256];
    for (int i = 0; i < strn; i++) { b[i] = c[i] = d[i] = h[i] = A[i] = l[i] = u[i] = z[i] = 0.0; }
    for (int i = 0; i <= n - 1; ++i) h[i] = x[i + 1] - x[i];
    for (int i = 1; i <= n - 1; ++i)
        A[i] = 3.0 * (a[i + 1] - a[i]) / h[i] - 3.0 * (a[i] - a[i - 1]) / h[i - 1];
    l[0] = 1.0; u[0] = 0.0; z[0] = 0.0;
    for (int i = 1; i <= n - 1; ++i) {
        l[i] = 2.0 * (x[i + 1] - x[i - 1]) - h[i - 1] * u[i - 1];
        u[i] = h[i] / l[i];
        z[i] = (A[i] - h[i - 1] * z[i - 1]) / l[i];
    }
    l[n] = 1.0; z[n] = 0.0; c[n] = 0.0;
    for (int j = n - 1; j >= 0; --j) {
        c[j] = z[j] - u[j] * c[j + 1];
        b[j] = (a[j + 1] - a[j]) / h[j] - h[j] * (c[j + 1] + 2.0 * c[j]) / 3.0;
        d[j] = (c[j + 1] - c[j]) / (3.0 * h[j]);
    }
}

static int spline_interval(int strn, const double *x, double xx)
{
    int nm = 0;
    for (int i = 0; i < strn; ++i) {
        if (x[i] < xx) nm = i;
        else break;
    }
    return nm;
}

void orc_spline_deriv_nodes(int strn, const double *x, const double *a, double *out)
{
    double c[256], b[256], d[256];
    orc_spline_coeffs(strn, x, a, c, b, d);
    for (int e = 0; e < strn; ++e) {
        const double xx = x[e];
        const int nm = spline_interval(strn, x, xx);
        const double dx = xx - x[nm];
        out[e] = b[nm] + 2.0 * dx * c[nm] + 3.0 * dx * dx * d[nm];
    }
}

void orc_spline_value_at(int strn, const double *x, const double *a, const double *xr,
                         double *out)
{
    double c[256], b[256], d[256];
    orc_spline_coeffs(strn, x, a, c, b, d);
    for (int e = 0; e < strn; ++e) {
        const double xx = xr[e];
        const int nm = spline_interval(strn, x, xx);
        const double dx = xx - x[nm];
        out[e] = a[nm] + dx * b[nm] + dx * dx * c[nm] + dx * dx * dx * d[nm];
    }
}

/* scft.cu:351-355 */
double orc_edist(int M, const double *Wa, const double *Wb)
{
    double dist = 0;
    for (int r = 0; r < M; r++) dist += pow(Wa[r] - Wb[r], 2.0);
    return sqrt(dist);
}

/* ------------------------------------------------------------------------------------ */
/* A11-A15: one solve_field iteration, single process, doii = -1 (scft.cu:592-757)      */
/* ------------------------------------------------------------------------------------ */
static void arc_length(int strn, int M, double **W, double *x)
{
    double Del[256], Delsum = 0;
    /* the string metric is over W- only (scft.cu:351-355); mutant 3 adds W+ */
    const int Mmetric = ORC_MUTANT == 3 ? 2 * M : M;
    for (int ii = 0; ii < strn - 1; ii++) { Del[ii] = orc_edist(Mmetric, W[ii], W[ii + 1]); Delsum += Del[ii]; }
    for (int ii = 0; ii < strn - 1; ii++) Del[ii] /= Delsum;
    x[0] = 0; x[strn - 1] = 1;
    for (int ii = 1; ii < strn - 1; ii++) x[ii] = x[ii - 1] + Del[ii - 1];
}

void orc_string_iter(orc_grid *g, const orc_run *run, double **W, double **DEV, int *doiis,
                     int k, double errTol, double *alf, orc_iter_report *rep)
{
    const int M = g->M, strn = run->nrep;
    const double lambda = ORC_MUTANT == 1 ? 0.5 : 0.05, eps = ORC_MUTANT == 6 ? 1.0 : 0.5, chi = run->chi;
    const double lambda_redist = ORC_MUTANT == 8 ? 0.5 : lambda;
    const double errs_norm = ORC_MUTANT == 2 ? M : 2.0 * M;
    double lnQ, phis[6], x[256], xref[256];
    int dospring = 1, doredist = 0, doperp = 0;
    if (run->dostring != 1) dospring = 0;
    if (k >= run->springsteps && run->dostring == 1) { dospring = 0; doredist = 1; doperp = 1; }

    for (int ii = 0; ii < strn; ii++) memset(DEV[ii], 0, sizeof(double) * 2 * M);
    for (int ii = 0; ii < strn; ii++) {
        rep->steps[ii] = 0;
        if (doiis[ii] == 1) {
            rep->steps[ii]++;
            rep->steps[ii] += orc_solve_field0(g, W[ii], DEV[ii], chi, run->f, run->N, run->alpha,
                                               run->phidb, run->ens, 20, errTol);
            orc_props(g, W[ii], &lnQ, phis, run->N, run->alpha, run->phidb, run->f, run->ens, NULL, NULL);
            const double *phiA = W[ii] + 2 * M, *phiB = W[ii] + 3 * M;
            for (int r = 0; r < M; r++) {
                DEV[ii][r] = chi * (phiB[r] - phiA[r]) - 2.0 * W[ii][r];
                if (ii > 0 && ii < strn - 1 && dospring == 1)
                    DEV[ii][r] += eps * (W[ii + 1][r] + W[ii - 1][r] - 2.0 * W[ii][r]);
                DEV[ii][r + M] = chi * (phiB[r] + phiA[r] - 1.0);
            }
        }
    }
    arc_length(strn, M, W, x);
    for (int ii = 0; ii < strn; ii++) xref[ii] = (1.0 * ii) / (strn - 1.0);

    for (int ii = 0; ii < strn; ii++) {
        double S1 = 0.0;
        for (int r = 0; r < 2 * M; r++) S1 += DEV[ii][r] * DEV[ii][r];
        rep->errs[ii] = pow(S1 / errs_norm, 0.5);
    }

    for (int ii = 0; ii < strn; ii++) {
        rep->normDER[ii] = rep->normDEV[ii] = rep->dotWs[ii] = rep->dotWs2[ii] = 0;
    }
    if (doperp == 1) {
        /* dWda lives at W[ii] + 6M (scft.cu:1050) */
        for (int ii = 0; ii < strn; ii++)
            if (doiis[ii] == 1) memset(W[ii] + 6 * M, 0, sizeof(double) * M);
#pragma omp parallel
        {
            double col[256], der[256];
#pragma omp for schedule(static)
            for (int r = 0; r < M; r++) {
                for (int ii = 0; ii < strn; ii++) col[ii] = W[ii][r];
                orc_spline_deriv_nodes(strn, x, col, der);
                for (int ii = 0; ii < strn; ii++)
                    if (doiis[ii] == 1) W[ii][6 * M + r] = der[ii];
            }
        }
        for (int ii = 0; ii < strn; ii++) {
            double s = 0;
            if (doiis[ii] == 1) for (int r = 0; r < M; r++) s += W[ii][6 * M + r] * W[ii][6 * M + r];
            rep->normDER[ii] = sqrt(s);
        }
        for (int ii = 0; ii < strn; ii++) {
            double s = 0;
            if (doiis[ii] == 1) for (int r = 0; r < M; r++) s += DEV[ii][r] * DEV[ii][r];
            rep->normDEV[ii] = sqrt(s);
        }
        for (int ii = 0; ii < strn; ii++) {
            double s = 0;
            if (doiis[ii] == 1)
                for (int r = 0; r < M; r++)
                    s += (DEV[ii][r] / rep->normDEV[ii]) * (W[ii][6 * M + r] / rep->normDER[ii]);
            rep->dotWs[ii] = s;
        }
        for (int ii = 0; ii < strn; ii++) rep->errs3[ii] = 0;
        for (int ii = 1; ii < strn - 1; ii++) {
            if (doiis[ii] == 1 && rep->normDER[ii] > 0 && rep->normDEV[ii] > 0) {
                const double sgn = ORC_MUTANT == 7 ? -1.0 : 1.0;
                for (int r = 0; r < M; r++)
                    DEV[ii][r] -= sgn * (rep->dotWs[ii] / rep->normDER[ii]) * W[ii][6 * M + r] * rep->normDEV[ii];
                double s = 0;
                for (int r = 0; r < 2 * M; r++) s += DEV[ii][r] * DEV[ii][r];
                rep->errs3[ii] = pow(s / errs_norm, 0.5);
            }
        }
        for (int ii = 0; ii < strn; ii++) {
            double s = 0;
            if (doiis[ii] == 1)
                for (int r = 0; r < M; r++)
                    s += (DEV[ii][r] / rep->normDEV[ii]) * (W[ii][6 * M + r] / rep->normDER[ii]);
            rep->dotWs2[ii] = s;
        }
    } else {
        for (int ii = 0; ii < strn; ii++) rep->errs3[ii] = rep->errs[ii];
    }
    rep->errs3[0] = rep->errs[0];
    rep->errs3[strn - 1] = rep->errs[strn - 1];

    /* mixing with fixed lambda (scft.cu:690-693) */
    for (int ii = 0; ii < strn; ii++)
        for (int r = 0; r < 2 * M; r++) W[ii][r] = W[ii][r] + lambda * DEV[ii][r];

    /* arc length again (scft.cu:725-743) */
    arc_length(strn, M, W, x);
    for (int ii = 0; ii < strn; ii++) alf[ii] = x[ii];
    {
        double fitup = 0;
        for (int ii = 1; ii < strn - 1; ii++) fitup += pow(x[ii] - xref[ii], 2.0);
        fitup /= strn - 2.0;
        fitup = sqrt(fitup);
        fitup /= strn - 1.0;
        rep->fitup = fitup;
    }
    for (int ii = 0; ii < strn; ii++) { rep->x[ii] = x[ii]; rep->xref[ii] = xref[ii]; }

    /* redistribution (scft.cu:745-757; density.h:511-528) */
    if (doredist == 1 && ORC_MUTANT != 4) {
        const int i0 = ORC_MUTANT == 5 ? 0 : 1, i1 = ORC_MUTANT == 5 ? strn : strn - 1;
#pragma omp parallel
        {
            double col[256], val[256];
#pragma omp for schedule(static)
            for (int r = 0; r < M; r++) {
                for (int ii = 0; ii < strn; ii++) col[ii] = W[ii][r];
                orc_spline_value_at(strn, x, col, xref, val);
                for (int ii = i0; ii < i1; ii++)
                    if (doiis[ii] == 1) W[ii][r] += (val[ii] - W[ii][r]) * lambda_redist;
            }
        }
    }
}

/* ------------------------------------------------------------------------------------ */
/* A16: FreeE, single process (scft.cu:812-853)                                          */
/* ------------------------------------------------------------------------------------ */
double orc_free_energy(orc_grid *g, const orc_run *run, double **W, double **DEV,
                       const int *doiis, double *alf, double *FEs0, double *FEs1, double *FEs,
                       double *phctot)
{
    const int M = g->M, strn = run->nrep;
    const double chi = run->chi, f = run->f;
    double FE = 0, lnQ, phis[6];
    alf[0] = 0;
    for (int ii = 0; ii < strn; ii++) {
        if (doiis[ii] == 1 || ii == 0 || ii == strn - 1) {
            orc_solve_field0(g, W[ii], DEV[ii], chi, f, run->N, run->alpha, run->phidb, run->ens, 100, 1e-4);
            orc_props(g, W[ii], &lnQ, phis, run->N, run->alpha, run->phidb, f, run->ens, NULL, NULL);
            FEs0[ii] = -lnQ;
            FEs1[ii] = 0;
            phctot[ii] = 0;
            const double *phiA = W[ii] + 2 * M, *phiB = W[ii] + 3 * M;
            for (int r = 0; r < M; r++) {
                const double wa = W[ii][r] + W[ii][r + M], wb = W[ii][r + M] - W[ii][r];
                const double sa = ORC_MUTANT == 9 ? -1.0 : 1.0;
                FEs1[ii] += (chi * phiA[r] * phiB[r] - sa * wa * phiA[r] - wb * phiB[r]) / M;
                phctot[ii] += phiA[r] / (M * f);
            }
            FEs[ii] = FEs0[ii] + FEs1[ii];
        }
        if (ii > 0) alf[ii] = alf[ii - 1] + orc_edist(M, W[ii], W[ii - 1]);
    }
    for (int ii = 0; ii < strn; ii++) FE += FEs[ii];
    return FE;
}

/* ------------------------------------------------------------------------------------ */
/* Box derivative of lnQ at fixed W (the adjoint of density.h:344-399)                  */
/* ------------------------------------------------------------------------------------ */
/* The step operator S = (4 E4 K12 E2 K12 E4 - E2 K6 E2)/3 (density.h:129-220) is symmetric,
 * so the adjoint of q1 after step s is q2[N-1-s] and
 *   dQC/dD_i = sum_s <q2[N-1-s], dS_s/dD_i q1[s]>,  <f, g> = mean(f g).
 * With v0 = E4 b, u0 = E4 a, v1 = E2 K12 v0, u1 = E2 K12 u0:
 *   <a, dS b> = (4 (<u0, dK12 v1> + <u1, dK12 v0>) - <E2 a, dK6 E2 b>) / 3,
 * and, from k^2 = sum_j (2 pi n_j / D_j)^2 (density.h:79-95),
 *   dK12/dD_i = K12 * k_i^2 ds / (6 D_i),  dK6/dD_i = K6 * k_i^2 ds / (3 D_i).
 * Each <u, dK v> is taken in real space: dK applied by an FFT sandwich, then a mean of
 * products. Q2 = <q2[Nh]> uses q2 itself as its adjoint the same way. */
static void kcomp(const orc_grid *g, int dir, double *out)
{
    const int *m = g->m;
    const double pi = 4 * atan(1.0);
    for (int k0 = -(m[0] - 1) / 2; k0 <= m[0] / 2; k0++) {
        const int K0 = (k0 < 0) ? (k0 + m[0]) : k0;
        for (int k1 = -(m[1] - 1) / 2; k1 <= m[1] / 2; k1++) {
            const int K1 = (k1 < 0) ? (k1 + m[1]) : k1;
            for (int k2 = 0; k2 <= m[2] / 2; k2++) {
                const int n = dir == 0 ? k0 : (dir == 1 ? k1 : k2);
                const int k = k2 + (m[2] / 2 + 1) * (K1 + m[1] * K0);
                out[k] = 4 * pi * pi * (double)n * n / (g->D[dir] * g->D[dir]);
            }
        }
    }
}

static double mean_prod(const double *a, const double *b, int M)
{
    double s = 0;
#pragma omp parallel for reduction(+ : s) schedule(static)
    for (int r = 0; r < M; r++) s += a[r] * b[r];
    return s / M;
}

/* adds <a, dS/dD_i b> for i = 0..2 to acc */
static void grad_term(orc_grid *g, const double *e4, const double *e2, double *const dk12[3],
                      double *const dk6[3], const double *a, const double *b, double acc[3])
{
    const int M = g->M;
    double *v0 = malloc(sizeof(double) * M), *u0 = malloc(sizeof(double) * M);
    double *v1 = malloc(sizeof(double) * M), *u1 = malloc(sizeof(double) * M);
    double *vf = malloc(sizeof(double) * M), *uf = malloc(sizeof(double) * M);
    double *t = malloc(sizeof(double) * M);
    for (int r = 0; r < M; r++) {
        v0[r] = e4[r] * b[r];
        u0[r] = e4[r] * a[r];
        vf[r] = e2[r] * b[r];
        uf[r] = e2[r] * a[r];
    }
    memcpy(v1, v0, sizeof(double) * M);
    memcpy(u1, u0, sizeof(double) * M);
    sandwich(g, v1, g->k12);
    sandwich(g, u1, g->k12);
    for (int r = 0; r < M; r++) { v1[r] *= e2[r]; u1[r] *= e2[r]; }
    for (int i = 0; i < 3; i++) {
        double th = 0, tf;
        memcpy(t, v1, sizeof(double) * M);
        sandwich(g, t, dk12[i]);
        th += mean_prod(u0, t, M);
        memcpy(t, v0, sizeof(double) * M);
        sandwich(g, t, dk12[i]);
        th += mean_prod(u1, t, M);
        memcpy(t, vf, sizeof(double) * M);
        sandwich(g, t, dk6[i]);
        tf = mean_prod(uf, t, M);
        acc[i] += (4 * th - tf) / 3;
    }
    free(v0); free(u0); free(v1); free(u1); free(vf); free(uf); free(t);
}

void orc_dlnq_dbox(orc_grid *g, double *w, int Ns, double alpha, double phic, double f, int ens,
                   double *lnQ, double dlnQ[3])
{
    const int M = g->M, Mk = g->Mk;
    const int n1 = (int)round(Ns * f), n2 = Ns - n1;
    const int Nh = (int)(alpha * Ns); /* density.h:346 */
    const double ds = 1.0 / Ns;
    double Q[3], phis[5];
    orc_props(g, w, lnQ, phis, Ns, alpha, phic, f, ens, Q, NULL);
    const double Q1 = Q[0], Q2 = Q[1], QC = Q[2];
#define Q1P(s) (g->qq + (size_t)2 * (s) * M)
#define Q2P(s) (g->qq + ((size_t)2 * (s) + 1) * M)
    double *dk12[3], *dk6[3], *kc = malloc(sizeof(double) * Mk);
    for (int i = 0; i < 3; i++) {
        dk12[i] = malloc(sizeof(double) * Mk);
        dk6[i] = malloc(sizeof(double) * Mk);
        kcomp(g, i, kc);
        for (int k = 0; k < Mk; k++) {
            dk12[i][k] = g->k12[k] * kc[k] * ds / (6 * g->D[i]);
            dk6[i][k] = g->k6[k] * kc[k] * ds / (3 * g->D[i]);
        }
    }
    double dQC[3] = {0, 0, 0}, dQ2[3] = {0, 0, 0}, dQ1[3] = {0, 0, 0};
    for (int s = 0; s < Ns; s++) {
        const int isA = s < n1; /* species of chain 1 on step s (density.h:136-146) */
        grad_term(g, isA ? g->e4A : g->e4B, isA ? g->e2A : g->e2B, dk12, dk6, Q2P(Ns - 1 - s),
                  Q1P(s), dQC);
    }
    for (int t = 0; t < Nh; t++) {
        const int isB = t < n2; /* species of chain 2 on step t */
        grad_term(g, isB ? g->e4B : g->e4A, isB ? g->e2B : g->e2A, dk12, dk6, Q2P(Nh - 1 - t),
                  Q2P(t), dQ2);
    }
    if (g->phihA != 0.0)
        for (int t = 0; t < Nh; t++) {
            const int isA = t < n1; /* the A homopolymer's steps: q1 up to Nh */
            grad_term(g, isA ? g->e4A : g->e4B, isA ? g->e2A : g->e2B, dk12, dk6, Q1P(Nh - 1 - t),
                      Q1P(t), dQ1);
        }
    for (int i = 0; i < 3; i++) {
        if (ens == 1) {
            dlnQ[i] = phic * dQC[i] / QC + ((1.0 - phic - g->phihA) / alpha) * dQ2[i] / Q2;
            if (g->phihA != 0.0) dlnQ[i] += (g->phihA / alpha) * dQ1[i] / Q1;
        } else {
            dlnQ[i] = dQ2[i] + phic * dQC[i] + g->phihA * dQ1[i];
        }
        free(dk12[i]);
        free(dk6[i]);
    }
    free(kc);
#undef Q1P
#undef Q2P
}

/* ------------------------------------------------------------------------------------ */
/* F3: Anderson mixing, an opt-in mode for dostring = 0 (the reference keeps only its    */
/* arrays: DIM, DDEV, WIN, scft.cu:30-31, 1029-1031)                                      */
/* ------------------------------------------------------------------------------------ */
struct orc_anderson {
    int M2, H, count, head, started;
    double *Wr, *Dr; /* [(H+1)][2M] rings */
    double *G;       /* [(H+1)^2] sums of D_i D_j over 2M */
};

orc_anderson *orc_anderson_create(int M, int H)
{
    orc_anderson *a = (orc_anderson *)calloc(1, sizeof(orc_anderson));
    a->M2 = 2 * M;
    a->H = H;
    a->Wr = (double *)calloc((size_t)(H + 1) * a->M2, sizeof(double));
    a->Dr = (double *)calloc((size_t)(H + 1) * a->M2, sizeof(double));
    a->G = (double *)calloc((size_t)(H + 1) * (H + 1), sizeof(double));
    return a;
}

void orc_anderson_destroy(orc_anderson *a)
{
    if (!a) return;
    free(a->Wr); free(a->Dr); free(a->G); free(a);
}

/* Solves U c = V (n x n, row-major) by Gaussian elimination with partial pivoting.
 * Returns 0 when a pivot falls below 1e-14 of the largest |U| entry. */
int orc_small_solve(int n, double *U, double *V, double *c)
{
    double umax = 0;
    for (int i = 0; i < n * n; i++) umax = fmax(umax, fabs(U[i]));
    for (int col = 0; col < n; col++) {
        int piv = col;
        for (int r = col + 1; r < n; r++)
            if (fabs(U[r * n + col]) > fabs(U[piv * n + col])) piv = r;
        if (!(fabs(U[piv * n + col]) > 1e-14 * umax)) return 0;
        if (piv != col) {
            for (int j = 0; j < n; j++) {
                const double t = U[col * n + j]; U[col * n + j] = U[piv * n + j]; U[piv * n + j] = t;
            }
            const double t = V[col]; V[col] = V[piv]; V[piv] = t;
        }
        for (int r = col + 1; r < n; r++) {
            const double fct = U[r * n + col] / U[col * n + col];
            for (int j = col; j < n; j++) U[r * n + j] -= fct * U[col * n + j];
            V[r] -= fct * V[col];
        }
    }
    for (int r = n - 1; r >= 0; r--) {
        double s = V[r];
        for (int j = r + 1; j < n; j++) s -= U[r * n + j] * c[j];
        c[r] = s / U[r * n + r];
    }
    return 1;
}

/* One iteration on one replica row W (7M; W-, W+ updated, densities written by props).
 * DEV = (chi(phiB - phiA) - 2W-, chi(phiA + phiB - 1)) as in solve_field (scft.cu:595-639),
 * errs = sqrt(sum DEV^2 / 2M). Until errs < start: the reference's simple mixing
 * W += 0.05 DEV (scft.cu:690-693). From then on, Anderson over the last H residuals:
 * with d_k the newest residual and d_{k-m} the older ones,
 *   U_mn = <d_k - d_{k-m}, d_k - d_{k-n}>,  V_m = <d_k - d_{k-m}, d_k>,  U c = V,
 *   W <- sum_j a_j (W_j + lambda d_j),  a_k = 1 - sum c_m,  a_{k-m} = c_m.
 * A singular U drops to the newest pair alone (W_k + lambda d_k). Returns errs. */
double orc_anderson_iter(orc_grid *g, double *W, orc_anderson *a, double chi, double f, int N,
                         double alpha, double phic, int ens, double lambda, double start)
{
    const int M = g->M, M2 = 2 * M, H = a->H;
    double lnQ, phis[6];
    orc_props(g, W, &lnQ, phis, N, alpha, phic, f, ens, NULL, NULL);
    const double *phiA = W + 2 * M, *phiB = W + 3 * M;
    double *d = (double *)malloc(sizeof(double) * M2);
    double S = 0;
    for (int r = 0; r < M; r++) {
        d[r] = chi * (phiB[r] - phiA[r]) - 2.0 * W[r];
        d[r + M] = chi * (phiB[r] + phiA[r] - 1.0);
    }
    for (int r = 0; r < M2; r++) S += d[r] * d[r];
    const double errs = sqrt(S / M2);
    if (!a->started && errs >= start) {
        for (int r = 0; r < M2; r++) W[r] += 0.05 * d[r];
        free(d);
        return errs;
    }
    a->started = 1;
    const int s = a->head, L = H + 1;
    memcpy(a->Wr + (size_t)s * M2, W, sizeof(double) * M2);
    memcpy(a->Dr + (size_t)s * M2, d, sizeof(double) * M2);
    if (a->count < L) a->count++;
    for (int m = 0; m < a->count; m++) {
        const int j = (s - m + L) % L;
        const double *Dj = a->Dr + (size_t)j * M2;
        double gsum = 0;
        for (int r = 0; r < M2; r++) gsum += d[r] * Dj[r];
        a->G[s * L + j] = a->G[j * L + s] = gsum;
    }
    int n = a->count - 1;
    double coef[64];
    if (n > 0) {
        double U[64 * 64], V[64], c[64];
        for (int m = 1; m <= n; m++) {
            const int jm = (s - m + L) % L;
            V[m - 1] = a->G[s * L + s] - a->G[jm * L + s];
            for (int q = 1; q <= n; q++) {
                const int jq = (s - q + L) % L;
                U[(m - 1) * n + (q - 1)] = a->G[s * L + s] - a->G[s * L + jq] - a->G[jm * L + s] + a->G[jm * L + jq];
            }
        }
        if (!orc_small_solve(n, U, V, c)) n = 0;
        for (int m = 0; m < n; m++) coef[m] = c[m];
    }
    double a0 = 1.0;
    for (int m = 0; m < n; m++) a0 -= coef[m];
    for (int r = 0; r < M2; r++) {
        double v = a0 * (a->Wr[(size_t)s * M2 + r] + lambda * a->Dr[(size_t)s * M2 + r]);
        for (int m = 1; m <= n; m++) {
            const int jm = (s - m + L) % L;
            v += coef[m - 1] * (a->Wr[(size_t)jm * M2 + r] + lambda * a->Dr[(size_t)jm * M2 + r]);
        }
        W[r] = v;
    }
    a->head = (s + 1) % L;
    free(d);
    return errs;
}
