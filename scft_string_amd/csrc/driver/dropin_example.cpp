// dropin_example.cpp -- a caller written the way scft.cu is: host rows W[ii] of 7M doubles
// (W-, W+, phiA, phiB, phih, phicB, dW-/dalpha; scft.cu:1037-1051), the field update in host
// loops, and the device work behind class density's methods only:
//   D2->props (density.h:344; called at scft.cu:526, 603),
//   D2->setupSplineBuffers, setSplineX, setSplineXRef (density.h:462-488; scft.cu:643, 746-747, 1011),
//   D2->splineDerivsChunked, splineRedistChunked (density.h:491-528; scft.cu:644-647, 748-751).
// Here density is scft_host::density (scft_host.hpp), so this shows the drop-in at the
// reference's own call sites: swapping the class is the only change such a caller needs.
//
//   scft_dropin_example <config> <W_in.bin> <W_out.bin>
// config: "m0 m1 m2 D0 D1 D2 N f alpha ens phidb chi nrep fix0 fix1 springsteps iters chunk"
// W_in / W_out: nrep rows of 2M doubles (W-, W+). Prints one line per iteration:
// "k errs[0..nrep) errs3[0..nrep) x[0..nrep)".
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <vector>

#include "scft_host.hpp"

namespace {

struct Run {
    int M = 0, R = 0, N = 0, ens = 1, springsteps = 0;
    double f = 0, alpha = 0, phidb = 0, chi = 0;
    int chunk = 0;
};

// solve_field0 (scft.cu:516-554) on replica ii
int relax_w_plus(double** W, double** DEV, scft_host::density* D2, const Run& run, int ii, int maxIter,
                 double errTol)
{
    const int M = run.M;
    double lambda = 0.05, err = 1.0, errt = 1.0, lnQ, phis[5];
    int k;
    for (k = 1; k < maxIter && err > errTol; k++) {
        D2->props(W[ii], &lnQ, ii, run.N, run.alpha, run.phidb, phis, run.f);
        const double *phiA = W[ii] + 2 * M, *phiB = W[ii] + 3 * M;
        double s = 0;
        for (int r = 0; r < M; r++) {
            DEV[ii][r] = 0;
            DEV[ii][r + M] = run.chi * (phiB[r] + phiA[r] - 1.0);
            s += DEV[ii][r + M] * DEV[ii][r + M];
        }
        err = std::sqrt(s / M);
        for (int r = 0; r < 2 * M; r++) W[ii][r] += lambda * DEV[ii][r];
        lambda = err < errt ? std::min(0.1, lambda * 1.09) : std::max(0.005, lambda / 1.10);
        errt = err;
    }
    return k;
}

// normalised cumulative W- distance (scft.cu:351-355, 612-622)
void positions(double** W, const Run& run, std::vector<double>& x)
{
    const int R = run.R, M = run.M;
    std::vector<double> del(R - 1);
    double sum = 0;
    for (int ii = 0; ii + 1 < R; ii++) {
        double d = 0;
        for (int r = 0; r < M; r++) d += (W[ii + 1][r] - W[ii][r]) * (W[ii + 1][r] - W[ii][r]);
        del[ii] = std::sqrt(d);
        sum += del[ii];
    }
    x.assign(R, 0.0);
    x[R - 1] = 1.0;
    for (int ii = 1; ii + 1 < R; ii++) x[ii] = x[ii - 1] + del[ii - 1] / sum;
}

// one iteration k of solve_field's body (scft.cu:592-757), single process
void iterate(double** W, double** DEV, scft_host::density* D2, const Run& run, int* doiis, int k, double errTol,
             std::vector<double>& errs, std::vector<double>& errs3, std::vector<double>& x)
{
    const int R = run.R, M = run.M;
    const double lambda = 0.05, eps = 0.5;
    const bool spring = k < run.springsteps;
    double lnQ, phis[5];
    for (int ii = 0; ii < R; ii++)
        for (int r = 0; r < 2 * M; r++) DEV[ii][r] = 0;
    for (int ii = 0; ii < R; ii++) {
        if (doiis[ii] != 1) continue;
        relax_w_plus(W, DEV, D2, run, ii, 20, errTol);
        D2->props(W[ii], &lnQ, ii, run.N, run.alpha, run.phidb, phis, run.f);
        const double *phiA = W[ii] + 2 * M, *phiB = W[ii] + 3 * M;
        for (int r = 0; r < M; r++) {
            DEV[ii][r] = run.chi * (phiB[r] - phiA[r]) - 2.0 * W[ii][r];
            if (spring && ii > 0 && ii < R - 1) DEV[ii][r] += eps * (W[ii + 1][r] + W[ii - 1][r] - 2.0 * W[ii][r]);
            DEV[ii][r + M] = run.chi * (phiB[r] + phiA[r] - 1.0);
        }
    }
    positions(W, run, x);
    std::vector<double> xref(R);
    for (int ii = 0; ii < R; ii++) xref[ii] = ii / (R - 1.0);
    errs.assign(R, 0.0);
    errs3.assign(R, 0.0);
    for (int ii = 0; ii < R; ii++) {
        double s = 0;
        for (int r = 0; r < 2 * M; r++) s += DEV[ii][r] * DEV[ii][r];
        errs[ii] = std::sqrt(s / (2.0 * M));
    }
    if (!spring) {
        std::vector<double*> dWda(R);
        for (int ii = 0; ii < R; ii++) dWda[ii] = W[ii] + 6 * M;
        for (int ii = 0; ii < R; ii++)
            if (doiis[ii] == 1) std::fill(dWda[ii], dWda[ii] + M, 0.0);
        D2->setSplineX(x.data());
        for (int r0 = 0; r0 < M; r0 += run.chunk)
            D2->splineDerivsChunked(W, x.data(), r0, std::min(run.chunk, M - r0), doiis, dWda.data());
        for (int ii = 1; ii < R - 1; ii++) {
            if (doiis[ii] != 1) continue;
            double nder = 0, ndev = 0, dot = 0;
            for (int r = 0; r < M; r++) {
                nder += dWda[ii][r] * dWda[ii][r];
                ndev += DEV[ii][r] * DEV[ii][r];
            }
            nder = std::sqrt(nder);
            ndev = std::sqrt(ndev);
            if (!(nder > 0 && ndev > 0)) continue;
            for (int r = 0; r < M; r++) dot += (DEV[ii][r] / ndev) * (dWda[ii][r] / nder);
            for (int r = 0; r < M; r++) DEV[ii][r] -= (dot / nder) * dWda[ii][r] * ndev;
            double s = 0;
            for (int r = 0; r < 2 * M; r++) s += DEV[ii][r] * DEV[ii][r];
            errs3[ii] = std::sqrt(s / (2.0 * M));
        }
    } else {
        errs3 = errs;
    }
    errs3[0] = errs[0];
    errs3[R - 1] = errs[R - 1];
    for (int ii = 0; ii < R; ii++)
        for (int r = 0; r < 2 * M; r++) W[ii][r] += lambda * DEV[ii][r];
    positions(W, run, x);
    if (!spring) {
        D2->setSplineX(x.data());
        D2->setSplineXRef(xref.data());
        for (int r0 = 0; r0 < M; r0 += run.chunk)
            D2->splineRedistChunked(W, x.data(), xref.data(), r0, std::min(run.chunk, M - r0), doiis, lambda);
    }
}

}  // namespace

int main(int argc, char** argv)
{
    if (argc != 4) {
        fprintf(stderr, "usage: %s config W_in.bin W_out.bin\n", argv[0]);
        return 1;
    }
    scft_params p{};
    Run run;
    int iters = 0;
    FILE* cf = fopen(argv[1], "r");
    if (!cf || fscanf(cf, "%d %d %d %lf %lf %lf %d %lf %lf %d %lf %lf %d %d %d %d %d %d", &p.m[0], &p.m[1], &p.m[2],
                      &p.D[0], &p.D[1], &p.D[2], &p.N, &p.f, &p.alpha, &p.ens, &p.phidb, &p.chi, &p.nrep, &p.fix[0],
                      &p.fix[1], &p.springsteps, &iters, &run.chunk) != 18) {
        fprintf(stderr, "bad config %s\n", argv[1]);
        return 1;
    }
    fclose(cf);
    p.precision = SCFT_FP64;
    p.dostring = 1;
    p.fe_fp64 = 1;
    run.M = p.m[0] * p.m[1] * p.m[2];
    run.R = p.nrep;
    run.N = p.N;
    run.f = p.f;
    run.alpha = p.alpha;
    run.ens = p.ens;
    run.phidb = p.phidb;
    run.chi = p.chi;
    run.springsteps = p.springsteps;
    const int M = run.M, R = run.R;
    std::vector<std::vector<double>> Wrows(R, std::vector<double>(7 * (size_t)M, 0.0)),
        DEVrows(R, std::vector<double>(2 * (size_t)M, 0.0));
    std::vector<double*> W(R), DEV(R);
    FILE* in = fopen(argv[2], "rb");
    for (int ii = 0; ii < R; ii++) {
        W[ii] = Wrows[ii].data();
        DEV[ii] = DEVrows[ii].data();
        if (!in || fread(W[ii], sizeof(double), 2 * (size_t)M, in) != 2 * (size_t)M) {
            fprintf(stderr, "short input %s\n", argv[2]);
            return 1;
        }
    }
    fclose(in);
    try {
        scft_host::density D2(p);
        D2.setupSplineBuffers(run.chunk);
        std::vector<int> doiis(R, 1);
        if (p.fix[0] == 1) doiis[0] = 0;
        if (p.fix[1] == 1) doiis[R - 1] = 0;
        std::vector<double> errs, errs3, x;
        for (int k = 0; k < iters; k++) {
            iterate(W.data(), DEV.data(), &D2, run, doiis.data(), k, 1e-3, errs, errs3, x);
            printf("%d", k);
            for (double v : errs) printf(" %.17g", v);
            for (double v : errs3) printf(" %.17g", v);
            for (double v : x) printf(" %.17g", v);
            printf("\n");
        }
    } catch (const scft_host::error& e) {
        fprintf(stderr, "scft_dropin_example: %s\n", e.what());
        return 2;
    }
    FILE* out = fopen(argv[3], "wb");
    for (int ii = 0; ii < R; ii++) fwrite(W[ii], sizeof(double), 2 * (size_t)M, out);
    fclose(out);
    return 0;
}
