// scft_api.hip -- C-ABI implementation: context, batched props, and the host control flow
// of solve_field0 / solve_field / FreeE with all field arithmetic on the device (FP64).
#include "scft_api.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <thread>
#include <vector>

#include "field_kernels.hpp"
#include "prop_engine.hpp"

using scft::ChainConst;
using scft::PropEngine;
using scft::SCAL_N;

struct scft_ctx {
    scft_params p{};
    long long M = 0;
    int R = 0;
    double f_used = 0, ds = 0;
    ChainConst cc{};
    hipStream_t st = nullptr;
    std::string err;
    std::string info;  // scft_engine_info
    std::string tuned;  // the autotuner's candidate timings when autotune = 1 chose the schedule
    std::unique_ptr<PropEngine<float>> pe32;
    std::unique_ptr<PropEngine<double>> pe64;
    // FP32 contexts: FP64 engine for the props that FreeE takes F from (created on first use)
    std::unique_ptr<PropEngine<double>> pe64_fe;
    // field state
    double *dW = nullptr, *dphi = nullptr, *dDEV = nullptr, *ddWda = nullptr;
    // scratch rows for scft_props_batch
    double *dWs = nullptr, *dphis_rows = nullptr;
    // reductions / small arrays
    double* dpart = nullptr;
    int nblk = 0;
    double *dsm = nullptr, *dlam = nullptr, *dcproj = nullptr, *dphisv = nullptr;
    int *dactive = nullptr, *dslot = nullptr;
    // pinned host staging
    double* hsm = nullptr;  // [R*SCAL_N]
    int* hslot = nullptr;   // [R]
    double* hsmall = nullptr;  // [4R] x, xref, lam, cproj
    // spline weights of one launch (spline_weights): device [R][R] and pinned staging
    double *dG = nullptr, *hG = nullptr;
    // chunked host-row spline calls (density.h:462-528): staging of [R][spl_chunk] grid
    // points (device in / out, pinned host), nodes and evaluation points
    int spl_chunk = 0;
    double *dspl_in = nullptr, *dspl_out = nullptr, *hspl = nullptr;
    std::vector<double> spl_x, spl_xref;
    int* hact = nullptr;    // [R]
    // Anderson mixing (opt-in, dostring = 0): per replica rings of the last H+1 (W, DEV)
    // pairs [R][H+1][2M], the Gram sums of their DEVs on the host, and small staging
    struct Anderson {
        int count = 0, head = 0;
        bool started = false;
        std::vector<double> G;  // [(H+1)^2]
    };
    std::vector<Anderson> and_;
    double *dAW = nullptr, *dAD = nullptr, *dApart = nullptr, *dAcoef = nullptr;
    int* dArows = nullptr;
    double* hAbuf = nullptr;  // [(H+1) (nblk + 1)] partials and coefficients
    int* hArows = nullptr;    // [H+1]
};

namespace {

scft_status fail(scft_ctx* c, scft_status s, const std::string& msg)
{
    if (c) c->err = msg;
    return s;
}

#define HIPCHK(ctx, expr)                                                                       \
    do {                                                                                        \
        hipError_t e_ = (expr);                                                                 \
        if (e_ != hipSuccess)                                                                   \
            return fail((ctx), SCFT_ERR_HIP, std::string(#expr ": ") + hipGetErrorString(e_)); \
    } while (0)

#define SCFTCHK(expr)                         \
    do {                                      \
        scft_status s_ = (expr);              \
        if (s_ != SCFT_OK) return s_;         \
    } while (0)

// props for the slots listed in h_slots (replica indices into W/phi rows). fe64: use the
// FP64 free-energy engine of an FP32 context.
scft_status run_props(scft_ctx* c, const double* dW, long long wstride, double* dphi, long long phistride,
                      const std::vector<int>& slots, double* lnQ_host, double* phis_host, bool fe64 = false)
{
    const int nb = (int)slots.size();
    if (nb == 0) return SCFT_OK;
    for (int i = 0; i < nb; ++i) c->hslot[i] = slots[i];
    HIPCHK(c, hipMemcpyAsync(c->dslot, c->hslot, sizeof(int) * nb, hipMemcpyHostToDevice, c->st));
    PropEngine<double>* p64 = (fe64 && c->pe64_fe) ? c->pe64_fe.get() : c->pe64.get();
    const bool use32 = c->pe32 && !(fe64 && c->pe64_fe);
    hipError_t e = use32 ? c->pe32->run(dW, wstride, dphi, phistride, c->dslot, nb, c->cc, c->dphisv, c->st)
                         : p64->run(dW, wstride, dphi, phistride, c->dslot, nb, c->cc, c->dphisv, c->st);
    if (e != hipSuccess) return fail(c, SCFT_ERR_HIP, std::string("propagator launch: ") + hipGetErrorString(e));
    if (lnQ_host || phis_host) {
        const double* scal = use32 ? c->pe32->scal() : p64->scal();
        HIPCHK(c, hipMemcpyAsync(c->hsm, scal, sizeof(double) * SCAL_N * nb, hipMemcpyDeviceToHost, c->st));
        std::vector<double> ph(5 * (size_t)nb);
        if (phis_host)
            HIPCHK(c, hipMemcpyAsync(ph.data(), c->dphisv, sizeof(double) * 5 * nb, hipMemcpyDeviceToHost, c->st));
        HIPCHK(c, hipStreamSynchronize(c->st));
        for (int i = 0; i < nb; ++i) {
            if (lnQ_host) lnQ_host[i] = c->hsm[SCAL_N * i + 0];
            if (phis_host)
                for (int j = 0; j < 5; ++j) phis_host[5 * i + j] = ph[5 * i + j];
        }
    }
    return SCFT_OK;
}

// finalize partial sums [nslots][nblk][ncomp] and bring them to host (hsm)
scft_status finalize_to_host(scft_ctx* c, int nslots, int ncomp, std::vector<double>& out)
{
    out.assign((size_t)nslots * ncomp, 0.0);
    if (nslots == 0) return SCFT_OK;
    scft::launch_finalize(c->dpart, nslots, c->nblk, ncomp, c->dsm, c->st);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipMemcpyAsync(c->hsm, c->dsm, sizeof(double) * nslots * ncomp, hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    std::memcpy(out.data(), c->hsm, sizeof(double) * nslots * ncomp);
    return SCFT_OK;
}

// HIP's current device belongs to the host thread. A context may be created on one thread
// and used from others (scft_iterate_concurrent runs each context on its own thread), so
// every entry point makes the context's device current before it touches the context's
// memory or streams.
scft_status bind_device(scft_ctx* c)
{
    int d = -1;
    if (hipGetDevice(&d) == hipSuccess && d == c->p.device) return SCFT_OK;
    const hipError_t e = hipSetDevice(c->p.device);
    if (e != hipSuccess) return fail(c, SCFT_ERR_HIP, std::string("hipSetDevice: ") + hipGetErrorString(e));
    return SCFT_OK;
}

// solve_field kills the run when its error norm is NaN (assert(err==err), scft.cu:770);
// here the iteration returns SCFT_ERR_NAN before the fields are mixed.
scft_status check_errs_finite(scft_ctx* c, const double* errs, int R)
{
    for (int ii = 0; ii < R; ++ii)
        if (!(errs[ii] == errs[ii])) {
            char b[128];
            snprintf(b, sizeof b, "solve_field: error norm is NaN for replica %d", ii);
            return fail(c, SCFT_ERR_NAN, b);
        }
    return SCFT_OK;
}

std::vector<int> fixed_doiis(const scft_ctx* c)
{
    std::vector<int> d(c->R, 1);  // single process owns every replica (scft.cu:952-975)
    if (c->p.fix[0] == 1) d[0] = 0;  // scft.cu:587-588
    if (c->p.fix[1] == 1) d[c->R - 1] = 0;
    return d;
}

// squared neighbour distances over W- (scft.cu:351-355) -> host del2[R-1]
scft_status edist2(scft_ctx* c, std::vector<double>& del2)
{
    del2.assign(c->R > 1 ? c->R - 1 : 0, 0.0);
    if (c->R < 2) return SCFT_OK;
    scft::launch_edist(c->R, c->M, c->dW, c->dpart, c->nblk, c->st);
    HIPCHK(c, hipGetLastError());
    return finalize_to_host(c, c->R - 1, 1, del2);
}

// Anderson histories describe the fields the iteration produced; any outside change of the
// fields or the box starts them again.
void anderson_reset(scft_ctx* c)
{
    for (auto& a : c->and_) {
        a.count = a.head = 0;
        a.started = false;
    }
}

// Solves U c = V (n x n, row-major) by Gaussian elimination with partial pivoting; false
// when a pivot falls below 1e-14 of the largest |U| entry.
bool small_solve(int n, std::vector<double>& U, std::vector<double>& V, std::vector<double>& x)
{
    double umax = 0;
    for (double u : U) umax = std::max(umax, std::fabs(u));
    for (int col = 0; col < n; ++col) {
        int piv = col;
        for (int r = col + 1; r < n; ++r)
            if (std::fabs(U[r * n + col]) > std::fabs(U[piv * n + col])) piv = r;
        if (!(std::fabs(U[piv * n + col]) > 1e-14 * umax)) return false;
        if (piv != col) {
            for (int j = 0; j < n; ++j) std::swap(U[col * n + j], U[piv * n + j]);
            std::swap(V[col], V[piv]);
        }
        for (int r = col + 1; r < n; ++r) {
            const double fct = U[r * n + col] / U[col * n + col];
            for (int j = col; j < n; ++j) U[r * n + j] -= fct * U[col * n + j];
            V[r] -= fct * V[col];
        }
    }
    x.assign(n, 0.0);
    for (int r = n - 1; r >= 0; --r) {
        double s = V[r];
        for (int j = r + 1; j < n; ++j) s -= U[r * n + j] * x[j];
        x[r] = s / U[r * n + r];
    }
    return true;
}

// scft.cu:612-621 / 725-736: normalised arc-length positions from W- distances
void arc_positions(int R, const std::vector<double>& del2, std::vector<double>& x)
{
    x.assign(R, 0.0);
    if (R < 2) return;
    std::vector<double> Del(R - 1);
    double Delsum = 0;
    for (int i = 0; i < R - 1; ++i) {
        Del[i] = std::sqrt(del2[i]);
        Delsum += Del[i];
    }
    for (int i = 0; i < R - 1; ++i) Del[i] /= Delsum;
    x[0] = 0;
    x[R - 1] = 1;
    for (int i = 1; i < R - 1; ++i) x[i] = x[i - 1] + Del[i - 1];
}

// One iteration of the opt-in Anderson mode (dostring = 0). Every active replica gets a
// props and its residual DEV = (chi(phiB - phiA) - 2W-, chi(phiA + phiB - 1)) (scft.cu:595-639
// without the spring). Then each replica takes the reference's simple mixing W += 0.05 DEV
// (scft.cu:690-693) while errs >= start, and Anderson mixing after. With d_k the newest
// residual and d_{k-m} the m-th older one in its ring:
//   U_mn = <d_k - d_{k-m}, d_k - d_{k-n}>,  V_m = <d_k - d_{k-m}, d_k>,  U c = V,
//   W <- sum_j a_j (W_j + lambda d_j),  a_k = 1 - sum c_m,  a_{k-m} = c_m,
// falling back to the newest pair alone when U is singular.
scft_status anderson_iterate(scft_ctx* c, double* alf, scft_iter_report* rep)
{
    const int R = c->R, L = c->p.anderson + 1;
    const long long M = c->M, M2 = 2 * M;
    const double lam = c->p.anderson_lambda > 0 ? c->p.anderson_lambda : 0.2;
    // Started too far from a solution, Anderson can jump to another saddle point (a lamella
    // at errs 0.1 went to the disordered state), so simple mixing runs until errs < 1e-2.
    const double start = c->p.anderson_start > 0 ? c->p.anderson_start : 1e-2;
    const std::vector<int> doiis = fixed_doiis(c);
    std::vector<int> active;
    for (int ii = 0; ii < R; ++ii)
        if (doiis[ii]) active.push_back(ii);
    const int nb = (int)active.size();
    std::vector<double> errs(R, 0.0);
    std::vector<int> steps(R, 0);
    if (nb > 0) {
        SCFTCHK(run_props(c, c->dW, 2 * M, c->dphi, 4 * M, active, nullptr, nullptr));
        scft::launch_dev_full(c->dslot, nb, R, M, c->p.chi, 0, 0.5, c->dphi, c->dW, c->dDEV, c->dpart, c->nblk,
                              c->st);
        HIPCHK(c, hipGetLastError());
        std::vector<double> devs;  // [slot][2]: sum DEV^2 (2M), sum DEV-^2 (M)
        SCFTCHK(finalize_to_host(c, nb, 2, devs));
        for (int j = 0; j < nb; ++j) {
            errs[active[j]] = std::sqrt(devs[2 * j] / (double)M2);
            steps[active[j]] = 1;
        }
        SCFTCHK(check_errs_finite(c, errs.data(), R));
    }
    for (const int ii : active) {
        double* W = c->dW + (size_t)ii * M2;
        const double* D = c->dDEV + (size_t)ii * M2;
        auto& A = c->and_[ii];
        if (!A.started && errs[ii] >= start) {
            scft::launch_axpy(W, D, 0.05, M2, c->nblk, c->st);
            HIPCHK(c, hipGetLastError());
            continue;
        }
        A.started = true;
        const int s = A.head;
        double* Wr = c->dAW + (size_t)ii * L * M2;
        double* Dr = c->dAD + (size_t)ii * L * M2;
        HIPCHK(c, hipMemcpyAsync(Wr + (size_t)s * M2, W, sizeof(double) * M2, hipMemcpyDeviceToDevice, c->st));
        HIPCHK(c, hipMemcpyAsync(Dr + (size_t)s * M2, D, sizeof(double) * M2, hipMemcpyDeviceToDevice, c->st));
        if (A.count < L) ++A.count;
        // Gram sums of the newest residual with every stored one, newest first
        for (int m = 0; m < A.count; ++m) c->hArows[m] = (s - m + L) % L;
        HIPCHK(c, hipMemcpyAsync(c->dArows, c->hArows, sizeof(int) * A.count, hipMemcpyHostToDevice, c->st));
        scft::launch_and_dots(Dr + (size_t)s * M2, Dr, c->dArows, A.count, M2, c->dApart, c->nblk, c->st);
        HIPCHK(c, hipGetLastError());
        HIPCHK(c, hipMemcpyAsync(c->hAbuf, c->dApart, sizeof(double) * A.count * c->nblk, hipMemcpyDeviceToHost,
                                 c->st));
        HIPCHK(c, hipStreamSynchronize(c->st));
        for (int m = 0; m < A.count; ++m) {
            double g = 0;
            for (int b = 0; b < c->nblk; ++b) g += c->hAbuf[(size_t)b * A.count + m];
            const int j = c->hArows[m];
            A.G[(size_t)s * L + j] = A.G[(size_t)j * L + s] = g;
        }
        int n = A.count - 1;
        std::vector<double> cm;
        if (n > 0) {
            std::vector<double> U((size_t)n * n), V(n);
            const double gss = A.G[(size_t)s * L + s];
            for (int m = 1; m <= n; ++m) {
                const int jm = c->hArows[m];
                V[m - 1] = gss - A.G[(size_t)jm * L + s];
                for (int q = 1; q <= n; ++q) {
                    const int jq = c->hArows[q];
                    U[(size_t)(m - 1) * n + (q - 1)] =
                        gss - A.G[(size_t)s * L + jq] - A.G[(size_t)jm * L + s] + A.G[(size_t)jm * L + jq];
                }
            }
            if (!small_solve(n, U, V, cm)) n = 0;
        }
        double a0 = 1.0;
        for (int m = 0; m < n; ++m) a0 -= cm[m];
        c->hAbuf[0] = a0;
        for (int m = 0; m < n; ++m) c->hAbuf[1 + m] = cm[m];
        HIPCHK(c, hipMemcpyAsync(c->dAcoef, c->hAbuf, sizeof(double) * (n + 1), hipMemcpyHostToDevice, c->st));
        scft::launch_and_combine(W, Wr, Dr, c->dArows, c->dAcoef, n + 1, lam, M2, c->nblk, c->st);
        HIPCHK(c, hipGetLastError());
        A.head = (s + 1) % L;
    }
    std::vector<double> del2, x;
    SCFTCHK(edist2(c, del2));
    arc_positions(R, del2, x);
    HIPCHK(c, hipStreamSynchronize(c->st));
    if (alf)
        for (int ii = 0; ii < R; ++ii) alf[ii] = x[ii];
    if (rep) {
        double stepav = 0;
        for (int ii = 0; ii < R; ++ii) stepav += steps[ii];
        auto put = [R](double* dst, const std::vector<double>& v) {
            if (dst) std::memcpy(dst, v.data(), sizeof(double) * R);
        };
        const std::vector<double> zero(R, 0.0);
        put(rep->errs, errs);
        put(rep->errs3, errs);
        put(rep->x, x);
        put(rep->xref, zero);
        put(rep->dotWs, zero);
        put(rep->dotWs2, zero);
        put(rep->normDER, zero);
        put(rep->normDEV, zero);
        if (rep->steps) std::memcpy(rep->steps, steps.data(), sizeof(int) * R);
        rep->fitup = 0;
        rep->stepav = stepav / R;
    }
    return SCFT_OK;
}

// "two-stage c64 nt512 s2", "two-pass c64 s2", "stockham s1", "pencil s3"
std::string sched_name(const scft_schedule& s)
{
    char b[96];
    if (s.family == SCFT_SCHED_TWO_STAGE)
        snprintf(b, sizeof b, "two-stage c%d nt%d s%d", s.coltile, s.plane_nt, s.streams);
    else if (s.family == SCFT_SCHED_TWO_PASS)
        snprintf(b, sizeof b, "two-pass c%d s%d", s.coltile, s.streams);
    else
        snprintf(b, sizeof b, "%s s%d",
                 s.family == SCFT_SCHED_STOCKHAM       ? "stockham"
                 : s.family == SCFT_SCHED_PENCIL_SPLIT ? "pencil-split"
                                                       : "pencil",
                 s.streams);
    return b;
}

}  // namespace

extern "C" {

const char* scft_last_error(const scft_ctx* ctx) { return ctx ? ctx->err.c_str() : scft::noctx_error(); }

const char* scft_fft_path(const scft_ctx* ctx)
{
    if (!ctx) return "";
    if (ctx->pe32) return ctx->pe32->fft_path();
    return ctx->pe64 ? ctx->pe64->fft_path() : "";
}

const char* scft_engine_info(scft_ctx* ctx)
{
    if (!ctx) return "";
    ctx->info = ctx->pe32 ? ctx->pe32->config() : ctx->pe64 ? ctx->pe64->config() : std::string();
    if (!ctx->tuned.empty()) ctx->info += " (autotuned); candidates: " + ctx->tuned;
    return ctx->info.c_str();
}

scft_status scft_ctx_create(const scft_params* p, scft_ctx** out)
{
    if (!p || !out) return SCFT_ERR_INVALID;
    *out = nullptr;
    auto c = std::make_unique<scft_ctx>();
    c->p = *p;
    c->R = p->nrep;
    c->M = (long long)p->m[0] * p->m[1] * p->m[2];
    if (p->m[0] <= 0 || p->m[1] <= 0 || p->m[2] <= 0 || p->N <= 1 || p->nrep <= 0)
        return SCFT_ERR_INVALID;
    if (p->ens != 1 && p->ens != 2) return SCFT_ERR_INVALID;
    if (p->precision != SCFT_FP64 && p->precision != SCFT_FP32) return SCFT_ERR_INVALID;
    if (p->nrep > scft::SPLINE_RMAX) return SCFT_ERR_UNSUPPORTED;
    // f <- round(N f)/N (scft.cu:1008); props: n1 = round(Ns f), n2 = Ns - n1, Nh = (int)(alpha Ns)
    c->f_used = (double)std::round(p->N * p->f) / p->N;
    c->ds = 1.0 / p->N;
    c->cc.N = p->N;
    c->cc.n1 = (int)std::round(p->N * c->f_used);
    c->cc.n2 = p->N - c->cc.n1;
    c->cc.Nh = (int)(p->alpha * p->N);
    c->cc.ens = p->ens;
    c->cc.alpha = p->alpha;
    c->cc.phic = p->phidb;
    c->cc.ds = c->ds;
    c->cc.phihA = p->phihA;
    if (c->cc.n1 <= 0 || c->cc.n1 >= p->N || c->cc.Nh < 0 || c->cc.Nh > c->cc.n2) return SCFT_ERR_INVALID;
    // the A homopolymer reuses q1 up to Nh, which is all-A only for Nh <= n1
    if (p->phihA < 0 || (p->phihA > 0 && c->cc.Nh > c->cc.n1)) return SCFT_ERR_INVALID;
    if (p->ens == 1 && p->phidb + p->phihA > 1.0) return SCFT_ERR_INVALID;
    if (p->anderson < 0 || p->anderson > 40 || (p->anderson > 0 && p->dostring != 0)) return SCFT_ERR_INVALID;
    if (p->ckpt < 0) return SCFT_ERR_INVALID;
    const scft_schedule& sc = p->sched;
    if (sc.family < SCFT_SCHED_DEFAULT || sc.family > SCFT_SCHED_PENCIL_SPLIT || sc.coltile < 0 || sc.plane_nt < 0 ||
        sc.streams < 0 || sc.streams > 4)
        return SCFT_ERR_INVALID;
    // autotune = 1: the schedule comes from scft_autotune, which times every candidate in a
    // context of its own before this one allocates anything
    if (p->autotune) {
        scft_params q = *p;
        q.autotune = 0;
        char rep[4096];
        const scft_status ts = scft_autotune(&q, nullptr, 2, &c->p.sched, rep, sizeof rep);
        if (ts != SCFT_OK) return ts;
        c->tuned = rep;
    }

    if (hipSetDevice(p->device) != hipSuccess) return SCFT_ERR_HIP;
    if (hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking) != hipSuccess) return SCFT_ERR_HIP;
    std::string msg;
    if (p->precision == SCFT_FP32) {
        c->pe32 = std::make_unique<PropEngine<float>>();
        msg = c->pe32->init(p->m, p->D, p->N, c->R, p->ckpt, &c->p.sched);
    } else {
        c->pe64 = std::make_unique<PropEngine<double>>();
        msg = c->pe64->init(p->m, p->D, p->N, c->R, p->ckpt, &c->p.sched);
    }
    if (!msg.empty()) {
        fprintf(stderr, "scft_ctx_create: %s\n", msg.c_str());
        (void)hipStreamDestroy(c->st);
        c->st = nullptr;
        (void)hipGetLastError();  // a failed allocation must not surface in a later context's launch check
        return msg.find("schedule:") == 0 ? SCFT_ERR_INVALID
               : msg.find("no ") == 0     ? SCFT_ERR_UNSUPPORTED
                                          : SCFT_ERR_NOMEM;
    }
    const long long M = c->M;
    const int R = c->R;
    c->nblk = (int)std::min<long long>(1024, (M + scft::FK_NT - 1) / scft::FK_NT);
    bool ok = hipMalloc((void**)&c->dW, sizeof(double) * 2 * M * R) == hipSuccess &&
              hipMalloc((void**)&c->dphi, sizeof(double) * 4 * M * R) == hipSuccess &&
              hipMalloc((void**)&c->dDEV, sizeof(double) * 2 * M * R) == hipSuccess &&
              hipMalloc((void**)&c->ddWda, sizeof(double) * M * R) == hipSuccess &&
              hipMalloc((void**)&c->dWs, sizeof(double) * 2 * M * R) == hipSuccess &&
              hipMalloc((void**)&c->dphis_rows, sizeof(double) * 4 * M * R) == hipSuccess &&
              hipMalloc((void**)&c->dpart, sizeof(double) * (size_t)c->nblk * 4 * R) == hipSuccess &&
              hipMalloc((void**)&c->dsm, sizeof(double) * SCAL_N * R) == hipSuccess &&
              hipMalloc((void**)&c->dlam, sizeof(double) * R) == hipSuccess &&
              hipMalloc((void**)&c->dcproj, sizeof(double) * R) == hipSuccess &&
              hipMalloc((void**)&c->dphisv, sizeof(double) * 5 * R) == hipSuccess &&
              hipMalloc((void**)&c->dactive, sizeof(int) * R) == hipSuccess &&
              hipMalloc((void**)&c->dslot, sizeof(int) * R) == hipSuccess &&
              hipHostMalloc((void**)&c->hsm, sizeof(double) * SCAL_N * R, hipHostMallocDefault) == hipSuccess &&
              hipHostMalloc((void**)&c->hslot, sizeof(int) * R, hipHostMallocDefault) == hipSuccess &&
              hipHostMalloc((void**)&c->hsmall, sizeof(double) * 4 * R, hipHostMallocDefault) == hipSuccess &&
              hipHostMalloc((void**)&c->hact, sizeof(int) * R, hipHostMallocDefault) == hipSuccess &&
              hipMalloc((void**)&c->dG, sizeof(double) * R * R) == hipSuccess &&
              hipHostMalloc((void**)&c->hG, sizeof(double) * R * R, hipHostMallocDefault) == hipSuccess &&
              scft::row_mix_setup() == hipSuccess;
    if (ok && p->anderson > 0) {
        const size_t L = (size_t)p->anderson + 1;
        ok = hipMalloc((void**)&c->dAW, sizeof(double) * L * 2 * M * R) == hipSuccess &&
             hipMalloc((void**)&c->dAD, sizeof(double) * L * 2 * M * R) == hipSuccess &&
             hipMalloc((void**)&c->dApart, sizeof(double) * L * c->nblk) == hipSuccess &&
             hipMalloc((void**)&c->dAcoef, sizeof(double) * L) == hipSuccess &&
             hipMalloc((void**)&c->dArows, sizeof(int) * L) == hipSuccess &&
             hipHostMalloc((void**)&c->hAbuf, sizeof(double) * L * (c->nblk + 1), hipHostMallocDefault) ==
                 hipSuccess &&
             hipHostMalloc((void**)&c->hArows, sizeof(int) * L, hipHostMallocDefault) == hipSuccess;
        c->and_.assign(R, scft_ctx::Anderson{});
        for (auto& a : c->and_) a.G.assign(L * L, 0.0);
    }
    if (!ok) {
        (void)hipGetLastError();
        scft_ctx_destroy(c.release());
        return SCFT_ERR_NOMEM;
    }
    if (hipMemsetAsync(c->dW, 0, sizeof(double) * 2 * M * R, c->st) != hipSuccess ||
        hipMemsetAsync(c->dphi, 0, sizeof(double) * 4 * M * R, c->st) != hipSuccess ||
        hipMemsetAsync(c->ddWda, 0, sizeof(double) * M * R, c->st) != hipSuccess ||
        hipStreamSynchronize(c->st) != hipSuccess) {
        scft_ctx_destroy(c.release());
        return SCFT_ERR_HIP;
    }
    *out = c.release();
    return SCFT_OK;
}

scft_status scft_ctx_destroy(scft_ctx* c)
{
    if (!c) return SCFT_OK;
    (void)bind_device(c);
    if (c->st) (void)hipStreamSynchronize(c->st);
    c->pe32.reset();
    c->pe64.reset();
    c->pe64_fe.reset();
    void* dptrs[] = {c->dW,    c->dphi,  c->dDEV,   c->ddWda,   c->dWs,    c->dphis_rows, c->dpart,
                     c->dsm,   c->dlam,  c->dcproj, c->dphisv,     c->dactive,
                     c->dslot, c->dAW,   c->dAD,    c->dApart,  c->dAcoef, c->dArows, c->dG};
    for (void* q : dptrs)
        if (q) (void)hipFree(q);
    void* hptrs[] = {c->hsm, c->hslot, c->hsmall, c->hact, c->hAbuf, c->hArows, c->hG, c->hspl};
    for (void* q : {(void*)c->dspl_in, (void*)c->dspl_out})
        if (q) (void)hipFree(q);
    for (void* q : hptrs)
        if (q) (void)hipHostFree(q);
    if (c->st) (void)hipStreamDestroy(c->st);
    delete c;
    return SCFT_OK;
}

scft_status scft_sync(scft_ctx* c)
{
    if (!c) return SCFT_ERR_INVALID;
    SCFTCHK(bind_device(c));
    HIPCHK(c, hipStreamSynchronize(c->st));
    return SCFT_OK;
}

scft_status scft_props_batch(scft_ctx* c, const double* W, double* phi, double* lnQ, double* phis, int nrep)
{
    if (!c || !W || nrep <= 0 || nrep > c->R) return fail(c, SCFT_ERR_INVALID, "scft_props_batch: bad arguments");
    SCFTCHK(bind_device(c));
    const long long M = c->M;
    HIPCHK(c, hipMemcpyAsync(c->dWs, W, sizeof(double) * 2 * M * nrep, hipMemcpyHostToDevice, c->st));
    std::vector<int> slots(nrep);
    for (int i = 0; i < nrep; ++i) slots[i] = i;
    std::vector<double> lq(nrep), ph(5 * (size_t)nrep);
    SCFTCHK(run_props(c, c->dWs, 2 * M, c->dphis_rows, 4 * M, slots, lq.data(), ph.data()));
    if (phi) HIPCHK(c, hipMemcpyAsync(phi, c->dphis_rows, sizeof(double) * 4 * M * nrep, hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    if (lnQ) std::memcpy(lnQ, lq.data(), sizeof(double) * nrep);
    if (phis) std::memcpy(phis, ph.data(), sizeof(double) * 5 * nrep);
    return SCFT_OK;
}

scft_status scft_props_batch_device(scft_ctx* c, const double* dW, double* dphi, double* lnQ, double* phis,
                                    int nrep)
{
    if (!c || !dW || !dphi || nrep <= 0 || nrep > c->R)
        return fail(c, SCFT_ERR_INVALID, "scft_props_batch_device: bad arguments");
    SCFTCHK(bind_device(c));
    std::vector<int> slots(nrep);
    for (int i = 0; i < nrep; ++i) slots[i] = i;
    std::vector<double> lq(nrep), ph(5 * (size_t)nrep);
    SCFTCHK(run_props(c, dW, 2 * c->M, dphi, 4 * c->M, slots, lq.data(), ph.data()));
    if (lnQ) std::memcpy(lnQ, lq.data(), sizeof(double) * nrep);
    if (phis) std::memcpy(phis, ph.data(), sizeof(double) * 5 * nrep);
    return SCFT_OK;
}

scft_status scft_set_fields(scft_ctx* c, const double* W)
{
    if (!c || !W) return fail(c, SCFT_ERR_INVALID, "scft_set_fields: null");
    SCFTCHK(bind_device(c));
    HIPCHK(c, hipMemcpyAsync(c->dW, W, sizeof(double) * 2 * c->M * c->R, hipMemcpyHostToDevice, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    anderson_reset(c);
    return SCFT_OK;
}

scft_status scft_get_fields(scft_ctx* c, double* W)
{
    if (!c || !W) return fail(c, SCFT_ERR_INVALID, "scft_get_fields: null");
    SCFTCHK(bind_device(c));
    HIPCHK(c, hipMemcpyAsync(W, c->dW, sizeof(double) * 2 * c->M * c->R, hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    return SCFT_OK;
}

scft_status scft_get_densities(scft_ctx* c, double* phi)
{
    if (!c || !phi) return fail(c, SCFT_ERR_INVALID, "scft_get_densities: null");
    SCFTCHK(bind_device(c));
    HIPCHK(c, hipMemcpyAsync(phi, c->dphi, sizeof(double) * 4 * c->M * c->R, hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    return SCFT_OK;
}

scft_status scft_get_tangent(scft_ctx* c, double* dWda)
{
    if (!c || !dWda) return fail(c, SCFT_ERR_INVALID, "scft_get_tangent: null");
    SCFTCHK(bind_device(c));
    HIPCHK(c, hipMemcpyAsync(dWda, c->ddWda, sizeof(double) * c->M * c->R, hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    return SCFT_OK;
}

// solve_field0 (scft.cu:516-554), every flagged replica in lockstep with its own lambda and exit
scft_status scft_solve_w_plus(scft_ctx* c, int maxIter, double errTol, const int* which, int* k_out)
{
    if (!c) return SCFT_ERR_INVALID;
    SCFTCHK(bind_device(c));
    anderson_reset(c);  // W+ changes outside the Anderson iteration
    const int R = c->R;
    const long long M = c->M;
    std::vector<int> act(R), k(R, 1);
    std::vector<double> err(R, 1.0), errt(R, 1.0), lam(R, 0.05);
    for (int ii = 0; ii < R; ++ii) act[ii] = which ? (which[ii] != 0) : 1;
    const double dlam1 = 1.09, dlam2 = 1.10;
    std::vector<double> s1;
    for (;;) {
        std::vector<int> slots;
        for (int ii = 0; ii < R; ++ii)
            if (act[ii] && k[ii] < maxIter && err[ii] > errTol) slots.push_back(ii);
        if (slots.empty()) break;
        const int nb = (int)slots.size();
        for (int ii = 0; ii < R; ++ii) c->hsmall[ii] = lam[ii];
        HIPCHK(c, hipMemcpyAsync(c->dlam, c->hsmall, sizeof(double) * R, hipMemcpyHostToDevice, c->st));
        SCFTCHK(run_props(c, c->dW, 2 * M, c->dphi, 4 * M, slots, nullptr, nullptr));
        scft::launch_wplus_mix(c->dslot, nb, M, c->p.chi, c->dlam, c->dphi, c->dW, c->dDEV, c->dpart, c->nblk,
                               c->st);
        HIPCHK(c, hipGetLastError());
        SCFTCHK(finalize_to_host(c, nb, 1, s1));
        for (int j = 0; j < nb; ++j) {
            const int ii = slots[j];
            err[ii] = std::pow(s1[j] / (M), 0.5);
            if (!(err[ii] == err[ii])) {
                char b[160];
                snprintf(b, sizeof b, "solve_field0: W+ error is NaN for replica %d at k=%d", ii, k[ii]);
                return fail(c, SCFT_ERR_NAN, b);
            }
            if (err[ii] < errt[ii]) lam[ii] = std::min(0.1, lam[ii] * dlam1);
            else lam[ii] = std::max(0.005, lam[ii] / dlam2);
            errt[ii] = err[ii];
            k[ii]++;
        }
    }
    if (k_out)
        for (int ii = 0; ii < R; ++ii) k_out[ii] = act[ii] ? k[ii] : 0;
    return SCFT_OK;
}

// Weights of the natural cubic spline at the points xe[erows[j]], j < ne (spline_weight_row):
// G [ne][R] row-major. R >= 2.
static void spline_weights(int R, const double* x, const double* xe, const int* erows, int ne, bool deriv, double* G)
{
    for (int j = 0; j < ne; ++j) scft::spline_weight_row(R, x, xe[erows[j]], deriv, G + (size_t)j * R);
}

scft_status scft_spline_weights(int nrep, const double* x, const double* xe, const int* erows, int ne, int deriv,
                                double* G)
{
    if (nrep < 2 || nrep > scft::SPLINE_RMAX || !x || !xe || !erows || ne < 0 || ne > nrep || !G)
        return SCFT_ERR_INVALID;
    for (int i = 0; i + 1 < nrep; ++i)
        if (!(x[i + 1] > x[i])) return SCFT_ERR_INVALID;
    for (int j = 0; j < ne; ++j)
        if (erows[j] < 0 || erows[j] >= nrep) return SCFT_ERR_INVALID;
    spline_weights(nrep, x, xe, erows, ne, deriv != 0, G);
    return SCFT_OK;
}

// The tangent dW-/dalpha at the nodes (doperp, scft.cu:643-650) of the replicas in rows, or the
// redistribution W- += lambda (spline(xref) - W-) of them (doredist, scft.cu:745-757,
// density.h:511-528): one weight matrix, one launch over the rows W [R][stride] (W- part,
// first M of each row). tangent: out rows of stride ostride; redistribution: in place.
static scft_status spline_launch(scft_ctx* c, const double* x, const double* xref, const std::vector<int>& rows,
                          bool tangent, double lambda, const double* in, long long istride, double* out,
                          long long ostride, long long M)
{
    const int R = c->R, ne = (int)rows.size();
    if (ne == 0) return SCFT_OK;
    // the staging buffers of an earlier launch may still be in flight
    HIPCHK(c, hipStreamSynchronize(c->st));
    for (int j = 0; j < ne; ++j) c->hact[j] = rows[j];
    spline_weights(R, x, tangent ? x : xref, rows.data(), ne, tangent, c->hG);
    if (!tangent)
        for (int j = 0; j < ne; ++j)
            for (int i = 0; i < R; ++i)
                c->hG[(size_t)j * R + i] = lambda * c->hG[(size_t)j * R + i] + (i == rows[j] ? 1.0 - lambda : 0.0);
    HIPCHK(c, hipMemcpyAsync(c->dG, c->hG, sizeof(double) * ne * R, hipMemcpyHostToDevice, c->st));
    HIPCHK(c, hipMemcpyAsync(c->dactive, c->hact, sizeof(int) * ne, hipMemcpyHostToDevice, c->st));
    scft::launch_row_mix(R, M, c->dG, c->dactive, ne, in, istride, out, ostride, c->st);
    HIPCHK(c, hipGetLastError());
    return SCFT_OK;
}

// ---------------------------------------------------------------- chunked host-row spline
scft_status scft_spline_setup(scft_ctx* c, int chunkM)
{
    if (!c || chunkM <= 0) return fail(c, SCFT_ERR_INVALID, "scft_spline_setup: chunkM must be positive");
    SCFTCHK(bind_device(c));
    const int R = c->R;
    const long long n = std::min<long long>(chunkM, c->M);
    HIPCHK(c, hipStreamSynchronize(c->st));
    for (void* q : {(void*)c->dspl_in, (void*)c->dspl_out})
        if (q) (void)hipFree(q);
    if (c->hspl) (void)hipHostFree(c->hspl);
    c->dspl_in = c->dspl_out = c->hspl = nullptr;
    c->spl_chunk = 0;
    const size_t bytes = sizeof(double) * (size_t)R * n;
    if (hipMalloc((void**)&c->dspl_in, bytes) != hipSuccess || hipMalloc((void**)&c->dspl_out, bytes) != hipSuccess ||
        hipHostMalloc((void**)&c->hspl, bytes, hipHostMallocDefault) != hipSuccess) {
        (void)hipGetLastError();
        return fail(c, SCFT_ERR_NOMEM, "scft_spline_setup: staging allocation failed");
    }
    c->spl_chunk = (int)n;
    return SCFT_OK;
}

scft_status scft_spline_set_x(scft_ctx* c, const double* x)
{
    if (!c || !x) return fail(c, SCFT_ERR_INVALID, "scft_spline_set_x: null");
    for (int i = 0; i + 1 < c->R; ++i)
        if (!(x[i + 1] > x[i])) return fail(c, SCFT_ERR_INVALID, "scft_spline_set_x: x must increase strictly");
    c->spl_x.assign(x, x + c->R);
    return SCFT_OK;
}

scft_status scft_spline_set_xref(scft_ctx* c, const double* xref)
{
    if (!c || !xref) return fail(c, SCFT_ERR_INVALID, "scft_spline_set_xref: null");
    c->spl_xref.assign(xref, xref + c->R);
    return SCFT_OK;
}

// Checks a chunk call and copies rows ww[ii][r0 .. r0+n) into the device staging [R][n].
static scft_status spline_chunk_in(scft_ctx* c, const double* const* ww, int r0, int n, const int* doiis,
                                   bool need_xref, const char* who)
{
    if (!c || !ww || !doiis) return fail(c, SCFT_ERR_INVALID, std::string(who) + ": null");
    if (c->R < 2) return fail(c, SCFT_ERR_INVALID, std::string(who) + ": the spline needs at least two replicas");
    if (c->spl_chunk == 0) return fail(c, SCFT_ERR_INVALID, std::string(who) + ": call scft_spline_setup first");
    if ((int)c->spl_x.size() != c->R || (need_xref && (int)c->spl_xref.size() != c->R))
        return fail(c, SCFT_ERR_INVALID, std::string(who) + ": nodes not set (scft_spline_set_x / _xref)");
    if (r0 < 0 || n <= 0 || n > c->spl_chunk || (long long)r0 + n > c->M)
        return fail(c, SCFT_ERR_INVALID, std::string(who) + ": chunk outside the grid or the staging");
    SCFTCHK(bind_device(c));
    HIPCHK(c, hipStreamSynchronize(c->st));  // the staging of an earlier call is free
    for (int ii = 0; ii < c->R; ++ii) {
        if (!ww[ii]) return fail(c, SCFT_ERR_INVALID, std::string(who) + ": null row");
        std::memcpy(c->hspl + (size_t)ii * n, ww[ii] + r0, sizeof(double) * n);
    }
    HIPCHK(c, hipMemcpyAsync(c->dspl_in, c->hspl, sizeof(double) * (size_t)c->R * n, hipMemcpyHostToDevice, c->st));
    return SCFT_OK;
}

scft_status scft_spline_derivs(scft_ctx* c, const double* const* ww, int r0, int n, const int* doiis,
                               double* const* out)
{
    SCFTCHK(spline_chunk_in(c, ww, r0, n, doiis, false, "scft_spline_derivs"));
    if (!out) return fail(c, SCFT_ERR_INVALID, "scft_spline_derivs: null output");
    std::vector<int> rows;
    for (int ii = 0; ii < c->R; ++ii)
        if (doiis[ii] == 1) {
            if (!out[ii]) return fail(c, SCFT_ERR_INVALID, "scft_spline_derivs: null output row");
            rows.push_back(ii);
        }
    SCFTCHK(spline_launch(c, c->spl_x.data(), nullptr, rows, true, 0.0, c->dspl_in, n, c->dspl_out, n, n));
    HIPCHK(c, hipMemcpyAsync(c->hspl, c->dspl_out, sizeof(double) * (size_t)c->R * n, hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    for (int ii : rows) std::memcpy(out[ii] + r0, c->hspl + (size_t)ii * n, sizeof(double) * n);
    return SCFT_OK;
}

scft_status scft_spline_redist(scft_ctx* c, double* const* ww, int r0, int n, const int* doiis, double lambda)
{
    SCFTCHK(spline_chunk_in(c, ww, r0, n, doiis, true, "scft_spline_redist"));
    std::vector<int> rows;
    for (int ii = 1; ii < c->R - 1; ++ii)
        if (doiis[ii] == 1) rows.push_back(ii);
    SCFTCHK(spline_launch(c, c->spl_x.data(), c->spl_xref.data(), rows, false, lambda, c->dspl_in, n, c->dspl_in, n,
                          n));
    HIPCHK(c, hipMemcpyAsync(c->hspl, c->dspl_in, sizeof(double) * (size_t)c->R * n, hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    for (int ii : rows) std::memcpy(ww[ii] + r0, c->hspl + (size_t)ii * n, sizeof(double) * n);
    return SCFT_OK;
}

// one solve_field iteration (scft.cu:592-757), single process, doii = -1
scft_status scft_string_iterate(scft_ctx* c, int k, double errTol, double* alf, scft_iter_report* rep)
{
    if (!c) return SCFT_ERR_INVALID;
    SCFTCHK(bind_device(c));
    if (c->p.anderson > 0) return anderson_iterate(c, alf, rep);
    const int R = c->R;
    const long long M = c->M;
    const double lambda = 0.05, eps = 0.5;  // scft.cu:562, 570 (lambda stays 0.05: A.1)
    int dospring = 1, doredist = 0, doperp = 0;
    if (c->p.dostring != 1) dospring = 0;
    if (k >= c->p.springsteps && c->p.dostring == 1) {
        dospring = 0;
        doredist = 1;
        doperp = 1;
    }
    const std::vector<int> doiis = fixed_doiis(c);
    std::vector<int> active;
    for (int ii = 0; ii < R; ++ii)
        if (doiis[ii]) active.push_back(ii);

    HIPCHK(c, hipMemsetAsync(c->dDEV, 0, sizeof(double) * 2 * M * R, c->st));
    std::vector<int> kk(R, 0), steps(R, 0);
    SCFTCHK(scft_solve_w_plus(c, 20, errTol, doiis.data(), kk.data()));
    for (int ii = 0; ii < R; ++ii) steps[ii] = doiis[ii] ? 1 + kk[ii] : 0;
    SCFTCHK(run_props(c, c->dW, 2 * M, c->dphi, 4 * M, active, nullptr, nullptr));
    const int nb = (int)active.size();
    scft::launch_dev_full(c->dslot, nb, R, M, c->p.chi, dospring, eps, c->dphi, c->dW, c->dDEV, c->dpart, c->nblk,
                          c->st);
    HIPCHK(c, hipGetLastError());
    std::vector<double> devs;  // [slot][2]: sum DEV^2 (2M), sum DEV-^2 (M)
    SCFTCHK(finalize_to_host(c, nb, 2, devs));
    std::vector<double> del2, x, xref(R);
    SCFTCHK(edist2(c, del2));
    arc_positions(R, del2, x);
    for (int ii = 0; ii < R; ++ii) xref[ii] = (1.0 * ii) / (R - 1.0);

    std::vector<double> errs(R, 0.0), errs3(R, 0.0), normDER(R, 0.0), normDEV(R, 0.0), dotWs(R, 0.0),
        dotWs2(R, 0.0);
    std::vector<double> sdev(R, 0.0), sdevm(R, 0.0);
    for (int j = 0; j < nb; ++j) {
        sdev[active[j]] = devs[2 * j];
        sdevm[active[j]] = devs[2 * j + 1];
    }
    for (int ii = 0; ii < R; ++ii) errs[ii] = std::pow(sdev[ii] / (2 * M), 0.5);
    SCFTCHK(check_errs_finite(c, errs.data(), R));

    std::vector<int> proj(R, 0);  // interior replicas whose DEV- is projected (scft.cu:670-678)
    for (int ii = 0; ii < R; ++ii) c->hsmall[3 * R + ii] = 0.0;  // cproj
    if (doperp == 1) {
        SCFTCHK(spline_launch(c, x.data(), nullptr, active, true, 0.0, c->dW, 2 * M, c->ddWda, M, M));
        scft::launch_tangent_dots(c->dslot, nb, M, c->ddWda, c->dDEV, c->dpart, c->nblk, c->st);
        HIPCHK(c, hipGetLastError());
        std::vector<double> td;
        SCFTCHK(finalize_to_host(c, nb, 2, td));
        for (int j = 0; j < nb; ++j) {
            const int ii = active[j];
            normDER[ii] = std::sqrt(td[2 * j]);
            normDEV[ii] = std::sqrt(sdevm[ii]);
            dotWs[ii] = td[2 * j + 1] / (normDEV[ii] * normDER[ii]);
        }
        for (int ii = 1; ii < R - 1; ++ii) {
            if (doiis[ii] == 1 && normDER[ii] > 0 && normDEV[ii] > 0) {
                proj[ii] = 1;
                c->hsmall[3 * R + ii] = (dotWs[ii] / normDER[ii]) * normDEV[ii];
            } else if (doiis[ii] == 1) {
                fprintf(stderr, "Fatal error occurred (0)... %d %d %lf %lf\n", k, ii, normDER[ii], normDEV[ii]);
            }
        }
    }
    HIPCHK(c, hipMemcpyAsync(c->dcproj, c->hsmall + 3 * R, sizeof(double) * R, hipMemcpyHostToDevice, c->st));
    scft::launch_project_mix(R, M, c->dcproj, lambda, c->ddWda, c->dDEV, c->dW, c->dpart, c->nblk, c->st);
    HIPCHK(c, hipGetLastError());
    std::vector<double> pm;
    SCFTCHK(finalize_to_host(c, R, 2, pm));
    if (doperp == 1) {
        for (int ii = 0; ii < R; ++ii) {
            errs3[ii] = proj[ii] ? std::pow(pm[2 * ii] / (2 * M), 0.5) : 0.0;
            // DEV- of a replica that was not projected (or projected by 0) is unchanged
            if (doiis[ii] == 1)
                dotWs2[ii] = (c->hsmall[3 * R + ii] != 0.0) ? pm[2 * ii + 1] / (normDEV[ii] * normDER[ii])
                                                             : dotWs[ii];
        }
    } else {
        for (int ii = 0; ii < R; ++ii) errs3[ii] = errs[ii];
    }
    errs3[0] = errs[0];
    errs3[R - 1] = errs[R - 1];

    // arc length of the mixed string (scft.cu:725-743)
    SCFTCHK(edist2(c, del2));
    arc_positions(R, del2, x);
    double fitup = 0;
    for (int ii = 1; ii < R - 1; ++ii) fitup += std::pow(x[ii] - xref[ii], 2.0);
    fitup /= R - 2.0;
    fitup = std::sqrt(fitup);
    fitup /= R - 1.0;
    if (alf)
        for (int ii = 0; ii < R; ++ii) alf[ii] = x[ii];

    if (doredist == 1) {
        std::vector<int> inner;  // interior replicas that move (density.h:524-527)
        for (int ii = 1; ii < R - 1; ++ii)
            if (doiis[ii]) inner.push_back(ii);
        SCFTCHK(spline_launch(c, x.data(), xref.data(), inner, false, lambda, c->dW, 2 * M, c->dW, 2 * M, M));
    }
    HIPCHK(c, hipStreamSynchronize(c->st));

    if (rep) {
        double stepav = 0;
        for (int ii = 0; ii < R; ++ii) stepav += steps[ii];
        stepav /= R;
        auto put = [R](double* dst, const std::vector<double>& v) {
            if (dst) std::memcpy(dst, v.data(), sizeof(double) * R);
        };
        put(rep->errs, errs);
        put(rep->errs3, errs3);
        put(rep->x, x);
        put(rep->xref, xref);
        put(rep->dotWs, dotWs);
        put(rep->dotWs2, dotWs2);
        put(rep->normDER, normDER);
        put(rep->normDEV, normDEV);
        if (rep->steps) std::memcpy(rep->steps, steps.data(), sizeof(int) * R);
        rep->fitup = fitup;
        rep->stepav = stepav;
    }
    return SCFT_OK;
}

// FreeE (scft.cu:812-853), single process
scft_status scft_free_energy(scft_ctx* c, double* FE, double* FEs, double* FEs0, double* FEs1, double* phctot,
                             double* alf)
{
    if (!c) return SCFT_ERR_INVALID;
    SCFTCHK(bind_device(c));
    const int R = c->R;
    const long long M = c->M;
    const std::vector<int> doiis = fixed_doiis(c);
    std::vector<int> which(R, 0), set;
    for (int ii = 0; ii < R; ++ii)
        if (doiis[ii] == 1 || ii == 0 || ii == R - 1) {
            which[ii] = 1;
            set.push_back(ii);
        }
    SCFTCHK(scft_solve_w_plus(c, 100, 1e-4, which.data(), nullptr));
    const int nb = (int)set.size();
    std::vector<double> lq(nb);
    // In an FP32 context the props that F is taken from runs in FP64: F is stationary in the
    // fields at a converged string point, so evaluating it in FP64 on the FP32-relaxed fields
    // leaves only a second-order error from the FP32 relaxation.
    const bool fe64 = c->pe32 && c->p.fe_fp64;
    if (fe64 && !c->pe64_fe) {
        c->pe64_fe = std::make_unique<PropEngine<double>>();
        const std::string msg = c->pe64_fe->init(c->p.m, c->p.D, c->p.N, c->R, c->p.ckpt);
        if (!msg.empty()) {
            c->pe64_fe.reset();
            (void)hipGetLastError();
            return fail(c, SCFT_ERR_NOMEM, "FP64 free-energy engine: " + msg);
        }
    }
    SCFTCHK(run_props(c, c->dW, 2 * M, c->dphi, 4 * M, set, lq.data(), nullptr, fe64));
    scft::launch_fe_sums(c->dslot, nb, M, c->p.chi, c->dW, c->dphi, c->dpart, c->nblk, c->st);
    HIPCHK(c, hipGetLastError());
    std::vector<double> fs;
    SCFTCHK(finalize_to_host(c, nb, 2, fs));
    std::vector<double> vFEs(R, 0.0), vFEs0(R, 0.0), vFEs1(R, 0.0), vph(R, 0.0), valf(R, 0.0);
    for (int j = 0; j < nb; ++j) {
        const int ii = set[j];
        vFEs0[ii] = -lq[j];
        vFEs1[ii] = fs[2 * j] / M;
        vph[ii] = fs[2 * j + 1] / (M * c->f_used);
        vFEs[ii] = vFEs0[ii] + vFEs1[ii];
    }
    std::vector<double> del2;
    SCFTCHK(edist2(c, del2));
    valf[0] = 0;
    for (int ii = 1; ii < R; ++ii) valf[ii] = valf[ii - 1] + std::sqrt(del2[ii - 1]);
    double tot = 0;
    for (int ii = 0; ii < R; ++ii) tot += vFEs[ii];
    auto put = [R](double* dst, const std::vector<double>& v) {
        if (dst) std::memcpy(dst, v.data(), sizeof(double) * R);
    };
    put(FEs, vFEs);
    put(FEs0, vFEs0);
    put(FEs1, vFEs1);
    put(phctot, vph);
    put(alf, valf);
    if (FE) *FE = tot;
    return SCFT_OK;
}

}  // extern "C"

namespace {

// d lnQ / dD_i of the nb slots of the last run_props on engine pe: from the box
// derivatives of the factored QC', Q2' and the ensemble's lnQ (density.h:388-399).
template <typename T>
scft_status dlnq_from_engine(scft_ctx* c, PropEngine<T>* pe, int nb, double* dl)
{
    std::vector<double> d(9 * (size_t)nb), sc(SCAL_N * (size_t)nb);
    hipError_t e = pe->box_gradient(nb, c->cc, d.data(), c->st);
    if (e == hipErrorNotSupported)
        return fail(c, SCFT_ERR_UNSUPPORTED, "box gradient: every axis must factor into 2, 3, 5 and be at most 1024");
    if (e != hipSuccess) return fail(c, SCFT_ERR_HIP, std::string("box gradient: ") + hipGetErrorString(e));
    HIPCHK(c, hipMemcpyAsync(sc.data(), pe->scal(), sizeof(double) * SCAL_N * nb, hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    const double phic = c->cc.phic, alpha = c->cc.alpha, phihA = c->cc.phihA, phihB = 1.0 - phic - phihA;
    for (int j = 0; j < nb; ++j) {
        const double* s = &sc[(size_t)SCAL_N * j];
        const double QC = s[3], Q2 = s[4], Q1 = s[5], QCp = s[6], Q2p = s[7], Q1p = s[9];
        for (int i = 0; i < 3; ++i) {
            // d ln QC/dD_i, d ln Q2/dD_i, d ln Q1/dD_i
            const double rc = d[9 * j + i] / QCp, r2 = d[9 * j + 3 + i] / Q2p, r1 = d[9 * j + 6 + i] / Q1p;
            double v = c->cc.ens == 1 ? phic * rc + (phihB / alpha) * r2 : Q2 * r2 + phic * QC * rc;
            if (phihA != 0.0) v += c->cc.ens == 1 ? (phihA / alpha) * r1 : phihA * Q1 * r1;
            dl[3 * j + i] = v;
        }
    }
    return SCFT_OK;
}

scft_status dlnq_slots(scft_ctx* c, int nb, double* dl)
{
    return c->pe32 ? dlnq_from_engine(c, c->pe32.get(), nb, dl) : dlnq_from_engine(c, c->pe64.get(), nb, dl);
}

}  // namespace

extern "C" {

scft_status scft_dlnq_dbox(scft_ctx* c, const double* W, int nrow, double* lnQ, double* dlnQ_dD)
{
    if (!c || !W || !dlnQ_dD || nrow <= 0 || nrow > c->R)
        return fail(c, SCFT_ERR_INVALID, "scft_dlnq_dbox: bad arguments");
    SCFTCHK(bind_device(c));
    const long long M = c->M;
    HIPCHK(c, hipMemcpyAsync(c->dWs, W, sizeof(double) * 2 * M * nrow, hipMemcpyHostToDevice, c->st));
    std::vector<int> slots(nrow);
    for (int i = 0; i < nrow; ++i) slots[i] = i;
    std::vector<double> lq(nrow);
    SCFTCHK(run_props(c, c->dWs, 2 * M, c->dphis_rows, 4 * M, slots, lq.data(), nullptr));
    SCFTCHK(dlnq_slots(c, nrow, dlnQ_dD));
    if (lnQ) std::memcpy(lnQ, lq.data(), sizeof(double) * nrow);
    return SCFT_OK;
}

scft_status scft_box_stress(scft_ctx* c, double* dFdD, double* dFdchi)
{
    if (!c) return SCFT_ERR_INVALID;
    SCFTCHK(bind_device(c));
    const int R = c->R;
    const long long M = c->M;
    std::vector<int> slots(R);
    for (int i = 0; i < R; ++i) slots[i] = i;
    SCFTCHK(run_props(c, c->dW, 2 * M, c->dphi, 4 * M, slots, nullptr, nullptr));
    if (dFdchi) {
        scft::launch_phiab_sums(c->dslot, R, M, c->dphi, c->dpart, c->nblk, c->st);
        HIPCHK(c, hipGetLastError());
        std::vector<double> s;
        SCFTCHK(finalize_to_host(c, R, 1, s));
        for (int ii = 0; ii < R; ++ii) dFdchi[ii] = s[ii] / (double)M;
    }
    if (dFdD) {
        std::vector<double> dl(3 * (size_t)R);
        SCFTCHK(dlnq_slots(c, R, dl.data()));
        for (int i = 0; i < 3 * R; ++i) dFdD[i] = -dl[i];
    }
    return SCFT_OK;
}

scft_status scft_backward(scft_ctx* c, const double* dL_dF, double* dL_dD, double* dL_dchi)
{
    if (!c) return SCFT_ERR_INVALID;
    if (!dL_dF) return fail(c, SCFT_ERR_INVALID, "scft_backward: dL_dF is NULL");
    const int R = c->R;
    for (int r = 0; r < R; ++r)
        if (!std::isfinite(dL_dF[r])) return fail(c, SCFT_ERR_INVALID, "scft_backward: dL_dF is not finite");
    std::vector<double> dD(dL_dD ? 3 * (size_t)R : 0), dchi(dL_dchi ? (size_t)R : 0);
    SCFTCHK(scft_box_stress(c, dL_dD ? dD.data() : nullptr, dL_dchi ? dchi.data() : nullptr));
    for (int r = 0; r < R; ++r) {
        if (dL_dD)
            for (int i = 0; i < 3; ++i) dL_dD[3 * r + i] = dL_dF[r] * dD[3 * r + i];
        if (dL_dchi) dL_dchi[r] = dL_dF[r] * dchi[r];
    }
    return SCFT_OK;
}

scft_status scft_set_box(scft_ctx* c, const double D[3])
{
    if (!c || !D || !(D[0] > 0 && D[1] > 0 && D[2] > 0))
        return fail(c, SCFT_ERR_INVALID, "scft_set_box: box lengths must be positive");
    SCFTCHK(bind_device(c));
    HIPCHK(c, hipStreamSynchronize(c->st));
    hipError_t e = hipSuccess;
    if (c->pe32) e = c->pe32->set_box(D);
    if (e == hipSuccess && c->pe64) e = c->pe64->set_box(D);
    if (e == hipSuccess && c->pe64_fe) e = c->pe64_fe->set_box(D);
    if (e != hipSuccess) return fail(c, SCFT_ERR_HIP, std::string("scft_set_box: ") + hipGetErrorString(e));
    for (int i = 0; i < 3; ++i) c->p.D[i] = D[i];
    anderson_reset(c);
    return SCFT_OK;
}

scft_status scft_iterate_concurrent(scft_ctx* const* ctxs, int n, const int* k, double errTol, double* const* alf,
                                    scft_iter_report* reps)
{
    if (!ctxs || !k || n <= 0) return SCFT_ERR_INVALID;
    for (int i = 0; i < n; ++i) {
        if (!ctxs[i]) return SCFT_ERR_INVALID;
        for (int j = 0; j < i; ++j)
            if (ctxs[j] == ctxs[i]) return fail(ctxs[i], SCFT_ERR_INVALID, "scft_iterate_concurrent: repeated context");
    }
    std::vector<scft_status> st(n, SCFT_OK);
    std::vector<std::thread> th;
    th.reserve(n);
    for (int i = 0; i < n; ++i)
        th.emplace_back([&, i]() {
            st[i] = scft_string_iterate(ctxs[i], k[i], errTol, alf ? alf[i] : nullptr, reps ? &reps[i] : nullptr);
        });
    for (auto& t : th) t.join();
    for (int i = 0; i < n; ++i)
        if (st[i] != SCFT_OK) return st[i];
    return SCFT_OK;
}

scft_status scft_relax_box(scft_ctx* c, int axes, double stressTol, double errTol, int maxFieldIter, int maxBoxes,
                           double* D, double* dFdD, int* boxes)
{
    if (!c || !D || c->R != 1 || c->p.dostring != 0 || (axes & 7) == 0 || maxFieldIter <= 0 || maxBoxes <= 0)
        return fail(c, SCFT_ERR_INVALID,
                    "scft_relax_box: needs one replica, dostring = 0, an axis in `axes` and positive limits");
    SCFTCHK(scft_set_box(c, D));
    double Dprev[3] = {0, 0, 0}, gprev[3] = {0, 0, 0}, g[3] = {0, 0, 0};
    bool have_prev = false;
    double errs = 0, alf = 0;
    int steps = 0, k = 0;
    scft_iter_report rep{};
    rep.errs = &errs;
    rep.steps = &steps;
    for (int b = 0; b < maxBoxes; ++b) {
        for (int it = 0; it < maxFieldIter; ++it) {
            SCFTCHK(scft_string_iterate(c, k++, 1e-3, &alf, &rep));
            if (errs <= errTol) break;
        }
        SCFTCHK(scft_box_stress(c, g, nullptr));
        if (dFdD) std::memcpy(dFdD, g, sizeof g);
        if (boxes) *boxes = b + 1;
        bool done = true;
        for (int i = 0; i < 3; ++i)
            if ((axes >> i & 1) && std::fabs(g[i]) > stressTol) done = false;
        if (done) return SCFT_OK;
        double Dn[3] = {D[0], D[1], D[2]};
        for (int i = 0; i < 3; ++i) {
            if (!(axes >> i & 1)) continue;
            double step = (g[i] > 0 ? -0.01 : 0.01) * D[i];
            if (have_prev && g[i] != gprev[i]) step = -g[i] * (D[i] - Dprev[i]) / (g[i] - gprev[i]);
            const double cap = 0.1 * D[i];
            step = std::max(-cap, std::min(cap, step));
            Dn[i] = D[i] + step;
        }
        for (int i = 0; i < 3; ++i) {
            Dprev[i] = D[i];
            gprev[i] = g[i];
            D[i] = Dn[i];
        }
        have_prev = true;
        SCFTCHK(scft_set_box(c, D));
    }
    return fail(c, SCFT_ERR_NOCONV, "scft_relax_box: stress tolerance not reached within maxBoxes boxes");
}

scft_status scft_debug_qq_invariant(scft_ctx* c, int slot, double* qq)
{
    if (!c || !qq || slot < 0 || slot >= c->R) return SCFT_ERR_INVALID;
    SCFTCHK(bind_device(c));
    hipError_t e = c->pe32 ? c->pe32->qq_invariant(slot, qq, c->st) : c->pe64->qq_invariant(slot, qq, c->st);
    if (e != hipSuccess) return fail(c, SCFT_ERR_HIP, hipGetErrorString(e));
    return SCFT_OK;
}

scft_status scft_autotune(const scft_params* p, const double* W, int reps, scft_schedule* best, char* report,
                          size_t report_len)
{
    if (!p || !best || reps < 1) return SCFT_ERR_INVALID;
    if (report && report_len) report[0] = '\0';
    if (p->m[0] <= 0 || p->m[1] <= 0 || p->m[2] <= 0 || p->nrep <= 0) return SCFT_ERR_INVALID;
    if (p->precision != SCFT_FP64 && p->precision != SCFT_FP32) return SCFT_ERR_INVALID;
    // the rows a string iteration propagates: every replica but the fixed ends
    int nact = p->nrep;
    if (p->dostring && !p->justFE) nact -= (p->fix[0] ? 1 : 0) + (p->fix[1] ? 1 : 0);
    nact = std::max(1, nact);
    const std::vector<scft_schedule> cand = p->precision == SCFT_FP32 ? PropEngine<float>::candidates(p->m, nact)
                                                                       : PropEngine<double>::candidates(p->m, nact);
    if (cand.empty()) return SCFT_ERR_UNSUPPORTED;
    const long long M = (long long)p->m[0] * p->m[1] * p->m[2];
    std::vector<double> trial;
    const double* Wt = W;
    if (!Wt) {
        // lamellar W- along x with a small per-replica modulation along y, W+ = 0
        trial.assign((size_t)p->nrep * 2 * (size_t)M, 0.0);
        const double tpi = 8 * atan(1.0);
        for (int i = 0; i < p->nrep; ++i)
            for (long long r = 0; r < M; ++r) {
                const long long x = r / ((long long)p->m[1] * p->m[2]), y = (r / p->m[2]) % p->m[1];
                trial[(size_t)i * 2 * M + r] = 1.5 * cos(tpi * x / p->m[0]) +
                                               0.1 * sin(tpi * (double)y / p->m[1] + tpi * i / p->nrep);
            }
        Wt = trial.data();
    }
    // one candidate in a context of its own: an untimed props, then the best of n timed ones
    auto time_sched = [&](const scft_schedule& sc, int n, double* ms) -> scft_status {
        scft_params q = *p;
        q.autotune = 0;
        q.sched = sc;
        scft_ctx* c = nullptr;
        scft_status st = scft_ctx_create(&q, &c);
        if (st == SCFT_OK) st = scft_set_fields(c, Wt);
        double t = 0;
        if (st == SCFT_OK) st = scft_debug_time_props(c, nact, 1, &t);
        for (int r = 0; r < n && st == SCFT_OK; ++r) {
            st = scft_debug_time_props(c, nact, 1, &t);
            if (st == SCFT_OK && (*ms < 0 || t < *ms)) *ms = t;
        }
        if (c) (void)scft_ctx_destroy(c);
        return st;
    };
    std::vector<double> ms(cand.size(), -1.0);
    std::vector<scft_status> sts(cand.size(), SCFT_OK);
    for (size_t i = 0; i < cand.size(); ++i) sts[i] = time_sched(cand[i], reps, &ms[i]);
    // Candidates within 3% of the fastest are timed again, each in a fresh context: the first
    // pass separates schedules that differ by more than run-to-run noise, the second picks
    // among the near ties.
    double best_ms = -1;
    for (size_t i = 0; i < cand.size(); ++i)
        if (sts[i] == SCFT_OK && (best_ms < 0 || ms[i] < best_ms)) best_ms = ms[i];
    if (best_ms > 0)
        for (size_t i = 0; i < cand.size(); ++i)
            if (sts[i] == SCFT_OK && ms[i] <= 1.03 * best_ms) sts[i] = time_sched(cand[i], reps + 1, &ms[i]);
    std::string rep;
    scft_status last = SCFT_OK;
    best_ms = -1;
    for (size_t i = 0; i < cand.size(); ++i) {
        char b[160];
        if (sts[i] == SCFT_OK) {
            snprintf(b, sizeof b, "%s %.4f; ", sched_name(cand[i]).c_str(), ms[i]);
            if (best_ms < 0 || ms[i] < best_ms) {
                best_ms = ms[i];
                *best = cand[i];
            }
        } else {
            snprintf(b, sizeof b, "%s failed (status %d); ", sched_name(cand[i]).c_str(), (int)sts[i]);
            last = sts[i];
        }
        rep += b;
    }
    if (best_ms < 0) return last != SCFT_OK ? last : SCFT_ERR_UNSUPPORTED;
    rep += "ms per contour step over " + std::to_string(nact) + " rows; chosen: " + sched_name(*best);
    if (report && report_len) {
        const size_t n = std::min(report_len - 1, rep.size());
        std::memcpy(report, rep.data(), n);
        report[n] = '\0';
    }
    return SCFT_OK;
}

scft_status scft_debug_time_props(scft_ctx* c, int nrep, int reps, double* ms_per_step)
{
    if (!c || !ms_per_step || nrep <= 0 || nrep > c->R || reps <= 0) return SCFT_ERR_INVALID;
    SCFTCHK(bind_device(c));
    std::vector<int> slots(nrep);
    for (int i = 0; i < nrep; ++i) slots[i] = i;
    hipError_t e = c->pe32 ? c->pe32->enable_timing(true) : c->pe64->enable_timing(true);
    if (e != hipSuccess) return fail(c, SCFT_ERR_HIP, hipGetErrorString(e));
    double tot = 0;
    for (int r = 0; r < reps; ++r) {
        SCFTCHK(run_props(c, c->dW, 2 * c->M, c->dphi, 4 * c->M, slots, nullptr, nullptr));
        HIPCHK(c, hipStreamSynchronize(c->st));
        tot += c->pe32 ? c->pe32->loop_ms() : c->pe64->loop_ms();
    }
    if (c->pe32) (void)c->pe32->enable_timing(false);
    else (void)c->pe64->enable_timing(false);
    *ms_per_step = tot / reps / c->p.N;
    return SCFT_OK;
}

}  // extern "C"
