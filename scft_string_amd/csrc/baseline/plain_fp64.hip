// plain_fp64.hip -- comparator for the benchmark (not the product): the "plain per-replica
// FP64 GPU path". It runs density::props (density.h:344-459) with the reference's
// execution structure on MI355X:
//   - one replica at a time;
//   - hipFFT D2Z/Z2D over the stacked {q, q-dagger} pair (batch 2, density.h:67-74);
//   - one launch per elementwise operation (14 per contour step, density.h:129-220);
//   - Simpson sums as one launch per contour index (density.h:401-422);
//   - single-block reductions (GPU_kernels.h:58-90), and blocking copies of W in and
//     phi out per call (density.h:362, 457).
// Written for this project (the kernels are plain one-thread-per-point FP64 loops), to
// measure what the batched, fused FP32 path gains over that structure.
#include <hip/hip_runtime.h>
#include <hipfft/hipfft.h>

#include <cmath>
#include <cstdlib>
#include <vector>

namespace {

constexpr int TPB = 256;

__global__ void pk_mul(double* a, const double* b, const double* c, long long n)
{
    long long i = (long long)blockIdx.x * TPB + threadIdx.x;
    if (i < n) a[i] = b[i] * c[i];
}
__global__ void pk_mul_self(double* a, const double* b, long long n)
{
    long long i = (long long)blockIdx.x * TPB + threadIdx.x;
    if (i < n) a[i] *= b[i];
}
__global__ void pk_mul_spec(hipfftDoubleComplex* h, const double* k, long long n)
{
    long long i = (long long)blockIdx.x * TPB + threadIdx.x;
    if (i < n) {
        h[i].x *= k[i];
        h[i].y *= k[i];
    }
}
__global__ void pk_rich(double* out, const double* h, const double* f, const double* e4, const double* e2, long long n)
{
    long long i = (long long)blockIdx.x * TPB + threadIdx.x;
    if (i < n) out[i] = (h[i] * e4[i] * 4 - f[i] * e2[i]) / 3;
}
__global__ void pk_kfac(double* k12, double* k6, const double* ksq, double ds, double Md, long long nk)
{
    long long i = (long long)blockIdx.x * TPB + threadIdx.x;
    if (i < nk) {
        k12[i] = exp(-ksq[i] * ds / 12) / Md;
        k6[i] = k12[i] * k12[i] * Md;
    }
}
__global__ void pk_wfac(double* e4A, double* e2A, double* e4B, double* e2B, const double* W, double ds, long long M)
{
    long long i = (long long)blockIdx.x * TPB + threadIdx.x;
    if (i < M) {
        e4A[i] = exp(-(W[i] + W[i + M]) * ds / 4);
        e2A[i] = e4A[i] * e4A[i];
        e4B[i] = exp(-(-W[i] + W[i + M]) * ds / 4);
        e2B[i] = e4B[i] * e4B[i];
    }
}
__global__ void pk_set(double* a, double v, long long n)
{
    long long i = (long long)blockIdx.x * TPB + threadIdx.x;
    if (i < n) a[i] = v;
}
__global__ void pk_sum1(const double* a, double* out, long long n)  // one block
{
    __shared__ double sh[1024];
    double v = 0;
    for (long long i = threadIdx.x; i < n; i += 1024) v += a[i];
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int o = 512; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = sh[0];
}
__global__ void pk_ends(double* rhB, double* rCA, double* rCB, const double* q2Nh, const double* q2N, const double* q1N,
                        const double* q1n1, const double* q2n2, long long M)
{
    long long i = (long long)blockIdx.x * TPB + threadIdx.x;
    if (i < M) {
        rhB[i] = 2.0 * q2Nh[i];
        rCA[i] = q2N[i] + q1n1[i] * q2n2[i];
        rCB[i] = q1N[i] + q1n1[i] * q2n2[i];
    }
}
__global__ void pk_acc(double* r, const double* a, const double* b, int w, long long M)
{
    long long i = (long long)blockIdx.x * TPB + threadIdx.x;
    if (i < M) r[i] += w * a[i] * b[i];
}
__global__ void pk_final(double* rA, double* rB, double* rh, double* rCBo, double* rC, const double* rhB, const double* rCA,
                         const double* rCB, double zB, double zC, double ds, long long M)
{
    long long i = (long long)blockIdx.x * TPB + threadIdx.x;
    if (i < M) {
        const double hB = rhB[i] * zB * ds / 3.0, cA = rCA[i] * zC * ds / 3.0, cB = rCB[i] * zC * ds / 3.0;
        rA[i] = cA;
        rB[i] = hB + cB;
        rh[i] = hB;
        rCBo[i] = cB;
        rC[i] = cA + cB;
    }
}

inline unsigned nblk(long long n) { return (unsigned)((n + TPB - 1) / TPB); }

}  // namespace

struct plain_ctx {
    int m[3];
    long long M, Mk;
    int N;
    hipfftHandle fwd, bwd;
    double *W, *ksq, *k12, *k6, *e4A, *e2A, *e4B, *e2B, *qs0, *qs1, *qq, *red, *out;  // out: rA rB rh rCB rC rhB rCA rCB'
    hipfftDoubleComplex* h;
};

extern "C" {

plain_ctx* plain_create(const int m[3], const double D[3], int N)
{
    plain_ctx* c = new plain_ctx();
    for (int i = 0; i < 3; ++i) c->m[i] = m[i];
    c->M = (long long)m[0] * m[1] * m[2];
    c->Mk = (long long)m[0] * m[1] * (m[2] / 2 + 1);
    c->N = N;
    const long long M = c->M, Mk = c->Mk;
    // k^2 on the half spectrum (setup_k_sq, density.h:79-95)
    std::vector<double> ks(Mk);
    const double pi = 4 * atan(1.0);
    for (int k0 = -(m[0] - 1) / 2; k0 <= m[0] / 2; k0++)
        for (int k1 = -(m[1] - 1) / 2; k1 <= m[1] / 2; k1++)
            for (int k2 = 0; k2 <= m[2] / 2; k2++) {
                const int K0 = k0 < 0 ? k0 + m[0] : k0, K1 = k1 < 0 ? k1 + m[1] : k1;
                ks[k2 + (m[2] / 2 + 1) * (K1 + m[1] * K0)] =
                    4 * pi * pi * (k0 * k0 / (D[0] * D[0]) + k1 * k1 / (D[1] * D[1]) + k2 * k2 / (D[2] * D[2]));
            }
    bool ok = hipMalloc(&c->W, 2 * M * 8) == hipSuccess && hipMalloc(&c->ksq, Mk * 8) == hipSuccess &&
              hipMalloc(&c->k12, Mk * 8) == hipSuccess && hipMalloc(&c->k6, Mk * 8) == hipSuccess &&
              hipMalloc(&c->e4A, M * 8) == hipSuccess && hipMalloc(&c->e2A, M * 8) == hipSuccess &&
              hipMalloc(&c->e4B, M * 8) == hipSuccess && hipMalloc(&c->e2B, M * 8) == hipSuccess &&
              hipMalloc(&c->qs0, 2 * M * 8) == hipSuccess && hipMalloc(&c->qs1, 2 * M * 8) == hipSuccess &&
              hipMalloc(&c->qq, 2 * (N + 1) * M * 8) == hipSuccess && hipMalloc(&c->red, 8) == hipSuccess &&
              hipMalloc(&c->out, 8 * M * 8) == hipSuccess &&
              hipMalloc(&c->h, 2 * Mk * sizeof(hipfftDoubleComplex)) == hipSuccess;
    if (!ok) return nullptr;
    hipMemcpy(c->ksq, ks.data(), Mk * 8, hipMemcpyHostToDevice);
    int n[3] = {m[0], m[1], m[2]};
    if (hipfftPlanMany(&c->fwd, 3, n, nullptr, 1, 0, nullptr, 1, 0, HIPFFT_D2Z, 2) != HIPFFT_SUCCESS ||
        hipfftPlanMany(&c->bwd, 3, n, nullptr, 1, 0, nullptr, 1, 0, HIPFFT_Z2D, 2) != HIPFFT_SUCCESS)
        return nullptr;
    return c;
}

void plain_destroy(plain_ctx* c)
{
    if (!c) return;
    hipfftDestroy(c->fwd);
    hipfftDestroy(c->bwd);
    void* p[] = {c->W, c->ksq, c->k12, c->k6, c->e4A, c->e2A, c->e4B, c->e2B, c->qs0, c->qs1, c->qq, c->red, c->out, c->h};
    for (void* q : p) hipFree(q);
    delete c;
}

// w[0:2M] in, w[2M:6M] = phiA, phiB, phih, phicB out (density.h:344-459); 0 on success
int plain_props(plain_ctx* c, double* w, double* lnQ, int Ns, double alpha, double phic, double f, int ens)
{
    const long long M = c->M, Mk = c->Mk;
    const int n1 = (int)std::round(Ns * f), n2 = Ns - n1, Nh = (int)(alpha * Ns);
    const double ds = 1.0 / Ns;
    if (hipMemcpy(c->W, w, 2 * M * 8, hipMemcpyHostToDevice) != hipSuccess) return 1;
    pk_kfac<<<nblk(Mk), TPB>>>(c->k12, c->k6, c->ksq, ds, (double)M, Mk);
    pk_wfac<<<nblk(M), TPB>>>(c->e4A, c->e2A, c->e4B, c->e2B, c->W, ds, M);
    auto q1 = [&](int s) { return c->qq + 2 * (long long)s * M; };
    auto q2 = [&](int s) { return c->qq + (2 * (long long)s + 1) * M; };
    pk_set<<<nblk(M), TPB>>>(q1(0), 1.0, M);
    pk_set<<<nblk(M), TPB>>>(q2(0), 1.0, M);
    for (int s = 0; s < Ns; ++s) {
        const double* in = q1(s);
        double* outp = q1(s + 1);
        const double *a4 = s < n1 ? c->e4A : c->e4B, *a2 = s < n1 ? c->e2A : c->e2B;
        const double *b4 = s < n2 ? c->e4B : c->e4A, *b2 = s < n2 ? c->e2B : c->e2A;
        pk_mul<<<nblk(M), TPB>>>(c->qs0, in, a4, M);
        pk_mul<<<nblk(M), TPB>>>(c->qs0 + M, in + M, b4, M);
        hipfftExecD2Z(c->fwd, c->qs0, c->h);
        pk_mul_spec<<<nblk(Mk), TPB>>>(c->h, c->k12, Mk);
        pk_mul_spec<<<nblk(Mk), TPB>>>(c->h + Mk, c->k12, Mk);
        hipfftExecZ2D(c->bwd, c->h, c->qs0);
        pk_mul_self<<<nblk(M), TPB>>>(c->qs0, a2, M);
        pk_mul_self<<<nblk(M), TPB>>>(c->qs0 + M, b2, M);
        hipfftExecD2Z(c->fwd, c->qs0, c->h);
        pk_mul_spec<<<nblk(Mk), TPB>>>(c->h, c->k12, Mk);
        pk_mul_spec<<<nblk(Mk), TPB>>>(c->h + Mk, c->k12, Mk);
        hipfftExecZ2D(c->bwd, c->h, c->qs0);
        pk_mul<<<nblk(M), TPB>>>(c->qs1, in, a2, M);
        pk_mul<<<nblk(M), TPB>>>(c->qs1 + M, in + M, b2, M);
        hipfftExecD2Z(c->fwd, c->qs1, c->h);
        pk_mul_spec<<<nblk(Mk), TPB>>>(c->h, c->k6, Mk);
        pk_mul_spec<<<nblk(Mk), TPB>>>(c->h + Mk, c->k6, Mk);
        hipfftExecZ2D(c->bwd, c->h, c->qs1);
        pk_rich<<<nblk(M), TPB>>>(outp, c->qs0, c->qs1, a4, a2, M);
        pk_rich<<<nblk(M), TPB>>>(outp + M, c->qs0 + M, c->qs1 + M, b4, b2, M);
    }
    double Q2 = 0, QC = 0;
    pk_sum1<<<1, 1024>>>(q2(Nh), c->red, M);
    hipMemcpy(&Q2, c->red, 8, hipMemcpyDeviceToHost);
    pk_sum1<<<1, 1024>>>(q1(Ns), c->red, M);
    hipMemcpy(&QC, c->red, 8, hipMemcpyDeviceToHost);
    Q2 /= M;
    QC /= M;
    double zB, zC;
    if (ens == 1) {
        *lnQ = phic * log(QC) + ((1.0 - phic) / alpha) * log(Q2);
        zB = (1.0 - phic) / (Q2 * alpha);
        zC = phic / QC;
    } else {
        zB = 1.0;
        zC = phic;
        *lnQ = Q2 + phic * QC;
    }
    double *rA = c->out, *rB = rA + M, *rh = rB + M, *rCBo = rh + M, *rC = rCBo + M, *rhB = rC + M, *rCA = rhB + M,
           *rCB = rCA + M;
    pk_ends<<<nblk(M), TPB>>>(rhB, rCA, rCB, q2(Nh), q2(Ns), q1(Ns), q1(n1), q2(n2), M);
    int sidx, wt;
    for (sidx = 1, wt = 4; sidx < Nh; sidx++, wt = 6 - wt) pk_acc<<<nblk(M), TPB>>>(rhB, q2(sidx), q2(Nh - sidx), wt, M);
    for (sidx = 1, wt = 4; sidx < n1; sidx++, wt = 6 - wt) pk_acc<<<nblk(M), TPB>>>(rCA, q1(sidx), q2(Ns - sidx), wt, M);
    for (sidx = n1 + 1, wt = 4; sidx < Ns; sidx++, wt = 6 - wt)
        pk_acc<<<nblk(M), TPB>>>(rCB, q1(sidx), q2(Ns - sidx), wt, M);
    pk_final<<<nblk(M), TPB>>>(rA, rB, rh, rCBo, rC, rhB, rCA, rCB, zB, zC, ds, M);
    double tmp;
    const double* means[] = {rC, rA, rB, rhB};
    for (const double* p : means) {
        pk_sum1<<<1, 1024>>>(p, c->red, M);
        hipMemcpy(&tmp, c->red, 8, hipMemcpyDeviceToHost);
    }
    return hipMemcpy(w + 2 * M, rA, 4 * M * 8, hipMemcpyDeviceToHost) == hipSuccess ? 0 : 2;
}

}  // extern "C"
