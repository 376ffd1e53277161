// field_kernels.hip -- FP64 string-method field update kernels (scft.cu:516-853).
#include "field_kernels.hpp"

namespace scft {

__device__ __forceinline__ double fk_block_sum(double v, double* sh)
{
    const int tid = threadIdx.x;
    sh[tid] = v;
    __syncthreads();
#pragma unroll
    for (int o = FK_NT / 2; o > 0; o >>= 1) {
        if (tid < o) sh[tid] += sh[tid + o];
        __syncthreads();
    }
    const double r = sh[0];
    __syncthreads();
    return r;
}

#define GRID_STRIDE(r, M) \
    for (long long r = (long long)blockIdx.x * FK_NT + threadIdx.x; r < (M); r += (long long)gridDim.x * FK_NT)

// ---------------------------------------------------------------- solve_field0 body
__global__ void __launch_bounds__(FK_NT) k_wplus_mix(const int* __restrict__ slot2rep, long long M, double chi,
                                                     const double* __restrict__ lambda,
                                                     const double* __restrict__ phi, double* __restrict__ W,
                                                     double* __restrict__ DEV, double* __restrict__ part)
{
    __shared__ double sh[FK_NT];
    const int slot = blockIdx.y, rep = slot2rep[slot];
    const double lam = lambda[rep];
    const double* ph = phi + (size_t)rep * 4 * M;
    double* w = W + (size_t)rep * 2 * M;
    double* dv = DEV + (size_t)rep * 2 * M;
    double s1 = 0;
    GRID_STRIDE(r, M)
    {
        const double dp = chi * (ph[M + r] + ph[r] - 1.0);  // phiB + phiA (scft.cu:529)
        dv[r] = 0;
        dv[r + M] = dp;
        s1 += dp * dp;
        w[r + M] = w[r + M] + lam * dp;  // W- + lam*0 is W- itself
    }
    s1 = fk_block_sum(s1, sh);
    if (threadIdx.x == 0) part[(size_t)slot * gridDim.x + blockIdx.x] = s1;
}

void launch_wplus_mix(const int* d_slot2rep, int nb, long long M, double chi, const double* d_lambda,
                      const double* phi, double* W, double* DEV, double* part, int nblk, hipStream_t st)
{
    if (nb <= 0) return;
    hipLaunchKernelGGL(k_wplus_mix, dim3(nblk, nb), dim3(FK_NT), 0, st, d_slot2rep, M, chi, d_lambda, phi, W, DEV,
                       part);
}

// ---------------------------------------------------------------- DEV of solve_field
__global__ void __launch_bounds__(FK_NT) k_dev_full(const int* __restrict__ slot2rep, int R, long long M, double chi,
                                                    int spring, double eps, const double* __restrict__ phi,
                                                    const double* __restrict__ W, double* __restrict__ DEV,
                                                    double* __restrict__ part)
{
    __shared__ double sh[FK_NT];
    const int slot = blockIdx.y, ii = slot2rep[slot];
    const double* ph = phi + (size_t)ii * 4 * M;
    const double* w = W + (size_t)ii * 2 * M;
    double* dv = DEV + (size_t)ii * 2 * M;
    const bool spr = spring && ii > 0 && ii < R - 1;
    const double* wl = W + (size_t)(spr ? ii - 1 : ii) * 2 * M;
    const double* wr = W + (size_t)(spr ? ii + 1 : ii) * 2 * M;
    double s2 = 0, sm = 0;
    GRID_STRIDE(r, M)
    {
        const double pa = ph[r], pb = ph[M + r];
        double dm = chi * (pb - pa) - 2.0 * w[r];
        if (spr) dm += eps * (wr[r] + wl[r] - 2.0 * w[r]);
        const double dp = chi * (pb + pa - 1.0);
        dv[r] = dm;
        dv[r + M] = dp;
        s2 += dm * dm + dp * dp;
        sm += dm * dm;
    }
    s2 = fk_block_sum(s2, sh);
    sm = fk_block_sum(sm, sh);
    if (threadIdx.x == 0) {
        part[((size_t)slot * gridDim.x + blockIdx.x) * 2 + 0] = s2;
        part[((size_t)slot * gridDim.x + blockIdx.x) * 2 + 1] = sm;
    }
}

void launch_dev_full(const int* d_slot2rep, int nb, int R, long long M, double chi, int spring, double eps,
                     const double* phi, const double* W, double* DEV, double* part, int nblk, hipStream_t st)
{
    if (nb <= 0) return;
    hipLaunchKernelGGL(k_dev_full, dim3(nblk, nb), dim3(FK_NT), 0, st, d_slot2rep, R, M, chi, spring, eps, phi, W,
                       DEV, part);
}

// ---------------------------------------------------------------- arc length
__global__ void __launch_bounds__(FK_NT) k_edist(long long M, const double* __restrict__ W, double* __restrict__ part)
{
    __shared__ double sh[FK_NT];
    const int i = blockIdx.y;
    const double* a = W + (size_t)i * 2 * M;
    const double* b = W + (size_t)(i + 1) * 2 * M;
    double s = 0;
    GRID_STRIDE(r, M)
    {
        const double d = a[r] - b[r];
        s += d * d;
    }
    s = fk_block_sum(s, sh);
    if (threadIdx.x == 0) part[(size_t)i * gridDim.x + blockIdx.x] = s;
}

void launch_edist(int R, long long M, const double* W, double* part, int nblk, hipStream_t st)
{
    if (R < 2) return;
    hipLaunchKernelGGL(k_edist, dim3(nblk, R - 1), dim3(FK_NT), 0, st, M, W, part);
}

__global__ void __launch_bounds__(FK_NT) k_finalize(const double* __restrict__ part, int nblk, int ncomp,
                                                    double* __restrict__ out)
{
    __shared__ double sh[FK_NT];
    const int slot = blockIdx.x;
    for (int c = 0; c < ncomp; ++c) {
        double v = 0;
        for (int b = threadIdx.x; b < nblk; b += FK_NT) v += part[((size_t)slot * nblk + b) * ncomp + c];
        v = fk_block_sum(v, sh);
        if (threadIdx.x == 0) out[(size_t)slot * ncomp + c] = v;
    }
}

void launch_finalize(const double* part, int nslots, int nblk, int ncomp, double* out, hipStream_t st)
{
    if (nslots <= 0) return;
    hipLaunchKernelGGL(k_finalize, dim3(nslots), dim3(FK_NT), 0, st, part, nblk, ncomp, out);
}

// ---------------------------------------------------------------- natural cubic spline
// The data-independent parts of cubefit_chunk_g (GPU_kernels.h:253-265): h, l, u, plus
// the evaluation intervals of eval_spline_*_chunk_g (GPU_kernels.h:291-297), shared by
// every grid point.
struct SplineShared {
    double x[SPLINE_RMAX], h[SPLINE_RMAX], l[SPLINE_RMAX], u[SPLINE_RMAX];
    int nm[SPLINE_RMAX];
    double dx[SPLINE_RMAX];
};

__device__ void spline_setup(SplineShared& S, int R, const double* __restrict__ x, const double* __restrict__ xe)
{
    const int n = R - 1;
    for (int i = threadIdx.x; i < R; i += FK_NT) S.x[i] = x[i];
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int i = 0; i < R; ++i) S.h[i] = S.l[i] = S.u[i] = 0.0;
        for (int i = 0; i <= n - 1; ++i) S.h[i] = S.x[i + 1] - S.x[i];
        S.l[0] = 1.0;
        S.u[0] = 0.0;
        for (int i = 1; i <= n - 1; ++i) {
            S.l[i] = 2.0 * (S.x[i + 1] - S.x[i - 1]) - S.h[i - 1] * S.u[i - 1];
            S.u[i] = S.h[i] / S.l[i];
        }
        S.l[n] = 1.0;
        for (int e = 0; e < R; ++e) {
            const double xx = xe[e];
            int nm = 0;
            for (int i = 0; i < R; ++i) {
                if (S.x[i] < xx) nm = i;
                else break;
            }
            if (nm > n - 1) nm = n - 1;  // only reachable for x that is not increasing
            if (nm < 0) nm = 0;
            S.nm[e] = nm;
            S.dx[e] = xx - S.x[nm];
        }
    }
    __syncthreads();
}

// c coefficients of one grid point's natural spline through a[0..R) (GPU_kernels.h:254-275)
__device__ __forceinline__ void spline_c(const SplineShared& S, int R, const double* a, double* c)
{
    const int n = R - 1;
    double z[SPLINE_RMAX];
    z[0] = 0.0;
    for (int i = 1; i <= n - 1; ++i) {
        const double A = 3.0 * (a[i + 1] - a[i]) / S.h[i] - 3.0 * (a[i] - a[i - 1]) / S.h[i - 1];
        z[i] = (A - S.h[i - 1] * z[i - 1]) / S.l[i];
    }
    c[n] = 0.0;
    for (int j = n - 1; j >= 0; --j) c[j] = z[j] - S.u[j] * c[j + 1];
}

__global__ void __launch_bounds__(FK_NT) k_spline_deriv(int R, long long M, const double* __restrict__ x,
                                                        const int* __restrict__ active, const double* __restrict__ W,
                                                        double* __restrict__ dWda)
{
    __shared__ SplineShared S;
    spline_setup(S, R, x, x);
    GRID_STRIDE(r, M)
    {
        double a[SPLINE_RMAX], c[SPLINE_RMAX];
        for (int i = 0; i < R; ++i) a[i] = W[(size_t)i * 2 * M + r];
        spline_c(S, R, a, c);
        for (int e = 0; e < R; ++e) {
            if (!active[e]) continue;
            const int nm = S.nm[e];
            const double h = S.h[nm], dx = S.dx[e];
            const double b = (a[nm + 1] - a[nm]) / h - h * (c[nm + 1] + 2.0 * c[nm]) / 3.0;
            const double d = (c[nm + 1] - c[nm]) / (3.0 * h);
            dWda[(size_t)e * M + r] = b + 2.0 * dx * c[nm] + 3.0 * dx * dx * d;
        }
    }
}

// part [slot][nblk][2] = (sum dWda^2, sum DEV- dWda)  (scft.cu:649-667)
__global__ void __launch_bounds__(FK_NT) k_tangent_dots(const int* __restrict__ slot2rep, long long M,
                                                        const double* __restrict__ dWda,
                                                        const double* __restrict__ DEV, double* __restrict__ part)
{
    __shared__ double sh[FK_NT];
    const int slot = blockIdx.y, ii = slot2rep[slot];
    const double* t = dWda + (size_t)ii * M;
    const double* dv = DEV + (size_t)ii * 2 * M;
    double s1 = 0, s2 = 0;
    GRID_STRIDE(r, M)
    {
        const double v = t[r];
        s1 += v * v;
        s2 += dv[r] * v;
    }
    s1 = fk_block_sum(s1, sh);
    s2 = fk_block_sum(s2, sh);
    if (threadIdx.x == 0) {
        part[((size_t)slot * gridDim.x + blockIdx.x) * 2 + 0] = s1;
        part[((size_t)slot * gridDim.x + blockIdx.x) * 2 + 1] = s2;
    }
}

void launch_spline_deriv(int R, long long M, const double* d_x, const int* d_active, const double* W,
                         double* dWda, hipStream_t st)
{
    const int nblk = (int)((M + FK_NT - 1) / FK_NT);
    hipLaunchKernelGGL(k_spline_deriv, dim3(nblk < 2048 ? nblk : 2048), dim3(FK_NT), 0, st, R, M, d_x, d_active, W,
                       dWda);
}

void launch_tangent_dots(const int* d_slot2rep, int nb, long long M, const double* dWda, const double* DEV,
                         double* part, int nblk, hipStream_t st)
{
    if (nb <= 0) return;
    hipLaunchKernelGGL(k_tangent_dots, dim3(nblk, nb), dim3(FK_NT), 0, st, d_slot2rep, M, dWda, DEV, part);
}

__global__ void __launch_bounds__(FK_NT) k_project_mix(long long M, const double* __restrict__ cproj, double lambda,
                                                       const double* __restrict__ dWda, double* __restrict__ DEV,
                                                       double* __restrict__ W, double* __restrict__ part)
{
    __shared__ double sh[FK_NT];
    const int ii = blockIdx.y;
    const double cp = cproj[ii];
    double* dv = DEV + (size_t)ii * 2 * M;
    double* w = W + (size_t)ii * 2 * M;
    const double* t = dWda + (size_t)ii * M;
    double s2 = 0, sd = 0;
    GRID_STRIDE(r, M)
    {
        double dm = dv[r];
        const double dp = dv[r + M];
        if (cp != 0.0) {
            const double tv = t[r];
            dm -= cp * tv;
            dv[r] = dm;
            sd += dm * tv;
        }
        s2 += dm * dm + dp * dp;
        w[r] = w[r] + lambda * dm;
        w[r + M] = w[r + M] + lambda * dp;
    }
    s2 = fk_block_sum(s2, sh);
    sd = fk_block_sum(sd, sh);
    if (threadIdx.x == 0) {
        part[((size_t)ii * gridDim.x + blockIdx.x) * 2 + 0] = s2;
        part[((size_t)ii * gridDim.x + blockIdx.x) * 2 + 1] = sd;
    }
}

void launch_project_mix(int R, long long M, const double* d_cproj, double lambda, const double* dWda, double* DEV,
                        double* W, double* part, int nblk, hipStream_t st)
{
    hipLaunchKernelGGL(k_project_mix, dim3(nblk, R), dim3(FK_NT), 0, st, M, d_cproj, lambda, dWda, DEV, W, part);
}

__global__ void __launch_bounds__(FK_NT) k_redistribute(int R, long long M, const double* __restrict__ x,
                                                        const double* __restrict__ xref, const int* __restrict__ active,
                                                        double lambda, double* __restrict__ W)
{
    __shared__ SplineShared S;
    spline_setup(S, R, x, xref);
    GRID_STRIDE(r, M)
    {
        double a[SPLINE_RMAX], c[SPLINE_RMAX];
        for (int i = 0; i < R; ++i) a[i] = W[(size_t)i * 2 * M + r];
        spline_c(S, R, a, c);
        for (int e = 1; e < R - 1; ++e) {
            if (!active[e]) continue;
            const int nm = S.nm[e];
            const double h = S.h[nm], dx = S.dx[e];
            const double b = (a[nm + 1] - a[nm]) / h - h * (c[nm + 1] + 2.0 * c[nm]) / 3.0;
            const double d = (c[nm + 1] - c[nm]) / (3.0 * h);
            const double v = a[nm] + dx * b + dx * dx * c[nm] + dx * dx * dx * d;
            W[(size_t)e * 2 * M + r] = a[e] + (v - a[e]) * lambda;
        }
    }
}

void launch_redistribute(int R, long long M, const double* d_x, const double* d_xref, const int* d_active,
                         double lambda, double* W, hipStream_t st)
{
    const int nblk = (int)((M + FK_NT - 1) / FK_NT);
    hipLaunchKernelGGL(k_redistribute, dim3(nblk < 2048 ? nblk : 2048), dim3(FK_NT), 0, st, R, M, d_x, d_xref,
                       d_active, lambda, W);
}

// ---------------------------------------------------------------- free energy sums
__global__ void __launch_bounds__(FK_NT) k_fe_sums(const int* __restrict__ slot2rep, long long M, double chi,
                                                   const double* __restrict__ W, const double* __restrict__ phi,
                                                   double* __restrict__ part)
{
    __shared__ double sh[FK_NT];
    const int slot = blockIdx.y, ii = slot2rep[slot];
    const double* w = W + (size_t)ii * 2 * M;
    const double* ph = phi + (size_t)ii * 4 * M;
    double s1 = 0, sa = 0;
    GRID_STRIDE(r, M)
    {
        const double wa = w[r] + w[r + M], wb = w[r + M] - w[r];
        const double pa = ph[r], pb = ph[M + r];
        s1 += chi * pa * pb - wa * pa - wb * pb;
        sa += pa;
    }
    s1 = fk_block_sum(s1, sh);
    sa = fk_block_sum(sa, sh);
    if (threadIdx.x == 0) {
        part[((size_t)slot * gridDim.x + blockIdx.x) * 2 + 0] = s1;
        part[((size_t)slot * gridDim.x + blockIdx.x) * 2 + 1] = sa;
    }
}

void launch_fe_sums(const int* d_slot2rep, int nb, long long M, double chi, const double* W, const double* phi,
                    double* part, int nblk, hipStream_t st)
{
    if (nb <= 0) return;
    hipLaunchKernelGGL(k_fe_sums, dim3(nblk, nb), dim3(FK_NT), 0, st, d_slot2rep, M, chi, W, phi, part);
}

// ---------------------------------------------------------------- Anderson mixing
// part[blk][q] = block sums of D . ring[rows[q]] over M2 entries, q < n
__global__ void __launch_bounds__(FK_NT) k_and_dots(const double* __restrict__ D, const double* __restrict__ ring,
                                                    const int* __restrict__ rows, int n, long long M2,
                                                    double* __restrict__ part)
{
    __shared__ double sh[FK_NT];
    for (int q = 0; q < n; ++q) {
        const double* Dj = ring + (size_t)rows[q] * (size_t)M2;
        double s = 0;
        GRID_STRIDE(r, M2) { s += D[r] * Dj[r]; }
        s = fk_block_sum(s, sh);
        if (threadIdx.x == 0) part[(size_t)blockIdx.x * n + q] = s;
    }
}

// W = sum_q coef[q] (Wr[rows[q]] + lam Dr[rows[q]]), in the order q = 0, 1, ...
__global__ void __launch_bounds__(FK_NT) k_and_combine(double* __restrict__ W, const double* __restrict__ Wr,
                                                       const double* __restrict__ Dr, const int* __restrict__ rows,
                                                       const double* __restrict__ coef, int n, double lam,
                                                       long long M2)
{
    GRID_STRIDE(r, M2)
    {
        double v = 0;
        for (int q = 0; q < n; ++q) {
            const size_t o = (size_t)rows[q] * (size_t)M2 + r;
            v += coef[q] * (Wr[o] + lam * Dr[o]);
        }
        W[r] = v;
    }
}

// W += lam D over M2 entries
__global__ void __launch_bounds__(FK_NT) k_axpy(double* __restrict__ W, const double* __restrict__ D, double lam,
                                                long long M2)
{
    GRID_STRIDE(r, M2) { W[r] += lam * D[r]; }
}

void launch_and_dots(const double* D, const double* ring, const int* rows, int n, long long M2, double* part,
                     int nblk, hipStream_t st)
{
    hipLaunchKernelGGL(k_and_dots, dim3(nblk), dim3(FK_NT), 0, st, D, ring, rows, n, M2, part);
}

void launch_and_combine(double* W, const double* Wr, const double* Dr, const int* rows, const double* coef, int n,
                        double lam, long long M2, int nblk, hipStream_t st)
{
    hipLaunchKernelGGL(k_and_combine, dim3(nblk), dim3(FK_NT), 0, st, W, Wr, Dr, rows, coef, n, lam, M2);
}

void launch_axpy(double* W, const double* D, double lam, long long M2, int nblk, hipStream_t st)
{
    hipLaunchKernelGGL(k_axpy, dim3(nblk), dim3(FK_NT), 0, st, W, D, lam, M2);
}

// part[slot][nblk] = block sums of phiA phiB (the chi derivative of FEs1, scft.cu:823-828)
__global__ void __launch_bounds__(FK_NT) k_phiab_sums(const int* __restrict__ slot2rep, long long M,
                                                      const double* __restrict__ phi, double* __restrict__ part)
{
    __shared__ double sh[FK_NT];
    const int slot = blockIdx.y;
    const double* ph = phi + (size_t)slot2rep[slot] * 4 * M;
    double s = 0;
    GRID_STRIDE(r, M) { s += ph[r] * ph[M + r]; }
    s = fk_block_sum(s, sh);
    if (threadIdx.x == 0) part[(size_t)slot * gridDim.x + blockIdx.x] = s;
}

void launch_phiab_sums(const int* d_slot2rep, int nb, long long M, const double* phi, double* part, int nblk,
                       hipStream_t st)
{
    if (nb <= 0) return;
    hipLaunchKernelGGL(k_phiab_sums, dim3(nblk, nb), dim3(FK_NT), 0, st, d_slot2rep, M, phi, part);
}

}  // namespace scft
