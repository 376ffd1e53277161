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
// At fixed nodes x the natural cubic spline of one grid point's column a[0..R) (cubefit_chunk_g,
// GPU_kernels.h:232-284) and its derivative or value at any point (eval_spline_*_chunk_g,
// GPU_kernels.h:286-329) are linear in a. The host turns x (and the evaluation points) into
// one R x R weight matrix per launch (spline_weights, scft_api.hip); here every column r of
// the replica rows is multiplied by it: out[erows[j]][r] = sum_i G[j][i] in[i][r].
// A workgroup stages RM_TR consecutive columns of all R rows in LDS (coalesced row loads),
// then each wave produces whole output rows from LDS, the weights broadcast across the wave.
// The whole tile is read before any of it is written, so out may be in (in-place update).
constexpr int RM_TR = 64;  // columns per tile: one wave's lanes

__global__ void __launch_bounds__(FK_NT) k_row_mix(int R, long long M, const double* __restrict__ G,
                                                   const int* __restrict__ erows, int ne, const double* in,
                                                   long long istride, double* out, long long ostride)
{
    extern __shared__ double lds[];
    double* S = lds;                 // [R][RM_TR]
    double* Gs = lds + R * RM_TR;    // [ne][R]
    for (int t = threadIdx.x; t < ne * R; t += FK_NT) Gs[t] = G[t];
    const int lane = threadIdx.x % RM_TR, grp = threadIdx.x / RM_TR;
    constexpr int NG = FK_NT / RM_TR;
    for (long long r0 = (long long)blockIdx.x * RM_TR; r0 < M; r0 += (long long)gridDim.x * RM_TR) {
        const long long r = r0 + lane;
        __syncthreads();  // the previous tile's LDS reads are done (and Gs is written)
        for (int i = grp; i < R; i += NG) S[i * RM_TR + lane] = r < M ? in[(size_t)i * istride + r] : 0.0;
        __syncthreads();
        if (r < M)
            for (int j = grp; j < ne; j += NG) {
                const double* g = Gs + j * R;
                double v = 0.0;
                for (int i = 0; i < R; ++i) v += g[i] * S[i * RM_TR + lane];
                out[(size_t)erows[j] * ostride + r] = v;
            }
    }
}

size_t row_mix_lds_bytes(int R, int ne) { return sizeof(double) * ((size_t)R * RM_TR + (size_t)ne * R); }

hipError_t row_mix_setup()
{
    return hipFuncSetAttribute((const void*)k_row_mix, hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)row_mix_lds_bytes(SPLINE_RMAX, SPLINE_RMAX));
}

void launch_row_mix(int R, long long M, const double* d_G, const int* d_erows, int ne, const double* in,
                    long long istride, double* out, long long ostride, hipStream_t st)
{
    if (ne <= 0 || M <= 0) return;
    const long long tiles = (M + RM_TR - 1) / RM_TR;
    const int grid = (int)(tiles < 8192 ? tiles : 8192);
    hipLaunchKernelGGL(k_row_mix, dim3(grid), dim3(FK_NT), row_mix_lds_bytes(R, ne), st, R, M, d_G, d_erows, ne, in,
                       istride, out, ostride);
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
