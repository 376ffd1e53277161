// field_kernels.hpp -- FP64 field update of the string method on the device
// (scft.cu:516-853). All replicas' fields live in one [R][2M] array; reductions are two-stage
// (block partials in a fixed order) so results are deterministic run to run.
#pragma once
#include <hip/hip_runtime.h>

namespace scft {

constexpr int FK_NT = 256;
constexpr int SPLINE_RMAX = 64;  // replicas supported by the per-point spline kernels

// solve_field0 body (scft.cu:527-543) for nb slots: DEV- = 0, DEV+ = chi (phiB + phiA - 1),
// W += lambda DEV, partial sums of DEV^2. part layout [slot][nblk].
void launch_wplus_mix(const int* d_slot2rep, int nb, long long M, double chi, const double* d_lambda,
                      const double* phi, double* W, double* DEV, double* part, int nblk, hipStream_t st);

// solve_field DEV (scft.cu:604-608) for nb slots; spring term when spring != 0.
// part [slot][nblk][2] = (sum DEV^2 over 2M, sum DEV-^2 over M)
void launch_dev_full(const int* d_slot2rep, int nb, int R, long long M, double chi, int spring, double eps,
                     const double* phi, const double* W, double* DEV, double* part, int nblk, hipStream_t st);

// partial sums of (W-[i+1] - W-[i])^2 for i < R-1 (scft.cu:351-355); part [i][nblk]
void launch_edist(int R, long long M, const double* W, double* part, int nblk, hipStream_t st);

// finalize: out[slot*ncomp + c] = sum over nblk of part[(slot*nblk + b)*ncomp + c]
void launch_finalize(const double* part, int nslots, int nblk, int ncomp, double* out, hipStream_t st);

// One row of the spline weights: g[i] = derivative (deriv) or value at xx of the natural
// cubic spline through (x[i], a[i]), i < R (cubefit_chunk_g, GPU_kernels.h:232-284), per unit
// a[i]; xx lies on the interval of the last node strictly below it (GPU_kernels.h:291-297,
// 313-319). Host and device share it. 2 <= R <= SPLINE_RMAX.
__host__ __device__ inline void spline_weight_row(int R, const double* x, double xx, bool deriv, double* g)
{
    const int n = R - 1;
    double h[SPLINE_RMAX], l[SPLINE_RMAX], u[SPLINE_RMAX], z[SPLINE_RMAX], c[SPLINE_RMAX];
    for (int i = 0; i < R; ++i) h[i] = l[i] = u[i] = 0.0;
    for (int i = 0; i < n; ++i) h[i] = x[i + 1] - x[i];
    l[0] = 1.0;
    for (int i = 1; i < n; ++i) {
        l[i] = 2.0 * (x[i + 1] - x[i - 1]) - h[i - 1] * u[i - 1];
        u[i] = h[i] / l[i];
    }
    int k = 0;
    for (int i = 0; i < R; ++i) {
        if (x[i] < xx) k = i;
        else break;
    }
    if (k > n - 1) k = n - 1;  // only an x that is not increasing gets here
    const double t = xx - x[k];
    for (int col = 0; col < R; ++col) {  // column col of the linear map: a = unit vector
        z[0] = 0.0;
        for (int i = 1; i < n; ++i) {
            const double a1 = i + 1 == col, a0 = i == col, am = i - 1 == col;
            const double A = 3.0 * (a1 - a0) / h[i] - 3.0 * (a0 - am) / h[i - 1];
            z[i] = (A - h[i - 1] * z[i - 1]) / l[i];
        }
        c[n] = 0.0;
        for (int j = n - 1; j >= 0; --j) c[j] = z[j] - u[j] * c[j + 1];
        const double ak = k == col, ak1 = k + 1 == col;
        const double b = (ak1 - ak) / h[k] - h[k] * (c[k + 1] + 2.0 * c[k]) / 3.0;
        const double d = (c[k + 1] - c[k]) / (3.0 * h[k]);
        g[col] = deriv ? b + 2.0 * t * c[k] + 3.0 * t * t * d : ak + t * b + t * t * c[k] + t * t * t * d;
    }
}

// out[erows[j]][r] = sum_i G[j][i] in[i][r] for j < ne, r < M (G: [ne][R] device, row
// strides istride / ostride; out may equal in). The spline tangent and redistribution of
// the string (GPU_kernels.h:232-329) with the weights of spline_weights. R <= SPLINE_RMAX.
void launch_row_mix(int R, long long M, const double* d_G, const int* d_erows, int ne, const double* in,
                    long long istride, double* out, long long ostride, hipStream_t st);
// its dynamic LDS bound (call once per process before the first launch)
hipError_t row_mix_setup();

// message of the calling thread's last failed call that has no context (scft_noalloc.hip)
const char* noctx_error();

// part [slot][nblk][2] = (sum dWda^2, sum DEV- dWda) per slot (scft.cu:649-667)
void launch_tangent_dots(const int* d_slot2rep, int nb, long long M, const double* dWda, const double* DEV,
                         double* part, int nblk, hipStream_t st);

// projection (scft.cu:670-674) + fixed-lambda mixing (scft.cu:690-693) for every replica:
// DEV- -= cproj[ii] dWda; W += lambda DEV; part [ii][nblk][2] = (sum DEV'^2 over 2M,
// sum DEV-' dWda)
void launch_project_mix(int R, long long M, const double* d_cproj, double lambda, const double* dWda,
                        double* DEV, double* W, double* part, int nblk, hipStream_t st);

// FreeE sums (scft.cu:823-828) for nb slots: part [slot][nblk][2] =
// (sum chi phiA phiB - wa phiA - wb phiB, sum phiA)
void launch_fe_sums(const int* d_slot2rep, int nb, long long M, double chi, const double* W, const double* phi,
                    double* part, int nblk, hipStream_t st);

// Anderson mixing (opt-in, dostring = 0): dots of D with listed ring rows -> part[nblk][n];
// W = sum_q coef[q] (Wr[rows[q]] + lam Dr[rows[q]]); W += lam D. All over M2 = 2M entries.
void launch_and_dots(const double* D, const double* ring, const int* rows, int n, long long M2, double* part,
                     int nblk, hipStream_t st);
void launch_and_combine(double* W, const double* Wr, const double* Dr, const int* rows, const double* coef, int n,
                        double lam, long long M2, int nblk, hipStream_t st);
void launch_axpy(double* W, const double* D, double lam, long long M2, int nblk, hipStream_t st);

// part [slot][nblk] = block sums of phiA phiB
void launch_phiab_sums(const int* d_slot2rep, int nb, long long M, const double* phi, double* part, int nblk,
                       hipStream_t st);

}  // namespace scft
