// prop_types.hpp -- types shared by the propagator kernels and their host driver.
#pragma once
#include <type_traits>
#include "fft_engine.hpp"

namespace scft {

constexpr int RED_NT = 256;  // threads of the reduction / pointwise kernels

template <typename T>
struct PropArgs {
    long long M;           // grid points
    int m0, m1, m2;        // grid (x slowest, z fastest)
    int N, n1, n2, Nh;     // contour steps, block split, homopolymer length
    int nb;                // active batch slots
    int pencil_ct;         // line kernels: use the compile-time-length ones where they exist
    int rlines;            // split line step: register-FFT y and z passes where compiled
    int coltile;           // two-stage column pass: tile width (0: the grid's default)
    int pgrid;             // planes with a persistent plane_mid: > 0 runs it on that many
                           // workgroups; 0 launches one workgroup per plane
    int planent;           // two-stage plane kernels: workgroup size (0: the grid's default)
    cpx<T>* Ch;            // [slot][M] half-step sandwich work field
    cpx<T>* Cf;            // [slot][M] full-step sandwich work field
    const T* e4A;          // [slot][M] quarter-step factor of species A (Boltz<T> form)
    const T* e4B;          // [slot][M] the same for species B
    cpx<T>* qs;            // [slot][qpitch][M] stored (q1, q2)
    int qpitch;            // q-store entries per slot (N + 1 when every contour index is kept)
    int qck;               // 0: every contour index is kept; K > 0: checkpoints (ck_index)
    int qlo, qhi, qbase;   // qck > 0 and qlo <= qhi: a recompute window, see q_index
    int qsrc;              // q-store index a resumed segment starts from
    const cpx<T>* tw0;     // twiddle tables, lengths m0, m1, m2
    const cpx<T>* tw1;
    const cpx<T>* tw2;
    const T* gx;           // per-axis kinetic tables (see Kinetic<T>)
    const T* gy;
    const T* gz;
    T invM;
};

// Store to the q store: written once per contour step and read only by the density sums
// after the contour loop. A non-temporal store keeps this stream (N+1 fields per props)
// from displacing the step's working fields (Ch, Cf, factors) from the caches.
template <typename T>
__device__ __forceinline__ void store_q(cpx<T>* p, cpx<T> v)
{
    using V = T __attribute__((ext_vector_type(2)));
    const V w = {v.x, v.y};
    __builtin_nontemporal_store(w, reinterpret_cast<V*>(p));
}

// Checkpointed q store (qck = K > 0), per slot: entries 0 .. N/K hold contour indices
// 0, K, 2K, ...; entry N/K + 1 holds index N and N/K + 2 index Nh (the partition
// functions read them); then two recompute windows of K + 1 entries each.
__host__ __device__ constexpr int ck_pitch(int N, int K) { return N / K + 3 + 2 * (K + 1); }
__host__ __device__ constexpr int ck_window(int N, int K, int w) { return N / K + 3 + w * (K + 1); }
__host__ __device__ constexpr int ck_index(int s, int N, int Nh, int K)
{
    return s % K == 0 ? s / K : s == N ? N / K + 1 : s == Nh ? N / K + 2 : -1;
}

// q-store index of contour index s, -1 when the store does not keep it: s itself (full
// store), a checkpoint entry (the forward pass of a checkpointed store), or, in a recompute
// window, qbase + s - qlo for qlo <= s <= qhi.
template <typename T>
__device__ __forceinline__ int q_index(const PropArgs<T>& a, int s)
{
    if (a.qck == 0) return s;
    if (a.qlo <= a.qhi) return (s >= a.qlo && s <= a.qhi) ? a.qbase + (s - a.qlo) : -1;
    return ck_index(s, a.N, a.Nh, a.qck);
}

// Per-slot scalars (density.h:376-399) in FP64 from the factored sums:
// scal[slot][SCAL_N] = {lnQ, sC, sH, QC, Q2, Q1, QC', Q2', sHA, Q1'}. sC, sH, sHA multiply
// the copolymer, B-homopolymer and A-homopolymer Simpson sums directly; primes are the
// partition functions of the mean-field-factored propagators.
constexpr int SCAL_N = 10;

// phihA: the A homopolymer the reference declares but leaves at zA = 0 (density.h:352,
// 390-397). Canonical: its volume fraction, the B homopolymer taking 1 - phic - phihA;
// grand canonical: its activity zA. It has the B homopolymer's length Nh and uses the
// copolymer's A-end propagator q1 up to Nh, so Nh <= n1 (density.h:405-407).
struct ChainConst {
    int N, n1, n2, Nh, ens;
    double alpha, phic, ds, phihA;
};

// Boltzmann-factor arithmetic per precision. a = (w - <w>) ds / 4 is the exponent of the
// quarter-step factor E1/4 = exp(-a) (GPU_kernels.h:220); E1/2 = (E1/4)^2 (GPU_kernels.h:221).
//
// FP64 stores E1/4 itself and multiplies exactly as the reference does.
// FP32 stores u = 1 - E1/4 = -expm1(-a) instead. |u| << 1, so rounding u to FP32 moves
// the factor by ~|u| 2^-24 rather than 2^-24. Storing E1/4 in FP32 would bias every
// point by the same relative error on each of the 4 uses per step, for all N steps. That
// systematic drift is ~1e-5 in q(N) and in the free energy at N = 128. q*E1/4 is then
// q - q*u, and q*E1/2 is q - q*v with v = u(2 - u), whose roundings vary with q and so
// average out over the grid.
template <typename T>
struct Boltz;

// x / 3 by multiplication: q = x c with c = RN(1/3), then one residual correction
// r = x - 3 q (exact under FMA), q + r c. This is the correctly rounded quotient: checked for
// every FP32 mantissa over 33 binades and for 2e8 random FP64 values against IEEE division,
// with no difference. Three instructions where the IEEE division sequence takes about ten
// (FP32) or twenty (FP64); the Richardson combine does one per point and part per step.
__host__ __device__ __forceinline__ float div3(float x)
{
    constexpr float c = 1.0f / 3.0f;
    const float q = x * c;
    return fmaf(fmaf(-3.0f, q, x), c, q);
}
__host__ __device__ __forceinline__ double div3(double x)
{
    constexpr double c = 1.0 / 3.0;
    const double q = x * c;
    return fma(fma(-3.0, q, x), c, q);
}

template <>
struct Boltz<double> {
    static __host__ __device__ __forceinline__ double from_arg(double a) { return exp(-a); }
    static __host__ __device__ __forceinline__ double mul4(double q, double e) { return q * e; }
    static __host__ __device__ __forceinline__ double mul2(double q, double e) { return q * (e * e); }
    // (4 qs0 E1/4 - qs1 E1/2) / 3  (GPU_kernels.h:202)
    static __host__ __device__ __forceinline__ double rich(double h0, double h1, double e)
    {
        return div3(h0 * e * 4.0 - h1 * (e * e));
    }
};

template <>
struct Boltz<float> {
    static __host__ __device__ __forceinline__ float from_arg(double a) { return (float)(-expm1(-a)); }
    static __host__ __device__ __forceinline__ float mul4(float q, float u) { return q - q * u; }
    static __host__ __device__ __forceinline__ float mul2(float q, float u)
    {
        const float v = u * (2.0f - u);
        return q - q * v;
    }
    static __host__ __device__ __forceinline__ float rich(float h0, float h1, float u)
    {
        const float v = u * (2.0f - u);
        return div3((h0 - h0 * u) * 4.0f - (h1 - h1 * v));
    }
};

// The inverse transforms' normalisation 1/L. In FP32 a rounded 1/L (L = 96, 9216, ...)
// scales the mean of q by (1 + d), |d| up to 3e-8, on every sandwich, the same way on every
// step: 3N sandwiches move lnQ by ~3N d. It is applied as inv.hi + inv.lo (inv.hi = RN(1/L),
// inv.lo = RN(1/L - inv.hi)), exact to FP32 rounding of the product. Powers of two have
// inv.lo = 0. FP64 takes RN(1/L), whose bias (~1e-16 per sandwich) is far below its bar.
template <typename T>
struct InvLen {
    T hi, lo;
    __device__ __forceinline__ explicit InvLen(int L)
    {
        const double r = 1.0 / (double)L;
        hi = (T)r;
        lo = std::is_same<T, float>::value ? (T)(r - (double)hi) : T(0);
    }
    __device__ __forceinline__ void apply(cpx<T>& v) const
    {
        if constexpr (std::is_same<T, float>::value) {
            v.x = fmaf(v.x, lo, v.x * hi);
            v.y = fmaf(v.y, lo, v.y * hi);
        } else {
            v.x *= hi;
            v.y *= hi;
        }
    }
};

// Kinetic-factor arithmetic per precision. With a_i = k_i^2 ds/12 per axis
// (setup_k_sq, density.h:79-95), the half-step factor is exp(-(ax+ay+az))/M and the
// full-step factor its square times M (GPU_kernels.h:207-213).
//
// FP64 tables hold g_i = exp(-a_i), so the factor is gx gy gz / M.
// FP32 tables hold kappa_i = 1 - exp(-a_i) = -expm1(-a_i). The factor is applied as
// (q - q kappa) / M, with kappa = 1 - prod(1 - kappa_i) assembled from sums of small
// positive terms. A low-k mode decays by kappa ~ 1e-3 per step, so rounding the factor
// itself to FP32 would misstate that decay by 2^-24 / kappa, identically on every step.
template <typename T>
struct Kinetic;

template <>
struct Kinetic<double> {
    static double table(double a) { return exp(-a); }
    // the factor of an axis left out (a = 0)
    static __host__ __device__ __forceinline__ double one() { return 1.0; }
    // per-column part (y, z)
    static __host__ __device__ __forceinline__ double col(double gy, double gz) { return gy * gz; }
    static __host__ __device__ __forceinline__ void apply(cpx<double>& v, double gx, double gyz, bool full,
                                                          double invM)
    {
        double g = gx * gyz;
        if (full) g = g * g;
        g = g * invM;
        v.x *= g;
        v.y *= g;
    }
};

template <>
struct Kinetic<float> {
    static float table(double a) { return (float)(-expm1(-a)); }
    static __host__ __device__ __forceinline__ float one() { return 0.0f; }
    static __host__ __device__ __forceinline__ float col(float ky, float kz) { return ky + kz - ky * kz; }
    static __host__ __device__ __forceinline__ void apply(cpx<float>& v, float kx, float kyz, bool full, float invM)
    {
        float k = kx + kyz - kx * kyz;
        if (full) k = k * (2.0f - k);
        v.x = (v.x - v.x * k) * invM;
        v.y = (v.y - v.y * k) * invM;
    }
};

// Box-derivative constants of the gradient kernels: ds, 2 pi / D_i and D_i.
struct GradConst {
    double ds;
    double kD[3];  // 2 pi / D_i
    double D[3];
};

}  // namespace scft
