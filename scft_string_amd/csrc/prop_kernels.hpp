// prop_kernels.hpp -- the chain-propagator hot path for gfx950.
//
// Layout. Each batch slot b holds one replica's pair {q, q-dagger} packed as ONE complex
// field z = q + i q-dagger over the m0 x m1 x m2 grid (x slowest, z fastest), so every
// FFT is complex-to-complex. This is exact for the reference's operator: the kinetic
// factor exp(-k^2 ds/..) is real and even in k, so F^-1 K F maps q + i q' to
// (F^-1 K F q) + i (F^-1 K F q') (density.h:149-205 applies the same sandwich to q and
// q-dagger separately, batch 2). The Boltzmann factors are per-chain real multiplies.
//
// One 3-D FFT = "plane" pass (2-D FFT over the (y,z) plane of one x, in LDS) + "column"
// pass (1-D FFT along x for a tile of 64 (y,z) columns, in LDS). Since the kinetic
// factor is diagonal in k, the column kernel does forward-x, multiply, inverse-x in one
// pass, and the plane kernels fuse inverse-2D -> Boltzmann/Richardson -> forward-2D.
// Per contour step s (density.h:129-220):
//   col(Ch,Cf) -> plane_mid(Ch) -> col(Ch) -> plane_end
// where Ch carries the two half steps and Cf the full step; plane_end forms
// q(s+1) = (4 qs0 e4 - qs1 e2)/3 (GPU_kernels.h:199-203), stores it, and starts step s+1.
//
// Mean-field factoring. Each species' Boltzmann factor is built from w - <w>
// (exp(-(w-c) ds/4)); the step operator scales exactly by exp(-c ds) under a uniform
// shift, so the stored q' relate to the reference's q by known exp(-sum c ds) factors that
// are applied in FP64 to Q and to the density prefactors. This keeps FP32 in range when W+
// drifts in the canonical ensemble.
#pragma once
#include "prop_types.hpp"

namespace scft {

template <int N1, int N2>
struct PlaneGeom {
    static constexpr int P2 = N2 + ((N2 % 2 == 0) ? 1 : 0);  // odd row pitch in LDS
    static constexpr int NE = N1 * N2;
};

template <typename T, int N1, int N2, int NT>
__device__ __forceinline__ void plane_fft(cpx<T>* buf, const cpx<T>* tw1, const cpx<T>* tw2,
                                          int tid, bool inv)
{
    constexpr int P2 = PlaneGeom<N1, N2>::P2;
    if (inv) {
        fft_lines<T, N1, true, N2, 1, P2, NT>(buf, tw1, tid);  // y columns
        fft_lines<T, N2, true, N1, P2, 1, NT>(buf, tw2, tid);  // z rows
    } else {
        fft_lines<T, N2, false, N1, P2, 1, NT>(buf, tw2, tid);
        fft_lines<T, N1, false, N2, 1, P2, NT>(buf, tw1, tid);
    }
}

template <typename T, int N1, int N2, int NT>
__device__ __forceinline__ void plane_load(cpx<T>* buf, const cpx<T>* __restrict__ g, int tid)
{
    constexpr int P2 = PlaneGeom<N1, N2>::P2, NE = N1 * N2, PE = (NE + NT - 1) / NT;
#pragma unroll
    for (int i = 0; i < PE; ++i) {
        const int e = tid + i * NT;
        if (NE % NT == 0 || e < NE) buf[(e / N2) * P2 + (e % N2)] = g[e];
    }
}

template <typename T, int N1, int N2, int NT>
__device__ __forceinline__ void plane_store(const cpx<T>* buf, cpx<T>* __restrict__ g, int tid)
{
    constexpr int P2 = PlaneGeom<N1, N2>::P2, NE = N1 * N2, PE = (NE + NT - 1) / NT;
#pragma unroll
    for (int i = 0; i < PE; ++i) {
        const int e = tid + i * NT;
        if (NE % NT == 0 || e < NE) g[e] = buf[(e / N2) * P2 + (e % N2)];
    }
}

template <typename T, int N1, int N2>
__device__ __forceinline__ void load_twiddles(cpx<T>* tw1, cpx<T>* tw2, const PropArgs<T>& a, int tid,
                                              int nt)
{
    for (int i = tid; i < N1; i += nt) tw1[i] = a.tw1[i];
    for (int i = tid; i < N2; i += nt) tw2[i] = a.tw2[i];
}

// PL_RESUME starts a segment of the contour loop from a stored q (q-store index a.qsrc) at
// contour index s, as PL_BEGIN starts it from q = 1 at index 0.
enum PlaneMode { PL_BEGIN = 0, PL_STEP = 1, PL_LAST = 2, PL_RESUME = 3 };

// plane_end: completes step s (PL_STEP/PL_LAST), starts the contour (PL_BEGIN) or resumes it
// from a stored q at index s (PL_RESUME), stores q where the q store keeps that index, then
// (unless PL_LAST) launches the next step's two forward 2-D transforms.
// grid: (m0 planes, nb slots), block NT, dynamic LDS = plane + twiddles.
template <typename T, int N1, int N2, int NT>
__global__ void __launch_bounds__(NT) k_plane_end(PropArgs<T> a, int s, int mode)
{
    constexpr int P2 = PlaneGeom<N1, N2>::P2, NE = N1 * N2, PE = (NE + NT - 1) / NT;
    extern __shared__ unsigned char smem_raw[];
    cpx<T>* buf = reinterpret_cast<cpx<T>*>(smem_raw);
    cpx<T>* tw1 = buf + N1 * P2;
    cpx<T>* tw2 = tw1 + N1;
    const int tid = threadIdx.x;
    const int x = blockIdx.x, slot = blockIdx.y;
    const size_t fo = (size_t)slot * (size_t)a.M + (size_t)x * NE;

    load_twiddles<T, N1, N2>(tw1, tw2, a, tid, NT);

    T ea[PE], eb[PE];
#pragma unroll
    for (int i = 0; i < PE; ++i) {
        const int e = tid + i * NT;
        if (NE % NT == 0 || e < NE) {
            ea[i] = a.e4A[fo + e];
            eb[i] = a.e4B[fo + e];
        } else {
            ea[i] = eb[i] = T(0);
        }
    }

    cpx<T> q[PE];
    int sout;
    if (mode == PL_BEGIN) {
#pragma unroll
        for (int i = 0; i < PE; ++i) q[i] = {T(1), T(1)};
        sout = 0;
    } else if (mode == PL_RESUME) {
        const cpx<T>* qi = a.qs + ((size_t)slot * (size_t)a.qpitch + (size_t)a.qsrc) * (size_t)a.M + (size_t)x * NE;
#pragma unroll
        for (int i = 0; i < PE; ++i) {
            const int e = tid + i * NT;
            if (NE % NT == 0 || e < NE) q[i] = qi[e];
        }
        sout = s;
    } else {
        cpx<T> h0[PE];
        plane_load<T, N1, N2, NT>(buf, a.Ch + fo, tid);
        __syncthreads();
        plane_fft<T, N1, N2, NT>(buf, tw1, tw2, tid, true);
#pragma unroll
        for (int i = 0; i < PE; ++i) {
            const int e = tid + i * NT;
            if (NE % NT == 0 || e < NE) h0[i] = buf[(e / N2) * P2 + (e % N2)];
        }
        __syncthreads();
        plane_load<T, N1, N2, NT>(buf, a.Cf + fo, tid);
        __syncthreads();
        plane_fft<T, N1, N2, NT>(buf, tw1, tw2, tid, true);
        const bool qa = s < a.n1, pb = s < a.n2;  // species of this step (density.h:136-146)
#pragma unroll
        for (int i = 0; i < PE; ++i) {
            const int e = tid + i * NT;
            if (NE % NT == 0 || e < NE) {
                const cpx<T> h1 = buf[(e / N2) * P2 + (e % N2)];
                const T ex = qa ? ea[i] : eb[i];
                const T ey = pb ? eb[i] : ea[i];
                q[i].x = Boltz<T>::rich(h0[i].x, h1.x, ex);
                q[i].y = Boltz<T>::rich(h0[i].y, h1.y, ey);
            }
        }
        sout = s + 1;
        __syncthreads();
    }

    const int qix = q_index(a, sout);
    if (qix >= 0) {
        cpx<T>* qo = a.qs + ((size_t)slot * (size_t)a.qpitch + (size_t)qix) * (size_t)a.M + (size_t)x * NE;
#pragma unroll
        for (int i = 0; i < PE; ++i) {
            const int e = tid + i * NT;
            if (NE % NT == 0 || e < NE) store_q(&qo[e], q[i]);
        }
    }
    if (mode == PL_LAST) return;

    const int sn = sout;  // next step index
    const bool qa = sn < a.n1, pb = sn < a.n2;
    // half-step chain: E1/4 q
#pragma unroll
    for (int i = 0; i < PE; ++i) {
        const int e = tid + i * NT;
        if (NE % NT == 0 || e < NE) {
            const T ex = qa ? ea[i] : eb[i];
            const T ey = pb ? eb[i] : ea[i];
            buf[(e / N2) * P2 + (e % N2)] = {Boltz<T>::mul4(q[i].x, ex), Boltz<T>::mul4(q[i].y, ey)};
        }
    }
    __syncthreads();
    plane_fft<T, N1, N2, NT>(buf, tw1, tw2, tid, false);
    plane_store<T, N1, N2, NT>(buf, a.Ch + fo, tid);
    __syncthreads();
    // full-step chain: E1/2 q
#pragma unroll
    for (int i = 0; i < PE; ++i) {
        const int e = tid + i * NT;
        if (NE % NT == 0 || e < NE) {
            const T ex = qa ? ea[i] : eb[i];
            const T ey = pb ? eb[i] : ea[i];
            buf[(e / N2) * P2 + (e % N2)] = {Boltz<T>::mul2(q[i].x, ex), Boltz<T>::mul2(q[i].y, ey)};
        }
    }
    __syncthreads();
    plane_fft<T, N1, N2, NT>(buf, tw1, tw2, tid, false);
    plane_store<T, N1, N2, NT>(buf, a.Cf + fo, tid);
}

// plane_mid: Ch <- F2( E1/2 * F2^-1(Ch) ), the middle of the two half steps
// (density.h:156-173).
template <typename T, int N1, int N2, int NT>
__global__ void __launch_bounds__(NT) k_plane_mid(PropArgs<T> a, int s)
{
    constexpr int P2 = PlaneGeom<N1, N2>::P2, NE = N1 * N2, PE = (NE + NT - 1) / NT;
    extern __shared__ unsigned char smem_raw[];
    cpx<T>* buf = reinterpret_cast<cpx<T>*>(smem_raw);
    cpx<T>* tw1 = buf + N1 * P2;
    cpx<T>* tw2 = tw1 + N1;
    const int tid = threadIdx.x;
    const int x = blockIdx.x, slot = blockIdx.y;
    const size_t fo = (size_t)slot * (size_t)a.M + (size_t)x * NE;
    load_twiddles<T, N1, N2>(tw1, tw2, a, tid, NT);
    const bool qa = s < a.n1, pb = s < a.n2;
    T ex[PE], ey[PE];
#pragma unroll
    for (int i = 0; i < PE; ++i) {
        const int e = tid + i * NT;
        if (NE % NT == 0 || e < NE) {
            const T va = a.e4A[fo + e], vb = a.e4B[fo + e];
            const T x4 = qa ? va : vb, y4 = pb ? vb : va;
            ex[i] = x4;
            ey[i] = y4;
        }
    }
    plane_load<T, N1, N2, NT>(buf, a.Ch + fo, tid);
    __syncthreads();
    plane_fft<T, N1, N2, NT>(buf, tw1, tw2, tid, true);
#pragma unroll
    for (int i = 0; i < PE; ++i) {
        const int e = tid + i * NT;
        if (NE % NT == 0 || e < NE) {
            cpx<T>& v = buf[(e / N2) * P2 + (e % N2)];
            v.x = Boltz<T>::mul2(v.x, ex[i]);
            v.y = Boltz<T>::mul2(v.y, ey[i]);
        }
    }
    __syncthreads();
    plane_fft<T, N1, N2, NT>(buf, tw1, tw2, tid, false);
    plane_store<T, N1, N2, NT>(buf, a.Ch + fo, tid);
}

// column pass: for a tile of TC consecutive (y,z) columns of one field, forward FFT along
// x, multiply by exp(-k^2 ds/12)/M (half) or exp(-k^2 ds/6)/M (full), inverse FFT along x
// (density.h:149-156 / 198-205 with GPU_kernels.h:175-181, 207-213).
// grid: (ceil(m1*m2/TC) tiles, nh + nf fields); fields [0,nh) are Ch[0..nh), [nh, nh+nf)
// are Cf[0..nf).
template <typename T, int N0, int N1, int N2, int TC, int NT>
__global__ void __launch_bounds__(NT) k_column(PropArgs<T> a, int nh, int nf)
{
    constexpr int NC = N1 * N2;           // columns per field
    constexpr int PE = (N0 * TC + NT - 1) / NT;
    extern __shared__ unsigned char smem_raw[];
    cpx<T>* buf = reinterpret_cast<cpx<T>*>(smem_raw);  // [N0][TC]
    cpx<T>* tw0 = buf + N0 * TC;
    T* gx = reinterpret_cast<T*>(tw0 + N0);
    const int tid = threadIdx.x;
    const int tile = blockIdx.x, f = blockIdx.y;
    const bool full = f >= nh;
    cpx<T>* fld = full ? a.Cf + (size_t)(f - nh) * (size_t)a.M : a.Ch + (size_t)f * (size_t)a.M;
    const int c0 = tile * TC;
    for (int i = tid; i < N0; i += NT) {
        tw0[i] = a.tw0[i];
        gx[i] = a.gx[i];
    }
#pragma unroll
    for (int i = 0; i < PE; ++i) {
        const int e = tid + i * NT;
        if ((N0 * TC) % NT == 0 || e < N0 * TC) {
            const int row = e / TC, c = e % TC;
            const int col = c0 + c;
            buf[e] = (NC % TC == 0 || col < NC) ? fld[(size_t)row * NC + col] : cpx<T>{T(0), T(0)};
        }
    }
    __syncthreads();
    fft_lines<T, N0, false, TC, 1, TC, NT>(buf, tw0, tid);
#pragma unroll
    for (int i = 0; i < PE; ++i) {
        const int e = tid + i * NT;
        if ((N0 * TC) % NT == 0 || e < N0 * TC) {
            const int row = e / TC, c = e % TC;
            const int col = c0 + c;
            const int y = col / N2, z = col % N2;
            Kinetic<T>::apply(buf[e], gx[row], Kinetic<T>::col(a.gy[y < N1 ? y : 0], a.gz[z]), full, a.invM);
        }
    }
    __syncthreads();
    fft_lines<T, N0, true, TC, 1, TC, NT>(buf, tw0, tid);
#pragma unroll
    for (int i = 0; i < PE; ++i) {
        const int e = tid + i * NT;
        if ((N0 * TC) % NT == 0 || e < N0 * TC) {
            const int row = e / TC, c = e % TC;
            const int col = c0 + c;
            if (NC % TC == 0 || col < NC) fld[(size_t)row * NC + col] = buf[e];
        }
    }
}

// ---------------------------------------------------------------- FP64 reductions
// Block tree sum of RED_NT doubles in LDS; thread 0 returns the total.
__device__ __forceinline__ double block_sum(double v, double* sh)
{
    const int tid = threadIdx.x;
    sh[tid] = v;
    __syncthreads();
#pragma unroll
    for (int o = RED_NT / 2; o > 0; o >>= 1) {
        if (tid < o) sh[tid] += sh[tid + o];
        __syncthreads();
    }
    const double r = sh[0];
    __syncthreads();
    return r;
}

// Per-slot sums of W- and W+ (the mean-field shifts). grid (nblk, nb); partial [slot][nblk][2]
__global__ void __launch_bounds__(RED_NT) k_wsum(const double* __restrict__ W, long long wstride,
                                                 const int* __restrict__ slot2rep, long long M,
                                                 double* __restrict__ part)
{
    __shared__ double sh[RED_NT];
    const int slot = blockIdx.y;
    const double* w = W + (size_t)slot2rep[slot] * (size_t)wstride;
    double sm = 0, sp = 0;
    for (long long r = (long long)blockIdx.x * RED_NT + threadIdx.x; r < M; r += (long long)gridDim.x * RED_NT) {
        sm += w[r];
        sp += w[r + M];
    }
    sm = block_sum(sm, sh);
    sp = block_sum(sp, sh);
    if (threadIdx.x == 0) {
        part[((size_t)slot * gridDim.x + blockIdx.x) * 2 + 0] = sm;
        part[((size_t)slot * gridDim.x + blockIdx.x) * 2 + 1] = sp;
    }
}

// Fixed-order sum of nblk partials with stride `ncomp`; one block per slot.
__device__ __forceinline__ double sum_partials(const double* part, int nblk, int ncomp, int comp, double* sh)
{
    double v = 0;
    for (int i = threadIdx.x; i < nblk; i += RED_NT) v += part[(size_t)i * ncomp + comp];
    return block_sum(v, sh);
}

// cAB[slot] = (<wA>, <wB>) with wA = W+ + W-, wB = W+ - W-
__global__ void __launch_bounds__(RED_NT) k_wmean(const double* __restrict__ part, int nblk, long long M,
                                                  double* __restrict__ cAB)
{
    __shared__ double sh[RED_NT];
    const int slot = blockIdx.x;
    const double* p = part + (size_t)slot * nblk * 2;
    const double sm = sum_partials(p, nblk, 2, 0, sh);
    const double sp = sum_partials(p, nblk, 2, 1, sh);
    if (threadIdx.x == 0) {
        const double mm = sm / (double)M, mp = sp / (double)M;
        cAB[2 * slot + 0] = mp + mm;
        cAB[2 * slot + 1] = mp - mm;
    }
}

// Boltzmann factors of A and B with a = (W- + W+ - cA) ds/4, (-W- + W+ - cB) ds/4
// (GPU_kernels.h:215-225): E1/4 itself in FP64, 1 - E1/4 in FP32 (see Boltz)
template <typename T>
__global__ void __launch_bounds__(RED_NT) k_factors(const double* __restrict__ W, long long wstride,
                                                    const int* __restrict__ slot2rep, long long M, double ds,
                                                    const double* __restrict__ cAB, T* __restrict__ e4A,
                                                    T* __restrict__ e4B)
{
    const int slot = blockIdx.y;
    const double* w = W + (size_t)slot2rep[slot] * (size_t)wstride;
    const double cA = cAB[2 * slot], cB = cAB[2 * slot + 1];
    for (long long r = (long long)blockIdx.x * RED_NT + threadIdx.x; r < M; r += (long long)gridDim.x * RED_NT) {
        const double wm = w[r], wp = w[r + M];
        e4A[(size_t)slot * M + r] = Boltz<T>::from_arg((wm + wp - cA) * ds / 4);
        e4B[(size_t)slot * M + r] = Boltz<T>::from_arg((-wm + wp - cB) * ds / 4);
    }
}

// partial sums of q1'[N], q2'[Nh] (density.h:377-381) and q1'[Nh] (Q1 of the A homopolymer)
template <typename T>
__global__ void __launch_bounds__(RED_NT) k_qsum(const cpx<T>* __restrict__ qs, long long M, int qpitch, int iN,
                                                 int iH, double* __restrict__ part)
{
    __shared__ double sh[RED_NT];
    const int slot = blockIdx.y;
    const cpx<T>* qN = qs + ((size_t)slot * qpitch + iN) * (size_t)M;  // contour index N
    const cpx<T>* qH = qs + ((size_t)slot * qpitch + iH) * (size_t)M;  // contour index Nh
    double sc = 0, sh2 = 0, sh1 = 0;
    for (long long r = (long long)blockIdx.x * RED_NT + threadIdx.x; r < M; r += (long long)gridDim.x * RED_NT) {
        sc += (double)qN[r].x;
        sh2 += (double)qH[r].y;
        sh1 += (double)qH[r].x;
    }
    sc = block_sum(sc, sh);
    sh2 = block_sum(sh2, sh);
    sh1 = block_sum(sh1, sh);
    if (threadIdx.x == 0) {
        double* o = part + ((size_t)slot * gridDim.x + blockIdx.x) * 3;
        o[0] = sc;
        o[1] = sh2;
        o[2] = sh1;
    }
}

__global__ void __launch_bounds__(RED_NT) k_scalars(const double* __restrict__ part, int nblk, long long M,
                                                    ChainConst cc, const double* __restrict__ cAB,
                                                    double* __restrict__ scal)
{
    __shared__ double sh[RED_NT];
    const int slot = blockIdx.x;
    const double* p = part + (size_t)slot * nblk * 3;
    const double sQC = sum_partials(p, nblk, 3, 0, sh);
    const double sQ2 = sum_partials(p, nblk, 3, 1, sh);
    const double sQ1 = sum_partials(p, nblk, 3, 2, sh);
    if (threadIdx.x != 0) return;
    const double cA = cAB[2 * slot], cB = cAB[2 * slot + 1];
    const double QCp = sQC / (double)M, Q2p = sQ2 / (double)M, Q1p = sQ1 / (double)M;
    const double LC = (cc.n1 * cA + cc.n2 * cB) * cc.ds;  // q1(N) = q1'(N) exp(-LC)
    const double LH = cc.Nh * cB * cc.ds;                  // q2(Nh) = q2'(Nh) exp(-LH)
    const double LA = cc.Nh * cA * cc.ds;                  // q1(Nh) = q1'(Nh) exp(-LA), Nh <= n1
    const double lQC = log(QCp) - LC, lQ2 = log(Q2p) - LH, lQ1 = log(Q1p) - LA;
    const double QC = exp(lQC), Q2 = exp(lQ2), Q1 = exp(lQ1);
    double lnQ, sC, sH, sHA;
    if (cc.ens == 1) {
        const double phihB = 1.0 - cc.phic - cc.phihA;
        lnQ = cc.phic * lQC + (phihB / cc.alpha) * lQ2;
        if (cc.phihA != 0.0) lnQ += (cc.phihA / cc.alpha) * lQ1;
        sC = cc.phic / QCp * cc.ds / 3.0;                    // zC e^{-LC} = phic / QC'
        sH = phihB / (Q2p * cc.alpha) * cc.ds / 3.0;         // zB e^{-LH}
        sHA = cc.phihA / (Q1p * cc.alpha) * cc.ds / 3.0;     // zA e^{-LA}
    } else {
        lnQ = cc.phihA * Q1 + 1.0 * Q2 + cc.phic * QC;
        sC = cc.phic * exp(-LC) * cc.ds / 3.0;
        sH = exp(-LH) * cc.ds / 3.0;
        sHA = cc.phihA * exp(-LA) * cc.ds / 3.0;
    }
    double* o = scal + (size_t)slot * SCAL_N;
    o[0] = lnQ; o[1] = sC; o[2] = sH; o[3] = QC; o[4] = Q2; o[5] = Q1; o[6] = QCp; o[7] = Q2p;
    o[8] = sHA; o[9] = Q1p;
}

// Simpson weights of density.h:401-422: end terms 1; interior 4,2,4,... from the first
// interior point whatever the parity (A.8).
__device__ __forceinline__ double wt_odd_first(int s_from_start) { return (s_from_start & 1) ? 4.0 : 2.0; }

// The density sums read each q-store slot once: non-temporal loads.
template <typename T>
__device__ __forceinline__ cpx<T> load_q(const cpx<T>* p)
{
    using V = T __attribute__((ext_vector_type(2)));
    const V v = __builtin_nontemporal_load(reinterpret_cast<const V*>(p));
    return {v.x, v.y};
}
constexpr int DU = 4;  // slot pairs per batch in k_density

// One point of the densities from the raw Simpson sums (density.h:423-430, GPU_kernels.h:
// 114-139): phi rows (phiA, phiB, phih, phicB), and the phis sums (rC, rA, rB, rhA, rhB).
__device__ __forceinline__ void density_point(double cA, double cB, double hA, double hB, double sC, double sH,
                                              double sHA, double* ph, long long r, long long M, double (&acc)[5])
{
    const double rCA = cA * sC, rCB = cB * sC, rhB = hB * sH, rhA = hA * sHA;
    const double rA = rhA + rCA, rB = rhB + rCB;
    ph[r] = rA;
    ph[r + M] = rB;
    ph[r + 2 * M] = rhA + rhB;
    ph[r + 3 * M] = rCB;
    acc[0] += rCA + rCB;
    acc[1] += rA;
    acc[2] += rB;
    acc[3] += rhA;
    acc[4] += rhB;
}

// the workgroup's phis partials to o[0..4]
__device__ __forceinline__ void density_partials(double (&acc)[5], double* sh, double* o)
{
    for (int c = 0; c < 5; ++c) acc[c] = block_sum(acc[c], sh);
    if (threadIdx.x == 0)
        for (int c = 0; c < 5; ++c) o[c] = acc[c];
}

// Weights of the copolymer term q1[i] q2[N-i] in the A-block and B-block sums
// (density.h:413-422): ends 1, the junction n1 in both, interior wt_odd_first from each
// block's start. A weight of 1 or 0 adds v or exactly nothing, as the branches did.
struct CoW {
    double a, b;
};
__device__ __forceinline__ CoW co_weight(int i, int n1, int N)
{
    if (i == 0) return {1.0, 0.0};
    if (i < n1) return {wt_odd_first(i), 0.0};
    if (i == n1) return {1.0, 1.0};
    if (i < N) return {0.0, wt_odd_first(i - n1)};
    return {0.0, 1.0};
}

// Densities (density.h:401-430): phi[rep] = (phiA, phiB, phih, phicB), and block partial
// sums of (rC, rA, rB, rhA, rhB) for phis (density.h:433-454).
template <typename T>
__global__ void __launch_bounds__(RED_NT) k_density(const cpx<T>* __restrict__ qs, long long M, ChainConst cc,
                                                    const double* __restrict__ scal, double* __restrict__ phi,
                                                    long long phistride, const int* __restrict__ slot2rep,
                                                    double* __restrict__ part)
{
    __shared__ double sh[RED_NT];
    const int slot = blockIdx.y;
    const int N = cc.N, n1 = cc.n1, Nh = cc.Nh;
    const cpx<T>* q = qs + (size_t)slot * (N + 1) * (size_t)M;
    const double sC = scal[(size_t)slot * SCAL_N + 1], sH = scal[(size_t)slot * SCAL_N + 2],
                 sHA = scal[(size_t)slot * SCAL_N + 8];
    double* ph = phi + (size_t)slot2rep[slot] * (size_t)phistride;
    double acc[5] = {0, 0, 0, 0, 0};
    for (long long r = (long long)blockIdx.x * RED_NT + threadIdx.x; r < M; r += (long long)gridDim.x * RED_NT) {
        // homopolymers: B = 2 q2[Nh] + sum_{s=1}^{Nh-1} w_s q2[s] q2[Nh-s] (density.h:409-412),
        // A the same over q1 (density.h:404-407), from the same slots' real parts. First, with
        // ordinary loads, so that the copolymer loop below finds slots 1..Nh in the cache
        const cpx<T> zh = q[(size_t)Nh * M + r];
        double hB = 2.0 * (double)zh.y, hA = 2.0 * (double)zh.x;
        for (int s = 1; 2 * s <= Nh; ++s) {
            const int t = Nh - s;
            if (t < 1) break;
            const cpx<T> zs = q[(size_t)s * M + r], zt = q[(size_t)t * M + r];
            const double w = (t != s) ? wt_odd_first(s) + wt_odd_first(t) : wt_odd_first(s);
            hB += w * ((double)zs.y * (double)zt.y);
            hA += w * ((double)zs.x * (double)zt.x);
        }
        double cA = 0, cB = 0;
        // read in pairs (s, N-s), DU pairs per batch: every load of a batch is in flight
        // before the first product (the slots are M apart, so nothing else orders them)
        const int S = N / 2;  // s = 0..S
        int s = 0;
        for (; s + DU - 1 <= S; s += DU) {
            cpx<T> zs[DU], zt[DU];
#pragma unroll
            for (int u = 0; u < DU; ++u) {
                zs[u] = load_q(q + (size_t)(s + u) * M + r);
                zt[u] = load_q(q + (size_t)(N - s - u) * M + r);
            }
#pragma unroll
            for (int u = 0; u < DU; ++u) {
                const int si = s + u, ti = N - s - u;
                const double v = (double)zs[u].x * (double)zt[u].y;
                const CoW w = co_weight(si, n1, N);
                cA += w.a * v;
                cB += w.b * v;
                if (ti != si) {
                    const double v2 = (double)zt[u].x * (double)zs[u].y;
                    const CoW w2 = co_weight(ti, n1, N);
                    cA += w2.a * v2;
                    cB += w2.b * v2;
                }
            }
        }
        for (; s <= S; ++s) {
            const cpx<T> zs = load_q(q + (size_t)s * M + r), zt = load_q(q + (size_t)(N - s) * M + r);
            const double v = (double)zs.x * (double)zt.y;
            const CoW w = co_weight(s, n1, N);
            cA += w.a * v;
            cB += w.b * v;
            if (N - s != s) {
                const double v2 = (double)zt.x * (double)zs.y;
                const CoW w2 = co_weight(N - s, n1, N);
                cA += w2.a * v2;
                cB += w2.b * v2;
            }
        }
        density_point(cA, cB, hA, hB, sC, sH, sHA, ph, r, M, acc);
    }
    density_partials(acc, sh, part + ((size_t)slot * gridDim.x + blockIdx.x) * 5);
}

// Checkpointed q store: the Simpson sums of k_density over recompute windows. The phi rows
// of a slot first hold the raw sums (cA, cB, hA, hB), zeroed by k_dens_zero. k_dens_win adds
// the pairs whose lower index s is in [s0, s1): copolymer (kind 0) pairs (s, N - s) or
// homopolymer (kind 1) pairs (s, Nh - s), with index s at q-store entry aBase + s - aLo and
// its partner t at bBase + t - bLo. k_dens_finish adds the homopolymer end terms 2 q[Nh] and
// turns the raw sums into phi and the phis partials as k_density does.
__global__ void __launch_bounds__(RED_NT) k_dens_zero(long long M, double* __restrict__ phi, long long phistride,
                                                      const int* __restrict__ slot2rep)
{
    double* ph = phi + (size_t)slot2rep[blockIdx.y] * (size_t)phistride;
    for (long long r = (long long)blockIdx.x * RED_NT + threadIdx.x; r < 4 * M; r += (long long)gridDim.x * RED_NT)
        ph[r] = 0.0;
}

template <typename T>
__global__ void __launch_bounds__(RED_NT) k_dens_win(const cpx<T>* __restrict__ qs, long long M, int qpitch,
                                                     ChainConst cc, int kind, int s0, int s1, int aLo, int aBase,
                                                     int bLo, int bBase, double* __restrict__ phi,
                                                     long long phistride, const int* __restrict__ slot2rep)
{
    const int slot = blockIdx.y;
    const int N = cc.N, n1 = cc.n1, L = kind == 0 ? cc.N : cc.Nh;
    const cpx<T>* q = qs + (size_t)slot * (size_t)qpitch * (size_t)M;
    double* ph = phi + (size_t)slot2rep[slot] * (size_t)phistride;
    for (long long r = (long long)blockIdx.x * RED_NT + threadIdx.x; r < M; r += (long long)gridDim.x * RED_NT) {
        double u0 = 0, u1 = 0;  // kind 0: cA, cB; kind 1: hA, hB
        for (int s = s0; s < s1; ++s) {
            const int t = L - s;
            const cpx<T> zs = load_q(q + (size_t)(aBase + s - aLo) * M + r);
            const cpx<T> zt = load_q(q + (size_t)(bBase + t - bLo) * M + r);
            if (kind == 0) {
                const double v = (double)zs.x * (double)zt.y;
                const CoW w = co_weight(s, n1, N);
                u0 += w.a * v;
                u1 += w.b * v;
                if (t != s) {
                    const double v2 = (double)zt.x * (double)zs.y;
                    const CoW w2 = co_weight(t, n1, N);
                    u0 += w2.a * v2;
                    u1 += w2.b * v2;
                }
            } else {
                const double w = (t != s) ? wt_odd_first(s) + wt_odd_first(t) : wt_odd_first(s);
                u0 += w * ((double)zs.x * (double)zt.x);
                u1 += w * ((double)zs.y * (double)zt.y);
            }
        }
        const long long o = kind == 0 ? 0 : 2 * M;
        ph[r + o] += u0;
        ph[r + o + M] += u1;
    }
}

template <typename T>
__global__ void __launch_bounds__(RED_NT) k_dens_finish(const cpx<T>* __restrict__ qs, long long M, int qpitch,
                                                        int iH, const double* __restrict__ scal,
                                                        double* __restrict__ phi, long long phistride,
                                                        const int* __restrict__ slot2rep, double* __restrict__ part)
{
    __shared__ double sh[RED_NT];
    const int slot = blockIdx.y;
    const cpx<T>* qh = qs + ((size_t)slot * (size_t)qpitch + (size_t)iH) * (size_t)M;
    const double sC = scal[(size_t)slot * SCAL_N + 1], sH = scal[(size_t)slot * SCAL_N + 2],
                 sHA = scal[(size_t)slot * SCAL_N + 8];
    double* ph = phi + (size_t)slot2rep[slot] * (size_t)phistride;
    double acc[5] = {0, 0, 0, 0, 0};
    for (long long r = (long long)blockIdx.x * RED_NT + threadIdx.x; r < M; r += (long long)gridDim.x * RED_NT) {
        const cpx<T> zh = load_q(qh + r);
        const double cA = ph[r], cB = ph[r + M];
        const double hA = 2.0 * (double)zh.x + ph[r + 2 * M], hB = 2.0 * (double)zh.y + ph[r + 3 * M];
        density_point(cA, cB, hA, hB, sC, sH, sHA, ph, r, M, acc);
    }
    density_partials(acc, sh, part + ((size_t)slot * gridDim.x + blockIdx.x) * 5);
}

__global__ void __launch_bounds__(RED_NT) k_phis(const double* __restrict__ part, int nblk, long long M,
                                                 double* __restrict__ phis)
{
    __shared__ double sh[RED_NT];
    const int slot = blockIdx.x;
    const double* p = part + (size_t)slot * nblk * 5;
    for (int c = 0; c < 5; ++c) {
        const double v = sum_partials(p, nblk, 5, c, sh);
        if (threadIdx.x == 0) phis[(size_t)slot * 5 + c] = v / (double)M;
    }
}

// Weight total 2 + sum_{s=1}^{n-1} w_s of one Simpson sum (density.h:405-422: end terms 1,
// interior 4, 2, 4, ... from the first interior point, whatever the parity).
__host__ __device__ inline double simpson_total(int n)
{
    double t = 2.0;
    for (int s = 1; s < n; ++s) t += (s & 1) ? 4.0 : 2.0;
    return t;
}

// FP32 normalisation compensation. Every Richardson step is a symmetric operator, so
// <q1'[s] q2'[L-s]> equals QC' (L = N) or Q2' (L = Nh) for every s, and each Simpson sum of
// k_density has an exact mean: <rCA> = sC QC' W(n1), <rCB> = sC QC' W(n2), <rhB> = sH Q2' W(Nh).
// With FP32 storage the invariant drifts with s by ~1e-7 (the plane spectra round the large
// low-k coefficients on every store), which shifts the density normalisation and the free
// energy by ~1e-6. fac[slot] = (exact mean / computed mean) of (rCA, rCB, rhB); in exact
// arithmetic all three are 1. phis (the computed means) are set to the exact ones.
// Used without the A homopolymer (phi rows then hold rCA, rhB and rCB separately).
__global__ void k_renorm_factors(ChainConst cc, const double* __restrict__ scal, double* __restrict__ phis,
                                 double* __restrict__ fac)
{
    const int slot = blockIdx.x;
    if (threadIdx.x != 0) return;
    const double* o = scal + (size_t)slot * SCAL_N;
    double* p = phis + (size_t)slot * 5;
    const double sC = o[1], sH = o[2], QCp = o[6], Q2p = o[7];
    const double tA = sC * QCp * simpson_total(cc.n1), tB = sC * QCp * simpson_total(cc.n2);
    const double tH = sH * Q2p * simpson_total(cc.Nh);
    const double mA = p[1] - p[3], mB = p[0] - mA, mH = p[4];  // <rCA>, <rCB>, <rhB>
    double* f = fac + (size_t)slot * 3;
    f[0] = mA != 0.0 ? tA / mA : 1.0;
    f[1] = mB != 0.0 ? tB / mB : 1.0;
    f[2] = mH != 0.0 ? tH / mH : 1.0;
    p[0] = tA + tB;
    p[1] = tA;
    p[2] = tH + tB;
    p[3] = 0.0;
    p[4] = tH;
}

// phi rows (phiA = rCA, phiB = rhB + rCB, phih = rhB, phicB = rCB) scaled by fac
__global__ void __launch_bounds__(RED_NT) k_renorm_apply(long long M, const double* __restrict__ fac,
                                                         double* __restrict__ phi, long long phistride,
                                                         const int* __restrict__ slot2rep)
{
    const int slot = blockIdx.y;
    const double fA = fac[3 * slot], fB = fac[3 * slot + 1], fH = fac[3 * slot + 2];
    double* ph = phi + (size_t)slot2rep[slot] * (size_t)phistride;
    for (long long r = (long long)blockIdx.x * RED_NT + threadIdx.x; r < M; r += (long long)gridDim.x * RED_NT) {
        const double h = ph[r + 2 * M] * fH, cB = ph[r + 3 * M] * fB;
        ph[r] *= fA;
        ph[r + M] = h + cB;
        ph[r + 2 * M] = h;
        ph[r + 3 * M] = cB;
    }
}

// Diagnostics: part [s][nblk] = sum_r q1'[s](r) q2'[N-s](r) for one slot (A4 invariant).
template <typename T>
__global__ void __launch_bounds__(RED_NT) k_qq(const cpx<T>* __restrict__ qs, long long M, int N, int slot,
                                               double* __restrict__ part)
{
    __shared__ double sh[RED_NT];
    const int s = blockIdx.y;
    const cpx<T>* a = qs + ((size_t)slot * (N + 1) + s) * (size_t)M;
    const cpx<T>* b = qs + ((size_t)slot * (N + 1) + (N - s)) * (size_t)M;
    double v = 0;
    for (long long r = (long long)blockIdx.x * RED_NT + threadIdx.x; r < M; r += (long long)gridDim.x * RED_NT)
        v += (double)a[r].x * (double)b[r].y;
    v = block_sum(v, sh);
    if (threadIdx.x == 0) part[(size_t)s * gridDim.x + blockIdx.x] = v;
}

}  // namespace scft
