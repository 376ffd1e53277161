// prop_types.hpp -- types shared by the propagator kernels and their host driver.
#pragma once
#include "fft_engine.hpp"

namespace scft {

constexpr int RED_NT = 256;  // threads of the reduction / pointwise kernels

template <typename T>
struct PropArgs {
    long long M;           // grid points
    int N, n1, n2, Nh;     // contour steps, block split, homopolymer length
    int nb;                // active batch slots
    cpx<T>* Ch;            // [slot][M] half-step sandwich work field
    cpx<T>* Cf;            // [slot][M] full-step sandwich work field
    const T* e4A;          // [slot][M] exp(-(wA - cA) ds / 4)
    const T* e4B;          // [slot][M] exp(-(wB - cB) ds / 4)
    cpx<T>* qs;            // [slot][N+1][M] stored (q1, q2)
    const cpx<T>* tw0;     // twiddle tables, lengths m0, m1, m2
    const cpx<T>* tw1;
    const cpx<T>* tw2;
    const T* gx;           // exp(-kx^2 ds/12) per axis (k from setup_k_sq, density.h:79-95)
    const T* gy;
    const T* gz;
    T invM;
};

// Per-slot scalars (density.h:376-399) in FP64 from the factored sums.
// scal[slot] = {lnQ, sC, sH, QC, Q2, Q1}; sC, sH multiply the Simpson sums directly.
struct ChainConst {
    int N, n1, n2, Nh, ens;
    double alpha, phic, ds;
};

}  // namespace scft
