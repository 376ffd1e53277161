// fft2.hpp -- register-resident DFTs and the two-stage ("four-step") line FFT.
//
// A length-N line with N = N1 * N2 is transformed as
//   stage 1: the owner of n1 holds x[n1 + N1 n2] (n2 < N2) and does an N2-point DFT in
//            registers, then multiplies by W_N^(n1 k2);
//   LDS transpose;
//   stage 2: the owner of k2 holds Y[n1][k2] (n1 < N1) and does an N1-point DFT in
//            registers, giving X[k2 + N2 k1] (k1 < N1).
// The inverse runs the same stages mirrored (conjugate roots), so it starts from the
// stage-2 ownership and ends in the stage-1 ownership. Hence inverse -> pointwise ->
// forward needs no data movement, and a forward -> diagonal multiply -> inverse needs no
// data movement either: only one LDS transpose per 1-D transform remains.
//
// Register DFTs use a compile-time Stockham over radices 2, 3, 4, 5, 8 with twiddles that
// are compile-time constants (constexpr Taylor series, rounded once to T).
#pragma once
#include <utility>

#include "fft_engine.hpp"

namespace scft {

// ---------------------------------------------------------------- constexpr trig
constexpr double kTwoPi = 6.283185307179586476925286766559;

constexpr double cx_reduce(double x)
{
    while (x > kTwoPi / 2) x -= kTwoPi;
    while (x < -kTwoPi / 2) x += kTwoPi;
    return x;
}
constexpr double cx_cos(double x)
{
    x = cx_reduce(x);
    double term = 1, sum = 1;
    for (int k = 1; k < 30; ++k) {
        term *= -x * x / ((2.0 * k - 1) * (2.0 * k));
        sum += term;
    }
    return sum;
}
constexpr double cx_sin(double x)
{
    x = cx_reduce(x);
    double term = x, sum = x;
    for (int k = 1; k < 30; ++k) {
        term *= -x * x / ((2.0 * k) * (2.0 * k + 1));
        sum += term;
    }
    return sum;
}

// W_L^t = exp(-2 pi i t / L) (forward) and its conjugate (inverse), times v
template <typename T, int L, int t, bool INV>
__host__ __device__ __forceinline__ cpx<T> mul_wconst(cpx<T> v)
{
    constexpr int tt = ((t % L) + L) % L;
    if constexpr (tt == 0) {
        return v;
    } else if constexpr (2 * tt == L) {
        return {-v.x, -v.y};
    } else if constexpr (4 * tt == L) {
        return mul_mi<INV>(v);  // W^(L/4) = -i (forward)
    } else if constexpr (4 * tt == 3 * L) {
        return mul_mi<!INV>(v);
    } else {
        constexpr T c = (T)cx_cos(kTwoPi * tt / L);
        constexpr T s = (T)(INV ? cx_sin(kTwoPi * tt / L) : -cx_sin(kTwoPi * tt / L));
        return {v.x * c - v.y * s, v.x * s + v.y * c};
    }
}

template <int... I, typename F>
__host__ __device__ __forceinline__ void sfor_impl(std::integer_sequence<int, I...>, F&& f)
{
    (f(std::integral_constant<int, I>{}), ...);
}
template <int N, typename F>
__host__ __device__ __forceinline__ void sfor(F&& f)
{
    sfor_impl(std::make_integer_sequence<int, N>{}, f);
}

// One compile-time Stockham pass over L values in registers.
template <typename T, int L, int R, int NS, bool INV>
__host__ __device__ __forceinline__ void rpass(const cpx<T>* in, cpx<T>* out)
{
    constexpr int NB = L / R;
    sfor<NB>([&](auto jc) {
        constexpr int j = decltype(jc)::value;
        constexpr int k = j % NS;
        cpx<T> v[R];
        sfor<R>([&](auto rc) {
            constexpr int r = decltype(rc)::value;
            v[r] = mul_wconst<T, NS * R, r * k, INV>(in[j + r * NB]);
        });
        dft<R, INV, T>::run(v);
        constexpr int idx = (j / NS) * NS * R + k;
        sfor<R>([&](auto rc) {
            constexpr int r = decltype(rc)::value;
            out[idx + r * NS] = v[r];
        });
    });
}

template <typename T, int L, int NS, bool INV>
struct rstock {
    static __host__ __device__ __forceinline__ void run(cpx<T>* v)
    {
        constexpr int R = pick_radix(L / NS);
        static_assert(R != 0, "register DFT length must factor into 2, 3, 4, 5, 8");
        cpx<T> t[L];
        rpass<T, L, R, NS, INV>(v, t);
        sfor<L>([&](auto ic) { v[decltype(ic)::value] = t[decltype(ic)::value]; });
        rstock<T, L, NS * R, INV>::run(v);
    }
};
template <typename T, int L, bool INV>
struct rstock<T, L, L, INV> {
    static __host__ __device__ __forceinline__ void run(cpx<T>*) {}
};

// In-register DFT of length L (natural order in and out, unnormalised).
template <typename T, int L, bool INV>
__host__ __device__ __forceinline__ void rdft(cpx<T>* v)
{
    if constexpr (L > 1) rstock<T, L, 1, INV>::run(v);
}

// Two-stage split N = N1 * N2 with N1 <= N2 <= 16 (N1 = 1 when N <= 16).
// N2 up to 25 (a 20- or 25-point register DFT) only when no split with N2 <= 16 exists, so
// that lengths such as 200 = 10 x 20 split too.
__host__ __device__ constexpr int split_n1(int N)
{
    if (N <= 16) return fft_len_supported(N) ? 1 : 0;
    for (int lim = 16; lim <= 25; lim += 9)
        for (int d = 16; d >= 2; --d)
            if (N % d == 0 && d * d <= N && N / d <= lim && fft_len_supported(d) && fft_len_supported(N / d))
                return d;
    return 0;
}

}  // namespace scft
