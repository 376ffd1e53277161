// fft_engine.hpp -- device-side building blocks for the batched mixed-radix FFTs (gfx950).
//
// A "line set" is NLINES complex lines of length N held in LDS, element i of line l at
// buf[l*LS + i*ES]. One Stockham radix-R pass is done by every thread of the workgroup:
// work item w -> (line = w % NLINES, butterfly j = w / NLINES), so consecutive lanes touch
// consecutive lines. For column sets (LS = 1) that is consecutive LDS words; for row sets
// the row pitch is chosen odd so the lanes spread over all banks.
//
// Pass p (radix R, NS = product of the radices before it), butterfly j:
//   v[r] = x[j + r*N/R] * W_{NS*R}^{r*(j % NS)},  v = DFT_R(v),
//   y[(j/NS)*NS*R + j%NS + r*NS] = v[r].
// Twiddles come from an N-entry table tw[t] = exp(-2 pi i t / N) built in long double on
// the host and rounded once to T.
#pragma once
#include <hip/hip_runtime.h>

namespace scft {

template <typename T>
struct alignas(2 * sizeof(T)) cpx {
    T x, y;
};

template <typename T>
__host__ __device__ __forceinline__ cpx<T> cadd(cpx<T> a, cpx<T> b) { return {a.x + b.x, a.y + b.y}; }
template <typename T>
__host__ __device__ __forceinline__ cpx<T> csub(cpx<T> a, cpx<T> b) { return {a.x - b.x, a.y - b.y}; }
template <typename T>
__host__ __device__ __forceinline__ cpx<T> cmul(cpx<T> a, cpx<T> b)
{
    return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
// a * conj(b)
template <typename T>
__host__ __device__ __forceinline__ cpx<T> cmulc(cpx<T> a, cpx<T> b)
{
    return {a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y};
}
template <typename T>
__host__ __device__ __forceinline__ cpx<T> cscale(cpx<T> a, T s) { return {a.x * s, a.y * s}; }
// multiply by -i (forward) or +i (inverse)
template <bool INV, typename T>
__host__ __device__ __forceinline__ cpx<T> mul_mi(cpx<T> a)
{
    if (INV) return {-a.y, a.x};
    return {a.y, -a.x};
}

// ---------------------------------------------------------------- small DFTs in registers
template <int R, bool INV, typename T>
struct dft;

template <bool INV, typename T>
struct dft<2, INV, T> {
    static __host__ __device__ __forceinline__ void run(cpx<T>* v)
    {
        cpx<T> a = v[0], b = v[1];
        v[0] = cadd(a, b);
        v[1] = csub(a, b);
    }
};

template <bool INV, typename T>
struct dft<3, INV, T> {
    static __host__ __device__ __forceinline__ void run(cpx<T>* v)
    {
        const T c = T(0.86602540378443864676372317075294);  // sqrt(3)/2
        cpx<T> s = cadd(v[1], v[2]), d = csub(v[1], v[2]);
        cpx<T> a = {v[0].x - T(0.5) * s.x, v[0].y - T(0.5) * s.y};
        cpx<T> t = cscale(mul_mi<INV>(d), c);
        v[0] = cadd(v[0], s);
        v[1] = cadd(a, t);
        v[2] = csub(a, t);
    }
};

template <bool INV, typename T>
struct dft<4, INV, T> {
    static __host__ __device__ __forceinline__ void run(cpx<T>* v)
    {
        cpx<T> a = cadd(v[0], v[2]), b = csub(v[0], v[2]);
        cpx<T> c = cadd(v[1], v[3]), d = mul_mi<INV>(csub(v[1], v[3]));
        v[0] = cadd(a, c);
        v[2] = csub(a, c);
        v[1] = cadd(b, d);
        v[3] = csub(b, d);
    }
};

template <bool INV, typename T>
struct dft<5, INV, T> {
    static __host__ __device__ __forceinline__ void run(cpx<T>* v)
    {
        const T c1 = T(0.30901699437494742410229341718282);   // cos(2pi/5)
        const T c2 = T(-0.80901699437494742410229341718282);  // cos(4pi/5)
        const T s1 = T(0.95105651629515357211643933337938);   // sin(2pi/5)
        const T s2 = T(0.58778525229247312916870595463907);   // sin(4pi/5)
        cpx<T> s14 = cadd(v[1], v[4]), d14 = csub(v[1], v[4]);
        cpx<T> s23 = cadd(v[2], v[3]), d23 = csub(v[2], v[3]);
        cpx<T> a1 = {v[0].x + c1 * s14.x + c2 * s23.x, v[0].y + c1 * s14.y + c2 * s23.y};
        cpx<T> a2 = {v[0].x + c2 * s14.x + c1 * s23.x, v[0].y + c2 * s14.y + c1 * s23.y};
        cpx<T> b1 = {s1 * d14.x + s2 * d23.x, s1 * d14.y + s2 * d23.y};
        cpx<T> b2 = {s2 * d14.x - s1 * d23.x, s2 * d14.y - s1 * d23.y};
        b1 = mul_mi<INV>(b1);
        b2 = mul_mi<INV>(b2);
        v[0] = cadd(v[0], cadd(s14, s23));
        v[1] = cadd(a1, b1);
        v[4] = csub(a1, b1);
        v[2] = cadd(a2, b2);
        v[3] = csub(a2, b2);
    }
};

template <bool INV, typename T>
struct dft<8, INV, T> {
    static __host__ __device__ __forceinline__ void run(cpx<T>* v)
    {
        const T h = T(0.70710678118654752440084436210485);
        cpx<T> a[4], b[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            a[k] = cadd(v[k], v[k + 4]);
            b[k] = csub(v[k], v[k + 4]);
        }
        // b[k] *= w8^k, w8 = exp(-+ i pi/4)
        {
            cpx<T> t = b[1];
            // (1 - i) h for forward, (1 + i) h for inverse
            b[1] = INV ? cpx<T>{(t.x - t.y) * h, (t.x + t.y) * h} : cpx<T>{(t.x + t.y) * h, (t.y - t.x) * h};
            b[2] = mul_mi<INV>(b[2]);
            t = b[3];
            // (-1 - i) h forward, (-1 + i) h inverse
            b[3] = INV ? cpx<T>{(-t.x - t.y) * h, (t.x - t.y) * h} : cpx<T>{(t.y - t.x) * h, (-t.x - t.y) * h};
        }
        dft<4, INV, T>::run(a);
        dft<4, INV, T>::run(b);
#pragma unroll
        for (int m = 0; m < 4; ++m) {
            v[2 * m] = a[m];
            v[2 * m + 1] = b[m];
        }
    }
};

// ---------------------------------------------------------------- radix plans
// Radix chosen for the remaining length n: prefer 8, then 4, then 2, 3, 5, never leaving
// a trailing 2 when a 4 can absorb it.
__host__ __device__ constexpr int pick_radix(int n)
{
    return (n % 8 == 0 && n != 16) ? 8
         : (n % 4 == 0) ? 4
         : (n % 2 == 0) ? 2
         : (n % 3 == 0) ? 3
         : (n % 5 == 0) ? 5
         : 0;
}

__host__ __device__ constexpr bool fft_len_supported(int n)
{
    while (n > 1) {
        int r = pick_radix(n);
        if (r == 0) return false;
        n /= r;
    }
    return true;
}

// One Stockham pass over a line set in LDS, in place: all reads to registers, barrier,
// all writes, barrier.
template <typename T, int N, int R, int NS, bool INV, int NLINES, int LS, int ES, int NT>
__device__ __forceinline__ void stockham_pass(cpx<T>* buf, const cpx<T>* tw, int tid)
{
    constexpr int NB = N / R;
    constexpr int TOTAL = NLINES * NB;
    constexpr int IT = (TOTAL + NT - 1) / NT;
    constexpr int TWS = N / (NS * R);
    cpx<T> v[IT][R];
#pragma unroll
    for (int it = 0; it < IT; ++it) {
        const int w = tid + it * NT;
        if (TOTAL % NT == 0 || w < TOTAL) {
            const int line = w % NLINES, j = w / NLINES;
            const cpx<T>* base = buf + line * LS;
#pragma unroll
            for (int r = 0; r < R; ++r) v[it][r] = base[(j + r * NB) * ES];
            if (NS > 1) {
                const int k = j % NS;
#pragma unroll
                for (int r = 1; r < R; ++r) {
                    const cpx<T> t = tw[r * k * TWS];
                    v[it][r] = INV ? cmulc(v[it][r], t) : cmul(v[it][r], t);
                }
            }
            dft<R, INV, T>::run(v[it]);
        }
    }
    __syncthreads();
#pragma unroll
    for (int it = 0; it < IT; ++it) {
        const int w = tid + it * NT;
        if (TOTAL % NT == 0 || w < TOTAL) {
            const int line = w % NLINES, j = w / NLINES;
            const int k = j % NS;
            cpx<T>* base = buf + line * LS;
            const int idx = (j / NS) * NS * R + k;
#pragma unroll
            for (int r = 0; r < R; ++r) base[(idx + r * NS) * ES] = v[it][r];
        }
    }
    __syncthreads();
}

// All passes for length N: recursion over the radix plan.
template <typename T, int N, int NS, bool INV, int NLINES, int LS, int ES, int NT>
struct line_fft {
    static __device__ __forceinline__ void run(cpx<T>* buf, const cpx<T>* tw, int tid)
    {
        constexpr int R = pick_radix(N / NS);
        static_assert(R != 0, "FFT length must factor into 2, 3, 4, 5, 8");
        stockham_pass<T, N, R, NS, INV, NLINES, LS, ES, NT>(buf, tw, tid);
        line_fft<T, N, NS * R, INV, NLINES, LS, ES, NT>::run(buf, tw, tid);
    }
};

template <typename T, int N, bool INV, int NLINES, int LS, int ES, int NT>
struct line_fft<T, N, N, INV, NLINES, LS, ES, NT> {
    static __device__ __forceinline__ void run(cpx<T>*, const cpx<T>*, int) {}
};

template <typename T, int N, bool INV, int NLINES, int LS, int ES, int NT>
__device__ __forceinline__ void fft_lines(cpx<T>* buf, const cpx<T>* tw, int tid)
{
    line_fft<T, N, 1, INV, NLINES, LS, ES, NT>::run(buf, tw, tid);
}

}  // namespace scft
