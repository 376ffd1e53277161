// prop_rlines.hpp -- the y and z passes of the split line step (prop_lines2.hpp) on the
// two-stage register FFT of the column pass (C2Geom, sc_x_sandwich) instead of LDS Stockham
// passes: one LDS exchange per transform instead of one per radix pass.
//
//   k_ry  (h, f):  h <- K1/2y h,  f <- Ky f           (or h alone)
//   k_rz  (h, f):  h <- K1/2z E1/2 K1/2z h,  f <- Kz f   (E1/2 of step s)
// and the x pass is the column pass k_scol2 of the two-pass step (prop_sandwich.hpp), which
// completes the same step from the same field state.
//
// The same one-axis sandwiches as k_ly and k_lz (density.h:129-220, GPU_kernels.h:207-213,
// the kinetic factor split by axis), on the same real-space natural-order fields, so either
// form of a pass can follow either form of the one before. In FP32 every line keeps its
// sum in FP64 through the sandwich (sc_x_sandwich).
#pragma once
#include "prop_pencil.hpp"
#include "prop_sandwich.hpp"

namespace scft {

// Axis lengths with register line passes (each must split into two factors of at most 16).
#define RL_LENGTHS(X) X(96) X(128) X(150) X(160) X(192) X(256)
// Lines per workgroup of the y pass (consecutive z: the width of each coalesced row read) and
// of the z pass (contiguous rows). Measured on the 200 x 150 x 150 box: 32 lines per
// workgroup left too few workgroups per CU (2 in FP64: 77 KB of LDS and 144 VGPRs each);
// FP64 gains from 8, FP32 does best at 16, except FP32 at length 150 (X1 = 10), where 16
// lines make 160 stage-1 items in a 192-thread workgroup and 24 make 240 in 256: 0.544
// against 0.574 ms per step at four replicas. FP64 at 150 on 12 lines (173 VGPRs, two waves
// per SIMD) was slower at one replica (0.256 against 0.243 ms).
template <typename T>
constexpr int rl_ty(int L)
{
    return sizeof(T) == 8 ? 8 : (L == 150 ? 24 : 16);
}
template <typename T>
constexpr int rl_tz(int L)
{
    return sizeof(T) == 8 ? 8 : (L == 150 ? 24 : 16);
}

// The x pass of the split step runs the column pass k_scol2 with the (y, z) extent given at
// run time (NY = NZ = 0), for these x lengths, on rl_tx<T>(L) columns per workgroup (256
// bytes of each x row: 128 was slower), except FP64 at 256, which needs 256 VGPRs. At x = 200
// (X1 = 20) the default tiles leave wave slots idle. FP64 takes 12 columns: its 218 VGPRs
// allow two waves per SIMD, so 16 columns (320 stage-1 items, five waves) run one workgroup
// per CU and 12 (240 items, four waves) run two; 0.243 against 0.261 ms per step on the
// 200 x 150 x 150 box (24 columns, one eight-wave workgroup, 0.241; both slower at x = 128).
// FP32 takes 24: at four waves per SIMD, 32 columns (640 items, ten waves) run one workgroup
// per CU and 24 (480 items, eight waves) run two; 0.168 against 0.180 ms per step at one
// replica, 0.576 against 0.606 at four (16 columns: 0.191, 0.615).
#define RX_LENGTHS(X) X(96) X(128) X(150) X(160) X(192) X(200) X(256)
template <typename T>
constexpr int rl_tx(int L)
{
    return sizeof(T) == 8 ? (L == 200 ? 12 : 16) : (L == 200 ? 24 : 32);
}
template <typename T>
constexpr bool rx_ok(int L)
{
    return !(sizeof(T) == 8 && L == 256);
}
// Split of the x pass: the larger factor first (200 = 20 x 10, 128 = 16 x 8), so that the
// stage-1 items, which hold q, f and the factor rows, take the shorter DFT and the longer
// one runs in stage 2 on one field at a time
constexpr int rl_x1(int L)
{
    return L / split_n1(L);
}

template <typename T, int L, int TC>
constexpr size_t rl_lds_bytes()
{
    return scol_lds_bytes<T, L, TC>();
}

// y-line pass on TC z-columns of one x. Field f < nh: Ch of slot f, K1/2y sandwich;
// f >= nh: Cf of slot f - nh, Ky sandwich. grid (m0 * ceil(m2 / TC), fields), block
// C2Geom<L, TC>::NT, LDS rl_lds_bytes<T, L, TC>(). Tiles run in XCD-contiguous order
// (pc_xcd_block): neighbouring z tiles share the cache lines at their boundary.
template <typename T, int L, int TC = rl_ty<T>(L)>
__global__ void __attribute__((amdgpu_flat_work_group_size(1, C2Geom<L, TC>::NT))) k_ry(PropArgs<T> a, int nh)
{
    using G = C2Geom<L, TC>;
    constexpr int NT = G::NT, X1 = G::X1, X2 = G::X2, I1 = G::I1;
    extern __shared__ unsigned char smem_raw[];
    cpx<T>* buf = reinterpret_cast<cpx<T>*>(smem_raw);
    cpx<T>* tw = buf + G::LDSN;
    T* g = reinterpret_cast<T*>(tw + L);
    double2* dsum = reinterpret_cast<double2*>(smem_raw + ((G::LDSN + L) * sizeof(cpx<T>) + L * sizeof(T) + 15) / 16 * 16);
    const int tid = threadIdx.x;
    const int m2 = a.m2;
    const int ntz = (m2 + TC - 1) / TC;
    const int bt = pc_xcd_block();
    const int x = bt / ntz, z0 = (bt - x * ntz) * TC;
    const bool full = (int)blockIdx.y >= nh;
    cpx<T>* fld = pc_field(a, blockIdx.y, nh) + (size_t)x * L * m2 + z0;
    for (int i = tid; i < L; i += NT) {
        tw[i] = a.tw1[i];
        g[i] = a.gy[i];
    }
    const InvLen<T> inv(L);
    auto item_ok = [&](int w) { return ((TC * X1) % NT == 0 || w < TC * X1) && z0 + w % TC < m2; };
    auto elem = [&](int w, int n2) { return (size_t)(w / TC + X1 * n2) * m2 + (size_t)(w % TC); };
    cpx<T> v[I1][X2];
#pragma unroll
    for (int it = 0; it < I1; ++it) {
        const int w = tid + it * NT;
        const bool ok = item_ok(w);
#pragma unroll
        for (int n2 = 0; n2 < X2; ++n2) v[it][n2] = ok ? fld[elem(w, n2)] : cpx<T>{T(0), T(0)};
    }
    __syncthreads();  // tables
    sc_x_sandwich<G, T, L, TC>(v, buf, tw, g, full, inv, dsum, tid);
#pragma unroll
    for (int it = 0; it < I1; ++it) {
        const int w = tid + it * NT;
        if (item_ok(w))
#pragma unroll
            for (int n2 = 0; n2 < X2; ++n2) fld[elem(w, n2)] = v[it][n2];
    }
}

// z-line pass of step s on TC contiguous z lines of slot blockIdx.y:
// h <- K1/2z E1/2 K1/2z h (Ch), f <- Kz f (Cf). Stage-1 items are row-major (consecutive
// lanes on consecutive z of one line). grid (ceil(m0 m1 / TC), nb slots), block
// C2Geom<L, TC>::NT, LDS rl_lds_bytes<T, L, TC>().
template <typename T, int L, int TC = rl_tz<T>(L)>
__global__ void __attribute__((amdgpu_flat_work_group_size(1, C2Geom<L, TC>::NT))) k_rz(PropArgs<T> a, int s)
{
    using G = C2Geom<L, TC>;
    constexpr int NT = G::NT, X1 = G::X1, X2 = G::X2, I1 = G::I1;
    extern __shared__ unsigned char smem_raw[];
    cpx<T>* buf = reinterpret_cast<cpx<T>*>(smem_raw);
    cpx<T>* tw = buf + G::LDSN;
    T* g = reinterpret_cast<T*>(tw + L);
    double2* dsum = reinterpret_cast<double2*>(smem_raw + ((G::LDSN + L) * sizeof(cpx<T>) + L * sizeof(T) + 15) / 16 * 16);
    const int tid = threadIdx.x, slot = blockIdx.y;
    const long long rows = (long long)a.m0 * a.m1, r0 = (long long)blockIdx.x * TC;
    const size_t fo = (size_t)slot * (size_t)a.M + (size_t)r0 * L;
    for (int i = tid; i < L; i += NT) {
        tw[i] = a.tw2[i];
        g[i] = a.gz[i];
    }
    const InvLen<T> inv(L);
    auto item_ok = [&](int w) { return ((TC * X1) % NT == 0 || w < TC * X1) && r0 + w / X1 < rows; };
    auto elem = [&](int w, int n2) { return (size_t)(w / X1) * L + (size_t)(w % X1 + X1 * n2); };
    auto load = [&](const cpx<T>* p, cpx<T> (&v)[I1][X2]) {
#pragma unroll
        for (int it = 0; it < I1; ++it) {
            const int w = tid + it * NT;
            const bool ok = item_ok(w);
#pragma unroll
            for (int n2 = 0; n2 < X2; ++n2) v[it][n2] = ok ? p[elem(w, n2)] : cpx<T>{T(0), T(0)};
        }
    };
    auto store = [&](cpx<T>* p, const cpx<T> (&v)[I1][X2]) {
#pragma unroll
        for (int it = 0; it < I1; ++it) {
            const int w = tid + it * NT;
            if (item_ok(w))
#pragma unroll
                for (int n2 = 0; n2 < X2; ++n2) p[elem(w, n2)] = v[it][n2];
        }
    };
    cpx<T> h[I1][X2];
    load(a.Ch + fo, h);
    __syncthreads();  // tables
    sc_x_sandwich<G, T, L, TC, true>(h, buf, tw, g, false, inv, dsum, tid);
    // E1/2 of step s: q1 (real part) takes species A while s < n1, q2 (imaginary part)
    // species B while s < n2 (density.h:136-146)
    const T* ex = s < a.n1 ? a.e4A : a.e4B;
    const T* ey = s < a.n2 ? a.e4B : a.e4A;
#pragma unroll
    for (int it = 0; it < I1; ++it) {
        const int w = tid + it * NT;
        if (item_ok(w))
#pragma unroll
            for (int n2 = 0; n2 < X2; ++n2) {
                const size_t e = fo + elem(w, n2);
                h[it][n2].x = Boltz<T>::mul2(h[it][n2].x, ex[e]);
                h[it][n2].y = Boltz<T>::mul2(h[it][n2].y, ey[e]);
            }
    }
    __syncthreads();  // buf is reused
    sc_x_sandwich<G, T, L, TC, true>(h, buf, tw, g, false, inv, dsum, tid);
    store(a.Ch + fo, h);
    cpx<T> f[I1][X2];
    load(a.Cf + fo, f);
    __syncthreads();
    sc_x_sandwich<G, T, L, TC, true>(f, buf, tw, g, true, inv, dsum, tid);
    store(a.Cf + fo, f);
}

}  // namespace scft
