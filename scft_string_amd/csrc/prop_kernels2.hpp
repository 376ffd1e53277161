// prop_kernels2.hpp -- propagator kernels built on the two-stage line FFT (fft2.hpp).
//
// Same contour-step schedule and data layout as prop_kernels.hpp (col(Ch,Cf) -> plane_mid ->
// col(Ch) -> plane_end, one complex field q + i q-dagger per slot), but every 1-D transform
// is two register DFTs around a single LDS transpose. The first and last stages read and
// write global memory directly, and inverse -> pointwise -> forward never leaves
// registers. Used for grids whose three axes all split as N1 x N2 with 1 < N1 <= N2 <= 16
// (24 ... 128); smaller axes use the LDS Stockham kernels.
#pragma once
#include <type_traits>

#include "fft2.hpp"
#include "prop_types.hpp"

namespace scft {

// Sum of v over the 64 lanes of a wave in FP64, the same value in every lane; a fixed
// butterfly order, so the result does not depend on timing.
__device__ __forceinline__ double wave_sum64(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// The L outputs of an inverse DFT sum to exactly L times its DC input. In FP32 the partial
// sums that carry the DC are rounded once and shared by all outputs, so the outputs' sum,
// and with it the mean of the inverse transform, is off by ~1 ulp of the DC per transform,
// coherently. Folding the FP64 residual of the sum (against L (tx, ty), the DC input or its
// FP64 value where the kernel has one) into u[0] leaves half an ulp of one output. Used on
// the items that carry the 3-D DC (the mean of q).
template <typename T, int L>
__device__ __forceinline__ void inv_sum_fix(cpx<T>* u, double tx, double ty)
{
    double sx = 0.0, sy = 0.0;
#pragma unroll
    for (int b = 0; b < L; ++b) {
        sx += (double)u[b].x;
        sy += (double)u[b].y;
    }
    u[0].x = (T)((double)u[0].x + ((double)L * tx - sx));
    u[0].y = (T)((double)u[0].y + ((double)L * ty - sy));
}

constexpr int cdiv(int a, int b) { return (a + b - 1) / b; }
constexpr int cmax(int a, int b) { return a > b ? a : b; }

// z split of a plane: from length 32 up, one that allows wave-local z exchanges (N1 divides
// N2 and 64) when the length has one, else split_n1. This changes 48 from 6 x 8 to 4 x 12:
// measured per contour step, 48^3 FP64 x 16 0.0925 -> 0.0802 ms, FP32 x 1 0.0294 -> 0.0281,
// FP32 x 16 0.0544 -> 0.0553. At 24 (2 x 12 against 4 x 6) it was 14-18% slower.
constexpr int split_wlz(int N)
{
    if (N < 32) return split_n1(N);
    for (int d = 16; d >= 2; --d)
        if (N % d == 0 && d * d <= N && N / d <= 16 && 64 % d == 0 && (N / d) % d == 0 && fft_len_supported(d) &&
            fft_len_supported(N / d))
            return d;
    return split_n1(N);
}

// Plane geometry: rows y (NY), columns z (NZ). Work items per stage:
//   Z1 (y, a1)  holds x[y][a1 + A1 a2], a2 < A2      (real-space ownership)
//   Z2 (y, k2)  holds Z[y][k2 + A2 k1], k1 < A1      (after the z transform)
//   Y1 (kz, b1) holds P[b1 + B1 b2][kz], b2 < B2
//   Y2 (kz, l2) holds Y[l2 + B2 l1][kz], l1 < B1     (k-space ownership, global layout)
// NTO > 0 sets the workgroup size (a multiple of 64); 0 takes one thread per Z1 item.
template <int NY, int NZ, int NTO = 0>
struct P2Geom {
    static constexpr int A1 = split_wlz(NZ), A2 = NZ / A1;
    static constexpr int B1 = split_n1(NY), B2 = NY / B1;
    static constexpr int NT0 = cmax(64, cdiv(NY * A1, 64) * 64) > 1024 ? 1024 : cmax(64, cdiv(NY * A1, 64) * 64);
    static constexpr int NT = NTO > 0 ? NTO : NT0;
    static_assert(NT % 64 == 0 && NT <= 1024, "plane workgroup size");
    static constexpr int IZ1 = cdiv(NY * A1, NT), IZ2 = cdiv(NY * A2, NT);
    static constexpr int IY1 = cdiv(NZ * B1, NT), IY2 = cdiv(NZ * B2, NT);
    // Wave-local z transposes: when A1 divides A2 and 64, each stage-1 item (y, a1) also
    // takes the stage-2 items (y, a1 + A1 j), j < A2 / A1 (z2_item). The lanes owning row y
    // are then the same A1 consecutive lanes of one wave in both stages, and the TZ and PL row
    // regions coincide, so the z-stage exchanges need no workgroup barrier.
    static constexpr bool WLZ = A2 % A1 == 0 && 64 % A1 == 0;
    static constexpr int IZ2E = WLZ ? IZ1 * (A2 / A1) : IZ2;  // stage-2 items per thread
    // LDS layouts in complex elements
    static constexpr int PZ = A2 + 1;                                        // TZ: y*RZ + a1*PZ + k2
    static constexpr int RZ = WLZ ? cmax(A1 * PZ, NZ + 8) : A1 * PZ + 4;
    static constexpr int PP = WLZ ? RZ : NZ + 8;                             // PL: y*PP + kz
    static constexpr int PY = B2, CY = B1 * PY + 1;                          // TY: kz*CY + b1*PY + l2
    static constexpr int LDSN = cmax(cmax(NY * RZ, NY * PP), NZ * CY);
    static constexpr bool ok = A1 > 1 && B1 > 1;
};

template <typename T, int NY, int NZ, int NTO = 0>
struct P2 : P2Geom<NY, NZ, NTO> {
    using G = P2Geom<NY, NZ, NTO>;
    using G::A1; using G::A2; using G::B1; using G::B2; using G::NT;
    using G::IZ1; using G::IY1; using G::IY2;
    using G::PZ; using G::RZ; using G::PP; using G::PY; using G::CY; using G::WLZ; using G::IZ2E;

    // z stage-2 item `it` of thread tid: its row y and k2, false when it has none
    static __device__ __forceinline__ bool z2_item(int tid, int it, int& y, int& k2)
    {
        if constexpr (WLZ) {
            constexpr int R = A2 / A1;
            const int w = tid + (it / R) * NT;
            y = w / A1;
            k2 = w % A1 + A1 * (it % R);
            return (NY * A1) % NT == 0 || w < NY * A1;
        } else {
            const int w = tid + it * NT;
            y = w / A2;
            k2 = w % A2;
            return (NY * A2) % NT == 0 || w < NY * A2;
        }
    }

    // exchange point inside the z stages: a wave-level ordering point when the exchange
    // stays inside one wave, else a workgroup barrier
    static __device__ __forceinline__ void zsync()
    {
        if constexpr (WLZ) {
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        } else {
            __syncthreads();
        }
    }

    // forward 2-D transform of the Z1-layout registers zr; result to global `out` (row-major
    // [y][z] of the plane). buf must be free on entry; it is in use on return.
    static __device__ __forceinline__ void fwd(cpx<T> (&zr)[IZ1][A2], cpx<T>* buf, const cpx<T>* twz,
                                               const cpx<T>* twy, cpx<T>* __restrict__ out, int tid)
    {
        // z stage 1
#pragma unroll
        for (int it = 0; it < IZ1; ++it) {
            const int w = tid + it * NT;
            if ((NY * A1) % NT == 0 || w < NY * A1) {
                const int y = w / A1, a1 = w % A1;
                rdft<T, A2, false>(zr[it]);
                sfor<A2>([&](auto kc) {
                    constexpr int k2 = decltype(kc)::value;
                    const cpx<T> v = k2 == 0 ? zr[it][0] : cmul(zr[it][k2], twz[(a1 * k2) % NZ]);
                    buf[y * RZ + a1 * PZ + k2] = v;
                });
            }
        }
        zsync();
        // z stage 2
        cpx<T> zz[IZ2E][A1];
#pragma unroll
        for (int it = 0; it < IZ2E; ++it) {
            int y, k2;
            if (z2_item(tid, it, y, k2)) {
                sfor<A1>([&](auto ac) {
                    constexpr int a1 = decltype(ac)::value;
                    zz[it][a1] = buf[y * RZ + a1 * PZ + k2];
                });
                rdft<T, A1, false>(zz[it]);
            }
        }
        zsync();
#pragma unroll
        for (int it = 0; it < IZ2E; ++it) {
            int y, k2;
            if (z2_item(tid, it, y, k2)) {
                sfor<A1>([&](auto kc) {
                    constexpr int k1 = decltype(kc)::value;
                    buf[y * PP + k2 + A2 * k1] = zz[it][k1];
                });
            }
        }
        __syncthreads();
        // y stage 1
        cpx<T> yy[IY1][B2];
#pragma unroll
        for (int it = 0; it < IY1; ++it) {
            const int w = tid + it * NT;
            if ((NZ * B1) % NT == 0 || w < NZ * B1) {
                const int kz = w % NZ, b1 = w / NZ;
                sfor<B2>([&](auto bc) {
                    constexpr int b2 = decltype(bc)::value;
                    yy[it][b2] = buf[(b1 + B1 * b2) * PP + kz];
                });
                rdft<T, B2, false>(yy[it]);
                sfor<B2>([&](auto lc) {
                    constexpr int l2 = decltype(lc)::value;
                    if (l2 != 0) yy[it][l2] = cmul(yy[it][l2], twy[(b1 * l2) % NY]);
                });
            }
        }
        __syncthreads();
#pragma unroll
        for (int it = 0; it < IY1; ++it) {
            const int w = tid + it * NT;
            if ((NZ * B1) % NT == 0 || w < NZ * B1) {
                const int kz = w % NZ, b1 = w / NZ;
                sfor<B2>([&](auto lc) {
                    constexpr int l2 = decltype(lc)::value;
                    buf[kz * CY + b1 * PY + l2] = yy[it][l2];
                });
            }
        }
        __syncthreads();
        // y stage 2 -> global
#pragma unroll
        for (int it = 0; it < IY2; ++it) {
            const int w = tid + it * NT;
            if ((NZ * B2) % NT == 0 || w < NZ * B2) {
                const int kz = w % NZ, l2 = w / NZ;
                cpx<T> u[B1];
                sfor<B1>([&](auto bc) {
                    constexpr int b1 = decltype(bc)::value;
                    u[b1] = buf[kz * CY + b1 * PY + l2];
                });
                rdft<T, B1, false>(u);
                sfor<B1>([&](auto lc) {
                    constexpr int l1 = decltype(lc)::value;
                    out[(l2 + B2 * l1) * NZ + kz] = u[l1];
                });
            }
        }
    }

    // forward 2-D transform of the Z1-layout registers zr, the spectrum left in registers in
    // Y2 ownership (u[it][l1] = Y[l2 + B2 l1][kz] of item it = (kz, l2)), the layout inv_regs
    // starts from. The same stages as fwd (kept as a separate copy: sharing the stages through
    // a common helper changed the register allocation of the FP64 64^2 plane_end, which then
    // spilled). buf must be free on entry; it is in use on return.
    static __device__ __forceinline__ void fwd_regs(cpx<T> (&zr)[IZ1][A2], cpx<T>* buf, const cpx<T>* twz,
                                                    const cpx<T>* twy, cpx<T> (&uo)[IY2][B1], int tid)
    {
        // z stage 1
#pragma unroll
        for (int it = 0; it < IZ1; ++it) {
            const int w = tid + it * NT;
            if ((NY * A1) % NT == 0 || w < NY * A1) {
                const int y = w / A1, a1 = w % A1;
                rdft<T, A2, false>(zr[it]);
                sfor<A2>([&](auto kc) {
                    constexpr int k2 = decltype(kc)::value;
                    const cpx<T> v = k2 == 0 ? zr[it][0] : cmul(zr[it][k2], twz[(a1 * k2) % NZ]);
                    buf[y * RZ + a1 * PZ + k2] = v;
                });
            }
        }
        zsync();
        // z stage 2
        cpx<T> zz[IZ2E][A1];
#pragma unroll
        for (int it = 0; it < IZ2E; ++it) {
            int y, k2;
            if (z2_item(tid, it, y, k2)) {
                sfor<A1>([&](auto ac) {
                    constexpr int a1 = decltype(ac)::value;
                    zz[it][a1] = buf[y * RZ + a1 * PZ + k2];
                });
                rdft<T, A1, false>(zz[it]);
            }
        }
        zsync();
#pragma unroll
        for (int it = 0; it < IZ2E; ++it) {
            int y, k2;
            if (z2_item(tid, it, y, k2)) {
                sfor<A1>([&](auto kc) {
                    constexpr int k1 = decltype(kc)::value;
                    buf[y * PP + k2 + A2 * k1] = zz[it][k1];
                });
            }
        }
        __syncthreads();
        // y stage 1
        cpx<T> yy[IY1][B2];
#pragma unroll
        for (int it = 0; it < IY1; ++it) {
            const int w = tid + it * NT;
            if ((NZ * B1) % NT == 0 || w < NZ * B1) {
                const int kz = w % NZ, b1 = w / NZ;
                sfor<B2>([&](auto bc) {
                    constexpr int b2 = decltype(bc)::value;
                    yy[it][b2] = buf[(b1 + B1 * b2) * PP + kz];
                });
                rdft<T, B2, false>(yy[it]);
                sfor<B2>([&](auto lc) {
                    constexpr int l2 = decltype(lc)::value;
                    if (l2 != 0) yy[it][l2] = cmul(yy[it][l2], twy[(b1 * l2) % NY]);
                });
            }
        }
        __syncthreads();
#pragma unroll
        for (int it = 0; it < IY1; ++it) {
            const int w = tid + it * NT;
            if ((NZ * B1) % NT == 0 || w < NZ * B1) {
                const int kz = w % NZ, b1 = w / NZ;
                sfor<B2>([&](auto lc) {
                    constexpr int l2 = decltype(lc)::value;
                    buf[kz * CY + b1 * PY + l2] = yy[it][l2];
                });
            }
        }
        __syncthreads();
        // y stage 2 -> registers
#pragma unroll
        for (int it = 0; it < IY2; ++it) {
            const int w = tid + it * NT;
            if ((NZ * B2) % NT == 0 || w < NZ * B2) {
                const int kz = w % NZ, l2 = w / NZ;
                sfor<B1>([&](auto bc) {
                    constexpr int b1 = decltype(bc)::value;
                    uo[it][b1] = buf[kz * CY + b1 * PY + l2];
                });
                rdft<T, B1, false>(uo[it]);
            }
        }
    }

    // k-space input of the inverse, in Y2 ownership, read straight from global memory
    static __device__ __forceinline__ void inv_load(const cpx<T>* __restrict__ in, cpx<T> (&u)[IY2][B1], int tid)
    {
#pragma unroll
        for (int it = 0; it < IY2; ++it) {
            const int w = tid + it * NT;
            if ((NZ * B2) % NT == 0 || w < NZ * B2) {
                const int kz = w % NZ, l2 = w / NZ;
                sfor<B1>([&](auto lc) {
                    constexpr int l1 = decltype(lc)::value;
                    u[it][l1] = in[(l2 + B2 * l1) * NZ + kz];
                });
            }
        }
    }

    // inverse 2-D transform from global `in` to Z1-layout registers zr.
    // buf must be free on entry; it is in use on return.
    static __device__ __forceinline__ void inv(const cpx<T>* __restrict__ in, cpx<T>* buf, const cpx<T>* twz,
                                               const cpx<T>* twy, cpx<T> (&zr)[IZ1][A2], int tid)
    {
        cpx<T> u[IY2][B1];
        inv_load(in, u, tid);
        inv_regs(u, buf, twz, twy, zr, tid);
    }

    // inverse 2-D transform of Y2-ownership registers u (from inv_load) to Z1-layout zr
    static __device__ __forceinline__ void inv_regs(cpx<T> (&uin)[IY2][B1], cpx<T>* buf, const cpx<T>* twz,
                                                    const cpx<T>* twy, cpx<T> (&zr)[IZ1][A2], int tid)
    {
        inv_y2(uin, buf, twy, tid);
        inv_tail(buf, twz, twy, zr, tid);
    }

    // first stage of the inverse (y stage 2'): consumes the registers u, result in buf
    static __device__ __forceinline__ void inv_y2(cpx<T> (&uin)[IY2][B1], cpx<T>* buf, const cpx<T>* twy, int tid)
    {
#pragma unroll
        for (int it = 0; it < IY2; ++it) {
            const int w = tid + it * NT;
            if ((NZ * B2) % NT == 0 || w < NZ * B2) {
                const int kz = w % NZ, l2 = w / NZ;
                cpx<T>* u = uin[it];
                rdft<T, B1, true>(u);
                sfor<B1>([&](auto bc) {
                    constexpr int b1 = decltype(bc)::value;
                    const cpx<T> v = (b1 == 0 || l2 == 0) ? u[b1] : cmulc(u[b1], twy[(b1 * l2) % NY]);
                    buf[kz * CY + b1 * PY + l2] = v;
                });
            }
        }
    }

    // remaining stages of the inverse, from buf (after inv_y2) to Z1-layout zr
    static __device__ __forceinline__ void inv_tail(cpx<T>* buf, const cpx<T>* twz, const cpx<T>* twy,
                                                    cpx<T> (&zr)[IZ1][A2], int tid)
    {
        __syncthreads();
        // y stage 1'
        cpx<T> yy[IY1][B2];
#pragma unroll
        for (int it = 0; it < IY1; ++it) {
            const int w = tid + it * NT;
            if ((NZ * B1) % NT == 0 || w < NZ * B1) {
                const int kz = w % NZ, b1 = w / NZ;
                sfor<B2>([&](auto lc) {
                    constexpr int l2 = decltype(lc)::value;
                    yy[it][l2] = buf[kz * CY + b1 * PY + l2];
                });
                rdft<T, B2, true>(yy[it]);
            }
        }
        __syncthreads();
#pragma unroll
        for (int it = 0; it < IY1; ++it) {
            const int w = tid + it * NT;
            if ((NZ * B1) % NT == 0 || w < NZ * B1) {
                const int kz = w % NZ, b1 = w / NZ;
                sfor<B2>([&](auto bc) {
                    constexpr int b2 = decltype(bc)::value;
                    buf[(b1 + B1 * b2) * PP + kz] = yy[it][b2];
                });
            }
        }
        __syncthreads();
        // z stage 2'
        cpx<T> zz[IZ2E][A1];
#pragma unroll
        for (int it = 0; it < IZ2E; ++it) {
            int y, k2;
            if (z2_item(tid, it, y, k2)) {
                sfor<A1>([&](auto kc) {
                    constexpr int k1 = decltype(kc)::value;
                    zz[it][k1] = buf[y * PP + k2 + A2 * k1];
                });
                rdft<T, A1, true>(zz[it]);
                sfor<A1>([&](auto ac) {
                    constexpr int a1 = decltype(ac)::value;
                    if (a1 != 0) zz[it][a1] = cmulc(zz[it][a1], twz[(a1 * k2) % NZ]);
                });
            }
        }
        zsync();
#pragma unroll
        for (int it = 0; it < IZ2E; ++it) {
            int y, k2;
            if (z2_item(tid, it, y, k2)) {
                sfor<A1>([&](auto ac) {
                    constexpr int a1 = decltype(ac)::value;
                    buf[y * RZ + a1 * PZ + k2] = zz[it][a1];
                });
            }
        }
        zsync();
        // z stage 1'
#pragma unroll
        for (int it = 0; it < IZ1; ++it) {
            const int w = tid + it * NT;
            if ((NY * A1) % NT == 0 || w < NY * A1) {
                const int y = w / A1, a1 = w % A1;
                sfor<A2>([&](auto kc) {
                    constexpr int k2 = decltype(kc)::value;
                    zr[it][k2] = buf[y * RZ + a1 * PZ + k2];
                });
                rdft<T, A2, true>(zr[it]);
            }
        }
    }

    // plane element index of Z1 item w, value a2
    static __device__ __forceinline__ int z1_elem(int w, int a2) { return (w / A1) * NZ + (w % A1) + A1 * a2; }
    static __device__ __forceinline__ bool z1_valid(int w) { return (NY * A1) % NT == 0 || w < NY * A1; }
};

template <typename T, int NY, int NZ>
__device__ __forceinline__ void p2_load_tw(cpx<T>* twz, cpx<T>* twy, const PropArgs<T>& a, int tid, int nt)
{
    for (int i = tid; i < NZ; i += nt) twz[i] = a.tw2[i];
    for (int i = tid; i < NY; i += nt) twy[i] = a.tw1[i];
}

// Minimum waves per SIMD k_plane_end2 is compiled for (1: the compiler's choice). The FP64
// 64^2 plane otherwise takes 132 VGPRs, one 512-thread workgroup per CU; capped at 4 waves it
// takes 126 with no spill and two workgroups fit (64^3 FP64 x 14: 0.1456 -> 0.1367 ms per
// contour step, bit-identical results).
template <typename T, int NE, int NT>
constexpr int end_wpe() { return (NE == 4096 && NT == 512 && sizeof(T) == 8) ? 4 : 1; }

// plane_end on the two-stage FFT (see k_plane_end in prop_kernels.hpp for the contract)
template <typename T, int NY, int NZ, int NTO = 0>
__global__ void __attribute__((amdgpu_flat_work_group_size(1, P2Geom<NY, NZ, NTO>::NT),
                               amdgpu_waves_per_eu(end_wpe<T, NY * NZ, P2Geom<NY, NZ, NTO>::NT>())))
k_plane_end2(PropArgs<T> a, int s, int mode)
{
    using P = P2<T, NY, NZ, NTO>;
    constexpr int NT = P::NT, IZ1 = P::IZ1, A2 = P::A2, NE = NY * NZ;
    extern __shared__ unsigned char smem_raw[];
    cpx<T>* buf = reinterpret_cast<cpx<T>*>(smem_raw);
    cpx<T>* twz = buf + P::LDSN;
    cpx<T>* twy = twz + NZ;
    const int tid = threadIdx.x;
    const int x = blockIdx.x, slot = blockIdx.y;
    const size_t fo = (size_t)slot * (size_t)a.M + (size_t)x * NE;
    p2_load_tw<T, NY, NZ>(twz, twy, a, tid, NT);

    // Below 64^2 every global load of the plane is issued up front (both k-space inputs and
    // the factors). From 64^2 up the kernel loads Cf after the Ch inverse and the factors
    // after both inverses: larger planes would spill otherwise, and at 64^2 the fewer live
    // registers (72 against 89 VGPRs in FP32) fit a third workgroup per CU. Measured per
    // contour step at 14 x 64^3: FP64 0.1551 -> 0.1456 ms, FP32 0.0768 -> 0.0760. At 32^2
    // late loads cost 12%.
    constexpr bool EARLY = NE < 4096;
    cpx<T> uh[P::IY2][P::B1], uf[P::IY2][P::B1];
    if (mode != PL_BEGIN && mode != PL_RESUME) {
        P::inv_load(a.Ch + fo, uh, tid);
        if constexpr (EARLY) P::inv_load(a.Cf + fo, uf, tid);
    }
    T ea[IZ1][A2], eb[IZ1][A2];
    auto load_factors = [&]() {
#pragma unroll
        for (int it = 0; it < IZ1; ++it) {
            const int w = tid + it * NT;
            if (P::z1_valid(w)) {
#pragma unroll
                for (int a2 = 0; a2 < A2; ++a2) {
                    const int e = P::z1_elem(w, a2);
                    ea[it][a2] = a.e4A[fo + e];
                    eb[it][a2] = a.e4B[fo + e];
                }
            }
        }
    };
    if constexpr (EARLY) load_factors();
    cpx<T> q[IZ1][A2];
    int sout;
    __syncthreads();  // twiddles
    if (mode == PL_BEGIN) {
        if constexpr (!EARLY) load_factors();
#pragma unroll
        for (int it = 0; it < IZ1; ++it)
#pragma unroll
            for (int a2 = 0; a2 < A2; ++a2) q[it][a2] = {T(1), T(1)};
        sout = 0;
    } else if (mode == PL_RESUME) {
        if constexpr (!EARLY) load_factors();
        const cpx<T>* qi = a.qs + ((size_t)slot * (size_t)a.qpitch + (size_t)a.qsrc) * (size_t)a.M + (size_t)x * NE;
#pragma unroll
        for (int it = 0; it < IZ1; ++it) {
            const int w = tid + it * NT;
#pragma unroll
            for (int a2 = 0; a2 < A2; ++a2) q[it][a2] = P::z1_valid(w) ? qi[P::z1_elem(w, a2)] : cpx<T>{T(0), T(0)};
        }
        sout = s;
    } else {
        cpx<T> h1[IZ1][A2];
        P::inv_regs(uh, buf, twz, twy, q, tid);  // q holds qs0 for now
        __syncthreads();
        if constexpr (!EARLY) P::inv_load(a.Cf + fo, uf, tid);
        P::inv_regs(uf, buf, twz, twy, h1, tid);
        if constexpr (!EARLY) load_factors();
        const bool qa = s < a.n1, pb = s < a.n2;  // species of this step (density.h:136-146)
#pragma unroll
        for (int it = 0; it < IZ1; ++it)
#pragma unroll
            for (int a2 = 0; a2 < A2; ++a2) {
                const T ex = qa ? ea[it][a2] : eb[it][a2];
                const T ey = pb ? eb[it][a2] : ea[it][a2];
                q[it][a2].x = Boltz<T>::rich(q[it][a2].x, h1[it][a2].x, ex);
                q[it][a2].y = Boltz<T>::rich(q[it][a2].y, h1[it][a2].y, ey);
            }
        sout = s + 1;
    }
    const int qix = q_index(a, sout);
    if (qix >= 0) {
        cpx<T>* qo = a.qs + ((size_t)slot * (size_t)a.qpitch + (size_t)qix) * (size_t)a.M + (size_t)x * NE;
#pragma unroll
        for (int it = 0; it < IZ1; ++it) {
            const int w = tid + it * NT;
            if (P::z1_valid(w))
#pragma unroll
                for (int a2 = 0; a2 < A2; ++a2) store_q(&qo[P::z1_elem(w, a2)], q[it][a2]);
        }
    }
    if (mode == PL_LAST) return;

    const bool qa = sout < a.n1, pb = sout < a.n2;
    cpx<T> v[IZ1][A2];
#pragma unroll
    for (int it = 0; it < IZ1; ++it)
#pragma unroll
        for (int a2 = 0; a2 < A2; ++a2) {
            const T ex = qa ? ea[it][a2] : eb[it][a2];
            const T ey = pb ? eb[it][a2] : ea[it][a2];
            v[it][a2] = {Boltz<T>::mul4(q[it][a2].x, ex), Boltz<T>::mul4(q[it][a2].y, ey)};
            q[it][a2] = {Boltz<T>::mul2(q[it][a2].x, ex), Boltz<T>::mul2(q[it][a2].y, ey)};
        }
    __syncthreads();
    P::fwd(v, buf, twz, twy, a.Ch + fo, tid);
    __syncthreads();
    P::fwd(q, buf, twz, twy, a.Cf + fo, tid);
}

// First pass of one box-gradient term (grad_kernels.hpp) on the two-stage FFT: b = part pbk
// of q-store entry ib and a = part pak of entry ia (0 real, 1 imaginary), both taking the
// step operator of species specA (0 A, 1 B), packed as Z = b + i a;
// Ch <- F2(E1/4 Z), Cf <- F2(E1/2 Z) for slot y. With specB >= 0 the workgroup also forms
// the mirror term, which reads the same two entries with their roles swapped (b = part pbk
// of entry ia, a = part pak of entry ib, species specB), into slot y + nb of Ch and Cf; its
// second read of the entries follows the first in the same workgroup, so it is served from
// cache. The factors are those of slot y for both terms.
template <typename T, int NY, int NZ>
__global__ void __attribute__((amdgpu_flat_work_group_size(1, P2Geom<NY, NZ>::NT)))
k_gplane_init2(PropArgs<T> a, int ib, int pbk, int ia, int pak, int specA, int specB)
{
    using P = P2<T, NY, NZ>;
    constexpr int NT = P::NT, IZ1 = P::IZ1, A2 = P::A2, NE = NY * NZ;
    extern __shared__ unsigned char smem_raw[];
    cpx<T>* buf = reinterpret_cast<cpx<T>*>(smem_raw);
    cpx<T>* twz = buf + P::LDSN;
    cpx<T>* twy = twz + NZ;
    const int tid = threadIdx.x;
    const int x = blockIdx.x, slot = blockIdx.y;
    const size_t fo = (size_t)slot * (size_t)a.M + (size_t)x * NE;
    p2_load_tw<T, NY, NZ>(twz, twy, a, tid, NT);
    const cpx<T>* q0 = a.qs + ((size_t)slot * (size_t)a.qpitch + (size_t)ib) * (size_t)a.M + (size_t)x * NE;
    const cpx<T>* q1 = a.qs + ((size_t)slot * (size_t)a.qpitch + (size_t)ia) * (size_t)a.M + (size_t)x * NE;
    for (int term = 0; term < (specB >= 0 ? 2 : 1); ++term) {
        // term 0: b from entry ib, a from entry ia; term 1 the other way round
        const cpx<T>* qb = term ? q1 : q0;
        const cpx<T>* qa = term ? q0 : q1;
        const T* ef = (term ? specB : specA) == 0 ? a.e4A : a.e4B;
        const size_t out = fo + (term ? (size_t)a.nb * (size_t)a.M : 0);
        cpx<T> v[IZ1][A2], h[IZ1][A2];
#pragma unroll
        for (int it = 0; it < IZ1; ++it) {
            const int w = tid + it * NT;
#pragma unroll
            for (int a2 = 0; a2 < A2; ++a2) {
                if (P::z1_valid(w)) {
                    const int e = P::z1_elem(w, a2);
                    const cpx<T> vb = qb[e], va = qa[e];
                    const T zb = pbk ? vb.y : vb.x, za = pak ? va.y : va.x, ev = ef[fo + e];
                    v[it][a2] = {Boltz<T>::mul4(zb, ev), Boltz<T>::mul4(za, ev)};
                    h[it][a2] = {Boltz<T>::mul2(zb, ev), Boltz<T>::mul2(za, ev)};
                } else {
                    v[it][a2] = h[it][a2] = {T(0), T(0)};
                }
            }
        }
        __syncthreads();  // twiddles; the previous term's last reads of buf
        P::fwd(v, buf, twz, twy, a.Ch + out, tid);
        __syncthreads();
        P::fwd(h, buf, twz, twy, a.Cf + out, tid);
    }
}

// plane_mid on the two-stage FFT: Ch <- F2( E1/2 * F2^-1(Ch) )
template <typename T, int NY, int NZ, int NTO = 0>
__global__ void __attribute__((amdgpu_flat_work_group_size(1, P2Geom<NY, NZ, NTO>::NT)))
k_plane_mid2(PropArgs<T> a, int s)
{
    using P = P2<T, NY, NZ, NTO>;
    constexpr int NT = P::NT, IZ1 = P::IZ1, A2 = P::A2, NE = NY * NZ;
    extern __shared__ unsigned char smem_raw[];
    cpx<T>* buf = reinterpret_cast<cpx<T>*>(smem_raw);
    cpx<T>* twz = buf + P::LDSN;
    cpx<T>* twy = twz + NZ;
    const int tid = threadIdx.x;
    const int x = blockIdx.x, slot = blockIdx.y;
    const size_t fo = (size_t)slot * (size_t)a.M + (size_t)x * NE;
    p2_load_tw<T, NY, NZ>(twz, twy, a, tid, NT);
    const bool qa = s < a.n1, pb = s < a.n2;
    T ex[IZ1][A2], ey[IZ1][A2];
    auto load_factors = [&]() {
#pragma unroll
        for (int it = 0; it < IZ1; ++it) {
            const int w = tid + it * NT;
            if (P::z1_valid(w)) {
#pragma unroll
                for (int a2 = 0; a2 < A2; ++a2) {
                    const int e = P::z1_elem(w, a2);
                    const T va = a.e4A[fo + e], vb = a.e4B[fo + e];
                    ex[it][a2] = qa ? va : vb;
                    ey[it][a2] = pb ? vb : va;
                }
            }
        }
    };
    // factors with the k-space input below 64^2, after the inverse from 64^2 up (as in
    // k_plane_end2: 71 against 91 VGPRs at 64^2 FP32)
    constexpr bool EARLY = NE < 4096;
    if constexpr (EARLY) load_factors();
    __syncthreads();  // twiddles
    cpx<T> z[IZ1][A2];
    P::inv(a.Ch + fo, buf, twz, twy, z, tid);
    if constexpr (!EARLY) load_factors();
#pragma unroll
    for (int it = 0; it < IZ1; ++it)
#pragma unroll
        for (int a2 = 0; a2 < A2; ++a2) {
            z[it][a2].x = Boltz<T>::mul2(z[it][a2].x, ex[it][a2]);
            z[it][a2].y = Boltz<T>::mul2(z[it][a2].y, ey[it][a2]);
        }
    __syncthreads();
    P::fwd(z, buf, twz, twy, a.Ch + fo, tid);
}

// Persistent form of plane_mid for planes that take one workgroup per CU (their LDS). A
// workgroup with its CU to itself cannot overlap one plane's global loads with another
// plane's transforms, so each workgroup here walks the planes w = blockIdx.x,
// blockIdx.x + gridDim.x, ... of the batch (plane w of slot w / m0 at offset w NE) and loads
// the next plane's k-space input during the current plane's forward transform. Barriers wait
// for LDS only, so those loads stay in flight through it. Requires gridDim.x <= nplanes =
// m0 * nb. Measured on MI355X: 96^3 FP32 x 8, 0.1697 -> 0.1632 ms per contour step. The same
// for plane_end, and for 128^2 planes (1024 threads, at most 128 VGPRs), spills registers and
// was slower (128^3 FP32 x 4: 0.21 -> 0.26 ms with mid, 0.41 with both).
template <typename T, int NY, int NZ>
__global__ void __attribute__((amdgpu_flat_work_group_size(1, P2Geom<NY, NZ>::NT)))
k_plane_mid2p(PropArgs<T> a, int s, int nplanes)
{
    using P = P2<T, NY, NZ>;
    constexpr int NT = P::NT, IZ1 = P::IZ1, A2 = P::A2, NE = NY * NZ;
    extern __shared__ unsigned char smem_raw[];
    cpx<T>* buf = reinterpret_cast<cpx<T>*>(smem_raw);
    cpx<T>* twz = buf + P::LDSN;
    cpx<T>* twy = twz + NZ;
    const int tid0 = threadIdx.x, G = gridDim.x;
    p2_load_tw<T, NY, NZ>(twz, twy, a, tid0, NT);
    const bool qa = s < a.n1, pb = s < a.n2;
    cpx<T> u[P::IY2][P::B1];
    P::inv_load(a.Ch + (size_t)blockIdx.x * NE, u, tid0);
    __syncthreads();  // twiddles
    for (int w = blockIdx.x; w < nplanes; w += G) {
        const size_t fo = (size_t)w * NE;
        // per-thread offsets are recomputed in every iteration rather than held in registers
        // across the loop
        int tid = threadIdx.x;
        asm volatile("" : "+v"(tid));
        P::inv_y2(u, buf, twy, tid);
        T ex[IZ1][A2], ey[IZ1][A2];
#pragma unroll
        for (int it = 0; it < IZ1; ++it) {
            const int wi = tid + it * NT;
            if (P::z1_valid(wi)) {
#pragma unroll
                for (int a2 = 0; a2 < A2; ++a2) {
                    const int e = P::z1_elem(wi, a2);
                    const T va = a.e4A[fo + e], vb = a.e4B[fo + e];
                    ex[it][a2] = qa ? va : vb;
                    ey[it][a2] = pb ? vb : va;
                }
            }
        }
        cpx<T> z[IZ1][A2];
        P::inv_tail(buf, twz, twy, z, tid);
#pragma unroll
        for (int it = 0; it < IZ1; ++it)
#pragma unroll
            for (int a2 = 0; a2 < A2; ++a2) {
                z[it][a2].x = Boltz<T>::mul2(z[it][a2].x, ex[it][a2]);
                z[it][a2].y = Boltz<T>::mul2(z[it][a2].y, ey[it][a2]);
            }
        if (w + G < nplanes) P::inv_load(a.Ch + fo + (size_t)G * NE, u, tid);
        __syncthreads();
        P::fwd(z, buf, twz, twy, a.Ch + fo, tid);
        __syncthreads();  // buf is reused by the next plane
    }
}

// Column pass on the two-stage FFT: for TC consecutive (y,z) columns of one field, forward
// FFT along x, multiply by the kinetic factor, inverse FFT along x (see k_column).
// X1O > 0 sets the split N0 = X1 X2 (X1 = X1O): stage-1 items hold X2 elements, stage-2
// items X1, so a larger X1 moves the longer register DFT to the stage where fewer fields
// are live.
template <int N0, int TC, int X1O = 0>
struct C2Geom {
    static constexpr int X1 = X1O > 0 ? X1O : split_n1(N0), X2 = N0 / X1;
    static constexpr int NT = cdiv(TC * X1, 64) * 64 > 1024 ? 1024 : cdiv(TC * X1, 64) * 64;
    static constexpr int I1 = cdiv(TC * X1, NT), I2 = cdiv(TC * X2, NT);
    static constexpr int PX = X2, CX = X1 * PX + 1;  // TX: c*CX + n1*PX + k2
    static constexpr int LDSN = TC * CX;
    static constexpr bool ok = X1 > 1;
};

template <typename T, int N0, int NY, int NZ, int TC>
__global__ void __attribute__((amdgpu_flat_work_group_size(1, C2Geom<N0, TC>::NT))) k_column2(PropArgs<T> a, int nh, int nf)
{
    using G = C2Geom<N0, TC>;
    constexpr int NT = G::NT, X1 = G::X1, X2 = G::X2, I1 = G::I1, I2 = G::I2, CX = G::CX, PX = G::PX;
    constexpr int NC = NY * NZ;
    extern __shared__ unsigned char smem_raw[];
    cpx<T>* buf = reinterpret_cast<cpx<T>*>(smem_raw);
    cpx<T>* tw0 = buf + G::LDSN;
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
    cpx<T> v[I1][X2];
    // stage 1: load rows n1 + X1 n2 straight from global
#pragma unroll
    for (int it = 0; it < I1; ++it) {
        const int w = tid + it * NT;
        if ((TC * X1) % NT == 0 || w < TC * X1) {
            const int c = w % TC, n1 = w / TC, col = c0 + c;
            const bool in = NC % TC == 0 || col < NC;
            sfor<X2>([&](auto nc) {
                constexpr int n2 = decltype(nc)::value;
                v[it][n2] = in ? fld[(size_t)(n1 + X1 * n2) * NC + col] : cpx<T>{T(0), T(0)};
            });
        }
    }
    // FP32: the x transform's DC coefficient of column (0, 0), i.e. the 3-D DC, is the FP64
    // sum of the column's FP32 values, rounded once. Wave 0 of tile 0 loads the column a
    // second time, one value per lane, with the loads above (so that its latency is theirs),
    // and adds it up after stage 1, when few registers are live.
    constexpr bool DC64 = std::is_same<T, float>::value;
    constexpr int NPL = (N0 + 63) / 64;
    __shared__ double dcx[2];
    cpx<T> dcv[NPL];
    if constexpr (DC64) {
        if (tile == 0 && tid < 64) {
#pragma unroll
            for (int j = 0; j < NPL; ++j) {
                const int x = tid + 64 * j;
                dcv[j] = x < N0 ? fld[(size_t)x * NC] : cpx<T>{T(0), T(0)};
            }
        }
    }
    __syncthreads();  // twiddle / gx tables
#pragma unroll
    for (int it = 0; it < I1; ++it) {
        const int w = tid + it * NT;
        if ((TC * X1) % NT == 0 || w < TC * X1) {
            const int c = w % TC, n1 = w / TC;
            rdft<T, X2, false>(v[it]);
            sfor<X2>([&](auto kc) {
                constexpr int k2 = decltype(kc)::value;
                buf[c * CX + n1 * PX + k2] = k2 == 0 ? v[it][0] : cmul(v[it][k2], tw0[(n1 * k2) % N0]);
            });
        }
    }
    if constexpr (DC64) {
        if (tile == 0 && tid < 64) {
            double sx = 0.0, sy = 0.0;
#pragma unroll
            for (int j = 0; j < NPL; ++j) {
                sx += (double)dcv[j].x;
                sy += (double)dcv[j].y;
            }
            sx = wave_sum64(sx);
            sy = wave_sum64(sy);
            if (tid == 0) {
                dcx[0] = sx;
                dcx[1] = sy;
            }
        }
    }
    __syncthreads();
    cpx<T> u[I2][X1];
#pragma unroll
    for (int it = 0; it < I2; ++it) {
        const int w = tid + it * NT;
        if ((TC * X2) % NT == 0 || w < TC * X2) {
            const int c = w % TC, k2 = w / TC, col = c0 + c;
            sfor<X1>([&](auto nc) {
                constexpr int n1 = decltype(nc)::value;
                u[it][n1] = buf[c * CX + n1 * PX + k2];
            });
            rdft<T, X1, false>(u[it]);
            if constexpr (DC64) {
                if (tile == 0 && w == 0) u[it][0] = {(T)dcx[0], (T)dcx[1]};  // column 0, k2 = 0: kx = 0
            }
            // kinetic factor at kx = k2 + X2 k1 (exp(-k^2 ds/12)/M, or its square * M)
            const int y = (col < NC ? col : 0) / NZ, z = (col < NC ? col : 0) % NZ;
            const T gyz = Kinetic<T>::col(a.gy[y], a.gz[z]);
            sfor<X1>([&](auto kc) {
                constexpr int k1 = decltype(kc)::value;
                Kinetic<T>::apply(u[it][k1], gx[k2 + X2 * k1], gyz, full, a.invM);
            });
            rdft<T, X1, true>(u[it]);
            if constexpr (DC64) {
                // column 0, k2 = 0: the target is the FP64 3-D DC times the kinetic factor at
                // k = 0 (1/M), not its FP32 rounding, which would walk the mean by ~0.3 ulp
                // per column pass
                if (tile == 0 && w == 0) inv_sum_fix<T, X1>(u[it], dcx[0] * (double)a.invM, dcx[1] * (double)a.invM);
            }
            sfor<X1>([&](auto nc) {
                constexpr int n1 = decltype(nc)::value;
                if (n1 != 0) u[it][n1] = cmulc(u[it][n1], tw0[(n1 * k2) % N0]);
            });
        }
    }
    __syncthreads();
#pragma unroll
    for (int it = 0; it < I2; ++it) {
        const int w = tid + it * NT;
        if ((TC * X2) % NT == 0 || w < TC * X2) {
            const int c = w % TC, k2 = w / TC;
            sfor<X1>([&](auto nc) {
                constexpr int n1 = decltype(nc)::value;
                buf[c * CX + n1 * PX + k2] = u[it][n1];
            });
        }
    }
    __syncthreads();
#pragma unroll
    for (int it = 0; it < I1; ++it) {
        const int w = tid + it * NT;
        if ((TC * X1) % NT == 0 || w < TC * X1) {
            const int c = w % TC, n1 = w / TC, col = c0 + c;
            sfor<X2>([&](auto kc) {
                constexpr int k2 = decltype(kc)::value;
                v[it][k2] = buf[c * CX + n1 * PX + k2];
            });
            const cpx<T> vin0 = v[it][0];
            rdft<T, X2, true>(v[it]);
            if constexpr (DC64) {
                if (tile == 0 && c == 0) inv_sum_fix<T, X2>(v[it], vin0.x, vin0.y);  // column 0
            }
            if (NC % TC == 0 || col < NC)
                sfor<X2>([&](auto nc) {
                    constexpr int n2 = decltype(nc)::value;
                    fld[(size_t)(n1 + X1 * n2) * NC + col] = v[it][n2];
                });
        }
    }
}

// Column pass of a box-gradient term on the two-stage FFT. Field f < nh is Ch of slot f,
// f >= nh is Cf of slot f - nh. Every field: forward FFT along x, the spectrum stored to
// K0 (f < nh) or Kf (f >= nh) in natural (kx, ky, kz) order; f < nh: then times the
// half-step kinetic factor / M and inverse along x, back to Ch (as k_column2).
template <typename T, int N0, int NY, int NZ, int TC>
__global__ void __attribute__((amdgpu_flat_work_group_size(1, C2Geom<N0, TC>::NT)))
k_gcol2(PropArgs<T> a, int nh, cpx<T>* __restrict__ K0, cpx<T>* __restrict__ Kf)
{
    using G = C2Geom<N0, TC>;
    constexpr int NT = G::NT, X1 = G::X1, X2 = G::X2, I1 = G::I1, I2 = G::I2, CX = G::CX, PX = G::PX;
    constexpr int NC = NY * NZ;
    extern __shared__ unsigned char smem_raw[];
    cpx<T>* buf = reinterpret_cast<cpx<T>*>(smem_raw);
    cpx<T>* tw0 = buf + G::LDSN;
    T* gx = reinterpret_cast<T*>(tw0 + N0);
    const int tid = threadIdx.x;
    const int tile = blockIdx.x, f = blockIdx.y;
    const bool isf = f >= nh;
    const size_t so = (size_t)(isf ? f - nh : f) * (size_t)a.M;
    cpx<T>* fld = (isf ? a.Cf : a.Ch) + so;
    cpx<T>* ko = (isf ? Kf : K0) + so;
    const int c0 = tile * TC;
    for (int i = tid; i < N0; i += NT) {
        tw0[i] = a.tw0[i];
        gx[i] = a.gx[i];
    }
    cpx<T> v[I1][X2];
#pragma unroll
    for (int it = 0; it < I1; ++it) {
        const int w = tid + it * NT;
        if ((TC * X1) % NT == 0 || w < TC * X1) {
            const int c = w % TC, n1 = w / TC, col = c0 + c;
            const bool in = NC % TC == 0 || col < NC;
            sfor<X2>([&](auto nc) {
                constexpr int n2 = decltype(nc)::value;
                v[it][n2] = in ? fld[(size_t)(n1 + X1 * n2) * NC + col] : cpx<T>{T(0), T(0)};
            });
        }
    }
    __syncthreads();  // twiddle / gx tables
#pragma unroll
    for (int it = 0; it < I1; ++it) {
        const int w = tid + it * NT;
        if ((TC * X1) % NT == 0 || w < TC * X1) {
            const int c = w % TC, n1 = w / TC;
            rdft<T, X2, false>(v[it]);
            sfor<X2>([&](auto kc) {
                constexpr int k2 = decltype(kc)::value;
                buf[c * CX + n1 * PX + k2] = k2 == 0 ? v[it][0] : cmul(v[it][k2], tw0[(n1 * k2) % N0]);
            });
        }
    }
    __syncthreads();
    const bool back = !isf;  // uniform over the workgroup
    cpx<T> u[I2][X1];
#pragma unroll
    for (int it = 0; it < I2; ++it) {
        const int w = tid + it * NT;
        if ((TC * X2) % NT == 0 || w < TC * X2) {
            const int c = w % TC, k2 = w / TC, col = c0 + c;
            sfor<X1>([&](auto nc) {
                constexpr int n1 = decltype(nc)::value;
                u[it][n1] = buf[c * CX + n1 * PX + k2];
            });
            rdft<T, X1, false>(u[it]);
            if (NC % TC == 0 || col < NC)
                sfor<X1>([&](auto kc) {
                    constexpr int k1 = decltype(kc)::value;
                    ko[(size_t)(k2 + X2 * k1) * NC + col] = u[it][k1];
                });
            if (back) {
                const int y = (col < NC ? col : 0) / NZ, z = (col < NC ? col : 0) % NZ;
                const T gyz = Kinetic<T>::col(a.gy[y], a.gz[z]);
                sfor<X1>([&](auto kc) {
                    constexpr int k1 = decltype(kc)::value;
                    Kinetic<T>::apply(u[it][k1], gx[k2 + X2 * k1], gyz, false, a.invM);
                });
                rdft<T, X1, true>(u[it]);
                sfor<X1>([&](auto nc) {
                    constexpr int n1 = decltype(nc)::value;
                    if (n1 != 0) u[it][n1] = cmulc(u[it][n1], tw0[(n1 * k2) % N0]);
                });
            }
        }
    }
    if (!back) return;
    __syncthreads();
#pragma unroll
    for (int it = 0; it < I2; ++it) {
        const int w = tid + it * NT;
        if ((TC * X2) % NT == 0 || w < TC * X2) {
            const int c = w % TC, k2 = w / TC;
            sfor<X1>([&](auto nc) {
                constexpr int n1 = decltype(nc)::value;
                buf[c * CX + n1 * PX + k2] = u[it][n1];
            });
        }
    }
    __syncthreads();
#pragma unroll
    for (int it = 0; it < I1; ++it) {
        const int w = tid + it * NT;
        if ((TC * X1) % NT == 0 || w < TC * X1) {
            const int c = w % TC, n1 = w / TC, col = c0 + c;
            sfor<X2>([&](auto kc) {
                constexpr int k2 = decltype(kc)::value;
                v[it][k2] = buf[c * CX + n1 * PX + k2];
            });
            rdft<T, X2, true>(v[it]);
            if (NC % TC == 0 || col < NC)
                sfor<X2>([&](auto nc) {
                    constexpr int n2 = decltype(nc)::value;
                    fld[(size_t)(n1 + X1 * n2) * NC + col] = v[it][n2];
                });
        }
    }
}

}  // namespace scft
