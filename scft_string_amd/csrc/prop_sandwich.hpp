// prop_sandwich.hpp -- the contour step as two passes: one plane pass and one column pass.
//
// The kinetic factor exp(-k^2 ds/12) of step::do_step (density.h:129-220,
// GPU_kernels.h:207-213) is a product of per-axis factors, so its sandwich F^-1 K F is the
// product of an x-sandwich and a (y, z)-sandwich, which act on different axes and commute:
// K = Kx Kyz. One Richardson step (density.h:129-220)
//
//   q_h = E1/4 K1/2 E1/2 K1/2 E1/4 q,   q_f = E1/2 K E1/2 q,   q' = (4 q_h - q_f) / 3
//
// then regroups as
//
//   column pass (step s-1):  h = K1/2x E1/4 q(s),  f = Kx E1/2 q(s)
//   plane pass  (step s):    h <- K1/2yz E1/2 K1/2yz h,  f <- Kyz f
//   column pass (step s):    q(s+1) = (4 E1/4 K1/2x h - E1/2 f) / 3  -> q store;
//                            then h, f of step s+1 as above.
//
// Each pass keeps its sandwiches in registers (a 2-D FFT of a plane, or 1-D FFTs of a tile of
// x-lines, around LDS transposes), so a step moves each replica's h and f through memory
// twice (once per pass), plus the q-store write and the Boltzmann factors in each pass: 11
// complex-field transfers instead of the 15 of the four-launch step in prop_kernels2.hpp,
// and two launches instead of four. The fields between passes are real-space, natural
// order: nothing is stored in k-space. The arithmetic is the same FFTs, factors and
// Richardson combination; only the order of commuting per-axis factors changes, which moves
// results by rounding.
#pragma once

#include <type_traits>

#include "prop_kernels.hpp"
#include "prop_kernels2.hpp"

namespace scft {

// plane of Z1-ownership registers from / to a real-space plane of a field
template <typename P, typename T>
__device__ __forceinline__ void sp_load(const cpx<T>* __restrict__ in, cpx<T> (&z)[P::IZ1][P::A2], int tid)
{
#pragma unroll
    for (int it = 0; it < P::IZ1; ++it) {
        const int w = tid + it * P::NT;
#pragma unroll
        for (int a2 = 0; a2 < P::A2; ++a2) z[it][a2] = P::z1_valid(w) ? in[P::z1_elem(w, a2)] : cpx<T>{T(0), T(0)};
    }
}

template <typename P, typename T>
__device__ __forceinline__ void sp_store(cpx<T>* __restrict__ out, const cpx<T> (&z)[P::IZ1][P::A2], int tid)
{
#pragma unroll
    for (int it = 0; it < P::IZ1; ++it) {
        const int w = tid + it * P::NT;
        if (P::z1_valid(w))
#pragma unroll
            for (int a2 = 0; a2 < P::A2; ++a2) out[P::z1_elem(w, a2)] = z[it][a2];
    }
}

// (y, z)-sandwich of the plane in registers: forward 2-D FFT, times the (y, z) part of the
// half-step kinetic factor (its square when full) and 1/(NY NZ), inverse 2-D FFT. buf must
// be free on entry; it is in use on return.
template <typename P, typename T>
__device__ __forceinline__ void sp_yz_sandwich(cpx<T> (&z)[P::IZ1][P::A2], cpx<T>* buf, const cpx<T>* twz,
                                               const cpx<T>* twy, const T* gy, const T* gz, bool full,
                                               const InvLen<T>& inv, int tid)
{
    constexpr int NZ = P::A1 * P::A2;
    cpx<T> u[P::IY2][P::B1];
    P::fwd_regs(z, buf, twz, twy, u, tid);
#pragma unroll
    for (int it = 0; it < P::IY2; ++it) {
        const int w = tid + it * P::NT;
        if ((NZ * P::B2) % P::NT == 0 || w < NZ * P::B2) {
            const int kz = w % NZ, l2 = w / NZ;
            const T gzk = gz[kz];
            sfor<P::B1>([&](auto lc) {
                constexpr int l1 = decltype(lc)::value;
                Kinetic<T>::apply(u[it][l1], Kinetic<T>::col(gy[l2 + P::B2 * l1], gzk), Kinetic<T>::one(), full, T(1));
                inv.apply(u[it][l1]);
            });
        }
    }
    __syncthreads();  // the forward's last reads of buf are done
    P::inv_regs(u, buf, twz, twy, z, tid);
}

// Plane pass of step s: h <- K1/2yz E1/2 K1/2yz h (Ch), f <- Kyz f (Cf), plane x of slot y.
template <typename T, int NY, int NZ>
__global__ void __attribute__((amdgpu_flat_work_group_size(1, P2Geom<NY, NZ>::NT))) k_splane2(PropArgs<T> a, int s)
{
    using P = P2<T, NY, NZ>;
    constexpr int NT = P::NT, IZ1 = P::IZ1, A2 = P::A2, NE = NY * NZ;
    extern __shared__ unsigned char smem_raw[];
    cpx<T>* buf = reinterpret_cast<cpx<T>*>(smem_raw);
    cpx<T>* twz = buf + P::LDSN;
    cpx<T>* twy = twz + NZ;
    T* gy = reinterpret_cast<T*>(twy + NY);
    T* gz = gy + NY;
    const int tid = threadIdx.x;
    const int x = blockIdx.x, slot = blockIdx.y;
    const size_t fo = (size_t)slot * (size_t)a.M + (size_t)x * NE;
    p2_load_tw<T, NY, NZ>(twz, twy, a, tid, NT);
    for (int i = tid; i < NY; i += NT) gy[i] = a.gy[i];
    for (int i = tid; i < NZ; i += NT) gz[i] = a.gz[i];
    const InvLen<T> inv(NE);
    cpx<T> z[IZ1][A2];
    sp_load<P>(a.Ch + fo, z, tid);
    __syncthreads();  // tables
    sp_yz_sandwich<P>(z, buf, twz, twy, gy, gz, false, inv, tid);
    // E1/2 of step s: q1 (real part) takes species A while s < n1, q2 (imaginary part) species
    // B while s < n2 (density.h:136-146)
    const bool qa = s < a.n1, pb = s < a.n2;
#pragma unroll
    for (int it = 0; it < IZ1; ++it) {
        const int w = tid + it * NT;
        if (P::z1_valid(w)) {
#pragma unroll
            for (int a2 = 0; a2 < A2; ++a2) {
                const int e = P::z1_elem(w, a2);
                const T va = a.e4A[fo + e], vb = a.e4B[fo + e];
                z[it][a2].x = Boltz<T>::mul2(z[it][a2].x, qa ? va : vb);
                z[it][a2].y = Boltz<T>::mul2(z[it][a2].y, pb ? vb : va);
            }
        }
    }
    // f is loaded before the second h sandwich so that its reads overlap that sandwich where
    // the registers allow: the extra field spills at 128^2 FP32 and costs a wave per SIMD at
    // 96^2 FP32, where the later load is faster
    constexpr bool early_f = !(std::is_same<T, float>::value && NE > 64 * 64);
    cpx<T> zf[IZ1][A2];
    if (early_f) sp_load<P>(a.Cf + fo, zf, tid);
    __syncthreads();
    sp_yz_sandwich<P>(z, buf, twz, twy, gy, gz, false, inv, tid);
    sp_store<P>(a.Ch + fo, z, tid);
    if (!early_f) sp_load<P>(a.Cf + fo, zf, tid);
    __syncthreads();
    sp_yz_sandwich<P>(zf, buf, twz, twy, gy, gz, true, inv, tid);
    sp_store<P>(a.Cf + fo, zf, tid);
}

// x-sandwich of a tile of TC x-lines in stage-1 ownership (item (c, n1) holds
// x = n1 + X1 n2, n2 < X2, of column c0 + c): forward FFT along x, times the x part of the
// half-step kinetic factor (its square when full) and 1/N0, inverse. buf must be free on
// entry; it is in use on return.
//
// FP32: the sandwich keeps each line's sum, which the kinetic factor (1 at k = 0) leaves
// unchanged, in FP64. In FP32 the forward transform's DC is a rounded tree sum, and the
// inverse hands it to every output through shared rounded partial sums; where many lines
// carry the same data (a lamella along x makes every x-line identical) those roundings are
// the same in every line and move the mean of q, and with it Q, coherently. So: each stage-1
// item's sum goes to dsum in FP64; after the inverse, the line's stage-2 DC item folds the
// FP64 residual of its outputs' sum against the FP64 line sum / N0 into one output
// (inv_sum_fix), as every stage-1' item does against its own DC input. The line sum is then
// exact to the rounding of single outputs, which no longer lines up across lines.
//
// ROW: stage-1 item w is (c, n1) = (w / X1, w % X1) instead of (w % TC, w / TC), so that
// consecutive lanes hold consecutive elements of one line (lines contiguous in memory).
template <typename G, typename T, int N0, int TC, bool ROW = false>
__device__ __forceinline__ void sc_x_sandwich(cpx<T> (&v)[G::I1][G::X2], cpx<T>* buf, const cpx<T>* tw0, const T* gx,
                                              bool full, const InvLen<T>& inv, double2* dsum, int tid)
{
    constexpr int NT = G::NT, X1 = G::X1, X2 = G::X2, I1 = G::I1, I2 = G::I2, CX = G::CX, PX = G::PX;
    constexpr bool F32 = std::is_same<T, float>::value;
#pragma unroll
    for (int it = 0; it < I1; ++it) {
        const int w = tid + it * NT;
        if ((TC * X1) % NT == 0 || w < TC * X1) {
            const int c = ROW ? w / X1 : w % TC, n1 = ROW ? w % X1 : w / TC;
            if constexpr (F32) {
                double sx = 0.0, sy = 0.0;
#pragma unroll
                for (int n2 = 0; n2 < X2; ++n2) {
                    sx += (double)v[it][n2].x;
                    sy += (double)v[it][n2].y;
                }
                dsum[c * X1 + n1] = make_double2(sx, sy);
            }
            rdft<T, X2, false>(v[it]);
            sfor<X2>([&](auto kc) {
                constexpr int k2 = decltype(kc)::value;
                buf[c * CX + n1 * PX + k2] = k2 == 0 ? v[it][0] : cmul(v[it][k2], tw0[(n1 * k2) % N0]);
            });
        }
    }
    __syncthreads();
    cpx<T> u[I2][X1];
#pragma unroll
    for (int it = 0; it < I2; ++it) {
        const int w = tid + it * NT;
        if ((TC * X2) % NT == 0 || w < TC * X2) {
            const int c = w % TC, k2 = w / TC;
            sfor<X1>([&](auto nc) {
                constexpr int n1 = decltype(nc)::value;
                u[it][n1] = buf[c * CX + n1 * PX + k2];
            });
            rdft<T, X1, false>(u[it]);
            sfor<X1>([&](auto kc) {
                constexpr int k1 = decltype(kc)::value;
                Kinetic<T>::apply(u[it][k1], gx[k2 + X2 * k1], Kinetic<T>::one(), full, T(1));
                inv.apply(u[it][k1]);
            });
            rdft<T, X1, true>(u[it]);
            if constexpr (F32) {
                // outputs of the DC item sum to X1 times its input: the line mean (the k = 0
                // factor is 1), here the FP64 line sum / N0
                if (k2 == 0) {
                    double dcx = 0.0, dcy = 0.0;
#pragma unroll
                    for (int n1 = 0; n1 < X1; ++n1) {
                        const double2 p = dsum[c * X1 + n1];
                        dcx += p.x;
                        dcy += p.y;
                    }
                    inv_sum_fix<T, X1>(u[it], dcx / (double)N0, dcy / (double)N0);
                }
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
            const int c = ROW ? w / X1 : w % TC, n1 = ROW ? w % X1 : w / TC;
            sfor<X2>([&](auto kc) {
                constexpr int k2 = decltype(kc)::value;
                v[it][k2] = buf[c * CX + n1 * PX + k2];
            });
            const cpx<T> vin0 = v[it][0];
            rdft<T, X2, true>(v[it]);
            if constexpr (F32) inv_sum_fix<T, X2>(v[it], (double)vin0.x, (double)vin0.y);
        }
    }
}

// Dynamic LDS of k_scol2: the transpose buffer, twiddles and kinetic table of the x axis,
// and in FP32 the FP64 item sums of sc_x_sandwich.
template <typename T, int N0, int TC, int X1O = 0>
constexpr size_t scol_lds_bytes()
{
    using G = C2Geom<N0, TC, X1O>;
    const size_t base = (size_t)(G::LDSN + N0) * sizeof(cpx<T>) + (size_t)N0 * sizeof(T);
    return std::is_same<T, float>::value ? (base + 15) / 16 * 16 + (size_t)TC * G::X1 * 16 : base;
}

// Column pass over TC (y, z) columns of slot blockIdx.y. PL_STEP / PL_LAST complete step s:
// h <- K1/2x h, q(s+1) = (4 E1/4 h - E1/2 f) / 3 to the q store; PL_BEGIN starts the contour
// from q(0) = 1, PL_RESUME from the stored q at index s (entry a.qsrc). Then, except for
// PL_LAST, the fields of the next step: Ch <- K1/2x E1/4 q, Cf <- Kx E1/2 q.
// NY = NZ = 0: any (y, z) extent, NC = m1 m2 at run time (the x pass of the split line step,
// whose fields hold the same state between passes).
template <typename T, int N0, int NY, int NZ, int TC, int X1O = 0>
__global__ void __attribute__((amdgpu_flat_work_group_size(1, C2Geom<N0, TC, X1O>::NT)))
k_scol2(PropArgs<T> a, int s, int mode)
{
    using G = C2Geom<N0, TC, X1O>;
    constexpr int NT = G::NT, X1 = G::X1, X2 = G::X2, I1 = G::I1;
    const int NC = NY > 0 ? NY * NZ : a.m1 * a.m2;
    constexpr bool ctile = NY > 0 && (NY * NZ) % TC == 0;  // every tile full
    // f is read before the h sandwich, which its reads overlap, where a line's f fits in
    // registers beside h (X2 elements of 8 bytes or fewer, or a compiled plane grid); longer
    // lines read it after the sandwich
    constexpr bool early_f = NY > 0 || X2 * sizeof(T) <= 64;
    extern __shared__ unsigned char smem_raw[];
    cpx<T>* buf = reinterpret_cast<cpx<T>*>(smem_raw);
    cpx<T>* tw0 = buf + G::LDSN;
    T* gx = reinterpret_cast<T*>(tw0 + N0);
    // FP32: per-item FP64 line sums (sc_x_sandwich), after the tables at a 16-byte boundary
    double2* dsum = reinterpret_cast<double2*>(
        smem_raw + ((G::LDSN + N0) * sizeof(cpx<T>) + N0 * sizeof(T) + 15) / 16 * 16);
    const int tid = threadIdx.x;
    const int c0 = blockIdx.x * TC, slot = blockIdx.y;
    const size_t so = (size_t)slot * (size_t)a.M;
    for (int i = tid; i < N0; i += NT) {
        tw0[i] = a.tw0[i];
        gx[i] = a.gx[i];
    }
    const InvLen<T> inv(N0);
    // item it of this thread: valid, and its column inside the grid
    auto item_ok = [&](int w) { return ((TC * X1) % NT == 0 || w < TC * X1) && (ctile || c0 + w % TC < NC); };
    auto elem = [&](int w, int n2) { return (size_t)(w / TC + X1 * n2) * NC + (size_t)(c0 + w % TC); };
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
    T ea[I1][X2], eb[I1][X2];
    auto load_factors = [&]() {
#pragma unroll
        for (int it = 0; it < I1; ++it) {
            const int w = tid + it * NT;
            const bool ok = item_ok(w);
#pragma unroll
            for (int n2 = 0; n2 < X2; ++n2) {
                ea[it][n2] = ok ? a.e4A[so + elem(w, n2)] : T(0);
                eb[it][n2] = ok ? a.e4B[so + elem(w, n2)] : T(0);
            }
        }
    };
    cpx<T> q[I1][X2];
    int sout;
    if (mode == PL_BEGIN) {
#pragma unroll
        for (int it = 0; it < I1; ++it)
#pragma unroll
            for (int n2 = 0; n2 < X2; ++n2) q[it][n2] = {T(1), T(1)};
        load_factors();
        sout = 0;
        __syncthreads();  // tables
    } else if (mode == PL_RESUME) {
        load(a.qs + ((size_t)slot * (size_t)a.qpitch + (size_t)a.qsrc) * (size_t)a.M, q);
        load_factors();
        sout = s;
        __syncthreads();  // tables
    } else {
        load(a.Ch + so, q);  // q holds h until the combination below
        cpx<T> f[I1][X2];
        if (early_f) load(a.Cf + so, f);
        __syncthreads();  // tables
        sc_x_sandwich<G, T, N0, TC>(q, buf, tw0, gx, false, inv, dsum, tid);
        if (!early_f) load(a.Cf + so, f);
        load_factors();
        const bool qa = s < a.n1, pb = s < a.n2;  // species of step s (density.h:136-146)
#pragma unroll
        for (int it = 0; it < I1; ++it)
#pragma unroll
            for (int n2 = 0; n2 < X2; ++n2) {
                const T ex = qa ? ea[it][n2] : eb[it][n2];
                const T ey = pb ? eb[it][n2] : ea[it][n2];
                q[it][n2].x = Boltz<T>::rich(q[it][n2].x, f[it][n2].x, ex);
                q[it][n2].y = Boltz<T>::rich(q[it][n2].y, f[it][n2].y, ey);
            }
        sout = s + 1;
        __syncthreads();  // buf is reused below
    }
    const int qix = q_index(a, sout);
    if (qix >= 0) {
        cpx<T>* qo = a.qs + ((size_t)slot * (size_t)a.qpitch + (size_t)qix) * (size_t)a.M;
#pragma unroll
        for (int it = 0; it < I1; ++it) {
            const int w = tid + it * NT;
            if (item_ok(w))
#pragma unroll
                for (int n2 = 0; n2 < X2; ++n2) store_q(&qo[elem(w, n2)], q[it][n2]);
        }
    }
    if (mode == PL_LAST) return;
    const bool qa = sout < a.n1, pb = sout < a.n2;
    cpx<T> f[I1][X2];
#pragma unroll
    for (int it = 0; it < I1; ++it)
#pragma unroll
        for (int n2 = 0; n2 < X2; ++n2) {
            const T ex = qa ? ea[it][n2] : eb[it][n2];
            const T ey = pb ? eb[it][n2] : ea[it][n2];
            f[it][n2] = {Boltz<T>::mul2(q[it][n2].x, ex), Boltz<T>::mul2(q[it][n2].y, ey)};
            q[it][n2] = {Boltz<T>::mul4(q[it][n2].x, ex), Boltz<T>::mul4(q[it][n2].y, ey)};
        }
    sc_x_sandwich<G, T, N0, TC>(q, buf, tw0, gx, false, inv, dsum, tid);
    store(a.Ch + so, q);
    __syncthreads();
    sc_x_sandwich<G, T, N0, TC>(f, buf, tw0, gx, true, inv, dsum, tid);
    store(a.Cf + so, f);
}

}  // namespace scft
