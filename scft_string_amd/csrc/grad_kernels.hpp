// grad_kernels.hpp -- d lnQ / dD (box derivative at fixed W) on the GPU: the adjoint of
// density::props (density.h:344-399).
//
// The Richardson step S_s = (4 E4 K12 E2 K12 E4 - E2 K6 E2)/3 is symmetric, so the adjoint
// of q1 after step s is q2[N-1-s], already in the q store:
//   dQC/dD_i = sum_s <q2[N-1-s], dS_s/dD_i q1[s]>,   <f, g> = mean(f g),
// and Q2 = <q2[Nh]> likewise with a = q2[Nh-1-t], b = q2[t]. Per step, with
// v0 = E4 b, u0 = E4 a, v1 = E2 K12 v0, u1 = E2 K12 u0, vf = E2 b, uf = E2 a,
//   <a, dS b> = (4 (<u0, dK12 v1> + <u1, dK12 v0>) - <uf, dK6 vf>) / 3.
// a and b take the same step operator, so they are packed as Z = b + i a and carried
// through one sandwich pipeline: Z0 = E4 Z, Z1 = E2 K12 Z0, Zf = E2 Z. For a real
// diagonal multiplier d(k),
//   Im mean(Z0 * dK Z1) = <v0, dK u1> + <u0, dK v1>  and  Im mean(Zf * dK Zf) = 2 <vf, dK uf>,
// and mean(X * dK Y) = M^-2 sum_k d(k) X^(-k) Y^(k) for the unnormalised transforms used
// here. So each step needs the full spectra of Z0, Z1 and Zf and one reduction over k:
//   k_gz(INIT)  Z0, Zf from the q store and E4 -> forward z       (A, B)
//   k_py        forward y on A, B
//   k_gx(SPLIT) A: forward x -> K0, times K12, inverse x -> A;  B: forward x -> Kf
//   k_py^-1 A, k_gz(MID) inverse z, times E2, forward z, k_py A
//   k_gx(FWD)   A: forward x -> K1
//   k_gacc      sum_k dK12/dD_i(k) Im(K0(-k) K1(k)) and dK6/dD_i(k) Im(Kf(-k) Kf(k))
// with dK12/dD_i = K12 k_i^2 ds / (6 D_i) and dK6/dD_i = K6 k_i^2 ds / (3 D_i), from
// k^2 = sum_j (2 pi n_j / D_j)^2 (density.h:79-95). That is the line pipeline (any 2-3-5
// grid). Grids with a two-stage kernel set take four passes per step instead:
//   k_gplane_init2   Z0, Zf from the q store and E4, E2 -> (y, z) spectra   (A, B)
//   k_gcol2          A: forward x -> K0, times K12, inverse x -> A;  B: forward x -> Kf
//   k_plane_mid2     A: inverse (y, z), times E2, forward (y, z)
//   k_gcol2_acc      A: forward x -> K1 in registers, the sum over k with K0, Kf read at +-k
#pragma once
#include "prop_kernels2.hpp"
#include "prop_pencil.hpp"

namespace scft {

enum GradMode { GZ_INIT = 0, GZ_MID = 1, GX_SPLIT = 2, GX_FWD = 3 };

// z-line pass of the gradient pipeline. grid (ceil(m0 m1 / nlz), nb), block PC_NT.
// GZ_INIT: b = part pb of q-store slot sb, a = part pa of slot sa (0 real, 1 imaginary);
//          A <- Fz(E4 (b + i a)), B <- Fz(E2 (b + i a)) with the factors of species spec.
// GZ_MID:  A <- Fz(E2 Fz^-1(A)).
// A, B are a.Ch, a.Cf. LDS as k_pz. L > 0: compiled for z length L (tile pc_ct_nlz).
template <typename T, int L = 0>
__global__ void __launch_bounds__(PC_NT) k_gz(PropArgs<T> a, int mode, int sb, int pb, int sa, int pa,
                                               int spec, int nlz_arg)
{
    constexpr int ZE = pc_ept<T, L>() / 2;  // elements per thread and field
    extern __shared__ unsigned char smem_raw[];
    const int m2 = L > 0 ? L : a.m2;
    const int nlz = L > 0 ? pc_ct_nlz<T>(L) : nlz_arg;
    const int nfl = mode == GZ_MID ? 1 : 2;
    const int nlines = nfl * nlz, P = pc_pitch(nlines);
    cpx<T>* buf = reinterpret_cast<cpx<T>*>(smem_raw);
    cpx<T>* tw = buf + (size_t)m2 * P;
    const int tid = threadIdx.x, slot = blockIdx.y;
    const long long l0 = (long long)blockIdx.x * nlz;
    const long long lrem = (long long)a.m0 * a.m1 - l0;
    const int nl = lrem < nlz ? (int)lrem : nlz;
    const size_t fo = (size_t)slot * (size_t)a.M + (size_t)l0 * m2;
    const int ne = nl * m2, nt = nlz * m2;
    const T* ef = spec == 0 ? a.e4A : a.e4B;

    T ev[ZE];
    cpx<T> c0[ZE];
    if (mode == GZ_INIT) {
        const cpx<T>* qb = a.qs + ((size_t)slot * (size_t)a.qpitch + (size_t)sb) * (size_t)a.M + (size_t)l0 * m2;
        const cpx<T>* qa = a.qs + ((size_t)slot * (size_t)a.qpitch + (size_t)sa) * (size_t)a.M + (size_t)l0 * m2;
#pragma unroll
        for (int i = 0; i < ZE; ++i) {
            const int e = tid + i * PC_NT;
            const bool ok = e < ne;
            ev[i] = ok ? ef[fo + e] : T(0);
            const cpx<T> vb = ok ? qb[e] : cpx<T>{T(0), T(0)};
            const cpx<T> va = ok ? qa[e] : cpx<T>{T(0), T(0)};
            c0[i] = {pb ? vb.y : vb.x, pa ? va.y : va.x};
        }
    } else {
#pragma unroll
        for (int i = 0; i < ZE; ++i) {
            const int e = tid + i * PC_NT;
            const bool ok = e < ne;
            ev[i] = ok ? ef[fo + e] : T(0);
            c0[i] = ok ? a.Ch[fo + e] : cpx<T>{T(0), T(0)};
        }
    }
    for (int i = tid; i < m2; i += PC_NT) tw[i] = a.tw2[i];

    if (mode == GZ_MID) {
#pragma unroll
        for (int i = 0; i < ZE; ++i) {
            const int e = tid + i * PC_NT;
            if (e < nt) {
                const int l = e / m2, z = e - l * m2;
                buf[z * P + l] = c0[i];
            }
        }
        __syncthreads();
        pz_fft<T, true, L>(buf, m2, nfl, nlz, P, tw, tid);
#pragma unroll
        for (int i = 0; i < ZE; ++i) {
            const int e = tid + i * PC_NT;
            if (e < ne) {
                const int l = e / m2, z = e - l * m2;
                cpx<T>& v = buf[z * P + l];
                v = {Boltz<T>::mul2(v.x, ev[i]), Boltz<T>::mul2(v.y, ev[i])};
            }
        }
    } else {
#pragma unroll
        for (int i = 0; i < ZE; ++i) {
            const int e = tid + i * PC_NT;
            if (e < nt) {
                const int l = e / m2, z = e - l * m2;
                const cpx<T> v = c0[i];
                buf[z * P + l] = {Boltz<T>::mul4(v.x, ev[i]), Boltz<T>::mul4(v.y, ev[i])};
                buf[z * P + nlz + l] = {Boltz<T>::mul2(v.x, ev[i]), Boltz<T>::mul2(v.y, ev[i])};
            }
        }
    }
    __syncthreads();
    pz_fft<T, false, L>(buf, m2, nfl, nlz, P, tw, tid);
#pragma unroll
    for (int i = 0; i < ZE; ++i) {
        const int e = tid + i * PC_NT;
        if (e < ne) {
            const int l = e / m2, z = e - l * m2;
            a.Ch[fo + e] = buf[z * P + l];
            if (nfl == 2) a.Cf[fo + e] = buf[z * P + nlz + l];
        }
    }
}

// x-line pass of the gradient pipeline on tx z-columns of one y.
// grid (m1 * ceil(m2 / tx), nh + nf fields), block PC_NT. LDS as k_px.
// GX_SPLIT: field f < nh (A of slot f): forward x -> K0, times K12, inverse x -> A;
//           field f >= nh (B of slot f - nh): forward x -> Kf.
// GX_FWD:   field f < nh: forward x -> K0 (the caller passes K1 there).
// L > 0: compiled for x length L (tile pc_ct_t).
template <typename T, int L = 0>
__global__ void __launch_bounds__(PC_NT) k_gx(PropArgs<T> a, int mode, int nh, int tx_arg, cpx<T>* __restrict__ K0,
                                               cpx<T>* __restrict__ Kf)
{
    constexpr int CE = pc_ept<T, L>();  // elements per thread
    extern __shared__ unsigned char smem_raw[];
    const int m0 = L > 0 ? L : a.m0, m2 = a.m2;
    const int tx = L > 0 ? pc_ct_t<T>(L) : tx_arg, P = pc_pitch(tx);
    const size_t sx = (size_t)a.m1 * m2;
    cpx<T>* buf = reinterpret_cast<cpx<T>*>(smem_raw);
    cpx<T>* tw = buf + (size_t)m0 * P;
    const int tid = threadIdx.x;
    const int ntz = (m2 + tx - 1) / tx;
    const int bt = pc_xcd_block();
    const int y = bt / ntz, z0 = (bt - y * ntz) * tx;
    const int nz = m2 - z0 < tx ? m2 - z0 : tx;
    const int f = blockIdx.y;
    const bool isB = f >= nh;
    const size_t off = (size_t)(isB ? f - nh : f) * (size_t)a.M + (size_t)y * m2 + z0;
    cpx<T>* g = (isB ? a.Cf : a.Ch) + off;
    cpx<T>* ko = (isB ? Kf : K0) + off;
    const int ne = m0 * tx;
    cpx<T> r[CE];
#pragma unroll
    for (int i = 0; i < CE; ++i) {
        const int e = tid + i * PC_NT;
        const int x = e / tx, zz = e - x * tx;
        r[i] = (e < ne && zz < nz) ? g[x * sx + zz] : cpx<T>{T(0), T(0)};
    }
    for (int i = tid; i < m0; i += PC_NT) tw[i] = a.tw0[i];
#pragma unroll
    for (int i = 0; i < CE; ++i) {
        const int e = tid + i * PC_NT;
        const int x = e / tx, zz = e - x * tx;
        if (e < ne) buf[x * P + zz] = r[i];
    }
    __syncthreads();
    pl_fft<T, false, L>(buf, m0, tx, P, tw, tid);
#pragma unroll
    for (int i = 0; i < CE; ++i) {
        const int e = tid + i * PC_NT;
        const int x = e / tx, zz = e - x * tx;
        if (e < ne && zz < nz) ko[x * sx + zz] = buf[x * P + zz];
    }
    if (mode != GX_SPLIT || isB) return;
    const T gy = a.gy[y];
#pragma unroll
    for (int i = 0; i < CE; ++i) {
        const int e = tid + i * PC_NT;
        const int x = e / tx, zz = e - x * tx;
        if (e < ne && zz < nz)
            Kinetic<T>::apply(buf[x * P + zz], a.gx[x], Kinetic<T>::col(gy, a.gz[z0 + zz]), false, a.invM);
    }
    __syncthreads();
    pl_fft<T, true, L>(buf, m0, tx, P, tw, tid);
#pragma unroll
    for (int i = 0; i < CE; ++i) {
        const int e = tid + i * PC_NT;
        const int x = e / tx, zz = e - x * tx;
        if (e < ne && zz < nz) g[x * sx + zz] = buf[x * P + zz];
    }
}

// Adds, per slot and direction i, sum_k [4/3 dK12_i(k) Im(K0(-k) K1(k)) - 1/6 dK6_i(k)
// Im(Kf(-k) Kf(k))] to part[slot][blk][3] (scaled by M^-2 on the host), where
// dK12_i = exp(-k^2 ds/12) k_i^2 ds / (6 D_i) and dK6_i = exp(-k^2 ds/6) k_i^2 ds / (3 D_i).
// grid (nblk, nb), block RED_NT. Launches run in order on one stream, so each block's
// read-modify-write of its own partial is deterministic.
template <typename T>
__global__ void __launch_bounds__(RED_NT) k_gacc(const cpx<T>* __restrict__ K0, const cpx<T>* __restrict__ K1,
                                                  const cpx<T>* __restrict__ Kf, long long M, int m0, int m1,
                                                  int m2, GradConst gc, double* __restrict__ part)
{
    const int slot = blockIdx.y;
    const size_t so = (size_t)slot * (size_t)M;
    double t[3] = {0, 0, 0};
    // one slot's field has M < 2^31 points: 32-bit index arithmetic (a 64-bit division per
    // point cost more than the loads)
    const int Mi = (int)M;
    for (int r = blockIdx.x * RED_NT + threadIdx.x; r < Mi; r += gridDim.x * RED_NT) {
        const int rq = r / m2;
        const int kz = r - rq * m2;
        const int kx = rq / m1;
        const int ky = rq - kx * m1;
        const int nx = kx <= m0 / 2 ? kx : kx - m0, ny = ky <= m1 / 2 ? ky : ky - m1,
                  nz = kz <= m2 / 2 ? kz : kz - m2;
        const int mx = kx == 0 ? 0 : m0 - kx, my = ky == 0 ? 0 : m1 - ky, mz = kz == 0 ? 0 : m2 - kz;
        const int rm = (mx * m1 + my) * m2 + mz;
        const double k2x = (gc.kD[0] * nx) * (gc.kD[0] * nx), k2y = (gc.kD[1] * ny) * (gc.kD[1] * ny),
                     k2z = (gc.kD[2] * nz) * (gc.kD[2] * nz);
        const double kh = exp(-(k2x + k2y + k2z) * gc.ds / 12.0);
        const cpx<T> a0 = K0[so + rm], b1 = K1[so + r], fm = Kf[so + rm], fp = Kf[so + r];
        const double bh = (double)a0.x * (double)b1.y + (double)a0.y * (double)b1.x;
        const double bf = (double)fm.x * (double)fp.y + (double)fm.y * (double)fp.x;
        // 4/3 * ds/(6 D_i) * kh * bh - 1/3 * ds/(3 D_i) * kh^2 * bf / 2, times k_i^2
        const double c = gc.ds * ((2.0 / 9.0) * kh * bh - (1.0 / 18.0) * kh * kh * bf);
        t[0] += c * k2x;
        t[1] += c * k2y;
        t[2] += c * k2z;
    }
    // the 1/D_i of every term, once per thread; then each wave sums its 64 lanes by
    // shuffles and one barrier hands the RED_NT / 64 wave sums to thread 0
    __shared__ double wsum[3][RED_NT / 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        double v = t[i] / gc.D[i];
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
        if (lane == 0) wsum[i][wave] = v;
    }
    __syncthreads();
    if (threadIdx.x < 3) {
        double v = 0;
        for (int w = 0; w < RED_NT / 64; ++w) v += wsum[threadIdx.x][w];
        part[((size_t)slot * gridDim.x + blockIdx.x) * 3 + threadIdx.x] += v;
    }
}

// Last pass of a box-gradient term on the two-stage FFT, fused with the reduction over k.
// Ch of each slot holds the (y, z) spectrum of Z1 = E1/2 K12 Z0 with x in natural order; the
// forward FFT along x gives K1(k) on the workgroup's TC columns, which is paired in registers
// with K0(-k), Kf(-k) and Kf(k) (written by k_gcol2) to add, per slot and
// direction i, the columns' share of
//   sum_k [4/3 dK12_i(k) Im(K0(-k) K1(k)) - 1/6 dK6_i(k) Im(Kf(-k) Kf(k))]
// to part[slot][tile][i] (scaled by M^-2 on the host), the sum k_gacc forms from stored
// spectra. K1 is never stored. grid (ntile, nb), block C2Geom<N0, TC>::NT. A slot's launches
// run in order on one stream, so each workgroup's read-modify-write of its partial is
// deterministic.
template <typename T, int N0, int NY, int NZ, int TC>
__global__ void __attribute__((amdgpu_flat_work_group_size(1, C2Geom<N0, TC>::NT)))
k_gcol2_acc(PropArgs<T> a, const cpx<T>* __restrict__ K0, const cpx<T>* __restrict__ Kf, GradConst gc,
            double* __restrict__ part)
{
    using G = C2Geom<N0, TC>;
    constexpr int NT = G::NT, X1 = G::X1, X2 = G::X2, I1 = G::I1, I2 = G::I2, CX = G::CX, PX = G::PX;
    constexpr int NC = NY * NZ;
    static_assert(NT % 64 == 0, "whole waves");
    extern __shared__ unsigned char smem_raw[];
    cpx<T>* buf = reinterpret_cast<cpx<T>*>(smem_raw);
    cpx<T>* tw0 = buf + G::LDSN;
    const int tid = threadIdx.x;
    const int tile = blockIdx.x, slot = blockIdx.y;
    const size_t so = (size_t)slot * (size_t)a.M;
    const cpx<T>* fld = a.Ch + so;
    const cpx<T>* k0 = K0 + so;
    const cpx<T>* kf = Kf + so;
    const int c0 = tile * TC;
    for (int i = tid; i < N0; i += NT) tw0[i] = a.tw0[i];
    // stage 1: rows n1 + X1 n2 of each column, forward along n2, twiddled into LDS
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
    __syncthreads();  // twiddle table
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
    // stage 2: forward along n1 gives K1 at kx = k2 + X2 k1; each point meets its mirror
    double t[3] = {0, 0, 0};
#pragma unroll
    for (int it = 0; it < I2; ++it) {
        const int w = tid + it * NT;
        if ((TC * X2) % NT == 0 || w < TC * X2) {
            const int c = w % TC, k2 = w / TC, col = c0 + c;
            cpx<T> u[X1];
            sfor<X1>([&](auto nc) {
                constexpr int n1 = decltype(nc)::value;
                u[n1] = buf[c * CX + n1 * PX + k2];
            });
            rdft<T, X1, false>(u);
            if (NC % TC == 0 || col < NC) {
                const int ky = col / NZ, kz = col - ky * NZ;
                const int ny = ky <= NY / 2 ? ky : ky - NY, nz = kz <= NZ / 2 ? kz : kz - NZ;
                const int colm = (ky == 0 ? 0 : NY - ky) * NZ + (kz == 0 ? 0 : NZ - kz);
                const double k2y = (gc.kD[1] * ny) * (gc.kD[1] * ny), k2z = (gc.kD[2] * nz) * (gc.kD[2] * nz);
                sfor<X1>([&](auto kc) {
                    constexpr int k1 = decltype(kc)::value;
                    const int kx = k2 + X2 * k1;
                    const int nx = kx <= N0 / 2 ? kx : kx - N0, kxm = kx == 0 ? 0 : N0 - kx;
                    const double k2x = (gc.kD[0] * nx) * (gc.kD[0] * nx);
                    // every point reads its own Kf and its mirror's: taking each +-k pair once
                    // (the term is even) from half the columns made the pass slower
                    const cpx<T> a0 = k0[(size_t)kxm * NC + colm], fm = kf[(size_t)kxm * NC + colm],
                                 fp = kf[(size_t)kx * NC + col];
                    const double bh = (double)a0.x * (double)u[k1].y + (double)a0.y * (double)u[k1].x;
                    const double bf = (double)fm.x * (double)fp.y + (double)fm.y * (double)fp.x;
                    const double kh = exp(-(k2x + k2y + k2z) * gc.ds / 12.0);
                    // 4/3 * ds/(6 D_i) * kh * bh - 1/3 * ds/(3 D_i) * kh^2 * bf / 2, times k_i^2
                    const double cf = gc.ds * ((2.0 / 9.0) * kh * bh - (1.0 / 18.0) * kh * kh * bf);
                    t[0] += cf * k2x;
                    t[1] += cf * k2y;
                    t[2] += cf * k2z;
                });
            }
        }
    }
    // the 1/D_i once per thread; each wave sums its lanes by shuffles, and the wave sums
    // meet in LDS (buf is free once every wave has read its stage-2 items)
    __syncthreads();
    double* wsum = reinterpret_cast<double*>(smem_raw);
    const int lane = tid & 63, wave = tid >> 6;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        double x = t[i] / gc.D[i];
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) x += __shfl_down(x, o, 64);
        if (lane == 0) wsum[i * (NT / 64) + wave] = x;
    }
    __syncthreads();
    if (tid < 3) {
        double x = 0;
        for (int w = 0; w < NT / 64; ++w) x += wsum[tid * (NT / 64) + w];
        part[((size_t)slot * gridDim.x + tile) * 3 + tid] += x;
    }
}

}  // namespace scft
