// prop_lines2.hpp -- the line (pencil) contour step with the kinetic factor split by axis.
//
// The kinetic factor exp(-k^2 ds/12) (density.h:129-220, GPU_kernels.h:207-213) is a product
// of per-axis factors and the 3-D DFT a product of per-axis DFTs, so every sandwich F^-1 K F
// is Kx Ky Kz, three commuting one-axis sandwiches, each of which a line pass completes in
// LDS. One Richardson step (q_h = E1/4 K1/2 E1/2 K1/2 E1/4 q, q_f = E1/2 K E1/2 q,
// q' = (4 q_h - q_f) / 3) then runs as four line passes:
//
//   k_ly  (h, f):  h <- K1/2y h,  f <- Ky f
//   k_lz  (h, f):  h <- K1/2z E1/2 K1/2z h,  f <- Kz f          (E1/2 of step s)
//   k_ly  (h):     h <- K1/2y h
//   k_lx  (h, f):  h <- K1/2x h; q(s+1) = (4 E1/4 h - E1/2 f) / 3 -> q store;
//                  then h = K1/2x E1/4 q(s+1), f = Kx E1/2 q(s+1) for step s + 1
//
// with h, f real-space in natural order between passes: 17 complex-field transfers per
// replica and step (factors and the q store included), against 27 for the pencil step of
// prop_pencil.hpp, whose y and z passes only transform. The tiles, LDS layout and line
// transforms (runtime radix plan or compiled length) are those of prop_pencil.hpp.
#pragma once

#include "prop_pencil.hpp"

namespace scft {

// Kinetic factor of one axis at spectral index k (table g, its square when full) and the
// inverse's 1/L, on one element
template <typename T>
__device__ __forceinline__ void lk_apply(cpx<T>& v, T g, bool full, const InvLen<T>& inv)
{
    Kinetic<T>::apply(v, g, Kinetic<T>::one(), full, T(1));
    inv.apply(v);
}

// y-line sandwich on tz z-columns of one x: forward y, K1/2y (fields f < nh, Ch) or Ky
// (f >= nh, Cf), inverse y. grid (m0 * ceil(m2 / tz), nh + nf fields), block PC_NT. LDS as
// k_py.
template <typename T, int L = 0>
__global__ void __launch_bounds__(PC_NT) k_ly(PropArgs<T> a, int nh, int tz_arg)
{
    constexpr int CE = pc_ept<T, L>();
    extern __shared__ unsigned char smem_raw[];
    const int m1 = L > 0 ? L : a.m1, m2 = a.m2;
    const int tz = L > 0 ? pc_ct_t<T>(L) : tz_arg, P = pc_pitch(tz);
    cpx<T>* buf = reinterpret_cast<cpx<T>*>(smem_raw);
    cpx<T>* tw = buf + (size_t)m1 * P;
    const int tid = threadIdx.x;
    const int ntz = (m2 + tz - 1) / tz;
    const int bt = pc_xcd_block();
    const int x = bt / ntz, z0 = (bt - x * ntz) * tz;
    const int nz = m2 - z0 < tz ? m2 - z0 : tz;
    const bool full = (int)blockIdx.y >= nh;
    cpx<T>* g = pc_field(a, blockIdx.y, nh) + (size_t)x * m1 * m2 + z0;
    const int ne = m1 * tz;
    const InvLen<T> inv(m1);
    cpx<T> r[CE];
#pragma unroll
    for (int i = 0; i < CE; ++i) {
        const int e = tid + i * PC_NT;
        const int y = e / tz, zz = e - y * tz;
        r[i] = (e < ne && zz < nz) ? g[(size_t)y * m2 + zz] : cpx<T>{T(0), T(0)};
    }
    for (int i = tid; i < m1; i += PC_NT) tw[i] = a.tw1[i];
#pragma unroll
    for (int i = 0; i < CE; ++i) {
        const int e = tid + i * PC_NT;
        const int y = e / tz, zz = e - y * tz;
        if (e < ne) buf[y * P + zz] = r[i];
    }
    __syncthreads();
    pl_fft<T, false, L>(buf, m1, tz, P, tw, tid);
#pragma unroll
    for (int i = 0; i < CE; ++i) {
        const int e = tid + i * PC_NT;
        const int y = e / tz, zz = e - y * tz;
        if (e < ne) lk_apply(buf[y * P + zz], a.gy[y], full, inv);
    }
    __syncthreads();
    pl_fft<T, true, L>(buf, m1, tz, P, tw, tid);
#pragma unroll
    for (int i = 0; i < CE; ++i) {
        const int e = tid + i * PC_NT;
        const int y = e / tz, zz = e - y * tz;
        if (e < ne && zz < nz) g[(size_t)y * m2 + zz] = buf[y * P + zz];
    }
}

// z-transform of the first nz of the buffer's lines (pitch P)
template <typename T, bool INV, int L>
__device__ __forceinline__ void lz_fft(cpx<T>* buf, int m2, int nz, int P, const cpx<T>* tw, int tid)
{
    if constexpr (L > 0) {
        constexpr int NZ = pc_ct_nlz<T>(L);
        if (nz == 2 * NZ) fft_lines<T, L, INV, 2 * NZ, 1, pc_pitch(2 * NZ), PC_NT>(buf, tw, tid);
        else fft_lines<T, L, INV, NZ, 1, pc_pitch(2 * NZ), PC_NT>(buf, tw, tid);
    } else {
        pc_fft<T, INV>(buf, m2, nz, P, tw, tid);
    }
}

// z-line pass of step s: h <- K1/2z E1/2 K1/2z h (Ch), f <- Kz f (Cf), on nlz z lines.
// grid (ceil(m0 m1 / nlz), nb slots), block PC_NT. LDS as k_pz: lines [0, nlz) hold h,
// [nlz, 2 nlz) hold f, pitch pc_pitch(2 nlz), then m2 twiddles.
template <typename T, int L = 0>
__global__ void __launch_bounds__(PC_NT) k_lz(PropArgs<T> a, int s, int nlz_arg)
{
    constexpr int ZE = pc_ept<T, L>() / 2;
    extern __shared__ unsigned char smem_raw[];
    const int m2 = L > 0 ? L : a.m2;
    const int nlz = L > 0 ? pc_ct_nlz<T>(L) : nlz_arg;
    const int P = pc_pitch(2 * nlz);
    cpx<T>* buf = reinterpret_cast<cpx<T>*>(smem_raw);
    cpx<T>* tw = buf + (size_t)m2 * P;
    const int tid = threadIdx.x, slot = blockIdx.y;
    const long long l0 = (long long)blockIdx.x * nlz;
    const long long lrem = (long long)a.m0 * a.m1 - l0;
    const int nl = lrem < nlz ? (int)lrem : nlz;
    const size_t fo = (size_t)slot * (size_t)a.M + (size_t)l0 * m2;
    const int ne = nl * m2, nt = nlz * m2;
    const InvLen<T> inv(m2);
    T va[ZE], vb[ZE];
    cpx<T> c0[ZE], c1[ZE];
#pragma unroll
    for (int i = 0; i < ZE; ++i) {
        const int e = tid + i * PC_NT;
        const bool ok = e < ne;
        va[i] = ok ? a.e4A[fo + e] : T(0);
        vb[i] = ok ? a.e4B[fo + e] : T(0);
        c0[i] = ok ? a.Ch[fo + e] : cpx<T>{T(0), T(0)};
        c1[i] = ok ? a.Cf[fo + e] : cpx<T>{T(0), T(0)};
    }
    for (int i = tid; i < m2; i += PC_NT) tw[i] = a.tw2[i];
#pragma unroll
    for (int i = 0; i < ZE; ++i) {
        const int e = tid + i * PC_NT;
        if (e < nt) {
            const int l = e / m2, z = e - l * m2;
            buf[z * P + l] = c0[i];
            buf[z * P + nlz + l] = c1[i];
        }
    }
    __syncthreads();
    lz_fft<T, false, L>(buf, m2, 2 * nlz, P, tw, tid);
    // K1/2z on h, Kz on f (the factor depends on the spectral index z only)
#pragma unroll
    for (int i = 0; i < ZE; ++i) {
        const int e = tid + i * PC_NT;
        if (e < nt) {
            const int l = e / m2, z = e - l * m2;
            const T gz = a.gz[z];
            lk_apply(buf[z * P + l], gz, false, inv);
            lk_apply(buf[z * P + nlz + l], gz, true, inv);
        }
    }
    __syncthreads();
    lz_fft<T, true, L>(buf, m2, 2 * nlz, P, tw, tid);
    // E1/2 of step s on h (density.h:136-146)
    const bool qa = s < a.n1, pb = s < a.n2;
#pragma unroll
    for (int i = 0; i < ZE; ++i) {
        const int e = tid + i * PC_NT;
        if (e < ne) {
            const int l = e / m2, z = e - l * m2;
            cpx<T>& v = buf[z * P + l];
            v.x = Boltz<T>::mul2(v.x, qa ? va[i] : vb[i]);
            v.y = Boltz<T>::mul2(v.y, pb ? vb[i] : va[i]);
        }
    }
    __syncthreads();
    lz_fft<T, false, L>(buf, m2, nlz, P, tw, tid);
#pragma unroll
    for (int i = 0; i < ZE; ++i) {
        const int e = tid + i * PC_NT;
        if (e < nt) {
            const int l = e / m2, z = e - l * m2;
            lk_apply(buf[z * P + l], a.gz[z], false, inv);
        }
    }
    __syncthreads();
    lz_fft<T, true, L>(buf, m2, nlz, P, tw, tid);
#pragma unroll
    for (int i = 0; i < ZE; ++i) {
        const int e = tid + i * PC_NT;
        if (e < ne) {
            const int l = e / m2, z = e - l * m2;
            a.Ch[fo + e] = buf[z * P + l];
            a.Cf[fo + e] = buf[z * P + nlz + l];
        }
    }
}

// x-line pass on tx z-columns of one y of slot blockIdx.y. PZ_STEP / PZ_LAST complete step s
// (h <- K1/2x h, q(s+1) = (4 E1/4 h - E1/2 f) / 3 to the q store); PZ_BEGIN starts from
// q(0) = 1, PZ_RESUME from the stored q at index s (entry a.qsrc). Then, except for PZ_LAST,
// h = K1/2x E1/4 q, f = Kx E1/2 q for the next step. grid (m1 * ceil(m2 / tx), nb slots),
// block PC_NT. LDS as k_px.
template <typename T, int L = 0>
__global__ void __launch_bounds__(PC_NT) k_lx(PropArgs<T> a, int s, int mode, int tx_arg)
{
    constexpr int CE = pc_ept<T, L>();
    extern __shared__ unsigned char smem_raw[];
    const int m0 = L > 0 ? L : a.m0, m2 = a.m2;
    const int tx = L > 0 ? pc_ct_t<T>(L) : tx_arg, P = pc_pitch(tx);
    const size_t sx = (size_t)a.m1 * m2;
    cpx<T>* buf = reinterpret_cast<cpx<T>*>(smem_raw);
    cpx<T>* tw = buf + (size_t)m0 * P;
    const int tid = threadIdx.x, slot = blockIdx.y;
    const int ntz = (m2 + tx - 1) / tx;
    const int bt = pc_xcd_block();
    const int y = bt / ntz, z0 = (bt - y * ntz) * tx;
    const int nz = m2 - z0 < tx ? m2 - z0 : tx;
    const size_t so = (size_t)slot * (size_t)a.M, lo = (size_t)y * m2 + z0;  // slot, element (0, y, z0)
    const int ne = m0 * tx;
    const InvLen<T> inv(m0);
    auto ok = [&](int e) { return e < ne && e - (e / tx) * tx < nz; };
    // offset of tile element e = (x, zz) inside a field
    auto gidx = [&](int e) { return lo + (size_t)(e / tx) * sx + (size_t)(e - (e / tx) * tx); };
    const cpx<T>* Ch = a.Ch + so;
    const cpx<T>* Cf = a.Cf + so;
    const T* eA = a.e4A + so;
    const T* eB = a.e4B + so;
    T va[CE], vb[CE];
    cpx<T> q[CE];
    const bool fresh = mode == PZ_BEGIN || mode == PZ_RESUME;
#pragma unroll
    for (int i = 0; i < CE; ++i) {
        const int e = tid + i * PC_NT;
        const bool in = ok(e);
        va[i] = in ? eA[gidx(e)] : T(0);
        vb[i] = in ? eB[gidx(e)] : T(0);
    }
    for (int i = tid; i < m0; i += PC_NT) tw[i] = a.tw0[i];
    int sout;
    if (fresh) {
        const cpx<T>* qi = a.qs + ((size_t)slot * (size_t)a.qpitch + (size_t)a.qsrc) * (size_t)a.M;
#pragma unroll
        for (int i = 0; i < CE; ++i) {
            const int e = tid + i * PC_NT;
            q[i] = mode == PZ_BEGIN ? cpx<T>{T(1), T(1)} : (ok(e) ? qi[gidx(e)] : cpx<T>{T(0), T(0)});
        }
        sout = mode == PZ_BEGIN ? 0 : s;
    } else {
        cpx<T> fr[CE];
#pragma unroll
        for (int i = 0; i < CE; ++i) {
            const int e = tid + i * PC_NT;
            const bool in = ok(e);
            q[i] = in ? Ch[gidx(e)] : cpx<T>{T(0), T(0)};
            fr[i] = in ? Cf[gidx(e)] : cpx<T>{T(0), T(0)};
        }
#pragma unroll
        for (int i = 0; i < CE; ++i) {
            const int e = tid + i * PC_NT;
            const int x = e / tx, zz = e - x * tx;
            if (e < ne) buf[x * P + zz] = q[i];
        }
        __syncthreads();
        pl_fft<T, false, L>(buf, m0, tx, P, tw, tid);
#pragma unroll
        for (int i = 0; i < CE; ++i) {
            const int e = tid + i * PC_NT;
            const int x = e / tx, zz = e - x * tx;
            if (e < ne) lk_apply(buf[x * P + zz], a.gx[x], false, inv);
        }
        __syncthreads();
        pl_fft<T, true, L>(buf, m0, tx, P, tw, tid);
        const bool qa = s < a.n1, pb = s < a.n2;  // species of step s (density.h:136-146)
#pragma unroll
        for (int i = 0; i < CE; ++i) {
            const int e = tid + i * PC_NT;
            const int x = e / tx, zz = e - x * tx;
            const cpx<T> h = e < ne ? buf[x * P + zz] : cpx<T>{T(0), T(0)};
            q[i].x = Boltz<T>::rich(h.x, fr[i].x, qa ? va[i] : vb[i]);
            q[i].y = Boltz<T>::rich(h.y, fr[i].y, pb ? vb[i] : va[i]);
        }
        sout = s + 1;
        __syncthreads();  // buf is rewritten below
    }
    const int qix = q_index(a, sout);
    if (qix >= 0) {
        cpx<T>* qo = a.qs + ((size_t)slot * (size_t)a.qpitch + (size_t)qix) * (size_t)a.M;
#pragma unroll
        for (int i = 0; i < CE; ++i) {
            const int e = tid + i * PC_NT;
            if (ok(e)) store_q(&qo[gidx(e)], q[i]);
        }
    }
    if (mode == PZ_LAST) return;
    const bool na = sout < a.n1, nbb = sout < a.n2;  // next step's species
    cpx<T> fq[CE];
#pragma unroll
    for (int i = 0; i < CE; ++i) {
        const int e = tid + i * PC_NT;
        const int x = e / tx, zz = e - x * tx;
        const bool in = ok(e);
        const T ex = na ? va[i] : vb[i], ey = nbb ? vb[i] : va[i];
        fq[i] = in ? cpx<T>{Boltz<T>::mul2(q[i].x, ex), Boltz<T>::mul2(q[i].y, ey)} : cpx<T>{T(0), T(0)};
        if (e < ne)
            buf[x * P + zz] = in ? cpx<T>{Boltz<T>::mul4(q[i].x, ex), Boltz<T>::mul4(q[i].y, ey)}
                                 : cpx<T>{T(0), T(0)};
    }
    __syncthreads();
    pl_fft<T, false, L>(buf, m0, tx, P, tw, tid);
#pragma unroll
    for (int i = 0; i < CE; ++i) {
        const int e = tid + i * PC_NT;
        const int x = e / tx, zz = e - x * tx;
        if (e < ne) lk_apply(buf[x * P + zz], a.gx[x], false, inv);
    }
    __syncthreads();
    pl_fft<T, true, L>(buf, m0, tx, P, tw, tid);
#pragma unroll
    for (int i = 0; i < CE; ++i) {
        const int e = tid + i * PC_NT;
        const int x = e / tx, zz = e - x * tx;
        if (ok(e)) a.Ch[so + gidx(e)] = buf[x * P + zz];
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < CE; ++i) {
        const int e = tid + i * PC_NT;
        const int x = e / tx, zz = e - x * tx;
        if (e < ne) buf[x * P + zz] = fq[i];
    }
    __syncthreads();
    pl_fft<T, false, L>(buf, m0, tx, P, tw, tid);
#pragma unroll
    for (int i = 0; i < CE; ++i) {
        const int e = tid + i * PC_NT;
        const int x = e / tx, zz = e - x * tx;
        if (e < ne) lk_apply(buf[x * P + zz], a.gx[x], true, inv);
    }
    __syncthreads();
    pl_fft<T, true, L>(buf, m0, tx, P, tw, tid);
#pragma unroll
    for (int i = 0; i < CE; ++i) {
        const int e = tid + i * PC_NT;
        const int x = e / tx, zz = e - x * tx;
        if (ok(e)) a.Cf[so + gidx(e)] = buf[x * P + zz];
    }
}

}  // namespace scft
