// prop_pencil.hpp -- propagator kernels for any grid whose axes factor into 2, 3, 5: one
// 1-D line pass per axis, with lengths known only at run time.
//
// The plane kernels need a whole (y, z) plane in LDS. For the reference's shipped box
// (m = 200 x 150 x 150, input.dat) a plane is 180 KB in FP32 and 360 KB in FP64, so it
// does not fit in 160 KB. Here every pass holds a tile of whole lines instead:
//   k_pz  z lines (contiguous). Inverse z -> real-space Boltzmann / Richardson work -> forward z
//   k_py  y lines of one x and a tile of z (stride m2), forward or inverse
//   k_px  x lines of one y and a tile of z (stride m1 m2): forward, kinetic factor, inverse
// The field layout and the packed z = q + i q-dagger convention are those of the plane
// kernels, so the q store, the Simpson densities and every reduction are shared. A contour
// step s (density.h:129-220) is
//   px(Ch, Cf) -> py^-1(Ch) -> pz_mid -> py(Ch) -> px(Ch) -> py^-1(Ch, Cf) -> pz_end -> py(Ch, Cf)
// where between launches Ch and Cf hold fields transformed along y and z but not x.
//
// Line transforms are Stockham autosort passes (radix 8, 4, 2, 3, 5, chosen at run time by
// pick_radix) over lines stored element-major in LDS: element e of line l sits at
// buf[e * P + l] with an odd pitch P, so consecutive lanes take consecutive lines.
#pragma once
#include "prop_types.hpp"

namespace scft {

// Tile size: measured on MI355X against 512 x 4096, 256 x 1024 and 128 x 1024 (threads x
// elements). Larger tiles serialise more work behind each barrier; smaller ones cut
// the contiguous run of the strided passes.
constexpr int PC_NT = 256;     // threads per workgroup of every pencil kernel
constexpr int PC_EMAX = 2048;  // complex elements (lines x length) per workgroup transform
constexpr int PC_LMAX = 1024;  // longest axis

enum PencilMode { PZ_BEGIN = 0, PZ_STEP = 1, PZ_LAST = 2, PZ_MID = 3, PZ_RESUME = 4 };

__host__ __device__ constexpr int pc_bpt(int R) { return (PC_EMAX / R + PC_NT - 1) / PC_NT; }
__host__ __device__ constexpr int pc_pitch(int nlines) { return nlines + ((nlines & 1) ? 0 : 1); }

// One Stockham pass of radix R on nl lines of length L, NS = product of earlier radices.
// All reads to registers, barrier, all writes, barrier. Needs nl * L <= PC_EMAX.
template <typename T, int R, bool INV>
__device__ __forceinline__ void pc_pass(cpx<T>* buf, int L, int NS, int nl, int P, const cpx<T>* tw, int tid)
{
    constexpr int BPT = pc_bpt(R);
    const int NB = L / R, total = nl * NB, tws = L / (NS * R);
    cpx<T> v[BPT][R];
#pragma unroll
    for (int it = 0; it < BPT; ++it) {
        const int w = tid + it * PC_NT;
        if (w < total) {
            const int line = w % nl, j = w / nl;
#pragma unroll
            for (int r = 0; r < R; ++r) v[it][r] = buf[(j + r * NB) * P + line];
            if (NS > 1) {
                const int k = j % NS;
#pragma unroll
                for (int r = 1; r < R; ++r) {
                    const cpx<T> t = tw[r * k * tws];
                    v[it][r] = INV ? cmulc(v[it][r], t) : cmul(v[it][r], t);
                }
            }
            dft<R, INV, T>::run(v[it]);
        }
    }
    __syncthreads();
#pragma unroll
    for (int it = 0; it < BPT; ++it) {
        const int w = tid + it * PC_NT;
        if (w < total) {
            const int line = w % nl, j = w / nl;
            const int k = j % NS;
            const int idx = (j / NS) * NS * R + k;
#pragma unroll
            for (int r = 0; r < R; ++r) buf[(idx + r * NS) * P + line] = v[it][r];
        }
    }
    __syncthreads();
}

// Unnormalised DFT of nl lines of length L in place (forward sign -1, inverse +1). The
// caller has synchronised after filling buf; the transform ends with a barrier.
template <typename T, bool INV>
__device__ void pc_fft(cpx<T>* buf, int L, int nl, int P, const cpx<T>* tw, int tid)
{
    for (int NS = 1; NS < L;) {
        const int R = pick_radix(L / NS);
        switch (R) {
            case 8: pc_pass<T, 8, INV>(buf, L, NS, nl, P, tw, tid); break;
            case 4: pc_pass<T, 4, INV>(buf, L, NS, nl, P, tw, tid); break;
            case 2: pc_pass<T, 2, INV>(buf, L, NS, nl, P, tw, tid); break;
            case 3: pc_pass<T, 3, INV>(buf, L, NS, nl, P, tw, tid); break;
            default: pc_pass<T, 5, INV>(buf, L, NS, nl, P, tw, tid); break;
        }
        NS *= R;
    }
}

// Axis lengths with compile-time kernels: each line pass depends on its own axis length
// only, so a few lengths cover the large grids (the reference's 200 x 150 x 150 box, FP64
// 128^3, ...). Their tiles are fixed: nlz = E / 2 / L z lines, t = E / L z columns with
// E = pc_ct_emax<T>(L), and the transform is the compile-time Stockham of fft_engine.hpp
// (same radix plan and butterflies as pc_fft). Other lengths take the runtime kernels
// (L = 0).
#define PC_CT_LENGTHS(X) X(96) X(128) X(150) X(160) X(192) X(200) X(256)
__host__ __device__ constexpr bool pc_ct_len(int L)
{
    return L == 96 || L == 128 || L == 150 || L == 160 || L == 192 || L == 200 || L == 256;
}
// Elements per workgroup transform of the compile-time kernels: PC_EMAX, except FP64
// length 96, which takes 1536. Its tiles are then 16 columns (256 B, two cache lines) and
// 26 KB of LDS, so five workgroups share a CU where four did (96^3 FP64: 0.520 -> 0.474 ms
// per contour step). Measured on MI355X: 1536 on the other FP64 lengths cost 4-12% at 128
// and 256 (runs of 192 and 96 B) and moved 150 x 200 by +-3%; 4096 for FP32 (twice the
// columns and lines per tile) doubled the registers, cut occupancy from 7 to 3 waves per
// SIMD and slowed the input.dat box 25%.
template <typename T>
__host__ __device__ constexpr int pc_ct_emax(int L)
{
    return sizeof(T) == 8 && L == 96 ? 1536 : PC_EMAX;
}
template <typename T>
__host__ __device__ constexpr int pc_ct_nlz(int L) { return pc_ct_emax<T>(L) / 2 / L > 0 ? pc_ct_emax<T>(L) / 2 / L : 1; }
template <typename T>
__host__ __device__ constexpr int pc_ct_t(int L) { return pc_ct_emax<T>(L) / L > 0 ? pc_ct_emax<T>(L) / L : 1; }

// z-pass transform of nfl fields x nlz lines (runtime, or compile-time for L > 0)
template <typename T, bool INV, int L>
__device__ __forceinline__ void pz_fft(cpx<T>* buf, int m2, int nfl, int nlz, int P, const cpx<T>* tw, int tid)
{
    if constexpr (L > 0) {
        constexpr int NZ = pc_ct_nlz<T>(L);
        if (nfl == 2) fft_lines<T, L, INV, 2 * NZ, 1, pc_pitch(2 * NZ), PC_NT>(buf, tw, tid);
        else fft_lines<T, L, INV, NZ, 1, pc_pitch(NZ), PC_NT>(buf, tw, tid);
    } else {
        pc_fft<T, INV>(buf, m2, nfl * nlz, P, tw, tid);
    }
}

// y- or x-pass transform of nl lines of length len (compile-time for L > 0: nl = pc_ct_t(L))
template <typename T, bool INV, int L>
__device__ __forceinline__ void pl_fft(cpx<T>* buf, int len, int nl, int P, const cpx<T>* tw, int tid)
{
    if constexpr (L > 0) fft_lines<T, L, INV, pc_ct_t<T>(L), 1, pc_pitch(pc_ct_t<T>(L)), PC_NT>(buf, tw, tid);
    else pc_fft<T, INV>(buf, len, nl, P, tw, tid);
}

// Workgroups go to the 8 XCDs round-robin by linear id, and each XCD has its own L2. The
// y and x passes renumber blockIdx.x so that each XCD takes one contiguous run of tiles:
// adjacent z tiles share the cache lines at their boundary and then fetch them once, not
// once per tile. The map is a bijection on [0, gridDim.x); blocks of equal blockIdx.x % 8
// share an XCD for every blockIdx.y.
constexpr int PC_XCD = 8;

__device__ __forceinline__ int pc_xcd_block()
{
    const int G = gridDim.x, b = blockIdx.x;
    const int q = G / PC_XCD, r = G - q * PC_XCD, j = b % PC_XCD;
    return j * q + (j < r ? j : r) + b / PC_XCD;
}

// Element e of a workgroup's tile is handled by thread e % PC_NT in slot e / PC_NT of
// its registers. Every global load of a kernel goes to registers first, in fully unrolled
// loops, so all of a thread's loads are in flight together before the first LDS write.
constexpr int PZ_EPT = PC_EMAX / 2 / PC_NT;  // per field in k_pz (nlz m2 <= PC_EMAX / 2)
constexpr int PC_EPT = PC_EMAX / PC_NT;      // in k_py, k_px (lines x length <= PC_EMAX)
// the same for a kernel of compile-time length L (L = 0: the runtime plan, PC_EMAX)
template <typename T, int L>
__host__ __device__ constexpr int pc_ept()
{
    return (L > 0 ? pc_ct_emax<T>(L) : PC_EMAX) / PC_NT;
}

// z-line pass. grid (ceil(m0 m1 / nlz), nb slots), block PC_NT.
//   PZ_MID:   Ch <- Fz(E1/2 Fz^-1(Ch))                         (density.h:156-173)
//   PZ_STEP:  q(s+1) = (4 Fz^-1(Ch) E1/4 - Fz^-1(Cf) E1/2)/3   (GPU_kernels.h:199-203), stored;
//             then Ch <- Fz(E1/4 q), Cf <- Fz(E1/2 q) with step s+1's species
//   PZ_LAST:  as PZ_STEP without the forward part
//   PZ_BEGIN: q(0) = 1 stored, Ch <- Fz(E1/4), Cf <- Fz(E1/2) for step 0 (density.h:351-355)
//   PZ_RESUME: as PZ_BEGIN from the stored q at q-store index a.qsrc, contour index s
// q is stored only where the q store keeps its contour index (q_index).
// LDS: lines [0, nlz) hold Ch, [nlz, 2 nlz) hold Cf, pitch pc_pitch(2 nlz), then m2 twiddles.
template <typename T, int L = 0>
__global__ void __launch_bounds__(PC_NT) k_pz(PropArgs<T> a, int s, int mode, int nlz_arg)
{
    constexpr int ZE = pc_ept<T, L>() / 2;  // elements per thread and field
    extern __shared__ unsigned char smem_raw[];
    const int m2 = L > 0 ? L : a.m2;
    const int nlz = L > 0 ? pc_ct_nlz<T>(L) : nlz_arg;
    const int nfl = mode == PZ_MID ? 1 : 2;
    const bool fresh = mode == PZ_BEGIN || mode == PZ_RESUME;  // q not computed from Ch, Cf
    const int nlines = nfl * nlz, P = pc_pitch(nlines);
    cpx<T>* buf = reinterpret_cast<cpx<T>*>(smem_raw);
    cpx<T>* tw = buf + (size_t)m2 * P;
    const int tid = threadIdx.x, slot = blockIdx.y;
    const long long l0 = (long long)blockIdx.x * nlz;
    const long long lrem = (long long)a.m0 * a.m1 - l0;
    const int nl = lrem < nlz ? (int)lrem : nlz;  // valid lines of this workgroup
    const size_t fo = (size_t)slot * (size_t)a.M + (size_t)l0 * m2;
    const int ne = nl * m2, nt = nlz * m2;  // valid elements, tile elements

    T va[ZE], vb[ZE];
    cpx<T> c0[ZE], c1[ZE];
#pragma unroll
    for (int i = 0; i < ZE; ++i) {
        const int e = tid + i * PC_NT;
        const bool ok = e < ne;
        va[i] = ok ? a.e4A[fo + e] : T(0);
        vb[i] = ok ? a.e4B[fo + e] : T(0);
        c0[i] = (ok && !fresh) ? a.Ch[fo + e] : cpx<T>{T(0), T(0)};
        c1[i] = (ok && nfl == 2 && !fresh) ? a.Cf[fo + e] : cpx<T>{T(0), T(0)};
    }
    for (int i = tid; i < m2; i += PC_NT) tw[i] = a.tw2[i];

    if (!fresh) {
#pragma unroll
        for (int i = 0; i < ZE; ++i) {
            const int e = tid + i * PC_NT;
            if (e < nt) {
                const int l = e / m2, z = e - l * m2;
                buf[z * P + l] = c0[i];
                if (nfl == 2) buf[z * P + nlz + l] = c1[i];
            }
        }
        __syncthreads();
        pz_fft<T, true, L>(buf, m2, nfl, nlz, P, tw, tid);
    }

    if (mode == PZ_MID) {
        const bool qa = s < a.n1, pb = s < a.n2;  // species of this step (density.h:136-146)
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
        pz_fft<T, false, L>(buf, m2, nfl, nlz, P, tw, tid);
#pragma unroll
        for (int i = 0; i < ZE; ++i) {
            const int e = tid + i * PC_NT;
            if (e < ne) {
                const int l = e / m2, z = e - l * m2;
                a.Ch[fo + e] = buf[z * P + l];
            }
        }
        return;
    }

    const int sout = mode == PZ_BEGIN ? 0 : mode == PZ_RESUME ? s : s + 1;
    const int qix = q_index(a, sout);
    cpx<T>* qo = a.qs + ((size_t)slot * (size_t)a.qpitch + (size_t)(qix >= 0 ? qix : 0)) * (size_t)a.M + (size_t)l0 * m2;
    const cpx<T>* qi = a.qs + ((size_t)slot * (size_t)a.qpitch + (size_t)a.qsrc) * (size_t)a.M + (size_t)l0 * m2;
    const bool qa = s < a.n1, pb = s < a.n2;        // this step's species
    const bool na = sout < a.n1, nbb = sout < a.n2;  // next step's species
#pragma unroll
    for (int i = 0; i < ZE; ++i) {
        const int e = tid + i * PC_NT;
        if (e < nt) {
            const int l = e / m2, z = e - l * m2;
            if (e < ne) {
                cpx<T> q;
                if (mode == PZ_BEGIN) {
                    q = {T(1), T(1)};
                } else if (mode == PZ_RESUME) {
                    q = qi[e];
                } else {
                    const cpx<T> h0 = buf[z * P + l], h1 = buf[z * P + nlz + l];
                    q.x = Boltz<T>::rich(h0.x, h1.x, qa ? va[i] : vb[i]);
                    q.y = Boltz<T>::rich(h0.y, h1.y, pb ? vb[i] : va[i]);
                }
                if (qix >= 0) store_q(&qo[e], q);
                if (mode != PZ_LAST) {
                    const T ex = na ? va[i] : vb[i], ey = nbb ? vb[i] : va[i];
                    buf[z * P + l] = {Boltz<T>::mul4(q.x, ex), Boltz<T>::mul4(q.y, ey)};
                    buf[z * P + nlz + l] = {Boltz<T>::mul2(q.x, ex), Boltz<T>::mul2(q.y, ey)};
                }
            } else if (fresh) {
                buf[z * P + l] = {T(0), T(0)};
                buf[z * P + nlz + l] = {T(0), T(0)};
            }
        }
    }
    if (mode == PZ_LAST) return;
    __syncthreads();
    pz_fft<T, false, L>(buf, m2, nfl, nlz, P, tw, tid);
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

// Field f of a pencil launch: f < nh is Ch of slot f, else Cf of slot f - nh.
template <typename T>
__device__ __forceinline__ cpx<T>* pc_field(const PropArgs<T>& a, int f, int nh)
{
    return f < nh ? a.Ch + (size_t)f * (size_t)a.M : a.Cf + (size_t)(f - nh) * (size_t)a.M;
}

// y-line pass, forward or inverse, on tz z-columns of one x.
// grid (m0 * ceil(m2 / tz), nh + nf fields), block PC_NT. LDS: m1 x pc_pitch(tz), m1 twiddles.
template <typename T, bool INV, int L = 0>
__global__ void __launch_bounds__(PC_NT) k_py(PropArgs<T> a, int nh, int tz_arg)
{
    constexpr int CE = pc_ept<T, L>();  // elements per thread
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
    cpx<T>* g = pc_field(a, blockIdx.y, nh) + (size_t)x * m1 * m2 + z0;
    const int ne = m1 * tz;
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
    pl_fft<T, INV, L>(buf, m1, tz, P, tw, tid);
#pragma unroll
    for (int i = 0; i < CE; ++i) {
        const int e = tid + i * PC_NT;
        const int y = e / tz, zz = e - y * tz;
        if (e < ne && zz < nz) g[(size_t)y * m2 + zz] = buf[y * P + zz];
    }
}

// x-line pass on tx z-columns of one y: forward x, times exp(-k^2 ds/12)/M (Ch) or
// exp(-k^2 ds/6)/M (Cf), inverse x (density.h:149-156, 198-205; GPU_kernels.h:175-181, 207-213).
// grid (m1 * ceil(m2 / tx), nh + nf fields), block PC_NT. LDS: m0 x pc_pitch(tx), m0 twiddles.
template <typename T, int L = 0>
__global__ void __launch_bounds__(PC_NT) k_px(PropArgs<T> a, int nh, int tx_arg)
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
    const bool full = (int)blockIdx.y >= nh;
    cpx<T>* g = pc_field(a, blockIdx.y, nh) + (size_t)y * m2 + z0;
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
    const T gy = a.gy[y];
#pragma unroll
    for (int i = 0; i < CE; ++i) {
        const int e = tid + i * PC_NT;
        const int x = e / tx, zz = e - x * tx;
        if (e < ne && zz < nz)
            Kinetic<T>::apply(buf[x * P + zz], a.gx[x], Kinetic<T>::col(gy, a.gz[z0 + zz]), full, a.invM);
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

// ---------------------------------------------------------------- host-side tiling
struct PencilTiles {
    int nlz;  // z lines per k_pz workgroup (two fields each)
    int ty;   // z columns per k_py workgroup
    int tx;   // z columns per k_px workgroup
};

// Tiling for an m0 x m1 x m2 grid; every transform stays within PC_EMAX elements. With ct,
// axes of a compile-time length take that kernel's fixed tile.
template <typename T>
inline PencilTiles pencil_tiles(int m0, int m1, int m2, bool ct)
{
    auto tile = [](int len, int cap) {
        int t = PC_EMAX / len;
        if (t > cap) t = cap;
        if (t < 1) t = 1;
        const int n = (cap + t - 1) / t;  // tiles along z, then the smallest even split
        return (cap + n - 1) / n;
    };
    PencilTiles p;
    p.nlz = (ct && pc_ct_len(m2)) ? pc_ct_nlz<T>(m2) : PC_EMAX / 2 / m2;
    if (p.nlz < 1) p.nlz = 1;
    p.ty = (ct && pc_ct_len(m1)) ? pc_ct_t<T>(m1) : tile(m1, m2);
    p.tx = (ct && pc_ct_len(m0)) ? pc_ct_t<T>(m0) : tile(m0, m2);
    return p;
}

inline bool pencil_supported(int m0, int m1, int m2)
{
    const int m[3] = {m0, m1, m2};
    for (int i = 0; i < 3; ++i)
        if (m[i] < 1 || m[i] > PC_LMAX || !fft_len_supported(m[i])) return false;
    return true;
}

// Upper bound of pencil_lds_bytes over every supported grid: a tile of at most PC_EMAX
// elements plus one pad element per line (at most PC_LMAX lines) plus PC_LMAX twiddles.
// The kernels' maximum dynamic LDS is a per-function attribute shared by every context in
// the process, so setup reserves this bound rather than one grid's need.
template <typename T>
constexpr size_t pencil_lds_max()
{
    return (size_t)(PC_EMAX + 2 * PC_LMAX) * sizeof(cpx<T>);
}

// dynamic LDS bytes of the three kernels for one grid (what a launch passes)
template <typename T>
inline size_t pencil_lds_bytes(int m0, int m1, int m2, bool ct)
{
    const PencilTiles p = pencil_tiles<T>(m0, m1, m2, ct);
    const size_t z = (size_t)m2 * pc_pitch(2 * p.nlz) + m2;
    const size_t y = (size_t)m1 * pc_pitch(p.ty) + m1;
    const size_t x = (size_t)m0 * pc_pitch(p.tx) + m0;
    size_t b = z > y ? z : y;
    b = b > x ? b : x;
    return b * sizeof(cpx<T>);
}

}  // namespace scft
