// prop_engine.hip -- kernel instantiation table and the batched props driver.
#include "prop_engine.hpp"
#include "prop_kernels.hpp"
#include "prop_kernels2.hpp"
#include "prop_pencil.hpp"
#include "prop_sandwich.hpp"
#include "prop_lines2.hpp"
#include "prop_rlines.hpp"
#include "grad_kernels.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <type_traits>
#include <vector>

namespace scft {

// ---------------------------------------------------------------- per-grid kernel sets
// Two implementations per grid: the two-stage register FFT (prop_kernels2.hpp) when every
// axis splits as N1 x N2 (1 < N1 <= N2 <= 16) and fits in LDS, and the LDS Stockham
// kernels (prop_kernels.hpp) otherwise or when SCFT_LEGACY_FFT=1 is set.
template <typename T, int N0, int N1, int N2>
struct KS {
    static constexpr int NE = N1 * N2;
    static constexpr int NTP = NE <= 256 ? 64 : NE <= 1024 ? 128 : NE <= 4096 ? 256 : NE <= 8192 ? 512 : 1024;
    // (y, z) columns per workgroup of the column pass. Measured on MI355X per contour step:
    // 32 for x-length 48 and 96 (96^3 FP64 0.977 -> 0.800 ms, FP32 -5%, 48^3 -3%; 16 is no
    // better), 64 elsewhere (64^3 is 1.5-2.5% slower with 32, 128^3 unchanged).
    static constexpr int TC = (N0 == 48 || N0 == 96) ? 32 : 64;
    // the other width, for the two-stage column kernel (a schedule may pick it)
    static constexpr int TC_ALT = TC == 64 ? 32 : 64;
    static constexpr int NTC = 256;
    static constexpr size_t lds_plane = (size_t)(N1 * PlaneGeom<N1, N2>::P2 + N1 + N2) * sizeof(cpx<T>);
    static constexpr size_t lds_col = (size_t)(N0 * TC + N0) * sizeof(cpx<T>) + (size_t)N0 * sizeof(T);
    static constexpr int ntile = (N1 * N2 + TC - 1) / TC;
    static constexpr bool legacy_ok = lds_plane <= 160 * 1024 && lds_col <= 160 * 1024;
    using PG = P2Geom<N1, N2>;
    using CG = C2Geom<N0, TC>;
    static constexpr size_t lds2_plane = (size_t)(PG::LDSN + N1 + N2) * sizeof(cpx<T>);
    static constexpr size_t lds2_col = (size_t)(CG::LDSN + N0) * sizeof(cpx<T>) + (size_t)N0 * sizeof(T);
    static constexpr bool two_ok = PG::ok && CG::ok && lds2_plane <= 160 * 1024 && lds2_col <= 160 * 1024;
    // planes whose LDS allows only one workgroup per CU and that run fewer than 1024
    // threads (96^2 FP32) use the persistent plane_mid (k_plane_mid2p)
    static constexpr bool persist_ok = two_ok && lds2_plane > 80 * 1024 && PG::NT < 1024;
    using CGA = C2Geom<N0, TC_ALT>;
    static constexpr size_t lds2_col_alt = (size_t)(CGA::LDSN + N0) * sizeof(cpx<T>) + (size_t)N0 * sizeof(T);
    static constexpr bool alt_ok = two_ok && CGA::ok && lds2_col_alt <= 160 * 1024;
    static constexpr int ntile_alt = (N1 * N2 + TC_ALT - 1) / TC_ALT;
    // the other plane workgroup size, for planes of 512 or 768 threads (64^2, 96^2): half as
    // many threads, each with twice the items
    static constexpr int NT_ALT = (PG::NT >= 512 && PG::NT < 1024) ? PG::NT / 2 : 0;
    static constexpr bool palt_ok = two_ok && NT_ALT > 0;
    // the two-pass step (prop_sandwich.hpp): its plane pass holds the (y, z) kinetic tables
    // in LDS beside the two-stage plane's buffers
    static constexpr size_t lds3_plane = lds2_plane + (size_t)(N1 + N2) * sizeof(T);
    static constexpr size_t lds3_col = scol_lds_bytes<T, N0, TC>();
    static constexpr size_t lds3_col_alt = scol_lds_bytes<T, N0, TC_ALT>();
    static constexpr bool tp_ok = two_ok && lds3_plane <= 160 * 1024 && lds3_col <= 160 * 1024;
    static constexpr bool tp_alt_ok = tp_ok && alt_ok && lds3_col_alt <= 160 * 1024;
    static_assert(legacy_ok || two_ok, "no kernel fits in LDS for this grid");
    static_assert(fft_len_supported(N0) && fft_len_supported(N1) && fft_len_supported(N2), "radix");

    static hipError_t setup_legacy()
    {
        if constexpr (legacy_ok) {
            hipError_t e;
            e = hipFuncSetAttribute((const void*)k_plane_end<T, N1, N2, NTP>,
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_plane);
            if (e != hipSuccess) return e;
            e = hipFuncSetAttribute((const void*)k_plane_mid<T, N1, N2, NTP>,
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_plane);
            if (e != hipSuccess) return e;
            return hipFuncSetAttribute((const void*)k_column<T, N0, N1, N2, TC, NTC>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_col);
        }
        return hipErrorNotSupported;
    }
    static void begin_legacy(const PropArgs<T>& a, hipStream_t st)
    {
        if constexpr (legacy_ok)
            hipLaunchKernelGGL((k_plane_end<T, N1, N2, NTP>), dim3(N0, a.nb), dim3(NTP), lds_plane, st, a, 0,
                               (int)PL_BEGIN);
    }
    static void step_legacy(const PropArgs<T>& a, int s, hipStream_t st)
    {
        if constexpr (legacy_ok) {
            hipLaunchKernelGGL((k_column<T, N0, N1, N2, TC, NTC>), dim3(ntile, 2 * a.nb), dim3(NTC), lds_col, st,
                               a, a.nb, a.nb);
            hipLaunchKernelGGL((k_plane_mid<T, N1, N2, NTP>), dim3(N0, a.nb), dim3(NTP), lds_plane, st, a, s);
            hipLaunchKernelGGL((k_column<T, N0, N1, N2, TC, NTC>), dim3(ntile, a.nb), dim3(NTC), lds_col, st, a,
                               a.nb, 0);
            hipLaunchKernelGGL((k_plane_end<T, N1, N2, NTP>), dim3(N0, a.nb), dim3(NTP), lds_plane, st, a, s,
                               (int)(s == a.N - 1 ? PL_LAST : PL_STEP));
        }
    }
    static void resume_legacy(const PropArgs<T>& a, int c, hipStream_t st)
    {
        if constexpr (legacy_ok)
            hipLaunchKernelGGL((k_plane_end<T, N1, N2, NTP>), dim3(N0, a.nb), dim3(NTP), lds_plane, st, a, c,
                               (int)PL_RESUME);
    }
    static hipError_t setup_two()
    {
        if constexpr (two_ok) {
            hipError_t e;
            e = hipFuncSetAttribute((const void*)k_plane_end2<T, N1, N2>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                    (int)lds2_plane);
            if (e != hipSuccess) return e;
            e = hipFuncSetAttribute((const void*)k_plane_mid2<T, N1, N2>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                    (int)lds2_plane);
            if (e != hipSuccess) return e;
            if constexpr (persist_ok) {
                e = hipFuncSetAttribute((const void*)k_plane_mid2p<T, N1, N2>,
                                        hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds2_plane);
                if (e != hipSuccess) return e;
            }
            if constexpr (alt_ok) {
                e = hipFuncSetAttribute((const void*)k_column2<T, N0, N1, N2, TC_ALT>,
                                        hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds2_col_alt);
                if (e != hipSuccess) return e;
            }
            if constexpr (palt_ok) {
                e = hipFuncSetAttribute((const void*)k_plane_end2<T, N1, N2, NT_ALT>,
                                        hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds2_plane);
                if (e != hipSuccess) return e;
                e = hipFuncSetAttribute((const void*)k_plane_mid2<T, N1, N2, NT_ALT>,
                                        hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds2_plane);
                if (e != hipSuccess) return e;
            }
            return hipFuncSetAttribute((const void*)k_column2<T, N0, N1, N2, TC>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds2_col);
        }
        return hipErrorNotSupported;
    }
    // column pass over rows fields (Ch of every slot, then nf Cf), at the tile width a.coltile
    // asks for (0: TC)
    static void column_two(const PropArgs<T>& a, int rows, int nf, hipStream_t st)
    {
        if constexpr (alt_ok) {
            if (a.coltile == TC_ALT) {
                hipLaunchKernelGGL((k_column2<T, N0, N1, N2, TC_ALT>), dim3(ntile_alt, rows), dim3(CGA::NT),
                                   lds2_col_alt, st, a, a.nb, nf);
                return;
            }
        }
        hipLaunchKernelGGL((k_column2<T, N0, N1, N2, TC>), dim3(ntile, rows), dim3(CG::NT), lds2_col, st, a, a.nb, nf);
    }
    // plane_end / plane_mid at the workgroup size a.planent asks for (0: PG::NT)
    static void plane_end_two(const PropArgs<T>& a, int s, int mode, hipStream_t st)
    {
        if constexpr (palt_ok) {
            if (a.planent == NT_ALT) {
                hipLaunchKernelGGL((k_plane_end2<T, N1, N2, NT_ALT>), dim3(N0, a.nb), dim3(NT_ALT), lds2_plane, st,
                                   a, s, mode);
                return;
            }
        }
        hipLaunchKernelGGL((k_plane_end2<T, N1, N2>), dim3(N0, a.nb), dim3(PG::NT), lds2_plane, st, a, s, mode);
    }
    static void plane_mid_two(const PropArgs<T>& a, int s, hipStream_t st)
    {
        if constexpr (palt_ok) {
            if (a.planent == NT_ALT) {
                hipLaunchKernelGGL((k_plane_mid2<T, N1, N2, NT_ALT>), dim3(N0, a.nb), dim3(NT_ALT), lds2_plane, st,
                                   a, s);
                return;
            }
        }
        if constexpr (persist_ok) {
            if (a.pgrid > 0) {
                const int np = N0 * a.nb;
                hipLaunchKernelGGL((k_plane_mid2p<T, N1, N2>), dim3(std::min(a.pgrid, np)), dim3(PG::NT), lds2_plane,
                                   st, a, s, np);
                return;
            }
        }
        hipLaunchKernelGGL((k_plane_mid2<T, N1, N2>), dim3(N0, a.nb), dim3(PG::NT), lds2_plane, st, a, s);
    }
    static void begin_two(const PropArgs<T>& a, hipStream_t st)
    {
        if constexpr (two_ok) plane_end_two(a, 0, (int)PL_BEGIN, st);
    }
    static void step_two(const PropArgs<T>& a, int s, hipStream_t st)
    {
        if constexpr (two_ok) {
            column_two(a, 2 * a.nb, a.nb, st);
            plane_mid_two(a, s, st);
            column_two(a, a.nb, 0, st);
            plane_end_two(a, s, (int)(s == a.N - 1 ? PL_LAST : PL_STEP), st);
        }
    }
    static void resume_two(const PropArgs<T>& a, int c, hipStream_t st)
    {
        if constexpr (two_ok) plane_end_two(a, c, (int)PL_RESUME, st);
    }
    static hipError_t setup_tp()
    {
        if constexpr (tp_ok) {
            hipError_t e = hipFuncSetAttribute((const void*)k_splane2<T, N1, N2>,
                                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds3_plane);
            if (e != hipSuccess) return e;
            if constexpr (tp_alt_ok) {
                e = hipFuncSetAttribute((const void*)k_scol2<T, N0, N1, N2, TC_ALT>,
                                        hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds3_col_alt);
                if (e != hipSuccess) return e;
            }
            return hipFuncSetAttribute((const void*)k_scol2<T, N0, N1, N2, TC>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds3_col);
        }
        return hipErrorNotSupported;
    }
    // column pass of the two-pass step at the tile width a.coltile asks for (0: TC)
    static void scol_tp(const PropArgs<T>& a, int s, int mode, hipStream_t st)
    {
        if constexpr (tp_alt_ok) {
            if (a.coltile == TC_ALT) {
                hipLaunchKernelGGL((k_scol2<T, N0, N1, N2, TC_ALT>), dim3(ntile_alt, a.nb), dim3(CGA::NT), lds3_col_alt,
                                   st, a, s, mode);
                return;
            }
        }
        hipLaunchKernelGGL((k_scol2<T, N0, N1, N2, TC>), dim3(ntile, a.nb), dim3(CG::NT), lds3_col, st, a, s, mode);
    }
    static void begin_tp(const PropArgs<T>& a, hipStream_t st)
    {
        if constexpr (tp_ok) scol_tp(a, 0, (int)PL_BEGIN, st);
    }
    static void step_tp(const PropArgs<T>& a, int s, hipStream_t st)
    {
        if constexpr (tp_ok) {
            hipLaunchKernelGGL((k_splane2<T, N1, N2>), dim3(N0, a.nb), dim3(PG::NT), lds3_plane, st, a, s);
            scol_tp(a, s, (int)(s == a.N - 1 ? PL_LAST : PL_STEP), st);
        }
    }
    static void resume_tp(const PropArgs<T>& a, int c, hipStream_t st)
    {
        if constexpr (tp_ok) scol_tp(a, c, (int)PL_RESUME, st);
    }
    static hipError_t gsetup_two()
    {
        if constexpr (two_ok) {
            hipError_t e = hipFuncSetAttribute((const void*)k_gplane_init2<T, N1, N2>,
                                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds2_plane);
            if (e != hipSuccess) return e;
            e = hipFuncSetAttribute((const void*)k_gcol2_acc<T, N0, N1, N2, TC>,
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds2_col);
            if (e != hipSuccess) return e;
            return hipFuncSetAttribute((const void*)k_gcol2<T, N0, N1, N2, TC>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds2_col);
        }
        return hipErrorNotSupported;
    }
    // the passes of one gradient term, or of a term and its mirror (specB >= 0), over a.nb
    // slots: Gh, Gf <- F2(E1/4 Z), F2(E1/2 Z); x: K0, Kf <- spectra, Gh <- K12 sandwich;
    // Gh <- F2(E1/2 F2^-1 Gh) (plane_mid with both parts on the term's species, once per
    // term: the factors are the replica's); x: K1 spectrum paired with K0, Kf and summed
    static void gterm_two(const PropArgs<T>& a, int ib, int pbk, int ia, int pak, int specA, int specB,
                          cpx<T>* Gh, cpx<T>* Gf, cpx<T>* K0, cpx<T>* Kf, const GradConst& gc, double* part,
                          hipStream_t st)
    {
        if constexpr (two_ok) {
            const int nb = a.nb, nterm = specB >= 0 ? 2 : 1, nt = nterm * nb;
            PropArgs<T> w = a;
            w.Ch = Gh;
            w.Cf = Gf;
            hipLaunchKernelGGL((k_gplane_init2<T, N1, N2>), dim3(N0, nb), dim3(PG::NT), lds2_plane, st, w, ib, pbk, ia,
                               pak, specA, specB);
            hipLaunchKernelGGL((k_gcol2<T, N0, N1, N2, TC>), dim3(ntile, 2 * nt), dim3(CG::NT), lds2_col, st, w, nt,
                               K0, Kf);
            for (int t = 0; t < nterm; ++t) {
                PropArgs<T> b = a;  // k_plane_mid2 takes the real part's species from s < n1, the
                b.Ch = Gh + (size_t)t * nb * a.M;  // imaginary part's from s < n2: at s = 0 both
                const int spec = t ? specB : specA;  // get spec
                b.n1 = spec == 0 ? 1 : 0;
                b.n2 = spec == 0 ? 0 : 1;
                hipLaunchKernelGGL((k_plane_mid2<T, N1, N2>), dim3(N0, nb), dim3(PG::NT), lds2_plane, st, b, 0);
            }
            hipLaunchKernelGGL((k_gcol2_acc<T, N0, N1, N2, TC>), dim3(ntile, nt), dim3(CG::NT), lds2_col, st, w,
                               (const cpx<T>*)K0, (const cpx<T>*)Kf, gc, part);
        }
    }
    static KernelSet<T> make()
    {
        KernelSet<T> k{N0,      N1,      N2,      &setup_legacy, &begin_legacy,  &step_legacy,
                       nullptr, nullptr, nullptr, "stockham",    &resume_legacy, nullptr,
                       lds_plane > 80 * 1024, TC, alt_ok ? TC_ALT : 0, persist_ok, nullptr, nullptr};
        k.plane_nt = PG::NT;
        k.plane_nt_alt = palt_ok ? NT_ALT : 0;
        // two-pass default tile: 32 columns wherever that kernel exists. The autotuner picks it
        // on every grid measured: 128^3 FP32 x 8, 0.369 against 0.404 ms per step at one
        // stream; 64^3 FP32 x 1, 0.0250 against 0.0265; 64^3 FP64 x 1, 0.0340 against 0.0376;
        // 32^3 FP32 x 1, 0.0166 against 0.0173; the 14-replica bench string at 3 streams
        k.tc3 = tp_alt_ok ? 32 : TC;
        k.tc_alt3 = tp_alt_ok ? (k.tc3 == TC ? TC_ALT : TC) : 0;
        if (tp_ok) {
            k.setup3 = &setup_tp;
            k.begin3 = &begin_tp;
            k.step3 = &step_tp;
            k.resume3 = &resume_tp;
        }
        if (two_ok) {
            k.setup2 = &setup_two;
            k.begin2 = &begin_two;
            k.step2 = &step_two;
            k.resume2 = &resume_two;
            k.gsetup2 = &gsetup_two;
            k.gterm2 = &gterm_two;
            k.gtiles = ntile;
        }
        if (!legacy_ok) {
            k.setup = &setup_two;
            k.begin = &begin_two;
            k.step = &step_two;
            k.resume = &resume_two;
        }
        return k;
    }
};

#define SCFT_GRID_COMMON(T)                                                                         \
    KS<T, 8, 8, 8>::make(), KS<T, 12, 12, 12>::make(), KS<T, 16, 16, 16>::make(),                   \
        KS<T, 24, 24, 24>::make(), KS<T, 32, 32, 32>::make(), KS<T, 48, 48, 48>::make(),            \
        KS<T, 64, 64, 64>::make(), KS<T, 96, 96, 96>::make(), KS<T, 12, 8, 16>::make(),             \
        KS<T, 8, 12, 10>::make(), KS<T, 16, 12, 20>::make()

template <>
const KernelSet<float>* find_kernels<float>(int m0, int m1, int m2)
{
    static const KernelSet<float> table[] = {SCFT_GRID_COMMON(float), KS<float, 128, 128, 128>::make()};
    for (const auto& k : table)
        if (k.m0 == m0 && k.m1 == m1 && k.m2 == m2) return &k;
    return nullptr;
}

template <>
const KernelSet<double>* find_kernels<double>(int m0, int m1, int m2)
{
    static const KernelSet<double> table[] = {SCFT_GRID_COMMON(double)};
    for (const auto& k : table)
        if (k.m0 == m0 && k.m1 == m1 && k.m2 == m2) return &k;
    return nullptr;
}

std::string supported_grids(bool fp64)
{
    std::string s = "plane/column kernels for 8^3 12^3 16^3 24^3 32^3 48^3 64^3 96^3 12x8x16 8x12x10 16x12x20";
    if (!fp64) s += " 128^3";
    s += "; line kernels for any grid whose axes are 1.." + std::to_string(PC_LMAX) + " and factor into 2, 3, 5";
    return s;
}

// ---------------------------------------------------------------- pencil (line) kernels
template <typename T>
static hipError_t pencil_setup()
{
    const int lds = (int)pencil_lds_max<T>();  // shared by every context: never one grid's need
    std::vector<const void*> fns = {(const void*)k_pz<T, 0>, (const void*)k_py<T, false, 0>,
                                    (const void*)k_py<T, true, 0>, (const void*)k_px<T, 0>};
#define PC_FNS(L)                                                                                                 \
    fns.push_back((const void*)k_pz<T, L>);                                                                       \
    fns.push_back((const void*)k_py<T, false, L>);                                                                \
    fns.push_back((const void*)k_py<T, true, L>);                                                                 \
    fns.push_back((const void*)k_px<T, L>);
    PC_CT_LENGTHS(PC_FNS)
#undef PC_FNS
    for (const void* f : fns) {
        hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, lds);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

// Each pass launches the kernel compiled for its axis length when there is one.
template <typename T>
static void pencil_pz(const PropArgs<T>& a, int s, int mode, const PencilTiles& p, size_t lds, hipStream_t st)
{
    const long long lines = (long long)a.m0 * a.m1;
    const dim3 g((unsigned)((lines + p.nlz - 1) / p.nlz), a.nb);
    if (a.pencil_ct) switch (a.m2) {
#define PC_CASE(L) \
    case L: hipLaunchKernelGGL((k_pz<T, L>), g, dim3(PC_NT), lds, st, a, s, mode, p.nlz); return;
            PC_CT_LENGTHS(PC_CASE)
#undef PC_CASE
            default: break;
        }
    hipLaunchKernelGGL((k_pz<T, 0>), g, dim3(PC_NT), lds, st, a, s, mode, p.nlz);
}

// y pass over nfields fields (the first nh are Ch)
template <typename T, bool INV>
static void pencil_py(const PropArgs<T>& a, int nh, int nfields, const PencilTiles& p, size_t lds, hipStream_t st)
{
    const int ntz = (a.m2 + p.ty - 1) / p.ty;
    const dim3 g(a.m0 * ntz, nfields);
    if (a.pencil_ct) switch (a.m1) {
#define PC_CASE(L) \
    case L: hipLaunchKernelGGL((k_py<T, INV, L>), g, dim3(PC_NT), lds, st, a, nh, p.ty); return;
            PC_CT_LENGTHS(PC_CASE)
#undef PC_CASE
            default: break;
        }
    hipLaunchKernelGGL((k_py<T, INV, 0>), g, dim3(PC_NT), lds, st, a, nh, p.ty);
}

template <typename T>
static void pencil_px(const PropArgs<T>& a, int nh, int nfields, const PencilTiles& p, size_t lds, hipStream_t st)
{
    const int ntz = (a.m2 + p.tx - 1) / p.tx;
    const dim3 g(a.m1 * ntz, nfields);
    if (a.pencil_ct) switch (a.m0) {
#define PC_CASE(L) \
    case L: hipLaunchKernelGGL((k_px<T, L>), g, dim3(PC_NT), lds, st, a, nh, p.tx); return;
            PC_CT_LENGTHS(PC_CASE)
#undef PC_CASE
            default: break;
        }
    hipLaunchKernelGGL((k_px<T, 0>), g, dim3(PC_NT), lds, st, a, nh, p.tx);
}

template <typename T>
static void pencil_begin(const PropArgs<T>& a, hipStream_t st)
{
    const PencilTiles p = pencil_tiles<T>(a.m0, a.m1, a.m2, a.pencil_ct);
    const size_t lds = pencil_lds_bytes<T>(a.m0, a.m1, a.m2, a.pencil_ct);
    pencil_pz<T>(a, 0, PZ_BEGIN, p, lds, st);
    pencil_py<T, false>(a, a.nb, 2 * a.nb, p, lds, st);
}

template <typename T>
static void pencil_resume(const PropArgs<T>& a, int c, hipStream_t st)
{
    const PencilTiles p = pencil_tiles<T>(a.m0, a.m1, a.m2, a.pencil_ct);
    const size_t lds = pencil_lds_bytes<T>(a.m0, a.m1, a.m2, a.pencil_ct);
    pencil_pz<T>(a, c, PZ_RESUME, p, lds, st);
    pencil_py<T, false>(a, a.nb, 2 * a.nb, p, lds, st);
}

template <typename T>
static void pencil_step(const PropArgs<T>& a, int s, hipStream_t st)
{
    const PencilTiles p = pencil_tiles<T>(a.m0, a.m1, a.m2, a.pencil_ct);
    const size_t lds = pencil_lds_bytes<T>(a.m0, a.m1, a.m2, a.pencil_ct);
    const bool last = s == a.N - 1;
    pencil_px<T>(a, a.nb, 2 * a.nb, p, lds, st);
    pencil_py<T, true>(a, a.nb, a.nb, p, lds, st);
    pencil_pz<T>(a, s, PZ_MID, p, lds, st);
    pencil_py<T, false>(a, a.nb, a.nb, p, lds, st);
    pencil_px<T>(a, a.nb, a.nb, p, lds, st);
    pencil_py<T, true>(a, a.nb, 2 * a.nb, p, lds, st);
    pencil_pz<T>(a, s, last ? PZ_LAST : PZ_STEP, p, lds, st);
    if (!last) pencil_py<T, false>(a, a.nb, 2 * a.nb, p, lds, st);
}

// ---------------------------------------------------------------- PropEngine
template <typename T>
PropEngine<T>::~PropEngine()
{
    void* ptrs[] = {Ch_,   Cf_,   e4A_,  e4B_,  qs_, tw_[0], tw_[1], tw_[2], g_[0],
                    g_[1], g_[2], part_, cAB_, scal_, K0_, K1_,    Kf_,    gpart_, rn_, Gh_, Gf_};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    for (hipEvent_t e : ev_)
        if (e) (void)hipEventDestroy(e);
    clear_graphs();
    if (fork_) (void)hipEventDestroy(fork_);
    for (int i = 0; i + 1 < kMaxStreams; ++i) {
        if (join_[i]) (void)hipEventDestroy(join_[i]);
        if (sx_[i]) (void)hipStreamDestroy(sx_[i]);
    }
}

// Extra streams (default priority) and join events for n replica-group streams, created when
// first needed. Streams take hardware queues in creation order, so an engine creates no
// more than it runs: with several contexts on one GPU, unused streams would put every
// context's working streams on the same few queues and serialise them.
template <typename T>
hipError_t PropEngine<T>::ensure_streams(int n)
{
    for (int i = 0; i + 1 < n && i + 1 < kMaxStreams; ++i) {
        hipError_t e = hipSuccess;
        if (!sx_[i]) e = hipStreamCreateWithFlags(&sx_[i], hipStreamNonBlocking);
        if (e == hipSuccess && !join_[i]) e = hipEventCreateWithFlags(&join_[i], hipEventDisableTiming);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

template <typename T>
hipError_t PropEngine<T>::enable_timing(bool on)
{
    timing_ = on;
    for (int i = 0; i < 2 && on; ++i)
        if (!ev_[i]) {
            hipError_t e = hipEventCreate(&ev_[i]);
            if (e != hipSuccess) return e;
        }
    return hipSuccess;
}

template <typename T>
double PropEngine<T>::loop_ms() const
{
    float ms = 0;
    if (!ev_[0] || !ev_[1] || hipEventElapsedTime(&ms, ev_[0], ev_[1]) != hipSuccess) return -1.0;
    return (double)ms;
}

// ---------------------------------------------------------------- split line step
// (prop_lines2.hpp): k_ly (h, f), k_lz, k_ly (h), k_lx per contour step
template <typename T>
static hipError_t lines_setup()
{
    const int lds = (int)pencil_lds_max<T>();
    std::vector<const void*> fns = {(const void*)k_ly<T, 0>, (const void*)k_lz<T, 0>, (const void*)k_lx<T, 0>};
#define PC_FNS(L)                                                                                                 \
    fns.push_back((const void*)k_ly<T, L>);                                                                       \
    fns.push_back((const void*)k_lz<T, L>);                                                                       \
    fns.push_back((const void*)k_lx<T, L>);
    PC_CT_LENGTHS(PC_FNS)
#undef PC_FNS
    for (const void* f : fns) {
        hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, lds);
        if (e != hipSuccess) return e;
    }
    hipError_t e = hipSuccess;
#define RL_ATTR(L)                                                                                               \
    static_assert(C2Geom<L, rl_ty<T>(L)>::ok && C2Geom<L, rl_tz<T>(L)>::ok, "register line pass " #L);        \
    static_assert(rl_lds_bytes<T, L, rl_ty<T>(L)>() <= 160 * 1024, "k_ry LDS " #L);                            \
    static_assert(rl_lds_bytes<T, L, rl_tz<T>(L)>() <= 160 * 1024, "k_rz LDS " #L);                            \
    if (e == hipSuccess)                                                                                          \
        e = hipFuncSetAttribute((const void*)k_ry<T, L>, hipFuncAttributeMaxDynamicSharedMemorySize,              \
                                (int)rl_lds_bytes<T, L, rl_ty<T>(L)>());                                           \
    if (e == hipSuccess)                                                                                          \
        e = hipFuncSetAttribute((const void*)k_rz<T, L>, hipFuncAttributeMaxDynamicSharedMemorySize,              \
                                (int)rl_lds_bytes<T, L, rl_tz<T>(L)>());
    RL_LENGTHS(RL_ATTR)
#undef RL_ATTR
#define RX_ATTR(L)                                                                                               \
    if constexpr (rx_ok<T>(L)) {                                                                                  \
        static_assert(C2Geom<L, rl_tx<T>(L), rl_x1(L)>::ok, "x pass " #L);                                      \
        static_assert(scol_lds_bytes<T, L, rl_tx<T>(L), rl_x1(L)>() <= 160 * 1024, "x pass LDS " #L);          \
        if (e == hipSuccess)                                                                                      \
            e = hipFuncSetAttribute((const void*)k_scol2<T, L, 0, 0, rl_tx<T>(L), rl_x1(L)>,                       \
                                    hipFuncAttributeMaxDynamicSharedMemorySize,                                   \
                                    (int)scol_lds_bytes<T, L, rl_tx<T>(L), rl_x1(L)>());                           \
    }
    RX_LENGTHS(RX_ATTR)
#undef RX_ATTR
    return e;
}

template <typename T>
static void lines_ly(const PropArgs<T>& a, int nh, int nfields, const PencilTiles& p, size_t lds, hipStream_t st)
{
    if (a.rlines) switch (a.m1) {
#define RL_CASE(L)                                                                                                \
    case L: {                                                                                                     \
        constexpr int TY = rl_ty<T>(L);                                                                            \
        using RG = C2Geom<L, TY>;                                                                                 \
        hipLaunchKernelGGL((k_ry<T, L>), dim3(a.m0 * ((a.m2 + TY - 1) / TY), nfields), dim3(RG::NT),             \
                           (rl_lds_bytes<T, L, TY>()), st, a, nh);                                                \
        return;                                                                                                   \
    }
            RL_LENGTHS(RL_CASE)
#undef RL_CASE
            default: break;
        }
    const int ntz = (a.m2 + p.ty - 1) / p.ty;
    const dim3 g(a.m0 * ntz, nfields);
    if (a.pencil_ct) switch (a.m1) {
#define PC_CASE(L) \
    case L: hipLaunchKernelGGL((k_ly<T, L>), g, dim3(PC_NT), lds, st, a, nh, p.ty); return;
            PC_CT_LENGTHS(PC_CASE)
#undef PC_CASE
            default: break;
        }
    hipLaunchKernelGGL((k_ly<T, 0>), g, dim3(PC_NT), lds, st, a, nh, p.ty);
}

template <typename T>
static void lines_lz(const PropArgs<T>& a, int s, const PencilTiles& p, size_t lds, hipStream_t st)
{
    const long long lines = (long long)a.m0 * a.m1;
    if (a.rlines) switch (a.m2) {
#define RL_CASE(L)                                                                                                \
    case L: {                                                                                                     \
        constexpr int TZ = rl_tz<T>(L);                                                                            \
        using RG = C2Geom<L, TZ>;                                                                                 \
        hipLaunchKernelGGL((k_rz<T, L>), dim3((unsigned)((lines + TZ - 1) / TZ), a.nb), dim3(RG::NT),            \
                           (rl_lds_bytes<T, L, TZ>()), st, a, s);                                                 \
        return;                                                                                                   \
    }
            RL_LENGTHS(RL_CASE)
#undef RL_CASE
            default: break;
        }
    const dim3 g((unsigned)((lines + p.nlz - 1) / p.nlz), a.nb);
    if (a.pencil_ct) switch (a.m2) {
#define PC_CASE(L) \
    case L: hipLaunchKernelGGL((k_lz<T, L>), g, dim3(PC_NT), lds, st, a, s, p.nlz); return;
            PC_CT_LENGTHS(PC_CASE)
#undef PC_CASE
            default: break;
        }
    hipLaunchKernelGGL((k_lz<T, 0>), g, dim3(PC_NT), lds, st, a, s, p.nlz);
}

template <typename T>
static void lines_lx(const PropArgs<T>& a, int s, int mode, const PencilTiles& p, size_t lds, hipStream_t st)
{
    if (a.rlines) {
        const int pm = mode == PZ_BEGIN ? PL_BEGIN : mode == PZ_RESUME ? PL_RESUME : mode == PZ_LAST ? PL_LAST : PL_STEP;
        const int nc = a.m1 * a.m2;
        switch (a.m0) {
#define RX_CASE(L)                                                                                                \
    case L:                                                                                                       \
        if constexpr (rx_ok<T>(L)) {                                                                              \
            constexpr int TX = rl_tx<T>(L), X1 = rl_x1(L);                                                         \
            using RG = C2Geom<L, TX, X1>;                                                                         \
            hipLaunchKernelGGL((k_scol2<T, L, 0, 0, TX, X1>), dim3((nc + TX - 1) / TX, a.nb), dim3(RG::NT),      \
                               (scol_lds_bytes<T, L, TX, X1>()), st, a, s, pm);                                   \
            return;                                                                                               \
        }                                                                                                         \
        break;
            RX_LENGTHS(RX_CASE)
#undef RX_CASE
            default: break;
        }
    }
    const int ntz = (a.m2 + p.tx - 1) / p.tx;
    const dim3 g(a.m1 * ntz, a.nb);
    if (a.pencil_ct) switch (a.m0) {
#define PC_CASE(L) \
    case L: hipLaunchKernelGGL((k_lx<T, L>), g, dim3(PC_NT), lds, st, a, s, mode, p.tx); return;
            PC_CT_LENGTHS(PC_CASE)
#undef PC_CASE
            default: break;
        }
    hipLaunchKernelGGL((k_lx<T, 0>), g, dim3(PC_NT), lds, st, a, s, mode, p.tx);
}

template <typename T>
static void lines_begin(const PropArgs<T>& a, hipStream_t st)
{
    const PencilTiles p = pencil_tiles<T>(a.m0, a.m1, a.m2, a.pencil_ct);
    lines_lx<T>(a, 0, PZ_BEGIN, p, pencil_lds_bytes<T>(a.m0, a.m1, a.m2, a.pencil_ct), st);
}

template <typename T>
static void lines_step(const PropArgs<T>& a, int s, hipStream_t st)
{
    const PencilTiles p = pencil_tiles<T>(a.m0, a.m1, a.m2, a.pencil_ct);
    const size_t lds = pencil_lds_bytes<T>(a.m0, a.m1, a.m2, a.pencil_ct);
    lines_ly<T>(a, a.nb, 2 * a.nb, p, lds, st);
    lines_lz<T>(a, s, p, lds, st);
    lines_ly<T>(a, a.nb, a.nb, p, lds, st);
    lines_lx<T>(a, s, s == a.N - 1 ? PZ_LAST : PZ_STEP, p, lds, st);
}

template <typename T>
static void lines_resume(const PropArgs<T>& a, int c, hipStream_t st)
{
    const PencilTiles p = pencil_tiles<T>(a.m0, a.m1, a.m2, a.pencil_ct);
    lines_lx<T>(a, c, PZ_RESUME, p, pencil_lds_bytes<T>(a.m0, a.m1, a.m2, a.pencil_ct), st);
}

template <typename T>
static hipError_t upload_twiddles(cpx<T>** dst, int n)
{
    std::vector<cpx<T>> h(n);
    const long double pi = acosl(-1.0L);
    for (int t = 0; t < n; ++t) {
        const long double a = -2.0L * pi * (long double)t / (long double)n;
        h[t].x = (T)cosl(a);
        h[t].y = (T)sinl(a);
    }
    hipError_t e = hipMalloc((void**)dst, sizeof(cpx<T>) * n);
    if (e != hipSuccess) return e;
    return hipMemcpy(*dst, h.data(), sizeof(cpx<T>) * n, hipMemcpyHostToDevice);
}

template <typename T>
std::string PropEngine<T>::init(const int m[3], const double D[3], int N, int nbmax, int ck,
                                const scft_schedule* sched)
{
    for (int i = 0; i < 3; ++i) { m_[i] = m[i]; D_[i] = D[i]; }
    N_ = N;
    nbmax_ = nbmax;
    ck_ = ck > 0 ? ck : 0;
    qpitch_ = ck_ ? ck_pitch(N, ck_) : N + 1;
    M_ = (long long)m[0] * m[1] * m[2];
    const scft_schedule sc = sched ? *sched : scft_schedule{0, 0, 0, 0};
    const char* pen_rt = getenv("SCFT_PENCIL_RT");  // =1: runtime-length line kernels only
    pencil_ct_ = !(pen_rt && pen_rt[0] == '1');
    const char* rl = getenv("SCFT_RLINES");  // =0: LDS Stockham y and z passes in the split step
    rlines_ = pencil_ct_ && !(rl && rl[0] == '0');
    const KernelSet<T>* kset = find_kernels<T>(m[0], m[1], m[2]);
    const bool pencil_ok = pencil_supported(m[0], m[1], m[2]);
    // persistent plane_mid (grids whose kernel set has it): one workgroup per CU;
    // SCFT_PLANE_GRID sets the count, 0 turns it off
    {
        int dev = 0, cus = 0;
        if (hipGetDevice(&dev) == hipSuccess &&
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess)
            pgrid_ = cus;
        if (const char* pg = getenv("SCFT_PLANE_GRID")) pgrid_ = std::max(0, atoi(pg));
    }
    // Kernel family. Default: plane/column kernels where this grid has them (the two-pass
    // step when it has that set: 16-22% faster per contour step than the four-launch
    // two-stage step on every grid measured), otherwise the split line step (3-14% faster
    // than the pencil step); SCFT_PENCIL=1 forces the pencil step, SCFT_LEGACY_FFT=1 keeps
    // the Stockham plane/column kernels, SCFT_FAMILY=two-stage|two-pass|stockham|pencil|
    // pencil-split names one. Without a two-stage set, a
    // Stockham plane that fills a CU's LDS runs one workgroup per CU, and the line kernels
    // are faster there (96^3 FP64: 0.529 against 0.805 ms per contour step for 8 replicas).
    int family = sc.family;
    if (family == SCFT_SCHED_DEFAULT) {
        const char* fe = getenv("SCFT_FAMILY");
        if (fe && *fe) {
            const std::string f(fe);
            family = f == "two-stage"      ? SCFT_SCHED_TWO_STAGE
                     : f == "two-pass"     ? SCFT_SCHED_TWO_PASS
                     : f == "stockham"     ? SCFT_SCHED_STOCKHAM
                     : f == "pencil"       ? SCFT_SCHED_PENCIL
                     : f == "pencil-split" ? SCFT_SCHED_PENCIL_SPLIT
                                           : SCFT_SCHED_DEFAULT;
        }
    }
    if (family == SCFT_SCHED_DEFAULT) {
        const char* pen = getenv("SCFT_PENCIL");
        const char* legacy = getenv("SCFT_LEGACY_FFT");
        const bool force_legacy = legacy && legacy[0] == '1';
        if (pen && pen[0] == '1') family = SCFT_SCHED_PENCIL;
        else if (!kset) family = SCFT_SCHED_PENCIL_SPLIT;
        else if (kset->step3 && !force_legacy) family = SCFT_SCHED_TWO_PASS;
        else if (kset->step2 && !force_legacy) family = SCFT_SCHED_TWO_STAGE;
        else if (!kset->step2 && kset->big_plane && !force_legacy && pencil_ok) family = SCFT_SCHED_PENCIL_SPLIT;
        else family = SCFT_SCHED_STOCKHAM;
    } else {
        // an explicit family must exist as such for this grid
        const bool have = family == SCFT_SCHED_TWO_STAGE    ? kset && kset->step2
                          : family == SCFT_SCHED_TWO_PASS     ? kset && kset->step3
                          : family == SCFT_SCHED_STOCKHAM     ? kset && kset->step != kset->step2
                          : family == SCFT_SCHED_PENCIL       ? pencil_ok
                          : family == SCFT_SCHED_PENCIL_SPLIT ? pencil_ok
                                                              : false;
        if (!have) return "schedule: kernel family " + std::to_string(family) + " is not available for this grid";
    }
    hipError_t e = hipSuccess;
    if (family == SCFT_SCHED_TWO_STAGE || family == SCFT_SCHED_TWO_PASS || family == SCFT_SCHED_STOCKHAM) {
        ks_ = kset;
        const bool use_two = family == SCFT_SCHED_TWO_STAGE, use_tp = family == SCFT_SCHED_TWO_PASS;
        setup_ = use_tp ? ks_->setup3 : use_two ? ks_->setup2 : ks_->setup;
        begin_ = use_tp ? ks_->begin3 : use_two ? ks_->begin2 : ks_->begin;
        step_ = use_tp ? ks_->step3 : use_two ? ks_->step2 : ks_->step;
        resume_ = use_tp ? ks_->resume3 : use_two ? ks_->resume2 : ks_->resume;
        path_ = use_tp ? "two-pass" : use_two ? "two-stage" : "stockham";
        // column tile: the schedule's, else SCFT_COLTILE=<width>, else the family's default
        const int dflt = use_tp ? ks_->tc3 : ks_->tc;
        int ct = sc.coltile;
        if (ct == 0 && (use_two || use_tp)) {
            const char* cte = getenv("SCFT_COLTILE");
            const int alt = use_tp ? ks_->tc_alt3 : ks_->tc_alt;
            if (cte && alt && atoi(cte) == alt) ct = alt;
        }
        if (ct != 0 && ct != dflt && !((use_two && ct == ks_->tc_alt) || (use_tp && ct == ks_->tc_alt3)))
            return "schedule: column tile " + std::to_string(ct) + " is not available for this grid";
        if (ct == 0) ct = dflt;
        // the kernels take coltile 0 as the grid's TC
        if (ct != ks_->tc) coltile_ = ct;
        if (ct != dflt)
            path_ = use_tp ? (ct == 32 ? "two-pass-c32" : "two-pass-c64") : (ct == 32 ? "two-stage-c32" : "two-stage-c64");
        // plane workgroup size: the schedule's, else SCFT_PLANE_NT=<threads>, else the default
        int pn = sc.plane_nt;
        if (pn == 0 && use_two) {
            const char* pne = getenv("SCFT_PLANE_NT");
            if (pne && ks_->plane_nt_alt && atoi(pne) == ks_->plane_nt_alt) pn = ks_->plane_nt_alt;
        }
        if (pn != 0 && !(use_two && (pn == ks_->plane_nt || pn == ks_->plane_nt_alt)) && !(use_tp && pn == ks_->plane_nt))
            return "schedule: plane workgroup size " + std::to_string(pn) + " is not available for this grid";
        if (pn != 0 && pn != ks_->plane_nt) planent_ = pn;
        e = setup_();
    } else {
        if (!pencil_ok) {
            char b[512];
            snprintf(b, sizeof b, "no %s propagator kernels for grid %dx%dx%d (supported: %s)",
                     sizeof(T) == 8 ? "FP64" : "FP32", m[0], m[1], m[2], supported_grids(sizeof(T) == 8).c_str());
            return b;
        }
        if (sc.coltile != 0 || sc.plane_nt != 0) return "schedule: tile sizes apply to the two-stage kernels only";
        ks_ = nullptr;
        const bool split = family == SCFT_SCHED_PENCIL_SPLIT;
        begin_ = split ? &lines_begin<T> : &pencil_begin<T>;
        step_ = split ? &lines_step<T> : &pencil_step<T>;
        resume_ = split ? &lines_resume<T> : &pencil_resume<T>;
        path_ = split ? "pencil-split" : "pencil";
        e = split ? lines_setup<T>() : pencil_setup<T>();
    }
    if (e != hipSuccess) return std::string("hipFuncSetAttribute: ") + hipGetErrorString(e);
    nblk_ = (int)std::min<long long>(1024, (M_ + RED_NT - 1) / RED_NT);
    // A failed hipMalloc also becomes the thread's last error, which a later
    // hipGetLastError() after kernel launches would report: it is read off here.
    auto alloc = [&](void** p, size_t bytes) -> bool {
        if (hipMalloc(p, bytes) != hipSuccess) {
            (void)hipGetLastError();
            return false;
        }
        bytes_ += bytes;
        return true;
    };
    const size_t fld = sizeof(cpx<T>) * (size_t)M_ * nbmax;
    if (!alloc((void**)&Ch_, fld) || !alloc((void**)&Cf_, fld) ||
        !alloc((void**)&e4A_, sizeof(T) * (size_t)M_ * nbmax) ||
        !alloc((void**)&e4B_, sizeof(T) * (size_t)M_ * nbmax) ||
        !alloc((void**)&qs_, sizeof(cpx<T>) * (size_t)M_ * (size_t)qpitch_ * nbmax) ||
        !alloc((void**)&part_, sizeof(double) * (size_t)nblk_ * 5 * nbmax) ||
        !alloc((void**)&cAB_, sizeof(double) * 2 * nbmax) || !alloc((void**)&scal_, sizeof(double) * SCAL_N * nbmax) ||
        !alloc((void**)&rn_, sizeof(double) * 3 * nbmax))
        return "device allocation failed (propagator buffers)";
    for (int i = 0; i < 3; ++i) {
        e = upload_twiddles<T>(&tw_[i], m[i]);
        if (e != hipSuccess) return std::string("twiddles: ") + hipGetErrorString(e);
        if (hipMalloc((void**)&g_[i], sizeof(T) * m[i]) != hipSuccess) return "device allocation failed (tables)";
    }
    e = set_box(D);
    if (e != hipSuccess) return std::string("kinetic tables: ") + hipGetErrorString(e);
    // Graph replay of the contour loop is opt-in: on MI355X it measured slower than direct
    // launches from two streams (0.0816 vs 0.0744 ms per contour step, 14 x 64^3 FP32).
    const char* gr = getenv("SCFT_GRAPH");
    use_graph_ = gr && gr[0] == '1';
    // replica-group streams: the schedule's, else SCFT_STREAMS, else the default
    if (sc.streams < 0 || sc.streams > kMaxStreams)
        return "schedule: " + std::to_string(sc.streams) + " replica-group streams (1 .. " +
               std::to_string(kMaxStreams) + ")";
    const char* ns = getenv("SCFT_STREAMS");
    nstreams_ = sc.streams > 0 ? sc.streams : std::max(1, std::min(kMaxStreams, ns ? atoi(ns) : kDefaultStreams));
    // fork/join events (no timing) and the extra streams this replica-group count uses
    if (hipEventCreateWithFlags(&fork_, hipEventDisableTiming) != hipSuccess) return "event creation failed";
    if (ensure_streams(nstreams_) != hipSuccess) return "stream / event creation failed";
    return "";
}

// Every schedule this grid can run for nb slots: the kernel families, for the two-stage
// kernels each column tile x plane workgroup size, each with 1 .. min(3, nb) replica-group
// streams.
template <typename T>
std::vector<scft_schedule> PropEngine<T>::candidates(const int m[3], int nb)
{
    std::vector<scft_schedule> out;
    const KernelSet<T>* k = find_kernels<T>(m[0], m[1], m[2]);
    std::vector<scft_schedule> fam;
    if (k && k->step2) {
        const int cts[2] = {k->tc, k->tc_alt}, pns[2] = {k->plane_nt, k->plane_nt_alt};
        for (int ct : cts)
            for (int pn : pns)
                if (ct && pn) fam.push_back({SCFT_SCHED_TWO_STAGE, ct, pn, 0});
    }
    if (k && k->step3) {
        const int cts[2] = {k->tc3, k->tc_alt3};
        for (int ct : cts)
            if (ct) fam.push_back({SCFT_SCHED_TWO_PASS, ct, k->plane_nt, 0});
    }
    if (k && k->step != k->step2) fam.push_back({SCFT_SCHED_STOCKHAM, 0, 0, 0});
    if (pencil_supported(m[0], m[1], m[2])) {
        fam.push_back({SCFT_SCHED_PENCIL_SPLIT, 0, 0, 0});
        fam.push_back({SCFT_SCHED_PENCIL, 0, 0, 0});
    }
    for (scft_schedule f : fam)
        for (int ns = 1; ns <= std::min(3, std::max(1, nb)); ++ns) {
            f.streams = ns;
            out.push_back(f);
        }
    return out;
}

template <typename T>
std::string PropEngine<T>::config() const
{
    char b[320];
    int n = snprintf(b, sizeof b, "%s, %d replica-group stream%s, %s", path_, nstreams_, nstreams_ > 1 ? "s" : "",
                     use_graph_ ? "graph replay" : "direct launches");
    if (ks_ && ks_->persist_mid && pgrid_ > 0 && planent_ == 0 && strncmp(path_, "two-stage", 9) == 0 && n > 0 &&
        n < (int)sizeof b)
        n += snprintf(b + n, sizeof b - n, ", persistent plane_mid on %d workgroups", pgrid_);
    if (planent_ > 0 && n > 0 && n < (int)sizeof b)
        n += snprintf(b + n, sizeof b - n, ", %d-thread plane workgroups", planent_);
    if (ck_ && n > 0 && n < (int)sizeof b)
        snprintf(b + n, sizeof b - n, ", q checkpointed every %d steps (%d of %d entries per replica)", ck_, qpitch_,
                 N_ + 1);
    return b;
}

// Per-axis kinetic tables for a_i = (2 pi k_i / D_i)^2 ds / 12 with k_i the signed frequency
// of index k (setup_k_sq, density.h:79-95): exp(-a_i) in FP64, 1 - exp(-a_i) in FP32.
template <typename T>
hipError_t PropEngine<T>::set_box(const double D[3])
{
    const double pi = 4 * atan(1.0);
    const double ds = 1.0 / N_;
    for (int i = 0; i < 3; ++i) {
        D_[i] = D[i];
        std::vector<T> h(m_[i]);
        for (int k = 0; k < m_[i]; ++k) {
            const int kk = (k <= m_[i] / 2) ? k : k - m_[i];
            const double k2 = 4 * pi * pi * (double)kk * kk / (D[i] * D[i]);
            h[k] = Kinetic<T>::table(k2 * ds / 12);
        }
        hipError_t e = hipMemcpy(g_[i], h.data(), sizeof(T) * m_[i], hipMemcpyHostToDevice);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

template <typename T>
hipError_t PropEngine<T>::run(const double* dW, long long wstride, double* dphi, long long phistride,
                              const int* d_slot2rep, int nb, const ChainConst& cc, double* d_phis, hipStream_t st)
{
    if (nb <= 0) return hipSuccess;
    PropArgs<T> a = args(nb, cc);

    const dim3 rg(nblk_, nb);
    hipLaunchKernelGGL(k_wsum, rg, dim3(RED_NT), 0, st, dW, wstride, d_slot2rep, M_, part_);
    hipLaunchKernelGGL(k_wmean, dim3(nb), dim3(RED_NT), 0, st, part_, nblk_, M_, cAB_);
    hipLaunchKernelGGL((k_factors<T>), rg, dim3(RED_NT), 0, st, dW, wstride, d_slot2rep, M_, cc.ds, cAB_, e4A_,
                       e4B_);
    if (timing_) (void)hipEventRecord(ev_[0], st);
    hipError_t e = contour(a, nb, cc, st);
    if (e != hipSuccess) return e;
    if (timing_) (void)hipEventRecord(ev_[1], st);
    hipLaunchKernelGGL((k_qsum<T>), rg, dim3(RED_NT), 0, st, qs_, M_, qpitch_, store_index(cc.N, cc),
                       store_index(cc.Nh, cc), part_);
    hipLaunchKernelGGL(k_scalars, dim3(nb), dim3(RED_NT), 0, st, part_, nblk_, M_, cc, cAB_, scal_);
    if (ck_ == 0) {
        hipLaunchKernelGGL((k_density<T>), rg, dim3(RED_NT), 0, st, qs_, M_, cc, scal_, dphi, phistride,
                           d_slot2rep, part_);
    } else {
        e = windowed_density(a, nb, cc, dphi, phistride, d_slot2rep, st);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(k_phis, dim3(nb), dim3(RED_NT), 0, st, part_, nblk_, M_, d_phis);
    if (std::is_same<T, float>::value && cc.phihA == 0.0) {
        hipLaunchKernelGGL(k_renorm_factors, dim3(nb), dim3(64), 0, st, cc, scal_, d_phis, rn_);
        hipLaunchKernelGGL(k_renorm_apply, rg, dim3(RED_NT), 0, st, M_, rn_, dphi, phistride, d_slot2rep);
    }
    return hipGetLastError();
}

// Recompute of contour indices lo..hi (hi <= N) into window w of a checkpointed store:
// resume from the checkpoint c at or below lo, then steps c .. hi-1. Only lo..hi are stored.
template <typename T>
void PropEngine<T>::window(PropArgs<T> a, int w, int lo, int hi, hipStream_t st)
{
    const int c = (lo / ck_) * ck_;
    a.qlo = lo;
    a.qhi = hi;
    a.qbase = ck_window(N_, ck_, w);
    a.qsrc = c / ck_;
    resume_(a, c, st);
    for (int s = c; s < hi; ++s) step_(a, s, st);
}

// The Simpson sums of k_density on a checkpointed store. Pairs (s, L - s) with L = N
// (copolymer) or Nh (homopolymer) and s <= L - s are taken K lower indices at a time:
// window 0 holds s in [lo, hi), window 1 the partners L - s in (L - hi, L - lo].
template <typename T>
hipError_t PropEngine<T>::windowed_density(const PropArgs<T>& a, int nb, const ChainConst& cc, double* dphi,
                                           long long phistride, const int* d_slot2rep, hipStream_t st)
{
    const dim3 rg(nblk_, nb);
    hipLaunchKernelGGL(k_dens_zero, rg, dim3(RED_NT), 0, st, M_, dphi, phistride, d_slot2rep);
    const int K = ck_;
    for (int kind = 0; kind < 2; ++kind) {
        const int L = kind == 0 ? cc.N : cc.Nh;
        const int first = kind == 0 ? 0 : 1;  // homopolymer pairs start at s = 1
        for (int lo = first; 2 * lo <= L && (kind == 0 || L - lo >= 1); lo += K) {
            int hi = std::min(lo + K, L / 2 + 1);
            if (hi <= lo) break;
            window(a, 0, lo, hi - 1, st);
            window(a, 1, L - hi + 1, L - lo, st);
            hipLaunchKernelGGL((k_dens_win<T>), rg, dim3(RED_NT), 0, st, qs_, M_, qpitch_, cc, kind, lo, hi, lo,
                               ck_window(N_, K, 0), L - hi + 1, ck_window(N_, K, 1), dphi, phistride, d_slot2rep);
        }
    }
    hipLaunchKernelGGL((k_dens_finish<T>), rg, dim3(RED_NT), 0, st, qs_, M_, qpitch_, store_index(cc.Nh, cc), scal_,
                       dphi, phistride, d_slot2rep, part_);
    return hipGetLastError();
}

template <typename T>
PropArgs<T> PropEngine<T>::args(int nb, const ChainConst& cc) const
{
    PropArgs<T> a;
    a.M = M_;
    a.m0 = m_[0];
    a.m1 = m_[1];
    a.m2 = m_[2];
    a.N = cc.N;
    a.n1 = cc.n1;
    a.n2 = cc.n2;
    a.Nh = cc.Nh;
    a.nb = nb;
    a.pencil_ct = pencil_ct_ ? 1 : 0;
    a.rlines = rlines_ ? 1 : 0;
    a.coltile = coltile_;
    a.pgrid = pgrid_;
    a.planent = planent_;
    a.Ch = Ch_;
    a.Cf = Cf_;
    a.e4A = e4A_;
    a.e4B = e4B_;
    a.qs = qs_;
    a.qpitch = qpitch_;
    a.qck = ck_;
    a.qlo = 1;  // no window
    a.qhi = 0;
    a.qbase = 0;
    a.qsrc = 0;
    a.tw0 = tw_[0];
    a.tw1 = tw_[1];
    a.tw2 = tw_[2];
    a.gx = g_[0];
    a.gy = g_[1];
    a.gz = g_[2];
    a.invM = (T)(1.0 / (double)M_);
    return a;
}

// Box derivative of the factored QC', Q2' (adjoint of props; the pipeline and its maths
// are described in grad_kernels.hpp). The q store and the Boltzmann factors have the same
// layout on every path, so either gradient pipeline reads what any props path stored.
template <typename T>
hipError_t PropEngine<T>::box_gradient(int nb, const ChainConst& cc, double* out, hipStream_t st)
{
    if (nb <= 0) return hipSuccess;
    // the two-stage gradient passes where the props run on the two-stage or the two-pass step
    // (the same discrete operator, to rounding); the line pipeline otherwise (any 2-3-5 grid)
    const bool two = ks_ && ks_->gterm2 && (step_ == ks_->step2 || (ks_->step3 && step_ == ks_->step3));
    if (!two && !pencil_supported(m_[0], m_[1], m_[2])) return hipErrorNotSupported;
    hipError_t e;
    if (two && (e = ks_->gsetup2()) != hipSuccess) return e;
    // partials per slot and chain: the column tiles of k_gcol2_acc, or k_gacc's blocks
    const int gt = two ? ks_->gtiles : 0;
    if (!gpart_) {
        // the two-stage pipeline carries a term and its mirror: work fields of 2 nbmax slots
        const size_t fld = sizeof(cpx<T>) * (size_t)M_ * nbmax_;
        const size_t pb = sizeof(double) * (size_t)nbmax_ * 2 * 3 * std::max(nblk_, gt) * 3;
        if (two) {
            if ((e = hipMalloc((void**)&Gh_, 2 * fld)) != hipSuccess ||
                (e = hipMalloc((void**)&Gf_, 2 * fld)) != hipSuccess ||
                (e = hipMalloc((void**)&K0_, 2 * fld)) != hipSuccess ||
                (e = hipMalloc((void**)&Kf_, 2 * fld)) != hipSuccess || (e = hipMalloc((void**)&gpart_, pb)) != hipSuccess) {
                (void)hipGetLastError();  // not left for the next launch check to report
                return e;
            }
            bytes_ += 8 * fld + pb;
        } else {
            if ((e = hipMalloc((void**)&K0_, fld)) != hipSuccess || (e = hipMalloc((void**)&K1_, fld)) != hipSuccess ||
                (e = hipMalloc((void**)&Kf_, fld)) != hipSuccess || (e = hipMalloc((void**)&gpart_, pb)) != hipSuccess) {
                (void)hipGetLastError();
                return e;
            }
            bytes_ += 3 * fld + pb;
            const int lds = (int)pencil_lds_max<T>();
            std::vector<const void*> fns = {(const void*)k_gz<T, 0>, (const void*)k_gx<T, 0>,
                                            (const void*)k_py<T, false>, (const void*)k_py<T, true>};
#define PC_FNS(L)                                                                                                 \
    fns.push_back((const void*)k_gz<T, L>);                                                                       \
    fns.push_back((const void*)k_gx<T, L>);                                                                       \
    fns.push_back((const void*)k_py<T, false, L>);                                                                \
    fns.push_back((const void*)k_py<T, true, L>);
            PC_CT_LENGTHS(PC_FNS)
#undef PC_FNS
            for (const void* f : fns)
                if ((e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, lds)) != hipSuccess)
                    return e;
        }
    }
    // k_gacc workgroups per slot: about 1024 over the launch (4 per CU), so that many
    // replicas take several points per thread and the block sums stay few (one point per
    // thread over 8 x 1024 workgroups cost 23 us per launch at 8 x 64^3, 128 workgroups
    // per slot 17 us; one replica on 128 workgroups left half the CUs idle)
    const int gblk = two ? gt : std::max(1, std::min(nblk_, std::max(128, 1024 / std::max(nb, 1))));
    const int nslot = two ? 2 * nb : nb;            // mirror terms add into slots nb .. 2nb-1
    const size_t np = (size_t)nslot * gblk * 3;  // partials per chain
    if ((e = hipMemsetAsync(gpart_, 0, sizeof(double) * 3 * np, st)) != hipSuccess) return e;

    // One launch per pass over all slots on st: at 64^3 the passes take 8-16 us each, and
    // replica groups on their own streams (as the props run) doubled the launches per term
    // and left the groups' kernels alternating rather than overlapping.
    const PropArgs<T> a = args(nb, cc);
    // The line passes run on the props tiling: the kernel compiled for the axis length where
    // one exists (its fixed tile), else the runtime plan.
    const PencilTiles p = pencil_tiles<T>(m_[0], m_[1], m_[2], pencil_ct_);
    const size_t lds = pencil_lds_bytes<T>(m_[0], m_[1], m_[2], pencil_ct_);
    const double twopi = 8 * atan(1.0);
    const GradConst gc{cc.ds, {twopi / D_[0], twopi / D_[1], twopi / D_[2]}, {D_[0], D_[1], D_[2]}};
    const long long lines = (long long)m_[0] * m_[1];
    const dim3 gzg((unsigned)((lines + p.nlz - 1) / p.nlz), nb);
    const int ntx = (m_[2] + p.tx - 1) / p.tx;
    auto gz = [&](int mode, int ib, int pbk, int ia, int pak, int spec) {
        if (pencil_ct_) switch (m_[2]) {
#define PC_CASE(L)                                                                                                \
    case L: hipLaunchKernelGGL((k_gz<T, L>), gzg, dim3(PC_NT), lds, st, a, mode, ib, pbk, ia, pak, spec, p.nlz);   \
        return;
                PC_CT_LENGTHS(PC_CASE)
#undef PC_CASE
                default: break;
            }
        hipLaunchKernelGGL((k_gz<T, 0>), gzg, dim3(PC_NT), lds, st, a, mode, ib, pbk, ia, pak, spec, p.nlz);
    };
    auto gx = [&](int mode, int nfields, cpx<T>* k0, cpx<T>* kf) {
        const dim3 g(m_[1] * ntx, nfields);
        if (pencil_ct_) switch (m_[0]) {
#define PC_CASE(L)                                                                                                \
    case L: hipLaunchKernelGGL((k_gx<T, L>), g, dim3(PC_NT), lds, st, a, mode, nb, p.tx, k0, kf); return;
                PC_CT_LENGTHS(PC_CASE)
#undef PC_CASE
                default: break;
            }
        hipLaunchKernelGGL((k_gx<T, 0>), g, dim3(PC_NT), lds, st, a, mode, nb, p.tx, k0, kf);
    };
    // The term of contour step s of a chain, b and a read from q-store entries ib, ia:
    //   chain 0: b = q1[s], a = q2[N-1-s], species of chain 1 on step s (density.h:136-146)
    //   chain 1: b = q2[t], a = q2[Nh-1-t], species of chain 2 on step t
    //   chain 2: b = q1[t], a = q1[Nh-1-t], species of chain 1 on step t (A for Nh <= n1)
    // and its mirror m = L-1-s (L = N or Nh) reads the same entries the other way round.
    auto species = [&](int chain, int s) { return chain == 1 ? (s < cc.n2 ? 1 : 0) : (s < cc.n1 ? 0 : 1); };
    auto term = [&](int chain, int s, int ib, int ia) {
        double* pc = gpart_ + (size_t)chain * np;
        const int pbk = chain == 1 ? 1 : 0, pak = chain == 2 ? 0 : 1;
        const int spec = species(chain, s);
        gz((int)GZ_INIT, ib, pbk, ia, pak, spec);
        pencil_py<T, false>(a, nb, 2 * nb, p, lds, st);
        gx((int)GX_SPLIT, 2 * nb, K0_, Kf_);
        pencil_py<T, true>(a, nb, nb, p, lds, st);
        gz((int)GZ_MID, 0, 0, 0, 0, spec);
        pencil_py<T, false>(a, nb, nb, p, lds, st);
        gx((int)GX_FWD, nb, K1_, (cpx<T>*)nullptr);
        hipLaunchKernelGGL((k_gacc<T>), dim3(gblk, nb), dim3(RED_NT), 0, st, K0_, K1_, Kf_, M_, m_[0], m_[1],
                           m_[2], gc, pc);
    };
    // terms s and m = L-1-s, from entries is (index s) and im (index m)
    auto pair = [&](int chain, int s, int m, int is, int im) {
        if (two) {
            const int pbk = chain == 1 ? 1 : 0, pak = chain == 2 ? 0 : 1;
            ks_->gterm2(a, is, pbk, im, pak, species(chain, s), m != s ? species(chain, m) : -1, Gh_, Gf_, K0_, Kf_,
                        gc, gpart_ + (size_t)chain * np, st);
            return;
        }
        term(chain, s, is, im);
        if (m != s) term(chain, m, im, is);
    };
    const bool hA = cc.phihA != 0.0;  // chain 2 (the A homopolymer's Q1) only when it is present
    if (ck_ == 0) {
        for (int s = 0; 2 * s <= cc.N - 1; ++s) pair(0, s, cc.N - 1 - s, s, cc.N - 1 - s);
        for (int chain = 1; chain <= (hA ? 2 : 1); ++chain)
            for (int t = 0; 2 * t <= cc.Nh - 1; ++t) pair(chain, t, cc.Nh - 1 - t, t, cc.Nh - 1 - t);
    } else {
        // Terms s and L-1-s (L = N for chain 0, Nh for chains 1, 2) read the same two
        // contour indices: window 0 holds s in [lo, hi), window 1 the mirrors (L-1-hi, L-1-lo].
        const int w0 = ck_window(N_, ck_, 0), w1 = ck_window(N_, ck_, 1);
        for (int kind = 0; kind < 2; ++kind) {
            const int L = kind == 0 ? cc.N : cc.Nh;
            for (int lo = 0; 2 * lo <= L - 1; lo += ck_) {
                const int hi = std::min(lo + ck_, (L - 1) / 2 + 1);
                PropArgs<T> aw = a;
                window(aw, 0, lo, hi - 1, st);
                window(aw, 1, L - hi, L - 1 - lo, st);
                for (int s = lo; s < hi; ++s) {
                    const int m = L - 1 - s, is = w0 + (s - lo), im = w1 + (m - (L - hi));
                    for (int chain = kind == 0 ? 0 : 1; chain <= (kind == 0 ? 0 : (hA ? 2 : 1)); ++chain)
                        pair(chain, s, m, is, im);
                }
            }
        }
    }
    if ((e = hipGetLastError()) != hipSuccess) return e;
    std::vector<double> h(3 * np);
    if ((e = hipMemcpyAsync(h.data(), gpart_, sizeof(double) * 3 * np, hipMemcpyDeviceToHost, st)) != hipSuccess)
        return e;
    if ((e = hipStreamSynchronize(st)) != hipSuccess) return e;
    const double scale = 1.0 / ((double)M_ * (double)M_);
    for (int slot = 0; slot < nb; ++slot)
        for (int chain = 0; chain < 3; ++chain)
            for (int i = 0; i < 3; ++i) {
                double v = 0;
                for (int sl = slot; sl < nslot; sl += nb)  // the slot's terms, then its mirror terms
                    for (int b = 0; b < gblk; ++b) v += h[(size_t)chain * np + ((size_t)sl * gblk + b) * 3 + i];
                out[(size_t)slot * 9 + chain * 3 + i] = v * scale;
            }
    return hipSuccess;
}

// The contour loop (begin + N steps) for nb slots. With graphs enabled (SCFT_GRAPH=1) each
// replica group's loop is captured once per batch size and replayed on its own stream.
template <typename T>
hipError_t PropEngine<T>::contour(PropArgs<T> a, int nb, const ChainConst& cc, hipStream_t st)
{
    if (!use_graph_) return contour_launch(a, nb, cc, st);
    PropArgs<T> g[kMaxStreams];
    hipStream_t gs[kMaxStreams];
    const int ng = groups(a, nb, st, g, gs);
    const GraphEntry* ge = nullptr;
    for (const auto& x : graphs_)
        if (x.nb == nb && x.st == st) ge = &x;
    hipError_t e = hipSuccess;
    if (!ge) {
        GraphEntry n{nb, st, ng, {}};
        for (int i = 0; i < ng && e == hipSuccess; ++i) {
            hipGraph_t graph = nullptr;
            e = hipStreamBeginCapture(gs[i], hipStreamCaptureModeThreadLocal);
            if (e != hipSuccess) break;
            begin_(g[i], gs[i]);
            for (int s = 0; s < cc.N; ++s) step_(g[i], s, gs[i]);
            const hipError_t el = hipGetLastError();
            e = hipStreamEndCapture(gs[i], &graph);
            if (e == hipSuccess) e = el;
            if (e == hipSuccess) e = hipGraphInstantiate(&n.exec[i], graph, nullptr, nullptr, 0);
            if (graph) (void)hipGraphDestroy(graph);
        }
        if (e != hipSuccess) {
            for (hipGraphExec_t x : n.exec)
                if (x) (void)hipGraphExecDestroy(x);
            return e;
        }
        graphs_.push_back(n);
        ge = &graphs_.back();
    }
    if (ng > 1) {
        e = hipEventRecord(fork_, st);
        for (int i = 1; i < ng && e == hipSuccess; ++i) e = hipStreamWaitEvent(gs[i], fork_, 0);
    }
    for (int i = 0; i < ng && e == hipSuccess; ++i) e = hipGraphLaunch(ge->exec[i], gs[i]);
    for (int i = 1; i < ng && e == hipSuccess; ++i) {
        e = hipEventRecord(join_[i - 1], gs[i]);
        if (e == hipSuccess) e = hipStreamWaitEvent(st, join_[i - 1], 0);
    }
    return e;
}

// Destroys every captured contour loop (after the work that used them is complete: the
// callers synchronise first).
template <typename T>
void PropEngine<T>::clear_graphs()
{
    for (auto& g : graphs_)
        for (hipGraphExec_t x : g.exec)
            if (x) (void)hipGraphExecDestroy(x);
    graphs_.clear();
}

template <typename T>
int PropEngine<T>::groups(const PropArgs<T>& a, int nb, hipStream_t st, PropArgs<T>* g, hipStream_t* gs) const
{
    const int ng = std::max(1, std::min(nstreams_, nb));
    int first = 0;
    for (int i = 0; i < ng; ++i) {
        const int cnt = nb / ng + (i < nb % ng ? 1 : 0);
        const size_t off = (size_t)first * (size_t)M_;
        g[i] = a;
        g[i].nb = cnt;
        g[i].Ch = Ch_ + off;
        g[i].Cf = Cf_ + off;
        g[i].e4A = e4A_ + off;
        g[i].e4B = e4B_ + off;
        g[i].qs = qs_ + off * (size_t)qpitch_;
        gs[i] = i == 0 ? st : sx_[i - 1];
        first += cnt;
    }
    return ng;
}

// Independent replica groups on their own streams: one group's LDS/latency-bound plane
// kernels co-run with another group's bandwidth-bound column kernels.
template <typename T>
hipError_t PropEngine<T>::contour_launch(PropArgs<T> a, int nb, const ChainConst& cc, hipStream_t st)
{
    PropArgs<T> g[kMaxStreams];
    hipStream_t gs[kMaxStreams];
    const int ng = groups(a, nb, st, g, gs);
    hipError_t e = hipSuccess;
    if (ng > 1) {
        e = hipEventRecord(fork_, st);
        for (int i = 1; i < ng && e == hipSuccess; ++i) e = hipStreamWaitEvent(gs[i], fork_, 0);
        if (e != hipSuccess) return e;
    }
    for (int i = 0; i < ng; ++i) begin_(g[i], gs[i]);
    for (int s = 0; s < cc.N; ++s)
        for (int i = 0; i < ng; ++i) step_(g[i], s, gs[i]);
    for (int i = 1; i < ng && e == hipSuccess; ++i) {
        e = hipEventRecord(join_[i - 1], gs[i]);
        if (e == hipSuccess) e = hipStreamWaitEvent(st, join_[i - 1], 0);
    }
    if (e != hipSuccess) return e;
    return hipGetLastError();
}

template <typename T>
hipError_t PropEngine<T>::qq_invariant(int slot, double* qq, hipStream_t st)
{
    if (ck_) return hipErrorNotSupported;  // needs every contour index
    std::vector<double> buf((size_t)(N_ + 1) * nblk_);
    double* dtmp = nullptr;
    hipError_t e = hipMalloc((void**)&dtmp, sizeof(double) * buf.size());
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return e;
    }
    hipLaunchKernelGGL((k_qq<T>), dim3(nblk_, N_ + 1), dim3(RED_NT), 0, st, qs_, M_, N_, slot, dtmp);
    e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(buf.data(), dtmp, sizeof(double) * buf.size(), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    (void)hipFree(dtmp);
    if (e != hipSuccess) return e;
    for (int s = 0; s <= N_; ++s) {
        double v = 0;
        for (int b = 0; b < nblk_; ++b) v += buf[(size_t)s * nblk_ + b];
        qq[s] = v / (double)M_;
    }
    return hipSuccess;
}

template class PropEngine<float>;
template class PropEngine<double>;

}  // namespace scft
