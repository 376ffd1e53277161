// prop_engine.hpp -- host-side driver of the batched propagator (one density::props per
// batch slot, density.h:344-459), templated on the propagator's arithmetic type.
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "prop_types.hpp"
#include "scft_api.h"

namespace scft {

template <typename T>
struct KernelSet {
    int m0, m1, m2;
    hipError_t (*setup)();
    void (*begin)(const PropArgs<T>&, hipStream_t);
    void (*step)(const PropArgs<T>&, int s, hipStream_t);
    // two-stage register-FFT variant (nullptr when the grid has none)
    hipError_t (*setup2)();
    void (*begin2)(const PropArgs<T>&, hipStream_t);
    void (*step2)(const PropArgs<T>&, int s, hipStream_t);
    const char* name;
    // start a recompute segment from the stored q at contour index c (PL_RESUME)
    void (*resume)(const PropArgs<T>&, int c, hipStream_t);
    void (*resume2)(const PropArgs<T>&, int c, hipStream_t);
    // the LDS Stockham plane needs more than half a CU's LDS (one workgroup per CU)
    bool big_plane;
    // two-stage column tile: the default width and the alternative (0: none)
    int tc, tc_alt;
    // the two-stage set has the persistent plane_mid (k_plane_mid2p)
    bool persist_mid;
    // one box-gradient term on the two-stage kernels (k_gplane_init2, k_gcol2, k_plane_mid2,
    // k_gcol2_acc), or with specB >= 0 a term and its mirror: b, a from q-store entries ib
    // (part pbk), ia (part pak), species specA (the mirror: entries swapped, species specB);
    // work fields Gh, Gf, K0, Kf of 2 nb slots; gtiles partials per slot (mirror: slots
    // nb .. 2nb-1) added to part (nullptr when the grid has no two-stage set)
    hipError_t (*gsetup2)();
    void (*gterm2)(const PropArgs<T>&, int ib, int pbk, int ia, int pak, int specA, int specB, cpx<T>* Gh,
                   cpx<T>* Gf, cpx<T>* K0, cpx<T>* Kf, const GradConst& gc, double* part, hipStream_t);
    // two-stage plane kernels: the default workgroup size and the alternative (0: none)
    int plane_nt, plane_nt_alt;
    // the two-pass step (prop_sandwich.hpp; nullptr when the grid has none), its default
    // column tile width and the other one (0: none)
    int tc3, tc_alt3;
    hipError_t (*setup3)();
    void (*begin3)(const PropArgs<T>&, hipStream_t);
    void (*step3)(const PropArgs<T>&, int s, hipStream_t);
    void (*resume3)(const PropArgs<T>&, int c, hipStream_t);
    int gtiles;
};

// kernel set for an (m0, m1, m2) grid, or nullptr when no instantiation exists
template <typename T>
const KernelSet<T>* find_kernels(int m0, int m1, int m2);

// list of supported grids as text (for error messages)
std::string supported_grids(bool fp64);

template <typename T>
class PropEngine {
public:
    PropEngine() = default;
    PropEngine(const PropEngine&) = delete;
    PropEngine& operator=(const PropEngine&) = delete;
    ~PropEngine();

    // Allocates buffers for nbmax slots. ck = 0 keeps q at every contour index; ck = K > 0
    // keeps every K-th (ck_index) and recomputes the rest in windows (run, box_gradient).
    // Returns "" on success, else an error message.
    // sched (nullptr: the defaults) picks the kernel family, tile sizes and streams; a
    // choice this grid cannot run is an error ("schedule: ...").
    std::string init(const int m[3], const double D[3], int N, int nbmax, int ck = 0,
                     const scft_schedule* sched = nullptr);
    // every schedule the grid can run for nb slots (the autotuner's candidates)
    static std::vector<scft_schedule> candidates(const int m[3], int nb);
    // Recomputes the kinetic factor tables for a new box D.
    hipError_t set_box(const double D[3]);

    // props for nb slots: W row of slot i = dW + slot2rep[i]*wstride (device), phi likewise.
    // Writes phis [nb][5] (device) and per-slot scalars (scal()).
    hipError_t run(const double* dW, long long wstride, double* dphi, long long phistride,
                   const int* d_slot2rep, int nb, const ChainConst& cc, double* d_phis,
                   hipStream_t st);

    // device [nbmax][SCAL_N]: lnQ, sC, sH, QC, Q2, Q1, QC', Q2', sHA, Q1'
    const double* scal() const { return scal_; }
    // qq[s] = <q1'[s] q2'[N-s]> for s = 0..N of one slot (host array of N+1)
    hipError_t qq_invariant(int slot, double* qq, hipStream_t st);
    // Box derivatives of the factored partition functions of the nb slots of the last run()
    // (same cc): out[slot][0..2] = dQC'/dD_i, out[slot][3..5] = dQ2'/dD_i, out[slot][6..8] =
    // dQ1'/dD_i (zero when cc.phihA == 0) (host array).
    // hipErrorNotSupported when an axis does not factor into 2, 3, 5 or exceeds 1024.
    hipError_t box_gradient(int nb, const ChainConst& cc, double* out, hipStream_t st);
    int checkpoint_interval() const { return ck_; }
    int qstore_entries() const { return qpitch_; }
    long long M() const { return M_; }
    int N() const { return N_; }
    int nbmax() const { return nbmax_; }
    size_t bytes_allocated() const { return bytes_; }
    const char* fft_path() const { return path_; }
    // "two-stage, 2 replica-group streams, direct launches"
    std::string config() const;
    // HIP events around the contour loop of the next run() calls (diagnostics)
    hipError_t enable_timing(bool on);
    // elapsed ms of the contour loop of the last run() (after the stream is synchronised)
    double loop_ms() const;

private:
    static constexpr int kMaxStreams = 4;
    static constexpr int kDefaultStreams = 2;
    PropArgs<T> args(int nb, const ChainConst& cc) const;
    hipError_t ensure_streams(int n);
    hipError_t contour(PropArgs<T> a, int nb, const ChainConst& cc, hipStream_t st);
    // direct launches of the contour loop, the replica-group streams ordered after st and st
    // after them with the engine's fork / join events
    hipError_t contour_launch(PropArgs<T> a, int nb, const ChainConst& cc, hipStream_t st);
    // checkpointed store: recompute contour indices lo..hi into window w (entries
    // ck_window(w) + s - lo), from the checkpoint at or below lo
    void window(PropArgs<T> a, int w, int lo, int hi, hipStream_t st);
    // q-store entry of contour index s outside the windows
    int store_index(int s, const ChainConst& cc) const { return ck_ ? ck_index(s, cc.N, cc.Nh, ck_) : s; }
    hipError_t windowed_density(const PropArgs<T>& a, int nb, const ChainConst& cc, double* dphi,
                                long long phistride, const int* d_slot2rep, hipStream_t st);
    // The replica groups of an nb-slot batch: group i's arguments and stream (group 0 runs
    // on st). Returns the group count.
    int groups(const PropArgs<T>& a, int nb, hipStream_t st, PropArgs<T>* g, hipStream_t* gs) const;
    // The captured contour loops of one batch size: one graph per replica group, each
    // captured alone on its group's stream, so no capture holds an event or a second stream.
    // Replay orders the group streams after st and st after them with the engine's own
    // fork / join events, outside any capture.
    struct GraphEntry {
        int nb;
        hipStream_t st;
        int ng;
        hipGraphExec_t exec[kMaxStreams];
    };
    std::vector<GraphEntry> graphs_;
    void clear_graphs();
    bool use_graph_ = true;
    bool pencil_ct_ = true;  // line kernels compiled for the axis length where one exists
    bool rlines_ = true;     // split line step: register-FFT y and z passes (prop_rlines.hpp)
    int m_[3] = {0, 0, 0};
    double D_[3] = {0, 0, 0};
    long long M_ = 0;
    int N_ = 0, nbmax_ = 0, nblk_ = 0;
    const KernelSet<T>* ks_ = nullptr;
    hipError_t (*setup_)() = nullptr;
    void (*begin_)(const PropArgs<T>&, hipStream_t) = nullptr;
    void (*step_)(const PropArgs<T>&, int s, hipStream_t) = nullptr;
    void (*resume_)(const PropArgs<T>&, int c, hipStream_t) = nullptr;
    int ck_ = 0;      // checkpoint interval (0: every contour index kept)
    int coltile_ = 0;  // two-stage column tile width (0: the grid's default)
    int pgrid_ = 0;    // workgroups of the persistent plane kernels (0: one per plane)
    int planent_ = 0;  // two-stage plane workgroup size (0: the grid's default)
    int qpitch_ = 0;  // q-store entries per slot
    const char* path_ = "";
    cpx<T>* Ch_ = nullptr;
    cpx<T>* Cf_ = nullptr;
    T* e4A_ = nullptr;
    T* e4B_ = nullptr;
    cpx<T>* qs_ = nullptr;
    cpx<T>* tw_[3] = {nullptr, nullptr, nullptr};
    T* g_[3] = {nullptr, nullptr, nullptr};
    double* part_ = nullptr;
    // box-gradient spectra and partial sums (allocated on first use)
    cpx<T>* Gh_ = nullptr;  // two-stage gradient work fields (2 nbmax slots each)
    cpx<T>* Gf_ = nullptr;
    cpx<T>* K0_ = nullptr;
    cpx<T>* K1_ = nullptr;
    cpx<T>* Kf_ = nullptr;
    double* gpart_ = nullptr;
    double* cAB_ = nullptr;
    double* scal_ = nullptr;
    double* rn_ = nullptr;  // FP32 normalisation factors [nbmax][3] (k_renorm_factors)
    size_t bytes_ = 0;
    bool timing_ = false;
    hipEvent_t ev_[2] = {nullptr, nullptr};
    int nstreams_ = 1;
    hipStream_t sx_[kMaxStreams - 1] = {};
    hipEvent_t fork_ = nullptr;
    hipEvent_t join_[kMaxStreams - 1] = {};
};

}  // namespace scft
