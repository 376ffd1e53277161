// scft_host.hpp -- C++ host-side mirror of the reference's operator interface, built on the
// C-ABI (include/scft_api.h). Names and argument meaning follow density.h / scft.cu; errors
// throw scft_host::error (the reference prints and exits, GPU_kernels.h:16-48).
#pragma once
#include <stdexcept>
#include <string>
#include <vector>

#include "scft_api.h"

namespace scft_host {

struct error : std::runtime_error {
    scft_status code;
    error(scft_status c, const std::string& m) : std::runtime_error(m), code(c) {}
};

inline void check(scft_ctx* ctx, scft_status s, const char* what)
{
    if (s != SCFT_OK) throw error(s, std::string(what) + ": " + scft_last_error(ctx));
}

// Mirrors class density (density.h:225-548): one device context for every replica. The
// reference's FreeE, solve_field0 and solve_field (scft.cu:516-853) call only props and the
// spline methods below, so they compile against this class unchanged
// (dropin_example.cpp in this folder is such a caller).
class density {
public:
    explicit density(const scft_params& p) : p_(p)
    {
        scft_status s = scft_ctx_create(&p_, &ctx_);
        if (s != SCFT_OK) throw error(s, "scft_ctx_create failed (grid/precision unsupported or no device)");
        M_ = (long long)p.m[0] * p.m[1] * p.m[2];
    }
    ~density() { scft_ctx_destroy(ctx_); }
    density(const density&) = delete;
    density& operator=(const density&) = delete;

    // density::props (density.h:344): w[0:2M] in, w[2M:6M] = phiA, phiB, phih, phicB out.
    // Ns, alpha, phic and f are fixed per context; they are checked against it here.
    void props(double* w, double* lnQ, int /*ii*/, int Ns, double alpha, double phic, double* phis, double f)
    {
        if (Ns != p_.N || alpha != p_.alpha || phic != p_.phidb || (int)(Ns * f + 0.5) != (int)(Ns * p_.f + 0.5))
            throw error(SCFT_ERR_INVALID, "props: chain parameters differ from the context's");
        std::vector<double> phi(4 * M_);
        double ph[5];
        check(ctx_, scft_props_batch(ctx_, w, phi.data(), lnQ, ph, 1), "props");
        std::copy(phi.begin(), phi.end(), w + 2 * M_);
        for (int i = 0; i < 5; ++i) phis[i] = ph[i];
    }

    // The string's spline on host rows (density.h:462-528), same signatures. As in the
    // reference the kernels take the nodes and evaluation points from setSplineX /
    // setSplineXRef; the x and xref arguments of the chunk calls are not read.
    void setupSplineBuffers(int chunkM) { check(ctx_, scft_spline_setup(ctx_, chunkM), "setupSplineBuffers"); }
    void setSplineX(double* x) { check(ctx_, scft_spline_set_x(ctx_, x), "setSplineX"); }
    void setSplineXRef(double* xref) { check(ctx_, scft_spline_set_xref(ctx_, xref), "setSplineXRef"); }
    void splineDerivsChunked(double** ww, double* /*x*/, int r0, int chunkSize, int* doiis, double** out)
    {
        check(ctx_, scft_spline_derivs(ctx_, ww, r0, chunkSize, doiis, out), "splineDerivsChunked");
    }
    void splineRedistChunked(double** ww, double* /*x*/, double* /*xref*/, int r0, int chunkSize, int* doiis,
                             const double lambda)
    {
        check(ctx_, scft_spline_redist(ctx_, ww, r0, chunkSize, doiis, lambda), "splineRedistChunked");
    }

    scft_ctx* ctx() { return ctx_; }
    long long M() const { return M_; }
    const scft_params& params() const { return p_; }

private:
    scft_params p_;
    scft_ctx* ctx_ = nullptr;
    long long M_ = 0;
};

}  // namespace scft_host
