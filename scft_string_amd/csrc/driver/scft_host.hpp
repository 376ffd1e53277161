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

// Mirrors class density (density.h:225-548): one device context for every replica.
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

    scft_ctx* ctx() { return ctx_; }
    long long M() const { return M_; }
    const scft_params& params() const { return p_; }

private:
    scft_params p_;
    scft_ctx* ctx_ = nullptr;
    long long M_ = 0;
};

}  // namespace scft_host
