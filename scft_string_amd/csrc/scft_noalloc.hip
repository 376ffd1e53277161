// scft_noalloc.hip -- the one entry point that allocates no device memory.
//
// scft_spline_derivs_noalloc: the spline tangent of scft_spline_derivs (splineDerivsChunked,
// density.h:491-509) on device rows, for callers that cannot call hipMalloc (code under HIP
// graph capture, serving loops, harnesses that forbid it). Its only device memory besides
// the caller's rows is `scratch` below: the weight matrix and row list of one call.
#include <mutex>
#include <string>

#include "field_kernels.hpp"
#include "scft_api.h"

namespace scft {

namespace {

// weights G [ne][R] then the row list (ints) of one call: at most 64 replicas
constexpr int NOALLOC_G = SPLINE_RMAX * SPLINE_RMAX;
constexpr int NOALLOC_ROWS = SPLINE_RMAX / 2;  // doubles holding SPLINE_RMAX ints
static __device__ __align__(256) double scratch[NOALLOC_G + NOALLOC_ROWS];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
double* g_scratch[kMaxDevices] = {};  // device address of `scratch` per device (set-up below)
thread_local std::string g_err;

scft_status refuse(scft_status s, const std::string& msg)
{
    g_err = msg;
    return s;
}

// Device address of this GPU's copy of `scratch`, looked up once per device with that
// device current (the host-side name is not a device address).
hipError_t scratch_of_current_device(double** out)
{
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_scratch[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        e = row_mix_setup();  // the row-mix kernel's LDS bound (an attribute, no memory)
        if (e != hipSuccess) return e;
        g_scratch[dev] = static_cast<double*>(p);
    }
    *out = g_scratch[dev];
    return hipSuccess;
}

struct SplineNodes {
    double x[SPLINE_RMAX];
    int rows[SPLINE_RMAX];
};

// one thread per listed row: its weights (derivative at its own node) and its row index
__global__ void k_spline_weights(SplineNodes nd, int R, int ne, double* G, int* erows)
{
    const int j = threadIdx.x;
    if (j >= ne) return;
    erows[j] = nd.rows[j];
    spline_weight_row(R, nd.x, nd.x[nd.rows[j]], true, G + (size_t)j * R);
}

}  // namespace

const char* noctx_error() { return g_err.c_str(); }

}  // namespace scft

extern "C" scft_status scft_spline_derivs_noalloc(int nrep, long long M, const double* x, const int* doiis,
                                                  const double* dW, long long wstride, double* dout, long long ostride,
                                                  void* stream)
{
    using namespace scft;
    if (nrep > SPLINE_RMAX)
        return refuse(SCFT_ERR_UNSUPPORTED, "scft_spline_derivs_noalloc: nrep = " + std::to_string(nrep) +
                                                " exceeds the static scratch (" + std::to_string(SPLINE_RMAX) +
                                                " replicas)");
    if (nrep < 2 || M <= 0 || !x || !doiis || !dW || !dout || wstride < M || ostride < M)
        return refuse(SCFT_ERR_INVALID, "scft_spline_derivs_noalloc: bad arguments");
    SplineNodes nd;
    int ne = 0;
    for (int i = 0; i < nrep; ++i) {
        if (i + 1 < nrep && !(x[i + 1] > x[i]))
            return refuse(SCFT_ERR_INVALID, "scft_spline_derivs_noalloc: x must increase strictly");
        nd.x[i] = x[i];
        if (doiis[i] == 1) nd.rows[ne++] = i;
    }
    if (ne == 0) return SCFT_OK;
    double* s = nullptr;
    hipError_t e = scratch_of_current_device(&s);
    if (e != hipSuccess)
        return refuse(SCFT_ERR_HIP, std::string("scft_spline_derivs_noalloc: scratch set-up: ") + hipGetErrorString(e));
    double* G = s;
    int* rows = reinterpret_cast<int*>(s + NOALLOC_G);
    hipStream_t st = static_cast<hipStream_t>(stream);
    hipLaunchKernelGGL(k_spline_weights, dim3(1), dim3(SPLINE_RMAX), 0, st, nd, nrep, ne, G, rows);
    launch_row_mix(nrep, M, G, rows, ne, dW, wstride, dout, ostride, st);
    e = hipGetLastError();
    if (e != hipSuccess)
        return refuse(SCFT_ERR_HIP, std::string("scft_spline_derivs_noalloc: launch: ") + hipGetErrorString(e));
    return SCFT_OK;
}
