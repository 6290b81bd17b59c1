// sin/cos of the host C library on the device (prh_libm.h), exposed for
// verification: the same source evaluated on the host (no GPU needed) and on
// the device must give the C library's bits, which generate_rays
// (porerad/view_factors.py:168-176) takes through NumPy.
#include "prh_internal.h"
#include "prh_libm.h"

namespace prh {
namespace {

__global__ __launch_bounds__(256) void k_libm_sincos(const double *x, int64_t n, double *s, double *c)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    double sv, cv;
    libm::sincos(x[i], sv, cv);
    s[i] = sv;
    c[i] = cv;
}

}  // namespace
}  // namespace prh

using namespace prh;

extern "C" int prh_libm_sincos(const double *x, int64_t n, int32_t on_device, double *sin_x, double *cos_x)
{
    PRH_CHECK_ARG(n >= 0 && (n == 0 || (x && sin_x && cos_x)), "prh_libm_sincos: null pointer or negative count");
    for (int64_t i = 0; i < n; ++i)
        PRH_CHECK_ARG(std::isfinite(x[i]) && fabs(x[i]) < libm::SINCOS_MAX_ARG,
                      "prh_libm_sincos: argument %lld (%g) outside |x| < %g", (long long)i, x[i],
                      libm::SINCOS_MAX_ARG);
    if (n == 0) return PRH_OK;
    if (!on_device) {
        for (int64_t i = 0; i < n; ++i) libm::sincos(x[i], sin_x[i], cos_x[i]);
        return PRH_OK;
    }
    hipStream_t st;
    int rc = current_stream(&st);
    if (rc) return rc;
    double *d = nullptr;
    PRH_HIP(hipMalloc((void **)&d, 3 * n * sizeof(double)));
    hipError_t e = hipMemcpyAsync(d, x, n * sizeof(double), hipMemcpyHostToDevice, st);
    if (e == hipSuccess) {
        k_libm_sincos<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(d, n, d + n, d + 2 * n);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(sin_x, d + n, n * sizeof(double), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipMemcpyAsync(cos_x, d + 2 * n, n * sizeof(double), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    (void)hipFree(d);
    PRH_HIP(e);
    return PRH_OK;
}
