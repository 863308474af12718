// predicates_dev.hip -- device build of the predicate code for the GPU
// predicate test: one thread per input row evaluates the exact paths on the
// GPU, so that device code generation is checked against exact rational
// arithmetic.  Test infrastructure (built by __graft_entry__.build).
#include <hip/hip_runtime.h>

#include "../../libqdt_amd/csrc/predicates.hpp"

namespace
{

// row i: ax ay bx by cx cy dx dy -> out[4i..4i+3] = orient2d_exact(a, b, c),
// orient_general(a, b, c), incircle_exact(a, b, c, d), incircle_general(a, b, c, d)
__global__ void k_pred(const double* in, int n, int* out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double* r = in + 8 * (size_t)i;
    out[4 * i] = qdt::orient2d_exact(r[0], r[1], r[2], r[3], r[4], r[5]);
    double hi[6], lo[6];
    qdt::two_sum(r[0], -r[4], hi[0], lo[0]);
    qdt::two_sum(r[1], -r[5], hi[1], lo[1]);
    qdt::two_sum(r[2], -r[4], hi[2], lo[2]);
    qdt::two_sum(r[3], -r[5], hi[3], lo[3]);
    double h4[4] = { hi[0], hi[1], hi[2], hi[3] }, l4[4] = { lo[0], lo[1], lo[2], lo[3] };
    out[4 * i + 1] = qdt::orient_general(h4, l4);
    out[4 * i + 2] = qdt::incircle_exact(r[0], r[1], r[2], r[3], r[4], r[5], r[6], r[7]);
    for (int k = 0; k < 3; ++k)
    {
        qdt::two_sum(r[2 * k], -r[6], hi[2 * k], lo[2 * k]);
        qdt::two_sum(r[2 * k + 1], -r[7], hi[2 * k + 1], lo[2 * k + 1]);
    }
    out[4 * i + 3] = qdt::incircle_general(hi, lo);
}

} // namespace

extern "C" __attribute__((visibility("default"))) int pd_run(const double* h_in, int n, int* h_out)
{
    if (n <= 0) return 0;
    double* d_in = nullptr;
    int* d_out = nullptr;
    if (hipMalloc(&d_in, 64 * (size_t)n) != hipSuccess) return 1;
    if (hipMalloc(&d_out, 16 * (size_t)n) != hipSuccess)
    {
        (void)hipFree(d_in);
        return 1;
    }
    int rc = hipMemcpy(d_in, h_in, 64 * (size_t)n, hipMemcpyHostToDevice) != hipSuccess;
    if (!rc)
    {
        k_pred<<<(n + 127) / 128, 128>>>(d_in, n, d_out);
        rc = hipGetLastError() != hipSuccess;
    }
    if (!rc) rc = hipMemcpy(h_out, d_out, 16 * (size_t)n, hipMemcpyDeviceToHost) != hipSuccess;
    (void)hipFree(d_in);
    (void)hipFree(d_out);
    return rc;
}
