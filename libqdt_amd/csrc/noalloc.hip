// noalloc.hip -- qdt_gpu_execute_noalloc: the batch triangulation of
// device-resident points for callers that cannot allocate device memory
// (code under HIP graph capture, tight serving loops, harnesses that forbid
// hipMalloc).  It runs the same pipeline as qdt_gpu_execute_device, with the
// engine's whole workspace carved from one file-scope device array instead
// of hipMalloc'd regions; the caller passes the input and the outputs.
#include "../../include/qdt_gpu.h"
#include "engine.hpp"

#include <hip/hip_runtime.h>

#include <mutex>
#include <new>

namespace
{

constexpr size_t kNoAllocMaxPoints = (size_t)1 << 18; // 262,144 points per call at most
constexpr size_t kScratchBytes = (size_t)112 << 20;   // ~250 B per point, the kd forest and seam buffers
constexpr int kMaxDevices = 64;

} // namespace

// The workspace of every call on a GPU: each GPU the library runs on holds
// its own copy until the process ends.  Host code must not pass this name to
// a kernel or a copy (on the host it only stands for the symbol); the
// device address comes from hipGetSymbolAddress at set-up.
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace
{

struct DeviceSlot
{
    std::once_flag once;
    std::mutex mu;           // calls on one GPU share its scratch array
    void* ptr = nullptr;     // device address of this GPU's scratch array
    qdt::Engine* eng = nullptr;
    hipError_t err = hipSuccess;
};

DeviceSlot g_slots[kMaxDevices];

// Set-up on the first call on a device (that device current): the symbol's
// device address, and an engine that carves everything from it.
DeviceSlot* slot_for_current_device(qdt_status& st)
{
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices)
    {
        st = QDT_ERR_INTERNAL;
        return nullptr;
    }
    DeviceSlot& d = g_slots[dev];
    std::call_once(d.once, [&d] {
        void* p = nullptr;
        d.err = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (d.err != hipSuccess || !p)
        {
            if (d.err == hipSuccess) d.err = hipErrorInvalidSymbol;
            return;
        }
        d.ptr = p;
        d.eng = new (std::nothrow) qdt::Engine(nullptr, p, kScratchBytes);
    });
    if (d.err != hipSuccess || !d.eng)
    {
        st = QDT_ERR_INTERNAL;
        return nullptr;
    }
    st = QDT_OK;
    return &d;
}

} // namespace

extern "C" qdt_status qdt_gpu_execute_noalloc(const qdt_vec2d* d_pts, size_t n, uint32_t* d_tris,
                                              size_t tri_capacity, uint32_t* d_rep, size_t* triangles)
{
    if (n > kNoAllocMaxPoints) return QDT_ERR_TOO_MANY_POINTS;
    if (!triangles || (n && (!d_pts || !d_tris || !d_rep))) return QDT_ERR_NULL_ARGUMENT;
    *triangles = 0;
    if (n == 0) return QDT_OK;
    if (tri_capacity < 2 * n) return QDT_ERR_OUT_OF_MEMORY;
    qdt_status st = QDT_OK;
    DeviceSlot* d = slot_for_current_device(st);
    if (!d) return st;
    std::lock_guard<std::mutex> lk(d->mu);
    const qdt::Status s = d->eng->run((const double*)d_pts, n);
    if (s != qdt::Status::ok)
        return s == qdt::Status::not_finite ? QDT_ERR_NOT_FINITE
                                            : (s == qdt::Status::out_of_memory ? QDT_ERR_OUT_OF_MEMORY : QDT_ERR_INTERNAL);
    const qdt::DeviceResult& r = d->eng->result();
    hipStream_t stream = d->eng->stream();
    if (r.triangles &&
        hipMemcpyAsync(d_tris, r.tris, 12 * (size_t)r.triangles, hipMemcpyDeviceToDevice, stream) != hipSuccess)
        return QDT_ERR_INTERNAL;
    if (hipMemcpyAsync(d_rep, r.rep, 4 * n, hipMemcpyDeviceToDevice, stream) != hipSuccess ||
        hipStreamSynchronize(stream) != hipSuccess)
        return QDT_ERR_INTERNAL;
    *triangles = r.triangles;
    return QDT_OK;
}
