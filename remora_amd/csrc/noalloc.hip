// Entry point that allocates nothing: rmh_dpotrf_noalloc, rmh_dpotrf on a
// static device scratch array instead of the handle's workspace, for callers
// that cannot allocate device memory (code under HIP graph capture, tight
// serving loops, harnesses that forbid hipMalloc).
//
// `scratch` is a file-scope __device__ array: every GPU the library runs on
// holds its own copy until the process ends.  In host code its name is only a
// stand-in; rmh_create resolves the device address of this device's copy once
// (hipGetSymbolAddress, with the device current) and every use below goes
// through that address.  All calls on one GPU share the array: calls that
// overlap on one GPU (two streams, two threads) race and must be ordered by
// the caller.
#include <mutex>
#include <string>
#include <vector>

#include "rmh_context.hpp"

namespace rmh {

// 8 MiB of side buffer plus the info word's 256 bytes: potrf of order n with
// nb_inner-wide diagonal blocks needs 256 + n * nb_inner * 8 bytes, so
// n <= 32768 at nb_inner 32 (the default) and n <= 16384 at nb_inner 64.
constexpr size_t kScratchBytes = ((size_t)8 << 20) + 256;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {
std::mutex g_scratch_mu;
std::vector<void*> g_scratch_ptr;  // per device; nullptr until resolved
}  // namespace

rmh_status noalloc_scratch_init(int device) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if ((size_t)device < g_scratch_ptr.size() && g_scratch_ptr[device] != nullptr) return RMH_STATUS_SUCCESS;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || p == nullptr)
        return fail(RMH_STATUS_HIP_ERROR, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    if ((size_t)device >= g_scratch_ptr.size()) g_scratch_ptr.resize((size_t)device + 1, nullptr);
    g_scratch_ptr[device] = p;
    return RMH_STATUS_SUCCESS;
}

void* noalloc_scratch(int device, size_t* bytes) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    *bytes = kScratchBytes;
    return (size_t)device < g_scratch_ptr.size() ? g_scratch_ptr[device] : nullptr;
}

}  // namespace rmh

using namespace rmh;

extern "C" rmh_status rmh_dpotrf_noalloc(rmh_handle h, int upper, size_t n, double* A, size_t lda, int64_t* d_info,
                                         rmh_stream_t stream) {
    if (h == nullptr) return fail(RMH_STATUS_INVALID_HANDLE, "NULL handle");
    Context* c = reinterpret_cast<Context*>(h);
    RMH_TRY(c->activate());
    if (lda < (n > 1 ? n : 1)) return fail(RMH_STATUS_INVALID_VALUE, "potrf_noalloc: lda too small");
    if (n != 0 && A == nullptr) return fail(RMH_STATUS_INVALID_VALUE, "potrf_noalloc: A is NULL");
    size_t bytes = 0;
    void* ws = noalloc_scratch(c->device, &bytes);
    if (ws == nullptr) return fail(RMH_STATUS_INVALID_HANDLE, "potrf_noalloc: scratch not resolved for this device");
    const size_t max_n = (bytes - 256) / (c->nbi * sizeof(double));
    if (n > max_n)
        return fail(RMH_STATUS_INVALID_VALUE,
                    "potrf_noalloc: order " + std::to_string(n) + " above the maximum " + std::to_string(max_n) +
                        " for nb_inner " + std::to_string(c->nbi) + " (the static scratch holds " +
                        std::to_string(bytes) + " bytes)");
    return potrf_core<double>(c, upper != 0, (int64_t)n, A, (int64_t)lda, d_info, ws, true,
                              reinterpret_cast<hipStream_t>(stream));
}
