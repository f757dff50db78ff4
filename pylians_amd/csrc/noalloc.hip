// No-allocation variant of pyl_pk (P(k) of a device field) for callers that cannot allocate
// device memory: all of its scratch lives in a fixed per-GPU array.
#include <map>
#include <mutex>

#include "../../include/pylians_hip.h"
#include "common.hpp"
#include "fft.hpp"
#include "pk.hpp"
#include "util.hpp"

using namespace pyl;

// ------------------------------------------------------------------ no-allocation power spectrum
// pyl_pk_noalloc: pyl_pk for callers that cannot allocate device memory.  Everything pyl_pk keeps
// in library workspaces and cached tables -- the half spectrum, the float64 sums and mode counts,
// the per-workgroup partial histograms, the twiddles, the window table and the mode geometry --
// is carved from `scratch`, a fixed per-GPU array, and the tables are uploaded into it on every
// call.  Its address on a device comes from hipGetSymbolAddress on the first call made with that
// device current (host code never passes the array's name to a kernel).  Every call on a GPU
// shares that GPU's array and leaves its contents there; calls that overlap on one GPU (another
// host thread, or an earlier call still running on a different stream) race on it.  Calls from
// host threads are serialised per device.
constexpr size_t SCRATCH_BYTES = (size_t)PYL_NOALLOC_SCRATCH_MB << 20;
static __device__ __align__(256) unsigned char scratch[SCRATCH_BYTES];

namespace {
struct ScratchDev {
    void* ptr = nullptr;
    WsArena* arena = nullptr;
    std::mutex call;
};
std::mutex g_smu;
std::map<int, ScratchDev*> g_scratch;
}  // namespace

static int scratch_of(int dev, ScratchDev** out) {
    std::lock_guard<std::mutex> lk(g_smu);
    auto it = g_scratch.find(dev);
    if (it != g_scratch.end()) {
        *out = it->second;
        return OK;
    }
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    PYL_CHECK(e == hipSuccess && ptr, E_LAUNCH, "hipGetSymbolAddress(scratch) failed on device %d: %s", dev,
              hipGetErrorString(e));
    ScratchDev* sd = new ScratchDev();
    sd->ptr = ptr;
    sd->arena = ws_arena_fixed(ptr, SCRATCH_BYTES);
    g_scratch[dev] = sd;
    *out = sd;
    return OK;
}

extern "C" int pyl_pk_noalloc(const void* delta, int dims, double BoxSize, int axis, int mas_p, int prec, int flags,
                              double* out, void* deltak_out, void* stream) {
    PYL_CHECK(prec == PYL_F32 || prec == PYL_F64, E_ARG, "prec must be 4 or 8");
    PYL_CHECK(fft_size_supported(dims), E_UNSUPPORTED, "grid size %d not supported", dims);
    PYL_CHECK(dims <= PYL_NOALLOC_MAX_DIMS, E_UNSUPPORTED,
              "pyl_pk_noalloc: dims %d exceeds the %d its %zu-byte scratch holds (use pyl_pk)", dims,
              PYL_NOALLOC_MAX_DIMS, SCRATCH_BYTES);
    PYL_CHECK(axis >= 0 && axis <= 2, E_ARG, "axis must be 0, 1 or 2");
    PYL_CHECK(mas_p >= 0 && mas_p <= 4, E_ARG, "window exponent must be in [0, 4]");
    PYL_CHECK(BoxSize > 0, E_ARG, "BoxSize must be positive");
    PYL_CHECK(delta && out, E_ARG, "null pointer");
    int dev = 0;
    PYL_HIP(hipGetDevice(&dev));
    ScratchDev* sd = nullptr;
    int rc = scratch_of(dev, &sd);
    if (rc) return rc;
    hipStream_t st = (hipStream_t)stream;
    std::lock_guard<std::mutex> lk(sd->call);
    ws_arena_fixed_reset(sd->arena, st);
    ArenaScope scope(sd->arena);
    if (prec == PYL_F32)
        return pk_from_delta<float>((const float*)delta, dims, BoxSize, axis, mas_p, flags, out,
                                    (cplx<float>*)deltak_out, false, nullptr, 1.0, st);
    return pk_from_delta<double>((const double*)delta, dims, BoxSize, axis, mas_p, flags, out,
                                 (cplx<double>*)deltak_out, false, nullptr, 1.0, st);
}

