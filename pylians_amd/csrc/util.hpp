// Zero-fill and workspace helpers.
#pragma once
#include "common.hpp"

namespace pyl {

int zero_fill(void* ptr, size_t bytes, hipStream_t st);
// one entry of a batched zero-fill (k_zero_b: buffer blockIdx.y)
struct ZeroArgs {
    void* ptr;
    size_t bytes;
};
__global__ void k_zero_b(const ZeroArgs* __restrict__ tab);

// out[0] = float64 sum of part[0 .. nb) by one workgroup (the second step of per-workgroup
// partial sums, in place of one same-address atomic per workgroup)
int sum_parts(const double* part, int nb, double* out, hipStream_t st);

// Library-owned grow-only device workspace, one buffer per (device, slot).
enum WsSlot { WS_SPEC = 0, WS_SPEC2 = 1, WS_SUMS = 2, WS_GRID = 3, WS_BINS = 4, WS_AUX = 5, WS_AUX2 = 6,
              WS_AUX3 = 7, WS_HOST_POS = 8, WS_HOST_GRID = 9, WS_HOST_W = 10, WS_OUT = 11, WS_MASS = 12,
              WS_PART = 13, WS_STAGE = 14, WS_SUP = 15, WS_C2R = 16,
              WS_RED = 17, WS_BATCH = 18, WS_BATCH_TABLES = 19 };
void* workspace(int slot, size_t bytes, int* rc);
// Library-owned grow-only pinned host buffer, one per device, for the results a call copies back
// before it returns (one asynchronous copy, then a stream synchronisation).  Growing synchronises the
// device before freeing the old buffer.
void* host_pinned(size_t bytes, int* rc);

// A private set of workspaces (one buffer per slot) owned by a captured hipGraph.  While an
// ArenaScope is alive on a thread, workspace() on that thread serves the arena's buffers instead of
// the shared ones, so a graph keeps valid pointers whatever later eager calls do to the shared slots.
// A sealed arena never grows (a capture must not reallocate a buffer an earlier node references):
// a request beyond its size fails with E_ALLOC.
struct WsArena;
WsArena* ws_arena_create();
void ws_arena_seal(WsArena* a);
// frees every buffer (synchronises the device first); the graph using it must be destroyed before
void ws_arena_destroy(WsArena* a);
size_t ws_arena_bytes(const WsArena* a);
// A fixed arena over device memory the caller owns (the no-allocation entry points): slots and the
// lookup tables (twiddles, windows, mode geometry) are carved from [base, base + cap) in request
// order, and nothing is allocated; a request that does not fit fails with E_ALLOC.  Reset before each
// call; tables are uploaded again on every call, on stream st.
WsArena* ws_arena_fixed(void* base, size_t cap);
void ws_arena_fixed_reset(WsArena* a, hipStream_t st);
// true while a fixed arena is the current one on this thread
bool arena_fixed();
// copies `bytes` of host data into a fresh region of the current fixed arena (hipMemcpyAsync on
// its stream; the host data must stay alive -- the table caches keep it)
void* arena_upload(const void* host, size_t bytes, int* rc);
struct ArenaScope {
    explicit ArenaScope(WsArena* a);
    ~ArenaScope();
    WsArena* prev;
};

}  // namespace pyl
