// Argument blocks of the cell-binned deposit kernels (binned.hip), shared with the batch planner.
#pragma once
#include "common.hpp"
#include "deposit.hpp"

namespace pyl {

constexpr int TILE = 16;    // tile edge (cells)
constexpr int XCDS = 8;     // XCDs of gfx950; workgroup i runs on XCD i % XCDS
constexpr int SUPMAX = 1024;   // presort: most supertiles

struct BinArgs {
    const float* pos;
    const float* W;
    long long n;
    int N, nt;           // grid edge, tiles per edge
    double inv;          // N / BoxSize
    unsigned* fill;      // [ntiles * 8] records per (tile, class)
    float* bucket;       // [ntiles * CT * R]
    long long CT;        // record slots per tile
    int cofs[8], cap[8]; // first slot and capacity of each class segment
    unsigned* ov_count;  // [1]
    long long* ov_idx;   // [n]
    unsigned* ov_key;    // [n]
    Rsd rsd;             // redshift-space shift of the positions (rsd.vel == nullptr: none)
    int agg = 0;         // runs of equal keys in consecutive lanes counted once (bin_agg)
};

struct TileArgs {
    const float* bucket;
    const unsigned* fill;
    long long CT;
    int cofs[8], cap[8];
    const float* grouped;
    const unsigned* gcls;
    const unsigned* ooff;
    const unsigned* ocnt;
    int N, nt;
    double inv;
    double mult;          // weight multiplier
    void* grid;
    const double* dmass;  // device mass (nullable) for OUT_DELTA
    double mass;
};

struct OvArgs {
    const unsigned* ov_count;   // [1] records in the overflow list
    const long long* ov_idx;    // particle index of each
    const unsigned* ov_key;     // (tile, class) key of each
    unsigned* ocnt;             // [ntiles] overflow records per tile
    unsigned* ooff;             // [ntiles + 1] first grouped slot of each tile
    unsigned* ocur;             // [ntiles] slots handed out
    const float* pos;
    const float* W;
    float* grouped;             // records grouped by tile
    unsigned* gcls;             // class of each grouped record
    Rsd rsd;
    int ntiles;
};

struct SupArgs {
    const float* pos;
    const float* W;
    long long n;
    int N, nt, tps, nsup;
    double inv;
    Rsd rsd;
    unsigned* cnt;   // [nsup] particles per supertile
    unsigned* off;   // [nsup] first slot of each supertile
    unsigned* cur;   // [nsup] slots handed out so far
    float* spos;     // [n * 3] grouped positions (redshift-space shift applied)
    float* sw;       // [n] grouped weights (HASW)
};

// One job of a batched binned deposit (overdensity output): the arguments of every stage and
// the block count of each 1-D launch.  binned_job_bytes gives the device memory one job needs;
// binned_job_plan carves it from `mem` and fills the arguments.  cnt / scnt: counters to zero
// before the deposit.
struct BinnedJob {
    BinArgs bin;
    TileArgs tile;
    OvArgs ov;
    SupArgs sup;
    unsigned bin_blocks, tile_blocks, sup_blocks, sup_sblocks;
    void* cnt;
    size_t cnt_bytes;
    void* scnt;
    size_t scnt_bytes;
};
size_t binned_job_bytes(int64_t n, int dims, int scheme, bool hasw, bool sorted);
int binned_job_plan(const float* pos, const float* W, int64_t n, int dims, double BoxSize, int scheme, void* grid,
                    const double* dmass, double mass, bool sorted, char* mem, BinnedJob* job);
// Device tables of one class of jobs (same scheme, weights, grid type, accumulation, presort):
// per-stage argument arrays and the block ranges of the 1-D launches (start[nj] = total).
struct BinnedTables {
    int nj;
    const SupArgs* sup;
    const unsigned *sup_start, *ssc_start;
    unsigned sup_total, ssc_total;
    const BinArgs* bin;
    const unsigned* bin_start;
    unsigned bin_total;
    const OvArgs* ov;
    const TileArgs* tile;
    const unsigned* tile_start;
    unsigned tile_total;
};
int binned_batch_launch(int scheme, bool hasw, bool f64, int acc, bool sorted, const BinnedTables& t, hipStream_t st);
// accumulation mode of a binned deposit: `acc` (0: float64 weights, 1: float32 weights), or 2 (float64
// weights with register runs of the stencil block) for CIC inputs with at least RUN_PPC
// particles per cell
int dense_acc(int acc, int scheme, int64_t n, int dims);
// 1 when k_bin counts runs of equal keys in consecutive lanes once (BinArgs::agg): n >= 2 dims^3
int bin_agg(int64_t n, int dims);

}  // namespace pyl
