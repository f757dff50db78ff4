// Batches of independent particles -> P(k) jobs (pyl_pk_batch, pyl_graph_create).
//
// A batch is planned once: every job gets its own grid, spectrum, histogram and deposit buffers,
// carved from one device allocation, and the jobs are grouped into classes that share kernel
// instantiations -- deposit classes (scheme, weights, precision, accumulation, presort) and
// transform classes (grid size, precision, line of sight along z or not).  Each stage then runs as
// ONE launch for all jobs of a class: the arguments of job j sit in a device table, and in the 1-D
// launches job j owns the block range [start[j], start[j+1]).  A batch of 64 uneven realizations
// (four grid sizes) runs in about 60 launches instead of about 16 per job.  Jobs whose deposit is
// the direct-atomic one (chosen by flags or by the autotuner), or whose grid is not a multiple of
// the 16-cell tile, run through the single-job path in the same stream order.
//
// The plan holds device pointers only into memory that outlives it: the allocation is a library
// workspace slot (the graph's own arena under pyl_graph_create), so a captured graph replays on
// buffers nothing else touches.  Builder-defined: the reference computes one field per call
// (Pk/Pk.py:17-27).
#include <functional>
#include <map>
#include <memory>
#include <tuple>
#include <vector>

#include "../../include/pylians_hip.h"
#include "binned.hpp"
#include "common.hpp"
#include "deposit.hpp"
#include "fft.hpp"
#include "pk.hpp"
#include "profile.hpp"
#include "util.hpp"

namespace pyl {
int weight_sum(const float* W, int64_t n, double* out, hipStream_t st);

namespace {

size_t al256(size_t b) { return (b + 255) / 256 * 256; }

struct JobPlan {
    pyl_pk_job job;
    int kind = DK_BINNED;
    bool batched = false;
    int64_t P = 0;
    void* grid = nullptr;
    void* spec = nullptr;
    double* sums_base = nullptr;
    double* part = nullptr;
    double* dmass = nullptr;
    char* dep = nullptr;
    PkSums sums{};
    const PkGeometry* geom = nullptr;
    BinnedJob bj{};
};

// one class of jobs: host copies of the per-job arguments, then device copies
template <typename A> struct Table {
    std::vector<A> args;
    std::vector<unsigned> start{0u};
    A* dev = nullptr;
    unsigned* dstart = nullptr;
    void add(const A& a, unsigned blocks) {
        args.push_back(a);
        start.push_back(start.back() + blocks);
    }
    unsigned total() const { return start.back(); }
    int nj() const { return (int)args.size(); }
    size_t bytes() const { return al256(sizeof(A) * args.size()) + al256(sizeof(unsigned) * start.size()); }
    char* place(char* p) {
        dev = (A*)p;
        p += al256(sizeof(A) * args.size());
        dstart = (unsigned*)p;
        return p + al256(sizeof(unsigned) * start.size());
    }
    int upload() const {
        if (args.empty()) return OK;
        PYL_HIP(hipMemcpy(dev, args.data(), sizeof(A) * args.size(), hipMemcpyHostToDevice));
        PYL_HIP(hipMemcpy(dstart, start.data(), sizeof(unsigned) * start.size(), hipMemcpyHostToDevice));
        return OK;
    }
};

struct DepClass {
    int scheme, prec, acc;
    bool hasw, sorted;
    Table<SupArgs> sup, ssc;
    Table<BinArgs> bin;
    Table<OvArgs> ov;
    Table<TileArgs> tile;
};

template <typename R> struct FftClass {
    int N;
    Table<R2cArgs<R>> r2c;
    Table<ColArgs<R>> cols;
};

template <typename R> struct BinClass {
    int N, axis;
    Table<LastBinArgs<R>> last;
    Table<ReduceArgs> red;
};

struct FinClass {
    int N;
    Table<FinArgs> fin;
};

}  // namespace

struct BatchPlan {
    std::vector<JobPlan> jobs;
    std::vector<std::function<int(hipStream_t)>> steps;
    // classes (kept alive: the steps hold pointers to their device tables)
    std::map<std::tuple<int, int, int, int, int>, DepClass> dep;
    std::map<int, FftClass<float>> fft32;
    std::map<int, FftClass<double>> fft64;
    std::map<std::pair<int, int>, BinClass<float>> bin32;
    std::map<std::pair<int, int>, BinClass<double>> bin64;
    std::map<int, FinClass> fin;
    Table<ZeroArgs> zero;
};

static std::map<int, FftClass<float>>& fft_map(BatchPlan* B, float*) { return B->fft32; }
static std::map<int, FftClass<double>>& fft_map(BatchPlan* B, double*) { return B->fft64; }
static std::map<std::pair<int, int>, BinClass<float>>& bin_map(BatchPlan* B, float*) { return B->bin32; }
static std::map<std::pair<int, int>, BinClass<double>>& bin_map(BatchPlan* B, double*) { return B->bin64; }

template <typename R>
static int plan_spectrum(BatchPlan* B, JobPlan& J, char*& p) {
    const pyl_pk_job& j = J.job;
    const int N = j.dims;
    constexpr int64_t ALIGN = 128 / (int64_t)sizeof(cplx<R>);
    J.P = (half_len(N) + ALIGN - 1) / ALIGN * ALIGN;
    J.spec = p;
    p += al256(sizeof(cplx<R>) * (size_t)N * N * J.P);
    int rc = OK;
    J.geom = pk_geometry(N, j.axis, &rc);
    if (rc) return rc;
    J.sums_base = (double*)p;
    p += al256(sizeof(double) * pk_sums_doubles(J.geom->d));
    pk_sums_carve(J.geom->d, J.sums_base, &J.sums);
    J.part = (double*)p;
    p += al256(sizeof(double) * last_bin_part_doubles(N, j.axis, j.prec));
    B->zero.add(ZeroArgs{J.sums_base, sizeof(double) * pk_sums_doubles(J.geom->d)}, 0);
    const int flags = (j.flags & PYL_PK_ALL) | PK_ASYNC;
    R2cArgs<R> ra;
    ColArgs<R> ca;
    LastBinArgs<R> la;
    ReduceArgs rd;
    unsigned b1 = 0, b2 = 0, b3 = 0;
    rc = r2c_job<R>((const R*)J.grid, (cplx<R>*)J.spec, N, J.P, false, nullptr, 1.0, &ra, &b1);
    if (rc) return rc;
    rc = cols_job<R>((const cplx<R>*)J.spec, (cplx<R>*)J.spec, N, J.P, pk_last_axis(j.axis) == 0 ? 1 : 0, &ca, &b2);
    if (rc) return rc;
    rc = last_bin_job<R>((const cplx<R>*)J.spec, N, J.P, j.axis, j.scheme, flags, J.sums, J.part, &la, &rd, &b3);
    if (rc) return rc;
    FftClass<R>& F = fft_map(B, (R*)nullptr)[N];
    F.N = N;
    F.r2c.add(ra, b1);
    F.cols.add(ca, b2);
    const std::pair<int, int> bk(N, j.axis == 2 ? 1 : 0);
    BinClass<R>& C = bin_map(B, (R*)nullptr)[bk];
    C.N = N;
    C.axis = j.axis;
    C.last.add(la, b3);
    C.red.add(rd, 0);
    auto& Fi = B->fin[N];
    Fi.N = N;
    Fi.fin.add(finalize_args(J.geom, J.sums, j.BoxSize, flags, j.out), 0);
    return OK;
}

template <typename R>
static void plan_spectrum_steps(BatchPlan* B, std::map<int, FftClass<R>>& fft,
                                std::map<std::pair<int, int>, BinClass<R>>& bins) {
    for (auto& kv : fft) {
        FftClass<R>* F = &kv.second;
        B->steps.push_back([F](hipStream_t st) {
            {
                StageTimer t("fft_z_r2c", st);
                int rc = r2c_batch<R>(F->N, F->r2c.dev, F->r2c.dstart, F->r2c.nj(), F->r2c.total(), st);
                if (rc) return rc;
            }
            StageTimer t("fft_mid", st);
            return cols_batch<R>(F->N, F->cols.dev, F->cols.dstart, F->cols.nj(), F->cols.total(), st);
        });
    }
    for (auto& kv : bins) {
        BinClass<R>* C = &kv.second;
        B->steps.push_back([C](hipStream_t st) {
            StageTimer t("fft_last_bin", st);
            return last_bin_batch<R>(C->N, C->axis, C->last.dev, C->last.dstart, C->last.nj(), C->last.total(),
                                     C->red.dev, st);
        });
    }
}

static int batch_build(const pyl_pk_job* jobs, int njobs, hipStream_t st, BatchPlan** out) {
    BatchPlan* B = new BatchPlan();
    std::unique_ptr<BatchPlan> guard(B);
    B->jobs.resize(njobs);
    int rc = OK;
    // 1. validate; deposit algorithm per job (flags, or the autotuner: its trials run now, eagerly)
    for (int i = 0; i < njobs; ++i) {
        JobPlan& J = B->jobs[i];
        J.job = jobs[i];
        const pyl_pk_job& j = J.job;
        PYL_CHECK(j.pos && j.out && j.n > 0, E_ARG, "job %d: bad particle arguments", i);
        PYL_CHECK(j.prec == PYL_F32 || j.prec == PYL_F64, E_ARG, "job %d: prec must be 4 or 8", i);
        PYL_CHECK(fft_size_supported(j.dims), E_UNSUPPORTED, "job %d: grid size %d not supported", i, j.dims);
        PYL_CHECK(j.axis >= 0 && j.axis <= 2, E_ARG, "job %d: axis must be 0, 1 or 2", i);
        PYL_CHECK(j.scheme >= 1 && j.scheme <= 4, E_ARG, "job %d: unknown scheme %d", i, j.scheme);
        const bool f64 = j.prec == PYL_F64;
        if (!binned_supported(3, j.dims) || (j.flags & DEP_ATOMIC)) {
            J.kind = DK_ATOMIC;
        } else {
            const size_t cells = (size_t)j.dims * j.dims * j.dims;
            void* g = workspace(WS_GRID, cells * (f64 ? 8 : 4), &rc);
            if (rc) return rc;
            double* dm = nullptr;
            if (j.W) {
                dm = (double*)workspace(WS_MASS, 64, &rc);
                if (rc) return rc;
                rc = weight_sum(j.W, j.n, dm, st);
                if (rc) return rc;
            }
            rc = deposit_kind(j.pos, j.W, j.n, j.dims, j.BoxSize, j.scheme, f64, j.flags, g, dm, st, &J.kind, Rsd());
            if (rc) return rc;
        }
        J.batched = J.kind != DK_ATOMIC && j.n < (1LL << 31);
    }
    // 2. memory of every batched job
    size_t bytes = 0;
    for (auto& J : B->jobs) {
        if (!J.batched) continue;
        const pyl_pk_job& j = J.job;
        const bool f64 = j.prec == PYL_F64;
        const size_t cells = (size_t)j.dims * j.dims * j.dims;
        const int64_t P = f64 ? (half_len(j.dims) + 7) / 8 * 8 : (half_len(j.dims) + 15) / 16 * 16;
        bytes += al256(cells * (f64 ? 8 : 4)) + al256((size_t)j.dims * j.dims * P * (f64 ? 16 : 8));
        int r = OK;
        const PkGeometry* g = pk_geometry(j.dims, j.axis, &r);
        if (r) return r;
        bytes += al256(sizeof(double) * pk_sums_doubles(g->d)) + al256(sizeof(double) * last_bin_part_doubles(j.dims, j.axis, j.prec));
        bytes += j.W ? 256 : 0;
        bytes += al256(binned_job_bytes(j.n, j.dims, j.scheme, j.W != nullptr, J.kind == DK_SORTED));
    }
    // 3. carve, fill the per-job arguments, group into classes
    char* mem = nullptr;
    if (bytes) {
        mem = (char*)workspace(WS_BATCH, bytes, &rc);
        if (rc) return rc;
    }
    char* p = mem;
    for (auto& J : B->jobs) {
        if (!J.batched) continue;
        const pyl_pk_job& j = J.job;
        const bool f64 = j.prec == PYL_F64;
        const size_t cells = (size_t)j.dims * j.dims * j.dims;
        J.grid = p;
        p += al256(cells * (f64 ? 8 : 4));
        if (j.W) {
            J.dmass = (double*)p;
            p += 256;
        }
        rc = f64 ? plan_spectrum<double>(B, J, p) : plan_spectrum<float>(B, J, p);
        if (rc) return rc;
        J.dep = p;
        p += al256(binned_job_bytes(j.n, j.dims, j.scheme, j.W != nullptr, J.kind == DK_SORTED));
        rc = binned_job_plan(j.pos, j.W, j.n, j.dims, j.BoxSize, j.scheme, J.grid, J.dmass, (double)j.n,
                             J.kind == DK_SORTED, J.dep, &J.bj);
        if (rc) return rc;
        B->zero.add(ZeroArgs{J.bj.cnt, J.bj.cnt_bytes}, 0);
        if (J.bj.scnt) B->zero.add(ZeroArgs{J.bj.scnt, J.bj.scnt_bytes}, 0);
        const int acc = dense_acc(f64 ? 0 : acc_mode(j.flags), j.scheme, j.n, j.dims);
        auto key = std::make_tuple(j.scheme, j.prec, acc, j.W ? 1 : 0, J.kind == DK_SORTED ? 1 : 0);
        DepClass& D = B->dep[key];
        D.scheme = j.scheme;
        D.prec = j.prec;
        D.acc = acc;
        D.hasw = j.W != nullptr;
        D.sorted = J.kind == DK_SORTED;
        if (D.sorted) {
            D.sup.add(J.bj.sup, J.bj.sup_blocks);
            D.ssc.add(J.bj.sup, J.bj.sup_sblocks);
        }
        D.bin.add(J.bj.bin, J.bj.bin_blocks);
        D.ov.add(J.bj.ov, 0);
        D.tile.add(J.bj.tile, J.bj.tile_blocks);
    }
    // 4. device tables (a workspace slot of their own), uploaded now, before any capture
    size_t tb = B->zero.bytes();
    for (auto& kv : B->dep) tb += kv.second.sup.bytes() + kv.second.ssc.bytes() + kv.second.bin.bytes() +
                                  kv.second.ov.bytes() + kv.second.tile.bytes();
    for (auto& kv : B->fft32) tb += kv.second.r2c.bytes() + kv.second.cols.bytes();
    for (auto& kv : B->fft64) tb += kv.second.r2c.bytes() + kv.second.cols.bytes();
    for (auto& kv : B->bin32) tb += kv.second.last.bytes() + kv.second.red.bytes();
    for (auto& kv : B->bin64) tb += kv.second.last.bytes() + kv.second.red.bytes();
    for (auto& kv : B->fin) tb += kv.second.fin.bytes();
    // the previous contents of the workspaces may still be in use by queued work
    PYL_HIP(hipStreamSynchronize(st));
    if (tb) {
        p = (char*)workspace(WS_BATCH_TABLES, tb, &rc);
        if (rc) return rc;
    }
    p = B->zero.place(p);
    rc = B->zero.upload();
    if (rc) return rc;
    for (auto& kv : B->dep) {
        DepClass& D = kv.second;
        p = D.sup.place(p);
        p = D.ssc.place(p);
        p = D.bin.place(p);
        p = D.ov.place(p);
        p = D.tile.place(p);
        if ((rc = D.sup.upload()) || (rc = D.ssc.upload()) || (rc = D.bin.upload()) || (rc = D.ov.upload()) ||
            (rc = D.tile.upload()))
            return rc;
    }
    auto place_fft = [&](auto& m) {
        for (auto& kv : m) {
            p = kv.second.r2c.place(p);
            p = kv.second.cols.place(p);
            int r = kv.second.r2c.upload();
            if (!r) r = kv.second.cols.upload();
            if (r) return r;
        }
        return (int)OK;
    };
    auto place_bin = [&](auto& m) {
        for (auto& kv : m) {
            p = kv.second.last.place(p);
            p = kv.second.red.place(p);
            int r = kv.second.last.upload();
            if (!r) r = kv.second.red.upload();
            if (r) return r;
        }
        return (int)OK;
    };
    if ((rc = place_fft(B->fft32)) || (rc = place_fft(B->fft64)) || (rc = place_bin(B->bin32)) ||
        (rc = place_bin(B->bin64)))
        return rc;
    for (auto& kv : B->fin) {
        p = kv.second.fin.place(p);
        if ((rc = kv.second.fin.upload())) return rc;
    }
    // 5. the launch sequence
    if (B->zero.nj() > 0) {
        Table<ZeroArgs>* Z = &B->zero;
        B->steps.push_back([Z](hipStream_t st) {
            hipLaunchKernelGGL(k_zero_b, dim3(64, (unsigned)Z->nj()), dim3(256), 0, st, Z->dev);
            PYL_HIP(hipGetLastError());
            return (int)OK;
        });
    }
    for (auto& J : B->jobs) {
        if (!J.batched || !J.job.W) continue;
        const float* W = J.job.W;
        const int64_t n = J.job.n;
        double* dm = J.dmass;
        B->steps.push_back([W, n, dm](hipStream_t st) { return weight_sum(W, n, dm, st); });
    }
    for (auto& kv : B->dep) {
        DepClass* D = &kv.second;
        B->steps.push_back([D](hipStream_t st) {
            BinnedTables t{D->bin.nj(), D->sup.dev, D->sup.dstart, D->ssc.dstart, D->sup.total(), D->ssc.total(),
                           D->bin.dev, D->bin.dstart, D->bin.total(), D->ov.dev, D->tile.dev, D->tile.dstart,
                           D->tile.total()};
            return binned_batch_launch(D->scheme, D->hasw, D->prec == PYL_F64, D->acc, D->sorted, t, st);
        });
    }
    plan_spectrum_steps<float>(B, B->fft32, B->bin32);
    plan_spectrum_steps<double>(B, B->fft64, B->bin64);
    for (auto& kv : B->fin) {
        FinClass* F = &kv.second;
        B->steps.push_back([F](hipStream_t st) {
            StageTimer t("finalize", st);
            return finalize_batch(F->N, F->fin.dev, F->fin.nj(), st);
        });
    }
    // jobs outside the classes: the single-job path, in order
    for (auto& J : B->jobs) {
        if (J.batched) continue;
        const pyl_pk_job j = J.job;
        B->steps.push_back([j](hipStream_t st) {
            return pk_particles(j.pos, j.W, j.n, j.dims, j.BoxSize, j.scheme, j.axis, j.prec,
                                j.flags | PYL_PK_ASYNC | (j.flags & DEP_ATOMIC ? 0 : DEP_ATOMIC), j.out, st, Rsd());
        });
    }
    *out = guard.release();
    return OK;
}

static int batch_run(const BatchPlan* B, hipStream_t st) {
    for (auto& f : B->steps) {
        int rc = f(st);
        if (rc) return rc;
    }
    return OK;
}

}  // namespace pyl

using namespace pyl;

// ------------------------------------------------------------------ batches and graphs
extern "C" int pyl_pk_batch(const pyl_pk_job* jobs, int njobs, void* stream) {
    PYL_CHECK(njobs >= 0 && (njobs == 0 || jobs), E_ARG, "bad job array");
    if (njobs == 0) return OK;
    hipStream_t st = (hipStream_t)stream;
    BatchPlan* B = nullptr;
    int rc = batch_build(jobs, njobs, st, &B);
    if (rc) return rc;
    rc = batch_run(B, st);
    delete B;
    return rc;
}

namespace {
struct GraphHandle {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    WsArena* arena = nullptr;
    BatchPlan* plan = nullptr;
};

void graph_free(GraphHandle* h) {
    if (h->exec) (void)hipGraphExecDestroy(h->exec);
    if (h->graph) (void)hipGraphDestroy(h->graph);
    ws_arena_destroy(h->arena);
    delete h->plan;
    delete h;
}
}  // namespace

extern "C" int pyl_graph_create(const pyl_pk_job* jobs, int njobs, void* stream, void** graph) {
    PYL_CHECK(graph, E_ARG, "null graph handle");
    PYL_CHECK(njobs > 0 && jobs, E_ARG, "bad job array");
    hipStream_t st = (hipStream_t)stream;
    PYL_CHECK(st != nullptr, E_ARG, "graph capture needs a non-default stream");
    GraphHandle* h = new GraphHandle();
    h->arena = ws_arena_create();
    int rc = OK;
    {
        ArenaScope scope(h->arena);
        // plan in the graph's own workspaces (autotuner trials, tables, geometry), one eager run that
        // sizes every remaining workspace, then the capture of the same launch sequence
        rc = batch_build(jobs, njobs, st, &h->plan);
        if (rc == OK) rc = batch_run(h->plan, st);
        hipError_t e = rc ? hipSuccess : hipStreamSynchronize(st);
        if (rc == OK && e != hipSuccess) {
            set_error("hipStreamSynchronize -> %s", hipGetErrorString(e));
            rc = (int)e + 1000;
        }
        if (rc == OK) {
            ws_arena_seal(h->arena);
            e = hipStreamBeginCapture(st, hipStreamCaptureModeRelaxed);
            if (e != hipSuccess) {
                set_error("hipStreamBeginCapture -> %s", hipGetErrorString(e));
                rc = (int)e + 1000;
            } else {
                rc = batch_run(h->plan, st);
                e = hipStreamEndCapture(st, &h->graph);
                if (rc == OK && e != hipSuccess) {
                    set_error("hipStreamEndCapture -> %s", hipGetErrorString(e));
                    rc = (int)e + 1000;
                }
            }
        }
    }
    if (rc == OK) {
        hipError_t e = hipGraphInstantiate(&h->exec, h->graph, nullptr, nullptr, 0);
        if (e != hipSuccess) {
            set_error("hipGraphInstantiate -> %s", hipGetErrorString(e));
            rc = (int)e + 1000;
        }
    }
    if (rc) {
        graph_free(h);
        return rc;
    }
    *graph = h;
    return OK;
}

extern "C" int pyl_graph_launch(void* graph, void* stream) {
    PYL_CHECK(graph, E_ARG, "null graph");
    PYL_HIP(hipGraphLaunch(((GraphHandle*)graph)->exec, (hipStream_t)stream));
    return OK;
}

extern "C" int pyl_graph_destroy(void* graph) {
    if (!graph) return OK;
    graph_free((GraphHandle*)graph);
    return OK;
}
