// batch.cpp -- driver of the batched 2D path: the 2D step of the reference
// main loop (2D/run_2D.py:772-909) for many scenes at once.
//
// Every phase is one persistent launch over a unit table that spans all
// scenes (pfm_batch.h).  The host keeps only the reinitialisation schedule,
// which every scene shares; dt, the frame clip and the PCG scalars are
// per-scene device state.  The launch sequence of a step is split into
// segments at the only host decisions left (the schedule, and whether any
// scene's PCG is still iterating); each segment is captured once into a HIP
// graph per variant and replayed afterwards.  pfm_batch2d_tune times every
// kernel kind under several persistent grid sizes (and the coarse multigrid
// kernel under several block sizes) and keeps the fastest.
#include <map>
#include <new>
#include <vector>

#include "pfm_batch.h"

namespace pfm {
namespace {

// kKFaceRk4 is the fused midpoint launch; kKAdvect is no longer launched and
// keeps its slot so that pfm_batch2d_tuning's index layout stays as it was.
enum Kind {
    kKRk4 = 0, kKG2p, kKSeed, kKFaceRk4, kKAdvect, kKFace, kKTile, kKBin, kKCell, kKDown, kKUp, kKCount
};
constexpr int kCapCand[] = {0, 4096, 2048, 1024, 512};
constexpr int kNCap = sizeof(kCapCand) / sizeof(kCapCand[0]);
constexpr int kCoarseThreadCand[] = {256, 512, 1024};
constexpr int kItersPerReplay = 4;  // even: the p ping-pong is back on pp[0] after a replay

struct List {
    int32_t off = 0, n = 0;
};

struct Batch {
    pfm_batch2d_desc d;
    std::vector<int32_t> res;
    std::vector<int64_t> seeds;
    int S = 0;
    int64_t P = 0, F[2] = {0, 0}, C = 0, T = 0, Lcells = 0, NP = 0;
    std::vector<BScene> hs;
    BScene* sc = nullptr;
    BUnit* units = nullptr;
    List parts, faces, tiles, cells0;
    std::vector<List> mgtiles;                // per tiled level: 32 x 32 tiles
    // compacted copies of cells0 (list 0) and mgtiles[l] (list 1 + l): the
    // units of the scenes whose PCG is still iterating
    int n_clists = 0;
    std::vector<List> clist;
    std::vector<BUnit*> cdst;
    int32_t* cbeg = nullptr;   // n_clists x (S + 1) scene offsets within each list
    int32_t* ccount = nullptr; // n_clists
    int Lt = 0;                               // tiled levels (max over scenes)
    BWalls wl{1, 1, 1, 1};
    BState* st = nullptr;
    BMg* mg = nullptr;
    int32_t* iters = nullptr;  // 2 per scene
    int32_t* conv = nullptr;
    int32_t* count = nullptr;
    int32_t* host_count = nullptr;  // pinned
    // level arrays: r, the pre-smoothed z of each tiled level (z) and the
    // V-cycle output of every level (zu; level 0 is the preconditioned r)
    float *r = nullptr, *z = nullptr, *zu = nullptr, *x = nullptr, *Ap = nullptr;
    float* pp[2] = {nullptr, nullptr};  // search direction, ping-pong (iteration k reads pp[k & 1])
    double *part = nullptr, *part2 = nullptr;
    double* zr = nullptr;               // z.r partials (BScene::zr_off)
    int64_t NZR = 0;
    int32_t *counts = nullptr, *scan_tiles = nullptr, *slot = nullptr;
    float *rec = nullptr, *p2g_partial = nullptr;
    size_t coarse_lds = 0;
    int64_t bytes = 0;
    hipStream_t bs = nullptr;
    hipEvent_t ev_in = nullptr, ev_out = nullptr;
    // launch shapes
    int cap[kKCount];
    int coarse_threads = 512;
    // graphs
    bool use_graphs = true;
    std::map<int, hipGraphExec_t> graphs;
    // tuning
    bool timing = false;
    std::vector<hipEvent_t> evpool;
    size_t evused = 0;
    std::vector<std::pair<int, size_t>> timed;  // (kind, first event index)
};

BUnits view(const Batch* b, const List& l) { return BUnits{b->sc, b->units + l.off, l.n, nullptr, b->st}; }

// Compacted list k (0: level-0 cells, 1 + l: tiles of level l).
BUnits cview(const Batch* b, int k) { return BUnits{b->sc, b->cdst[k], b->clist[k].n, b->ccount + k, b->st}; }

int compact(Batch* b) {
    std::vector<const BUnit*> src(b->n_clists);
    std::vector<int32_t> len(b->n_clists);
    for (int k = 0; k < b->n_clists; ++k) {
        src[k] = b->units + b->clist[k].off;
        len[k] = b->clist[k].n;
    }
    return bm_compact(b->S, b->n_clists, b->mg, src.data(), len.data(), b->cbeg, b->cdst.data(), b->ccount, b->bs);
}

unsigned blocks_for(const Batch* b, int kind, const List& l) {
    const int c = b->cap[kind];
    int n = l.n;
    if (c > 0 && n > c) n = c;
    return (unsigned)(n < 1 ? 1 : n);
}

#define PFM_B_TRY(expr)          \
    do {                         \
        int _rc = (expr);        \
        if (_rc < 0) return _rc; \
    } while (0)

// Runs one launch (or launch group) of a kernel kind; while tuning, brackets
// it with events.
template <class F>
int timed_launch(Batch* b, int kind, F&& f) {
    if (!b->timing) return f();
    while (b->evused + 2 > b->evpool.size()) {
        hipEvent_t e;
        PFM_HIP_TRY(hipEventCreate(&e));
        b->evpool.push_back(e);
    }
    const size_t i0 = b->evused;
    b->evused += 2;
    PFM_HIP_TRY(hipEventRecord(b->evpool[i0], b->bs));
    const int rc = f();
    PFM_HIP_TRY(hipEventRecord(b->evpool[i0 + 1], b->bs));
    b->timed.push_back({kind, i0});
    return rc;
}

int copy_f(Batch* b, float* dst, const float* src, int64_t n) {
    PFM_HIP_TRY(hipMemcpyAsync(dst, src, sizeof(float) * (size_t)n, hipMemcpyDeviceToDevice, b->bs));
    return PFM_OK;
}

// ---- multigrid ----------------------------------------------------------------

// apply_preconditioner (3D/mgpcg.py:182-200): per tiled level one launch for
// the pre-smoother + restriction and one for the prolongation +
// post-smoother; the coarse levels in one workgroup per scene.
int vcycle(Batch* b) {
    hipStream_t s = b->bs;
    for (int l = 0; l < b->Lt; ++l) {
        PFM_B_TRY(timed_launch(b, kKDown, [&] {
            return bm_down(cview(b, 1 + l), l, b->wl, b->z, b->r, b->mg, blocks_for(b, kKDown, b->mgtiles[l]), s);
        }));
    }
    PFM_B_TRY(timed_launch(b, kKCount, [&] {
        return bm_coarse(b->S, b->sc, b->wl, b->zu, b->r, b->mg, b->zr, b->coarse_threads, b->coarse_lds, s);
    }));
    for (int l = b->Lt - 1; l >= 0; --l) {
        PFM_B_TRY(timed_launch(b, kKUp, [&] {
            return bm_up(cview(b, 1 + l), l, b->wl, b->z, b->zu, b->r, b->mg, b->zr,
                         blocks_for(b, kKUp, b->mgtiles[l]), s);
        }));
    }
    return PFM_OK;
}

// A level-0 cell launch over every scene, or (active) over the compacted
// list of the scenes still iterating.
int cells_launch(Batch* b, int (*fn)(Batch*, BUnits, unsigned), bool active = true) {
    return timed_launch(b, kKCell, [&] {
        return fn(b, active ? cview(b, 0) : view(b, b->cells0), blocks_for(b, kKCell, b->cells0));
    });
}

int finalize(Batch* b, int mode) {
    return bm_finalize(b->S, b->sc, mode, 1e-12, b->d.max_iters, b->part, b->part2, b->zr, b->mg, b->bs);
}

// apply_bc, divergence, init and the PCG set-up through the first
// preconditioner application (3D/mgpcg.py:202-228, 420-430).
int project_setup(Batch* b) {
    hipStream_t s = b->bs;
    PFM_B_TRY(timed_launch(b, kKFace, [&] {
        return bm_zero_walls(view(b, b->faces), b->d.u[0], b->d.u[1], blocks_for(b, kKFace, b->faces), s);
    }));
    PFM_B_TRY(cells_launch(b, [](Batch* b, BUnits t, unsigned nb) {
        return bm_div_init(t, b->d.u[0], b->d.u[1], b->r, b->x, b->pp[0], b->Ap, nb, b->bs);
    }, false));
    PFM_B_TRY(bm_mg_reset(b->S, b->mg, b->st, b->d.tol_mode, b->d.tol, b->d.max_iters, s));
    PFM_B_TRY(compact(b));
    // all walls Neumann: recentre r (3D/mgpcg.py:215)
    PFM_B_TRY(cells_launch(b, [](Batch* b, BUnits t, unsigned nb) {
        return bm_reduce(t, kBRedSum, b->r, nullptr, b->mg, b->part, nb, b->bs);
    }));
    PFM_B_TRY(finalize(b, kBFinMean));  // the level-0 readers subtract the mean as they load r
    if (b->d.tol_mode == 1) {
        PFM_B_TRY(cells_launch(b, [](Batch* b, BUnits t, unsigned nb) {
            return bm_reduce(t, kBRedRR, b->r, nullptr, b->mg, b->part, nb, b->bs);
        }));
        PFM_B_TRY(finalize(b, kBFinR0));
        PFM_B_TRY(bm_set_thresh(b->S, b->mg, b->d.tol, s));
    }
    PFM_B_TRY(vcycle(b));  // leaves the z.r partials
    return finalize(b, kBFinZTr);
}

// PCG iteration k of every active scene (3D/mgpcg.py:229-275): update_p and
// Ap in one launch from pp[k & 1] into pp[(k + 1) & 1].
int pcg_iteration(Batch* b, int k) {
    float* pin = b->pp[k & 1];
    float* pout = b->pp[(k + 1) & 1];
    PFM_B_TRY(timed_launch(b, kKCell, [&] {
        return bm_ap_dot(cview(b, 0), b->wl, b->zu, pin, pout, b->Ap, b->mg, b->part,
                         blocks_for(b, kKCell, b->cells0), b->bs);
    }));
    PFM_B_TRY(finalize(b, kBFinAlpha));
    PFM_B_TRY(timed_launch(b, kKCell, [&] {
        return bm_update_xr(cview(b, 0), b->x, b->r, pout, b->Ap, b->mg, b->part, b->part2,
                            blocks_for(b, kKCell, b->cells0), b->bs);
    }));
    PFM_B_TRY(finalize(b, kBFinRTr));
    PFM_B_TRY(compact(b));
    PFM_B_TRY(vcycle(b));  // recentres r as it loads it; leaves the z.r partials
    PFM_B_TRY(finalize(b, kBFinBeta));
    return compact(b);
}

// get_result + subtract_grad_p + apply_bc; records the solve in slot k.
int project_finish(Batch* b, int k) {
    hipStream_t s = b->bs;
    PFM_B_TRY(timed_launch(b, kKFace, [&] {
        return bm_subtract_grad_p(view(b, b->faces), b->x, b->d.u[0], b->d.u[1], blocks_for(b, kKFace, b->faces), s);
    }));
    PFM_B_TRY(timed_launch(b, kKFace, [&] {
        return bm_zero_walls(view(b, b->faces), b->d.u[0], b->d.u[1], blocks_for(b, kKFace, b->faces), s);
    }));
    return bm_save_solve(b->S, b->mg, k, b->iters, b->conv, s);
}

// ---- graphs -------------------------------------------------------------------

template <class F>
int run_segment(Batch* b, int key, F&& f) {
    if (!b->use_graphs || b->timing) return f();
    auto it = b->graphs.find(key);
    if (it == b->graphs.end()) {
        PFM_HIP_TRY(hipStreamBeginCapture(b->bs, hipStreamCaptureModeThreadLocal));
        const int rc = f();
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(b->bs, &g);
        if (rc < 0) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        if (e != hipSuccess) {
            set_error("graph capture failed: %s", hipGetErrorString(e));
            return PFM_ERR_HIP;
        }
        hipGraphExec_t ex = nullptr;
        const hipError_t e2 = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e2 != hipSuccess) {
            set_error("graph instantiation failed: %s", hipGetErrorString(e2));
            return PFM_ERR_HIP;
        }
        it = b->graphs.emplace(key, ex).first;
    }
    PFM_HIP_TRY(hipGraphLaunch(it->second, b->bs));
    return PFM_OK;
}

void drop_graphs(Batch* b) {
    for (auto& kv : b->graphs) (void)hipGraphExecDestroy(kv.second);
    b->graphs.clear();
}

enum SegKey { kSegIter = 1, kSegMid = 2, kSegFinish0 = 3, kSegFinish1 = 4, kSegProject = 5, kSegInit = 6,
              kSegPreBase = 16 };

// PCG iterations until no scene is active; returns the lockstep count.
int solve_loop(Batch* b, int* lockstep) {
    int n = 0;
    const int limit = b->d.max_iters + kItersPerReplay;
    while (true) {
        PFM_B_TRY(run_segment(b, kSegIter, [&] {
            for (int k = 0; k < kItersPerReplay; ++k) PFM_B_TRY(pcg_iteration(b, k));
            return bm_count_active(b->S, b->mg, b->count, b->bs);
        }));
        n += kItersPerReplay;
        PFM_HIP_TRY(hipMemcpyAsync(b->host_count, b->count, sizeof(int32_t), hipMemcpyDeviceToHost, b->bs));
        PFM_HIP_TRY(hipStreamSynchronize(b->bs));
        if (*b->host_count == 0 || n > limit) break;
    }
    if (lockstep) *lockstep = n;
    return PFM_OK;
}

// ---- step phases ----------------------------------------------------------------

int seed(Batch* b) {
    return timed_launch(b, kKSeed, [&] {
        return b_seed(view(b, b->parts), b->d.ppc_axis, b->d.particle_order, b->P, b->d.pos,
                      blocks_for(b, kKSeed, b->parts), b->bs);
    });
}

int g2p(Batch* b, float* imp, float* Cx, float* Cy) {
    return timed_launch(b, kKG2p, [&] {
        return b_g2p(view(b, b->parts), b->P, b->F, b->d.pos, b->d.u[0], b->d.u[1], imp, Cx, Cy,
                     blocks_for(b, kKG2p, b->parts), b->bs);
    });
}

int reset_T(Batch* b, float* const* T) { return pfm_reset_T_identity(2, b->P, T[0], T[1], nullptr, b->bs); }

int compute_dt(Batch* b) {
    PFM_B_TRY(b_ms_init(b->S, b->st, b->bs));
    PFM_B_TRY(timed_launch(b, kKCell, [&] {
        return b_max_speed(view(b, b->cells0), b->d.u[0], b->d.u[1], b->st, blocks_for(b, kKCell, b->cells0), b->bs);
    }));
    return b_dt(b->S, b->sc, b->st, b->d.CFL, b->d.visualize_dt, b->bs);
}

// 2D/run_2D.py:804-823 for the decisions of this step.
int reinit(Batch* b, bool long_r, bool reseed, bool short_r) {
    const pfm_batch2d_desc& d = b->d;
    if (long_r) {
        if (reseed) PFM_B_TRY(seed(b));
        PFM_B_TRY(g2p(b, d.imp, nullptr, nullptr));
        PFM_B_TRY(copy_f(b, d.init_imp, d.imp, 2 * b->P));
        PFM_B_TRY(reset_T(b, d.T));
        PFM_B_TRY(reset_T(b, d.T_init));
    }
    if (short_r) {
        PFM_B_TRY(g2p(b, d.imp_grad_m, d.init_C[0], d.init_C[1]));
        PFM_B_TRY(reset_T(b, d.T_grad_m));
        for (int a = 0; a < 2; ++a) PFM_B_TRY(copy_f(b, d.T_init[a], d.T[a], 2 * b->P));
    }
    return PFM_OK;
}

// The midpoint velocity up to its projection (2D/run_2D.py:865-872).
int midpoint_pre(Batch* b) {
    const pfm_batch2d_desc& d = b->d;
    const BUnits fu = view(b, b->faces);
    // the copy of u first: one launch then resets, runs RK4_grid and advects
    // every face, sampling the copy for both (the values RK4_grid reads as u)
    for (int a = 0; a < 2; ++a) PFM_B_TRY(copy_f(b, d.p2g_weight[a], d.u[a], b->F[a]));
    return timed_launch(b, kKFaceRk4, [&] {
        return b_midpoint_advect(fu, b->F, d.p2g_weight[0], d.p2g_weight[1], d.u[0], d.u[1], d.psi_grid[0],
                                 d.psi_grid[1], d.T_grid[0], d.T_grid[1], b->st, blocks_for(b, kKFaceRk4, b->faces),
                                 b->bs);
    });
}

// RK4_T_forward with the T / C / impulse updates and the P2G record in its
// epilogue, binning, P2G through the cell permutation (2D/run_2D.py:874-903).
int particles_to_grid(Batch* b) {
    const pfm_batch2d_desc& d = b->d;
    hipStream_t s = b->bs;
    PFM_B_TRY(timed_launch(b, kKRk4, [&] {
        const float* Tin[2] = {d.T_init[0], d.T_init[1]};
        const float* iC[2] = {d.init_C[0], d.init_C[1]};
        return b_rk4_update(view(b, b->parts), b->P, d.pos, d.T_grad_m[0], d.T_grad_m[1], d.u[0], d.u[1], b->st, Tin,
                            iC, d.init_imp, d.C, d.T, d.imp, b->rec, blocks_for(b, kKRk4, b->parts), s);
    }));
    PFM_B_TRY(timed_launch(b, kKBin, [&] {
        return b_bin(view(b, b->parts), b->P, b->C, d.pos, d.cell_of, d.cell_start, d.perm, b->counts, b->scan_tiles,
                     b->slot, blocks_for(b, kKBin, b->parts), s);
    }));
    const unsigned nb[2] = {blocks_for(b, kKTile, b->tiles), blocks_for(b, kKFace, b->faces)};
    return timed_launch(b, kKTile, [&] {
        return b_p2g_records(view(b, b->tiles), view(b, b->faces), d.cell_start, d.perm, b->rec, b->p2g_partial,
                             d.u[0], d.u[1], d.p2g_weight[0], d.p2g_weight[1], nb, s);
    });
}

int begin(Batch* b, hipStream_t caller) {
    PFM_HIP_TRY(hipEventRecord(b->ev_in, caller));
    PFM_HIP_TRY(hipStreamWaitEvent(b->bs, b->ev_in, 0));
    return PFM_OK;
}

int end(Batch* b, hipStream_t caller) {
    PFM_HIP_TRY(hipEventRecord(b->ev_out, b->bs));
    PFM_HIP_TRY(hipStreamWaitEvent(caller, b->ev_out, 0));
    return PFM_OK;
}

int step(Batch* b, pfm_batch2d_state* st) {
    const pfm_batch2d_desc& d = b->d;
    st->i += 1;
    const int64_t i = st->i;
    const bool long_r = i % d.reinit_every == 0;
    const bool reseed = long_r && (d.reinit_particle_pos || i == 0);
    if (long_r) st->ik = i;  // the 2D driver moves ik at every long reinit
    const bool short_r = (i - st->ik) % d.reinit_every_grad_m == 0;
    const int pre_key = kSegPreBase + (long_r ? 1 : 0) + (reseed ? 2 : 0) + (short_r ? 4 : 0);
    st->n_solves = 0;
    PFM_B_TRY(run_segment(b, pre_key, [&] {
        PFM_B_TRY(compute_dt(b));
        PFM_B_TRY(reinit(b, long_r, reseed, short_r));
        if (d.use_midpoint_vel)
            PFM_B_TRY(midpoint_pre(b));
        else
            PFM_B_TRY(particles_to_grid(b));
        return project_setup(b);
    }));
    PFM_B_TRY(solve_loop(b, &st->lockstep_iters[0]));
    st->n_solves = 1;
    if (d.use_midpoint_vel) {
        PFM_B_TRY(run_segment(b, kSegMid, [&] {
            PFM_B_TRY(project_finish(b, 0));
            PFM_B_TRY(particles_to_grid(b));
            return project_setup(b);
        }));
        PFM_B_TRY(solve_loop(b, &st->lockstep_iters[1]));
        st->n_solves = 2;
        return run_segment(b, kSegFinish1, [&] { return project_finish(b, 1); });
    }
    return run_segment(b, kSegFinish0, [&] { return project_finish(b, 0); });
}

int64_t levels_for(int n0, int n1, int base_level) {
    const int mn = n0 < n1 ? n0 : n1;
    int lg = 0;
    while ((2 << lg) <= mn) ++lg;
    int L = lg - base_level + 1;
    if (L < 1) L = 1;
    if (L > kBatchMaxLevels) L = kBatchMaxLevels;
    return L;
}

void free_batch(Batch* b) {
    if (!b) return;
    drop_graphs(b);
    for (hipEvent_t e : b->evpool) (void)hipEventDestroy(e);
    void* dev[] = {b->sc, b->units, b->st, b->mg, b->iters, b->conv, b->count, b->r, b->z, b->zu, b->x, b->pp[0],
                   b->pp[1], b->Ap, b->part, b->part2, b->zr, b->counts, b->scan_tiles, b->slot, b->rec,
                   b->p2g_partial, b->cbeg, b->ccount};
    for (void* q : dev)
        if (q) (void)hipFree(q);
    for (BUnit* q : b->cdst)
        if (q) (void)hipFree(q);
    if (b->host_count) (void)hipHostFree(b->host_count);
    if (b->ev_in) (void)hipEventDestroy(b->ev_in);
    if (b->ev_out) (void)hipEventDestroy(b->ev_out);
    if (b->bs) (void)hipStreamDestroy(b->bs);
    delete b;
}

}  // namespace
}  // namespace pfm

using namespace pfm;

extern "C" {

int pfm_batch2d_sizes(int32_t n_scenes, const int32_t* res, int32_t ppc_axis, int64_t* out) {
    PFM_CHECK_ARG(n_scenes >= 1 && res && out && ppc_axis >= 1, "bad arguments");
    int64_t P = 0, F0 = 0, F1 = 0, C = 0;
    for (int s = 0; s < n_scenes; ++s) {
        const int64_t a = res[2 * s], c = res[2 * s + 1];
        PFM_CHECK_ARG(a >= 4 && c >= 4 && a <= 65535 && c <= 65535, "scene %d: res (%lld, %lld) out of [4, 65535]", s,
                      (long long)a, (long long)c);
        P += a * c * ppc_axis * ppc_axis;
        F0 += (a + 1) * c;
        F1 += a * (c + 1);
        C += a * c;
    }
    out[0] = P;
    out[1] = F0;
    out[2] = F1;
    out[3] = C;
    return PFM_OK;
}

int pfm_batch2d_create(const pfm_batch2d_desc* desc, void** handle, int64_t* bytes_out) {
    PFM_CHECK_ARG(handle && desc, "null descriptor or handle out-pointer");
    *handle = nullptr;
    const pfm_batch2d_desc& d = *desc;
    PFM_CHECK_ARG(d.n_scenes >= 1 && d.res && d.seed, "need n_scenes >= 1, res and seed");
    PFM_CHECK_ARG(d.reinit_every >= 1 && d.reinit_every_grad_m >= 1, "reinit intervals must be >= 1");
    PFM_CHECK_ARG(d.CFL > 0.0 && d.visualize_dt > 0.0, "CFL and visualize_dt must be positive");
    PFM_CHECK_ARG(d.max_iters >= 0, "max_iters must be >= 0 for a batch");
    PFM_CHECK_ARG(d.tol_mode == 0 || d.tol_mode == 1, "tol_mode must be 0 or 1");
    PFM_CHECK_ARG(d.particle_order == 0 || d.particle_order == 1, "particle_order must be 0 or 1");
    for (int a = 0; a < 2; ++a)
        PFM_CHECK_ARG(d.u[a] && d.p2g_weight[a] && d.psi_grid[a] && d.T_grid[a] && d.T[a] && d.T_init[a] &&
                          d.T_grad_m[a] && d.C[a] && d.init_C[a],
                      "null field pointer (axis %d)", a);
    PFM_CHECK_ARG(d.pos && d.imp && d.init_imp && d.cell_of && d.cell_start && d.perm, "null field pointer");
    int64_t tot[4];
    PFM_B_TRY(pfm_batch2d_sizes(d.n_scenes, d.res, d.ppc_axis, tot));
    PFM_CHECK_ARG(tot[0] < (int64_t)INT32_MAX && tot[3] < (int64_t)INT32_MAX, "batch too large for int32 ids");
    PFM_CHECK_ARG(d.n_scenes <= kBatchCompactMaxScenes, "at most %d scenes per batch", kBatchCompactMaxScenes);
    Batch* b = new (std::nothrow) Batch();
    if (!b) {
        set_error("host allocation failed");
        return PFM_ERR_OOM;
    }
    b->d = d;
    b->S = d.n_scenes;
    b->res.assign(d.res, d.res + 2 * d.n_scenes);
    b->seeds.assign(d.seed, d.seed + d.n_scenes);
    b->d.res = b->res.data();
    b->d.seed = b->seeds.data();
    b->P = tot[0];
    b->F[0] = tot[1];
    b->F[1] = tot[2];
    b->C = tot[3];
    for (int k = 0; k < kKCount; ++k) b->cap[k] = 0;
    // scene table
    std::vector<BUnit> U;
    auto add = [&](List& l, int s, int64_t chunks) {
        for (int64_t c = 0; c < chunks; ++c) U.push_back(BUnit{s, (int32_t)c});
        l.n += (int32_t)chunks;
    };
    b->hs.resize(b->S);
    int64_t p_off = 0, f_off0 = 0, f_off1 = 0, c_off = 0, t_off = 0, lev = b->C;
    int32_t part_off = 0;
    size_t lds_max = 0;
    for (int s = 0; s < b->S; ++s) {
        BScene& S = b->hs[s];
        S = BScene{};
        pfm_grid pg;
        pg.dim = 2;
        pg.res[0] = b->res[2 * s];
        pg.res[1] = b->res[2 * s + 1];
        pg.res[2] = 1;
        const double dx = 1.0 / pg.res[1];
        pg.dx = (float)dx;
        int rc = make_grid(&pg, &S.g);
        if (rc) {
            free_batch(b);
            return rc;
        }
        S.dist = (float)(dx / d.ppc_axis);
        S.seed = b->seeds[s];
        const int n0 = pg.res[0], n1 = pg.res[1];
        S.n = (int64_t)n0 * n1 * d.ppc_axis * d.ppc_axis;
        S.p_off = p_off;
        S.nf[0] = (int64_t)(n0 + 1) * n1;
        S.nf[1] = (int64_t)n0 * (n1 + 1);
        S.f_off[0] = f_off0;
        S.f_off[1] = f_off1;
        S.nc = (int64_t)n0 * n1;
        S.c_off = c_off;
        S.nt[0] = (n0 + 15) / 16;
        S.nt[1] = (n1 + 15) / 16;
        S.t_off = t_off;
        S.n_levels = (int32_t)levels_for(n0, n1, d.base_level);
        S.coarse = -1;
        size_t lds = 0;
        for (int l = 0; l < S.n_levels; ++l) {
            S.ld[l][0] = n0 >> l > 0 ? n0 >> l : 1;
            S.ld[l][1] = n1 >> l > 0 ? n1 >> l : 1;
            const int64_t nl = (int64_t)S.ld[l][0] * S.ld[l][1];
            if (l == 0) {
                S.l_off[0] = c_off;
            } else {
                S.l_off[l] = lev;
                lev += nl;
            }
            if (S.coarse < 0 && nl <= kBatchCoarseCells) S.coarse = l;
            if (S.coarse >= 0) lds += 2 * sizeof(float) * (size_t)nl;
        }
        if (S.coarse < 0) {
            free_batch(b);
            PFM_CHECK_ARG(false, "scene %d: bottom level %dx%d exceeds %d cells (raise base_level's level count)", s,
                          S.ld[S.n_levels - 1][0], S.ld[S.n_levels - 1][1], kBatchCoarseCells);
        }
        if (lds > lds_max) lds_max = lds;
        if (S.coarse > b->Lt) b->Lt = S.coarse;
        S.part_off = part_off;
        S.n_part = (int32_t)((S.nc + kBlock - 1) / kBlock);
        part_off += S.n_part;
        S.zr_off = (int32_t)b->NZR;
        S.zr_n = S.coarse == 0 ? 1
                               : (int32_t)(((n0 + kBatchTile - 1) / kBatchTile) * ((n1 + kBatchTile - 1) / kBatchTile));
        b->NZR += S.zr_n;
        p_off += S.n;
        f_off0 += S.nf[0];
        f_off1 += S.nf[1];
        c_off += S.nc;
        t_off += (int64_t)S.nt[0] * S.nt[1];
    }
    b->T = t_off;
    b->Lcells = lev;
    b->NP = part_off;
    b->coarse_lds = lds_max;
    // unit tables
    b->parts.off = (int32_t)U.size();
    for (int s = 0; s < b->S; ++s) add(b->parts, s, (b->hs[s].n + kBlock - 1) / kBlock);
    b->faces.off = (int32_t)U.size();
    for (int s = 0; s < b->S; ++s) add(b->faces, s, (b->hs[s].nf[0] + b->hs[s].nf[1] + kBlock - 1) / kBlock);
    b->tiles.off = (int32_t)U.size();
    for (int s = 0; s < b->S; ++s) add(b->tiles, s, (int64_t)b->hs[s].nt[0] * b->hs[s].nt[1]);
    b->cells0.off = (int32_t)U.size();
    for (int s = 0; s < b->S; ++s) add(b->cells0, s, b->hs[s].n_part);
    b->n_clists = 1 + b->Lt;
    std::vector<int32_t> hbeg((size_t)b->n_clists * (b->S + 1), 0);
    for (int s = 0; s <= b->S; ++s) hbeg[s] = s < b->S ? b->hs[s].part_off : (int32_t)b->NP;
    b->mgtiles.resize(b->Lt);
    for (int l = 0; l < b->Lt; ++l) {
        b->mgtiles[l].off = (int32_t)U.size();
        int32_t* bl = hbeg.data() + (size_t)(1 + l) * (b->S + 1);
        for (int s = 0; s < b->S; ++s) {
            bl[s] = b->mgtiles[l].n;
            if (l < b->hs[s].coarse) {
                const int64_t t0 = (b->hs[s].ld[l][0] + kBatchTile - 1) / kBatchTile;
                const int64_t t1 = (b->hs[s].ld[l][1] + kBatchTile - 1) / kBatchTile;
                add(b->mgtiles[l], s, t0 * t1);
            }
        }
        bl[b->S] = b->mgtiles[l].n;
    }
    b->clist.push_back(b->cells0);
    for (int l = 0; l < b->Lt; ++l) b->clist.push_back(b->mgtiles[l]);
    // device memory
    int64_t bytes = 0;
    auto alloc = [&](void** ptr, size_t n) {
        if (hipMalloc(ptr, n < 256 ? 256 : n) != hipSuccess) return false;
        bytes += (int64_t)n;
        return true;
    };
    const int64_t ntiles = (b->C + 16 * kBlock - 1) / (16 * kBlock);
    bool ok = alloc((void**)&b->sc, sizeof(BScene) * b->S) && alloc((void**)&b->units, sizeof(BUnit) * U.size()) &&
              alloc((void**)&b->st, sizeof(BState) * b->S) && alloc((void**)&b->mg, sizeof(BMg) * b->S) &&
              alloc((void**)&b->iters, sizeof(int32_t) * 2 * b->S) &&
              alloc((void**)&b->conv, sizeof(int32_t) * 2 * b->S) && alloc((void**)&b->count, sizeof(int32_t)) &&
              alloc((void**)&b->r, sizeof(float) * b->Lcells) && alloc((void**)&b->z, sizeof(float) * b->Lcells) &&
              alloc((void**)&b->zu, sizeof(float) * b->Lcells) &&
              alloc((void**)&b->x, sizeof(float) * b->C) && alloc((void**)&b->pp[0], sizeof(float) * b->C) &&
              alloc((void**)&b->pp[1], sizeof(float) * b->C) && alloc((void**)&b->zr, sizeof(double) * b->NZR) &&
              alloc((void**)&b->Ap, sizeof(float) * b->C) && alloc((void**)&b->part, sizeof(double) * b->NP * kBatchWaves) &&
              alloc((void**)&b->part2, sizeof(double) * b->NP * kBatchWaves) &&
              alloc((void**)&b->counts, sizeof(int32_t) * (b->C + 1)) &&
              alloc((void**)&b->scan_tiles, sizeof(int32_t) * (ntiles + 1)) &&
              alloc((void**)&b->slot, sizeof(int32_t) * b->P) && alloc((void**)&b->rec, sizeof(float) * 8 * b->P) &&
              alloc((void**)&b->p2g_partial, sizeof(float) * b_p2g_partial_floats(b->T)) &&
              alloc((void**)&b->cbeg, sizeof(int32_t) * hbeg.size()) &&
              alloc((void**)&b->ccount, sizeof(int32_t) * b->n_clists);
    b->cdst.assign(b->n_clists, nullptr);
    for (int k = 0; ok && k < b->n_clists; ++k) ok = alloc((void**)&b->cdst[k], sizeof(BUnit) * b->clist[k].n);
    if (!ok) {
        free_batch(b);
        set_error("hipMalloc failed for the batch tables and workspaces");
        return PFM_ERR_OOM;
    }
    if (hipHostMalloc((void**)&b->host_count, sizeof(int32_t), hipHostMallocDefault) != hipSuccess) {
        free_batch(b);
        set_error("hipHostMalloc failed");
        return PFM_ERR_OOM;
    }
    b->bytes = bytes;
    auto fail_hip = [&](hipError_t e) {
        set_error("batch set-up failed: %s", hipGetErrorString(e));
        free_batch(b);
        return PFM_ERR_HIP;
    };
    hipError_t e;
    if ((e = hipStreamCreateWithFlags(&b->bs, hipStreamNonBlocking)) != hipSuccess) return fail_hip(e);
    if ((e = hipEventCreateWithFlags(&b->ev_in, hipEventDisableTiming)) != hipSuccess) return fail_hip(e);
    if ((e = hipEventCreateWithFlags(&b->ev_out, hipEventDisableTiming)) != hipSuccess) return fail_hip(e);
    // set-up copies on the batch's own non-blocking stream: the legacy stream
    // may not be used while another thread captures a graph
    hipStream_t bs = b->bs;
    if ((e = hipMemcpyAsync(b->sc, b->hs.data(), sizeof(BScene) * b->S, hipMemcpyHostToDevice, bs)) != hipSuccess)
        return fail_hip(e);
    if ((e = hipMemcpyAsync(b->units, U.data(), sizeof(BUnit) * U.size(), hipMemcpyHostToDevice, bs)) != hipSuccess)
        return fail_hip(e);
    if ((e = hipMemcpyAsync(b->cbeg, hbeg.data(), sizeof(int32_t) * hbeg.size(), hipMemcpyHostToDevice, bs)) !=
        hipSuccess)
        return fail_hip(e);
    if ((e = hipMemsetAsync(b->ccount, 0, sizeof(int32_t) * b->n_clists, bs)) != hipSuccess) return fail_hip(e);
    std::vector<BState> st0(b->S);
    for (auto& x : st0) {
        x = BState{};
        x.enabled = 1;
    }
    if ((e = hipMemcpyAsync(b->st, st0.data(), sizeof(BState) * b->S, hipMemcpyHostToDevice, bs)) != hipSuccess)
        return fail_hip(e);
    if ((e = hipMemsetAsync(b->mg, 0, sizeof(BMg) * b->S, bs)) != hipSuccess) return fail_hip(e);
    if ((e = hipMemsetAsync(b->iters, 0, sizeof(int32_t) * 2 * b->S, bs)) != hipSuccess) return fail_hip(e);
    if ((e = hipMemsetAsync(b->conv, 0, sizeof(int32_t) * 2 * b->S, bs)) != hipSuccess) return fail_hip(e);
    if ((e = hipStreamSynchronize(bs)) != hipSuccess) return fail_hip(e);
    if (bytes_out) *bytes_out = bytes;
    *handle = b;
    return PFM_OK;
}

int pfm_batch2d_project(void* h, pfm_stream_t stream) {
    PFM_CHECK_ARG(h, "null batch");
    Batch* b = (Batch*)h;
    hipStream_t caller = as_stream(stream);
    PFM_B_TRY(begin(b, caller));
    PFM_B_TRY(run_segment(b, kSegProject, [&] { return project_setup(b); }));
    PFM_B_TRY(solve_loop(b, nullptr));
    PFM_B_TRY(run_segment(b, kSegFinish0, [&] { return project_finish(b, 0); }));
    return end(b, caller);
}

int pfm_batch2d_init_particles(void* h, pfm_stream_t stream) {
    PFM_CHECK_ARG(h, "null batch");
    Batch* b = (Batch*)h;
    const pfm_batch2d_desc& d = b->d;
    hipStream_t caller = as_stream(stream);
    PFM_B_TRY(begin(b, caller));
    PFM_B_TRY(run_segment(b, kSegInit, [&] {
        PFM_B_TRY(seed(b));
        PFM_B_TRY(g2p(b, d.imp_grad_m, d.init_C[0], d.init_C[1]));
        PFM_B_TRY(reset_T(b, d.T_grad_m));
        PFM_B_TRY(g2p(b, d.imp, nullptr, nullptr));
        return copy_f(b, d.init_imp, d.imp, 2 * b->P);
    }));
    return end(b, caller);
}

int pfm_batch2d_step(void* h, pfm_batch2d_state* st, pfm_stream_t stream) {
    PFM_CHECK_ARG(h && st, "null batch or state");
    Batch* b = (Batch*)h;
    hipStream_t caller = as_stream(stream);
    PFM_B_TRY(begin(b, caller));
    PFM_B_TRY(step(b, st));
    return end(b, caller);
}

int pfm_batch2d_scene_info(void* h, double* dt, double* sub_t, int32_t* frame, int32_t* iters, int32_t* converged) {
    PFM_CHECK_ARG(h, "null batch");
    Batch* b = (Batch*)h;
    std::vector<BState> st(b->S);
    PFM_HIP_TRY(hipMemcpyAsync(st.data(), b->st, sizeof(BState) * b->S, hipMemcpyDeviceToHost, b->bs));
    if (iters) PFM_HIP_TRY(hipMemcpyAsync(iters, b->iters, sizeof(int32_t) * 2 * b->S, hipMemcpyDeviceToHost, b->bs));
    if (converged)
        PFM_HIP_TRY(hipMemcpyAsync(converged, b->conv, sizeof(int32_t) * 2 * b->S, hipMemcpyDeviceToHost, b->bs));
    PFM_HIP_TRY(hipStreamSynchronize(b->bs));
    for (int s = 0; s < b->S; ++s) {
        if (dt) dt[s] = st[s].dt;
        if (sub_t) sub_t[s] = st[s].sub_t;
        if (frame) frame[s] = st[s].frame;
    }
    return PFM_OK;
}

int pfm_batch2d_tune(void* h, pfm_batch2d_state* st, pfm_stream_t stream) {
    PFM_CHECK_ARG(h && st, "null batch or state");
    Batch* b = (Batch*)h;
    hipStream_t caller = as_stream(stream);
    PFM_B_TRY(begin(b, caller));
    constexpr int kNT = sizeof(kCoarseThreadCand) / sizeof(kCoarseThreadCand[0]);
    double best[kKCount + 1];
    int pick[kKCount + 1];
    for (int k = 0; k <= kKCount; ++k) {
        best[k] = 1e300;
        pick[k] = 0;
    }
    drop_graphs(b);
    b->timing = true;
    int rc = PFM_OK;
    // one step per candidate; the coarse kernel's block size cycles alongside
    const int rounds = kNCap > kNT ? kNCap : kNT;
    for (int c = 0; c < rounds && rc == PFM_OK; ++c) {
        for (int k = 0; k < kKCount; ++k) b->cap[k] = kCapCand[c % kNCap];
        b->coarse_threads = kCoarseThreadCand[c % kNT];
        b->evused = 0;
        b->timed.clear();
        rc = step(b, st);
        if (rc != PFM_OK) break;
        if (hipStreamSynchronize(b->bs) != hipSuccess) {
            rc = PFM_ERR_HIP;
            break;
        }
        double acc[kKCount + 1] = {0};
        for (auto& t : b->timed) {
            float ms = 0.f;
            if (hipEventElapsedTime(&ms, b->evpool[t.second], b->evpool[t.second + 1]) == hipSuccess)
                acc[t.first] += ms;
        }
        for (int k = 0; k <= kKCount; ++k) {
            const bool cand = k < kKCount ? c < kNCap : c < kNT;
            if (cand && acc[k] > 0.0 && acc[k] < best[k]) {
                best[k] = acc[k];
                pick[k] = c;
            }
        }
    }
    b->timing = false;
    for (int k = 0; k < kKCount; ++k) b->cap[k] = kCapCand[pick[k] % kNCap];
    b->coarse_threads = kCoarseThreadCand[pick[kKCount] % kNT];
    if (rc != PFM_OK) return rc;
    // fresh per-scene clocks and PCG scalars for the run that follows
    {
        std::vector<BState> st0(b->S);
        PFM_HIP_TRY(hipMemcpyAsync(st0.data(), b->st, sizeof(BState) * b->S, hipMemcpyDeviceToHost, b->bs));
        PFM_HIP_TRY(hipStreamSynchronize(b->bs));
        for (auto& x : st0) {
            const int32_t en = x.enabled;
            x = BState{};
            x.enabled = en;
        }
        PFM_HIP_TRY(hipMemcpyAsync(b->st, st0.data(), sizeof(BState) * b->S, hipMemcpyHostToDevice, b->bs));
        PFM_HIP_TRY(hipStreamSynchronize(b->bs));
    }
    PFM_HIP_TRY(hipMemsetAsync(b->mg, 0, sizeof(BMg) * b->S, b->bs));
    PFM_HIP_TRY(hipMemsetAsync(b->iters, 0, sizeof(int32_t) * 2 * b->S, b->bs));
    PFM_HIP_TRY(hipMemsetAsync(b->conv, 0, sizeof(int32_t) * 2 * b->S, b->bs));
    return end(b, caller);
}

int pfm_batch2d_tuning(void* h, int32_t* caps, int32_t n) {
    PFM_CHECK_ARG(h && caps && n >= 0, "bad arguments");
    Batch* b = (Batch*)h;
    int w = 0;
    for (int k = 0; k < kKCount && w < n; ++k) caps[w++] = b->cap[k];
    if (w < n) caps[w++] = b->coarse_threads;
    return w;
}

int pfm_batch2d_set_enabled(void* h, int32_t scene, int32_t enabled) {
    PFM_CHECK_ARG(h, "null batch");
    Batch* b = (Batch*)h;
    PFM_CHECK_ARG(scene >= 0 && scene < b->S, "scene %d out of [0, %d)", scene, b->S);
    const int32_t v = enabled ? 1 : 0;
    PFM_HIP_TRY(hipMemcpyAsync(&b->st[scene].enabled, &v, sizeof(int32_t), hipMemcpyHostToDevice, b->bs));
    PFM_HIP_TRY(hipStreamSynchronize(b->bs));
    return PFM_OK;
}

int pfm_batch2d_set_graphs(void* h, int32_t enable) {
    PFM_CHECK_ARG(h, "null batch");
    Batch* b = (Batch*)h;
    b->use_graphs = enable != 0;
    if (!b->use_graphs) drop_graphs(b);
    return PFM_OK;
}

int pfm_batch2d_destroy(void* h) {
    if (!h) return PFM_OK;
    Batch* b = (Batch*)h;
    (void)hipStreamSynchronize(b->bs);
    free_batch(b);
    return PFM_OK;
}

}  // extern "C"
