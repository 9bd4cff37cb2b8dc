// pfm_batch.h -- internal interface of the batched 2D path: many independent
// 2D scenes of different sizes advanced together, every phase of the step as
// one persistent launch over a table of work units.
//
// All scenes share concatenated arrays: scene s owns particles
// [p_off, p_off + n) of every particle array (SoA, component stride = the
// batch's particle total), faces [f_off[a], f_off[a] + nf[a]) of every face
// array of axis a (component stride = that axis' face total), cells
// [c_off, c_off + cells) of the level-0 arrays and [l_off[l], ...) of the
// multigrid level arrays.  A work unit is (scene, chunk): a chunk of
// kBlock items (particles, faces, cells, red/black pairs) of one scene, or
// one P2G tile.  A persistent kernel walks the units with a grid stride, so
// one launch covers every scene whatever its size.
//
// The per-item arithmetic of every batched kernel is the one of the
// single-scene kernel of the same phase (the bodies are shared), so a scene
// advanced in a batch reproduces its single-scene step except for the order
// of the PCG dot-product sums.
#pragma once

#include "pfm_host.h"

namespace pfm {

constexpr int kBatchMaxLevels = 12;
// Levels with at most this many cells run inside one workgroup per scene.
constexpr int kBatchCoarseCells = 4096;
// waves of a kBlock workgroup: level-0 reduction partials per unit
constexpr int kBatchWaves = kBlock / 64;

struct BScene {
    GridDev g;                 // the scene's 2D grid
    float dist;                // particle lattice spacing dx / ppc_axis
    int64_t seed;              // jitter seed (< 0: the reference lattice)
    int64_t p_off, n;          // particles
    int64_t f_off[2], nf[2];   // faces per axis
    int64_t c_off, nc;         // level-0 cells
    int32_t nt[2];             // P2G tiles per axis
    int64_t t_off;             // first P2G tile partial of the scene
    // multigrid: levels [0, coarse) run as tiled launches, [coarse,
    // n_levels) in the per-scene coarse kernel
    int32_t n_levels, coarse;
    int32_t ld[kBatchMaxLevels][2];
    int64_t l_off[kBatchMaxLevels];  // offset of level l in the level arrays
    int32_t part_off, n_part;        // level-0 reduction partials of the scene
    int32_t zr_off, zr_n;            // its z.r partials (level-0 V-cycle tiles, or 1 from the coarse kernel)
};

struct BUnit {
    int32_t scene, chunk;
};

// Per-scene dynamic state of the step.
struct BState {
    double dt;        // this step's dt (host double arithmetic of the reference, on device)
    double sub_t;     // time since the last frame
    int32_t frame;    // frame index
    uint32_t ms_bits; // max speed (float bits, exact atomic max)
    int32_t enabled;  // 0: the scene is left as it is (a served job that has finished)
    int32_t pad;
};

// Per-scene scalars of the PCG (the MGScalars of the single-scene solver).
struct BMg {
    double zTr, pAp, rTr, r0, sum_r, thresh;
    float alpha, beta, mean;
    int32_t active, it, converged, iters;
};

// Kernel-side view of one unit table.  A compacted table (the units of the
// scenes whose PCG is still running) keeps its length in device memory
// (nu_dev), written by the compaction kernel that ran before the launch.
struct BUnits {
    const BScene* sc;
    const BUnit* u;
    int32_t nu;
    const int32_t* nu_dev;
    const BState* st;  // per-scene enable flags (nullptr: every scene)
};

// Whether the unit's scene is stepped (a disabled scene keeps its fields).
__device__ __forceinline__ bool unit_on(const BUnits& t, int32_t scene) { return !t.st || t.st[scene].enabled; }

// Persistent loop over the units of a launch: `UNIT` is the unit index.
#define PFM_FOR_UNITS(T, UNIT)                                                              \
    for (int32_t UNIT = blockIdx.x, UNIT##_n = (T).nu_dev ? *(T).nu_dev : (T).nu; UNIT < UNIT##_n; \
         UNIT += gridDim.x)

// ---- launchers (implemented beside the single-scene kernels they share) ----
// Every launcher takes the grid size of the persistent launch (`blocks`).

// particles.hip
int b_seed(BUnits t, int ppc_axis, int order, int64_t P, float* pos, unsigned blocks, hipStream_t s);
int b_g2p(BUnits t, int64_t P, const int64_t F[2], const float* pos, const float* ux, const float* uy,
          float* imp, float* Cx, float* Cy, unsigned blocks, hipStream_t s);

int b_rk4_update(BUnits t, int64_t P, float* pos, float* Tx, float* Ty, const float* ux, const float* uy,
                 const BState* st, const float* const* Tin, const float* const* iC, const float* init_imp, float* const* C,
                 float* const* T, float* imp, float* rec, unsigned blocks, hipStream_t s);

// grid.hip
int b_max_speed(BUnits cells, const float* ux, const float* uy, BState* st, unsigned blocks, hipStream_t s);
int b_dt(int n_scenes, const BScene* sc, BState* st, double CFL, double visualize_dt, hipStream_t s);
int b_ms_init(int n_scenes, BState* st, hipStream_t s);
int b_midpoint_advect(BUnits faces, const int64_t F[2], const float* u0x, const float* u0y, float* u1x, float* u1y,
                      float* psi_x, float* psi_y, float* T_x, float* T_y, const BState* st, unsigned blocks,
                      hipStream_t s);

// p2g.hip
int b_bin(BUnits parts, int64_t P, int64_t C, const float* pos, int32_t* cell_of, int32_t* cell_start,
          int32_t* perm, int32_t* counts, int32_t* tiles, int32_t* slot, unsigned blocks, hipStream_t s);
// P2G from records in particle order (b_rk4_update), read through perm.
int b_p2g_records(BUnits tiles, BUnits faces, const int32_t* cell_start, const int32_t* perm, const float* rec,
                  float* partial, float* ux, float* uy, float* Wx, float* Wy, const unsigned blocks[2], hipStream_t s);
size_t b_p2g_partial_floats(int64_t tiles);

// batch_mg.hip: the batched projection (apply_bc, divergence, MGPCG,
// subtract_grad_p) of 2D scenes with Neumann walls.
struct BWalls {
    int lo0, hi0, lo1, hi1;  // walls that remove a neighbour from the diagonal
};
enum BRedMode { kBRedSum = 0, kBRedRR = 1, kBRedZR = 2 };
enum BFinMode { kBFinMean = 0, kBFinR0 = 1, kBFinZTr = 2, kBFinAlpha = 3, kBFinRTr = 4, kBFinBeta = 5 };

int bm_zero_walls(BUnits faces, float* ux, float* uy, unsigned blocks, hipStream_t s);
int bm_div_init(BUnits cells, const float* ux, const float* uy, float* r, float* x, float* p, float* Ap,
                unsigned blocks, hipStream_t s);
int bm_subtract_grad_p(BUnits faces, const float* x, float* ux, float* uy, unsigned blocks, hipStream_t s);
int bm_reduce(BUnits cells, int mode, const float* r, const float* z, const BMg* mg, double* part, unsigned blocks,
              hipStream_t s);
int bm_ap_dot(BUnits cells, BWalls wl, const float* z, const float* pin, float* pout, float* Ap, const BMg* mg,
              double* part, unsigned blocks, hipStream_t s);
int bm_update_xr(BUnits cells, float* x, float* r, const float* p, const float* Ap, const BMg* mg, double* part,
                 double* part2, unsigned blocks, hipStream_t s);
int bm_finalize(int n_scenes, const BScene* sc, int mode, double eps, int max_iters, const double* part,
                const double* part2, const double* zr, BMg* mg, hipStream_t s);
int bm_mg_reset(int n_scenes, BMg* mg, const BState* st, int tol_mode, double tol, int max_iters, hipStream_t s);
int bm_set_thresh(int n_scenes, BMg* mg, double tol, hipStream_t s);
int bm_save_solve(int n_scenes, const BMg* mg, int k, int32_t* iters, int32_t* conv, hipStream_t s);
int bm_count_active(int n_scenes, const BMg* mg, int32_t* out, hipStream_t s);
// fused tiled V-cycle levels: units are 32 x 32 tiles of one scene's level
constexpr int kBatchTile = 32;
int bm_down(BUnits tiles, int l, BWalls wl, float* z, float* r, const BMg* mg, unsigned blocks, hipStream_t s);
int bm_up(BUnits tiles, int l, BWalls wl, const float* zpre, float* zpost, const float* r, const BMg* mg,
          double* zr, unsigned blocks, hipStream_t s);
// Compaction of unit tables to the active scenes: list k (len[k] units,
// grouped by scene in scene order) holds the units of scene s at
// [beg[k * (S + 1) + s], beg[k * (S + 1) + s + 1]); the active scenes' units
// go to dst[k] (same capacity) and their count to count[k].
constexpr int kBatchCompactMaxScenes = 8192;
int bm_compact(int n_scenes, int n_lists, const BMg* mg, const BUnit* const* src, const int32_t* len,
               const int32_t* beg, BUnit* const* dst, int32_t* count, hipStream_t s);
int bm_coarse(int n_scenes, const BScene* sc, BWalls wl, float* z, float* r, const BMg* mg, double* zr, int threads,
              size_t lds_bytes, hipStream_t s);

}  // namespace pfm
