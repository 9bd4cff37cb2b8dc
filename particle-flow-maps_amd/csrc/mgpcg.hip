// mgpcg.hip -- multigrid-preconditioned CG for the MAC pressure Poisson
// problem (3D/mgpcg.py, 2D/mgpcg.py), restated for gfx950:
//   - every level is a dense array (the reference's pointer SNodes are fully
//     activated by the first solve, so they behave as dense); a 2D level of
//     n0 x n1 cells is addressed as the 3D level (1, n0, n1), so one set of
//     stencil kernels serves both dimensions;
//   - red/black Gauss-Seidel, Ap, restriction and prolongation use the
//     reference's operation order in fp32 without contraction, so each pass
//     is bit-identical to a sequential restatement; restriction is a gather of
//     the 2^d children in fixed order instead of float atomics;
//   - inside the V-cycle two colour sweeps run per launch on an LDS tile with
//     a two-cell halo (the first sweep also updates the inner halo ring, so
//     the second needs nothing from other tiles); the zero fill before the
//     pre-smoother, the prolongation before the post-smoother, the residual
//     recentring and the z.r product are folded into those launches, and the
//     residual + restriction runs on an LDS tile as well;
//   - the search-direction update p = z + beta p is folded into the Ap launch
//     (p ping-pongs between two buffers so neighbouring tiles read old p);
//   - dot products accumulate in FP64 (wave shuffles, then a block tree, then
//     a fixed-order sum of the per-block partials); alpha, beta and the
//     recentring mean are formed on the device, so the only host round trip
//     per iteration is one 8-byte read of r.r for the convergence test;
//   - the coarsest level's 50 red/black sweeps run in one workgroup out of LDS.
#include <algorithm>
#include <cstddef>
#include <new>
#include <type_traits>

#include <hip/hip_bf16.h>
#include <hip/hip_fp16.h>

#include "pfm_host.h"
#include "pfm_div.h"

namespace pfm {

constexpr int kMaxLevels = 16;
constexpr int kMgRedBlocks = 1024;

struct MGScalars {
    double zTr;     // old z.r
    double pAp;
    // read by the host in one copy: r.r, x.x (the fp32 floor of the true
    // residual), and the iteration counter and stop flags of solve()
    double rTr;
    double xTx;
    int it, done, conv;
    double r0;      // r.r of the recentred initial residual
    double sum_r;
    double new_zTr;
    float alpha, beta, mean;
    // 16-bit V-cycle storage: the preconditioner is linear, so it runs on
    // vs * r (vs a power of two near 1/rms(r), exact to apply and undo) and
    // its output is read back times vinv = 1/vs; keeps values in fp16 range
    float vs, vinv;
    // Convergence on the device (solve): the threshold on r.r (tol, or
    // tol^2 r0 for a relative tolerance).  Once `done` is set, the kernels of
    // the iterations still enqueued change nothing the solve returns: k_finalize
    // leaves every scalar alone and k_update_xr leaves x and r alone (the
    // V-cycle and Ap launches only write scratch), so the host can enqueue
    // several iterations (or replay them from a graph) and read the outcome once.
    double tol, thresh;
    int tol_mode, max_it;
};

// Whether the solve has stopped (checked by the kernels that write its state).
__device__ __forceinline__ bool pcg_stopped(const MGScalars* sc) { return sc != nullptr && sc->done != 0; }

struct Level {
    int d[3];      // (1, n0, n1) for a 2D level
    int64_t n;
};

struct MGHandle {
    int dim;
    int N[3];
    int n_levels;
    int coarse_level;       // first level run by k_coarse, -1 if none
    int bc[3][2];
    int dirichlet_counts;
    int precision;
    int vshift;  // extra binades of V-cycle input down-scaling (fp16 storage)
    Level lv[kMaxLevels];
    // level arrays: r[0] is fp32 (the PCG residual); z, z2 and r[l >= 1] are
    // fp32 at precision 0, __half (1) or __hip_bfloat16 (2) otherwise
    void* r[kMaxLevels];
    void* z[kMaxLevels];
    void* z2[kMaxLevels];   // ping-pong partner of z inside the fused sweeps
    float* b;               // right-hand side kept for true-residual replacement (precision 1/2)
    float* x;
    float* p;               // current search direction
    float* p2;              // its ping-pong partner
    float* Ap;
    double* partial;        // max(2 * kMgRedBlocks, tiles of level 0)
    int n_partial;
    MGScalars* sc;          // device
    double* host_rtr;       // pinned
    // solid obstacle (extension; nullptr without one): per level, 1 = solid,
    // and the diagonal (0 for a solid cell)
    uint8_t* solid[kMaxLevels];
    uint8_t* dg[kMaxLevels];
    int64_t n_fluid;        // cells that carry pressure (the recentring mean)
    // PCG iterations on the device: the first group of a solve enqueues as
    // many iterations as the shorter of the previous two solves took (a step
    // alternates two projections of different lengths), then one per group,
    // with one host read per group.  Each iteration can be replayed from a HIP
    // graph (one per p-buffer parity and replacement step), captured on a
    // private non-blocking stream.
    int last_iters[2];      // iterations of the previous two solves
    int use_graphs;
    float* p_base;          // the allocation first held by p (graph parity key)
    hipStream_t cap_stream;
    hipGraphExec_t iter_graph[2][2];
    int graph_failed;
    double graph_eps;       // the eps baked into the captured launches
};

// Captured iterations hold the level descriptors (obstacle masks), the fluid
// cell count and eps by value: drop them when any of those changes.
static void drop_graphs(MGHandle* h) {
    for (auto& row : h->iter_graph)
        for (auto& ex : row)
            if (ex) {
                (void)hipGraphExecDestroy(ex);
                ex = nullptr;
            }
}

// Stencil description of one level, always in 3D index form.
struct LvlDev {
    int nb2;                 // 2 * dim: neighbour count of an interior cell
    int d[3];
    int bc_lo[3], bc_hi[3];  // 1 when that wall decrements the diagonal
    // With a solid obstacle: the diagonal of every cell (open walls and solid
    // neighbours removed), 0 for a solid cell; nullptr without one.  Solid
    // cells hold z = r = p = x = 0, so as neighbours they add +0.
    const uint8_t* dg;
};

__device__ __forceinline__ void cell_ijk(const LvlDev& L, int64_t q, int I[3]) {
    I[2] = (int)(q % L.d[2]);
    I[1] = (int)((q / L.d[2]) % L.d[1]);
    I[0] = (int)(q / ((int64_t)L.d[2] * L.d[1]));
}

// Thread -> cell mapping of the per-cell kernels: 64 x 4 threads per block,
// the contiguous axis on threadIdx.x, axis 0 on blockIdx.z.  With `pairs`, x
// counts pairs of cells along the contiguous axis and the thread takes the
// member whose colour (I.sum() & 1) equals `phase`.
constexpr int kTx = 64, kTy = 4;

__device__ __forceinline__ bool cell_of_thread(const LvlDev& L, int phase, bool pairs, int I[3]) {
    const int fx = blockIdx.x * kTx + threadIdx.x;
    I[0] = blockIdx.z;
    I[1] = blockIdx.y * kTy + threadIdx.y;
    I[2] = pairs ? 2 * fx + ((I[0] + I[1] + phase) & 1) : fx;
    return I[0] < L.d[0] && I[1] < L.d[1] && I[2] < L.d[2];
}

static inline dim3 cell_grid(const int d[3], bool pairs) {
    const int nx = pairs ? (d[2] + 1) / 2 : d[2];
    return dim3((unsigned)((nx + kTx - 1) / kTx), (unsigned)((d[1] + kTy - 1) / kTy), (unsigned)d[0]);
}

__device__ __forceinline__ int64_t cell_index(const LvlDev& L, const int I[3]) {
    return ((int64_t)I[0] * L.d[1] + I[1]) * L.d[2] + I[2];
}

__device__ __forceinline__ bool inside(const LvlDev& L, int g0, int g1, int g2) {
    return (unsigned)g0 < (unsigned)L.d[0] && (unsigned)g1 < (unsigned)L.d[1] && (unsigned)g2 < (unsigned)L.d[2];
}

// num_fluid_neighbors: 3D/mgpcg.py:106-115 (with an obstacle, the stored diagonal).
__device__ __forceinline__ float nfn(const LvlDev& L, int i0, int i1, int i2) {
    if (L.dg) return (float)L.dg[((int64_t)i0 * L.d[1] + i1) * L.d[2] + i2];
    float num = (float)L.nb2;
    const int I[3] = {i0, i1, i2};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        if (I[a] <= 0 && L.bc_lo[a]) num -= 1.0f;
        if (I[a] >= L.d[a] - 1 && L.bc_hi[a]) num -= 1.0f;
    }
    return num;
}

__device__ __forceinline__ float nfn(const LvlDev& L, const int I[3]) { return nfn(L, I[0], I[1], I[2]); }

__device__ __forceinline__ bool solid_at(const LvlDev& L, int64_t q) { return L.dg && L.dg[q] == 0; }

// neighbor_sum: 3D/mgpcg.py:92-104 (per axis: +e then -e).  `x` points at
// the cell; s0 / s1 are the strides of axes 0 / 1 (axis 2 has stride 1).
__device__ __forceinline__ float nsum_at(const LvlDev& L, const float* x, int64_t s0, int64_t s1, int i0, int i1,
                                         int i2) {
    float ret = 0.0f;
    if (i0 < L.d[0] - 1) ret = __fadd_rn(ret, x[s0]);
    if (i0 > 0) ret = __fadd_rn(ret, x[-s0]);
    if (i1 < L.d[1] - 1) ret = __fadd_rn(ret, x[s1]);
    if (i1 > 0) ret = __fadd_rn(ret, x[-s1]);
    if (i2 < L.d[2] - 1) ret = __fadd_rn(ret, x[1]);
    if (i2 > 0) ret = __fadd_rn(ret, x[-1]);
    return ret;
}

__device__ __forceinline__ float nsum(const LvlDev& L, const float* __restrict__ x, int64_t q, const int I[3]) {
    return nsum_at(L, x + q, (int64_t)L.d[1] * L.d[2], L.d[2], I[0], I[1], I[2]);
}

__global__ void __launch_bounds__(kBlock) k_fill_f32(float* p, int64_t n, float v) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

// ---- single-pass kernels (the reference's kernels one by one) -------------

// smooth: 3D/mgpcg.py:165-172 (one colour; one thread per cell of that colour).
__global__ void __launch_bounds__(kBlock) k_smooth(LvlDev L, float* __restrict__ z,
                                                   const float* __restrict__ r, int phase) {
    int I[3];
    if (!cell_of_thread(L, phase, true, I)) return;
    const int64_t q = cell_index(L, I);
    if (solid_at(L, q)) return;
    z[q] = __fdiv_rn(__fadd_rn(r[q], nsum(L, z, q, I)), nfn(L, I));
}

// compute_Ap (3D/mgpcg.py:117-122) without the reduction (get_Ap).
__global__ void __launch_bounds__(kBlock) k_apply_A(LvlDev L, const float* __restrict__ p,
                                                    float* __restrict__ Ap) {
    int I[3];
    if (!cell_of_thread(L, 0, false, I)) return;
    const int64_t q = cell_index(L, I);
    Ap[q] = solid_at(L, q) ? 0.0f : __fsub_rn(__fmul_rn(nfn(L, I), p[q]), nsum(L, p, q, I));
}

// restrict (3D/mgpcg.py:152-158) as a gather over the 2^d children.  Any
// extents: the last coarse cell of an odd fine extent also owns the trailing
// fine cell (fine I maps to min(I/2, coarse-1)).
__global__ void __launch_bounds__(kBlock) k_restrict(LvlDev Lf, LvlDev Lc, float scale,
                                                     const float* __restrict__ z,
                                                     const float* __restrict__ r,
                                                     float* __restrict__ rc) {
    int C[3];
    if (!cell_of_thread(Lc, 0, false, C)) return;
    const int64_t qc = cell_index(Lc, C);
    int hi[3];
    for (int a = 0; a < 3; ++a) hi[a] = (C[a] == Lc.d[a] - 1) ? Lf.d[a] - 1 - 2 * C[a] : 1;
    float acc = 0.0f;
    for (int a = 0; a <= hi[0]; ++a)
        for (int b = 0; b <= hi[1]; ++b)
            for (int c = 0; c <= hi[2]; ++c) {
                int I[3] = {2 * C[0] + a, 2 * C[1] + b, 2 * C[2] + c};
                const int64_t q = cell_index(Lf, I);
                if (solid_at(Lf, q)) continue;
                const float res = __fsub_rn(r[q], __fsub_rn(__fmul_rn(nfn(Lf, I), z[q]), nsum(Lf, z, q, I)));
                acc = __fadd_rn(acc, __fdiv_rn(__fmul_rn(res, 1.0f), scale));
            }
    rc[qc] = acc;
}

__device__ __forceinline__ int64_t parent_index(const LvlDev& Lc, int i0, int i1, int i2) {
    const int c0 = min(i0 >> 1, Lc.d[0] - 1), c1 = min(i1 >> 1, Lc.d[1] - 1), c2 = min(i2 >> 1, Lc.d[2] - 1);
    return ((int64_t)c0 * Lc.d[1] + c1) * Lc.d[2] + c2;
}

// prolongate: 3D/mgpcg.py:160-163.
__global__ void __launch_bounds__(kBlock) k_prolongate(LvlDev Lf, LvlDev Lc, float* __restrict__ z,
                                                       const float* __restrict__ zc) {
    int I[3];
    if (!cell_of_thread(Lf, 0, false, I)) return;
    const int64_t q = cell_index(Lf, I);
    if (solid_at(Lf, q)) return;  // solid cells carry no pressure
    z[q] = __fadd_rn(z[q], zc[parent_index(Lc, I[0], I[1], I[2])]);
}

// ---- LDS-tiled V-cycle kernels --------------------------------------------

// Interior tile of the sweep and Ap kernels along (axis 0, axis 1, axis 2).
template <int DIM> struct MgTile;
template <> struct MgTile<3> { static constexpr int I = 8, J = 16, K = 32; };
template <> struct MgTile<2> { static constexpr int I = 1, J = 32, K = 64; };
// Coarse tile of the restriction (the fine tile is twice as large per axis,
// except axis 0 of a 2D level, which has extent 1 on every level).
template <int DIM> struct RsTile;
template <> struct RsTile<3> { static constexpr int I = 4, J = 8, K = 16; };
template <> struct RsTile<2> { static constexpr int I = 1, J = 16, K = 32; };

struct TileMap {
    int n0, n1, n2;  // tiles along each axis
    int total;       // tiles in the launch (= blocks)
};

// Tile of this block.  Workgroups are dealt to the 8 XCDs round robin, so the
// block index is permuted to give each XCD a contiguous run of tiles, and
// tiles are numbered with axis 0 fastest: the tiles that share halo cells
// with a tile ran just before it on the same XCD, so the halo is still in
// that XCD's L2.
//
// When the axis-0 tile count is a multiple of 8, XCD x instead takes the slab
// of axis-0 tiles [x q, x q + q) (q = n0 / 8) and walks it with axis 0
// fastest, then axis 2, then axis 1, so a tile's six neighbours are at most
// q n2 dispatches apart on the same XCD and their shared halo lines are still
// in its L2 (with axis 1 at n0 n2 dispatches apart, its halos were fetched
// twice from memory); only the slab faces are fetched by two XCDs.
__device__ __forceinline__ void tile_of_block(const TileMap& tm, int t[3]) {
    if ((tm.n0 & 7) == 0) {
        const int q = tm.n0 >> 3, x = blockIdx.x & 7, k = blockIdx.x >> 3;
        const int rest = k / q;
        t[0] = x * q + (k - rest * q);
        t[2] = rest % tm.n2;
        t[1] = rest / tm.n2;
        return;
    }
    int b = blockIdx.x;
    if ((tm.total & 7) == 0) b = (b & 7) * (tm.total >> 3) + (b >> 3);
    t[0] = b % tm.n0;
    b /= tm.n0;
    t[2] = b % tm.n2;
    t[1] = b / tm.n2;
}

__device__ __forceinline__ float to_f(float v) { return v; }
__device__ __forceinline__ float to_f(__half v) { return __half2float(v); }
__device__ __forceinline__ float to_f(__hip_bfloat16 v) { return __bfloat162float(v); }
template <typename V> __device__ __forceinline__ V from_f(float v);
template <> __device__ __forceinline__ float from_f<float>(float v) { return v; }
// fp16 stores saturate at the largest finite half, so an outlier can never
// turn the preconditioner's output into inf / NaN.
template <> __device__ __forceinline__ __half from_f<__half>(float v) {
    return __float2half(fminf(fmaxf(v, -65504.0f), 65504.0f));
}
template <> __device__ __forceinline__ __hip_bfloat16 from_f<__hip_bfloat16>(float v) { return __float2bfloat16(v); }

// restrict (3D/mgpcg.py:152-158) for any extents with 16-bit z and coarse r
// (the mixed-precision solver; the tiled kernel needs exactly halving
// extents): the gather of k_restrict over typed storage.  On level 0 (fp32 r,
// 16-bit z) r enters as vs * r, as in the tiled kernels.
template <typename ZT, typename RT>
__global__ void __launch_bounds__(kBlock) k_restrict_any(LvlDev Lf, LvlDev Lc, float inv_scale,
                                                         const ZT* __restrict__ z, const RT* __restrict__ r,
                                                         ZT* __restrict__ rc, const MGScalars* __restrict__ sc) {
    int C[3];
    if (!cell_of_thread(Lc, 0, false, C)) return;
    const int64_t qc = cell_index(Lc, C);
    const float vs = (sizeof(ZT) == 2 && sizeof(RT) == 4) ? sc->vs : 1.0f;
    const int64_t s0 = (int64_t)Lf.d[1] * Lf.d[2], s1 = Lf.d[2];
    int hi[3];
    for (int a = 0; a < 3; ++a) hi[a] = (C[a] == Lc.d[a] - 1) ? Lf.d[a] - 1 - 2 * C[a] : 1;
    float acc = 0.0f;
    for (int a = 0; a <= hi[0]; ++a)
        for (int b = 0; b <= hi[1]; ++b)
            for (int c = 0; c <= hi[2]; ++c) {
                int I[3] = {2 * C[0] + a, 2 * C[1] + b, 2 * C[2] + c};
                const int64_t q = cell_index(Lf, I);
                if (solid_at(Lf, q)) continue;
                const ZT* x = z + q;
                float ns = 0.0f;
                if (I[0] < Lf.d[0] - 1) ns = __fadd_rn(ns, to_f(x[s0]));
                if (I[0] > 0) ns = __fadd_rn(ns, to_f(x[-s0]));
                if (I[1] < Lf.d[1] - 1) ns = __fadd_rn(ns, to_f(x[s1]));
                if (I[1] > 0) ns = __fadd_rn(ns, to_f(x[-s1]));
                if (I[2] < Lf.d[2] - 1) ns = __fadd_rn(ns, to_f(x[1]));
                if (I[2] > 0) ns = __fadd_rn(ns, to_f(x[-1]));
                float rq = to_f(r[q]);
                if (sizeof(ZT) == 2 && sizeof(RT) == 4) rq *= vs;
                const float res = __fsub_rn(rq, __fsub_rn(__fmul_rn(nfn(Lf, I), to_f(x[0])), ns));
                acc = __fadd_rn(acc, __fmul_rn(__fmul_rn(res, 1.0f), inv_scale));
            }
    rc[qc] = from_f<ZT>(acc);
}

// Buffer view of a level array: the tiled kernels load through 32-bit byte
// offsets, and an offset past the end (kOOB) reads 0, which is the value the
// LDS tiles hold for cells outside the level.  Loads are issued without
// branches, so all of a thread's loads can be in flight at once.  kOOB plus
// the widest load (16 bytes) stays below 2^32, so the range check cannot wrap,
// and it exceeds every level array's size (pfm_mg_create caps the grid).
constexpr uint32_t kOOB = 0xFFFFFFC0u;

// Level arrays are stored as V = float, or (mixed precision) __half or
// __hip_bfloat16; every kernel computes in fp32 and rounds to nearest on store.
template <typename V>
__device__ __forceinline__ __amdgpu_buffer_rsrc_t lvl_rsrc(const V* p, const LvlDev& L) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<V*>(p), (short)0,
                                             (int)((uint32_t)L.d[0] * (uint32_t)L.d[1] * (uint32_t)L.d[2] *
                                                   (uint32_t)sizeof(V)),
                                             0x00020000);
}

template <typename V>
__device__ __forceinline__ float ldv(__amdgpu_buffer_rsrc_t rs, uint32_t boff) {
    if constexpr (sizeof(V) == 4) {
        return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rs, boff, 0, 0));
    } else {
        const unsigned short b = __builtin_amdgcn_raw_buffer_load_b16(rs, boff, 0, 0);
        return to_f(__builtin_bit_cast(V, b));
    }
}

// Byte offset of cell g in an array of V, or kOOB outside the level.
template <typename V>
__device__ __forceinline__ uint32_t cell_off(const LvlDev& L, int g0, int g1, int g2) {
    return inside(L, g0, g1, g2) ? (uint32_t)((g0 * L.d[1] + g1) * L.d[2] + g2) * (uint32_t)sizeof(V) : kOOB;
}

// The obstacle's diagonal byte of cell g (0 = solid; 0 outside the level).
__device__ __forceinline__ __amdgpu_buffer_rsrc_t dg_rsrc(const LvlDev& L) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<uint8_t*>(L.dg), (short)0,
                                             (int)((uint32_t)L.d[0] * (uint32_t)L.d[1] * (uint32_t)L.d[2]),
                                             0x00020000);
}

__device__ __forceinline__ int ld_dg(__amdgpu_buffer_rsrc_t rs, const LvlDev& L, int g0, int g1, int g2) {
    const uint32_t off = inside(L, g0, g1, g2) ? (uint32_t)((g0 * L.d[1] + g1) * L.d[2] + g2) : kOOB;
    return (int)__builtin_amdgcn_raw_buffer_load_b8(rs, off, 0, 0);
}

// Walls next to cell g that decrement its diagonal (num_fluid_neighbors).
__device__ __forceinline__ int walls(const LvlDev& L, int g0, int g1, int g2) {
    return (g0 <= 0 ? L.bc_lo[0] : 0) + (g0 >= L.d[0] - 1 ? L.bc_hi[0] : 0) + (g1 <= 0 ? L.bc_lo[1] : 0) +
           (g1 >= L.d[1] - 1 ? L.bc_hi[1] : 0) + (g2 <= 0 ? L.bc_lo[2] : 0) + (g2 >= L.d[2] - 1 ? L.bc_hi[2] : 0);
}

// Does every cell of the box [lo, hi) lie strictly inside the level (no
// wall and all six neighbours present)?
__device__ __forceinline__ bool box_interior(const LvlDev& L, int lo0, int lo1, int lo2, int hi0, int hi1, int hi2,
                                             bool axis0) {
    return (!axis0 || (lo0 >= 1 && hi0 <= L.d[0] - 1)) && lo1 >= 1 && hi1 <= L.d[1] - 1 && lo2 >= 1 &&
           hi2 <= L.d[2] - 1;
}

// neighbor_sum (3D/mgpcg.py:92-104) on an LDS tile.  A neighbour outside the
// level is an LDS cell holding +0, and adding +0 never changes the running
// sum (it starts at +0 and so can never be -0), so the reference's skipped
// terms need no test.  2D levels have no neighbours along axis 0.
template <int DIM>
__device__ __forceinline__ float nsum_lds(const float* x, int sp, int sk) {
    float ret = 0.0f;
    if (DIM == 3) {
        ret = __fadd_rn(ret, x[sp]);
        ret = __fadd_rn(ret, x[-sp]);
    }
    ret = __fadd_rn(ret, x[sk]);
    ret = __fadd_rn(ret, x[-sk]);
    ret = __fadd_rn(ret, x[1]);
    ret = __fadd_rn(ret, x[-1]);
    return ret;
}

// Diagonal n (as float) and RN(1/n) of cell g; `interior` (block uniform)
// skips the wall tests.
struct Diag {
    float n, y;
};

__device__ __forceinline__ Diag diag(const LvlDev& L, bool interior, float n_in, float y_in, int g0, int g1,
                                     int g2) {
    Diag dg{n_in, y_in};
    if (!interior) {
        const int n = L.nb2 - walls(L, g0, g1, g2);
        dg.n = (float)n;
        dg.y = pfm_rcp_small(n);
    }
    return dg;
}

// Threads of a sweep block: 5 waves, so the 306 (j, k-pair) columns of the
// first pass of a 3D tile take one round.
constexpr int kSweepThreads = 320;

// Two red/black passes (colours c0, then c0 ^ 1) over one tile: the
// reference's smooth(l, c0); smooth(l, c0 ^ 1) (3D/mgpcg.py:165-172,
// 186-188/197-199).  The first pass updates its colour on the tile plus a
// one-cell ring, reading the two-cell halo of z_in; the second updates the
// other colour on the tile, so every cell sees exactly the values a
// sequential sweep would.  Options:
//   ZERO_IN  z_in is all zero (the pre-smoother after z.fill(0), :185,190);
//   PROLONG  z_in += zc[parent] on load (prolongate, :160-163,196);
//   SUB_MEAN use r - mean (the recentring of solve, 3D/mgpcg.py:174-180);
//   DOT      write per-block partials of z_out . r.
// ZT: storage of z (float, or 16-bit in mixed precision); RT: storage of r
// (float on level 0, ZT on coarser levels).
//
// Work layout: every thread owns fixed columns along axis 0 -- a (j, k)
// position of the tile face for the loads and stores, a (j, pair of k)
// position for the passes -- and walks down the column with incremental
// addresses, so no per-cell index decoding is needed.  The colour member of
// a pair alternates from one axis-0 plane to the next.  All global loads
// (z tile, r of the first pass) are issued before the first barrier, the
// second pass's r right after it.
template <int DIM, typename ZT, typename RT, bool ZERO_IN, bool PROLONG, bool SUB_MEAN, bool DOT>
__global__ void __launch_bounds__(kSweepThreads, 3) k_rbgs(LvlDev L, LvlDev Lc, TileMap tm,
                                                        const ZT* __restrict__ zin, const ZT* __restrict__ zc,
                                                        const RT* __restrict__ r, ZT* __restrict__ zout, int c0,
                                                        const MGScalars* __restrict__ sc,
                                                        double* __restrict__ partial) {
    using T = MgTile<DIM>;
    constexpr int NT = kSweepThreads;
    constexpr int H = 2, HI = DIM == 3 ? H : 0, A0 = DIM == 3 ? 1 : 0;
    constexpr int SI = T::I + 2 * HI, SJ = T::J + 2 * H, SK = T::K + 2 * H, SP = SJ * SK, NS = SI * SP;
    constexpr int ZC = SJ * SK, ZR = (ZC + NT - 1) / NT;                      // z tile columns, rounds
    constexpr int P0I = T::I + 2 * A0, P0P = (T::K + 2) / 2, P0C = (T::J + 2) * P0P;
    constexpr int P0R = (P0C + NT - 1) / NT;                                  // pass 0: tile + ring
    constexpr int P1I = T::I, P1P = T::K / 2, P1C = T::J * P1P, P1R = (P1C + NT - 1) / NT;  // pass 1: tile
    constexpr int WC = T::J * T::K, WR = (WC + NT - 1) / NT;                  // interior columns
    static_assert(T::K % 2 == 0, "pairs along axis 2");
    __shared__ float zs[NS];
    int t[3];
    tile_of_block(tm, t);
    const int o0 = t[0] * T::I, o1 = t[1] * T::J, o2 = t[2] * T::K;
    const int b0 = o0 - HI, b1 = o1 - H, b2 = o2 - H;  // global cell of zs[0]
    const int d0 = L.d[0], d1 = L.d[1], d2 = L.d[2];
    const int plane = d1 * d2;
    const int tid = threadIdx.x;
    const int c1 = c0 ^ 1;
    const __amdgpu_buffer_rsrc_t rr = lvl_rsrc(r, L);
    // ---- z tile into registers
    float zv[ZR][SI];
    if (!ZERO_IN) {
        const __amdgpu_buffer_rsrc_t rz = lvl_rsrc(zin, L);
        const __amdgpu_buffer_rsrc_t rc = lvl_rsrc(zc, Lc);
#pragma unroll
        for (int m = 0; m < ZR; ++m) {
            const int p = tid + m * NT;
            const int lj = p / SK, lk = p - lj * SK;
            const int g1 = b1 + lj, g2 = b2 + lk;
            const bool col = p < ZC && (unsigned)g1 < (unsigned)d1 && (unsigned)g2 < (unsigned)d2;
            const int base = g1 * d2 + g2;
#pragma unroll
            for (int li = 0; li < SI; ++li) {
                const int g0 = b0 + li;
                const bool in = col && (unsigned)g0 < (unsigned)d0;
                zv[m][li] = ldv<ZT>(rz, in ? (uint32_t)(g0 * plane + base) * (uint32_t)sizeof(ZT) : kOOB);
            }
            if (PROLONG) {
                // b0 is even, so cells li = 2h and 2h + 1 share the coarse
                // parent along axis 0: one coarse load per pair
                static_assert(DIM == 2 || (MgTile<3>::I % 2 == 0 && HI % 2 == 0), "even axis-0 tile origin");
                const int e1 = Lc.d[1], e2 = Lc.d[2];
                const int pcol = min(g1 >> 1, e1 - 1) * e2 + min(g2 >> 1, e2 - 1);
                const int pplane = e1 * e2;
                constexpr int SH = (SI + 1) / 2;
                float zcv[SH];
#pragma unroll
                for (int h = 0; h < SH; ++h) {
                    const int g0 = b0 + 2 * h;
                    const bool in = col && (unsigned)g0 < (unsigned)d0;
                    const int pc = min(g0 >> 1, Lc.d[0] - 1) * pplane + pcol;
                    zcv[h] = ldv<ZT>(rc, in ? (uint32_t)pc * (uint32_t)sizeof(ZT) : kOOB);
                }
                // outside the level the z load reads 0 and 0 + 0 stays +0
#pragma unroll
                for (int li = 0; li < SI; ++li) {
                    const bool in = col && (unsigned)(b0 + li) < (unsigned)d0;
                    zv[m][li] = __fadd_rn(zv[m][li], in ? zcv[li >> 1] : 0.f);
                }
                if (L.dg) {
                    // solid cells take no prolongation (they carry no pressure)
                    const __amdgpu_buffer_rsrc_t rd = dg_rsrc(L);
                    int sv[SI];
#pragma unroll
                    for (int li = 0; li < SI; ++li) {
                        const int g0 = b0 + li;
                        const bool in = col && (unsigned)g0 < (unsigned)d0;
                        sv[li] = (int)__builtin_amdgcn_raw_buffer_load_b8(rd, in ? (uint32_t)(g0 * plane + base)
                                                                                 : kOOB, 0, 0);
                    }
#pragma unroll
                    for (int li = 0; li < SI; ++li) zv[m][li] = sv[li] == 0 ? 0.f : zv[m][li];
                }
            }
        }
    }
    // ---- r of the first pass
    float rv0[P0R][P0I];
#pragma unroll
    for (int m = 0; m < P0R; ++m) {
        const int p = tid + m * NT;
        const int rj = p / P0P, pk = p - rj * P0P;
        const int g1 = b1 + 1 + rj, gk0 = b2 + 1 + 2 * pk;
        const bool col = p < P0C && (unsigned)g1 < (unsigned)d1;
#pragma unroll
        for (int ri = 0; ri < P0I; ++ri) {
            const int g0 = b0 + HI - A0 + ri;
            const int g2 = gk0 + ((c0 + g0 + g1 + gk0) & 1);
            const bool in = col && (unsigned)g0 < (unsigned)d0 && (unsigned)g2 < (unsigned)d2;
            rv0[m][ri] = ldv<RT>(rr, in ? (uint32_t)(g0 * plane + g1 * d2 + g2) * (uint32_t)sizeof(RT) : kOOB);
        }
    }
    // ---- z tile into LDS (LDS index = li * SP + p with p = lj * SK + lk)
#pragma unroll
    for (int m = 0; m < ZR; ++m) {
        const int p = tid + m * NT;
        if (p < ZC) {
#pragma unroll
            for (int li = 0; li < SI; ++li) zs[li * SP + p] = ZERO_IN ? 0.f : zv[m][li];
        }
    }
    const float mean = SUB_MEAN ? sc->mean : 0.f;
    // 16-bit z: level 0 works on vs * r (coarser r arrive scaled)
    constexpr bool kScale = sizeof(ZT) == 2 && sizeof(RT) == 4;
    const float vs = kScale ? sc->vs : 1.0f;
    // the pass-0 region [o - A0, o + T + A0) x ... decides the interior case
    const bool interior = !L.dg && box_interior(L, o0 - A0, o1 - 1, o2 - 1, o0 + T::I + A0, o1 + T::J + 1,
                                                o2 + T::K + 1, DIM == 3);
    const bool interior1 = !L.dg && box_interior(L, o0, o1, o2, o0 + T::I, o1 + T::J, o2 + T::K, DIM == 3);
    const float n_int = (float)L.nb2, y_int = pfm_rcp_small(L.nb2);
    double s = 0.0;
    __syncthreads();
    // the thread index is hidden from the optimiser past the barrier so the
    // column coordinates are recomputed rather than held in registers
    int tid2 = tid;
    asm volatile("" : "+v"(tid2));
    // ---- r of the second pass arrives while the first one runs
    float rv1[P1R][P1I];
#pragma unroll
    for (int m = 0; m < P1R; ++m) {
        const int p = tid2 + m * NT;
        const int rj = p / P1P, pk = p - rj * P1P;
        const int g1 = b1 + 2 + rj, gk0 = b2 + 2 + 2 * pk;
        const bool col = p < P1C && (unsigned)g1 < (unsigned)d1;
#pragma unroll
        for (int ri = 0; ri < P1I; ++ri) {
            const int g0 = b0 + HI + ri;
            const int g2 = gk0 + ((c1 + g0 + g1 + gk0) & 1);
            const bool in = col && (unsigned)g0 < (unsigned)d0 && (unsigned)g2 < (unsigned)d2;
            rv1[m][ri] = ldv<RT>(rr, in ? (uint32_t)(g0 * plane + g1 * d2 + g2) * (uint32_t)sizeof(RT) : kOOB);
        }
    }
    // one cell update of a pass, given its neighbour sum; returns whether the
    // cell was updated.  A pass writes only cells of its colour and reads only
    // the other colour, so a column's neighbour sums are all gathered first
    // (their LDS reads overlap) and the stores follow.
    auto update = [&](int le, int g0, int g1, int g2, bool in, bool intr, float rin, float ns, float& zval,
                      float& rraw) -> bool {
        rraw = SUB_MEAN ? __fsub_rn(rin, mean) : rin;
        const float rv = kScale ? rraw * vs : rraw;
        Diag dg = diag(L, intr, n_int, y_int, g0, g1, g2);
        bool fluid = true;
        if (L.dg) {
            const int n = ld_dg(dg_rsrc(L), L, g0, g1, g2);
            dg = Diag{(float)n, pfm_rcp_small(n)};
            fluid = n != 0;
        }
        zval = pfm_div_small(__fadd_rn(rv, ns), dg.n, dg.y);
        const bool upd = in && fluid;
        if (upd) zs[le] = zval;
        return upd;
    };
    // ---- pass 0 (colour c0) on the tile and its one-cell ring
#pragma unroll
    for (int m = 0; m < P0R; ++m) {
        const int p = tid2 + m * NT;
        const int rj = p / P0P, pk = p - rj * P0P;
        const int lj = 1 + rj, g1 = b1 + lj;
        const int lk0 = 1 + 2 * pk, gk0 = b2 + lk0;
        const bool col = p < P0C && (unsigned)g1 < (unsigned)d1;
        float ns[P0I];
#pragma unroll
        for (int ri = 0; ri < P0I; ++ri) {
            const int li = HI - A0 + ri, g0 = b0 + li;
            const int off = (c0 + g0 + g1 + gk0) & 1;
            ns[ri] = nsum_lds<DIM>(zs + (li * SJ + lj) * SK + lk0 + off, SP, SK);
        }
#pragma unroll
        for (int ri = 0; ri < P0I; ++ri) {
            const int li = HI - A0 + ri, g0 = b0 + li;
            const int off = (c0 + g0 + g1 + gk0) & 1;
            const int g2 = gk0 + off;
            const bool in = col && (unsigned)g0 < (unsigned)d0 && (unsigned)g2 < (unsigned)d2;
            float v, rraw;
            const bool upd = update((li * SJ + lj) * SK + lk0 + off, g0, g1, g2, in, interior, rv0[m][ri], ns[ri], v,
                                    rraw);
            // cells of colour c0 inside the tile are final after this pass (the
            // product uses z as stored)
            if (DOT && upd && g0 >= o0 && g0 < o0 + T::I && g1 >= o1 && g1 < o1 + T::J && g2 >= o2 &&
                g2 < o2 + T::K)
                s += (double)to_f(from_f<ZT>(v)) * (double)rraw;
        }
    }
    __syncthreads();
    // ---- pass 1 (colour c1) on the tile
#pragma unroll
    for (int m = 0; m < P1R; ++m) {
        const int p = tid2 + m * NT;
        const int rj = p / P1P, pk = p - rj * P1P;
        const int lj = 2 + rj, g1 = b1 + lj;
        const int lk0 = 2 + 2 * pk, gk0 = b2 + lk0;
        const bool col = p < P1C && (unsigned)g1 < (unsigned)d1;
        float ns[P1I];
#pragma unroll
        for (int ri = 0; ri < P1I; ++ri) {
            const int li = HI + ri, g0 = b0 + li;
            const int off = (c1 + g0 + g1 + gk0) & 1;
            ns[ri] = nsum_lds<DIM>(zs + (li * SJ + lj) * SK + lk0 + off, SP, SK);
        }
#pragma unroll
        for (int ri = 0; ri < P1I; ++ri) {
            const int li = HI + ri, g0 = b0 + li;
            const int off = (c1 + g0 + g1 + gk0) & 1;
            const int g2 = gk0 + off;
            const bool in = col && (unsigned)g0 < (unsigned)d0 && (unsigned)g2 < (unsigned)d2;
            float v, rraw;
            const bool upd = update((li * SJ + lj) * SK + lk0 + off, g0, g1, g2, in, interior1, rv1[m][ri], ns[ri],
                                    v, rraw);
            if (DOT && upd) s += (double)to_f(from_f<ZT>(v)) * (double)rraw;
        }
    }
    __syncthreads();
    // ---- interior to z_out
#pragma unroll
    for (int m = 0; m < WR; ++m) {
        const int p = tid + m * NT;
        const int tj = p / T::K, tk = p - tj * T::K;
        const int g1 = o1 + tj, g2 = o2 + tk;
        if (p < WC && g1 < d1 && g2 < d2) {
            const int base = g1 * d2 + g2;
#pragma unroll
            for (int ti = 0; ti < T::I; ++ti) {
                const int g0 = o0 + ti;
                if (g0 < d0) zout[g0 * plane + base] = from_f<ZT>(zs[((ti + HI) * SJ + tj + H) * SK + tk + H]);
            }
        }
    }
    if (DOT) {
        __shared__ double sh[NT / 64];
        s = block_sum(s, sh);
        if (tid == 0) partial[blockIdx.x] = kScale ? s * (double)sc->vinv : s;  // z was stored scaled
    }
}

// ---- 3D sweeps on groups of 4 cells ---------------------------------------
// Four consecutive cells along axis 2 starting at a multiple of 4 (a "group")
// move as one 16-byte load or store (8 bytes at 16-bit storage).  On a level
// whose axis-2 extent is a multiple of 4 a group lies wholly inside or wholly
// outside the level, so one bounds test covers its four cells.
//
// The 16- and 8-byte loads call the LLVM buffer-load intrinsics by name: this
// compiler lowers __builtin_amdgcn_raw_buffer_load_b64 / _b128 to a single
// dword load.
typedef float pfm_v4f __attribute__((ext_vector_type(4)));
typedef float pfm_v2f __attribute__((ext_vector_type(2)));
__device__ pfm_v4f pfm_buffer_load_v4f32(__amdgpu_buffer_rsrc_t rs, int voffset, int soffset, int aux)
    __asm("llvm.amdgcn.raw.ptr.buffer.load.v4f32");
__device__ pfm_v2f pfm_buffer_load_v2f32(__amdgpu_buffer_rsrc_t rs, int voffset, int soffset, int aux)
    __asm("llvm.amdgcn.raw.ptr.buffer.load.v2f32");

__device__ __forceinline__ unsigned short lo16(float w) { return (unsigned short)(__builtin_bit_cast(uint32_t, w)); }
__device__ __forceinline__ unsigned short hi16(float w) {
    return (unsigned short)(__builtin_bit_cast(uint32_t, w) >> 16);
}

template <typename V>
__device__ __forceinline__ float4 ld4(__amdgpu_buffer_rsrc_t rs, uint32_t boff) {
    float4 v;
    if constexpr (sizeof(V) == 4) {
        const pfm_v4f q = pfm_buffer_load_v4f32(rs, (int)boff, 0, 0);
        v.x = q[0];
        v.y = q[1];
        v.z = q[2];
        v.w = q[3];
    } else {
        const pfm_v2f q = pfm_buffer_load_v2f32(rs, (int)boff, 0, 0);
        v.x = to_f(__builtin_bit_cast(V, lo16(q[0])));
        v.y = to_f(__builtin_bit_cast(V, hi16(q[0])));
        v.z = to_f(__builtin_bit_cast(V, lo16(q[1])));
        v.w = to_f(__builtin_bit_cast(V, hi16(q[1])));
    }
    return v;
}

// Two consecutive values (the coarse parents of a group).
template <typename V>
__device__ __forceinline__ float2 ld2(__amdgpu_buffer_rsrc_t rs, uint32_t boff) {
    float2 v;
    if constexpr (sizeof(V) == 4) {
        const pfm_v2f q = pfm_buffer_load_v2f32(rs, (int)boff, 0, 0);
        v.x = q[0];
        v.y = q[1];
    } else {
        const uint32_t q = __builtin_amdgcn_raw_buffer_load_b32(rs, boff, 0, 0);
        v.x = to_f(__builtin_bit_cast(V, (unsigned short)(q & 0xffffu)));
        v.y = to_f(__builtin_bit_cast(V, (unsigned short)(q >> 16)));
    }
    return v;
}

template <typename V>
__device__ __forceinline__ void st4(V* p, uint32_t idx, float4 v) {
    if constexpr (sizeof(V) == 4) {
        *reinterpret_cast<float4*>(p + idx) = v;
    } else {
        uint2 w;
        w.x = (uint32_t)__builtin_bit_cast(unsigned short, from_f<V>(v.x)) |
              ((uint32_t)__builtin_bit_cast(unsigned short, from_f<V>(v.y)) << 16);
        w.y = (uint32_t)__builtin_bit_cast(unsigned short, from_f<V>(v.z)) |
              ((uint32_t)__builtin_bit_cast(unsigned short, from_f<V>(v.w)) << 16);
        *reinterpret_cast<uint2*>(p + idx) = w;
    }
}

// One colour pass over one group held in an LDS tile: the group's cells of
// colour c are the two at offsets s and s + 2 (s = (c + g0 + g1) & 1, the
// group starting at an even cell).  Their axis-0 / axis-1 neighbours come
// from the neighbouring groups (16-byte LDS reads), the axis-2 neighbours
// from the group itself and the single cells just outside it.
struct GroupPass {
    float4 v;        // the group after the pass
    float za, zb;    // the new values of the two cells
    float ra, rb;    // their r as used (recentred, unscaled)
    bool ua, ub;     // whether each was updated
};

constexpr int kSw3Threads = 256;

// Items of the first pass of k_sweep3: the tile's groups first, numbered as
// the second pass numbers them (so a thread's r of its first four items is
// also the r of its second-pass groups), then the ring along axes 0 / 1 at
// the tile's axis-2 groups, then the edge groups along axis 2.  (ri, rj, lg)
// are the item's offsets from the tile origin minus one cell on axes 0 / 1
// and minus one group on axis 2.
//
// Tile item e (block of 256 threads: round e >> 8, wave (e >> 6) & 3, lane
// e & 63) is group 1 + (lane & 7) of row 1 + 2 (lane >> 3) + (wave & 1) of
// plane 1 + (wave >> 1) + 2 round: the rows of one wave have one parity, so
// every tile group of a wave has the same (axis 0 + axis 1) parity and picks
// the same two cells of a colour.
//
// Tile item e (block of 256 threads: round e >> 8, wave (e >> 6) & 3, lane
// e & 63) is group 1 + (lane & 7) of row 1 + 2 (lane >> 3) + (wave & 1) of
// plane 1 + (wave >> 1) + 2 round: the rows of one wave have one parity, so
// every tile group of a wave has the same (axis 0 + axis 1) parity and picks
// the same two cells of a colour.
template <int TI, int TJ, int TG>
__device__ __forceinline__ void sweep3_item(int e, int& ri, int& rj, int& lg) {
    static_assert(TI == 8 && TJ == 16 && TG == 8, "tile item layout assumes an 8 x 16 x 32 tile and 256 threads");
    constexpr int NTILE = TI * TJ * TG, RW = 2 * (TJ + 2) + 2 * TI, NRING = RW * TG;
    if (e < NTILE) {
        const int w = (e >> 6) & 3, l = e & 63;
        lg = 1 + (l & 7);
        rj = 1 + 2 * (l >> 3) + (w & 1);
        ri = 1 + (w >> 1) + 2 * (e >> 8);
    } else if (e < NTILE + NRING) {
        const int u = e - NTILE;
        lg = 1 + u % TG;
        const int w = u / TG;  // position on the ring of the (TI + 2) x (TJ + 2) rectangle
        const int wa = w - (TJ + 2), wb = w - 2 * (TJ + 2), wc = wb - TI;
        ri = w < TJ + 2 ? 0 : (wa < TJ + 2 ? TI + 1 : 1 + (wb < TI ? wb : wc));
        rj = w < TJ + 2 ? w : (wa < TJ + 2 ? wa : (wb < TI ? 0 : TJ + 1));
    } else {
        const int u = e - NTILE - NRING;
        lg = (u & 1) ? TG + 1 : 0;
        const int pos = u >> 1;
        rj = pos % (TJ + 2);
        ri = pos / (TJ + 2);
    }
}

// The two colour passes of k_rbgs for 3D levels with an axis-2 extent that is
// a multiple of 4, on groups.  The LDS tile holds z on the 8 x 16 x 32 tile
// plus 2 planes and 2 rows on axes 0 and 1 and one group (4 cells) on each
// side along axis 2.  The first pass covers the tile plus its one-cell ring
// along axes 0 and 1 and one group each side along axis 2; of those edge
// groups only the cell next to the tile is needed (the others are computed
// from the edge of the LDS tile and never read).  Each group of the second
// pass (the tile) is written to z_out from registers by the thread that
// updated it.  With a zero input the first pass reads nothing from LDS (its
// neighbours are all +0) and writes every group the second pass reads, so
// the tile is not filled first.  Per cell the arithmetic is exactly k_rbgs's
// (and so the reference's smooth, 3D/mgpcg.py:165-172): both give the same bits.
template <typename ZT, typename RT, bool ZERO_IN, bool PROLONG, bool SUB_MEAN, bool DOT>
__global__ void __launch_bounds__(kSw3Threads, 4) k_sweep3(LvlDev L, LvlDev Lc, TileMap tm,
                                                       const ZT* __restrict__ zin, const ZT* __restrict__ zc,
                                                       const RT* __restrict__ r, ZT* __restrict__ zout, int c0,
                                                       const MGScalars* __restrict__ sc,
                                                       double* __restrict__ partial) {
    using T = MgTile<3>;
    constexpr int NT = kSw3Threads;
    constexpr int TG = T::K / 4;                            // groups per tile row
    constexpr int GK = TG + 2;                              // groups per LDS row
    constexpr int SI = T::I + 4, SJ = T::J + 4, SK = 4 * GK, SP = SJ * SK;
    constexpr int NG = SI * SJ * GK, ZR = (NG + NT - 1) / NT;               // z tile groups, rounds
    constexpr int P0G = (T::I + 2) * (T::J + 2) * GK, P0R = (P0G + NT - 1) / NT;  // first pass
    constexpr int P1G = T::I * T::J * TG, P1R = P1G / NT;  // second pass: the tile
    static_assert(P1G % NT == 0 && P1R <= P0R, "the tile fills whole rounds of the first pass");
    // 4 floats of slack each side (the LDS tile's outer groups are never first-pass items, so
    // every LDS read stays inside the tile; the slack keeps that true for any tile change)
    __shared__ __align__(16) float zsm[SI * SP + 8];
    float* const zs = zsm + 4;
    int t[3];
    tile_of_block(tm, t);
    const int o0 = t[0] * T::I, o1 = t[1] * T::J, o2 = t[2] * T::K;
    const int b0 = o0 - 2, b1 = o1 - 2, bk = o2 - 4;  // global cell of LDS (0, 0, 0)
    const int d0 = L.d[0], d1 = L.d[1], d2 = L.d[2];
    const int tid = threadIdx.x;
    const int c1 = c0 ^ 1;
    const bool solid = L.dg != nullptr;
    const __amdgpu_buffer_rsrc_t rr = lvl_rsrc(r, L);
    const __amdgpu_buffer_rsrc_t rd = solid ? dg_rsrc(L) : rr;
    auto in3 = [&](int g0, int g1, int gk) {
        return (unsigned)g0 < (unsigned)d0 && (unsigned)g1 < (unsigned)d1 && (unsigned)gk < (unsigned)d2;
    };
    auto cell = [&](int g0, int g1, int gk) { return (uint32_t)((g0 * d1 + g1) * d2 + gk); };
    // ---- z tile
    float4 zv[ZERO_IN ? 1 : ZR];
    if (!ZERO_IN) {
        const __amdgpu_buffer_rsrc_t rz = lvl_rsrc(zin, L);
        const __amdgpu_buffer_rsrc_t rc = lvl_rsrc(zc, Lc);
#pragma unroll
        for (int m = 0; m < ZR; ++m) {
            const int e = tid + m * NT;
            const int lg = e % GK, q = e / GK, lj = q % SJ, li = q / SJ;
            const int g0 = b0 + li, g1 = b1 + lj, gk = bk + 4 * lg;
            const bool in = e < NG && in3(g0, g1, gk);
            const uint32_t off = cell(g0, g1, gk);
            zv[m] = ld4<ZT>(rz, in ? off * (uint32_t)sizeof(ZT) : kOOB);
            if (PROLONG) {
                // prolongate (3D/mgpcg.py:160-163): cells 4q..4q+3 have the
                // coarse parents 2q, 2q, 2q+1, 2q+1 along axis 2
                const int p0 = min(g0 >> 1, Lc.d[0] - 1), p1 = min(g1 >> 1, Lc.d[1] - 1);
                const uint32_t co = (uint32_t)((p0 * Lc.d[1] + p1) * Lc.d[2] + (gk >> 1));
                const float2 zp = ld2<ZT>(rc, in ? co * (uint32_t)sizeof(ZT) : kOOB);
                zv[m].x = __fadd_rn(zv[m].x, zp.x);
                zv[m].y = __fadd_rn(zv[m].y, zp.x);
                zv[m].z = __fadd_rn(zv[m].z, zp.y);
                zv[m].w = __fadd_rn(zv[m].w, zp.y);
            }
        }
        if (PROLONG && solid) {
            // solid cells take no prolongation (they carry no pressure)
#pragma unroll
            for (int m = 0; m < ZR; ++m) {
                const int e = tid + m * NT;
                const int lg = e % GK, q = e / GK, lj = q % SJ, li = q / SJ;
                const int g0 = b0 + li, g1 = b1 + lj, gk = bk + 4 * lg;
                const bool in = e < NG && in3(g0, g1, gk);
                const uint32_t dw = __builtin_amdgcn_raw_buffer_load_b32(rd, in ? cell(g0, g1, gk) : kOOB, 0, 0);
                zv[m].x = (dw & 0xffu) == 0 ? 0.f : zv[m].x;
                zv[m].y = (dw & 0xff00u) == 0 ? 0.f : zv[m].y;
                zv[m].z = (dw & 0xff0000u) == 0 ? 0.f : zv[m].z;
                zv[m].w = (dw & 0xff000000u) == 0 ? 0.f : zv[m].w;
            }
        }
    }
    // ---- r (and the obstacle's diagonals) of the first pass; items < P1G are the tile.
    // The r of the ring items is requested only once the z tile is in LDS, so the
    // z and r registers in flight stay within four waves per SIMD.
    float4 rv[P0R];
    uint32_t dgv[P0R];
    auto load_r = [&](int m) {
        const int e = tid + m * NT;
        int ri, rj, lg;
        sweep3_item<T::I, T::J, TG>(e, ri, rj, lg);
        const int g0 = o0 - 1 + ri, g1 = o1 - 1 + rj, gk = bk + 4 * lg;
        const bool in = e < P0G && in3(g0, g1, gk);
        rv[m] = ld4<RT>(rr, in ? cell(g0, g1, gk) * (uint32_t)sizeof(RT) : kOOB);
        dgv[m] = 0u;
    };
#pragma unroll
    for (int m = 0; m < P1R; ++m) load_r(m);
    if (ZERO_IN) {
#pragma unroll
        for (int m = P1R; m < P0R; ++m) load_r(m);
    }
    if (!ZERO_IN) {
#pragma unroll
        for (int m = 0; m < ZR; ++m) {
            const int e = tid + m * NT;
            if (e < NG) {
                const int lg = e % GK, q = e / GK, lj = q % SJ, li = q / SJ;
                *reinterpret_cast<float4*>(zs + li * SP + lj * SK + 4 * lg) = zv[m];
            }
        }
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int m = P1R; m < P0R; ++m) load_r(m);
    }
    if (solid) {
#pragma unroll
        for (int m = 0; m < P0R; ++m) {
            const int e = tid + m * NT;
            int ri, rj, lg;
            sweep3_item<T::I, T::J, TG>(e, ri, rj, lg);
            const int g0 = o0 - 1 + ri, g1 = o1 - 1 + rj, gk = bk + 4 * lg;
            const bool in = e < P0G && in3(g0, g1, gk);
            dgv[m] = __builtin_amdgcn_raw_buffer_load_b32(rd, in ? cell(g0, g1, gk) : kOOB, 0, 0);
        }
    }
    const float mean = SUB_MEAN ? sc->mean : 0.f;
    // 16-bit z: level 0 works on vs * r (coarser r arrive scaled)
    constexpr bool kScale = sizeof(ZT) == 2 && sizeof(RT) == 4;
    const float vs = kScale ? sc->vs : 1.0f;
    const bool interior0 = !solid && box_interior(L, o0 - 1, o1 - 1, o2 - 1, o0 + T::I + 1, o1 + T::J + 1,
                                                  o2 + T::K + 1, true);
    const bool interior1 = !solid && box_interior(L, o0, o1, o2, o0 + T::I, o1 + T::J, o2 + T::K, true);
    const float n_int = (float)L.nb2, y_int = pfm_rcp_small(L.nb2);
    // `zero`: every z is +0 (the first pass of a zero input), nothing is read from LDS.
    // S: the offset of the group's first cell of colour c, either fixed at
    // compile time (an std::integral_constant, for waves whose groups share it)
    // or an int taken from the group.
    auto pass = [&](auto S, int base, int g0, int g1, int gk, int c, bool in, bool intr, float4 rg, uint32_t dw,
                    bool zero) -> GroupPass {
        bool s1;
        if constexpr (std::is_same<decltype(S), int>::value) s1 = ((c + g0 + g1) & 1) != 0;
        else s1 = decltype(S)::value != 0;
        const float4 z4 = make_float4(0.f, 0.f, 0.f, 0.f);
        const float4 own = zero ? z4 : *reinterpret_cast<const float4*>(zs + base);
        const float4 ip = zero ? z4 : *reinterpret_cast<const float4*>(zs + base + SP);
        const float4 im = zero ? z4 : *reinterpret_cast<const float4*>(zs + base - SP);
        const float4 jp = zero ? z4 : *reinterpret_cast<const float4*>(zs + base + SK);
        const float4 jm = zero ? z4 : *reinterpret_cast<const float4*>(zs + base - SK);
        const float lf = zero ? 0.f : zs[base - 1], rt = zero ? 0.f : zs[base + 4];
        const float kmid = s1 ? own.z : own.y;  // axis-2 +1 of cell a and -1 of cell b
        // neighbor_sum order (3D/mgpcg.py:92-104): +e then -e per axis
        float na = 0.0f, nb = 0.0f;
        na = __fadd_rn(na, s1 ? ip.y : ip.x);
        nb = __fadd_rn(nb, s1 ? ip.w : ip.z);
        na = __fadd_rn(na, s1 ? im.y : im.x);
        nb = __fadd_rn(nb, s1 ? im.w : im.z);
        na = __fadd_rn(na, s1 ? jp.y : jp.x);
        nb = __fadd_rn(nb, s1 ? jp.w : jp.z);
        na = __fadd_rn(na, s1 ? jm.y : jm.x);
        nb = __fadd_rn(nb, s1 ? jm.w : jm.z);
        na = __fadd_rn(na, kmid);
        nb = __fadd_rn(nb, s1 ? rt : own.w);
        na = __fadd_rn(na, s1 ? own.x : lf);
        nb = __fadd_rn(nb, kmid);
        GroupPass o;
        o.ra = s1 ? rg.y : rg.x;
        o.rb = s1 ? rg.w : rg.z;
        if (SUB_MEAN) {
            o.ra = __fsub_rn(o.ra, mean);
            o.rb = __fsub_rn(o.rb, mean);
        }
        const float va = kScale ? o.ra * vs : o.ra, vb = kScale ? o.rb * vs : o.rb;
        const int ka = gk + (s1 ? 1 : 0), kb = ka + 2;
        float nfa = n_int, ya = y_int, nfb = n_int, yb = y_int;
        bool fa = true, fb = true;
        if (solid) {
            // a group of four fluid cells with the interior diagonal keeps n_int / y_int
            if (dw != (uint32_t)L.nb2 * 0x01010101u) {
                const int sa = s1 ? 8 : 0;
                const int na_ = (int)((dw >> sa) & 0xffu), nb_ = (int)((dw >> (sa + 16)) & 0xffu);
                nfa = (float)na_;
                ya = pfm_rcp_small(na_);
                fa = na_ != 0;
                nfb = (float)nb_;
                yb = pfm_rcp_small(nb_);
                fb = nb_ != 0;
            }
        } else if (!intr) {
            // walls (num_fluid_neighbors): axes 0 / 1 once per group, axis 2 per cell;
            // only the groups on a wall layer leave the interior diagonal
            const bool edge = g0 <= 0 || g0 >= d0 - 1 || g1 <= 0 || g1 >= d1 - 1 || gk <= 0 || gk + 4 >= d2;
            if (edge) {
                const int wij = (g0 <= 0 ? L.bc_lo[0] : 0) + (g0 >= d0 - 1 ? L.bc_hi[0] : 0) +
                                (g1 <= 0 ? L.bc_lo[1] : 0) + (g1 >= d1 - 1 ? L.bc_hi[1] : 0);
                const int wa = L.nb2 - wij - (ka <= 0 ? L.bc_lo[2] : 0) - (ka >= d2 - 1 ? L.bc_hi[2] : 0);
                const int wb = L.nb2 - wij - (kb <= 0 ? L.bc_lo[2] : 0) - (kb >= d2 - 1 ? L.bc_hi[2] : 0);
                nfa = (float)wa;
                ya = pfm_rcp_small(wa);
                nfb = (float)wb;
                yb = pfm_rcp_small(wb);
            }
        }
        o.za = pfm_div_small(__fadd_rn(va, na), nfa, ya);
        o.zb = pfm_div_small(__fadd_rn(vb, nb), nfb, yb);
        o.ua = in && fa;
        o.ub = in && fb;
        const float xa = o.ua ? o.za : (s1 ? own.y : own.x);
        const float xb = o.ub ? o.zb : (s1 ? own.w : own.z);
        o.v.x = s1 ? own.x : xa;
        o.v.y = s1 ? xa : own.y;
        o.v.z = s1 ? own.z : xb;
        o.v.w = s1 ? xb : own.w;
        return o;
    };
    double s = 0.0;
    auto dot = [&](const GroupPass& o) {
        if (o.ua) s += (double)to_f(from_f<ZT>(o.za)) * (double)o.ra;
        if (o.ub) s += (double)to_f(from_f<ZT>(o.zb)) * (double)o.rb;
    };
    if (!ZERO_IN) __syncthreads();
    // ---- pass 0 (colour c0): the tile, its ring along axes 0 / 1 and the edge groups
#pragma unroll
    for (int m = 0; m < P0R; ++m) {
        const int e = tid + m * NT;
        if (e < P0G) {
            int ri, rj, lg;
            sweep3_item<T::I, T::J, TG>(e, ri, rj, lg);
            const int g0 = o0 - 1 + ri, g1 = o1 - 1 + rj, gk = bk + 4 * lg;
            const int base = (ri + 1) * SP + (rj + 1) * SK + 4 * lg;
            const bool in = in3(g0, g1, gk);
            GroupPass o;
            if (m < P1R) {
                // tile items: the wave's groups share the colour offset
                if (__builtin_amdgcn_readfirstlane((c0 + g0 + g1) & 1))
                    o = pass(std::integral_constant<int, 1>{}, base, g0, g1, gk, c0, in, interior0, rv[m], dgv[m],
                             ZERO_IN);
                else
                    o = pass(std::integral_constant<int, 0>{}, base, g0, g1, gk, c0, in, interior0, rv[m], dgv[m],
                             ZERO_IN);
            } else {
                o = pass(0, base, g0, g1, gk, c0, in, interior0, rv[m], dgv[m], ZERO_IN);
            }
            *reinterpret_cast<float4*>(zs + base) = o.v;
            // cells of colour c0 inside the tile are final after this pass
            if (DOT && m < P1R) dot(o);
        }
    }
    __syncthreads();
    // ---- pass 1 (colour c1) on the tile, stored from registers
#pragma unroll
    for (int m = 0; m < P1R; ++m) {
        const int e = tid + m * NT;
        int ri, rj, lg;
        sweep3_item<T::I, T::J, TG>(e, ri, rj, lg);
        const int g0 = o0 - 1 + ri, g1 = o1 - 1 + rj, gk = bk + 4 * lg;
        const int base = (ri + 1) * SP + (rj + 1) * SK + 4 * lg;
        const bool in = in3(g0, g1, gk);
        GroupPass o;
        if (__builtin_amdgcn_readfirstlane((c1 + g0 + g1) & 1))
            o = pass(std::integral_constant<int, 1>{}, base, g0, g1, gk, c1, in, interior1, rv[m], dgv[m], false);
        else
            o = pass(std::integral_constant<int, 0>{}, base, g0, g1, gk, c1, in, interior1, rv[m], dgv[m], false);
        if (DOT) dot(o);
        if (in) st4<ZT>(zout, cell(g0, g1, gk), o.v);
    }
    if (DOT) {
        __shared__ double sh[NT / 64];
        s = block_sum(s, sh);
        if (tid == 0) partial[blockIdx.x] = kScale ? s * (double)sc->vinv : s;  // z was stored scaled
    }
}

// Residual r - A z on an LDS tile and its restriction (3D/mgpcg.py:152-158):
// each coarse cell sums its 2^d children in fixed order.  Requires every fine
// extent to be twice the coarse one (axis 0 of 2D levels: both 1).  With
// SUB_MEAN the recentred r (3D/mgpcg.py:174-180) is used and written back;
// each fine r is touched by exactly one block, so this is race free.
// `inv_scale` is 1 / (dim - 1) (0.5 or 1), so the product equals the
// reference's division exactly.
// ZT: storage of z and of the coarse r; RT: storage of the fine r.
template <int DIM, typename ZT, typename RT, bool SUB_MEAN>
__global__ void __launch_bounds__(kBlock) k_restrict_tiled(LvlDev Lf, LvlDev Lc, TileMap tm, float inv_scale,
                                                           const ZT* __restrict__ z, RT* __restrict__ r,
                                                           ZT* __restrict__ rc,
                                                           const MGScalars* __restrict__ sc) {
    using C = RsTile<DIM>;
    constexpr int FI = DIM == 3 ? 2 * C::I : 1, FJ = 2 * C::J, FK = 2 * C::K, NF = FI * FJ * FK;
    constexpr int HI = DIM == 3 ? 1 : 0;
    constexpr int SI = FI + 2 * HI, SJ = FJ + 2, SK = FK + 2, SP = SJ * SK, NS = SI * SP;
    constexpr int NL = (NS + kBlock - 1) / kBlock, NR = (NF + kBlock - 1) / kBlock;
    static_assert(NF % kBlock == 0, "whole fine tile per thread sweep");
    __shared__ float zs[NS];
    __shared__ float ts[NF];
    int t[3];
    tile_of_block(tm, t);
    const int c0 = t[0] * C::I, c1 = t[1] * C::J, c2 = t[2] * C::K;
    const int o0 = DIM == 3 ? 2 * c0 : 0, o1 = 2 * c1, o2 = 2 * c2;
    const int b0 = o0 - HI, b1 = o1 - 1, b2 = o2 - 1;
    const int tid = threadIdx.x;
    const __amdgpu_buffer_rsrc_t rz = lvl_rsrc(z, Lf), rr = lvl_rsrc(r, Lf);
    float zv[NL], rv[NR];
#pragma unroll
    for (int k = 0; k < NL; ++k) {
        const int e = tid + k * kBlock;
        const int lk = e % SK, lj = (e / SK) % SJ, li = e / SP;
        zv[k] = ldv<ZT>(rz, e < NS ? cell_off<ZT>(Lf, b0 + li, b1 + lj, b2 + lk) : kOOB);
    }
#pragma unroll
    for (int k = 0; k < NR; ++k) {
        const int e = tid + k * kBlock;
        const int fk = e % FK, fj = (e / FK) % FJ, fi = e / (FK * FJ);
        rv[k] = ldv<RT>(rr, cell_off<RT>(Lf, o0 + fi, o1 + fj, o2 + fk));
    }
#pragma unroll
    for (int k = 0; k < NL; ++k) {
        const int e = tid + k * kBlock;
        if (e < NS) zs[e] = zv[k];
    }
    const float mean = SUB_MEAN ? sc->mean : 0.f;
    const bool interior = box_interior(Lf, o0, o1, o2, o0 + FI, o1 + FJ, o2 + FK, DIM == 3);
    const float n_int = (float)Lf.nb2;
    __syncthreads();
#pragma unroll
    for (int k = 0; k < NR; ++k) {
        const int e = tid + k * kBlock;
        const int fk = e % FK, fj = (e / FK) % FJ, fi = e / (FK * FJ);
        const int g0 = o0 + fi, g1 = o1 + fj, g2 = o2 + fk;
        float nf = interior ? n_int : (float)(Lf.nb2 - walls(Lf, g0, g1, g2));
        bool fluid = true;
        if (Lf.dg) {
            const int n = ld_dg(dg_rsrc(Lf), Lf, g0, g1, g2);
            nf = (float)n;
            fluid = n != 0;
        }
        const bool in = inside(Lf, g0, g1, g2) && fluid;
        float rq = rv[k];
        if (SUB_MEAN) {
            rq = __fsub_rn(rq, mean);
            if (in) r[(g0 * Lf.d[1] + g1) * Lf.d[2] + g2] = from_f<RT>(rq);
        }
        const int le = ((fi + HI) * SJ + fj + 1) * SK + fk + 1;
        // 16-bit z: level 0 works on vs * r (coarser r arrive scaled)
        if constexpr (sizeof(ZT) == 2 && sizeof(RT) == 4) rq *= sc->vs;
        const float res = __fsub_rn(rq, __fsub_rn(__fmul_rn(nf, zs[le]), nsum_lds<DIM>(zs + le, SP, SK)));
        // solid children add nothing (the oracle skips them; acc starts at +0)
        ts[e] = in ? __fmul_rn(__fmul_rn(res, 1.0f), inv_scale) : 0.f;
    }
    __syncthreads();
    constexpr int NA = DIM == 3 ? 2 : 1;
    for (int e = tid; e < C::I * C::J * C::K; e += kBlock) {
        const int ck = e % C::K, cj = (e / C::K) % C::J, ci = e / (C::K * C::J);
        const int g0 = c0 + ci, g1 = c1 + cj, g2 = c2 + ck;
        if (!inside(Lc, g0, g1, g2)) continue;
        float acc = 0.0f;
#pragma unroll
        for (int a = 0; a < NA; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b)
#pragma unroll
                for (int c = 0; c < 2; ++c)
                    acc = __fadd_rn(acc, ts[((NA * ci + a) * FJ + 2 * cj + b) * FK + 2 * ck + c]);
        rc[(g0 * Lc.d[1] + g1) * Lc.d[2] + g2] = from_f<ZT>(acc);
    }
}

// ---- fused visits of the coarser tiled levels --------------------------------
// On levels whose launches are latency bound, each half of a V-cycle visit
// runs as one launch: down = the zero-input sweep, the plain sweep and the
// restriction (3D/mgpcg.py:184-191); up = the prolongation sweep and the plain
// sweep (:196-200).  A block loads its tile with a halo of H cells; each
// red/black pass updates the region one cell narrower than the last, so after
// p passes every cell within H - p of the tile holds exactly the value of p
// sequential passes over the whole level (an update reads only its face
// neighbours, which were exact after the previous pass).  z is rounded to its
// storage type after the second and the fourth pass, where the separate
// launches store it, so the results are those of the separate launches bit
// for bit.  A cell's diagonal byte is kNoUpd outside the level and on solid
// cells, which are never updated and hold 0.
template <int D, int TI, int TJ, int TK>
struct FsTile {
    static constexpr int DIM = D, I = TI, J = TJ, K = TK;
};
// 3D: small tiles, so each block's chain (box load, four passes, stores) is
// short; the launches of these levels are latency bound (measured: 32^3 down
// and up 10.2 / 8.3 us with 4x8x8 against 15.4 / 13.4 us with 8x8x16; 64^3
// 15.5 / 12.8 us with 4x8x16)
using FsTile3 = FsTile<3, 4, 8, 16>;
using FsTile3S = FsTile<3, 4, 8, 8>;  // levels up to kFuseSmallCells
constexpr int64_t kFuseSmallCells = (int64_t)1 << 15;
// 2D likewise (2D Poisson 512^2: 0.707 ms with 8x32 up to 2^18 cells, 0.778
// with 32x32; at 1024^2 16x32 tiles on level 0 beat 8x32, 0.873 / 0.916 ms)
using FsTile2 = FsTile<2, 1, 16, 32>;
using FsTile2S = FsTile<2, 1, 8, 32>;  // levels up to kFuseSmallCells2D
constexpr int64_t kFuseSmallCells2D = (int64_t)1 << 18;
constexpr int kFuseThreads = 256;
constexpr uint8_t kNoUpd = 0xFF;

template <typename T, int H>
struct FsGeom {
    static constexpr int DIM = T::DIM;
    static constexpr int HI = DIM == 3 ? H : 0;
    static constexpr int SI = T::I + 2 * HI, SJ = T::J + 2 * H, SK = T::K + 2 * H;
    static constexpr int SP = SJ * SK, NS = SI * SP;
};

// Round every LDS value to ZT (a no-op for fp32 storage).
template <typename ZT, int NS>
__device__ __forceinline__ void fs_round(float* zs) {
    if constexpr (sizeof(ZT) == 2) {
        for (int e = threadIdx.x; e < NS; e += kFuseThreads) zs[e] = to_f(from_f<ZT>(zs[e]));
    }
}

// One colour pass over the tile grown by RING cells.  Tile origin (o0, o1,
// o2) sits at LDS (HI, H, H); cells come in pairs along axis 2, of which the
// member of colour c (global (g0 + g1 + g2) & 1 == c) is updated.  A thread
// takes SEG consecutive pairs of a row (the member's parity is the same along
// a row): it gathers their neighbour sums first, so the LDS reads overlap, and
// stores after (a pass reads only the other colour).  ny[n] = (n, RN(1/n)).
template <typename T, int H, int RING>
__device__ __forceinline__ void fs_pass(float* zs, const float* rs, const uint8_t* nds, const float2* ny, int o0,
                                        int o1, int o2, int c) {
    using G = FsGeom<T, H>;
    constexpr int DIM = T::DIM;
    constexpr int RI = DIM == 3 ? RING : 0;
    constexpr int NI = T::I + 2 * RI, NJ = T::J + 2 * RING, NK = T::K + 2 * RING, HP = NK / 2;
    static_assert(NK % 2 == 0, "whole pairs along axis 2");
    constexpr int SEG = 4, NSEG = (HP + SEG - 1) / SEG, NU = NI * NJ * NSEG;
    for (int u = threadIdx.x; u < NU; u += kFuseThreads) {
        const int row = u / NSEG, sg = u - row * NSEG;
        const int ri = row / NJ, rj = row - ri * NJ;
        const int li = G::HI - RI + ri, lj = H - RING + rj, lk0 = H - RING + 2 * SEG * sg;
        const int g0 = o0 - G::HI + li, g1 = o1 - H + lj, gk0 = o2 - H + lk0;
        const int base = (li * G::SJ + lj) * G::SK + lk0 + ((g0 + g1 + gk0 + c) & 1);
        int nn[SEG];
        float ns[SEG], rv[SEG];
#pragma unroll
        for (int m = 0; m < SEG; ++m) {
            const bool ok = HP % SEG == 0 || SEG * sg + m < HP;
            const int le = ok ? base + 2 * m : base;
            nn[m] = ok ? (int)nds[le] : (int)kNoUpd;
            ns[m] = nsum_lds<DIM>(zs + le, G::SP, G::SK);
            rv[m] = rs[le];
        }
#pragma unroll
        for (int m = 0; m < SEG; ++m) {
            if (nn[m] == kNoUpd) continue;
            const float2 t = ny[nn[m]];
            zs[base + 2 * m] = pfm_div_small(__fadd_rn(rv[m], ns[m]), t.x, t.y);
        }
    }
    __syncthreads();
}

// Diagonal byte of global cell g (dv: the obstacle's byte, when there is
// one): kNoUpd outside the level or on a solid cell.
__device__ __forceinline__ uint8_t fs_diag(const LvlDev& L, int dv, int g0, int g1, int g2) {
    if (!inside(L, g0, g1, g2)) return kNoUpd;
    if (L.dg) return dv == 0 ? kNoUpd : (uint8_t)dv;
    return (uint8_t)(L.nb2 - walls(L, g0, g1, g2));
}

// The LDS box is loaded by columns: a thread owns positions (lj, lk) of the
// box face and walks down axis 0 with incremental offsets, all loads issued
// before any is used (kOOB reads 0 for cells outside the level).
template <typename T, int H>
struct FsCols {
    using G = FsGeom<T, H>;
    static constexpr int ZC = G::SJ * G::SK, ZR = (ZC + kFuseThreads - 1) / kFuseThreads;
};

// The obstacle bytes of a thread's columns (0 without an obstacle).
template <typename T, int H>
__device__ __forceinline__ void fs_load_dg(const LvlDev& L, int o0, int o1, int o2,
                                           int (&dv)[FsCols<T, H>::ZR][FsGeom<T, H>::SI]) {
    using G = FsGeom<T, H>;
    using C = FsCols<T, H>;
#pragma unroll
    for (int m = 0; m < C::ZR; ++m)
#pragma unroll
        for (int li = 0; li < G::SI; ++li) dv[m][li] = 0;
    if (!L.dg) return;
    const __amdgpu_buffer_rsrc_t rd = dg_rsrc(L);
    const int d1 = L.d[1], d2 = L.d[2], plane = d1 * d2;
#pragma unroll
    for (int m = 0; m < C::ZR; ++m) {
        const int p = threadIdx.x + m * kFuseThreads;
        const int lj = p / G::SK, lk = p - lj * G::SK;
        const int g1 = o1 - H + lj, g2 = o2 - H + lk;
        const bool col = p < C::ZC && (unsigned)g1 < (unsigned)d1 && (unsigned)g2 < (unsigned)d2;
        const int base = g1 * d2 + g2;
#pragma unroll
        for (int li = 0; li < G::SI; ++li) {
            const int g0 = o0 - G::HI + li;
            const bool in = col && (unsigned)g0 < (unsigned)L.d[0];
            dv[m][li] = (int)__builtin_amdgcn_raw_buffer_load_b8(rd, in ? (uint32_t)(g0 * plane + base) : kOOB, 0, 0);
        }
    }
}

// Down half of a level visit: z = 0, colours 0, 1, 0, 1; writes the tile's z
// to zout and r[l + 1] of the coarse cells whose first child lies in the tile.
// z is stored as ZT, r as RT (ZT below level 0; fp32 on level 0, where it
// enters as vs * r with 16-bit z, and has been recentred by its own launch).
// The halo of 5 keeps z exact one cell around the tile after the four passes:
// the trailing child of an odd extent lies one cell past the tile, and its
// outer neighbour is outside the level.
template <typename T, typename ZT, typename RT>
__global__ void __launch_bounds__(kFuseThreads) k_fuse_down(LvlDev L, LvlDev Lc, TileMap tm, float inv_scale,
                                                            const RT* __restrict__ r, ZT* __restrict__ zout,
                                                            ZT* __restrict__ rc, const MGScalars* __restrict__ sc) {
    constexpr int H = 5;
    using G = FsGeom<T, H>;
    constexpr int DIM = T::DIM;
    __shared__ float zs[G::NS], rs[G::NS];
    __shared__ uint8_t nds[G::NS];
    __shared__ float2 ny[8];
    if (threadIdx.x < 8) ny[threadIdx.x] = make_float2((float)threadIdx.x, pfm_rcp_small(threadIdx.x));
    int t[3];
    tile_of_block(tm, t);
    const int o0 = t[0] * T::I, o1 = t[1] * T::J, o2 = t[2] * T::K;
    const __amdgpu_buffer_rsrc_t rr = lvl_rsrc(r, L);
    using C = FsCols<T, H>;
    const int d1 = L.d[1], d2 = L.d[2], plane = d1 * d2;
    float rv[C::ZR][G::SI];
    int dv[C::ZR][G::SI];
#pragma unroll
    for (int m = 0; m < C::ZR; ++m) {
        const int p = threadIdx.x + m * kFuseThreads;
        const int lj = p / G::SK, lk = p - lj * G::SK;
        const int g1 = o1 - H + lj, g2 = o2 - H + lk;
        const bool col = p < C::ZC && (unsigned)g1 < (unsigned)d1 && (unsigned)g2 < (unsigned)d2;
        const int base = g1 * d2 + g2;
#pragma unroll
        for (int li = 0; li < G::SI; ++li) {
            const int g0 = o0 - G::HI + li;
            const bool in = col && (unsigned)g0 < (unsigned)L.d[0];
            rv[m][li] = ldv<RT>(rr, in ? (uint32_t)(g0 * plane + base) * (uint32_t)sizeof(RT) : kOOB);
        }
    }
    fs_load_dg<T, H>(L, o0, o1, o2, dv);
    // 16-bit z with an fp32 r (level 0): work on vs * r, as the sweeps do
    constexpr bool kScale = sizeof(ZT) == 2 && sizeof(RT) == 4;
    const float vs = kScale ? sc->vs : 1.0f;
#pragma unroll
    for (int m = 0; m < C::ZR; ++m) {
        const int p = threadIdx.x + m * kFuseThreads;
        if (p >= C::ZC) continue;
        const int lj = p / G::SK, lk = p - lj * G::SK;
#pragma unroll
        for (int li = 0; li < G::SI; ++li) {
            const int e = li * G::SP + p;
            rs[e] = kScale ? rv[m][li] * vs : rv[m][li];
            zs[e] = 0.f;
            nds[e] = fs_diag(L, dv[m][li], o0 - G::HI + li, o1 - H + lj, o2 - H + lk);
        }
    }
    __syncthreads();
    fs_pass<T, H, H - 1>(zs, rs, nds, ny, o0, o1, o2, 0);
    fs_pass<T, H, H - 2>(zs, rs, nds, ny, o0, o1, o2, 1);
    fs_round<ZT, G::NS>(zs);
    __syncthreads();
    fs_pass<T, H, H - 3>(zs, rs, nds, ny, o0, o1, o2, 0);
    fs_pass<T, H, H - 4>(zs, rs, nds, ny, o0, o1, o2, 1);
    fs_round<ZT, G::NS>(zs);
    __syncthreads();
    for (int e = threadIdx.x; e < T::I * T::J * T::K; e += kFuseThreads) {
        const int ti = e / (T::J * T::K), rem = e - ti * (T::J * T::K), tj = rem / T::K, tk = rem - tj * T::K;
        const int g0 = o0 + ti, g1 = o1 + tj, g2 = o2 + tk;
        if (!inside(L, g0, g1, g2)) continue;
        zout[((int64_t)g0 * L.d[1] + g1) * L.d[2] + g2] = from_f<ZT>(zs[((ti + G::HI) * G::SJ + tj + H) * G::SK + tk + H]);
    }
    // restriction (3D/mgpcg.py:152-158) of the coarse cells owned by the tile
    constexpr int CI = DIM == 3 ? T::I / 2 : 1, CJ = T::J / 2, CK = T::K / 2;
    const int c0 = DIM == 3 ? o0 / 2 : 0, c1 = o1 / 2, c2 = o2 / 2;
    for (int e = threadIdx.x; e < CI * CJ * CK; e += kFuseThreads) {
        const int ci = e / (CJ * CK), rem = e - ci * (CJ * CK), cj = rem / CK, ck = rem - cj * CK;
        const int C0 = c0 + ci, C1 = c1 + cj, C2 = c2 + ck;
        if (!inside(Lc, C0, C1, C2)) continue;
        const int h0 = C0 == Lc.d[0] - 1 ? L.d[0] - 1 - 2 * C0 : 1;
        const int h1 = C1 == Lc.d[1] - 1 ? L.d[1] - 1 - 2 * C1 : 1;
        const int h2 = C2 == Lc.d[2] - 1 ? L.d[2] - 1 - 2 * C2 : 1;
        float acc = 0.0f;
        for (int a = 0; a <= h0; ++a)
            for (int b = 0; b <= h1; ++b)
                for (int c = 0; c <= h2; ++c) {
                    const int li = 2 * C0 + a - (o0 - G::HI), lj = 2 * C1 + b - (o1 - H), lk = 2 * C2 + c - (o2 - H);
                    const int le = (li * G::SJ + lj) * G::SK + lk;
                    const int n = nds[le];
                    if (n == kNoUpd) continue;  // solid children add nothing
                    const float res = __fsub_rn(
                        rs[le], __fsub_rn(__fmul_rn((float)n, zs[le]), nsum_lds<DIM>(zs + le, G::SP, G::SK)));
                    acc = __fadd_rn(acc, __fmul_rn(__fmul_rn(res, 1.0f), inv_scale));
                }
        rc[((int64_t)C0 * Lc.d[1] + C1) * Lc.d[2] + C2] = from_f<ZT>(acc);
    }
}

// Up half of a level visit: z = zin + zc[parent] (solid cells keep 0),
// colours 1, 0, 1, 0; writes the tile's z to zout.  Tiles read their halo from
// zin, which no block of the launch writes.  Storage types as k_fuse_down.
// DOT (level 0): also the z.r partial of the tile (z as stored, r unscaled),
// the PCG's z.r after the preconditioner (3D/mgpcg.py:231-235, 264-268).
template <typename T, typename ZT, typename RT, bool DOT>
__global__ void __launch_bounds__(kFuseThreads) k_fuse_up(LvlDev L, LvlDev Lc, TileMap tm,
                                                          const ZT* __restrict__ zin, const ZT* __restrict__ zc,
                                                          const RT* __restrict__ r, ZT* __restrict__ zout,
                                                          const MGScalars* __restrict__ sc,
                                                          double* __restrict__ partial) {
    constexpr int H = 4;
    using G = FsGeom<T, H>;
    constexpr int DIM = T::DIM;
    __shared__ float zs[G::NS], rs[G::NS];
    __shared__ uint8_t nds[G::NS];
    __shared__ float2 ny[8];
    if (threadIdx.x < 8) ny[threadIdx.x] = make_float2((float)threadIdx.x, pfm_rcp_small(threadIdx.x));
    int t[3];
    tile_of_block(tm, t);
    const int o0 = t[0] * T::I, o1 = t[1] * T::J, o2 = t[2] * T::K;
    const __amdgpu_buffer_rsrc_t rr = lvl_rsrc(r, L), rz = lvl_rsrc(zin, L), rc = lvl_rsrc(zc, Lc);
    using C = FsCols<T, H>;
    static_assert((G::HI % 2 == 0 && T::I % 2 == 0) || DIM == 2, "even axis-0 box origin");
    constexpr int SH = (G::SI + 1) / 2;  // axis-0 pairs sharing a coarse parent
    const int d1 = L.d[1], d2 = L.d[2], plane = d1 * d2;
    const int e1 = Lc.d[1], e2 = Lc.d[2], pplane = e1 * e2;
    float rv[C::ZR][G::SI], zv[C::ZR][G::SI], zp[C::ZR][SH];
    int dv[C::ZR][G::SI];
#pragma unroll
    for (int m = 0; m < C::ZR; ++m) {
        const int p = threadIdx.x + m * kFuseThreads;
        const int lj = p / G::SK, lk = p - lj * G::SK;
        const int g1 = o1 - H + lj, g2 = o2 - H + lk;
        const bool col = p < C::ZC && (unsigned)g1 < (unsigned)d1 && (unsigned)g2 < (unsigned)d2;
        const int base = g1 * d2 + g2;
        const int pcol = min(g1 >> 1, e1 - 1) * e2 + min(g2 >> 1, e2 - 1);
#pragma unroll
        for (int li = 0; li < G::SI; ++li) {
            const int g0 = o0 - G::HI + li;
            const bool in = col && (unsigned)g0 < (unsigned)L.d[0];
            const uint32_t off = in ? (uint32_t)(g0 * plane + base) * (uint32_t)sizeof(ZT) : kOOB;
            rv[m][li] = ldv<RT>(rr, in ? (uint32_t)(g0 * plane + base) * (uint32_t)sizeof(RT) : kOOB);
            zv[m][li] = ldv<ZT>(rz, off);
        }
#pragma unroll
        for (int h = 0; h < SH; ++h) {
            const int g0 = o0 - G::HI + 2 * h;
            const bool in = col && (unsigned)g0 < (unsigned)L.d[0];
            const int pc = min(g0 >> 1, Lc.d[0] - 1) * pplane + pcol;
            zp[m][h] = ldv<ZT>(rc, in ? (uint32_t)pc * (uint32_t)sizeof(ZT) : kOOB);
        }
    }
    fs_load_dg<T, H>(L, o0, o1, o2, dv);
    constexpr bool kScale = sizeof(ZT) == 2 && sizeof(RT) == 4;
    const float vs = kScale ? sc->vs : 1.0f;
#pragma unroll
    for (int m = 0; m < C::ZR; ++m) {
        const int p = threadIdx.x + m * kFuseThreads;
        if (p >= C::ZC) continue;
        const int lj = p / G::SK, lk = p - lj * G::SK;
        const int g1 = o1 - H + lj, g2 = o2 - H + lk;
#pragma unroll
        for (int li = 0; li < G::SI; ++li) {
            const int e = li * G::SP + p, g0 = o0 - G::HI + li;
            const uint8_t n = fs_diag(L, dv[m][li], g0, g1, g2);
            const bool in = inside(L, g0, g1, g2);
            rs[e] = kScale ? rv[m][li] * vs : rv[m][li];
            nds[e] = n;
            // solid cells take no prolongation (they carry no pressure); outside
            // the level both loads read 0 and 0 + 0 stays +0
            zs[e] = (in && n == kNoUpd) ? 0.f : __fadd_rn(zv[m][li], in ? zp[m][li >> 1] : 0.f);
        }
    }
    __syncthreads();
    fs_pass<T, H, H - 1>(zs, rs, nds, ny, o0, o1, o2, 1);
    fs_pass<T, H, H - 2>(zs, rs, nds, ny, o0, o1, o2, 0);
    fs_round<ZT, G::NS>(zs);
    __syncthreads();
    fs_pass<T, H, H - 3>(zs, rs, nds, ny, o0, o1, o2, 1);
    fs_pass<T, H, H - 4>(zs, rs, nds, ny, o0, o1, o2, 0);
    double dsum = 0.0;
    for (int e = threadIdx.x; e < T::I * T::J * T::K; e += kFuseThreads) {
        const int ti = e / (T::J * T::K), rem = e - ti * (T::J * T::K), tj = rem / T::K, tk = rem - tj * T::K;
        const int g0 = o0 + ti, g1 = o1 + tj, g2 = o2 + tk;
        if (!inside(L, g0, g1, g2)) continue;
        const int le = ((ti + G::HI) * G::SJ + tj + H) * G::SK + tk + H;
        const int64_t q = ((int64_t)g0 * L.d[1] + g1) * L.d[2] + g2;
        const ZT zq = from_f<ZT>(zs[le]);
        zout[q] = zq;
        if (DOT && nds[le] != kNoUpd) dsum += (double)to_f(zq) * (double)to_f(r[q]);
    }
    if (DOT) {
        __shared__ double sh[kFuseThreads / 64];
        dsum = block_sum(dsum, sh);
        // z was stored scaled when the first level is
        if (threadIdx.x == 0) partial[blockIdx.x] = kScale ? dsum * (double)sc->vinv : dsum;
    }
}

// update_p (3D/mgpcg.py:147-150) fused with compute_Ap (:117-122) and the
// p.Ap reduction: p_new = z + beta p_old on the tile and a one-cell halo,
// then Ap from LDS.  p_new goes to a separate buffer because neighbouring
// tiles still read p_old.
template <int DIM, typename ZT>
__global__ void __launch_bounds__(kBlock) k_ap_dot_p(LvlDev L, TileMap tm, const ZT* __restrict__ z,
                                                     const float* __restrict__ p_old, float* __restrict__ p_new,
                                                     float* __restrict__ Ap, const MGScalars* __restrict__ sc,
                                                     double* __restrict__ partial) {
    using T = MgTile<DIM>;
    constexpr int HI = DIM == 3 ? 1 : 0;
    constexpr int SI = T::I + 2 * HI, SJ = T::J + 2, SK = T::K + 2, SP = SJ * SK, NS = SI * SP;
    constexpr int NL = (NS + kBlock - 1) / kBlock;
    constexpr int NT = T::I * T::J * T::K;
    static_assert(NT % kBlock == 0, "whole tile per thread sweep");
    __shared__ float ps[NS];
    __shared__ double sh[kBlock / 64];
    // past convergence the search direction stays as it is (a second idle
    // iteration would otherwise write its partner buffer, the current p) and
    // nothing reads Ap or the partials
    if (pcg_stopped(sc)) return;
    int t[3];
    tile_of_block(tm, t);
    const int o0 = t[0] * T::I, o1 = t[1] * T::J, o2 = t[2] * T::K;
    const int b0 = o0 - HI, b1 = o1 - 1, b2 = o2 - 1;
    const int tid = threadIdx.x;
    const __amdgpu_buffer_rsrc_t rz = lvl_rsrc(z, L), rp = lvl_rsrc(p_old, L);
    float zv[NL], pv[NL];
#pragma unroll
    for (int k = 0; k < NL; ++k) {
        const int e = tid + k * kBlock;
        const int lk = e % SK, lj = (e / SK) % SJ, li = e / SP;
        const bool in = e < NS && inside(L, b0 + li, b1 + lj, b2 + lk);
        const uint32_t q = (uint32_t)(((b0 + li) * L.d[1] + b1 + lj) * L.d[2] + b2 + lk);
        zv[k] = ldv<ZT>(rz, in ? q * (uint32_t)sizeof(ZT) : kOOB);
        pv[k] = ldv<float>(rp, in ? q * 4u : kOOB);
    }
    const float beta = sc->beta;
    const float zi = sizeof(ZT) == 2 ? sc->vinv : 1.0f;  // 16-bit z is stored scaled
#pragma unroll
    for (int k = 0; k < NL; ++k) {
        const int e = tid + k * kBlock;
        const int lk = e % SK, lj = (e / SK) % SJ, li = e / SP;
        const float zk = sizeof(ZT) == 2 ? zv[k] * zi : zv[k];
        if (e < NS) ps[e] = inside(L, b0 + li, b1 + lj, b2 + lk) ? __fadd_rn(zk, __fmul_rn(beta, pv[k])) : 0.f;
    }
    const bool interior = box_interior(L, o0, o1, o2, o0 + T::I, o1 + T::J, o2 + T::K, DIM == 3);
    const float n_int = (float)L.nb2;
    __syncthreads();
    double s = 0.0;
#pragma unroll 4
    for (int e = tid; e < NT; e += kBlock) {
        const int tk = e % T::K, tj = (e / T::K) % T::J, ti = e / (T::K * T::J);
        const int g0 = o0 + ti, g1 = o1 + tj, g2 = o2 + tk;
        const int le = ((ti + HI) * SJ + tj + 1) * SK + tk + 1;
        const float pq = ps[le];
        float nf = interior ? n_int : (float)(L.nb2 - walls(L, g0, g1, g2));
        bool fluid = true;
        if (L.dg) {
            const int n = ld_dg(dg_rsrc(L), L, g0, g1, g2);
            nf = (float)n;
            fluid = n != 0;
        }
        const float a = fluid ? __fsub_rn(__fmul_rn(nf, pq), nsum_lds<DIM>(ps + le, SP, SK)) : 0.f;
        if (inside(L, g0, g1, g2)) {
            const int q = (g0 * L.d[1] + g1) * L.d[2] + g2;
            p_new[q] = pq;
            Ap[q] = a;
            s += (double)pq * (double)a;
        }
    }
    s = block_sum(s, sh);
    if (tid == 0) partial[blockIdx.x] = s;
}

// k_ap_dot_p for 3D levels whose axis-2 extent is a multiple of 4, on groups
// of 4 cells (see k_sweep3): p_new = z + beta p_old on the tile plus a
// one-cell ring along axes 0 / 1 and one group each side along axis 2 goes to
// LDS (the items and their order are k_sweep3's first pass, the tile first),
// then each tile group forms Ap from its own p_new (kept in registers) and
// its neighbours in LDS.  Per cell the arithmetic is k_ap_dot_p's.
template <typename ZT>
__global__ void __launch_bounds__(kSw3Threads, 4) k_ap_dot_p3(LvlDev L, TileMap tm, const ZT* __restrict__ z,
                                                          const float* __restrict__ p_old,
                                                          float* __restrict__ p_new, float* __restrict__ Ap,
                                                          const MGScalars* __restrict__ sc,
                                                          double* __restrict__ partial) {
    using T = MgTile<3>;
    constexpr int NT = kSw3Threads;
    constexpr int TG = T::K / 4, GK = TG + 2;
    constexpr int SI = T::I + 2, SJ = T::J + 2, SK = 4 * GK, SP = SJ * SK;
    constexpr int NI = SI * SJ * GK, NR = (NI + NT - 1) / NT;  // p_new items (the tile first), rounds
    constexpr int P1G = T::I * T::J * TG, P1R = P1G / NT;     // the tile's groups
    static_assert(P1G % NT == 0 && P1R <= NR, "the tile fills whole rounds");
    __shared__ __align__(16) float ps[SI * SP];
    __shared__ double sh[NT / 64];
    // past convergence the search direction stays as it is (a second idle
    // iteration would otherwise write its partner buffer, the current p) and
    // nothing reads Ap or the partials
    if (pcg_stopped(sc)) return;
    int t[3];
    tile_of_block(tm, t);
    const int o0 = t[0] * T::I, o1 = t[1] * T::J, o2 = t[2] * T::K, bk = o2 - 4;
    const int d0 = L.d[0], d1 = L.d[1], d2 = L.d[2];
    const int tid = threadIdx.x;
    const bool solid = L.dg != nullptr;
    const __amdgpu_buffer_rsrc_t rz = lvl_rsrc(z, L), rp = lvl_rsrc(p_old, L);
    auto in3 = [&](int g0, int g1, int gk) {
        return (unsigned)g0 < (unsigned)d0 && (unsigned)g1 < (unsigned)d1 && (unsigned)gk < (unsigned)d2;
    };
    auto cell = [&](int g0, int g1, int gk) { return (uint32_t)((g0 * d1 + g1) * d2 + gk); };
    const float beta = sc->beta;
    const float zi = sizeof(ZT) == 2 ? sc->vinv : 1.0f;  // 16-bit z is stored scaled
    float4 pn[P1R];
    auto item = [&](int m) {
        const int e = tid + m * NT;
        int ri, rj, lg;
        sweep3_item<T::I, T::J, TG>(e, ri, rj, lg);
        const int g0 = o0 - 1 + ri, g1 = o1 - 1 + rj, gk = bk + 4 * lg;
        const bool in = e < NI && in3(g0, g1, gk);
        const uint32_t q = cell(g0, g1, gk);
        float4 zv = ld4<ZT>(rz, in ? q * (uint32_t)sizeof(ZT) : kOOB);
        const float4 pv = ld4<float>(rp, in ? q * 4u : kOOB);
        if (sizeof(ZT) == 2) {
            zv.x = zv.x * zi;
            zv.y = zv.y * zi;
            zv.z = zv.z * zi;
            zv.w = zv.w * zi;
        }
        float4 v;
        v.x = in ? __fadd_rn(zv.x, __fmul_rn(beta, pv.x)) : 0.f;
        v.y = in ? __fadd_rn(zv.y, __fmul_rn(beta, pv.y)) : 0.f;
        v.z = in ? __fadd_rn(zv.z, __fmul_rn(beta, pv.z)) : 0.f;
        v.w = in ? __fadd_rn(zv.w, __fmul_rn(beta, pv.w)) : 0.f;
        if (e < NI) *reinterpret_cast<float4*>(ps + ri * SP + rj * SK + 4 * lg) = v;
        return v;
    };
#pragma unroll
    for (int m = 0; m < P1R; ++m) pn[m] = item(m);
#pragma unroll
    for (int m = P1R; m < NR; ++m) (void)item(m);
    uint32_t dgv[P1R];
#pragma unroll
    for (int m = 0; m < P1R; ++m) dgv[m] = 0u;
    if (solid) {
        const __amdgpu_buffer_rsrc_t rd = dg_rsrc(L);
#pragma unroll
        for (int m = 0; m < P1R; ++m) {
            int ri, rj, lg;
            sweep3_item<T::I, T::J, TG>(tid + m * NT, ri, rj, lg);
            const int g0 = o0 - 1 + ri, g1 = o1 - 1 + rj, gk = bk + 4 * lg;
            dgv[m] = __builtin_amdgcn_raw_buffer_load_b32(rd, in3(g0, g1, gk) ? cell(g0, g1, gk) : kOOB, 0, 0);
        }
    }
    const bool interior = !solid && box_interior(L, o0, o1, o2, o0 + T::I, o1 + T::J, o2 + T::K, true);
    const float n_int = (float)L.nb2;
    __syncthreads();
    double s = 0.0;
#pragma unroll
    for (int m = 0; m < P1R; ++m) {
        int ri, rj, lg;
        sweep3_item<T::I, T::J, TG>(tid + m * NT, ri, rj, lg);
        const int g0 = o0 - 1 + ri, g1 = o1 - 1 + rj, gk = bk + 4 * lg;
        const int base = ri * SP + rj * SK + 4 * lg;
        const float4 own = pn[m];
        const float4 ip = *reinterpret_cast<const float4*>(ps + base + SP);
        const float4 im = *reinterpret_cast<const float4*>(ps + base - SP);
        const float4 jp = *reinterpret_cast<const float4*>(ps + base + SK);
        const float4 jm = *reinterpret_cast<const float4*>(ps + base - SK);
        const float lf = ps[base - 1], rt = ps[base + 4];
        const float pc[4] = {own.x, own.y, own.z, own.w};
        const float ipc[4] = {ip.x, ip.y, ip.z, ip.w}, imc[4] = {im.x, im.y, im.z, im.w};
        const float jpc[4] = {jp.x, jp.y, jp.z, jp.w}, jmc[4] = {jm.x, jm.y, jm.z, jm.w};
        float a[4];
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            // neighbor_sum order (3D/mgpcg.py:92-104): +e then -e per axis
            float ns = 0.0f;
            ns = __fadd_rn(ns, ipc[c]);
            ns = __fadd_rn(ns, imc[c]);
            ns = __fadd_rn(ns, jpc[c]);
            ns = __fadd_rn(ns, jmc[c]);
            ns = __fadd_rn(ns, c < 3 ? pc[c + 1] : rt);
            ns = __fadd_rn(ns, c > 0 ? pc[c - 1] : lf);
            float nf = n_int;
            bool fluid = true;
            if (solid) {
                const int n = (int)((dgv[m] >> (8 * c)) & 0xffu);
                nf = (float)n;
                fluid = n != 0;
            } else if (!interior) {
                nf = (float)(L.nb2 - walls(L, g0, g1, gk + c));
            }
            a[c] = fluid ? __fsub_rn(__fmul_rn(nf, pc[c]), ns) : 0.f;
        }
        if (in3(g0, g1, gk)) {
            const uint32_t q = cell(g0, g1, gk);
            *reinterpret_cast<float4*>(p_new + q) = own;
            *reinterpret_cast<float4*>(Ap + q) = make_float4(a[0], a[1], a[2], a[3]);
#pragma unroll
            for (int c = 0; c < 4; ++c) s += (double)pc[c] * (double)a[c];
        }
    }
    s = block_sum(s, sh);
    if (tid == 0) partial[blockIdx.x] = s;
}

// ---- coarse levels in one workgroup --------------------------------------
// The levels from `coarse_level` down to the bottom are small; launching their
// sweeps one by one costs a kernel's fixed latency each.  k_coarse runs that
// whole tail of the V-cycle (pre-smoothing, restriction, the bottom solve,
// prolongation, post-smoothing) in one workgroup out of LDS, with the same
// per-cell arithmetic and colour order as the level kernels.  z is kept with
// a one-cell zero border (axes 1 and 2 only for 2D levels) so the neighbour
// sums need no tests.
constexpr int kCoarseMaxLevels = 4;
constexpr int kCoarseLds = 14336;  // floats (56 KiB)
constexpr int kCoarseSlots = 4;    // cells per colour per thread on every level
constexpr int kCoarseThreads = 1024;
constexpr int kCoarseNd = kCoarseLds / 2;  // diagonal bytes: the levels hold at most this many cells

constexpr int kCoarseBottom64 = 4096;  // padded bottom cells of the FP64 bottom solve

// floor(x / d) for 0 <= x, d < 2^16 as a multiply-high by m = ceil(2^32 / d):
// x m / 2^32 = x / d + x (m - 2^32 / d) / 2^32, and the second term is below
// x / 2^32 < 1 / d, too small to carry x / d past the next integer.  m = 0
// stands for d = 1.
__host__ __device__ __forceinline__ uint32_t magic16(int d) { return d > 1 ? 0xFFFFFFFFu / (uint32_t)d + 1u : 0u; }
__device__ __forceinline__ int div16(int x, uint32_t m) { return m ? (int)__umulhi((uint32_t)x, m) : x; }

struct CoarseArgs {
    int nl;         // levels handled; the last one is the bottom
    int sweeps;     // bottom red/black iterations
    float inv_scale;
    int zoff[kCoarseMaxLevels], roff[kCoarseMaxLevels];
    int noff[kCoarseMaxLevels];  // offset of each level's diagonal bytes
    // division magics of each level: by d[1], by d[2] and by (d[2] + 1) / 2
    uint32_t m1[kCoarseMaxLevels], m2[kCoarseMaxLevels], mh[kCoarseMaxLevels];
    LvlDev L[kCoarseMaxLevels];
    void* z[kCoarseMaxLevels];  // ZT arrays
    void* r[kCoarseMaxLevels];  // R0T on the first level, ZT below
};

template <int DIM>
struct Padded {
    int s1, s0;  // LDS strides of axes 1 and 0 (axis 2 has stride 1)
    __device__ __forceinline__ explicit Padded(const LvlDev& L) {
        s1 = L.d[2] + 2;
        s0 = (L.d[1] + 2) * s1;
    }
    __device__ __forceinline__ int at(int i, int j, int k) const {
        return (DIM == 3 ? (i + 1) * s0 : 0) + (j + 1) * s1 + k + 1;
    }
};

template <int DIM>
__host__ __device__ __forceinline__ int padded_size(const int d[3]) {
    return (DIM == 3 ? d[0] + 2 : 1) * (d[1] + 2) * (d[2] + 2);
}

// The cells of one level dealt to this thread, per colour: slot m holds
// entry tid + m * kCoarseThreads of the colour's (row, k / 2) enumeration,
// decoded once per level visit (its LDS offset, r and diagonal kept in
// registers for the visit's passes; zq = -1 for no cell or a solid one).
struct CoarseCells {
    int zq[2][kCoarseSlots];
    float r[2][kCoarseSlots], n[2][kCoarseSlots], y[2][kCoarseSlots];
};

template <int DIM>
__device__ __forceinline__ void coarse_cells(const LvlDev& L, uint32_t m1, uint32_t mh, const float* rs,
                                             const uint8_t* nd, CoarseCells& C) {
    const Padded<DIM> P(L);
    const int h2 = (L.d[2] + 1) / 2, nslot = L.d[0] * L.d[1] * h2;
#pragma unroll
    for (int c = 0; c < 2; ++c)
#pragma unroll
        for (int m = 0; m < kCoarseSlots; ++m) {
            const int s = (int)threadIdx.x + m * kCoarseThreads;
            C.zq[c][m] = -1;
            C.r[c][m] = C.n[c][m] = C.y[c][m] = 0.f;
            if (s < nslot) {
                const int row = div16(s, mh), pk = s - row * h2;
                const int i = div16(row, m1), j = row - i * L.d[1];
                const int k = 2 * pk + ((c + i + j) & 1);
                const int q = row * L.d[2] + k;
                const int n = k < L.d[2] ? (int)nd[q] : 0;
                if (n != 0) {
                    C.zq[c][m] = P.at(i, j, k);
                    C.r[c][m] = rs[q];
                    C.n[c][m] = (float)n;
                    C.y[c][m] = pfm_rcp_small(n);
                }
            }
        }
}

// smooth(l, c) (3D/mgpcg.py:165-172) over the decoded cells of colour c.
template <int DIM, int c>
__device__ __forceinline__ void coarse_pass(float* zs, const Padded<DIM>& P, const CoarseCells& C) {
#pragma unroll
    for (int m = 0; m < kCoarseSlots; ++m)
        if (C.zq[c][m] >= 0)
            zs[C.zq[c][m]] = pfm_div_small(__fadd_rn(C.r[c][m], nsum_lds<DIM>(zs + C.zq[c][m], P.s0, P.s1)),
                                           C.n[c][m], C.y[c][m]);
}

// Residual and restriction (3D/mgpcg.py:152-158) from LDS level f to LDS
// level c, one coarse cell per thread iteration, children summed in fixed
// order, odd extents as in k_restrict.
template <int DIM>
__device__ void coarse_restrict(const LvlDev& Lf, const LvlDev& Lc, uint32_t mc1, uint32_t mc2, const float* zs,
                                const float* rs, const uint8_t* ndf, float* rcs, float inv_scale) {
    const Padded<DIM> P(Lf);
    const int nc = Lc.d[0] * Lc.d[1] * Lc.d[2];
    for (int e = threadIdx.x; e < nc; e += blockDim.x) {
        const int row = div16(e, mc2), C2 = e - row * Lc.d[2];
        const int C0 = div16(row, mc1), C1 = row - C0 * Lc.d[1];
        const int h0 = C0 == Lc.d[0] - 1 ? Lf.d[0] - 1 - 2 * C0 : 1;
        const int h1 = C1 == Lc.d[1] - 1 ? Lf.d[1] - 1 - 2 * C1 : 1;
        const int h2 = C2 == Lc.d[2] - 1 ? Lf.d[2] - 1 - 2 * C2 : 1;
        float acc = 0.0f;
        for (int a = 0; a <= h0; ++a)
            for (int b = 0; b <= h1; ++b)
                for (int c = 0; c <= h2; ++c) {
                    const int i = 2 * C0 + a, j = 2 * C1 + b, k = 2 * C2 + c;
                    const int zq = P.at(i, j, k);
                    const int fq = (i * Lf.d[1] + j) * Lf.d[2] + k;
                    const int n = ndf[fq];
                    if (n == 0) continue;  // solid children add nothing
                    const float nf = (float)n;
                    const float res =
                        __fsub_rn(rs[fq], __fsub_rn(__fmul_rn(nf, zs[zq]), nsum_lds<DIM>(zs + zq, P.s0, P.s1)));
                    acc = __fadd_rn(acc, __fmul_rn(__fmul_rn(res, 1.0f), inv_scale));
                }
        rcs[e] = acc;
    }
}

// prolongate (3D/mgpcg.py:160-163) between padded LDS levels, one fine cell
// per thread iteration.
template <int DIM>
__device__ void coarse_prolong(const LvlDev& Lf, const LvlDev& Lc, uint32_t mf1, uint32_t mf2, float* zs,
                               const float* zcs, const uint8_t* ndf) {
    const Padded<DIM> P(Lf), Pc(Lc);
    const int nf = Lf.d[0] * Lf.d[1] * Lf.d[2];
    for (int e = threadIdx.x; e < nf; e += blockDim.x) {
        if (Lf.dg && ndf[e] == 0) continue;  // solid cells carry no pressure
        const int row = div16(e, mf2), k = e - row * Lf.d[2];
        const int i = div16(row, mf1), j = row - i * Lf.d[1];
        const int c0 = min(i >> 1, Lc.d[0] - 1), c1 = min(j >> 1, Lc.d[1] - 1), c2 = min(k >> 1, Lc.d[2] - 1);
        const int q = P.at(i, j, k);
        zs[q] = __fadd_rn(zs[q], zcs[Pc.at(c0, c1, c2)]);
    }
}

// ZT: storage of z and of the coarse r in HBM; R0T: storage of r on the
// first level handled; BT: arithmetic of the bottom solve (float, or double
// for the mixed-precision solver).
template <int DIM, typename ZT, typename R0T, typename BT>
__global__ void __launch_bounds__(kCoarseThreads) k_coarse(CoarseArgs a, int sub_mean, int dot,
                                                           const MGScalars* __restrict__ sc, double* __restrict__ partial) {
    __shared__ float sm[kCoarseLds];
    __shared__ uint8_t nd[kCoarseNd];
    const int tid = threadIdx.x;
    for (int l = 0; l < a.nl; ++l) {
        const LvlDev& L = a.L[l];
        for (int e = tid; e < padded_size<DIM>(L.d); e += kCoarseThreads) sm[a.zoff[l] + e] = 0.f;
        // the diagonal of every cell (num_fluid_neighbors, or the obstacle's)
        const int n = L.d[0] * L.d[1] * L.d[2];
        for (int q = tid; q < n; q += kCoarseThreads) {
            int v;
            if (L.dg) {
                v = L.dg[q];
            } else {
                const int row = div16(q, a.m2[l]), k = q - row * L.d[2];
                const int i = div16(row, a.m1[l]), j = row - i * L.d[1];
                v = L.nb2 - walls(L, i, j, k);
            }
            nd[a.noff[l] + q] = (uint8_t)v;
        }
    }
    {
        const LvlDev& L0 = a.L[0];
        const int n0 = L0.d[0] * L0.d[1] * L0.d[2];
        const float mean = sub_mean ? sc->mean : 0.f;
        R0T* r0 = (R0T*)a.r[0];
        // 16-bit z with an fp32 first level (level 0): work on vs * r
        const float vs = (sizeof(ZT) == 2 && sizeof(R0T) == 4) ? sc->vs : 1.0f;
        for (int q = tid; q < n0; q += kCoarseThreads) {
            float v = to_f(r0[q]);
            if (sub_mean && !(L0.dg && L0.dg[q] == 0)) {
                v = __fsub_rn(v, mean);
                r0[q] = from_f<R0T>(v);
            }
            sm[a.roff[0] + q] = (sizeof(ZT) == 2 && sizeof(R0T) == 4) ? v * vs : v;
        }
    }
    __syncthreads();
    CoarseCells C;
    for (int l = 0; l < a.nl - 1; ++l) {
        const Padded<DIM> P(a.L[l]);
        float* zs = sm + a.zoff[l];
        coarse_cells<DIM>(a.L[l], a.m1[l], a.mh[l], sm + a.roff[l], nd + a.noff[l], C);
        for (int it = 0; it < 2; ++it) {
            coarse_pass<DIM, 0>(zs, P, C);
            __syncthreads();
            coarse_pass<DIM, 1>(zs, P, C);
            __syncthreads();
        }
        coarse_restrict<DIM>(a.L[l], a.L[l + 1], a.m1[l + 1], a.m2[l + 1], zs, sm + a.roff[l], nd + a.noff[l],
                             sm + a.roff[l + 1], a.inv_scale);
        __syncthreads();
    }
    // bottom: 50 red/black iterations from z = 0
    {
        const int b = a.nl - 1;
        const LvlDev& L = a.L[b];
        const Padded<DIM> P(L);
        float* zs = sm + a.zoff[b];
        coarse_cells<DIM>(L, a.m1[b], a.mh[b], sm + a.roff[b], nd + a.noff[b], C);
        if constexpr (sizeof(BT) == 4) {
            if (L.d[0] * L.d[1] * ((L.d[2] + 1) / 2) <= 64) {
                // every bottom cell is wave 0's (one per lane and colour): a wave's
                // LDS operations take effect in program order, so its passes need
                // no block barrier, only a fence against compiler reordering
                if (tid < 64) {
                    for (int it = 0; it < a.sweeps; ++it) {
                        coarse_pass<DIM, 0>(zs, P, C);
                        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
                        coarse_pass<DIM, 1>(zs, P, C);
                        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
                    }
                }
                __syncthreads();
            } else {
                for (int it = 0; it < a.sweeps; ++it) {
                    coarse_pass<DIM, 0>(zs, P, C);
                    __syncthreads();
                    coarse_pass<DIM, 1>(zs, P, C);
                    __syncthreads();
                }
            }
        } else {
            // the coarsest-level solve in double, rounded to fp32 once
            __shared__ double bz[kCoarseBottom64];
            const int np = padded_size<DIM>(L.d);
            for (int e = tid; e < np; e += kCoarseThreads) bz[e] = 0.0;
            __syncthreads();
            for (int it = 0; it < a.sweeps; ++it)
#pragma unroll
                for (int c = 0; c < 2; ++c) {
#pragma unroll
                    for (int m = 0; m < kCoarseSlots; ++m)
                        if (C.zq[c][m] >= 0) {
                            const double* x = bz + C.zq[c][m];
                            double ns = 0.0;
                            if (DIM == 3) ns = (ns + x[P.s0]) + x[-P.s0];
                            ns = (((ns + x[P.s1]) + x[-P.s1]) + x[1]) + x[-1];
                            bz[C.zq[c][m]] = ((double)C.r[c][m] + ns) / (double)C.n[c][m];
                        }
                    __syncthreads();
                }
            for (int e = tid; e < np; e += kCoarseThreads) zs[e] = (float)bz[e];
            __syncthreads();
        }
    }
    for (int l = a.nl - 2; l >= 0; --l) {
        const Padded<DIM> P(a.L[l]);
        float* zs = sm + a.zoff[l];
        coarse_prolong<DIM>(a.L[l], a.L[l + 1], a.m1[l], a.m2[l], zs, sm + a.zoff[l + 1], nd + a.noff[l]);
        coarse_cells<DIM>(a.L[l], a.m1[l], a.mh[l], sm + a.roff[l], nd + a.noff[l], C);
        __syncthreads();
        for (int it = 0; it < 2; ++it) {
            coarse_pass<DIM, 1>(zs, P, C);
            __syncthreads();
            coarse_pass<DIM, 0>(zs, P, C);
            __syncthreads();
        }
    }
    // write every level back, as the level-by-level V-cycle leaves them
    double s = 0.0;
    for (int l = 0; l < a.nl; ++l) {
        const LvlDev& L = a.L[l];
        const Padded<DIM> P(L);
        const int n = L.d[0] * L.d[1] * L.d[2];
        for (int q = tid; q < n; q += kCoarseThreads) {
            const int row = div16(q, a.m2[l]), k = q - row * L.d[2];
            const int i = div16(row, a.m1[l]), j = row - i * L.d[1];
            const ZT zq = from_f<ZT>(sm[a.zoff[l] + P.at(i, j, k)]);
            ((ZT*)a.z[l])[q] = zq;
            if (l > 0) ((ZT*)a.r[l])[q] = from_f<ZT>(sm[a.roff[l] + q]);
            if (dot && l == 0) s += (double)to_f(zq) * (double)sm[a.roff[0] + q];
        }
    }
    if (dot) {
        __shared__ double sh[kCoarseThreads / 64];
        s = block_sum(s, sh);
        // both factors carry vs when the first level is scaled
        if (tid == 0) partial[0] = (sizeof(ZT) == 2 && sizeof(R0T) == 4) ? s * (double)sc->vinv * sc->vinv : s;
    }
}

// Bottom solve: `sweeps` red/black iterations on a level small enough for LDS
// (apply_preconditioner's bottom loop, 3D/mgpcg.py:192-194), starting from
// z = 0 (the fill at :190).  Each thread owns up to kBottomPer cells whose
// indices, colours, diagonals and right-hand sides are decoded once; the
// sweeps then only touch LDS.
constexpr int kBottomMax = 8192;
constexpr int kBottomThreads = 1024;
constexpr int kBottomPer = kBottomMax / kBottomThreads;
template <typename ZT>
__global__ void __launch_bounds__(kBottomThreads) k_bottom(LvlDev L, int n, ZT* __restrict__ z,
                                                           const ZT* __restrict__ r, int sweeps) {
    __shared__ float zs[kBottomMax];
    int qi[kBottomPer], col[kBottomPer];
    int nb[kBottomPer][6];
    float rv[kBottomPer], mv[kBottomPer];
#pragma unroll
    for (int k = 0; k < kBottomPer; ++k) {
        const int q = threadIdx.x + k * kBottomThreads;
        qi[k] = q;
        col[k] = -1;
        if (q < n) {
            int I[3];
            cell_ijk(L, q, I);
            col[k] = solid_at(L, q) ? -1 : (I[0] + I[1] + I[2]) & 1;
            rv[k] = to_f(r[q]);
            mv[k] = nfn(L, I);
            const int st[3] = {L.d[1] * L.d[2], L.d[2], 1};
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                // neighbour order of neighbor_sum: +e then -e per axis; -1 = absent
                nb[k][2 * a] = I[a] < L.d[a] - 1 ? q + st[a] : -1;
                nb[k][2 * a + 1] = I[a] > 0 ? q - st[a] : -1;
            }
            zs[q] = 0.f;
        }
    }
    __syncthreads();
    for (int it = 0; it < sweeps; ++it) {
        for (int phase = 0; phase < 2; ++phase) {
#pragma unroll
            for (int k = 0; k < kBottomPer; ++k) {
                if (col[k] != phase) continue;
                float sum = 0.0f;
#pragma unroll
                for (int e = 0; e < 6; ++e)
                    if (nb[k][e] >= 0) sum = __fadd_rn(sum, zs[nb[k][e]]);
                zs[qi[k]] = __fdiv_rn(__fadd_rn(rv[k], sum), mv[k]);
            }
            __syncthreads();
        }
    }
#pragma unroll
    for (int k = 0; k < kBottomPer; ++k)
        if (qi[k] < n) z[qi[k]] = from_f<ZT>(zs[qi[k]]);
}

// ---- vector kernels of the CG recurrence ----------------------------------

// Partial sums of a.b (and, if b == nullptr, of a) in double.
template <typename AT>
__global__ void __launch_bounds__(kBlock) k_dot(int64_t n, const AT* __restrict__ a,
                                                const float* __restrict__ b, double* __restrict__ partial,
                                                const float* __restrict__ ascale) {
    __shared__ double sh[kBlock / 64];
    double s = 0.0;
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n;
         q += (int64_t)gridDim.x * blockDim.x)
        s += b ? (double)to_f(a[q]) * (double)b[q] : (double)to_f(a[q]);
    s = block_sum(s, sh);
    if (threadIdx.x == 0) partial[blockIdx.x] = ascale ? s * (double)*ascale : s;
}

// The fixed-order sum of k_finalize (kFinThreads = 1024 threads: thread t adds
// partials t, t + 1024, ... in turn, each wave sums its lanes, thread 0 adds
// the 16 wave sums in order) formed by a 256-thread block: physical wave w
// holds the lanes of the virtual waves w, w + 4, w + 8 and w + 12, so every
// block gets the bits k_finalize would.  Every thread returns the sum.
__device__ __forceinline__ double fin_sum(const double* __restrict__ part, int nparts, double* sh16) {
    static_assert(kBlock * 4 == 1024 && kBlock % 64 == 0, "four virtual finalize threads per thread");
    double v[4];
#pragma unroll
    for (int m = 0; m < 4; ++m) {
        double a = 0.0;
        for (int b = (int)threadIdx.x + m * kBlock; b < nparts; b += 1024) a += part[b];
        v[m] = wave_sum(a);
    }
    const int pw = (int)threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
#pragma unroll
        for (int m = 0; m < 4; ++m) sh16[pw + 4 * m] = v[m];
    }
    __syncthreads();
    double r = 0.0;
#pragma unroll
    for (int i = 0; i < 16; ++i) r += sh16[i];
    __syncthreads();
    return r;
}

// update_x + update_r (3D/mgpcg.py:137-145) fused with the r.r and sum(r)
// reductions (convergence test :249-255 and recenter :174-180), and with
// alpha = zTr / (p.Ap + eps) (:239-243) formed in every block from the Ap
// launch's partials as k_finalize forms it (block 0 also stores it).
// XX: also the partials of x.x (third block of kMgRedBlocks partials).
template <bool XX>
__global__ void __launch_bounds__(kBlock) k_update_xr(int64_t n, float* __restrict__ x,
                                                      float* __restrict__ r, const float* __restrict__ p,
                                                      const float* __restrict__ Ap, MGScalars* __restrict__ sc,
                                                      const double* __restrict__ ap_part, int ap_n, double eps,
                                                      double* __restrict__ partial) {
    __shared__ double sh[kBlock / 64];
    __shared__ double sh16[16];
    // past convergence x and r stay as they are (and nothing reads the
    // partials: k_finalize leaves the scalars alone)
    if (pcg_stopped(sc)) return;
    const double pAp = fin_sum(ap_part, ap_n, sh16);
    const float alpha = (float)(sc->zTr / (pAp + eps));
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        sc->pAp = pAp;
        sc->alpha = alpha;
    }
    double s2 = 0.0, s1 = 0.0, s3 = 0.0;
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n;
         q += (int64_t)gridDim.x * blockDim.x) {
        const float xq = __fadd_rn(x[q], __fmul_rn(alpha, p[q]));
        const float rq = __fsub_rn(r[q], __fmul_rn(alpha, Ap[q]));
        x[q] = xq;
        r[q] = rq;
        s2 += (double)rq * (double)rq;
        s1 += (double)rq;
        if (XX) s3 += (double)xq * (double)xq;
    }
    s2 = block_sum(s2, sh);
    s1 = block_sum(s1, sh);
    if (XX) s3 = block_sum(s3, sh);
    if (threadIdx.x == 0) {
        partial[blockIdx.x] = s2;
        partial[kMgRedBlocks + blockIdx.x] = s1;
        if (XX) partial[2 * kMgRedBlocks + blockIdx.x] = s3;
    }
}

// update_p: 3D/mgpcg.py:147-150 (in place; used where no Ap follows).
template <typename ZT>
__global__ void __launch_bounds__(kBlock) k_update_p(int64_t n, float* __restrict__ p,
                                                     const ZT* __restrict__ z,
                                                     const MGScalars* __restrict__ sc) {
    if (pcg_stopped(sc)) return;
    const float beta = sc->beta;
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n;
         q += (int64_t)gridDim.x * blockDim.x)
        p[q] = __fadd_rn(sizeof(ZT) == 2 ? to_f(z[q]) * sc->vinv : to_f(z[q]), __fmul_rn(beta, p[q]));
}

__global__ void __launch_bounds__(kBlock) k_sub_mean(int64_t n, float* __restrict__ r,
                                                     const MGScalars* __restrict__ sc,
                                                     const uint8_t* __restrict__ dg) {
    if (pcg_stopped(sc)) return;  // past convergence r stays as it is
    const float m = sc->mean;
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n;
         q += (int64_t)gridDim.x * blockDim.x)
        if (!dg || dg[q] != 0) r[q] = __fsub_rn(r[q], m);  // solid cells stay 0
}

// vs = 2^(round(-log2(rms r)) - vshift) from r.r over n cells (1 when r
// vanishes).
__device__ __forceinline__ void set_vscale(MGScalars* sc, double rtr, double n, int vshift) {
    if (rtr > 0.0 && n > 0.0) {
        const double e = rint(-0.5 * log2(rtr / n)) - vshift;
        const double c = fmin(fmax(e, -100.0), 100.0);
        sc->vs = (float)exp2(c);
        sc->vinv = (float)exp2(-c);
    } else {
        sc->vs = sc->vinv = 1.0f;
    }
}

enum FinalizeMode { kFinMean = 0, kFinR0 = 1, kFinZTr = 2, kFinAlpha = 3, kFinRTr = 4, kFinBeta = 5, kFinRTrX = 6 };

// Fixed-order reduction of the per-block partials and the scalar recurrences
// of solve() (3D/mgpcg.py:229-275), all in double.
// test: after r.r, stop the solve when it is below the threshold (the
// convergence test of 3D/mgpcg.py:255); kFinBeta counts the iteration and
// stops the solve at max_it.
// One block of kFinThreads: the partial sums (up to one per level-0 tile)
// spread over more lanes shorten the serial part of this one-block launch.
constexpr int kFinThreads = 1024;

__global__ void __launch_bounds__(kFinThreads) k_finalize(const double* __restrict__ partial, int nparts, int mode,
                                                     double eps, double ncell, int vshift, int test, MGScalars* sc) {
    if (pcg_stopped(sc)) return;
    __shared__ double sh[kFinThreads / 64];
    double s = 0.0;
    for (int b = threadIdx.x; b < nparts; b += blockDim.x) s += partial[b];
    s = block_sum(s, sh);
    double s1 = 0.0, s3 = 0.0;
    if (mode == kFinRTr || mode == kFinRTrX) {
        for (int b = threadIdx.x; b < kMgRedBlocks; b += blockDim.x) s1 += partial[kMgRedBlocks + b];
        s1 = block_sum(s1, sh);
    }
    if (mode == kFinRTrX) {
        for (int b = threadIdx.x; b < kMgRedBlocks; b += blockDim.x) s3 += partial[2 * kMgRedBlocks + b];
        s3 = block_sum(s3, sh);
    }
    if (threadIdx.x != 0) return;
    switch (mode) {
        case kFinMean:
            sc->sum_r = s;
            sc->mean = (float)(s / ncell);
            break;
        case kFinR0:
            sc->r0 = s;
            sc->rTr = s;
            if (sc->tol_mode == 1) sc->thresh = sc->tol * sc->tol * s;
            set_vscale(sc, s, ncell, vshift);
            break;
        case kFinZTr:
            sc->zTr = s;
            break;
        case kFinAlpha:
            sc->pAp = s;
            sc->alpha = (float)(sc->zTr / (s + eps));
            break;
        case kFinRTrX:
            sc->xTx = s3;
            [[fallthrough]];
        case kFinRTr:
            sc->rTr = s;
            sc->sum_r = s1;
            sc->mean = (float)(s1 / ncell);
            set_vscale(sc, s, ncell, vshift);
            if (test && s < sc->thresh) {
                sc->done = 1;
                sc->conv = 1;
            }
            break;
        case kFinBeta: {
            sc->new_zTr = s;
            sc->beta = (float)(s / (sc->zTr + eps));
            sc->zTr = s;
            sc->it += 1;
            if (sc->max_it >= 0 && sc->it >= sc->max_it) sc->done = 1;
            break;
        }
    }
}

__global__ void k_reset_scalars(MGScalars* sc, double tol, int tol_mode, int max_it) {
    float beta = sc->beta;  // the reference's beta field persists across solves
    *sc = MGScalars{};
    sc->beta = beta;
    sc->vs = sc->vinv = 1.0f;
    sc->tol = tol;
    sc->thresh = tol;  // tol_mode 1: tol^2 r0, set with r0
    sc->tol_mode = tol_mode;
    sc->max_it = max_it;
}

// The solve has ended: kernels run normally again.
__global__ void k_clear_stop(MGScalars* sc) { sc->done = 0; }


// ---- field kernels (velocity <-> level 0) ---------------------------------

// divergence (3D/mgpcg.py:380-390) fused with init(r, k) (:71-79).
template <int DIM>
__global__ void __launch_bounds__(kBlock) k_div_init(LvlDev L, const float* __restrict__ ux,
                                                     const float* __restrict__ uy,
                                                     const float* __restrict__ uz, float k,
                                                     float* __restrict__ r, float* __restrict__ z,
                                                     float* __restrict__ Ap, float* __restrict__ p,
                                                     float* __restrict__ x) {
    int I[3];
    if (!cell_of_thread(L, 0, false, I)) return;
    const int64_t q = cell_index(L, I);
    float div;
    if (DIM == 3) {
        const int ry = L.d[1], rz = L.d[2];
        const float vl = ux[((int64_t)I[0] * ry + I[1]) * rz + I[2]];
        const float vr = ux[((int64_t)(I[0] + 1) * ry + I[1]) * rz + I[2]];
        const float vb = uy[((int64_t)I[0] * (ry + 1) + I[1]) * rz + I[2]];
        const float vt = uy[((int64_t)I[0] * (ry + 1) + I[1] + 1) * rz + I[2]];
        const float va = uz[((int64_t)I[0] * ry + I[1]) * (rz + 1) + I[2]];
        const float vc = uz[((int64_t)I[0] * ry + I[1]) * (rz + 1) + I[2] + 1];
        div = __fsub_rn(__fadd_rn(__fsub_rn(__fadd_rn(__fsub_rn(vr, vl), vt), vb), vc), va);
    } else {
        // 2D cell (i, j) sits at (0, i, j); u_x is (n0 + 1, n1), u_y (n0, n1 + 1)
        const int i = I[1], j = I[2], n1 = L.d[2];
        const float vl = ux[(int64_t)i * n1 + j];
        const float vr = ux[(int64_t)(i + 1) * n1 + j];
        const float vb = uy[(int64_t)i * (n1 + 1) + j];
        const float vt = uy[(int64_t)i * (n1 + 1) + j + 1];
        div = __fsub_rn(__fadd_rn(__fsub_rn(vr, vl), vt), vb);
    }
    r[q] = solid_at(L, q) ? 0.f : __fmul_rn(div, k);
    if (z) z[q] = 0.f;
    Ap[q] = 0.f;
    p[q] = 0.f;
    x[q] = 0.f;
}

__global__ void __launch_bounds__(kBlock) k_init_rhs(int64_t n, const float* __restrict__ rhs, float k,
                                                     float* __restrict__ r, float* __restrict__ z,
                                                     float* __restrict__ Ap, float* __restrict__ p,
                                                     float* __restrict__ x, const uint8_t* __restrict__ dg) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    r[q] = (dg && dg[q] == 0) ? 0.f : __fmul_rn(rhs[q], k);
    if (z) z[q] = 0.f;
    Ap[q] = 0.f;
    p[q] = 0.f;
    x[q] = 0.f;
}

// subtract_grad_p: 3D/mgpcg.py:392-418 (pressure outside the domain = 0).
// F describes the face array of `axis` (3D index form, extent d[axis] + 1).
__global__ void __launch_bounds__(kBlock) k_subtract_grad_p(LvlDev L, LvlDev F, int axis,
                                                            const float* __restrict__ pr,
                                                            float* __restrict__ u) {
    int I[3];
    if (!cell_of_thread(F, 0, false, I)) return;
    const int64_t f = cell_index(F, I);
    int R[3] = {I[0], I[1], I[2]}, Lf[3] = {I[0], I[1], I[2]};
    Lf[axis] -= 1;
    const float prv = R[axis] >= L.d[axis] ? 0.0f : pr[cell_index(L, R)];
    const float plv = Lf[axis] < 0 ? 0.0f : pr[cell_index(L, Lf)];
    u[f] = __fsub_rn(u[f], __fsub_rn(prv, plv));
}

// apply_bc (3D/mgpcg.py:359-378, 2D/mgpcg.py:292-324): zero boundary planes
// of the face arrays, every plane of one call in one launch.  Plane k is the
// face index idx[k] along plane_axis[k] of the array u[k] of extents S[k];
// thread t takes entry t - off[k] of the plane with off[k] <= t < off[k + 1]
// (zeroing is idempotent, so planes that share an edge need no order).
constexpr int kMaxBcPlanes = 12;
struct BcPlanes {
    int n;
    float* u[kMaxBcPlanes];
    int S[kMaxBcPlanes][3];
    int plane_axis[kMaxBcPlanes], idx[kMaxBcPlanes];
    int64_t off[kMaxBcPlanes + 1];
};

__global__ void __launch_bounds__(kBlock) k_zero_planes(BcPlanes bp) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= bp.off[bp.n]) return;
    int k = 0;
    while (k + 1 < bp.n && t >= bp.off[k + 1]) ++k;
    const int* S = bp.S[k];
    const int pa = bp.plane_axis[k];
    const int a1 = pa == 0 ? 1 : 0, a2 = pa == 2 ? 1 : 2;
    const int64_t e = t - bp.off[k];
    int I[3];
    I[pa] = bp.idx[k];
    I[a1] = (int)(e / S[a2]);
    I[a2] = (int)(e % S[a2]);
    bp.u[k][((int64_t)I[0] * S[1] + I[1]) * S[2] + I[2]] = 0.0f;
}

// r = b - A x: the true-residual replacement of the mixed-precision PCG,
// made only while r.r exceeds floor2 x.x (see kFloorMargin).
__global__ void __launch_bounds__(kBlock) k_true_residual(LvlDev L, const float* __restrict__ b,
                                                          const float* __restrict__ x, float* __restrict__ r,
                                                          const MGScalars* __restrict__ sc, double floor2) {
    if (pcg_stopped(sc) || !(sc->rTr > floor2 * sc->xTx)) return;
    int I[3];
    if (!cell_of_thread(L, 0, false, I)) return;
    const int64_t q = cell_index(L, I);
    if (solid_at(L, q)) {
        r[q] = 0.0f;
        return;
    }
    // FP64: the stencil terms are large and nearly cancel, so in fp32 the
    // rounding of A x would swamp the residual; in double the result is the
    // residual of the stored x rounded once.
    const int64_t s0 = (int64_t)L.d[1] * L.d[2], s1 = L.d[2];
    const float* xq = x + q;
    double nb = 0.0;
    if (I[0] < L.d[0] - 1) nb += xq[s0];
    if (I[0] > 0) nb += xq[-s0];
    if (I[1] < L.d[1] - 1) nb += xq[s1];
    if (I[1] > 0) nb += xq[-s1];
    if (I[2] < L.d[2] - 1) nb += xq[1];
    if (I[2] > 0) nb += xq[-1];
    r[q] = (float)((double)b[q] - ((double)nfn(L, I) * (double)xq[0] - nb));
}

// ---- solid obstacle (extension of the reference solver) ----------------------

// Free slip at the obstacle: zero the faces of `axis` (F = that face array in
// 3D index form) that have a solid cell on either side.
__global__ void __launch_bounds__(kBlock) k_zero_solid_faces(LvlDev L, LvlDev F, int axis, float* __restrict__ u) {
    int I[3];
    if (!cell_of_thread(F, 0, false, I)) return;
    int R[3] = {I[0], I[1], I[2]}, Lc[3] = {I[0], I[1], I[2]};
    Lc[axis] -= 1;
    const bool sr = R[axis] < L.d[axis] && solid_at(L, cell_index(L, R));
    const bool sl = Lc[axis] >= 0 && solid_at(L, cell_index(L, Lc));
    if (sr || sl) u[cell_index(F, I)] = 0.0f;
}

__global__ void __launch_bounds__(kBlock) k_mask01(int64_t n, const uint8_t* __restrict__ in, uint8_t* __restrict__ out) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) out[q] = in[q] != 0;
}

// A coarse cell is solid when every fine cell it owns is (the last coarse
// cell of an odd fine extent also owns the trailing fine cell).
__global__ void __launch_bounds__(kBlock) k_solid_coarsen(LvlDev Lf, LvlDev Lc, const uint8_t* __restrict__ sf,
                                                          uint8_t* __restrict__ sc) {
    int C[3];
    if (!cell_of_thread(Lc, 0, false, C)) return;
    int hi[3];
    for (int a = 0; a < 3; ++a) hi[a] = (C[a] == Lc.d[a] - 1) ? Lf.d[a] - 1 - 2 * C[a] : 1;
    uint8_t all = 1;
    for (int a = 0; a <= hi[0]; ++a)
        for (int b = 0; b <= hi[1]; ++b)
            for (int c = 0; c <= hi[2]; ++c) {
                int I[3] = {2 * C[0] + a, 2 * C[1] + b, 2 * C[2] + c};
                all &= sf[cell_index(Lf, I)];
            }
    sc[cell_index(Lc, C)] = all;
}

// Diagonal of every cell: 2 dim minus the counting walls and the solid
// neighbours; 0 for a solid cell.  L carries no dg yet (walls only).
__global__ void __launch_bounds__(kBlock) k_solid_diag(LvlDev L, const uint8_t* __restrict__ sol,
                                                       uint8_t* __restrict__ dg) {
    int I[3];
    if (!cell_of_thread(L, 0, false, I)) return;
    const int64_t q = cell_index(L, I);
    if (sol[q]) {
        dg[q] = 0;
        return;
    }
    int n = L.nb2;
    for (int a = 0; a < 3; ++a) {
        if (L.d[a] == 1 && !L.bc_lo[a] && !L.bc_hi[a]) continue;  // the unused axis of a 2D level
        int J[3] = {I[0], I[1], I[2]};
        if (I[a] <= 0) n -= L.bc_lo[a];
        else {
            J[a] = I[a] - 1;
            n -= sol[cell_index(L, J)];
        }
        J[a] = I[a];
        if (I[a] >= L.d[a] - 1) n -= L.bc_hi[a];
        else {
            J[a] = I[a] + 1;
            n -= sol[cell_index(L, J)];
        }
    }
    dg[q] = (uint8_t)n;
}

// ---- host side ------------------------------------------------------------

// Axis of the 3D index form that holds axis a of the problem.
static inline int map_axis(const MGHandle* h, int a) { return a + (3 - h->dim); }

static LvlDev lvl_dev(const MGHandle* h, int l) {
    LvlDev L;
    L.nb2 = 2 * h->dim;
    L.dg = h->dg[l];
    for (int m = 0; m < 3; ++m) {
        L.d[m] = h->lv[l].d[m];
        L.bc_lo[m] = L.bc_hi[m] = 0;
    }
    for (int a = 0; a < h->dim; ++a) {
        const int m = map_axis(h, a), lo = h->bc[a][0], hi = h->bc[a][1];
        L.bc_lo[m] = lo == 2 || (h->dirichlet_counts && lo == 1);
        L.bc_hi[m] = hi == 2 || (h->dirichlet_counts && hi == 1);
    }
    return L;
}

static inline float restrict_scale(const MGHandle* h) { return (float)h->dim - 1.0f; }

// 1 / (dim - 1): 0.5 or 1, exact, so x * inv equals x / (dim - 1) bit for bit.
static inline float restrict_inv_scale(const MGHandle* h) { return h->dim == 3 ? 0.5f : 1.0f; }

static TileMap tile_map(const int d[3], int ti, int tj, int tk) {
    TileMap tm;
    tm.n0 = (d[0] + ti - 1) / ti;
    tm.n1 = (d[1] + tj - 1) / tj;
    tm.n2 = (d[2] + tk - 1) / tk;
    tm.total = tm.n0 * tm.n1 * tm.n2;
    return tm;
}

static TileMap sweep_tiles(const MGHandle* h, int l) {
    return h->dim == 3 ? tile_map(h->lv[l].d, MgTile<3>::I, MgTile<3>::J, MgTile<3>::K)
                       : tile_map(h->lv[l].d, MgTile<2>::I, MgTile<2>::J, MgTile<2>::K);
}

static TileMap restrict_tiles(const MGHandle* h, int l) {
    return h->dim == 3 ? tile_map(h->lv[l + 1].d, RsTile<3>::I, RsTile<3>::J, RsTile<3>::K)
                       : tile_map(h->lv[l + 1].d, RsTile<2>::I, RsTile<2>::J, RsTile<2>::K);
}

// The tiled restriction needs fine extents of exactly twice the coarse ones.
static bool restrict_tiled_ok(const MGHandle* h, int l) {
    for (int m = 0; m < 3; ++m) {
        const int f = h->lv[l].d[m], c = h->lv[l + 1].d[m];
        if (!(f == 2 * c || (f == 1 && c == 1 && m < 3 - h->dim))) return false;
    }
    return true;
}

static int fill(float* p, int64_t n, float v, hipStream_t s) {
    unsigned b = grid_for(n);
    if (b > 4096) b = 4096;
    hipLaunchKernelGGL(k_fill_f32, dim3(b), dim3(kBlock), 0, s, p, n, v);
    PFM_LAUNCH_CHECK();
    return PFM_OK;
}

static const dim3 kCellBlock(kTx, kTy, 1);

// The per-pass entry points (smooth, restrict, prolongate one level) work on
// fp32 level arrays, i.e. precision 0.
static int smooth(MGHandle* h, int l, int phase, hipStream_t s) {
    LvlDev L = lvl_dev(h, l);
    hipLaunchKernelGGL(k_smooth, cell_grid(L.d, true), kCellBlock, 0, s, L, (float*)h->z[l], (const float*)h->r[l],
                       phase);
    PFM_LAUNCH_CHECK();
    return PFM_OK;
}

static int restrict_level(MGHandle* h, int l, hipStream_t s) {
    LvlDev Lf = lvl_dev(h, l), Lc = lvl_dev(h, l + 1);
    hipLaunchKernelGGL(k_restrict, cell_grid(Lc.d, false), kCellBlock, 0, s, Lf, Lc, restrict_scale(h),
                       (const float*)h->z[l], (const float*)h->r[l], (float*)h->r[l + 1]);
    PFM_LAUNCH_CHECK();
    return PFM_OK;
}

static int prolongate_level(MGHandle* h, int l, hipStream_t s) {
    LvlDev Lf = lvl_dev(h, l), Lc = lvl_dev(h, l + 1);
    hipLaunchKernelGGL(k_prolongate, cell_grid(Lf.d, false), kCellBlock, 0, s, Lf, Lc, (float*)h->z[l],
                       (const float*)h->z[l + 1]);
    PFM_LAUNCH_CHECK();
    return PFM_OK;
}

#define PFM_TRY(expr)            \
    do {                         \
        int _rc = (expr);        \
        if (_rc) return _rc;     \
    } while (0)

enum SweepKind { kSweepPlain = 0, kSweepZero = 1, kSweepProlong = 2 };

template <int DIM, typename ZT, typename RT>
static void launch_rbgs(int kind, bool sub_mean, bool dot, dim3 g, hipStream_t s, LvlDev L, LvlDev Lc, TileMap tm,
                        const ZT* zin, const ZT* zc, const RT* r, ZT* zout, int c0, const MGScalars* sc,
                        double* partial) {
#define PFM_RBGS(Z, P, M, D)                                                                                     \
    hipLaunchKernelGGL((k_rbgs<DIM, ZT, RT, Z, P, M, D>), g, dim3(kSweepThreads), 0, s, L, Lc, tm, zin, zc, r, zout, c0, \
                       sc, partial)
    if (kind == kSweepZero) {
        if (sub_mean) PFM_RBGS(true, false, true, false);
        else PFM_RBGS(true, false, false, false);
    } else if (kind == kSweepProlong) {
        PFM_RBGS(false, true, false, false);
    } else if (sub_mean) {
        PFM_RBGS(false, false, true, false);
    } else if (dot) {
        PFM_RBGS(false, false, false, true);
    } else {
        PFM_RBGS(false, false, false, false);
    }
#undef PFM_RBGS
}

template <typename ZT, typename RT>
static void launch_sweep3(int kind, bool sub_mean, bool dot, dim3 g, hipStream_t s, LvlDev L, LvlDev Lc, TileMap tm,
                          const ZT* zin, const ZT* zc, const RT* r, ZT* zout, int c0, const MGScalars* sc,
                          double* partial) {
#define PFM_SW3(Z, P, M, D)                                                                                   \
    hipLaunchKernelGGL((k_sweep3<ZT, RT, Z, P, M, D>), g, dim3(kSw3Threads), 0, s, L, Lc, tm, zin, zc, r, zout, \
                       c0, sc, partial)
    if (kind == kSweepZero) {
        if (sub_mean) PFM_SW3(true, false, true, false);
        else PFM_SW3(true, false, false, false);
    } else if (kind == kSweepProlong) {
        PFM_SW3(false, true, false, false);
    } else if (sub_mean) {
        PFM_SW3(false, false, true, false);
    } else if (dot) {
        PFM_SW3(false, false, false, true);
    } else {
        PFM_SW3(false, false, false, false);
    }
#undef PFM_SW3
}

// Two colour passes (c0, c0 ^ 1) on level l, z_in -> z_out (ZT arrays).
// 3D levels whose axis-2 extent is a multiple of 4 run the group kernel
// k_sweep3, the others k_rbgs (same tiles, same bits).
template <typename ZT>
static int rbgs(MGHandle* h, int l, int kind, int c0, const ZT* zin, ZT* zout, bool sub_mean, bool dot,
                hipStream_t s) {
    const LvlDev L = lvl_dev(h, l);
    const LvlDev Lc = kind == kSweepProlong ? lvl_dev(h, l + 1) : L;
    const ZT* zc = kind == kSweepProlong ? (const ZT*)h->z[l + 1] : nullptr;
    const TileMap tm = sweep_tiles(h, l);
    const dim3 g((unsigned)tm.total);
    // r is fp32 on level 0 (the PCG residual) and ZT below
    const bool groups = h->dim == 3 && L.d[2] % 4 == 0;
    if (l == 0 || std::is_same<ZT, float>::value) {
        const float* r = (const float*)h->r[l];
        if (groups)
            launch_sweep3<ZT, float>(kind, sub_mean, dot, g, s, L, Lc, tm, zin, zc, r, zout, c0, h->sc, h->partial);
        else if (h->dim == 3)
            launch_rbgs<3, ZT, float>(kind, sub_mean, dot, g, s, L, Lc, tm, zin, zc, r, zout, c0, h->sc, h->partial);
        else
            launch_rbgs<2, ZT, float>(kind, sub_mean, dot, g, s, L, Lc, tm, zin, zc, r, zout, c0, h->sc, h->partial);
    } else {
        const ZT* r = (const ZT*)h->r[l];
        if (groups)
            launch_sweep3<ZT, ZT>(kind, false, false, g, s, L, Lc, tm, zin, zc, r, zout, c0, h->sc, h->partial);
        else if (h->dim == 3)
            launch_rbgs<3, ZT, ZT>(kind, false, false, g, s, L, Lc, tm, zin, zc, r, zout, c0, h->sc, h->partial);
        else
            launch_rbgs<2, ZT, ZT>(kind, false, false, g, s, L, Lc, tm, zin, zc, r, zout, c0, h->sc, h->partial);
    }
    PFM_LAUNCH_CHECK();
    return PFM_OK;
}

// Residual + restriction of level l into r[l + 1] (tiled when the extents
// halve exactly; otherwise the per-cell kernel, typed for 16-bit storage).  With
// sub_mean, r[l] is recentred first (written back).
template <typename ZT>
static int restrict_fused(MGHandle* h, int l, bool sub_mean, hipStream_t s) {
    if (!restrict_tiled_ok(h, l)) {
        if (sub_mean) {
            hipLaunchKernelGGL(k_sub_mean, dim3(kMgRedBlocks), dim3(kBlock), 0, s, h->lv[l].n, (float*)h->r[l],
                               h->sc, h->dg[l]);
            PFM_LAUNCH_CHECK();
        }
        if (std::is_same<ZT, float>::value) return restrict_level(h, l, s);
        // 16-bit storage: the typed per-cell gather
        const LvlDev Lf = lvl_dev(h, l), Lc = lvl_dev(h, l + 1);
        const float isc = restrict_inv_scale(h);
        if (l == 0)
            hipLaunchKernelGGL((k_restrict_any<ZT, float>), cell_grid(Lc.d, false), kCellBlock, 0, s, Lf, Lc, isc,
                               (const ZT*)h->z[l], (const float*)h->r[l], (ZT*)h->r[l + 1], h->sc);
        else
            hipLaunchKernelGGL((k_restrict_any<ZT, ZT>), cell_grid(Lc.d, false), kCellBlock, 0, s, Lf, Lc, isc,
                               (const ZT*)h->z[l], (const ZT*)h->r[l], (ZT*)h->r[l + 1], h->sc);
        PFM_LAUNCH_CHECK();
        return PFM_OK;
    }
    const LvlDev Lf = lvl_dev(h, l), Lc = lvl_dev(h, l + 1);
    const TileMap tm = restrict_tiles(h, l);
    const dim3 g((unsigned)tm.total);
    const float isc = restrict_inv_scale(h);
    const ZT* z = (const ZT*)h->z[l];
    ZT* rc = (ZT*)h->r[l + 1];
#define PFM_RS(D, RT, M) \
    hipLaunchKernelGGL((k_restrict_tiled<D, ZT, RT, M>), g, dim3(kBlock), 0, s, Lf, Lc, tm, isc, z, (RT*)h->r[l], rc, h->sc)
    if (l == 0 || std::is_same<ZT, float>::value) {
        if (h->dim == 3) {
            if (sub_mean) PFM_RS(3, float, true);
            else PFM_RS(3, float, false);
        } else {
            if (sub_mean) PFM_RS(2, float, true);
            else PFM_RS(2, float, false);
        }
    } else {
        if (h->dim == 3) PFM_RS(3, ZT, false);
        else PFM_RS(2, ZT, false);
    }
#undef PFM_RS
    PFM_LAUNCH_CHECK();
    return PFM_OK;
}

// 50 red/black iterations from z = 0 on the coarsest level (3D/mgpcg.py:190-194).
template <typename ZT>
static int bottom(MGHandle* h, int b, hipStream_t s) {
    if (h->lv[b].n <= kBottomMax && (b > 0 || std::is_same<ZT, float>::value)) {
        const LvlDev Lb = lvl_dev(h, b);
        hipLaunchKernelGGL(k_bottom<ZT>, dim3(1), dim3(kBottomThreads), 0, s, Lb, (int)h->lv[b].n, (ZT*)h->z[b],
                           (const ZT*)h->r[b], 50);
        PFM_LAUNCH_CHECK();
        return PFM_OK;
    }
    // one launch per red/black iteration, alternating buffers so the last lands in z[b]
    const ZT* zin = nullptr;
    for (int j = 0; j < 50; ++j) {
        ZT* zout = (ZT*)(((49 - j) & 1) ? h->z2[b] : h->z[b]);
        PFM_TRY(rbgs<ZT>(h, b, j == 0 ? kSweepZero : kSweepPlain, 0, zin, zout, false, false, s));
        zin = zout;
    }
    return PFM_OK;
}

// Level from which k_coarse can run the rest of the V-cycle (all those
// levels fit its LDS and the register slots of its cell decode), or -1.
static int pick_coarse_level(const MGHandle* h) {
    const int L = h->n_levels;
    const int* db = h->lv[L - 1].d;
    const int pb = h->dim == 3 ? padded_size<3>(db) : padded_size<2>(db);
    if (h->precision != 0 && pb > kCoarseBottom64) return -1;
    int best = -1;
    int64_t used = 0;
    for (int l = L - 1; l >= 0 && L - l <= kCoarseMaxLevels; --l) {
        const int* d = h->lv[l].d;
        if ((int64_t)d[0] * d[1] * ((d[2] + 1) / 2) > (int64_t)kCoarseSlots * kCoarseThreads) break;
        used += (h->dim == 3 ? padded_size<3>(d) : padded_size<2>(d)) + h->lv[l].n;
        if (used > kCoarseLds) break;
        best = l;
    }
    return best;
}

template <typename ZT>
static int coarse(MGHandle* h, bool sub_mean, bool dot, hipStream_t s) {
    CoarseArgs a{};
    const int lc = h->coarse_level;
    a.nl = h->n_levels - lc;
    a.sweeps = 50;
    a.inv_scale = restrict_inv_scale(h);
    int off = 0, noff = 0;
    for (int i = 0; i < a.nl; ++i) {
        const int l = lc + i;
        const int* d = h->lv[l].d;
        a.L[i] = lvl_dev(h, l);
        a.zoff[i] = off;
        off += h->dim == 3 ? padded_size<3>(d) : padded_size<2>(d);
        a.roff[i] = off;
        off += (int)h->lv[l].n;
        a.noff[i] = noff;
        noff += (int)h->lv[l].n;
        a.m1[i] = magic16(d[1]);
        a.m2[i] = magic16(d[2]);
        a.mh[i] = magic16((d[2] + 1) / 2);
        a.z[i] = h->z[l];
        a.r[i] = h->r[l];
    }
    // every level's padded z holds at least its cells, so the LDS bound also
    // bounds the diagonal bytes and keeps indices and extents below 2^16
    if (off > kCoarseLds || noff > kCoarseNd) return PFM_ERR_ARG;
#define PFM_CO(D, R0T, BT) \
    hipLaunchKernelGGL((k_coarse<D, ZT, R0T, BT>), dim3(1), dim3(kCoarseThreads), 0, s, a, (int)sub_mean, (int)dot, h->sc, h->partial)
    if constexpr (std::is_same<ZT, float>::value) {
        if (h->dim == 3) PFM_CO(3, float, float);
        else PFM_CO(2, float, float);
    } else {
        // 16-bit levels: the bottom solve runs in double
        if (lc == 0) {
            if (h->dim == 3) PFM_CO(3, float, double);
            else PFM_CO(2, float, double);
        } else {
            if (h->dim == 3) PFM_CO(3, ZT, double);
            else PFM_CO(2, ZT, double);
        }
    }
#undef PFM_CO
    PFM_LAUNCH_CHECK();
    return PFM_OK;
}

// Levels whose visits run as one fused down and one fused up launch: the
// levels below level 0 up to 2^18 cells, where the launches' fixed latency
// dominates, and level 0 of a 2D grid (whose tiles need a small halo box).
// 3D level 0 keeps the two-pass group sweeps, which are bandwidth bound.
constexpr int64_t kFuseMaxCells = (int64_t)1 << 18;
constexpr int64_t kFuseMaxCells2D0 = (int64_t)1 << 22;

static bool fuse_level(const MGHandle* h, int l) {
    if (l == 0) return h->dim == 2 && h->lv[0].n <= kFuseMaxCells2D0;
    return h->lv[l].n <= kFuseMaxCells;
}

template <typename T>
static TileMap fuse_tiles(const MGHandle* h, int l) {
    return tile_map(h->lv[l].d, T::I, T::J, T::K);
}

template <typename T, typename ZT>
static void fuse_down_t(MGHandle* h, int l, hipStream_t s) {
    const LvlDev L = lvl_dev(h, l), Lc = lvl_dev(h, l + 1);
    const TileMap tm = fuse_tiles<T>(h, l);
    const dim3 g((unsigned)tm.total);
    const float isc = restrict_inv_scale(h);
    // r is fp32 on level 0 (the PCG residual) and ZT below
    if (l == 0)
        hipLaunchKernelGGL((k_fuse_down<T, ZT, float>), g, dim3(kFuseThreads), 0, s, L, Lc, tm, isc,
                           (const float*)h->r[0], (ZT*)h->z2[0], (ZT*)h->r[1], h->sc);
    else
        hipLaunchKernelGGL((k_fuse_down<T, ZT, ZT>), g, dim3(kFuseThreads), 0, s, L, Lc, tm, isc,
                           (const ZT*)h->r[l], (ZT*)h->z2[l], (ZT*)h->r[l + 1], h->sc);
}

template <typename T, typename ZT>
static int fuse_up_t(MGHandle* h, int l, bool dot, hipStream_t s) {
    const LvlDev L = lvl_dev(h, l), Lc = lvl_dev(h, l + 1);
    const TileMap tm = fuse_tiles<T>(h, l);
    const dim3 g((unsigned)tm.total);
    const ZT* zi = (const ZT*)h->z2[l];
    const ZT* zc = (const ZT*)h->z[l + 1];
    ZT* zo = (ZT*)h->z[l];
    if (l == 0 && dot) {
        if (tm.total > h->n_partial) return PFM_ERR_ARG;  // sized at create
        hipLaunchKernelGGL((k_fuse_up<T, ZT, float, true>), g, dim3(kFuseThreads), 0, s, L, Lc, tm, zi, zc,
                           (const float*)h->r[0], zo, h->sc, h->partial);
    } else if (l == 0) {
        hipLaunchKernelGGL((k_fuse_up<T, ZT, float, false>), g, dim3(kFuseThreads), 0, s, L, Lc, tm, zi, zc,
                           (const float*)h->r[0], zo, h->sc, h->partial);
    } else {
        hipLaunchKernelGGL((k_fuse_up<T, ZT, ZT, false>), g, dim3(kFuseThreads), 0, s, L, Lc, tm, zi, zc,
                           (const ZT*)h->r[l], zo, h->sc, h->partial);
    }
    return PFM_OK;
}

// z2[l] <- pre-smoothed z of level l, r[l + 1] <- its restricted residual.
template <typename ZT>
static int fuse_down(MGHandle* h, int l, hipStream_t s) {
    if (h->dim == 2 && h->lv[l].n <= kFuseSmallCells2D) fuse_down_t<FsTile2S, ZT>(h, l, s);
    else if (h->dim == 2) fuse_down_t<FsTile2, ZT>(h, l, s);
    else if (h->lv[l].n <= kFuseSmallCells) fuse_down_t<FsTile3S, ZT>(h, l, s);
    else fuse_down_t<FsTile3, ZT>(h, l, s);
    PFM_LAUNCH_CHECK();
    return PFM_OK;
}

// z[l] <- post-smoothed z2[l] + z[l + 1][parent]; with dot (level 0) the z.r
// partials go to `partial`, their count to *nparts.
template <typename ZT>
static int fuse_up(MGHandle* h, int l, bool dot, int* nparts, hipStream_t s) {
    if (h->dim == 2 && h->lv[l].n <= kFuseSmallCells2D) {
        PFM_TRY((fuse_up_t<FsTile2S, ZT>(h, l, dot, s)));
        if (nparts) *nparts = fuse_tiles<FsTile2S>(h, l).total;
    } else if (h->dim == 2) {
        PFM_TRY((fuse_up_t<FsTile2, ZT>(h, l, dot, s)));
        if (nparts) *nparts = fuse_tiles<FsTile2>(h, l).total;
    } else if (h->lv[l].n <= kFuseSmallCells) {
        PFM_TRY((fuse_up_t<FsTile3S, ZT>(h, l, dot, s)));
        if (nparts) *nparts = fuse_tiles<FsTile3S>(h, l).total;
    } else {
        PFM_TRY((fuse_up_t<FsTile3, ZT>(h, l, dot, s)));
        if (nparts) *nparts = fuse_tiles<FsTile3>(h, l).total;
    }
    PFM_LAUNCH_CHECK();
    return PFM_OK;
}

// apply_preconditioner: 3D/mgpcg.py:182-200.  z[l] holds each level's final
// z afterwards (z2[l] is scratch).  Levels above coarse_level run as tiled
// launches, the rest in one k_coarse launch.  With sub_mean, r[0] is
// recentred by the V-cycle's own launches (the recentring of solve,
// :174-180); with dot_parts, the last level-0 launch leaves partials of z.r
// in `partial` and their count in *dot_parts (0 when no launch did).
template <typename ZT>
static int vcycle_t(MGHandle* h, hipStream_t s, bool sub_mean, int* dot_parts) {
    const int L = h->n_levels;
    const int lc = h->coarse_level;
    const int top = lc >= 0 ? lc : L - 1;  // levels [0, top) run as tiled launches
    const bool want_dot = dot_parts != nullptr;
    if (dot_parts) *dot_parts = 0;
    // a fused level 0 reads r with a halo, so r is recentred by its own launch first
    const bool tiled_mean = sub_mean && top > 0 && restrict_tiled_ok(h, 0) && !fuse_level(h, 0);
    const bool coarse_mean = sub_mean && lc == 0;
    if (sub_mean && !tiled_mean && !coarse_mean) {
        hipLaunchKernelGGL(k_sub_mean, dim3(kMgRedBlocks), dim3(kBlock), 0, s, h->lv[0].n, (float*)h->r[0], h->sc,
                           h->dg[0]);
        PFM_LAUNCH_CHECK();
    }
    for (int l = 0; l < top; ++l) {
        if (fuse_level(h, l)) {
            PFM_TRY(fuse_down<ZT>(h, l, s));
            continue;
        }
        const bool sm = tiled_mean && l == 0;
        PFM_TRY(rbgs<ZT>(h, l, kSweepZero, 0, nullptr, (ZT*)h->z2[l], sm, false, s));
        PFM_TRY(rbgs<ZT>(h, l, kSweepPlain, 0, (const ZT*)h->z2[l], (ZT*)h->z[l], sm, false, s));
        PFM_TRY(restrict_fused<ZT>(h, l, sm, s));
    }
    if (lc >= 0) {
        PFM_TRY(coarse<ZT>(h, coarse_mean, want_dot && lc == 0, s));
        if (want_dot && lc == 0) *dot_parts = 1;
    } else {
        PFM_TRY(bottom<ZT>(h, L - 1, s));
    }
    for (int l = top - 1; l >= 0; --l) {
        const bool dot = want_dot && l == 0;
        if (fuse_level(h, l)) {
            PFM_TRY(fuse_up<ZT>(h, l, dot, dot ? dot_parts : nullptr, s));
            continue;
        }
        PFM_TRY(rbgs<ZT>(h, l, kSweepProlong, 1, (const ZT*)h->z[l], (ZT*)h->z2[l], false, false, s));
        PFM_TRY(rbgs<ZT>(h, l, kSweepPlain, 1, (const ZT*)h->z2[l], (ZT*)h->z[l], false, dot, s));
        if (dot) *dot_parts = sweep_tiles(h, 0).total;
    }
    return PFM_OK;
}

static int vcycle(MGHandle* h, hipStream_t s, bool sub_mean, int* dot_parts) {
    switch (h->precision) {
        case 1: return vcycle_t<__half>(h, s, sub_mean, dot_parts);
        case 2: return vcycle_t<__hip_bfloat16>(h, s, sub_mean, dot_parts);
        default: return vcycle_t<float>(h, s, sub_mean, dot_parts);
    }
}

static int finalize(MGHandle* h, int mode, double eps, hipStream_t s, int nparts = kMgRedBlocks, int test = 0) {
    hipLaunchKernelGGL(k_finalize, dim3(1), dim3(kFinThreads), 0, s, h->partial, nparts, mode, eps,
                       (double)h->n_fluid, h->vshift, test, h->sc);
    PFM_LAUNCH_CHECK();
    return PFM_OK;
}

template <typename AT>
static int dot_t(MGHandle* h, const AT* a, const float* b, hipStream_t s, double* out,
                 const float* ascale = nullptr) {
    hipLaunchKernelGGL(k_dot<AT>, dim3(kMgRedBlocks), dim3(kBlock), 0, s, h->lv[0].n, a, b, out, ascale);
    PFM_LAUNCH_CHECK();
    return PFM_OK;
}

static int dot(MGHandle* h, const float* a, const float* b, hipStream_t s) { return dot_t(h, a, b, s, h->partial); }

// z[0] . r[0] for whatever z[0] stores
static int dot_zr(MGHandle* h, hipStream_t s) {
    switch (h->precision) {
        case 1: return dot_t(h, (const __half*)h->z[0], (const float*)h->r[0], s, h->partial, &h->sc->vinv);
        case 2: return dot_t(h, (const __hip_bfloat16*)h->z[0], (const float*)h->r[0], s, h->partial, &h->sc->vinv);
        default: return dot_t(h, (const float*)h->z[0], (const float*)h->r[0], s, h->partial);
    }
}

// p <- z + beta p (into the partner buffer), Ap and p.Ap partials; returns
// the partial count through *parts.
// The Ap launch's partials follow the three blocks of k_update_xr's, which
// reads them while writing its own.
constexpr int kApPartOff = 3 * kMgRedBlocks;

static int ap_dot_p(MGHandle* h, hipStream_t s, int* parts) {
    const LvlDev L0 = lvl_dev(h, 0);
    const TileMap tm = sweep_tiles(h, 0);
    const dim3 g((unsigned)tm.total);
    // 3D levels with an axis-2 extent that is a multiple of 4 run the group kernel
    const bool groups = h->dim == 3 && L0.d[2] % 4 == 0;
#define PFM_AP(D, ZT)                                                                                           \
    if (D == 3 && groups)                                                                                       \
        hipLaunchKernelGGL((k_ap_dot_p3<ZT>), g, dim3(kSw3Threads), 0, s, L0, tm, (const ZT*)h->z[0], h->p, h->p2, \
                           h->Ap, h->sc, h->partial + kApPartOff);                                              \
    else                                                                                                        \
        hipLaunchKernelGGL((k_ap_dot_p<D, ZT>), g, dim3(kBlock), 0, s, L0, tm, (const ZT*)h->z[0], h->p, h->p2,  \
                           h->Ap, h->sc, h->partial + kApPartOff)
    switch (h->precision) {
        case 1:
            if (h->dim == 3) PFM_AP(3, __half);
            else PFM_AP(2, __half);
            break;
        case 2:
            if (h->dim == 3) PFM_AP(3, __hip_bfloat16);
            else PFM_AP(2, __hip_bfloat16);
            break;
        default:
            if (h->dim == 3) PFM_AP(3, float);
            else PFM_AP(2, float);
    }
#undef PFM_AP
    PFM_LAUNCH_CHECK();
    std::swap(h->p, h->p2);
    *parts = tm.total;
    return PFM_OK;
}

static int update_p(MGHandle* h, hipStream_t s) {
    const int64_t n = h->lv[0].n;
    switch (h->precision) {
        case 1:
            hipLaunchKernelGGL(k_update_p<__half>, dim3(kMgRedBlocks), dim3(kBlock), 0, s, n, h->p,
                               (const __half*)h->z[0], h->sc);
            break;
        case 2:
            hipLaunchKernelGGL(k_update_p<__hip_bfloat16>, dim3(kMgRedBlocks), dim3(kBlock), 0, s, n, h->p,
                               (const __hip_bfloat16*)h->z[0], h->sc);
            break;
        default:
            hipLaunchKernelGGL(k_update_p<float>, dim3(kMgRedBlocks), dim3(kBlock), 0, s, n, h->p,
                               (const float*)h->z[0], h->sc);
    }
    PFM_LAUNCH_CHECK();
    return PFM_OK;
}

static bool all_neumann(const MGHandle* h) {
    for (int a = 0; a < h->dim; ++a)
        if (h->bc[a][0] != 2 || h->bc[a][1] != 2) return false;
    return true;
}

// V-cycle followed by z.r into the partials; returns the partial count.
static int precondition_dot(MGHandle* h, bool sub_mean, hipStream_t s, int* parts) {
    PFM_TRY(vcycle(h, s, sub_mean, parts));
    if (*parts == 0) {
        PFM_TRY(dot_zr(h, s));
        *parts = kMgRedBlocks;
    }
    return PFM_OK;
}

// Iterations between true-residual replacements (r = b - A x) in the
// mixed-precision solver.
constexpr int kReplaceEvery = 8;
// A replacement is made only while |r| exceeds kFloorMargin times the
// residual that the fp32 representation of x alone produces,
// 2 dim 2^-24 |x| (each stencil term of A x carries a rounding of x).  Below
// that the true residual is that representation noise, which no fp32
// correction of x can remove, and replacing r with it would stall the
// iteration; the recursive residual, which the fp32 solver also relies on,
// is then kept.
constexpr double kFloorMargin = 16.0;

// solve: 3D/mgpcg.py:202-280.  The reference's update_p after each
// preconditioner application is carried out at the start of the next Ap
// launch (same arithmetic, same beta), or explicitly when the loop ends
// without converging.  At precision 1/2 (16-bit V-cycle storage) the PCG
// residual is replaced by the true residual b - A x every kReplaceEvery
// iterations.  Convergence is tested on the device (k_finalize): the host
// reads the outcome once per group of iterations.

// Host copy of the device outcome of a solve: r.r, x.x, iterations, flags.
struct SolveOutcome {
    double rtr, xtx;
    int it, done, conv;
};

static int read_outcome(MGHandle* h, hipStream_t s, SolveOutcome* o) {
    static_assert(offsetof(MGScalars, xTx) == offsetof(MGScalars, rTr) + sizeof(double) &&
                      offsetof(MGScalars, it) == offsetof(MGScalars, rTr) + 2 * sizeof(double) &&
                      offsetof(MGScalars, conv) == offsetof(MGScalars, rTr) + 2 * sizeof(double) + 2 * sizeof(int),
                  "the outcome fields are contiguous");
    PFM_HIP_TRY(hipMemcpyAsync(h->host_rtr, &h->sc->rTr, 2 * sizeof(double) + 3 * sizeof(int), hipMemcpyDeviceToHost,
                               s));
    PFM_HIP_TRY(hipStreamSynchronize(s));
    const int* iv = (const int*)(h->host_rtr + 2);
    *o = SolveOutcome{h->host_rtr[0], h->host_rtr[1], iv[0], iv[1], iv[2]};
    return PFM_OK;
}

// One PCG iteration (3D/mgpcg.py:236-275) as launches only.  Head: Ap and
// alpha, the x / r update and the convergence test on the device, every
// kReplaceEvery-th iteration of the mixed-precision solver the true-residual
// replacement (taken on the device when r.r is above the representation
// floor).  Tail: the V-cycle, z.r and beta.  `k` is the iteration's index.
static int pcg_head(MGHandle* h, int k, double eps, hipStream_t s) {
    const int64_t n = h->lv[0].n;
    float* r0 = (float*)h->r[0];
    const bool replace = h->precision != 0;
    int parts = 0;
    PFM_TRY(ap_dot_p(h, s, &parts));
    const double* ap_part = h->partial + kApPartOff;
    if (replace)
        hipLaunchKernelGGL(k_update_xr<true>, dim3(kMgRedBlocks), dim3(kBlock), 0, s, n, h->x, r0, h->p, h->Ap,
                           h->sc, ap_part, parts, eps, h->partial);
    else
        hipLaunchKernelGGL(k_update_xr<false>, dim3(kMgRedBlocks), dim3(kBlock), 0, s, n, h->x, r0, h->p, h->Ap,
                           h->sc, ap_part, parts, eps, h->partial);
    PFM_LAUNCH_CHECK();
    const bool due = replace && (k + 1) % kReplaceEvery == 0;
    PFM_TRY(finalize(h, replace ? kFinRTrX : kFinRTr, eps, s, kMgRedBlocks, due ? 0 : 1));
    if (due) {
        const LvlDev L0 = lvl_dev(h, 0);
        const double floor_c = kFloorMargin * 2.0 * h->dim * 0x1p-24;
        hipLaunchKernelGGL(k_true_residual, cell_grid(L0.d, false), kCellBlock, 0, s, L0, h->b, h->x, r0, h->sc,
                           floor_c * floor_c);
        PFM_LAUNCH_CHECK();
        PFM_TRY(dot_t(h, r0, r0, s, h->partial));
        PFM_TRY(dot_t(h, r0, (const float*)nullptr, s, h->partial + kMgRedBlocks));
        PFM_TRY(finalize(h, kFinRTr, eps, s, kMgRedBlocks, 1));
    }
    return PFM_OK;
}

static int pcg_tail(MGHandle* h, double eps, bool neumann, hipStream_t s) {
    int parts = 0;
    PFM_TRY(precondition_dot(h, neumann, s, &parts));
    PFM_TRY(finalize(h, kFinBeta, eps, s, parts));
    return PFM_OK;
}

static int pcg_iteration(MGHandle* h, int k, double eps, bool neumann, hipStream_t s) {
    PFM_TRY(pcg_head(h, k, eps, s));
    return pcg_tail(h, eps, neumann, s);
}

// The iteration from a HIP graph: captured once per (p-buffer parity,
// replacement step) on the handle's own non-blocking stream in thread-local
// capture mode (other threads' HIP calls do not disturb it), then replayed on
// `s`.  The launches hold no memset, so replays match eager runs bit for bit.
// Returns false (and the caller runs the iteration eagerly) when capture is
// not possible.
static bool pcg_iteration_graph(MGHandle* h, int k, double eps, bool neumann, hipStream_t s) {
    if (!h->use_graphs || h->graph_failed) return false;
    if (h->graph_eps != eps) {
        drop_graphs(h);
        h->graph_eps = eps;
    }
    const int par = h->p == h->p_base ? 0 : 1;
    const int due = h->precision != 0 && (k + 1) % kReplaceEvery == 0;
    hipGraphExec_t& ex = h->iter_graph[par][due];
    if (!ex) {
        if (!h->cap_stream && hipStreamCreateWithFlags(&h->cap_stream, hipStreamNonBlocking) != hipSuccess) {
            h->cap_stream = nullptr;
            h->graph_failed = 1;
            return false;
        }
        float* p0 = h->p;
        float* p1 = h->p2;
        hipGraph_t g = nullptr;
        bool ok = hipStreamBeginCapture(h->cap_stream, hipStreamCaptureModeThreadLocal) == hipSuccess;
        if (ok) {
            const int rc = pcg_iteration(h, k, eps, neumann, h->cap_stream);
            const bool ended = hipStreamEndCapture(h->cap_stream, &g) == hipSuccess;
            ok = rc == PFM_OK && ended && g != nullptr;
        }
        // capturing launched nothing: the buffers are as before
        h->p = p0;
        h->p2 = p1;
        if (ok && hipGraphInstantiate(&ex, g, nullptr, nullptr, 0) != hipSuccess) {
            ex = nullptr;
            ok = false;
        }
        if (g) (void)hipGraphDestroy(g);
        (void)hipGetLastError();
        if (!ok) {
            h->graph_failed = 1;
            return false;
        }
    }
    if (hipGraphLaunch(ex, s) != hipSuccess) {
        (void)hipGetLastError();
        h->graph_failed = 1;
        return false;
    }
    std::swap(h->p, h->p2);  // the iteration's ap_dot_p wrote the partner buffer
    return true;
}

static int solve(MGHandle* h, int max_iters, double eps, double tol, int tol_mode, int* iters_out,
                 double* rtr_out, hipStream_t s) {
    const int64_t n = h->lv[0].n;
    float* r0 = (float*)h->r[0];
    const bool neumann = all_neumann(h);
    hipLaunchKernelGGL(k_reset_scalars, dim3(1), dim3(1), 0, s, h->sc, tol, tol_mode, max_iters);
    PFM_LAUNCH_CHECK();
    if (neumann) {
        PFM_TRY(dot(h, r0, nullptr, s));
        PFM_TRY(finalize(h, kFinMean, eps, s));
        hipLaunchKernelGGL(k_sub_mean, dim3(kMgRedBlocks), dim3(kBlock), 0, s, n, r0, h->sc, h->dg[0]);
        PFM_LAUNCH_CHECK();
    }
    const bool replace = h->precision != 0;
    if (replace) PFM_HIP_TRY(hipMemcpyAsync(h->b, r0, sizeof(float) * n, hipMemcpyDeviceToDevice, s));
    if (tol_mode == 1 || replace) {  // r0 (the relative threshold, and the V-cycle scale of 16-bit storage)
        PFM_TRY(dot(h, r0, r0, s));
        PFM_TRY(finalize(h, kFinR0, eps, s));
    }
    int parts = 0;
    PFM_TRY(precondition_dot(h, false, s, &parts));
    PFM_TRY(finalize(h, kFinZTr, eps, s, parts));
    int enq = 0;
    SolveOutcome o{0.0, 0.0, 0, 0, 0};
    if (!h->use_graphs || h->graph_failed) {
        // The reference's loop shape: the host reads the convergence test of
        // each iteration before issuing its V-cycle.
        for (; max_iters == -1 || enq < max_iters; ++enq) {
            PFM_TRY(pcg_head(h, enq, eps, s));
            PFM_TRY(read_outcome(h, s, &o));
            if (o.done) {
                ++enq;
                break;
            }
            PFM_TRY(pcg_tail(h, eps, neumann, s));
        }
    } else {
        // Whole iterations replayed from graphs in groups, each ending with one
        // read of the outcome: the first group as long as the shorter of the
        // previous two solves (so it rarely runs past convergence), then one at
        // a time.  Enqueued iterations past convergence change nothing the
        // solve returns.
        const int predicted = std::max(1, std::min(h->last_iters[0], h->last_iters[1]));
        while (max_iters == -1 || enq < max_iters) {
            int g = enq == 0 ? predicted : 1;
            if (max_iters != -1) g = std::min(g, max_iters - enq);
            for (int j = 0; j < g; ++j, ++enq)
                if (!pcg_iteration_graph(h, enq, eps, neumann, s)) PFM_TRY(pcg_iteration(h, enq, eps, neumann, s));
            PFM_TRY(read_outcome(h, s, &o));
            if (o.done) break;
        }
    }
    if (enq == 0 || !o.done) PFM_TRY(read_outcome(h, s, &o));
    // Every enqueued iteration swapped p and its partner on the host; the
    // device ran ap_dot_p only until the solve stopped.
    const int ran = o.conv ? o.it + 1 : o.it;
    if ((enq - ran) & 1) std::swap(h->p, h->p2);
    h->last_iters[1] = h->last_iters[0];
    h->last_iters[0] = ran;
    hipLaunchKernelGGL(k_clear_stop, dim3(1), dim3(1), 0, s, h->sc);
    PFM_LAUNCH_CHECK();
    if (!o.conv) PFM_TRY(update_p(h, s));
    if (iters_out) *iters_out = o.it;
    if (rtr_out) *rtr_out = o.rtr;
    return o.conv ? PFM_OK : PFM_WARN_NOT_CONVERGED;
}

static int apply_bc(MGHandle* h, float* const* u, hipStream_t s) {
    BcPlanes bp{};
    auto add = [&](float* f, const int S[3], int pa, int idx) {
        const int k = bp.n++;
        bp.u[k] = f;
        for (int c = 0; c < 3; ++c) bp.S[k][c] = S[c];
        bp.plane_axis[k] = pa;
        bp.idx[k] = idx;
        const int a1 = pa == 0 ? 1 : 0, a2 = pa == 2 ? 1 : 2;
        bp.off[k + 1] = bp.off[k] + (int64_t)S[a1] * S[a2];
    };
    for (int a = 0; a < h->dim; ++a) {
        int S[3] = {h->N[0], h->N[1], h->N[2]};
        S[a] += 1;
        // normal faces on Neumann walls (and on Dirichlet walls in 2D)
        for (int side = 0; side < 2; ++side) {
            int bt = h->bc[a][side];
            bool zero = bt == 2 || (h->dim == 2 && bt == 1);
            if (zero) add(u[a], S, a, side == 0 ? 0 : S[a] - 1);
        }
        // 2D Dirichlet walls also zero the tangential faces on them
        if (h->dim == 2) {
            for (int b = 0; b < 2; ++b) {
                if (b == a) continue;
                for (int side = 0; side < 2; ++side)
                    if (h->bc[b][side] == 1) add(u[a], S, b, side == 0 ? 0 : S[b] - 1);
            }
        }
    }
    if (bp.n > 0) {
        hipLaunchKernelGGL(k_zero_planes, dim3(grid_for(bp.off[bp.n])), dim3(kBlock), 0, s, bp);
        PFM_LAUNCH_CHECK();
    }
    // free slip at the obstacle (an extension: the reference has no solids)
    if (h->dg[0]) {
        const LvlDev L0 = lvl_dev(h, 0);
        for (int a = 0; a < h->dim; ++a) {
            const int m = map_axis(h, a);
            LvlDev F = L0;
            F.d[m] += 1;
            hipLaunchKernelGGL(k_zero_solid_faces, cell_grid(F.d, false), kCellBlock, 0, s, L0, F, m, u[a]);
            PFM_LAUNCH_CHECK();
        }
    }
    return PFM_OK;
}

static int div_init(MGHandle* h, const float* const* u, float k, hipStream_t s) {
    LvlDev L0 = lvl_dev(h, 0);
    const dim3 gd = cell_grid(L0.d, false);
    if (h->dim == 3)
        hipLaunchKernelGGL(k_div_init<3>, gd, kCellBlock, 0, s, L0, u[0], u[1], u[2], k, (float*)h->r[0], h->precision == 0 ? (float*)h->z[0] : nullptr,
                           h->Ap, h->p, h->x);
    else
        hipLaunchKernelGGL(k_div_init<2>, gd, kCellBlock, 0, s, L0, u[0], u[1], u[2], k, (float*)h->r[0], h->precision == 0 ? (float*)h->z[0] : nullptr,
                           h->Ap, h->p, h->x);
    PFM_LAUNCH_CHECK();
    return PFM_OK;
}

static int subtract_grad_p(MGHandle* h, float* const* u, hipStream_t s) {
    LvlDev L0 = lvl_dev(h, 0);
    for (int a = 0; a < h->dim; ++a) {
        const int m = map_axis(h, a);
        LvlDev F = L0;
        F.d[m] += 1;
        hipLaunchKernelGGL(k_subtract_grad_p, cell_grid(F.d, false), kCellBlock, 0, s, L0, F, m, h->x, u[a]);
        PFM_LAUNCH_CHECK();
    }
    return PFM_OK;
}

static int check_u(const MGHandle* h, const float* ux, const float* uy, const float* uz) {
    PFM_CHECK_ARG(h, "null solver handle");
    PFM_CHECK_ARG(ux && uy && (h->dim == 2 || uz), "null velocity field");
    return PFM_OK;
}

}  // namespace pfm

using namespace pfm;

extern "C" {

int pfm_mg_create(int dim, const int32_t* N, int base_level, const int32_t* bc, int precision,
                  void** handle) {
    PFM_CHECK_ARG(handle, "null handle out-pointer");
    *handle = nullptr;
    PFM_CHECK_ARG(dim == 2 || dim == 3, "dim must be 2 or 3");
    PFM_CHECK_ARG(N && bc, "null N or bc");
    PFM_CHECK_ARG(precision >= 0 && precision <= 2, "precision %d: 0 = fp32, 1 = fp16, 2 = bf16 level storage",
                  precision);
    int mn = N[0];
    for (int a = 0; a < dim; ++a) {
        PFM_CHECK_ARG(N[a] >= 2 && N[a] <= 65535, "N[%d]=%d out of range [2, 65535]", a, N[a]);
        if (N[a] < mn) mn = N[a];
    }
    int64_t cells = 1;
    for (int a = 0; a < dim; ++a) cells *= N[a];
    // level arrays are addressed with 32-bit byte offsets, and an fp32 level
    // stays below the out-of-range offset of the tiled loads
    PFM_CHECK_ARG(cells <= ((int64_t)1 << 30) - 64, "grid too large for one solver (%lld cells, limit 2^30 - 64)",
                  (long long)cells);
    MGHandle* h = new (std::nothrow) MGHandle();
    if (!h) {
        set_error("host allocation failed");
        return PFM_ERR_OOM;
    }
    h->dim = dim;
    for (int a = 0; a < 3; ++a) h->N[a] = a < dim ? N[a] : 1;
    // n_mg_levels = int(log2(min N)) - base_level + 1 (3D/mgpcg.py:28)
    int lg = 0;
    while ((2 << lg) <= mn) ++lg;
    int L = lg - base_level + 1;
    if (L < 1) L = 1;
    if (L > kMaxLevels) L = kMaxLevels;
    h->n_levels = L;
    h->dirichlet_counts = dim == 2;  // 2D/mgpcg.py:111-113
    h->precision = precision;
    {
        // The V-cycle output z grows like r / lambda_min with lambda_min ~
        // dim pi^2 / N^2 for the smoothest mode of the undivided Laplacian:
        // with fp16 storage, scale the input down so that rms(r) * N^2 /
        // (dim pi^2) stays below 2^10, keeping 2^5 of headroom for peaks.
        int nmax = 1;
        for (int a = 0; a < dim; ++a) nmax = N[a] > nmax ? N[a] : nmax;
        const double growth = (double)nmax * nmax / (dim * M_PI * M_PI);
        h->vshift = precision == 1 ? (int)fmax(0.0, ceil(log2(growth)) - 10.0) : 0;
    }
    for (int a = 0; a < 3; ++a) {
        h->bc[a][0] = a < dim ? bc[2 * a] : 0;
        h->bc[a][1] = a < dim ? bc[2 * a + 1] : 0;
    }
    for (int l = 0; l < L; ++l) {
        int64_t n = 1;
        for (int m = 0; m < 3; ++m) {
            const int a = m - (3 - dim);  // problem axis held by index axis m
            int e = a >= 0 ? h->N[a] >> l : 1;
            if (e < 1) e = 1;
            h->lv[l].d[m] = e;
            n *= e;
        }
        h->lv[l].n = n;
    }
    h->coarse_level = pick_coarse_level(h);
    h->n_fluid = h->lv[0].n;
    auto fail = [&](const char* what) {
        set_error("hipMalloc failed for %s", what);
        pfm_mg_destroy(h);
        return PFM_ERR_OOM;
    };
    const size_t esz = precision == 0 ? sizeof(float) : 2;  // z, z2, r[l >= 1]
    for (int l = 0; l < L; ++l) {
        const size_t b = esz * h->lv[l].n;
        if (hipMalloc(&h->r[l], l == 0 ? sizeof(float) * h->lv[l].n : b) != hipSuccess) return fail("r");
        if (hipMalloc(&h->z[l], b) != hipSuccess) return fail("z");
        if (hipMalloc(&h->z2[l], b) != hipSuccess) return fail("z scratch");
    }
    if (precision != 0 && hipMalloc(&h->b, sizeof(float) * h->lv[0].n) != hipSuccess) return fail("rhs copy");
    const size_t b0 = sizeof(float) * h->lv[0].n;
    if (hipMalloc(&h->x, b0) != hipSuccess) return fail("x");
    if (hipMalloc(&h->p, b0) != hipSuccess) return fail("p");
    h->p_base = h->p;
    // graph replay is opt-in: measured on the served job stream, capturing and
    // instantiating per job cost more than the launches it saves
    h->use_graphs = 0;
    if (hipMalloc(&h->p2, b0) != hipSuccess) return fail("p scratch");
    if (hipMalloc(&h->Ap, b0) != hipSuccess) return fail("Ap");
    h->n_partial = std::max(3 * kMgRedBlocks, sweep_tiles(h, 0).total);
    h->n_partial = std::max(h->n_partial, kApPartOff + sweep_tiles(h, 0).total);
    if (fuse_level(h, 0))
        h->n_partial = std::max(h->n_partial, h->lv[0].n <= kFuseSmallCells2D ? fuse_tiles<FsTile2S>(h, 0).total
                                                                               : fuse_tiles<FsTile2>(h, 0).total);
    if (hipMalloc(&h->partial, sizeof(double) * h->n_partial) != hipSuccess) return fail("partial");
    if (hipMalloc(&h->sc, sizeof(MGScalars)) != hipSuccess) return fail("scalars");
    if (hipHostMalloc(&h->host_rtr, sizeof(double) * 4, hipHostMallocDefault) != hipSuccess)
        return fail("pinned scalars");
    // The handle's own non-blocking stream (also the graph-capture stream)
    // clears the arrays: nothing here touches the legacy stream, which other
    // threads' graph captures forbid.
    if (hipStreamCreateWithFlags(&h->cap_stream, hipStreamNonBlocking) != hipSuccess) {
        h->cap_stream = nullptr;
        return fail("stream");
    }
    hipStream_t cs = h->cap_stream;
    PFM_HIP_TRY(hipMemsetAsync(h->sc, 0, sizeof(MGScalars), cs));
    for (int l = 0; l < L; ++l) {
        const size_t b = esz * h->lv[l].n;
        PFM_HIP_TRY(hipMemsetAsync(h->r[l], 0, l == 0 ? sizeof(float) * h->lv[l].n : b, cs));
        PFM_HIP_TRY(hipMemsetAsync(h->z[l], 0, b, cs));
        PFM_HIP_TRY(hipMemsetAsync(h->z2[l], 0, b, cs));
    }
    PFM_HIP_TRY(hipMemsetAsync(h->x, 0, b0, cs));
    PFM_HIP_TRY(hipMemsetAsync(h->p, 0, b0, cs));
    PFM_HIP_TRY(hipMemsetAsync(h->p2, 0, b0, cs));
    PFM_HIP_TRY(hipMemsetAsync(h->Ap, 0, b0, cs));
    PFM_HIP_TRY(hipStreamSynchronize(cs));
    *handle = h;
    return PFM_OK;
}

int pfm_mg_destroy(void* hv) {
    MGHandle* h = (MGHandle*)hv;
    if (!h) return PFM_OK;
    drop_graphs(h);
    if (h->cap_stream) (void)hipStreamDestroy(h->cap_stream);
    for (int l = 0; l < kMaxLevels; ++l) {
        if (h->r[l]) (void)hipFree(h->r[l]);
        if (h->z[l]) (void)hipFree(h->z[l]);
        if (h->z2[l]) (void)hipFree(h->z2[l]);
        if (h->solid[l]) (void)hipFree(h->solid[l]);
        if (h->dg[l]) (void)hipFree(h->dg[l]);
    }
    if (h->b) (void)hipFree(h->b);
    if (h->x) (void)hipFree(h->x);
    if (h->p) (void)hipFree(h->p);
    if (h->p2) (void)hipFree(h->p2);
    if (h->Ap) (void)hipFree(h->Ap);
    if (h->partial) (void)hipFree(h->partial);
    if (h->sc) (void)hipFree(h->sc);
    if (h->host_rtr) (void)hipHostFree(h->host_rtr);
    delete h;
    return PFM_OK;
}

int pfm_mg_n_levels(void* hv) { return hv ? ((MGHandle*)hv)->n_levels : 0; }

int pfm_mg_set_graphs(void* hv, int enable) {
    MGHandle* h = (MGHandle*)hv;
    PFM_CHECK_ARG(h, "null solver handle");
    h->use_graphs = enable != 0;
    h->graph_failed = 0;
    if (!enable) drop_graphs(h);
    return PFM_OK;
}

int pfm_mg_set_solid(void* hv, const uint8_t* solid, pfm_stream_t stream) {
    MGHandle* h = (MGHandle*)hv;
    PFM_CHECK_ARG(h, "null solver handle");
    hipStream_t s = as_stream(stream);
    drop_graphs(h);  // the captured launches hold the old masks and fluid count
    if (!solid) {
        PFM_HIP_TRY(hipStreamSynchronize(s));
        for (int l = 0; l < kMaxLevels; ++l) {
            if (h->solid[l]) (void)hipFree(h->solid[l]);
            if (h->dg[l]) (void)hipFree(h->dg[l]);
            h->solid[l] = h->dg[l] = nullptr;
        }
        h->n_fluid = h->lv[0].n;
        return PFM_OK;
    }
    for (int l = 0; l < h->n_levels; ++l) {
        const size_t b = (size_t)h->lv[l].n;
        if (!h->solid[l] && hipMalloc(&h->solid[l], b) != hipSuccess) {
            set_error("hipMalloc failed for the solid mask");
            return PFM_ERR_OOM;
        }
        if (!h->dg[l] && hipMalloc(&h->dg[l], b) != hipSuccess) {
            set_error("hipMalloc failed for the solid diagonal");
            return PFM_ERR_OOM;
        }
    }
    const int64_t n0 = h->lv[0].n;
    hipLaunchKernelGGL(k_mask01, dim3(grid_for(n0)), dim3(kBlock), 0, s, n0, solid, h->solid[0]);
    PFM_LAUNCH_CHECK();
    for (int l = 0; l + 1 < h->n_levels; ++l) {
        LvlDev Lf = lvl_dev(h, l), Lc = lvl_dev(h, l + 1);
        hipLaunchKernelGGL(k_solid_coarsen, cell_grid(Lc.d, false), kCellBlock, 0, s, Lf, Lc, h->solid[l],
                           h->solid[l + 1]);
        PFM_LAUNCH_CHECK();
    }
    for (int l = 0; l < h->n_levels; ++l) {
        LvlDev L = lvl_dev(h, l);
        L.dg = nullptr;
        hipLaunchKernelGGL(k_solid_diag, cell_grid(L.d, false), kCellBlock, 0, s, L, h->solid[l], h->dg[l]);
        PFM_LAUNCH_CHECK();
    }
    // fluid cells of the finest level (the recentring mean divides by them)
    uint8_t* host = (uint8_t*)malloc((size_t)n0);
    if (!host) {
        set_error("host allocation failed");
        return PFM_ERR_OOM;
    }
    hipError_t e = hipMemcpyAsync(host, h->solid[0], (size_t)n0, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) {
        free(host);
        set_error("HIP error %s reading the solid mask", hipGetErrorString(e));
        return PFM_ERR_HIP;
    }
    int64_t nf = 0;
    for (int64_t q = 0; q < n0; ++q) nf += host[q] == 0;
    free(host);
    PFM_CHECK_ARG(nf > 0, "the obstacle fills the whole domain");
    h->n_fluid = nf;
    return PFM_OK;
}

void* pfm_mg_field(void* hv, int which, int level) {
    MGHandle* h = (MGHandle*)hv;
    if (!h || level < 0 || level >= h->n_levels) return nullptr;
    // 16-bit level arrays have no fp32 view
    if (h->precision != 0 && (which == 1 || (which == 0 && level > 0))) return nullptr;
    switch (which) {
        case 0: return h->r[level];
        case 1: return h->z[level];
        case 2: return level == 0 ? h->x : nullptr;
        case 3: return level == 0 ? h->p : nullptr;
        case 4: return level == 0 ? h->Ap : nullptr;
        case 6: return level == 0 ? h->x : nullptr;  // get_result(self.p) copies x
        default: return nullptr;
    }
}

int pfm_mg_apply_bc(void* hv, float* u_x, float* u_y, float* u_z, pfm_stream_t stream) {
    MGHandle* h = (MGHandle*)hv;
    PFM_TRY(check_u(h, u_x, u_y, u_z));
    float* u[3] = {u_x, u_y, u_z};
    return apply_bc(h, u, as_stream(stream));
}

int pfm_mg_divergence_init(void* hv, const float* u_x, const float* u_y, const float* u_z, float k,
                           pfm_stream_t stream) {
    MGHandle* h = (MGHandle*)hv;
    PFM_TRY(check_u(h, u_x, u_y, u_z));
    const float* u[3] = {u_x, u_y, u_z};
    return div_init(h, u, k, as_stream(stream));
}

int pfm_mg_init(void* hv, const float* rhs, float k, pfm_stream_t stream) {
    MGHandle* h = (MGHandle*)hv;
    PFM_CHECK_ARG(h && rhs, "null argument");
    const int64_t n = h->lv[0].n;
    hipLaunchKernelGGL(k_init_rhs, dim3(grid_for(n)), dim3(kBlock), 0, as_stream(stream), n, rhs, k,
                       (float*)h->r[0], h->precision == 0 ? (float*)h->z[0] : nullptr, h->Ap, h->p, h->x, h->dg[0]);
    PFM_LAUNCH_CHECK();
    return PFM_OK;
}

int pfm_mg_solve(void* hv, int max_iters, double eps, double tol, int tol_mode, int* iters_out,
                 double* rTr_out, pfm_stream_t stream) {
    MGHandle* h = (MGHandle*)hv;
    PFM_CHECK_ARG(h, "null solver handle");
    PFM_CHECK_ARG(tol_mode == 0 || tol_mode == 1, "tol_mode must be 0 or 1");
    return solve(h, max_iters, eps, tol, tol_mode, iters_out, rTr_out, as_stream(stream));
}

int pfm_mg_subtract_grad_p(void* hv, float* u_x, float* u_y, float* u_z, pfm_stream_t stream) {
    MGHandle* h = (MGHandle*)hv;
    PFM_TRY(check_u(h, u_x, u_y, u_z));
    float* u[3] = {u_x, u_y, u_z};
    return subtract_grad_p(h, u, as_stream(stream));
}

int pfm_mg_poisson(void* hv, float* u_x, float* u_y, float* u_z, int max_iters, double tol,
                   int tol_mode, int* iters_out, double* rTr_out, pfm_stream_t stream) {
    MGHandle* h = (MGHandle*)hv;
    PFM_TRY(check_u(h, u_x, u_y, u_z));
    PFM_CHECK_ARG(tol_mode == 0 || tol_mode == 1, "tol_mode must be 0 or 1");
    hipStream_t s = as_stream(stream);
    float* u[3] = {u_x, u_y, u_z};
    const float* uc[3] = {u_x, u_y, u_z};
    PFM_TRY(apply_bc(h, u, s));
    PFM_TRY(div_init(h, uc, -1.0f, s));
    int rc = solve(h, max_iters, 1e-12, tol, tol_mode, iters_out, rTr_out, s);
    if (rc < 0) return rc;
    PFM_TRY(subtract_grad_p(h, u, s));
    PFM_TRY(apply_bc(h, u, s));
    return rc;
}

int pfm_mg_vcycle(void* hv, pfm_stream_t stream) {
    MGHandle* h = (MGHandle*)hv;
    PFM_CHECK_ARG(h, "null solver handle");
    return vcycle(h, as_stream(stream), false, nullptr);
}

int pfm_mg_smooth(void* hv, int level, int phase, pfm_stream_t stream) {
    MGHandle* h = (MGHandle*)hv;
    PFM_CHECK_ARG(h && level >= 0 && level < h->n_levels && (phase == 0 || phase == 1), "bad arguments");
    PFM_CHECK_ARG(h->precision == 0, "per-pass entry points need precision 0 (fp32 levels)");
    return smooth(h, level, phase, as_stream(stream));
}

int pfm_mg_restrict(void* hv, int level, pfm_stream_t stream) {
    MGHandle* h = (MGHandle*)hv;
    PFM_CHECK_ARG(h && level >= 0 && level < h->n_levels - 1, "bad level");
    PFM_CHECK_ARG(h->precision == 0, "per-pass entry points need precision 0 (fp32 levels)");
    return restrict_level(h, level, as_stream(stream));
}

int pfm_mg_prolongate(void* hv, int level, pfm_stream_t stream) {
    MGHandle* h = (MGHandle*)hv;
    PFM_CHECK_ARG(h && level >= 0 && level < h->n_levels - 1, "bad level");
    PFM_CHECK_ARG(h->precision == 0, "per-pass entry points need precision 0 (fp32 levels)");
    return prolongate_level(h, level, as_stream(stream));
}

int pfm_mg_apply_A(void* hv, const float* p, float* Ap, pfm_stream_t stream) {
    MGHandle* h = (MGHandle*)hv;
    PFM_CHECK_ARG(h && p && Ap, "null argument");
    LvlDev L0 = lvl_dev(h, 0);
    hipLaunchKernelGGL(k_apply_A, cell_grid(L0.d, false), kCellBlock, 0, as_stream(stream), L0, p, Ap);
    PFM_LAUNCH_CHECK();
    return PFM_OK;
}

}  // extern "C"
