// p2g.hip -- deterministic particle binning (stable counting sort by cell)
// and gather-form APIC P2G.
//
// Binning: cell = int(pos/dx) per axis (clamped), integer atomics give the
// per-cell counts (exact and order independent), a three-phase scan gives the
// cell offsets, and a per-cell sort of the scattered ids restores ascending
// particle id inside each cell, so the permutation equals a stable argsort by
// cell on every run.
//
// P2G: instead of the reference's float atomics (81 face splats per particle),
// a prep pass composes each particle's impulse and affine matrix once and
// writes them, in sorted order, as a 64-byte record; a tiled pass then
// scatters each cell's particles into per-thread face windows and sums the
// windows of a tile in LDS in a fixed order; a combine pass adds the <= 2^d
// tile partials of every face in tile order.  No float atomics anywhere, so
// the result is bit-reproducible run to run.
#include "pfm_batch.h"

namespace pfm {

// particles.hip
int launch_rk4_fold_bin(const GridDev& g, int64_t n, float* pos, const float* ux, const float* uy, const float* uz,
                        float dt, float* const Tin[3], const float* const iC[3], const float* init_imp, float* rec,
                        int32_t* cell_of, int32_t* counts, int32_t* slot, hipStream_t s);

constexpr int kScanItems = 16;                       // per thread
constexpr int kScanTile = kBlock * kScanItems;       // per block

__global__ void __launch_bounds__(kBlock) k_fill_i32(int32_t* p, int64_t n, int32_t v) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

// get_particles_id_in_every_cell (3D/run_3D.py:342-351): cell index and a
// slot from the count atomic (reordered later).  Runs of equal cells among
// consecutive lanes -- the common case, particles are numbered cell by cell
// and move < 1 cell per step -- are merged into one atomic per run.
template <int DIM>
__global__ void __launch_bounds__(kBlock) k_bin_count(GridDev g, int64_t n,
                                                      const float* __restrict__ pos,
                                                      int32_t* __restrict__ cell_of,
                                                      int32_t* __restrict__ counts,
                                                      int32_t* __restrict__ slot) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = i < n;
    int32_t cell = -1;
    if (valid) {
        cell = particle_cell<DIM>(g, pos, i, n);
        cell_of[i] = cell;
    }
    count_slots(cell, valid, counts, slot + (valid ? i : 0));
}

// Batched: the cell index of a scene's particle is global (the scene's cell
// offset added), so one counting sort over all scenes' cells orders every
// scene's particles inside that scene's cell range.
__global__ void __launch_bounds__(kBlock) kb_bin_count(BUnits t, int64_t P, const float* __restrict__ pos,
                                                       int32_t* __restrict__ cell_of, int32_t* __restrict__ counts,
                                                       int32_t* __restrict__ slot) {
    // every scene's particles are binned, enabled or not: the counting sort
    // places all particles of the batch
    PFM_FOR_UNITS(t, u) {
        const BUnit w = t.u[u];
        const BScene& S = t.sc[w.scene];
        const int64_t loc = (int64_t)w.chunk * kBlock + threadIdx.x;
        const bool valid = loc < S.n;
        const int64_t i = S.p_off + (valid ? loc : 0);
        int32_t cell = -1;
        if (valid) {
            cell = (int32_t)(S.c_off + particle_cell<2>(S.g, pos, i, P));
            cell_of[i] = cell;
        }
        count_slots(cell, valid, counts, slot + i);
    }
}

// Exclusive scan, phase 1: per-tile totals.
__global__ void __launch_bounds__(kBlock) k_scan_reduce(const int32_t* __restrict__ in, int64_t n,
                                                        int32_t* __restrict__ tile_sums) {
    __shared__ int32_t sh[kBlock / 64];
    const int64_t base = (int64_t)blockIdx.x * kScanTile;
    int32_t s = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        int64_t i = base + (int64_t)k * kBlock + threadIdx.x;
        if (i < n) s += in[i];
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        int32_t t = 0;
        for (int w = 0; w < kBlock / 64; ++w) t += sh[w];
        tile_sums[blockIdx.x] = t;
    }
}

// Block-wide exclusive scan of one value per thread; returns the prefix and
// the block total.
__device__ __forceinline__ int32_t block_exclusive_scan(int32_t v, int32_t* sh, int32_t& total) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    int32_t x = v;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        int32_t y = __shfl_up(x, off, 64);
        if (lane >= off) x += y;
    }
    if (lane == 63) sh[wid] = x;
    __syncthreads();
    int32_t wave_off = 0, tot = 0;
    for (int w = 0; w < kBlock / 64; ++w) {
        int32_t t = sh[w];
        if (w < wid) wave_off += t;
        tot += t;
    }
    __syncthreads();
    total = tot;
    return wave_off + x - v;
}

// Phase 2: exclusive scan of the tile totals in one block (sequential over
// chunks of kBlock).
__global__ void __launch_bounds__(kBlock) k_scan_tiles(int32_t* __restrict__ tile_sums, int64_t ntiles) {
    __shared__ int32_t sh[kBlock / 64];
    int32_t carry = 0;
    for (int64_t base = 0; base < ntiles; base += kBlock) {
        int64_t i = base + threadIdx.x;
        int32_t v = i < ntiles ? tile_sums[i] : 0;
        int32_t total;
        int32_t ex = block_exclusive_scan(v, sh, total);
        if (i < ntiles) tile_sums[i] = carry + ex;
        carry += total;
    }
}

// Phase 3: per-tile exclusive scan plus the tile offset; out[n] = grand total.
__global__ void __launch_bounds__(kBlock) k_scan_apply(const int32_t* __restrict__ in, int64_t n,
                                                       const int32_t* __restrict__ tile_off,
                                                       int32_t* __restrict__ out) {
    __shared__ int32_t sh[kBlock / 64];
    const int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanItems;
    int32_t v[kScanItems];
    int32_t s = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        int64_t i = base + k;
        v[k] = i < n ? in[i] : 0;
        s += v[k];
    }
    int32_t total;
    int32_t ex = block_exclusive_scan(s, sh, total) + tile_off[blockIdx.x];
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        int64_t i = base + k;
        if (i < n) out[i] = ex;
        ex += v[k];
    }
    if (base <= n - 1 && base + kScanItems >= n) out[n] = ex;
}

__global__ void __launch_bounds__(kBlock) k_bin_scatter(int64_t n, const int32_t* __restrict__ cell_of,
                                                        const int32_t* __restrict__ slot,
                                                        const int32_t* __restrict__ start,
                                                        int32_t* __restrict__ perm) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    perm[start[cell_of[i]] + slot[i]] = (int32_t)i;
}

// Ascending ids inside each cell (restores the stable order the atomics lost).
// A cell of up to kSortReg particles is held in registers under static
// indices: if its ids already ascend (the usual case: a run of equal cells
// in one wave gets its slots in lane order) nothing is written; otherwise
// each id goes to its rank among the cell's ids (distinct, so the ranks are
// a permutation).  Larger cells sort in place in global memory.
constexpr int kSortReg = 16;

__global__ void __launch_bounds__(kBlock) k_sort_cells(int64_t C, const int32_t* __restrict__ start,
                                                       int32_t* __restrict__ perm) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= C) return;
    const int32_t s = start[c], e = start[c + 1];
    const int m = e - s;
    if (m <= 1) return;
    if (m <= kSortReg) {
        int32_t v[kSortReg];
#pragma unroll
        for (int q = 0; q < kSortReg; ++q) v[q] = q < m ? perm[s + q] : 0x7fffffff;
        bool sorted = true;
#pragma unroll
        for (int q = 1; q < kSortReg; ++q) sorted = sorted && (q >= m || v[q - 1] < v[q]);
        if (sorted) return;
        int rank[kSortReg];
#pragma unroll
        for (int q = 0; q < kSortReg; ++q) {
            int k = 0;
#pragma unroll
            for (int r = 0; r < kSortReg; ++r) k += v[r] < v[q] ? 1 : 0;
            rank[q] = k;
        }
#pragma unroll
        for (int q = 0; q < kSortReg; ++q)
            if (q < m) perm[s + rank[q]] = v[q];
    } else {
        for (int q = 1; q < m; ++q) {
            int32_t key = perm[s + q];
            int r = q - 1;
            while (r >= 0 && perm[s + r] > key) {
                perm[s + r + 1] = perm[s + r];
                --r;
            }
            perm[s + r + 1] = key;
        }
    }
}

// Reference layout of the per-cell lists (3D/run_3D.py:345-351).
__global__ void __launch_bounds__(kBlock) k_cell_lists(int64_t C, const int32_t* __restrict__ start,
                                                       const int32_t* __restrict__ perm, int cell_max,
                                                       int32_t* __restrict__ ids,
                                                       int32_t* __restrict__ counts) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= C) return;
    const int32_t s = start[c], m = start[c + 1] - s;
    counts[c] = m;
    for (int q = 0; q < cell_max; ++q) ids[c * cell_max + q] = q < m ? perm[s + q] : -1;
}

struct PrepArgs {
    const float* pos;
    const float* imp;
    const float* C[3];
    const float* Tgm[3];
    const float* Tin[3];
};

// Record pass of P2G: one thread per sorted slot q reads particle perm[q] and
// writes record q (write_p2g_record).
template <int DIM, bool FUSED>
__device__ __forceinline__ void prep_record(const GridDev& g, int64_t q, int64_t n, const PrepArgs& a,
                                            const int32_t* __restrict__ perm, float* __restrict__ pay) {
    const int64_t i = perm[q];
    float pos[3] = {0.f, 0.f, 0.f}, A[3][3], B[3][3], C[3][3], imp[3];
#pragma unroll
    for (int r = 0; r < DIM; ++r) {
        pos[r] = a.pos[r * n + i];
        imp[r] = a.imp[r * n + i];
#pragma unroll
        for (int c = 0; c < DIM; ++c) {
            C[r][c] = a.C[r][c * n + i];
            if (FUSED) {
                A[r][c] = a.Tgm[r][c * n + i];
                B[r][c] = a.Tin[c][r * n + i];
            }
        }
    }
    write_p2g_record<DIM>(pay + q * Pay<DIM>::kRec, pos, g.inv_dx, FUSED, A, B, C, imp);
}

template <int DIM, bool FUSED>
__global__ void __launch_bounds__(kBlock) k_p2g_prep(GridDev g, int64_t n, PrepArgs a,
                                                     const int32_t* __restrict__ perm,
                                                     float* __restrict__ pay) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    prep_record<DIM, FUSED>(g, q, n, a, perm, pay);
}


// ---- tiled P2G --------------------------------------------------------------
// A block owns a tile of cells (one thread per cell).  For each face axis A,
// every thread scatters its cell's particles (sorted, ascending id) into a
// private window of the faces they reach -- 4 along A, 3 along the other
// axes (|x_p - x_f| < 3/2) -- with separable B-spline weights.  The windows
// are then summed into an LDS box in 4*3*3 fixed "offset phases": in phase
// (i,j,k) every thread adds its window entry (i,j,k) to the face at its cell
// + (i,j,k) - 1, so no two threads touch the same face within a phase and the
// order of additions is fixed.  Only the phases of different i need a block
// barrier between them: a wave holds whole rows of the tile along axis 0
// (3D: one row, 2D: four), so for one i the waves add to disjoint faces, and
// one wave's phases (i, j, k) run in program order (a wave's LDS operations
// are performed in order; a wavefront-scope fence keeps the compiler from
// reordering them).  Each tile writes its whole box (faces of its
// cells plus the halo its particles reach) as a partial; k_p2g_combine sums
// the partials of the <= 2^d tiles that reach a face in tile order and
// normalises by the weight.  Every step is deterministic.
template <int DIM>
struct TileCfg;
template <>
struct TileCfg<3> {
    static constexpr int T0 = 4, T1 = 8, T2 = 8;
};
template <>
struct TileCfg<2> {
    static constexpr int T0 = 16, T1 = 16, T2 = 1;
};

template <int DIM, int A, int C>
struct Win {
    static constexpr int value = C >= DIM ? 1 : (C == A ? 4 : 3);
};

template <int DIM, int A>
struct Box {
    using T = TileCfg<DIM>;
    static constexpr int B0 = T::T0 + Win<DIM, A, 0>::value - 1;
    static constexpr int B1 = T::T1 + Win<DIM, A, 1>::value - 1;
    static constexpr int B2 = DIM == 3 ? T::T2 + Win<DIM, A, 2>::value - 1 : 1;
    static constexpr int size = B0 * B1 * B2;
};

template <int DIM, int A>
struct BoxOff {
    static constexpr int value = A == 0 ? 0 : BoxOff<DIM, A - 1>::value + 2 * Box<DIM, A - 1>::size;
};
template <int DIM>
struct BoxOff<DIM, 0> {
    static constexpr int value = 0;
};

template <int DIM>
struct TileStride {
    static constexpr int value =
        DIM == 3 ? BoxOff<3, 2>::value + 2 * Box<3, 2>::size : BoxOff<2, 1>::value + 2 * Box<2, 1>::size;
};

constexpr int kTileLds = 704;  // >= the largest box (3D x-faces: 7*10*10)

// Orders one wave's LDS accesses before and after it in program order (the
// hardware performs them in order; this keeps the compiler from moving them).
__device__ __forceinline__ void wave_lds_order() {
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
}

template <int DIM, int A>
__device__ __forceinline__ void tile_axis(const GridDev& g, const float* __restrict__ pay,
                                          const int32_t* __restrict__ perm, int32_t s, int32_t e,
                                          const int cc[3], const int tl[3], float* lu, float* lw,
                                          float* __restrict__ part) {
    using P = Pay<DIM>;
    constexpr int W0 = Win<DIM, A, 0>::value, W1 = Win<DIM, A, 1>::value, W2 = Win<DIM, A, 2>::value;
    using BX = Box<DIM, A>;
    constexpr int Wn[3] = {W0, W1, W2};
    float au[W0][W1][W2], aw[W0][W1][W2];
#pragma unroll
    for (int i = 0; i < W0; ++i)
#pragma unroll
        for (int j = 0; j < W1; ++j)
#pragma unroll
            for (int k = 0; k < W2; ++k) {
                au[i][j][k] = 0.f;
                aw[i][j][k] = 0.f;
            }
    for (int32_t q = s; q < e; ++q) {
        // records in sorted order, or in particle order through the permutation
        const float* rec = pay + (int64_t)(perm ? perm[q] : q) * P::kRec;
        float p[3], crow[3];
        const float impA = rec[P::imp(A)];
#pragma unroll
        for (int c = 0; c < DIM; ++c) {
            p[c] = rec[P::kPos + c];
            crow[c] = rec[P::C(A, c)];
        }
        float wt[3][4], dd[3][4];
#pragma unroll
        for (int c = 0; c < DIM; ++c) {
#pragma unroll
            for (int w = 0; w < Wn[c]; ++w) {
                const float f = (float)(cc[c] - 1 + w);
                // reference: N_2(pos - face) along A, N_2(pos - face - 0.5)
                // elsewhere; dpos = (face (+0.5) - pos) * dx (3D/run_3D.py:486-489)
                const float d = c == A ? p[c] - f : (p[c] - f) - 0.5f;
                const float o = c == A ? f - p[c] : (f + 0.5f) - p[c];
                wt[c][w] = N2(d);
                dd[c][w] = crow[c] * (o * g.dx);
            }
        }
        if (DIM == 3) {
            // w (imp_A + C_A . dpos) with w = w0 w1 w2 and the affine term
            // separable per axis: (imp_A + d0 + d1) w0 w1 times w2, plus
            // w0 w1 times (w2 d2) -- three FMAs per window face
            float x2[W2];
#pragma unroll
            for (int k = 0; k < W2; ++k) x2[k] = wt[2][k] * dd[2][k];
#pragma unroll
            for (int i = 0; i < W0; ++i)
#pragma unroll
                for (int j = 0; j < W1; ++j) {
                    const float wij = wt[0][i] * wt[1][j];
                    const float vij = wij * ((impA + dd[0][i]) + dd[1][j]);
#pragma unroll
                    for (int k = 0; k < W2; ++k) {
                        au[i][j][k] = __builtin_fmaf(vij, wt[2][k], __builtin_fmaf(wij, x2[k], au[i][j][k]));
                        aw[i][j][k] = __builtin_fmaf(wij, wt[2][k], aw[i][j][k]);
                    }
                }
        } else {
#pragma unroll
            for (int i = 0; i < W0; ++i)
#pragma unroll
                for (int j = 0; j < W1; ++j) {
                    const float wij = wt[0][i] * wt[1][j];
                    const float dij = dd[0][i] + dd[1][j];
                    au[i][j][0] += (impA + dij) * wij;
                    aw[i][j][0] += wij;
                }
        }
    }
    // zero the box, then the offset phases
    for (int t = threadIdx.x; t < BX::size; t += blockDim.x) {
        lu[t] = 0.f;
        lw[t] = 0.f;
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < W0; ++i)
#pragma unroll
        for (int j = 0; j < W1; ++j)
#pragma unroll
            for (int k = 0; k < W2; ++k) {
                const int slot = ((tl[0] + i) * BX::B1 + (tl[1] + j)) * BX::B2 + (DIM == 3 ? tl[2] + k : 0);
                lu[slot] += au[i][j][k];
                lw[slot] += aw[i][j][k];
                if (j + 1 < W1 || k + 1 < W2)
                    wave_lds_order();
                else
                    __syncthreads();
            }
    float* dst = part + BoxOff<DIM, A>::value;
    for (int t = threadIdx.x; t < BX::size; t += blockDim.x) {
        dst[t] = lu[t];
        dst[BX::size + t] = lw[t];
    }
    __syncthreads();
}

template <int DIM>
__device__ __forceinline__ void p2g_tile_body(const GridDev& g, const int32_t* __restrict__ start,
                                              const float* __restrict__ pay, const int32_t* __restrict__ perm,
                                              int tile, int nt1, int nt2, float* __restrict__ part, float* lu,
                                              float* lw) {
    using TC = TileCfg<DIM>;
    const int tb[3] = {tile / (nt1 * nt2), (tile / nt2) % nt1, tile % nt2};
    int tl[3];
    if (DIM == 3) {
        tl[0] = threadIdx.x / (TC::T1 * TC::T2);
        tl[1] = (threadIdx.x / TC::T2) % TC::T1;
        tl[2] = threadIdx.x % TC::T2;
    } else {
        tl[0] = threadIdx.x / TC::T1;
        tl[1] = threadIdx.x % TC::T1;
        tl[2] = 0;
    }
    const int T[3] = {TC::T0, TC::T1, TC::T2};
    int cc[3];
    bool valid = true;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        cc[c] = tb[c] * T[c] + tl[c];
        if (c < DIM && cc[c] >= g.res[c]) valid = false;
    }
    int32_t s = 0, e = 0;
    if (valid) {
        const int64_t cell = ((int64_t)cc[0] * g.res[1] + cc[1]) * g.res[2] + (DIM == 3 ? cc[2] : 0);
        s = start[cell];
        e = start[cell + 1];
    }
    tile_axis<DIM, 0>(g, pay, perm, s, e, cc, tl, lu, lw, part);
    tile_axis<DIM, 1>(g, pay, perm, s, e, cc, tl, lu, lw, part);
    if (DIM == 3) tile_axis<DIM, DIM == 3 ? 2 : 1>(g, pay, perm, s, e, cc, tl, lu, lw, part);
}

template <int DIM>
__global__ void __launch_bounds__(kBlock) k_p2g_tile(GridDev g, const int32_t* __restrict__ start,
                                                     const float* __restrict__ pay,
                                                     const int32_t* __restrict__ perm, int nt1, int nt2,
                                                     float* __restrict__ partial) {
    __shared__ float lu[kTileLds], lw[kTileLds];
    const int tile = blockIdx.x;
    p2g_tile_body<DIM>(g, start, pay, perm, tile, nt1, nt2, partial + (int64_t)tile * TileStride<DIM>::value, lu,
                       lw);
}

// Batched: one unit per tile; records are in global sorted order, the
// scene's cell_start entries start at its cell offset.
__global__ void __launch_bounds__(kBlock) kb_p2g_tile(BUnits t, const int32_t* __restrict__ start,
                                                      const float* __restrict__ pay, const int32_t* __restrict__ perm,
                                                      float* __restrict__ partial) {
    __shared__ float lu[kTileLds], lw[kTileLds];
    PFM_FOR_UNITS(t, u) {
        const BUnit w = t.u[u];
        if (!unit_on(t, w.scene)) continue;
        const BScene& S = t.sc[w.scene];
        p2g_tile_body<2>(S.g, start + S.c_off, pay, perm, w.chunk, S.nt[1], 1,
                         partial + (S.t_off + w.chunk) * (int64_t)TileStride<2>::value, lu, lw);
    }
}

// ---- 3D tiled P2G, one block per (tile, face axis) ---------------------------
// Same tiles, windows, offset phases and partial boxes as tile_axis, with
//  - the three face axes of a tile in three blocks, numbered 24 g + 8 A + x
//    for tile 8 g + x: workgroups are dealt to the 8 XCDs round-robin by id,
//    so the three blocks of a tile run on one XCD at about the same time and
//    the second and third reads of the tile's records hit that XCD's L2
//    (one block walking the three axes in turn re-fetches them from HBM);
//  - the spline weights and offsets of a particle formed from its offset in
//    its cell, t = p - cc (exact: cc is floor(p), or p's clamped cell), as
//    the closed-form pieces of N_2 on [0, 1), with the general N_2 kept for
//    a particle outside its clamped cell (t outside [0, 1));
//  - the next record loaded while the current one is scattered.
// The face values and their summation order are those of tile_axis: every
// d = t - (w - 1) and o = (w - 1) - t here is the same real number, rounded
// once, as p - f and f - p there.
constexpr int kRec3Fields = 7;  // p[3], imp_A, row A of C

template <int A>
__device__ __forceinline__ void load_rec3(const float* __restrict__ pay, int32_t id, float r[kRec3Fields]) {
    using P = Pay<3>;
    static_assert(P::kPos == 0 && P::imp(A) == 4 + 4 * A && P::C(A, 0) == P::imp(A) + 1, "record quads");
    const float4* rec = reinterpret_cast<const float4*>(pay + (int64_t)id * P::kRec);
    const float4 a = rec[0], b = rec[1 + A];
    r[0] = a.x;
    r[1] = a.y;
    r[2] = a.z;
    r[3] = b.x;
    r[4] = b.y;
    r[5] = b.z;
    r[6] = b.w;
}

// N_2 at the four faces cc-1 .. cc+2 along the face axis (d = t+1, t, t-1,
// t-2) and at the three cell centres cc-1 .. cc+1 across it (d = t+1/2,
// t-1/2, t-3/2), for t in [0, 1).
__device__ __forceinline__ void face_weights01(float t, float w[4]) {
    const float h = t - 0.5f;
    const float q = 0.5f * (h * h);  // 0.5 (1.5 - |t+1|)^2 and 0.5 (1.5 - |t-2|)^2
    const float a = 1.5f - t;
    const float b = 1.0f - t;
    const float c = 0.5f + t;
    const bool lo = t < 0.5f, hi = t > 0.5f;
    w[0] = lo ? q : 0.f;
    w[1] = lo ? 0.75f - t * t : 0.5f * (a * a);
    w[2] = hi ? 0.75f - b * b : 0.5f * (c * c);
    w[3] = hi ? q : 0.f;
}

__device__ __forceinline__ void centre_weights01(float t, float w[3]) {
    const float b = 1.0f - t;
    const float h = t - 0.5f;
    w[0] = 0.5f * (b * b);
    w[1] = 0.75f - h * h;
    w[2] = 0.5f * (t * t);
}

// Scatter of one record (p, imp_A, row A of C) of a particle of cell cc
// into the face window of axis A.
template <int A>
__device__ __forceinline__ void scatter3(const float r[kRec3Fields], const float fc[3], float dx,
                                         float (&au)[Win<3, A, 0>::value][Win<3, A, 1>::value][Win<3, A, 2>::value],
                                         float (&aw)[Win<3, A, 0>::value][Win<3, A, 1>::value][Win<3, A, 2>::value]) {
    constexpr int W0 = Win<3, A, 0>::value, W1 = Win<3, A, 1>::value, W2 = Win<3, A, 2>::value;
    constexpr int Wn[3] = {W0, W1, W2};
    float t[3];
    bool inside = true;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        t[c] = r[c] - fc[c];
        // 0 <= t < 1 as one unsigned compare of the bits (t is never -0)
        inside = inside && __float_as_uint(t[c]) < 0x3f800000u;
    }
    float wt[3][4];
    if (inside) {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            if (c == A)
                face_weights01(t[c], wt[c]);
            else
                centre_weights01(t[c], wt[c]);
        }
    } else {
#pragma unroll
        for (int c = 0; c < 3; ++c)
#pragma unroll
            for (int w = 0; w < Wn[c]; ++w) {
                const float d = t[c] - (float)(w - 1);
                wt[c][w] = N2(c == A ? d : d - 0.5f);
            }
    }
    // imp_A + C_A . dpos with dpos_c = (face (+1/2) - p)_c dx (3D/run_3D.py:
    // 486-489) is affine in the window offsets: with o_c = w - 1 along A and
    // w - 1/2 across, it is b + sum_c (C_A,c dx) o_c, b = imp_A - sum_c
    // (C_A,c dx) t_c.  e(i, j, k) = e01(i, j) + g2(k), and the face gets
    // w0 w1 w2 e = w0 w1 (w2 e01 + w2 g2): three FMAs per window face.
    float Cd[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) Cd[c] = r[4 + c] * dx;
    float b = r[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) b = __builtin_fmaf(-Cd[c], t[c], b);
    auto off = [](int c, int w) { return c == A ? (float)(w - 1) : (float)w - 0.5f; };
    float e0[W0], x2[W2];
#pragma unroll
    for (int i = 0; i < W0; ++i) e0[i] = __builtin_fmaf(off(0, i), Cd[0], b);
#pragma unroll
    for (int k = 0; k < W2; ++k) x2[k] = wt[2][k] * (off(2, k) * Cd[2]);
#pragma unroll
    for (int i = 0; i < W0; ++i)
#pragma unroll
        for (int j = 0; j < W1; ++j) {
            const float wij = wt[0][i] * wt[1][j];
            const float eij = __builtin_fmaf(off(1, j), Cd[1], e0[i]);
#pragma unroll
            for (int k = 0; k < W2; ++k) {
                au[i][j][k] = __builtin_fmaf(wij, __builtin_fmaf(wt[2][k], eij, x2[k]), au[i][j][k]);
                aw[i][j][k] = __builtin_fmaf(wij, wt[2][k], aw[i][j][k]);
            }
        }
}

template <int A>
__device__ __forceinline__ void p2g_axis3(const GridDev& g, const float* __restrict__ pay,
                                          const int32_t* __restrict__ perm, int32_t s, int32_t e, const int cc[3],
                                          const int tl[3], float* lu, float* lw, float* __restrict__ dst) {
    constexpr int W0 = Win<3, A, 0>::value, W1 = Win<3, A, 1>::value, W2 = Win<3, A, 2>::value;
    using BX = Box<3, A>;
    float au[W0][W1][W2], aw[W0][W1][W2];
#pragma unroll
    for (int i = 0; i < W0; ++i)
#pragma unroll
        for (int j = 0; j < W1; ++j)
#pragma unroll
            for (int k = 0; k < W2; ++k) {
                au[i][j][k] = 0.f;
                aw[i][j][k] = 0.f;
            }
    const float fc[3] = {(float)cc[0], (float)cc[1], (float)cc[2]};
    auto id_of = [&](int32_t q) { return perm ? perm[q] : q; };
    // record q+1 and the id of q+2 are in flight while q is scattered; two
    // record buffers alternate, so no copies
    float ra[kRec3Fields], rb[kRec3Fields];
    int32_t id2 = 0;
    if (s < e) load_rec3<A>(pay, id_of(s), ra);
    if (s + 1 < e) id2 = id_of(s + 1);
    int32_t q = s;
    while (q < e) {
        if (q + 1 < e) load_rec3<A>(pay, id2, rb);
        if (q + 2 < e) id2 = id_of(q + 2);
        scatter3<A>(ra, fc, g.dx, au, aw);
        if (++q >= e) break;
        if (q + 1 < e) load_rec3<A>(pay, id2, ra);
        if (q + 2 < e) id2 = id_of(q + 2);
        scatter3<A>(rb, fc, g.dx, au, aw);
        ++q;
    }
    for (int t = threadIdx.x; t < BX::size; t += blockDim.x) {
        lu[t] = 0.f;
        lw[t] = 0.f;
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < W0; ++i)
#pragma unroll
        for (int j = 0; j < W1; ++j)
#pragma unroll
            for (int k = 0; k < W2; ++k) {
                const int slot = ((tl[0] + i) * BX::B1 + (tl[1] + j)) * BX::B2 + tl[2] + k;
                lu[slot] += au[i][j][k];
                lw[slot] += aw[i][j][k];
                if (j + 1 < W1 || k + 1 < W2)
                    wave_lds_order();
                else
                    __syncthreads();
            }
    float* box = dst + BoxOff<3, A>::value;
    for (int t = threadIdx.x; t < BX::size; t += blockDim.x) {
        box[t] = lu[t];
        box[BX::size + t] = lw[t];
    }
}

__global__ void __launch_bounds__(kBlock) k_p2g_tile3(GridDev g, const int32_t* __restrict__ start,
                                                      const float* __restrict__ pay,
                                                      const int32_t* __restrict__ perm, int nt1, int nt2,
                                                      int64_t ntiles, float* __restrict__ partial) {
    using TC = TileCfg<3>;
    __shared__ float lu[kTileLds], lw[kTileLds];
    const int64_t b = blockIdx.x;
    const int A = (int)((b % 24) / 8);
    const int64_t tile = (b / 24) * 8 + (b % 8);
    if (tile >= ntiles) return;  // the grid is rounded up to whole groups of 24
    const int tb[3] = {(int)(tile / ((int64_t)nt1 * nt2)), (int)((tile / nt2) % nt1), (int)(tile % nt2)};
    const int tl[3] = {(int)threadIdx.x / (TC::T1 * TC::T2), ((int)threadIdx.x / TC::T2) % TC::T1,
                       (int)threadIdx.x % TC::T2};
    const int T[3] = {TC::T0, TC::T1, TC::T2};
    int cc[3];
    bool valid = true;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        cc[c] = tb[c] * T[c] + tl[c];
        if (cc[c] >= g.res[c]) valid = false;
    }
    int32_t s = 0, e = 0;
    if (valid) {
        const int64_t cell = ((int64_t)cc[0] * g.res[1] + cc[1]) * g.res[2] + cc[2];
        s = start[cell];
        e = start[cell + 1];
    }
    float* dst = partial + tile * TileStride<3>::value;
    if (A == 0)
        p2g_axis3<0>(g, pay, perm, s, e, cc, tl, lu, lw, dst);
    else if (A == 1)
        p2g_axis3<1>(g, pay, perm, s, e, cc, tl, lu, lw, dst);
    else
        p2g_axis3<2>(g, pay, perm, s, e, cc, tl, lu, lw, dst);
}

struct CombineArgs {
    float* u[3];
    float* W[3];
    int64_t nf[3];
    int nt[3];
};

template <int DIM, int A>
__device__ __forceinline__ void combine_face(const GridDev& g, const CombineArgs& a, const int I[3],
                                             const float* __restrict__ partial, float& su, float& sw) {
    using TC = TileCfg<DIM>;
    using BX = Box<DIM, A>;
    const int T[3] = {TC::T0, TC::T1, TC::T2};
    int tlo[3] = {0, 0, 0}, thi[3] = {0, 0, 0};
#pragma unroll
    for (int c = 0; c < DIM; ++c) {
        const int lo = max(0, I[c] - (c == A ? 2 : 1));
        const int hi = min(g.res[c] - 1, I[c] + 1);
        tlo[c] = lo / T[c];
        thi[c] = hi / T[c];
    }
    su = 0.f;
    sw = 0.f;
    for (int t0 = tlo[0]; t0 <= thi[0]; ++t0)
        for (int t1 = tlo[1]; t1 <= thi[1]; ++t1)
            for (int t2 = tlo[2]; t2 <= thi[2]; ++t2) {
                const int64_t tile = ((int64_t)t0 * a.nt[1] + t1) * a.nt[2] + t2;
                const int l0 = I[0] - (t0 * T[0] - 1);
                const int l1 = I[1] - (t1 * T[1] - 1);
                const int l2 = DIM == 3 ? I[2] - (t2 * T[2] - 1) : 0;
                const float* box = partial + tile * TileStride<DIM>::value + BoxOff<DIM, A>::value;
                const int idx = (l0 * BX::B1 + l1) * BX::B2 + l2;
                su += box[idx];
                sw += box[BX::size + idx];
            }
}

template <int DIM>
__device__ __forceinline__ void combine_one(const GridDev& g, const CombineArgs& a, int A, int64_t f,
                                            const float* __restrict__ partial) {
    int S[3], I[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) S[c] = g.res[c] + (c == A ? 1 : 0);
    // make_grid rejects face fields of 2 GiB or more, so f < 2^29 and the
    // decode is 32-bit
    const uint32_t fu = (uint32_t)f;
    if (DIM == 3) {
        const uint32_t r = fu / (uint32_t)S[2];
        I[2] = (int)(fu - r * (uint32_t)S[2]);
        I[1] = (int)(r % (uint32_t)S[1]);
        I[0] = (int)(r / (uint32_t)S[1]);
    } else {
        I[2] = 0;
        I[1] = (int)(fu % (uint32_t)S[1]);
        I[0] = (int)(fu / (uint32_t)S[1]);
    }
    float su, sw;
    if (A == 0)
        combine_face<DIM, 0>(g, a, I, partial, su, sw);
    else if (A == 1 || DIM == 2)
        combine_face<DIM, 1>(g, a, I, partial, su, sw);
    else
        combine_face<DIM, DIM == 3 ? 2 : 1>(g, a, I, partial, su, sw);
    // u /= W where W > 0 (3D/run_3D.py:530-552)
    if (sw > 0.f) su = su / sw;
    a.u[A][f] = su;
    if (a.W[A]) a.W[A][f] = sw;
}

// Blocks in XCD-contiguous order: a 128-byte line of a tile's partial box
// holds about three face rows of it, read by consecutive blocks.
template <int DIM>
__global__ void __launch_bounds__(kBlock) k_p2g_combine(GridDev g, CombineArgs a,
                                                        const float* __restrict__ partial) {
    int64_t f = xcd_contiguous_block() * blockDim.x + threadIdx.x;
    int A = 0;
    while (A < DIM - 1 && f >= a.nf[A]) {
        f -= a.nf[A];
        ++A;
    }
    if (f >= a.nf[A]) return;
    combine_one<DIM>(g, a, A, f, partial);
}

__global__ void __launch_bounds__(kBlock) kb_p2g_combine(BUnits t, float* ux, float* uy, float* Wx, float* Wy,
                                                         const float* __restrict__ partial) {
    PFM_FOR_UNITS(t, u) {
        const BUnit w = t.u[u];
        if (!unit_on(t, w.scene)) continue;
        const BScene& S = t.sc[w.scene];
        int64_t f = (int64_t)w.chunk * kBlock + threadIdx.x;
        int A = 0;
        if (f >= S.nf[0]) {
            f -= S.nf[0];
            A = 1;
        }
        if (f >= S.nf[A]) continue;
        CombineArgs a;
        a.u[0] = ux + S.f_off[0];
        a.u[1] = uy + S.f_off[1];
        a.u[2] = nullptr;
        a.W[0] = Wx ? Wx + S.f_off[0] : nullptr;
        a.W[1] = Wy ? Wy + S.f_off[1] : nullptr;
        a.W[2] = nullptr;
        a.nf[0] = S.nf[0];
        a.nf[1] = S.nf[1];
        a.nf[2] = 0;
        a.nt[0] = S.nt[0];
        a.nt[1] = S.nt[1];
        a.nt[2] = 1;
        combine_one<2>(S.g, a, A, f, partial + S.t_off * (int64_t)TileStride<2>::value);
    }
}

struct Workspace {
    int32_t* counts;     // C
    int32_t* tiles;      // scan tiles
    int32_t* slot;       // n
    int32_t* cell_tmp;   // n
    float* pay;          // kRec * n
    float* partial;      // P2G tile partials
};

static size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

static void tile_counts(const GridDev& g, int nt[3]) {
    const int T3[3] = {TileCfg<3>::T0, TileCfg<3>::T1, TileCfg<3>::T2};
    const int T2[3] = {TileCfg<2>::T0, TileCfg<2>::T1, 1};
    for (int c = 0; c < 3; ++c) {
        const int T = g.dim == 3 ? T3[c] : T2[c];
        nt[c] = c < g.dim ? (g.res[c] + T - 1) / T : 1;
    }
}

static size_t ws_layout(const GridDev& g, int64_t n, char* base, Workspace* w) {
    const int64_t C = cell_count(g);
    const int64_t ntiles = (C + kScanTile - 1) / kScanTile;
    const int rec = g.dim == 3 ? Pay<3>::kRec : Pay<2>::kRec;
    int nt[3];
    tile_counts(g, nt);
    const int64_t ptiles = (int64_t)nt[0] * nt[1] * nt[2];
    const int64_t stride = g.dim == 3 ? TileStride<3>::value : TileStride<2>::value;
    size_t off = 0;
    auto take = [&](size_t bytes) {
        char* p = base ? base + off : nullptr;
        off += align256(bytes);
        return p;
    };
    Workspace t;
    t.counts = (int32_t*)take(sizeof(int32_t) * (size_t)(C + 1));
    t.tiles = (int32_t*)take(sizeof(int32_t) * (size_t)(ntiles + 1));
    t.slot = (int32_t*)take(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
    t.cell_tmp = (int32_t*)take(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
    t.pay = (float*)take(sizeof(float) * (size_t)rec * (size_t)(n > 0 ? n : 1));
    t.partial = (float*)take(sizeof(float) * (size_t)(ptiles * stride));
    if (w) *w = t;
    return off;
}

static int tile_and_combine(const GridDev& g, const int32_t* cell_start, const int32_t* perm,
                            const Workspace& w, float* u_x, float* u_y, float* u_z, float* W_x, float* W_y,
                            float* W_z, hipStream_t s);

// Exclusive scan of the per-cell counts into cell_start, scatter of the
// particles by (cell, slot), and the per-cell sort that restores ascending ids.
static int finish_binning(const GridDev& g, int64_t n, const Workspace& w, const int32_t* cell_of,
                          int32_t* cell_start, int32_t* perm, hipStream_t s) {
    const int64_t C = cell_count(g);
    const int64_t ntiles = (C + kScanTile - 1) / kScanTile;
    hipLaunchKernelGGL(k_scan_reduce, dim3((unsigned)ntiles), dim3(kBlock), 0, s, w.counts, C, w.tiles);
    PFM_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_scan_tiles, dim3(1), dim3(kBlock), 0, s, w.tiles, ntiles);
    PFM_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_scan_apply, dim3((unsigned)ntiles), dim3(kBlock), 0, s, w.counts, C, w.tiles,
                       cell_start);
    PFM_LAUNCH_CHECK();
    if (n > 0) {
        hipLaunchKernelGGL(k_bin_scatter, dim3(grid_for(n)), dim3(kBlock), 0, s, n, cell_of, w.slot,
                           cell_start, perm);
        PFM_LAUNCH_CHECK();
        hipLaunchKernelGGL(k_sort_cells, dim3(grid_for(C)), dim3(kBlock), 0, s, C, cell_start, perm);
        PFM_LAUNCH_CHECK();
    }
    return PFM_OK;
}

}  // namespace pfm

using namespace pfm;

extern "C" {

size_t pfm_p2g_workspace_size(const pfm_grid* gp, int64_t n) {
    GridDev g;
    if (make_grid(gp, &g) != PFM_OK || n < 0) return 0;
    return ws_layout(g, n, nullptr, nullptr);
}

int pfm_bin_particles(const pfm_grid* gp, int64_t n, const float* pos, int32_t* cell_of,
                      int32_t* cell_start, int32_t* perm, void* workspace, size_t ws_bytes,
                      pfm_stream_t stream) {
    GridDev g;
    int rc = make_grid(gp, &g);
    if (rc) return rc;
    PFM_CHECK_ARG(n >= 0 && n < (int64_t)INT32_MAX, "n out of range for int32 ids");
    PFM_CHECK_ARG(cell_start && perm && workspace, "null pointer");
    PFM_CHECK_ARG(n == 0 || pos, "null positions");
    const size_t need = ws_layout(g, n, nullptr, nullptr);
    PFM_CHECK_ARG(ws_bytes >= need, "workspace too small: %zu < %zu bytes", ws_bytes, need);
    Workspace w;
    ws_layout(g, n, (char*)workspace, &w);
    if (!cell_of) cell_of = w.cell_tmp;
    hipStream_t s = as_stream(stream);
    const int64_t C = cell_count(g);
    const int64_t ntiles = (C + kScanTile - 1) / kScanTile;
    unsigned fb = grid_for(C);
    if (fb > 4096) fb = 4096;
    hipLaunchKernelGGL(k_fill_i32, dim3(fb), dim3(kBlock), 0, s, w.counts, C, 0);
    PFM_LAUNCH_CHECK();
    if (n > 0) {
        if (g.dim == 3)
            hipLaunchKernelGGL(k_bin_count<3>, dim3(grid_for(n)), dim3(kBlock), 0, s, g, n, pos, cell_of,
                               w.counts, w.slot);
        else
            hipLaunchKernelGGL(k_bin_count<2>, dim3(grid_for(n)), dim3(kBlock), 0, s, g, n, pos, cell_of,
                               w.counts, w.slot);
        PFM_LAUNCH_CHECK();
    }
    return finish_binning(g, n, w, cell_of, cell_start, perm, s);
}

int pfm_rk4_particles_p2g_record_fold_count(const pfm_grid* gp, int64_t n, float* pos, const float* u_x,
                                            const float* u_y, const float* u_z, float dt, float* Tin_x, float* Tin_y,
                                            float* Tin_z, const float* iC_x, const float* iC_y, const float* iC_z,
                                            const float* init_imp, int32_t* cell_of, void* workspace, size_t ws_bytes,
                                            pfm_stream_t stream) {
    GridDev g;
    int rc = make_grid(gp, &g);
    if (rc) return rc;
    PFM_CHECK_ARG(g.dim == 3, "the folded RK4 + record path is 3D");
    PFM_CHECK_ARG(n >= 1 && n < (int64_t)INT32_MAX, "n out of range for int32 ids");
    PFM_CHECK_ARG(pos && u_x && u_y && u_z && Tin_x && Tin_y && Tin_z && iC_x && iC_y && iC_z && init_imp &&
                      workspace,
                  "null pointer");
    const size_t need = ws_layout(g, n, nullptr, nullptr);
    PFM_CHECK_ARG(ws_bytes >= need, "workspace too small: %zu < %zu bytes", ws_bytes, need);
    Workspace w;
    ws_layout(g, n, (char*)workspace, &w);
    if (!cell_of) cell_of = w.cell_tmp;
    hipStream_t s = as_stream(stream);
    const int64_t C = cell_count(g);
    unsigned fb = grid_for(C);
    if (fb > 4096) fb = 4096;
    hipLaunchKernelGGL(k_fill_i32, dim3(fb), dim3(kBlock), 0, s, w.counts, C, 0);
    PFM_LAUNCH_CHECK();
    float* const Tin[3] = {Tin_x, Tin_y, Tin_z};
    const float* const iC[3] = {iC_x, iC_y, iC_z};
    return launch_rk4_fold_bin(g, n, pos, u_x, u_y, u_z, dt, Tin, iC, init_imp, w.pay, cell_of, w.counts, w.slot, s);
}

int pfm_bin_finish(const pfm_grid* gp, int64_t n, const int32_t* cell_of, int32_t* cell_start, int32_t* perm,
                   void* workspace, size_t ws_bytes, pfm_stream_t stream) {
    GridDev g;
    int rc = make_grid(gp, &g);
    if (rc) return rc;
    PFM_CHECK_ARG(n >= 0 && n < (int64_t)INT32_MAX, "n out of range for int32 ids");
    PFM_CHECK_ARG(cell_start && perm && workspace, "null pointer");
    const size_t need = ws_layout(g, n, nullptr, nullptr);
    PFM_CHECK_ARG(ws_bytes >= need, "workspace too small: %zu < %zu bytes", ws_bytes, need);
    Workspace w;
    ws_layout(g, n, (char*)workspace, &w);
    return finish_binning(g, n, w, cell_of ? cell_of : w.cell_tmp, cell_start, perm, as_stream(stream));
}

int pfm_cell_lists(const pfm_grid* gp, const int32_t* cell_start, const int32_t* perm,
                   int32_t cell_max, int32_t* ids, int32_t* counts, pfm_stream_t stream) {
    GridDev g;
    int rc = make_grid(gp, &g);
    if (rc) return rc;
    PFM_CHECK_ARG(cell_start && perm && ids && counts && cell_max >= 1, "bad arguments");
    const int64_t C = cell_count(g);
    hipLaunchKernelGGL(k_cell_lists, dim3(grid_for(C)), dim3(kBlock), 0, as_stream(stream), C, cell_start,
                       perm, cell_max, ids, counts);
    PFM_LAUNCH_CHECK();
    return PFM_OK;
}

int pfm_p2g(const pfm_grid* gp, int64_t n, const float* pos, const float* imp, const float* C_x,
            const float* C_y, const float* C_z, const float* Tgm_x, const float* Tgm_y,
            const float* Tgm_z, const float* Tin_x, const float* Tin_y, const float* Tin_z,
            const int32_t* cell_start, const int32_t* perm, float* u_x, float* u_y, float* u_z,
            float* W_x, float* W_y, float* W_z, void* workspace, size_t ws_bytes,
            pfm_stream_t stream) {
    GridDev g;
    int rc = make_grid(gp, &g);
    if (rc) return rc;
    PFM_CHECK_ARG(n >= 0 && n < (int64_t)INT32_MAX, "n out of range");
    PFM_CHECK_ARG(cell_start && perm && workspace && u_x && u_y && (g.dim == 2 || u_z), "null pointer");
    PFM_CHECK_ARG(n == 0 || (pos && imp && C_x && C_y && (g.dim == 2 || C_z)), "null particle field");
    const bool fused = Tgm_x != nullptr;
    PFM_CHECK_ARG(!fused || (Tgm_y && Tin_x && Tin_y && (g.dim == 2 || (Tgm_z && Tin_z))),
                  "incomplete T_grad_m / T_init columns");
    const size_t need = ws_layout(g, n, nullptr, nullptr);
    PFM_CHECK_ARG(ws_bytes >= need, "workspace too small: %zu < %zu bytes", ws_bytes, need);
    Workspace w;
    ws_layout(g, n, (char*)workspace, &w);
    hipStream_t s = as_stream(stream);
    if (n > 0) {
        PrepArgs pa;
        pa.pos = pos;
        pa.imp = imp;
        const float* Cs[3] = {C_x, C_y, C_z};
        const float* Gs[3] = {Tgm_x, Tgm_y, Tgm_z};
        const float* Is[3] = {Tin_x, Tin_y, Tin_z};
        for (int c = 0; c < 3; ++c) {
            pa.C[c] = Cs[c];
            pa.Tgm[c] = Gs[c];
            pa.Tin[c] = Is[c];
        }
        dim3 gr(grid_for(n)), bl(kBlock);
        if (g.dim == 3) {
            if (fused)
                hipLaunchKernelGGL((k_p2g_prep<3, true>), gr, bl, 0, s, g, n, pa, perm, w.pay);
            else
                hipLaunchKernelGGL((k_p2g_prep<3, false>), gr, bl, 0, s, g, n, pa, perm, w.pay);
        } else {
            if (fused)
                hipLaunchKernelGGL((k_p2g_prep<2, true>), gr, bl, 0, s, g, n, pa, perm, w.pay);
            else
                hipLaunchKernelGGL((k_p2g_prep<2, false>), gr, bl, 0, s, g, n, pa, perm, w.pay);
        }
        PFM_LAUNCH_CHECK();
    }
    return tile_and_combine(g, cell_start, nullptr, w, u_x, u_y, u_z, W_x, W_y, W_z, s);
}

int pfm_p2g_records(const pfm_grid* gp, int64_t n, const int32_t* cell_start, const int32_t* perm,
                    float* u_x, float* u_y, float* u_z, float* W_x, float* W_y, float* W_z, void* workspace,
                    size_t ws_bytes, pfm_stream_t stream) {
    GridDev g;
    int rc = make_grid(gp, &g);
    if (rc) return rc;
    PFM_CHECK_ARG(n >= 0 && n < (int64_t)INT32_MAX, "n out of range");
    PFM_CHECK_ARG(cell_start && perm && workspace && u_x && u_y && (g.dim == 2 || u_z), "null pointer");
    const size_t need = ws_layout(g, n, nullptr, nullptr);
    PFM_CHECK_ARG(ws_bytes >= need, "workspace too small: %zu < %zu bytes", ws_bytes, need);
    Workspace w;
    ws_layout(g, n, (char*)workspace, &w);
    return tile_and_combine(g, cell_start, perm, w, u_x, u_y, u_z, W_x, W_y, W_z, as_stream(stream));
}

void* pfm_p2g_record_buffer(const pfm_grid* gp, int64_t n, void* workspace) {
    GridDev g;
    if (make_grid(gp, &g) != PFM_OK || n < 0 || !workspace) return nullptr;
    Workspace w;
    ws_layout(g, n, (char*)workspace, &w);
    return w.pay;
}

}  // extern "C"

namespace pfm {

static int tile_and_combine(const GridDev& g, const int32_t* cell_start, const int32_t* perm,
                            const Workspace& w, float* u_x, float* u_y, float* u_z, float* W_x, float* W_y,
                            float* W_z, hipStream_t s) {
    int nt[3];
    tile_counts(g, nt);
    const unsigned ptiles = (unsigned)((int64_t)nt[0] * nt[1] * nt[2]);
    if (g.dim == 3)
        hipLaunchKernelGGL(k_p2g_tile3, dim3((ptiles + 7) / 8 * 24), dim3(kBlock), 0, s, g, cell_start, w.pay,
                           perm, nt[1], nt[2], (int64_t)ptiles, w.partial);
    else
        hipLaunchKernelGGL(k_p2g_tile<2>, dim3(ptiles), dim3(kBlock), 0, s, g, cell_start, w.pay, perm, nt[1],
                           nt[2], w.partial);
    PFM_LAUNCH_CHECK();
    CombineArgs ca;
    float* us[3] = {u_x, u_y, u_z};
    float* Ws[3] = {W_x, W_y, W_z};
    int64_t total = 0;
    for (int c = 0; c < 3; ++c) {
        ca.u[c] = us[c];
        ca.W[c] = Ws[c];
        ca.nf[c] = c < g.dim ? face_count_h(g, c) : 0;
        ca.nt[c] = nt[c];
        total += ca.nf[c];
    }
    if (g.dim == 3)
        hipLaunchKernelGGL(k_p2g_combine<3>, dim3(grid_for(total)), dim3(kBlock), 0, s, g, ca, w.partial);
    else
        hipLaunchKernelGGL(k_p2g_combine<2>, dim3(grid_for(total)), dim3(kBlock), 0, s, g, ca, w.partial);
    PFM_LAUNCH_CHECK();
    return PFM_OK;
}

}  // namespace pfm

// ---- launchers of the batched 2D binning and P2G -------------------------------
namespace pfm {

size_t b_p2g_partial_floats(int64_t tiles) { return (size_t)tiles * TileStride<2>::value; }

int b_bin(BUnits parts, int64_t P, int64_t C, const float* pos, int32_t* cell_of, int32_t* cell_start,
          int32_t* perm, int32_t* counts, int32_t* tiles, int32_t* slot, unsigned blocks, hipStream_t s) {
    const int64_t ntiles = (C + kScanTile - 1) / kScanTile;
    unsigned fb = grid_for(C);
    if (fb > 4096) fb = 4096;
    hipLaunchKernelGGL(k_fill_i32, dim3(fb), dim3(kBlock), 0, s, counts, C, 0);
    PFM_LAUNCH_CHECK();
    hipLaunchKernelGGL(kb_bin_count, dim3(blocks), dim3(kBlock), 0, s, parts, P, pos, cell_of, counts, slot);
    PFM_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_scan_reduce, dim3((unsigned)ntiles), dim3(kBlock), 0, s, counts, C, tiles);
    PFM_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_scan_tiles, dim3(1), dim3(kBlock), 0, s, tiles, ntiles);
    PFM_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_scan_apply, dim3((unsigned)ntiles), dim3(kBlock), 0, s, counts, C, tiles, cell_start);
    PFM_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_bin_scatter, dim3(grid_for(P)), dim3(kBlock), 0, s, P, cell_of, slot, cell_start, perm);
    PFM_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_sort_cells, dim3(grid_for(C)), dim3(kBlock), 0, s, C, cell_start, perm);
    PFM_LAUNCH_CHECK();
    return PFM_OK;
}

int b_p2g_records(BUnits tiles, BUnits faces, const int32_t* cell_start, const int32_t* perm, const float* rec,
                  float* partial, float* ux, float* uy, float* Wx, float* Wy, const unsigned blocks[2], hipStream_t s) {
    hipLaunchKernelGGL(kb_p2g_tile, dim3(blocks[0]), dim3(kBlock), 0, s, tiles, cell_start, rec, perm, partial);
    PFM_LAUNCH_CHECK();
    hipLaunchKernelGGL(kb_p2g_combine, dim3(blocks[1]), dim3(kBlock), 0, s, faces, ux, uy, Wx, Wy, partial);
    PFM_LAUNCH_CHECK();
    return PFM_OK;
}


}  // namespace pfm
