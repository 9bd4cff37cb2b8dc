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
// each face gathers from the particles of its 4x3x3 neighbourhood of cells in
// a fixed order (cells lexicographic, then sorted particle order), so the
// result is bit-reproducible.  A prep pass composes each particle's impulse
// and affine matrix once and writes them, in sorted order, into a SoA payload
// that the gather reads contiguously per cell.
#include "pfm_host.h"

namespace pfm {

constexpr int kScanItems = 16;                       // per thread
constexpr int kScanTile = kBlock * kScanItems;       // per block

__global__ void __launch_bounds__(kBlock) k_fill_i32(int32_t* p, int64_t n, int32_t v) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

// get_particles_id_in_every_cell (3D/run_3D.py:342-351): cell index and an
// arbitrary slot from the count atomic (reordered later).
template <int DIM>
__global__ void __launch_bounds__(kBlock) k_bin_count(GridDev g, int64_t n,
                                                      const float* __restrict__ pos,
                                                      int32_t* __restrict__ cell_of,
                                                      int32_t* __restrict__ counts,
                                                      int32_t* __restrict__ slot) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int ci[3] = {0, 0, 0};
#pragma unroll
    for (int c = 0; c < DIM; ++c) {
        int v = (int)(pos[c * n + i] * g.inv_dx);
        ci[c] = max(0, min(v, g.res[c] - 1));
    }
    const int32_t cell = (int32_t)(((int64_t)ci[0] * g.res[1] + ci[1]) * g.res[2] + ci[2]);
    cell_of[i] = cell;
    slot[i] = atomicAdd(&counts[cell], 1);
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
__global__ void __launch_bounds__(kBlock) k_sort_cells(int64_t C, const int32_t* __restrict__ start,
                                                       int32_t* __restrict__ perm) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= C) return;
    const int32_t s = start[c], e = start[c + 1];
    const int m = e - s;
    if (m <= 1) return;
    constexpr int kReg = 32;
    if (m <= kReg) {
        int32_t v[kReg];
        for (int q = 0; q < m; ++q) v[q] = perm[s + q];
        for (int q = 1; q < m; ++q) {
            int32_t key = v[q];
            int r = q - 1;
            while (r >= 0 && v[r] > key) {
                v[r + 1] = v[r];
                --r;
            }
            v[r + 1] = key;
        }
        for (int q = 0; q < m; ++q) perm[s + q] = v[q];
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

// Payload (SoA, sorted order): pos/dx (DIM), imp (DIM), C rows (DIM*DIM).
template <int DIM>
struct Pay {
    static constexpr int kPos = 0;
    static constexpr int kImp = DIM;
    static constexpr int kC = 2 * DIM;
    static constexpr int kCount = 2 * DIM + DIM * DIM;
};

struct PrepArgs {
    const float* pos;
    const float* imp;
    const float* C[3];
    const float* Tgm[3];
    const float* Tin[3];
};

// P2G per-particle composition (3D/run_3D.py:460-479): T_init_C_T = cols(T_gm)
// @ (rows(init_C) @ rows(T_gm)), imp = cols(T_gm) @ cols(T_init) @ init_imp.
template <int DIM, bool FUSED>
__global__ void __launch_bounds__(kBlock) k_p2g_prep(GridDev g, int64_t n, PrepArgs a,
                                                     const int32_t* __restrict__ perm,
                                                     float* __restrict__ pay) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const int64_t i = perm[q];
    using P = Pay<DIM>;
#pragma unroll
    for (int c = 0; c < DIM; ++c) pay[(P::kPos + c) * n + q] = a.pos[c * n + i] * g.inv_dx;
    float C[3][3], imp[3];
#pragma unroll
    for (int r = 0; r < DIM; ++r) {
#pragma unroll
        for (int c = 0; c < DIM; ++c) C[r][c] = a.C[r][c * n + i];
        imp[r] = a.imp[r * n + i];
    }
    if (FUSED) {
        float A[3][3], B[3][3], P2[3][3], Q[3][3];
#pragma unroll
        for (int r = 0; r < DIM; ++r)
#pragma unroll
            for (int c = 0; c < DIM; ++c) {
                A[r][c] = a.Tgm[r][c * n + i];
                B[r][c] = a.Tin[c][r * n + i];
            }
#pragma unroll
        for (int r = 0; r < DIM; ++r)
#pragma unroll
            for (int c = 0; c < DIM; ++c) {
                float acc = C[r][0] * A[0][c];
#pragma unroll
                for (int k = 1; k < DIM; ++k) acc += C[r][k] * A[k][c];
                P2[r][c] = acc;
            }
#pragma unroll
        for (int r = 0; r < DIM; ++r)
#pragma unroll
            for (int c = 0; c < DIM; ++c) {
                float acc = A[0][r] * P2[0][c];
#pragma unroll
                for (int k = 1; k < DIM; ++k) acc += A[k][r] * P2[k][c];
                Q[r][c] = acc;
            }
        float Tt[3][3];
#pragma unroll
        for (int r = 0; r < DIM; ++r)
#pragma unroll
            for (int c = 0; c < DIM; ++c) {
                float acc = A[0][r] * B[0][c];
#pragma unroll
                for (int k = 1; k < DIM; ++k) acc += A[k][r] * B[k][c];
                Tt[r][c] = acc;
            }
        float m2[3];
#pragma unroll
        for (int r = 0; r < DIM; ++r) {
            float acc = Tt[r][0] * imp[0];
#pragma unroll
            for (int c = 1; c < DIM; ++c) acc += Tt[r][c] * imp[c];
            m2[r] = acc;
        }
#pragma unroll
        for (int r = 0; r < DIM; ++r) {
            imp[r] = m2[r];
#pragma unroll
            for (int c = 0; c < DIM; ++c) C[r][c] = Q[r][c];
        }
    }
#pragma unroll
    for (int r = 0; r < DIM; ++r) {
        pay[(P::kImp + r) * n + q] = imp[r];
#pragma unroll
        for (int c = 0; c < DIM; ++c) pay[(P::kC + r * DIM + c) * n + q] = C[r][c];
    }
}

struct GatherArgs {
    float* u[3];
    float* W[3];
    int64_t nf[3];
};

// Gather P2G for one face (3D/run_3D.py:477-552 restated as a gather): sum
// over the particles of cells [I_A-2, I_A+1] x [I_c-1, I_c+1] of
// (imp_A + C_A . dpos) * w and w, then u = sum / W where W > 0.
template <int DIM>
__global__ void __launch_bounds__(kBlock) k_p2g_gather(GridDev g, int64_t n, GatherArgs a,
                                                       const int32_t* __restrict__ start,
                                                       const float* __restrict__ pay) {
    int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int A = 0;
    while (A < DIM - 1 && f >= a.nf[A]) {
        f -= a.nf[A];
        ++A;
    }
    if (f >= a.nf[A]) return;
    using P = Pay<DIM>;
    int S[3], I[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) S[c] = g.res[c] + (c == A ? 1 : 0);
    if (DIM == 3) {
        I[2] = (int)(f % S[2]);
        I[1] = (int)((f / S[2]) % S[1]);
        I[0] = (int)(f / ((int64_t)S[2] * S[1]));
    } else {
        I[2] = 0;
        I[1] = (int)(f % S[1]);
        I[0] = (int)(f / S[1]);
    }
    // face position in grid units and dpos offsets (reference: face_id[c] for
    // c == A, face_id[c] + 0.5 otherwise)
    float fc[3];
    int lo[3], hi[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        fc[c] = (float)I[c];
        if (c < DIM) {
            lo[c] = max(0, I[c] - (c == A ? 2 : 1));
            hi[c] = min(g.res[c] - 1, I[c] + 1);
        } else {
            lo[c] = 0;
            hi[c] = 0;
        }
    }
    const float* pimp = pay + (int64_t)(P::kImp + A) * n;
    const float* pC = pay + (int64_t)(P::kC + A * DIM) * n;
    float su = 0.f, sw = 0.f;
    for (int ci = lo[0]; ci <= hi[0]; ++ci)
        for (int cj = lo[1]; cj <= hi[1]; ++cj) {
            const int64_t rowc = ((int64_t)ci * g.res[1] + cj) * g.res[2];
            const int32_t s = start[rowc + lo[2]];
            const int32_t e = start[rowc + hi[2] + 1];
            for (int32_t q = s; q < e; ++q) {
                float w = 1.f, delta = 0.f;
#pragma unroll
                for (int c = 0; c < DIM; ++c) {
                    const float pc = pay[(int64_t)(P::kPos + c) * n + q];
                    const float d = c == A ? pc - fc[c] : (pc - fc[c]) - 0.5f;
                    const float e2 = c == A ? fc[c] - pc : (fc[c] + 0.5f) - pc;
                    w = c == 0 ? N2(d) : w * N2(d);
                    const float dp = e2 * g.dx;
                    delta = c == 0 ? pC[q] * dp : delta + pC[(int64_t)c * n + q] * dp;
                }
                su += (pimp[q] + delta) * w;
                sw += w;
            }
        }
    if (sw > 0.f) su = su / sw;
    a.u[A][f] = su;
    if (a.W[A]) a.W[A][f] = sw;
}

struct Workspace {
    int32_t* counts;     // C
    int32_t* tiles;      // ntiles
    int32_t* slot;       // n
    int32_t* cell_tmp;   // n
    float* pay;          // kCount * n
};

static size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

static size_t ws_layout(const GridDev& g, int64_t n, char* base, Workspace* w) {
    const int64_t C = cell_count(g);
    const int64_t ntiles = (C + kScanTile - 1) / kScanTile;
    const int npay = g.dim == 3 ? Pay<3>::kCount : Pay<2>::kCount;
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
    t.pay = (float*)take(sizeof(float) * (size_t)npay * (size_t)(n > 0 ? n : 1));
    if (w) *w = t;
    return off;
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
    GatherArgs ga;
    float* us[3] = {u_x, u_y, u_z};
    float* Ws[3] = {W_x, W_y, W_z};
    int64_t total = 0;
    for (int c = 0; c < 3; ++c) {
        ga.u[c] = us[c];
        ga.W[c] = Ws[c];
        ga.nf[c] = c < g.dim ? face_count_h(g, c) : 0;
        total += ga.nf[c];
    }
    if (g.dim == 3)
        hipLaunchKernelGGL(k_p2g_gather<3>, dim3(grid_for(total)), dim3(kBlock), 0, s, g, n, ga, cell_start,
                           w.pay);
    else
        hipLaunchKernelGGL(k_p2g_gather<2>, dim3(grid_for(total)), dim3(kBlock), 0, s, g, n, ga, cell_start,
                           w.pay);
    PFM_LAUNCH_CHECK();
    return PFM_OK;
}

}  // extern "C"
