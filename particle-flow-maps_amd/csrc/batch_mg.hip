// batch_mg.hip -- the projection of the batched 2D path: apply_bc,
// divergence, MGPCG and subtract_grad_p for many 2D scenes at once
// (2D/mgpcg.py, the MGPCG_2 solver of 3D/mgpcg.py:337-347 with the PCG of
// :202-280 and the V-cycle of :182-200).
//
// Every scene keeps its own PCG scalars (BMg) and runs its own iteration
// count: all scenes advance in lockstep, one persistent launch per PCG
// phase, and a scene that has converged (or hit max_iters) is skipped by
// every later launch.  The per-cell arithmetic is that of the single-scene
// solver (same operation order, every operation rounded on its own: this
// file is compiled with -ffp-contract=off); the dot products are summed per
// unit of 256 cells and then per scene in unit order, so a scene's result
// does not depend on which other scenes share its batch.
//
// Levels of a scene above kBatchCoarseCells cells run as tiled persistent
// launches (one per red/black colour); the remaining levels, the bottom
// solve included, run in one workgroup per scene out of LDS.
#include "pfm_batch.h"

namespace pfm {

// Wall flags (BWalls): a Neumann (or, in 2D, Dirichlet) wall removes a
// neighbour from the diagonal (2D/mgpcg.py:106-115).

__device__ __forceinline__ float bnfn(const BWalls& w, int n0, int n1, int i, int j) {
    float num = 4.0f;
    if (i <= 0 && w.lo0) num -= 1.0f;
    if (i >= n0 - 1 && w.hi0) num -= 1.0f;
    if (j <= 0 && w.lo1) num -= 1.0f;
    if (j >= n1 - 1 && w.hi1) num -= 1.0f;
    return num;
}

// neighbor_sum (3D/mgpcg.py:92-104) at x (the cell), axis 0 then axis 1,
// + then - (the order of the single-scene solver).
__device__ __forceinline__ float bnsum(const float* x, int n0, int n1, int i, int j) {
    float ret = 0.0f;
    if (i < n0 - 1) ret = ret + x[n1];
    if (i > 0) ret = ret + x[-n1];
    if (j < n1 - 1) ret = ret + x[1];
    if (j > 0) ret = ret + x[-1];
    return ret;
}

__device__ __forceinline__ bool scene_active(const BMg* mg, int s) { return mg[s].active != 0; }

// ---- field kernels -------------------------------------------------------------

// apply_bc with Neumann walls everywhere: the normal faces on the domain
// boundary are zero (2D/mgpcg.py:292-324).
__global__ void __launch_bounds__(kBlock) kb_zero_walls(BUnits t, float* __restrict__ ux, float* __restrict__ uy) {
    PFM_FOR_UNITS(t, u) {
        const BUnit w = t.u[u];
        if (!unit_on(t, w.scene)) continue;
        const BScene& S = t.sc[w.scene];
        int64_t f = (int64_t)w.chunk * kBlock + threadIdx.x;
        const int n0 = S.g.res[0], n1 = S.g.res[1];
        if (f < S.nf[0]) {
            const int i = (int)(f / n1);
            if (i == 0 || i == n0) ux[S.f_off[0] + f] = 0.0f;
        } else {
            f -= S.nf[0];
            if (f >= S.nf[1]) continue;
            const int j = (int)(f % (n1 + 1));
            if (j == 0 || j == n1) uy[S.f_off[1] + f] = 0.0f;
        }
    }
}

// divergence (3D/mgpcg.py:380-390) fused with init(r, -1) (:71-79).
__global__ void __launch_bounds__(kBlock) kb_div_init(BUnits t, const float* __restrict__ ux,
                                                      const float* __restrict__ uy, float* __restrict__ r,
                                                      float* __restrict__ x, float* __restrict__ p,
                                                      float* __restrict__ Ap) {
    PFM_FOR_UNITS(t, u) {
        const BUnit w = t.u[u];
        if (!unit_on(t, w.scene)) continue;
        const BScene& S = t.sc[w.scene];
        const int64_t q = (int64_t)w.chunk * kBlock + threadIdx.x;
        if (q >= S.nc) continue;
        const int n1 = S.g.res[1];
        const int i = (int)(q / n1), j = (int)(q % n1);
        const float* vx = ux + S.f_off[0];
        const float* vy = uy + S.f_off[1];
        const float vl = vx[(int64_t)i * n1 + j];
        const float vr = vx[(int64_t)(i + 1) * n1 + j];
        const float vb = vy[(int64_t)i * (n1 + 1) + j];
        const float vt = vy[(int64_t)i * (n1 + 1) + j + 1];
        const float div = ((vr - vl) + vt) - vb;
        const int64_t g = S.c_off + q;
        r[g] = div * -1.0f;
        x[g] = 0.0f;
        p[g] = 0.0f;
        Ap[g] = 0.0f;
    }
}

// subtract_grad_p (3D/mgpcg.py:392-418), pressure outside the domain = 0.
__global__ void __launch_bounds__(kBlock) kb_subtract_grad_p(BUnits t, const float* __restrict__ x,
                                                             float* __restrict__ ux, float* __restrict__ uy) {
    PFM_FOR_UNITS(t, u) {
        const BUnit w = t.u[u];
        if (!unit_on(t, w.scene)) continue;
        const BScene& S = t.sc[w.scene];
        int64_t f = (int64_t)w.chunk * kBlock + threadIdx.x;
        const int n0 = S.g.res[0], n1 = S.g.res[1];
        const float* pr = x + S.c_off;
        if (f < S.nf[0]) {
            const int i = (int)(f / n1), j = (int)(f % n1);
            const float prv = i >= n0 ? 0.0f : pr[(int64_t)i * n1 + j];
            const float plv = i - 1 < 0 ? 0.0f : pr[(int64_t)(i - 1) * n1 + j];
            float* v = ux + S.f_off[0] + f;
            *v = *v - (prv - plv);
        } else {
            f -= S.nf[0];
            if (f >= S.nf[1]) continue;
            const int i = (int)(f / (n1 + 1)), j = (int)(f % (n1 + 1));
            const float prv = j >= n1 ? 0.0f : pr[(int64_t)i * n1 + j];
            const float plv = j - 1 < 0 ? 0.0f : pr[(int64_t)i * n1 + j - 1];
            float* v = uy + S.f_off[1] + f;
            *v = *v - (prv - plv);
        }
    }
}

// ---- level-0 PCG kernels --------------------------------------------------------

// Block sum of one double over kBlock threads (valid in thread 0).
__device__ __forceinline__ double bblock_sum(double v, double* sh) { return block_sum(v, sh); }

// Level-0 partial sums are kept per wave (kBatchWaves per unit of kBlock
// cells): a wave sum needs no block barrier, and the per-scene sum over
// (unit, wave) in order is still fixed.
__device__ __forceinline__ void wave_partial(double v, double* __restrict__ part, int32_t unit) {
    v = wave_sum(v);
    if ((threadIdx.x & 63) == 0) part[(int64_t)unit * kBatchWaves + (threadIdx.x >> 6)] = v;
}

// Per-unit partial sums of r (kBRedSum), r.r or z.r on level 0.  Level-0 r is
// kept as update_xr leaves it; every reader of the recentred r (3D/mgpcg.py:
// 174-180, 215) forms r - mean as it loads (the same rounding as storing it),
// so no separate recentring pass runs.  The sum itself reads r as stored.
__global__ void __launch_bounds__(kBlock) kb_reduce(BUnits t, int mode, const float* __restrict__ r,
                                                    const float* __restrict__ z, const BMg* __restrict__ mg,
                                                    double* __restrict__ part) {
    PFM_FOR_UNITS(t, u) {
        const BUnit w = t.u[u];
        if (!unit_on(t, w.scene)) continue;
        if (!scene_active(mg, w.scene)) continue;
        const BScene& S = t.sc[w.scene];
        const int64_t q = (int64_t)w.chunk * kBlock + threadIdx.x;
        double v = 0.0;
        if (q < S.nc) {
            const int64_t g = S.c_off + q;
            const double rv = mode == kBRedSum ? (double)r[g] : (double)(r[g] - mg[w.scene].mean);
            v = mode == kBRedSum ? rv : (mode == kBRedRR ? rv * rv : (double)z[g] * rv);
        }
        wave_partial(v, part, S.part_off + w.chunk);
    }
}

// update_p (3D/mgpcg.py:147-150) fused with compute_Ap (:117-122) and the
// p.Ap partials: p_new = z + beta p_old is formed for the cell and its four
// neighbours from p_old (pin) and written to pout, so no block reads a p that
// another block of the launch has already replaced.
__global__ void __launch_bounds__(kBlock) kb_ap_dot(BUnits t, BWalls wl, const float* __restrict__ z,
                                                    const float* __restrict__ pin, float* __restrict__ pout,
                                                    float* __restrict__ Ap, const BMg* __restrict__ mg,
                                                    double* __restrict__ part) {
    PFM_FOR_UNITS(t, u) {
        const BUnit w = t.u[u];
        if (!unit_on(t, w.scene)) continue;
        if (!scene_active(mg, w.scene)) continue;
        const BScene& S = t.sc[w.scene];
        const int64_t q = (int64_t)w.chunk * kBlock + threadIdx.x;
        const float beta = mg[w.scene].beta;
        double v = 0.0;
        if (q < S.nc) {
            const int n0 = S.g.res[0], n1 = S.g.res[1];
            const int i = (int)(q / n1), j = (int)(q % n1);
            const int64_t g = S.c_off + q;
            const float pc = z[g] + beta * pin[g];
            // neighbor_sum of p_new, axis 0 then axis 1, + then - (as bnsum)
            float ns = 0.0f;
            if (i < n0 - 1) ns = ns + (z[g + n1] + beta * pin[g + n1]);
            if (i > 0) ns = ns + (z[g - n1] + beta * pin[g - n1]);
            if (j < n1 - 1) ns = ns + (z[g + 1] + beta * pin[g + 1]);
            if (j > 0) ns = ns + (z[g - 1] + beta * pin[g - 1]);
            const float a = bnfn(wl, n0, n1, i, j) * pc - ns;
            pout[g] = pc;
            Ap[g] = a;
            v = (double)pc * (double)a;
        }
        wave_partial(v, part, S.part_off + w.chunk);
    }
}

// update_x + update_r (3D/mgpcg.py:137-145) with the r.r and sum(r) partials.
__global__ void __launch_bounds__(kBlock) kb_update_xr(BUnits t, float* __restrict__ x, float* __restrict__ r,
                                                       const float* __restrict__ p, const float* __restrict__ Ap,
                                                       const BMg* __restrict__ mg, double* __restrict__ part,
                                                       double* __restrict__ part2) {
    PFM_FOR_UNITS(t, u) {
        const BUnit w = t.u[u];
        if (!unit_on(t, w.scene)) continue;
        if (!scene_active(mg, w.scene)) continue;
        const BScene& S = t.sc[w.scene];
        const int64_t q = (int64_t)w.chunk * kBlock + threadIdx.x;
        const float alpha = mg[w.scene].alpha;
        double s2 = 0.0, s1 = 0.0;
        if (q < S.nc) {
            const int64_t g = S.c_off + q;
            x[g] = x[g] + alpha * p[g];
            const float rr = r[g] - mg[w.scene].mean;  // the recentred r
            const float rq = rr - alpha * Ap[g];
            r[g] = rq;
            s2 = (double)rq * (double)rq;
            s1 = (double)rq;
        }
        wave_partial(s2, part, S.part_off + w.chunk);
        wave_partial(s1, part2, S.part_off + w.chunk);
    }
}

// The scalar recurrences of solve (3D/mgpcg.py:229-275) in double, one block
// per scene; the scene's partials are summed in unit order.
__global__ void __launch_bounds__(kBlock) kb_finalize(const BScene* __restrict__ sc, int mode, double eps,
                                                      int max_iters, const double* __restrict__ part,
                                                      const double* __restrict__ part2,
                                                      const double* __restrict__ zr, BMg* __restrict__ mg) {
    __shared__ double sh[kBlock / 64];
    const int s = blockIdx.x;
    if (!scene_active(mg, s)) return;
    const BScene& S = sc[s];
    double v = 0.0, v1 = 0.0;
    if (mode == kBFinZTr || mode == kBFinBeta) {
        // z.r from the V-cycle's last launch (level-0 tiles, or the coarse kernel)
        for (int b = threadIdx.x; b < S.zr_n; b += blockDim.x) v += zr[S.zr_off + b];
    } else {
        const int64_t o = (int64_t)S.part_off * kBatchWaves;
        for (int b = threadIdx.x; b < S.n_part * kBatchWaves; b += blockDim.x) v += part[o + b];
    }
    v = block_sum(v, sh);
    if (mode == kBFinRTr) {
        const int64_t o = (int64_t)S.part_off * kBatchWaves;
        for (int b = threadIdx.x; b < S.n_part * kBatchWaves; b += blockDim.x) v1 += part2[o + b];
        v1 = block_sum(v1, sh);
    }
    if (threadIdx.x != 0) return;
    BMg& m = mg[s];
    const double ncell = (double)S.nc;
    switch (mode) {
        case kBFinMean:
            m.sum_r = v;
            m.mean = (float)(v / ncell);
            break;
        case kBFinR0:
            m.r0 = v;
            m.rTr = v;
            break;
        case kBFinZTr:
            m.zTr = v;
            break;
        case kBFinAlpha:
            m.pAp = v;
            m.alpha = (float)(m.zTr / (v + eps));
            break;
        case kBFinRTr:
            m.rTr = v;
            m.sum_r = v1;
            m.mean = (float)(v1 / ncell);
            if (v < m.thresh) {  // converged (3D/mgpcg.py:257-258)
                m.active = 0;
                m.converged = 1;
                m.iters = m.it;
            }
            break;
        case kBFinBeta:
            m.beta = (float)(v / (m.zTr + eps));
            m.zTr = v;
            m.it += 1;
            if (max_iters >= 0 && m.it >= max_iters) {
                m.active = 0;
                m.iters = m.it;
            }
            break;
    }
}

// Start of a solve: fresh scalars (the reference's beta persists across
// solves), every scene active unless max_iters is 0.
__global__ void kb_mg_reset(int n, BMg* __restrict__ mg, const BState* __restrict__ st, int tol_mode, double tol,
                            int max_iters) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const float beta = mg[s].beta;
    BMg m{};
    m.beta = beta;
    m.active = max_iters != 0 && st[s].enabled;
    m.thresh = tol_mode == 1 ? 0.0 : tol;
    mg[s] = m;
}

// Relative tolerance: thresh = tol^2 r0 (as the host computes it).
__global__ void kb_set_thresh(int n, BMg* __restrict__ mg, double tol) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < n) mg[s].thresh = tol * tol * mg[s].r0;
}

// Iterations and convergence of the solve just finished, into slot k.
__global__ void kb_save_solve(int n, const BMg* __restrict__ mg, int k, int32_t* __restrict__ iters,
                              int32_t* __restrict__ conv) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    iters[2 * s + k] = mg[s].converged ? mg[s].iters : mg[s].it;
    conv[2 * s + k] = mg[s].converged;
}

__global__ void __launch_bounds__(kBlock) kb_count_active(int n, const BMg* __restrict__ mg, int32_t* out) {
    __shared__ int32_t sh[kBlock / 64];
    int32_t c = 0;
    for (int s = threadIdx.x; s < n; s += blockDim.x) c += mg[s].active;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) c += __shfl_xor(c, off, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
        int32_t tot = 0;
        for (int k = 0; k < kBlock / 64; ++k) tot += sh[k];
        *out = tot;
    }
}

// ---- shared restriction gather ----------------------------------------------

// restrict (3D/mgpcg.py:152-158) as a gather of the children of coarse cell
// (ci, cj); the last coarse cell of an odd extent also owns the trailing
// fine cell.  2D scale 1 / (dim - 1) = 1.
__device__ __forceinline__ float restrict_cell(const BWalls& wl, const float* z, const float* r, int f0, int f1,
                                               int c0, int c1, int ci, int cj) {
    const int h0 = ci == c0 - 1 ? f0 - 1 - 2 * ci : 1;
    const int h1 = cj == c1 - 1 ? f1 - 1 - 2 * cj : 1;
    float acc = 0.0f;
    for (int b = 0; b <= h0; ++b)
        for (int c = 0; c <= h1; ++c) {
            const int i = 2 * ci + b, j = 2 * cj + c;
            const int64_t q = (int64_t)i * f1 + j;
            const float res = r[q] - (bnfn(wl, f0, f1, i, j) * z[q] - bnsum(z + q, f0, f1, i, j));
            acc = acc + (res * 1.0f) / 1.0f;
        }
    return acc;
}

// ---- fused tiled V-cycle passes ---------------------------------------------
// A unit is a kBT x kBT tile of one scene's level.  The tile is loaded with
// a halo of H cells; each red/black colour pass then updates the region one
// cell narrower than the last, so after k passes every cell within H - k of
// the tile holds exactly the value of k sequential sweeps over the whole
// level (a cell's update reads only its four neighbours, all of which were
// exact after the previous pass).  Cells outside the level are never loaded
// as neighbours (the wall tests of bnsum) and never updated.
constexpr int kBT = 32;

__device__ __forceinline__ float bnsum_s(const float* x, int stride, int n0, int n1, int i, int j) {
    float ret = 0.0f;
    if (i < n0 - 1) ret = ret + x[stride];
    if (i > 0) ret = ret + x[-stride];
    if (j < n1 - 1) ret = ret + x[1];
    if (j > 0) ret = ret + x[-1];
    return ret;
}

// One colour over the cells within RING of the tile (LDS arrays of width W,
// tile origin (i0, j0) at local (H, H)): thread t takes row t / HP and the
// member of pair t % HP of that row whose colour is `phase`.
template <int H, int RING>
__device__ __forceinline__ void tile_sweep(const BWalls& wl, float* z, const float* r, int n0, int n1, int i0,
                                           int j0, int phase) {
    constexpr int W = kBT + 2 * H, lo = H - RING, R = kBT + 2 * RING, HP = (R + 1) / 2;
    const int jstart = j0 - H + lo;
    for (int t = threadIdx.x; t < R * HP; t += blockDim.x) {
        const int li = lo + t / HP;
        const int i = i0 - H + li;
        const int lj = lo + 2 * (t % HP) + ((i + jstart + phase) & 1);
        const int j = j0 - H + lj;
        if (lj >= lo + R || i < 0 || j < 0 || i >= n0 || j >= n1) continue;
        const int q = li * W + lj;
        z[q] = (r[q] + bnsum_s(z + q, W, n0, n1, i, j)) / bnfn(wl, n0, n1, i, j);
    }
    __syncthreads();
}

__device__ __forceinline__ void tile_of_unit(const BScene& S, int l, int32_t chunk, int& i0, int& j0) {
    const int nt1 = (S.ld[l][1] + kBT - 1) / kBT;
    i0 = (chunk / nt1) * kBT;
    j0 = (chunk % nt1) * kBT;
}

// Pre-smoothing (z = 0, then colours 0, 1, 0, 1) and restriction of one tile
// of level l (3D/mgpcg.py:182-192): writes z[l] on the tile and r[l + 1] on the
// coarse cells whose first child lies in the tile.  The halo of 6 keeps z
// exact two cells around the tile after the four passes, which the residual
// of the trailing child of an odd extent (one cell past the tile) needs.
__global__ void __launch_bounds__(kBlock) kb_down(BUnits t, int l, BWalls wl, float* __restrict__ zg,
                                                  float* __restrict__ rg, const BMg* __restrict__ mg) {
    constexpr int H = 6, W = kBT + 2 * H;
    __shared__ float z[W * W], r[W * W];
    PFM_FOR_UNITS(t, u) {
        const BUnit w = t.u[u];
        if (!unit_on(t, w.scene)) continue;
        if (!scene_active(mg, w.scene)) continue;
        const BScene& S = t.sc[w.scene];
        const int n0 = S.ld[l][0], n1 = S.ld[l][1];
        int i0, j0;
        tile_of_unit(S, l, w.chunk, i0, j0);
        const float* rl = rg + S.l_off[l];
        const float mean = mg[w.scene].mean;
        for (int q = threadIdx.x; q < W * W; q += blockDim.x) {
            const int i = i0 - H + q / W, j = j0 - H + q % W;
            const bool in = i >= 0 && j >= 0 && i < n0 && j < n1;
            float rv = 0.0f;
            if (in) {
                rv = rl[(int64_t)i * n1 + j];
                if (l == 0) rv = rv - mean;  // level 0: the recentred r
            }
            r[q] = rv;
            z[q] = 0.0f;
        }
        __syncthreads();
        tile_sweep<H, H - 1>(wl, z, r, n0, n1, i0, j0, 0);
        tile_sweep<H, H - 2>(wl, z, r, n0, n1, i0, j0, 1);
        tile_sweep<H, H - 3>(wl, z, r, n0, n1, i0, j0, 0);
        tile_sweep<H, H - 4>(wl, z, r, n0, n1, i0, j0, 1);
        float* zl = zg + S.l_off[l];
        for (int q = threadIdx.x; q < kBT * kBT; q += blockDim.x) {
            const int i = i0 + q / kBT, j = j0 + q % kBT;
            if (i < n0 && j < n1) zl[(int64_t)i * n1 + j] = z[(q / kBT + H) * W + q % kBT + H];
        }
        // restriction (3D/mgpcg.py:152-158) of the coarse cells owned by the tile
        const int c0 = S.ld[l + 1][0], c1 = S.ld[l + 1][1];
        float* rc = rg + S.l_off[l + 1];
        constexpr int CT = kBT / 2;
        for (int q = threadIdx.x; q < CT * CT; q += blockDim.x) {
            const int ci = i0 / 2 + q / CT, cj = j0 / 2 + q % CT;
            if (ci >= c0 || cj >= c1) continue;
            const int h0 = ci == c0 - 1 ? n0 - 1 - 2 * ci : 1;
            const int h1 = cj == c1 - 1 ? n1 - 1 - 2 * cj : 1;
            float acc = 0.0f;
            for (int b = 0; b <= h0; ++b)
                for (int c = 0; c <= h1; ++c) {
                    const int i = 2 * ci + b, j = 2 * cj + c;
                    const int lq = (i - i0 + H) * W + (j - j0 + H);
                    const float res = r[lq] - (bnfn(wl, n0, n1, i, j) * z[lq] - bnsum_s(z + lq, W, n0, n1, i, j));
                    acc = acc + (res * 1.0f) / 1.0f;
                }
            rc[(int64_t)ci * c1 + cj] = acc;
        }
        __syncthreads();
    }
}

// Prolongation and post-smoothing (colours 1, 0, 1, 0) of one tile of level l
// (3D/mgpcg.py:194-200): z[l] += z[l + 1][parent] on the tile and a halo of
// 4, then the four passes.  The pre-smoothed z comes from zpre, the coarse
// correction from zpost[l + 1], and the result goes to zpost[l]: tiles read
// their halo from an array no block of the launch writes.
// DOT (level 0): also the z.r partial of the tile (the PCG's z.r after the
// preconditioner, 3D/mgpcg.py:231-235 and :264-268), into zr[S.zr_off + tile].
template <bool DOT>
__global__ void __launch_bounds__(kBlock) kb_up(BUnits t, int l, BWalls wl, const float* __restrict__ zpre,
                                                float* __restrict__ zpost, const float* __restrict__ rg,
                                                const BMg* __restrict__ mg, double* __restrict__ zr) {
    constexpr int H = 4, W = kBT + 2 * H;
    __shared__ float z[W * W], r[W * W];
    __shared__ double sh[kBlock / 64];
    PFM_FOR_UNITS(t, u) {
        const BUnit w = t.u[u];
        if (!unit_on(t, w.scene)) continue;
        if (!scene_active(mg, w.scene)) continue;
        const BScene& S = t.sc[w.scene];
        const int n0 = S.ld[l][0], n1 = S.ld[l][1];
        const int c0 = S.ld[l + 1][0], c1 = S.ld[l + 1][1];
        int i0, j0;
        tile_of_unit(S, l, w.chunk, i0, j0);
        const float* rl = rg + S.l_off[l];
        const float* zl = zpre + S.l_off[l];
        const float* zc = zpost + S.l_off[l + 1];
        for (int q = threadIdx.x; q < W * W; q += blockDim.x) {
            const int i = i0 - H + q / W, j = j0 - H + q % W;
            const bool in = i >= 0 && j >= 0 && i < n0 && j < n1;
            float zv = 0.0f, rv = 0.0f;
            if (in) {
                const int64_t g = (int64_t)i * n1 + j;
                zv = zl[g] + zc[(int64_t)min(i >> 1, c0 - 1) * c1 + min(j >> 1, c1 - 1)];
                rv = rl[g];
                if (l == 0) rv = rv - mg[w.scene].mean;  // level 0: the recentred r
            }
            z[q] = zv;
            r[q] = rv;
        }
        __syncthreads();
        tile_sweep<H, H - 1>(wl, z, r, n0, n1, i0, j0, 1);
        tile_sweep<H, H - 2>(wl, z, r, n0, n1, i0, j0, 0);
        tile_sweep<H, H - 3>(wl, z, r, n0, n1, i0, j0, 1);
        tile_sweep<H, H - 4>(wl, z, r, n0, n1, i0, j0, 0);
        float* zo = zpost + S.l_off[l];
        double v = 0.0;
        for (int q = threadIdx.x; q < kBT * kBT; q += blockDim.x) {
            const int i = i0 + q / kBT, j = j0 + q % kBT;
            const int lq = (q / kBT + H) * W + q % kBT + H;
            if (i < n0 && j < n1) {
                zo[(int64_t)i * n1 + j] = z[lq];
                if (DOT) v += (double)z[lq] * (double)r[lq];
            }
        }
        if (DOT) {
            v = bblock_sum(v, sh);
            if (threadIdx.x == 0) zr[S.zr_off + w.chunk] = v;
        }
        __syncthreads();
    }
}

// ---- per-scene coarse V-cycle out of LDS ------------------------------------

__device__ __forceinline__ void lds_smooth(const BWalls& wl, float* z, const float* r, int n0, int n1, int phase,
                                           bool zero) {
    const int hp = (n1 + 1) >> 1;
    for (int pr = threadIdx.x; pr < n0 * hp; pr += blockDim.x) {
        const int i = pr / hp;
        const int j = 2 * (pr % hp) + ((i + phase) & 1);
        if (j >= n1) continue;
        const int q = i * n1 + j;
        const float ns = zero ? 0.0f : bnsum(z + q, n0, n1, i, j);
        z[q] = (r[q] + ns) / bnfn(wl, n0, n1, i, j);
    }
    __syncthreads();
}

// Levels [coarse, n_levels) of each active scene: the rest of the V-cycle
// (3D/mgpcg.py:182-200) from r[coarse] to z[coarse], bottom solve included.
__global__ void kb_coarse(const BScene* __restrict__ sc, BWalls wl, float* __restrict__ zg,
                          float* __restrict__ rg, const BMg* __restrict__ mg, double* __restrict__ zr) {
    extern __shared__ float lds[];
    __shared__ double wsum[1024 / 64];
    const int s = blockIdx.x;
    if (!scene_active(mg, s)) return;
    const BScene& S = sc[s];
    const int lc = S.coarse, L = S.n_levels;
    int zo[kBatchMaxLevels], ro[kBatchMaxLevels];
    int off = 0;
    for (int l = lc; l < L; ++l) {
        const int n = S.ld[l][0] * S.ld[l][1];
        zo[l] = off;
        ro[l] = off + n;
        off += 2 * n;
    }
    {
        const int n = S.ld[lc][0] * S.ld[lc][1];
        // a scene whose level 0 is coarse reads the recentred r
        const float mean = mg[s].mean;
        for (int q = threadIdx.x; q < n; q += blockDim.x) {
            float rv = rg[S.l_off[lc] + q];
            if (lc == 0) rv = rv - mean;
            lds[ro[lc] + q] = rv;
        }
    }
    __syncthreads();
    for (int l = lc; l < L - 1; ++l) {
        const int n0 = S.ld[l][0], n1 = S.ld[l][1];
        float* z = lds + zo[l];
        const float* r = lds + ro[l];
        lds_smooth(wl, z, r, n0, n1, 0, true);
        lds_smooth(wl, z, r, n0, n1, 1, false);
        lds_smooth(wl, z, r, n0, n1, 0, false);
        lds_smooth(wl, z, r, n0, n1, 1, false);
        const int c0 = S.ld[l + 1][0], c1 = S.ld[l + 1][1];
        for (int q = threadIdx.x; q < c0 * c1; q += blockDim.x)
            lds[ro[l + 1] + q] = restrict_cell(wl, z, r, n0, n1, c0, c1, q / c1, q % c1);
        __syncthreads();
    }
    {
        const int b = L - 1;
        const int n0 = S.ld[b][0], n1 = S.ld[b][1];
        float* z = lds + zo[b];
        const float* r = lds + ro[b];
        for (int it = 0; it < 50; ++it) {
            lds_smooth(wl, z, r, n0, n1, 0, it == 0);
            lds_smooth(wl, z, r, n0, n1, 1, false);
        }
    }
    for (int l = L - 2; l >= lc; --l) {
        const int n0 = S.ld[l][0], n1 = S.ld[l][1];
        const int c0 = S.ld[l + 1][0], c1 = S.ld[l + 1][1];
        float* z = lds + zo[l];
        const float* zc = lds + zo[l + 1];
        for (int q = threadIdx.x; q < n0 * n1; q += blockDim.x) {
            const int i = q / n1, j = q % n1;
            z[q] = z[q] + zc[min(i >> 1, c0 - 1) * c1 + min(j >> 1, c1 - 1)];
        }
        __syncthreads();
        const float* r = lds + ro[l];
        lds_smooth(wl, z, r, n0, n1, 1, false);
        lds_smooth(wl, z, r, n0, n1, 0, false);
        lds_smooth(wl, z, r, n0, n1, 1, false);
        lds_smooth(wl, z, r, n0, n1, 0, false);
    }
    {
        const int n = S.ld[lc][0] * S.ld[lc][1];
        double v = 0.0;
        for (int q = threadIdx.x; q < n; q += blockDim.x) {
            zg[S.l_off[lc] + q] = lds[zo[lc] + q];
            if (lc == 0) v += (double)lds[zo[lc] + q] * (double)lds[ro[lc] + q];
        }
        if (lc == 0) {
            // the scene's z.r: wave sums, then the waves in order
            v = wave_sum(v);
            if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = v;
            __syncthreads();
            if (threadIdx.x == 0) {
                double tot = 0.0;
                for (int k = 0; k < (int)(blockDim.x >> 6); ++k) tot += wsum[k];
                zr[S.zr_off] = tot;
            }
        }
    }
}

// ---- launchers ---------------------------------------------------------------

#define PFM_B_LAUNCH(kern, blocks, ...)                                                   \
    do {                                                                                  \
        hipLaunchKernelGGL(kern, dim3(blocks), dim3(kBlock), 0, s, __VA_ARGS__);          \
        PFM_LAUNCH_CHECK();                                                               \
    } while (0)

int bm_zero_walls(BUnits faces, float* ux, float* uy, unsigned blocks, hipStream_t s) {
    PFM_B_LAUNCH(kb_zero_walls, blocks, faces, ux, uy);
    return PFM_OK;
}
int bm_div_init(BUnits cells, const float* ux, const float* uy, float* r, float* x, float* p, float* Ap,
                unsigned blocks, hipStream_t s) {
    PFM_B_LAUNCH(kb_div_init, blocks, cells, ux, uy, r, x, p, Ap);
    return PFM_OK;
}
int bm_subtract_grad_p(BUnits faces, const float* x, float* ux, float* uy, unsigned blocks, hipStream_t s) {
    PFM_B_LAUNCH(kb_subtract_grad_p, blocks, faces, x, ux, uy);
    return PFM_OK;
}
int bm_reduce(BUnits cells, int mode, const float* r, const float* z, const BMg* mg, double* part, unsigned blocks,
              hipStream_t s) {
    PFM_B_LAUNCH(kb_reduce, blocks, cells, mode, r, z, mg, part);
    return PFM_OK;
}
int bm_ap_dot(BUnits cells, BWalls wl, const float* z, const float* pin, float* pout, float* Ap, const BMg* mg,
              double* part, unsigned blocks, hipStream_t s) {
    PFM_B_LAUNCH(kb_ap_dot, blocks, cells, wl, z, pin, pout, Ap, mg, part);
    return PFM_OK;
}
int bm_update_xr(BUnits cells, float* x, float* r, const float* p, const float* Ap, const BMg* mg, double* part,
                 double* part2, unsigned blocks, hipStream_t s) {
    PFM_B_LAUNCH(kb_update_xr, blocks, cells, x, r, p, Ap, mg, part, part2);
    return PFM_OK;
}
int bm_finalize(int n_scenes, const BScene* sc, int mode, double eps, int max_iters, const double* part,
                const double* part2, const double* zr, BMg* mg, hipStream_t s) {
    PFM_B_LAUNCH(kb_finalize, n_scenes, sc, mode, eps, max_iters, part, part2, zr, mg);
    return PFM_OK;
}
int bm_mg_reset(int n_scenes, BMg* mg, const BState* st, int tol_mode, double tol, int max_iters, hipStream_t s) {
    hipLaunchKernelGGL(kb_mg_reset, dim3(grid_for(n_scenes)), dim3(kBlock), 0, s, n_scenes, mg, st, tol_mode, tol,
                       max_iters);
    PFM_LAUNCH_CHECK();
    return PFM_OK;
}
int bm_set_thresh(int n_scenes, BMg* mg, double tol, hipStream_t s) {
    hipLaunchKernelGGL(kb_set_thresh, dim3(grid_for(n_scenes)), dim3(kBlock), 0, s, n_scenes, mg, tol);
    PFM_LAUNCH_CHECK();
    return PFM_OK;
}
int bm_save_solve(int n_scenes, const BMg* mg, int k, int32_t* iters, int32_t* conv, hipStream_t s) {
    hipLaunchKernelGGL(kb_save_solve, dim3(grid_for(n_scenes)), dim3(kBlock), 0, s, n_scenes, mg, k, iters, conv);
    PFM_LAUNCH_CHECK();
    return PFM_OK;
}
int bm_count_active(int n_scenes, const BMg* mg, int32_t* out, hipStream_t s) {
    hipLaunchKernelGGL(kb_count_active, dim3(1), dim3(kBlock), 0, s, n_scenes, mg, out);
    PFM_LAUNCH_CHECK();
    return PFM_OK;
}
// Compaction of unit tables to the scenes still iterating.  Every block
// forms the exclusive scan over scenes of the active scenes' unit counts of
// its list (blockIdx.y) in LDS, then moves its share of the list's units:
// unit u of scene s goes to off[s] + (u - beg[s]), so scene order and the
// order inside a scene are kept.
constexpr int kMaxCompactLists = kBatchMaxLevels + 1;
constexpr int kCompactThreads = 256;
struct CompactArgs {
    const BUnit* src[kMaxCompactLists];
    BUnit* dst[kMaxCompactLists];
    int32_t len[kMaxCompactLists];
};

__global__ void __launch_bounds__(kCompactThreads) kb_compact(int S, const BMg* __restrict__ mg, CompactArgs a,
                                                              const int32_t* __restrict__ beg,
                                                              int32_t* __restrict__ count) {
    __shared__ int32_t off[kBatchCompactMaxScenes + 1];
    __shared__ int32_t wsum[kCompactThreads / 64];
    const int k = blockIdx.y;
    const int32_t* bk = beg + (int64_t)k * (S + 1);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    int32_t carry = 0;
    for (int base = 0; base < S; base += blockDim.x) {
        const int s = base + threadIdx.x;
        const int32_t n = (s < S && mg[s].active) ? bk[s + 1] - bk[s] : 0;
        int32_t x = n;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const int32_t y = __shfl_up(x, o, 64);
            if (lane >= o) x += y;
        }
        if (lane == 63) wsum[wid] = x;
        __syncthreads();
        int32_t woff = 0, tot = 0;
        for (int w2 = 0; w2 < kCompactThreads / 64; ++w2) {
            if (w2 < wid) woff += wsum[w2];
            tot += wsum[w2];
        }
        if (s < S) off[s] = carry + woff + x - n;
        carry += tot;
        __syncthreads();
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) count[k] = carry;
    for (int32_t u = blockIdx.x * blockDim.x + threadIdx.x; u < a.len[k]; u += gridDim.x * blockDim.x) {
        const BUnit w = a.src[k][u];
        if (mg[w.scene].active) a.dst[k][off[w.scene] + (u - bk[w.scene])] = w;
    }
}

int bm_compact(int n_scenes, int n_lists, const BMg* mg, const BUnit* const* src, const int32_t* len,
               const int32_t* beg, BUnit* const* dst, int32_t* count, hipStream_t s) {
    PFM_CHECK_ARG(n_lists >= 1 && n_lists <= kMaxCompactLists, "bad list count %d", n_lists);
    PFM_CHECK_ARG(n_scenes <= kBatchCompactMaxScenes, "too many scenes for the compaction (%d)", n_scenes);
    CompactArgs a{};
    int32_t mx = 1;
    for (int k = 0; k < n_lists; ++k) {
        a.src[k] = src[k];
        a.dst[k] = dst[k];
        a.len[k] = len[k];
        mx = len[k] > mx ? len[k] : mx;
    }
    unsigned gx = (unsigned)((mx + kCompactThreads - 1) / kCompactThreads);
    if (gx > 256) gx = 256;
    hipLaunchKernelGGL(kb_compact, dim3(gx, (unsigned)n_lists), dim3(kCompactThreads), 0, s, n_scenes, mg, a, beg,
                       count);
    PFM_LAUNCH_CHECK();
    return PFM_OK;
}

int bm_down(BUnits tiles, int l, BWalls wl, float* z, float* r, const BMg* mg, unsigned blocks, hipStream_t s) {
    PFM_B_LAUNCH(kb_down, blocks, tiles, l, wl, z, r, mg);
    return PFM_OK;
}
int bm_up(BUnits tiles, int l, BWalls wl, const float* zpre, float* zpost, const float* r, const BMg* mg,
          double* zr, unsigned blocks, hipStream_t s) {
    if (l == 0)
        PFM_B_LAUNCH(kb_up<true>, blocks, tiles, l, wl, zpre, zpost, r, mg, zr);
    else
        PFM_B_LAUNCH(kb_up<false>, blocks, tiles, l, wl, zpre, zpost, r, mg, zr);
    return PFM_OK;
}
int bm_coarse(int n_scenes, const BScene* sc, BWalls wl, float* z, float* r, const BMg* mg, double* zr, int threads,
              size_t lds_bytes, hipStream_t s) {
    hipLaunchKernelGGL(kb_coarse, dim3(n_scenes), dim3(threads), lds_bytes, s, sc, wl, z, r, mg, zr);
    PFM_LAUNCH_CHECK();
    return PFM_OK;
}

}  // namespace pfm
