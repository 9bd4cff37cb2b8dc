// mgpcg.hip -- multigrid-preconditioned CG for the MAC pressure Poisson
// problem (3D/mgpcg.py, 2D/mgpcg.py), restated for gfx950:
//   - every level is a dense array (the reference's pointer SNodes are fully
//     activated by the first solve, so they behave as dense);
//   - red/black Gauss-Seidel, Ap, restriction and prolongation use the
//     reference's operation order in fp32 without contraction, so each pass
//     is bit-identical to a sequential restatement; restriction is a gather of
//     the 2^d children in fixed order instead of float atomics;
//   - dot products accumulate in FP64 (wave shuffles, then a block tree, then
//     a fixed-order sum of the per-block partials); alpha, beta and the
//     recentring mean are formed on the device, so the only host round trip
//     per iteration is one 8-byte read of r.r for the convergence test;
//   - the coarsest level's 50 red/black sweeps run in one workgroup out of LDS.
#include <algorithm>
#include <new>

#include "pfm_host.h"

namespace pfm {

constexpr int kMaxLevels = 16;
constexpr int kMgRedBlocks = 1024;

struct MGScalars {
    double zTr;     // old z.r
    double pAp;
    double rTr;
    double r0;      // r.r of the recentred initial residual
    double sum_r;
    double new_zTr;
    float alpha, beta, mean;
    int pad;
};

struct Level {
    int d[3];
    int64_t n;
};

struct MGHandle {
    int dim;
    int N[3];
    int n_levels;
    int bc[3][2];
    int dirichlet_counts;
    int precision;
    Level lv[kMaxLevels];
    float* r[kMaxLevels];
    float* z[kMaxLevels];
    float* x;
    float* p;
    float* Ap;
    double* partial;     // max(2 * kMgRedBlocks, blocks of a level-0 stencil launch)
    MGScalars* sc;       // device
    double* host_rtr;    // pinned
};

// Stencil coefficients shared by all kernels of one level.
struct LvlDev {
    int dim;
    int d[3];
    int bc_lo[3], bc_hi[3];  // 1 when that wall decrements the diagonal
};

__device__ __forceinline__ void cell_ijk(const LvlDev& L, int64_t q, int I[3]) {
    I[2] = (int)(q % L.d[2]);
    I[1] = (int)((q / L.d[2]) % L.d[1]);
    I[0] = (int)(q / ((int64_t)L.d[2] * L.d[1]));
}

// Thread -> cell mapping for the stencil kernels: a 3D launch (64 x 4
// threads per block) with the fastest array axis on threadIdx.x, so no
// integer division is needed.  3D: (i, j, k) = (block.z, y, x); 2D: (i, j) =
// (y, x).  With `pairs`, x counts pairs of cells along the fast axis and the
// thread takes the member whose colour (I.sum() & 1) equals `phase`.
constexpr int kTx = 64, kTy = 4;

__device__ __forceinline__ bool cell_of_thread(const LvlDev& L, int phase, bool pairs, int I[3]) {
    const int fx = blockIdx.x * kTx + threadIdx.x;
    const int fy = blockIdx.y * kTy + threadIdx.y;
    if (L.dim == 3) {
        I[0] = blockIdx.z;
        I[1] = fy;
        I[2] = pairs ? 2 * fx + ((I[0] + I[1] + phase) & 1) : fx;
    } else {
        I[0] = fy;
        I[1] = pairs ? 2 * fx + ((I[0] + phase) & 1) : fx;
        I[2] = 0;
    }
    return I[0] < L.d[0] && I[1] < L.d[1] && I[2] < L.d[2];
}

static inline dim3 cell_grid(const int d[3], int dim, bool pairs) {
    const int fast = dim == 3 ? d[2] : d[1];
    const int mid = dim == 3 ? d[1] : d[0];
    const int nx = pairs ? (fast + 1) / 2 : fast;
    return dim3((unsigned)((nx + kTx - 1) / kTx), (unsigned)((mid + kTy - 1) / kTy),
                (unsigned)(dim == 3 ? d[0] : 1));
}

__device__ __forceinline__ int64_t cell_index(const LvlDev& L, const int I[3]) {
    return ((int64_t)I[0] * L.d[1] + I[1]) * L.d[2] + I[2];
}

// num_fluid_neighbors: 3D/mgpcg.py:106-115.
__device__ __forceinline__ float nfn(const LvlDev& L, const int I[3]) {
    float num = 2.0f * (float)L.dim;
    for (int a = 0; a < L.dim; ++a) {
        if (I[a] <= 0 && L.bc_lo[a]) num -= 1.0f;
        if (I[a] >= L.d[a] - 1 && L.bc_hi[a]) num -= 1.0f;
    }
    return num;
}

// neighbor_sum: 3D/mgpcg.py:92-104 (per axis: +e then -e).
__device__ __forceinline__ float nsum(const LvlDev& L, const float* __restrict__ x, int64_t q,
                                      const int I[3]) {
    float ret = 0.0f;
    int64_t stride[3] = {(int64_t)L.d[1] * L.d[2], (int64_t)L.d[2], 1};
    for (int a = 0; a < L.dim; ++a) {
        if (I[a] < L.d[a] - 1) ret = __fadd_rn(ret, x[q + stride[a]]);
        if (I[a] > 0) ret = __fadd_rn(ret, x[q - stride[a]]);
    }
    return ret;
}

__global__ void __launch_bounds__(kBlock) k_fill_f32(float* p, int64_t n, float v) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

// smooth: 3D/mgpcg.py:165-172 (one colour; one thread per cell of that colour).
__global__ void __launch_bounds__(kBlock) k_smooth(LvlDev L, float* __restrict__ z,
                                                   const float* __restrict__ r, int phase) {
    int I[3];
    if (!cell_of_thread(L, phase, true, I)) return;
    const int64_t q = cell_index(L, I);
    z[q] = __fdiv_rn(__fadd_rn(r[q], nsum(L, z, q, I)), nfn(L, I));
}

// compute_Ap (3D/mgpcg.py:117-122) fused with the p.Ap reduction; one
// partial per block, blocks numbered linearly.
__global__ void __launch_bounds__(kBlock) k_ap_dot(LvlDev L, const float* __restrict__ p,
                                                   float* __restrict__ Ap, double* __restrict__ partial) {
    __shared__ double sh[kBlock / 64];
    double s = 0.0;
    // 3D: blockIdx.z strides over the x-planes so the grid stays <= kMgRedBlocks
    const int planes = L.dim == 3 ? L.d[0] : 1;
    for (int pl = blockIdx.z; pl < planes; pl += gridDim.z) {
        int I[3];
        const int fx = blockIdx.x * kTx + threadIdx.x;
        const int fy = blockIdx.y * kTy + threadIdx.y;
        if (L.dim == 3) {
            I[0] = pl; I[1] = fy; I[2] = fx;
        } else {
            I[0] = fy; I[1] = fx; I[2] = 0;
        }
        if (I[0] < L.d[0] && I[1] < L.d[1] && I[2] < L.d[2]) {
            const int64_t q = cell_index(L, I);
            const float pq = p[q];
            const float a = __fsub_rn(__fmul_rn(nfn(L, I), pq), nsum(L, p, q, I));
            Ap[q] = a;
            s += (double)pq * (double)a;
        }
    }
    const int tid = threadIdx.y * kTx + threadIdx.x;
    // block_sum expects a 1D block index
    s = wave_sum(s);
    if ((tid & 63) == 0) sh[tid >> 6] = s;
    __syncthreads();
    if (tid == 0) {
        double t = 0.0;
        for (int w = 0; w < kBlock / 64; ++w) t += sh[w];
        partial[((int64_t)blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x] = t;
    }
}

// get_Ap for tests (no reduction).
__global__ void __launch_bounds__(kBlock) k_apply_A(LvlDev L, const float* __restrict__ p,
                                                    float* __restrict__ Ap) {
    int I[3];
    if (!cell_of_thread(L, 0, false, I)) return;
    const int64_t q = cell_index(L, I);
    Ap[q] = __fsub_rn(__fmul_rn(nfn(L, I), p[q]), nsum(L, p, q, I));
}

// restrict (3D/mgpcg.py:152-158) as a gather over the 2^d children.
__global__ void __launch_bounds__(kBlock) k_restrict(LvlDev Lf, LvlDev Lc,
                                                     const float* __restrict__ z,
                                                     const float* __restrict__ r,
                                                     float* __restrict__ rc) {
    int C[3];
    if (!cell_of_thread(Lc, 0, false, C)) return;
    const int64_t qc = cell_index(Lc, C);
    const float scale = (float)Lf.dim - 1.0f;
    // children 2C, 2C+1 per axis; the last coarse cell of an odd fine extent
    // also owns the trailing fine cell (fine I maps to min(I/2, coarse-1))
    int hi[3];
    for (int a = 0; a < 3; ++a) hi[a] = (C[a] == Lc.d[a] - 1) ? Lf.d[a] - 1 - 2 * C[a] : 1;
    if (Lf.dim == 2) hi[2] = 0;
    float acc = 0.0f;
    for (int a = 0; a <= hi[0]; ++a)
        for (int b = 0; b <= hi[1]; ++b)
            for (int c = 0; c <= hi[2]; ++c) {
                int I[3] = {2 * C[0] + a, 2 * C[1] + b, Lf.dim == 3 ? 2 * C[2] + c : 0};
                const int64_t q = ((int64_t)I[0] * Lf.d[1] + I[1]) * Lf.d[2] + I[2];
                const float res = __fsub_rn(r[q], __fsub_rn(__fmul_rn(nfn(Lf, I), z[q]), nsum(Lf, z, q, I)));
                acc = __fadd_rn(acc, __fdiv_rn(__fmul_rn(res, 1.0f), scale));
            }
    rc[qc] = acc;
}

// prolongate: 3D/mgpcg.py:160-163.
__global__ void __launch_bounds__(kBlock) k_prolongate(LvlDev Lf, LvlDev Lc, float* __restrict__ z,
                                                       const float* __restrict__ zc) {
    int I[3];
    if (!cell_of_thread(Lf, 0, false, I)) return;
    const int64_t q = cell_index(Lf, I);
    const int c0 = min(I[0] / 2, Lc.d[0] - 1), c1 = min(I[1] / 2, Lc.d[1] - 1);
    const int c2 = Lf.dim == 3 ? min(I[2] / 2, Lc.d[2] - 1) : 0;
    const int64_t qc = ((int64_t)c0 * Lc.d[1] + c1) * Lc.d[2] + c2;
    z[q] = __fadd_rn(z[q], zc[qc]);
}

// Bottom solve: `sweeps` red/black iterations on a level small enough for LDS
// (apply_preconditioner's bottom loop, 3D/mgpcg.py:192-194).  Each thread owns
// up to kBottomPer cells whose indices, colours, diagonals and right-hand
// sides are decoded once; the sweeps then only touch LDS.
constexpr int kBottomMax = 8192;
constexpr int kBottomThreads = 1024;
constexpr int kBottomPer = kBottomMax / kBottomThreads;
__global__ void __launch_bounds__(kBottomThreads) k_bottom(LvlDev L, int n, float* __restrict__ z,
                                                           const float* __restrict__ r, int sweeps) {
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
            col[k] = (I[0] + I[1] + I[2]) & 1;
            rv[k] = r[q];
            mv[k] = nfn(L, I);
            const int st[3] = {L.d[1] * L.d[2], L.d[2], 1};
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                // neighbour order of neighbor_sum: +e then -e per axis; -1 = absent
                nb[k][2 * a] = (a < L.dim && I[a] < L.d[a] - 1) ? q + st[a] : -1;
                nb[k][2 * a + 1] = (a < L.dim && I[a] > 0) ? q - st[a] : -1;
            }
            zs[q] = z[q];
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
        if (qi[k] < n) z[qi[k]] = zs[qi[k]];
}

// Partial sums of a.b (and, if b == nullptr, of a) in double.
__global__ void __launch_bounds__(kBlock) k_dot(int64_t n, const float* __restrict__ a,
                                                const float* __restrict__ b, double* __restrict__ partial) {
    __shared__ double sh[kBlock / 64];
    double s = 0.0;
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n;
         q += (int64_t)gridDim.x * blockDim.x)
        s += b ? (double)a[q] * (double)b[q] : (double)a[q];
    s = block_sum(s, sh);
    if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// update_x + update_r (3D/mgpcg.py:137-145) fused with the r.r and sum(r)
// reductions (convergence test :249-255 and recenter :174-180).
__global__ void __launch_bounds__(kBlock) k_update_xr(int64_t n, float* __restrict__ x,
                                                      float* __restrict__ r, const float* __restrict__ p,
                                                      const float* __restrict__ Ap,
                                                      const MGScalars* __restrict__ sc,
                                                      double* __restrict__ partial) {
    __shared__ double sh[kBlock / 64];
    const float alpha = sc->alpha;
    double s2 = 0.0, s1 = 0.0;
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n;
         q += (int64_t)gridDim.x * blockDim.x) {
        x[q] = __fadd_rn(x[q], __fmul_rn(alpha, p[q]));
        const float rq = __fsub_rn(r[q], __fmul_rn(alpha, Ap[q]));
        r[q] = rq;
        s2 += (double)rq * (double)rq;
        s1 += (double)rq;
    }
    s2 = block_sum(s2, sh);
    s1 = block_sum(s1, sh);
    if (threadIdx.x == 0) {
        partial[blockIdx.x] = s2;
        partial[kMgRedBlocks + blockIdx.x] = s1;
    }
}

// update_p: 3D/mgpcg.py:147-150.
__global__ void __launch_bounds__(kBlock) k_update_p(int64_t n, float* __restrict__ p,
                                                     const float* __restrict__ z,
                                                     const MGScalars* __restrict__ sc) {
    const float beta = sc->beta;
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n;
         q += (int64_t)gridDim.x * blockDim.x)
        p[q] = __fadd_rn(z[q], __fmul_rn(beta, p[q]));
}

__global__ void __launch_bounds__(kBlock) k_sub_mean(int64_t n, float* __restrict__ r,
                                                     const MGScalars* __restrict__ sc) {
    const float m = sc->mean;
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n;
         q += (int64_t)gridDim.x * blockDim.x)
        r[q] = __fsub_rn(r[q], m);
}

enum FinalizeMode { kFinMean = 0, kFinR0 = 1, kFinZTr = 2, kFinAlpha = 3, kFinRTr = 4, kFinBeta = 5 };

// Fixed-order reduction of the per-block partials and the scalar recurrences
// of solve() (3D/mgpcg.py:229-275), all in double.
__global__ void __launch_bounds__(kBlock) k_finalize(const double* __restrict__ partial, int nparts, int mode,
                                                     double eps, double ncell, MGScalars* sc) {
    __shared__ double sh[kBlock / 64];
    double s = 0.0;
    for (int b = threadIdx.x; b < nparts; b += blockDim.x) s += partial[b];
    s = block_sum(s, sh);
    double s1 = 0.0;
    if (mode == kFinRTr) {
        for (int b = threadIdx.x; b < kMgRedBlocks; b += blockDim.x) s1 += partial[kMgRedBlocks + b];
        s1 = block_sum(s1, sh);
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
            break;
        case kFinZTr:
            sc->zTr = s;
            break;
        case kFinAlpha:
            sc->pAp = s;
            sc->alpha = (float)(sc->zTr / (s + eps));
            break;
        case kFinRTr:
            sc->rTr = s;
            sc->sum_r = s1;
            sc->mean = (float)(s1 / ncell);
            break;
        case kFinBeta: {
            sc->new_zTr = s;
            sc->beta = (float)(s / (sc->zTr + eps));
            sc->zTr = s;
            break;
        }
    }
}

__global__ void k_reset_scalars(MGScalars* sc) {
    float beta = sc->beta;  // the reference's beta field persists across solves
    *sc = MGScalars{};
    sc->beta = beta;
}

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
    const int rx = L.d[0], ry = L.d[1], rz = L.d[2];
    (void)rx;
    float div;
    if (DIM == 3) {
        const float vl = ux[((int64_t)I[0] * ry + I[1]) * rz + I[2]];
        const float vr = ux[((int64_t)(I[0] + 1) * ry + I[1]) * rz + I[2]];
        const float vb = uy[((int64_t)I[0] * (ry + 1) + I[1]) * rz + I[2]];
        const float vt = uy[((int64_t)I[0] * (ry + 1) + I[1] + 1) * rz + I[2]];
        const float va = uz[((int64_t)I[0] * ry + I[1]) * (rz + 1) + I[2]];
        const float vc = uz[((int64_t)I[0] * ry + I[1]) * (rz + 1) + I[2] + 1];
        div = __fsub_rn(__fadd_rn(__fsub_rn(__fadd_rn(__fsub_rn(vr, vl), vt), vb), vc), va);
    } else {
        const float vl = ux[(int64_t)I[0] * ry + I[1]];
        const float vr = ux[(int64_t)(I[0] + 1) * ry + I[1]];
        const float vb = uy[(int64_t)I[0] * (ry + 1) + I[1]];
        const float vt = uy[(int64_t)I[0] * (ry + 1) + I[1] + 1];
        div = __fsub_rn(__fadd_rn(__fsub_rn(vr, vl), vt), vb);
    }
    r[q] = __fmul_rn(div, k);
    z[q] = 0.f;
    Ap[q] = 0.f;
    p[q] = 0.f;
    x[q] = 0.f;
}

__global__ void __launch_bounds__(kBlock) k_init_rhs(int64_t n, const float* __restrict__ rhs, float k,
                                                     float* __restrict__ r, float* __restrict__ z,
                                                     float* __restrict__ Ap, float* __restrict__ p,
                                                     float* __restrict__ x) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    r[q] = __fmul_rn(rhs[q], k);
    z[q] = 0.f;
    Ap[q] = 0.f;
    p[q] = 0.f;
    x[q] = 0.f;
}

// subtract_grad_p: 3D/mgpcg.py:392-418 (pressure outside the domain = 0).
template <int DIM>
__global__ void __launch_bounds__(kBlock) k_subtract_grad_p(LvlDev L, LvlDev F, int axis,
                                                            const float* __restrict__ pr,
                                                            float* __restrict__ u) {
    // F describes the face array of `axis` (extent d[axis] + 1)
    int I[3];
    if (!cell_of_thread(F, 0, false, I)) return;
    const int64_t f = cell_index(F, I);
    int R[3] = {I[0], I[1], I[2]}, Lf[3] = {I[0], I[1], I[2]};
    Lf[axis] -= 1;
    const float prv = R[axis] >= L.d[axis]
                          ? 0.0f
                          : pr[((int64_t)R[0] * L.d[1] + R[1]) * L.d[2] + R[2]];
    const float plv = Lf[axis] < 0 ? 0.0f : pr[((int64_t)Lf[0] * L.d[1] + Lf[1]) * L.d[2] + Lf[2]];
    u[f] = __fsub_rn(u[f], __fsub_rn(prv, plv));
}

// apply_bc: zero one boundary plane (face index `idx` along `plane_axis`) of
// the faces of `axis` (3D/mgpcg.py:359-378, 2D/mgpcg.py:292-324).
__global__ void __launch_bounds__(kBlock) k_zero_plane(int S0, int S1, int S2, int plane_axis,
                                                       int idx, float* __restrict__ u) {
    int S[3] = {S0, S1, S2};
    int a1 = plane_axis == 0 ? 1 : 0;
    int a2 = plane_axis == 2 ? 1 : 2;
    const int64_t m = (int64_t)S[a1] * S[a2];
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= m) return;
    int I[3];
    I[plane_axis] = idx;
    I[a1] = (int)(t / S[a2]);
    I[a2] = (int)(t % S[a2]);
    u[((int64_t)I[0] * S[1] + I[1]) * S[2] + I[2]] = 0.0f;
}

static LvlDev lvl_dev(const MGHandle* h, int l) {
    LvlDev L;
    L.dim = h->dim;
    for (int a = 0; a < 3; ++a) {
        L.d[a] = h->lv[l].d[a];
        int lo = a < h->dim ? h->bc[a][0] : 0, hi = a < h->dim ? h->bc[a][1] : 0;
        L.bc_lo[a] = lo == 2 || (h->dirichlet_counts && lo == 1);
        L.bc_hi[a] = hi == 2 || (h->dirichlet_counts && hi == 1);
    }
    return L;
}

static unsigned red_blocks(int64_t n) {
    (void)n;
    return kMgRedBlocks;
}

static int fill(float* p, int64_t n, float v, hipStream_t s) {
    unsigned b = grid_for(n);
    if (b > 4096) b = 4096;
    hipLaunchKernelGGL(k_fill_f32, dim3(b), dim3(kBlock), 0, s, p, n, v);
    PFM_LAUNCH_CHECK();
    return PFM_OK;
}

static const dim3 kCellBlock(kTx, kTy, 1);

static int smooth(MGHandle* h, int l, int phase, hipStream_t s) {
    LvlDev L = lvl_dev(h, l);
    hipLaunchKernelGGL(k_smooth, cell_grid(L.d, h->dim, true), kCellBlock, 0, s, L, h->z[l], h->r[l], phase);
    PFM_LAUNCH_CHECK();
    return PFM_OK;
}

static int restrict_level(MGHandle* h, int l, hipStream_t s) {
    LvlDev Lf = lvl_dev(h, l), Lc = lvl_dev(h, l + 1);
    hipLaunchKernelGGL(k_restrict, cell_grid(Lc.d, h->dim, false), kCellBlock, 0, s, Lf, Lc, h->z[l],
                       h->r[l], h->r[l + 1]);
    PFM_LAUNCH_CHECK();
    return PFM_OK;
}

static int prolongate_level(MGHandle* h, int l, hipStream_t s) {
    LvlDev Lf = lvl_dev(h, l), Lc = lvl_dev(h, l + 1);
    hipLaunchKernelGGL(k_prolongate, cell_grid(Lf.d, h->dim, false), kCellBlock, 0, s, Lf, Lc, h->z[l],
                       h->z[l + 1]);
    PFM_LAUNCH_CHECK();
    return PFM_OK;
}

#define PFM_TRY(expr)            \
    do {                         \
        int _rc = (expr);        \
        if (_rc) return _rc;     \
    } while (0)

// apply_preconditioner: 3D/mgpcg.py:182-200.
static int vcycle(MGHandle* h, hipStream_t s) {
    const int L = h->n_levels;
    PFM_TRY(fill(h->z[0], h->lv[0].n, 0.f, s));
    for (int l = 0; l < L - 1; ++l) {
        for (int i = 0; i < 2; ++i) {
            PFM_TRY(smooth(h, l, 0, s));
            PFM_TRY(smooth(h, l, 1, s));
        }
        PFM_TRY(fill(h->z[l + 1], h->lv[l + 1].n, 0.f, s));
        PFM_TRY(restrict_level(h, l, s));
    }
    const int b = L - 1;
    if (h->lv[b].n <= kBottomMax) {
        LvlDev Lb = lvl_dev(h, b);
        hipLaunchKernelGGL(k_bottom, dim3(1), dim3(kBottomThreads), 0, s, Lb, (int)h->lv[b].n, h->z[b], h->r[b], 50);
        PFM_LAUNCH_CHECK();
    } else {
        for (int i = 0; i < 50; ++i) {
            PFM_TRY(smooth(h, b, 0, s));
            PFM_TRY(smooth(h, b, 1, s));
        }
    }
    for (int l = L - 2; l >= 0; --l) {
        PFM_TRY(prolongate_level(h, l, s));
        for (int i = 0; i < 2; ++i) {
            PFM_TRY(smooth(h, l, 1, s));
            PFM_TRY(smooth(h, l, 0, s));
        }
    }
    return PFM_OK;
}

static int finalize(MGHandle* h, int mode, double eps, hipStream_t s, int nparts = kMgRedBlocks) {
    hipLaunchKernelGGL(k_finalize, dim3(1), dim3(kBlock), 0, s, h->partial, nparts, mode, eps,
                       (double)h->lv[0].n, h->sc);
    PFM_LAUNCH_CHECK();
    return PFM_OK;
}

static dim3 ap_dot_grid(const MGHandle* h) {
    dim3 gd = cell_grid(h->lv[0].d, h->dim, false);
    if (h->dim == 3) {
        const unsigned per_plane = gd.x * gd.y;
        unsigned z = per_plane >= (unsigned)kMgRedBlocks ? 1u : (unsigned)kMgRedBlocks / per_plane;
        if (z > gd.z) z = gd.z;
        gd.z = z;
    }
    return gd;
}

static int ap_dot_parts(const MGHandle* h) {
    const dim3 gd = ap_dot_grid(h);
    return (int)(gd.x * gd.y * gd.z);
}

static int dot(MGHandle* h, const float* a, const float* b, hipStream_t s) {
    hipLaunchKernelGGL(k_dot, dim3(red_blocks(h->lv[0].n)), dim3(kBlock), 0, s, h->lv[0].n, a, b, h->partial);
    PFM_LAUNCH_CHECK();
    return PFM_OK;
}

static bool all_neumann(const MGHandle* h) {
    for (int a = 0; a < h->dim; ++a)
        if (h->bc[a][0] != 2 || h->bc[a][1] != 2) return false;
    return true;
}

// solve: 3D/mgpcg.py:202-280.
static int solve(MGHandle* h, int max_iters, double eps, double tol, int tol_mode, int* iters_out,
                 double* rtr_out, hipStream_t s) {
    const int64_t n = h->lv[0].n;
    const unsigned rb = red_blocks(n);
    const bool neumann = all_neumann(h);
    hipLaunchKernelGGL(k_reset_scalars, dim3(1), dim3(1), 0, s, h->sc);
    PFM_LAUNCH_CHECK();
    if (neumann) {
        PFM_TRY(dot(h, h->r[0], nullptr, s));
        PFM_TRY(finalize(h, kFinMean, eps, s));
        hipLaunchKernelGGL(k_sub_mean, dim3(rb), dim3(kBlock), 0, s, n, h->r[0], h->sc);
        PFM_LAUNCH_CHECK();
    }
    double thresh = tol;
    if (tol_mode == 1) {
        PFM_TRY(dot(h, h->r[0], h->r[0], s));
        PFM_TRY(finalize(h, kFinR0, eps, s));
        PFM_HIP_TRY(hipMemcpyAsync(h->host_rtr, &h->sc->r0, sizeof(double), hipMemcpyDeviceToHost, s));
        PFM_HIP_TRY(hipStreamSynchronize(s));
        thresh = tol * tol * h->host_rtr[0];
    }
    PFM_TRY(vcycle(h, s));
    hipLaunchKernelGGL(k_update_p, dim3(rb), dim3(kBlock), 0, s, n, h->p, h->z[0], h->sc);
    PFM_LAUNCH_CHECK();
    PFM_TRY(dot(h, h->z[0], h->r[0], s));
    PFM_TRY(finalize(h, kFinZTr, eps, s));
    LvlDev L0 = lvl_dev(h, 0);
    int it = 0;
    double rtr = 0.0;
    bool converged = false;
    while (max_iters == -1 || it < max_iters) {
        hipLaunchKernelGGL(k_ap_dot, ap_dot_grid(h), kCellBlock, 0, s, L0, h->p, h->Ap, h->partial);
        PFM_LAUNCH_CHECK();
        PFM_TRY(finalize(h, kFinAlpha, eps, s, ap_dot_parts(h)));
        hipLaunchKernelGGL(k_update_xr, dim3(rb), dim3(kBlock), 0, s, n, h->x, h->r[0], h->p, h->Ap, h->sc,
                           h->partial);
        PFM_LAUNCH_CHECK();
        PFM_TRY(finalize(h, kFinRTr, eps, s));
        PFM_HIP_TRY(hipMemcpyAsync(h->host_rtr, &h->sc->rTr, sizeof(double), hipMemcpyDeviceToHost, s));
        PFM_HIP_TRY(hipStreamSynchronize(s));
        rtr = h->host_rtr[0];
        if (rtr < thresh) {
            converged = true;
            break;
        }
        if (neumann) {
            hipLaunchKernelGGL(k_sub_mean, dim3(rb), dim3(kBlock), 0, s, n, h->r[0], h->sc);
            PFM_LAUNCH_CHECK();
        }
        PFM_TRY(vcycle(h, s));
        PFM_TRY(dot(h, h->z[0], h->r[0], s));
        PFM_TRY(finalize(h, kFinBeta, eps, s));
        hipLaunchKernelGGL(k_update_p, dim3(rb), dim3(kBlock), 0, s, n, h->p, h->z[0], h->sc);
        PFM_LAUNCH_CHECK();
        ++it;
    }
    if (iters_out) *iters_out = it;
    if (rtr_out) *rtr_out = rtr;
    return converged ? PFM_OK : PFM_WARN_NOT_CONVERGED;
}

static int apply_bc(MGHandle* h, float* const* u, hipStream_t s) {
    for (int a = 0; a < h->dim; ++a) {
        int S[3] = {h->N[0], h->N[1], h->N[2]};
        S[a] += 1;
        // normal faces on Neumann walls (and on Dirichlet walls in 2D)
        for (int side = 0; side < 2; ++side) {
            int bt = h->bc[a][side];
            bool zero = bt == 2 || (h->dim == 2 && bt == 1);
            if (!zero) continue;
            int idx = side == 0 ? 0 : S[a] - 1;
            int a1 = a == 0 ? 1 : 0, a2 = a == 2 ? 1 : 2;
            int64_t m = (int64_t)S[a1] * S[a2];
            hipLaunchKernelGGL(k_zero_plane, dim3(grid_for(m)), dim3(kBlock), 0, s, S[0], S[1], S[2], a, idx, u[a]);
            PFM_LAUNCH_CHECK();
        }
        // 2D Dirichlet walls also zero the tangential faces on them
        if (h->dim == 2) {
            for (int b = 0; b < 2; ++b) {
                if (b == a) continue;
                for (int side = 0; side < 2; ++side) {
                    if (h->bc[b][side] != 1) continue;
                    int idx = side == 0 ? 0 : S[b] - 1;
                    int a1 = b == 0 ? 1 : 0, a2 = b == 2 ? 1 : 2;
                    int64_t m = (int64_t)S[a1] * S[a2];
                    hipLaunchKernelGGL(k_zero_plane, dim3(grid_for(m)), dim3(kBlock), 0, s, S[0], S[1], S[2], b,
                                       idx, u[a]);
                    PFM_LAUNCH_CHECK();
                }
            }
        }
    }
    return PFM_OK;
}

static int div_init(MGHandle* h, const float* const* u, float k, hipStream_t s) {
    LvlDev L0 = lvl_dev(h, 0);
    const dim3 gd = cell_grid(L0.d, h->dim, false);
    if (h->dim == 3)
        hipLaunchKernelGGL(k_div_init<3>, gd, kCellBlock, 0, s, L0, u[0], u[1], u[2], k, h->r[0], h->z[0],
                           h->Ap, h->p, h->x);
    else
        hipLaunchKernelGGL(k_div_init<2>, gd, kCellBlock, 0, s, L0, u[0], u[1], u[2], k, h->r[0], h->z[0],
                           h->Ap, h->p, h->x);
    PFM_LAUNCH_CHECK();
    return PFM_OK;
}

static int subtract_grad_p(MGHandle* h, float* const* u, hipStream_t s) {
    LvlDev L0 = lvl_dev(h, 0);
    for (int a = 0; a < h->dim; ++a) {
        LvlDev F = L0;
        F.d[a] += 1;
        const dim3 gd = cell_grid(F.d, h->dim, false);
        if (h->dim == 3)
            hipLaunchKernelGGL(k_subtract_grad_p<3>, gd, kCellBlock, 0, s, L0, F, a, h->x, u[a]);
        else
            hipLaunchKernelGGL(k_subtract_grad_p<2>, gd, kCellBlock, 0, s, L0, F, a, h->x, u[a]);
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
    PFM_CHECK_ARG(precision == 0, "precision %d not supported by this build", precision);
    int mn = N[0];
    for (int a = 0; a < dim; ++a) {
        PFM_CHECK_ARG(N[a] >= 2, "N[%d] must be >= 2", a);
        if (N[a] < mn) mn = N[a];
    }
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
    for (int a = 0; a < 3; ++a) {
        h->bc[a][0] = a < dim ? bc[2 * a] : 0;
        h->bc[a][1] = a < dim ? bc[2 * a + 1] : 0;
    }
    for (int l = 0; l < L; ++l) {
        int64_t n = 1;
        for (int a = 0; a < 3; ++a) {
            h->lv[l].d[a] = a < dim ? h->N[a] >> l : 1;
            if (h->lv[l].d[a] < 1) h->lv[l].d[a] = 1;
            n *= h->lv[l].d[a];
        }
        h->lv[l].n = n;
    }
    auto fail = [&](const char* what) {
        set_error("hipMalloc failed for %s", what);
        pfm_mg_destroy(h);
        return PFM_ERR_OOM;
    };
    for (int l = 0; l < L; ++l) {
        if (hipMalloc(&h->r[l], sizeof(float) * h->lv[l].n) != hipSuccess) return fail("r");
        if (hipMalloc(&h->z[l], sizeof(float) * h->lv[l].n) != hipSuccess) return fail("z");
    }
    const size_t b0 = sizeof(float) * h->lv[0].n;
    if (hipMalloc(&h->x, b0) != hipSuccess) return fail("x");
    if (hipMalloc(&h->p, b0) != hipSuccess) return fail("p");
    if (hipMalloc(&h->Ap, b0) != hipSuccess) return fail("Ap");
    const size_t nparts = std::max<size_t>(2 * kMgRedBlocks, (size_t)ap_dot_parts(h));
    if (hipMalloc(&h->partial, sizeof(double) * nparts) != hipSuccess) return fail("partial");
    if (hipMalloc(&h->sc, sizeof(MGScalars)) != hipSuccess) return fail("scalars");
    if (hipHostMalloc(&h->host_rtr, sizeof(double) * 4, hipHostMallocDefault) != hipSuccess)
        return fail("pinned scalars");
    PFM_HIP_TRY(hipMemset(h->sc, 0, sizeof(MGScalars)));
    for (int l = 0; l < L; ++l) {
        PFM_HIP_TRY(hipMemset(h->r[l], 0, sizeof(float) * h->lv[l].n));
        PFM_HIP_TRY(hipMemset(h->z[l], 0, sizeof(float) * h->lv[l].n));
    }
    PFM_HIP_TRY(hipMemset(h->x, 0, b0));
    PFM_HIP_TRY(hipMemset(h->p, 0, b0));
    PFM_HIP_TRY(hipMemset(h->Ap, 0, b0));
    PFM_HIP_TRY(hipDeviceSynchronize());
    *handle = h;
    return PFM_OK;
}

int pfm_mg_destroy(void* hv) {
    MGHandle* h = (MGHandle*)hv;
    if (!h) return PFM_OK;
    for (int l = 0; l < kMaxLevels; ++l) {
        if (h->r[l]) (void)hipFree(h->r[l]);
        if (h->z[l]) (void)hipFree(h->z[l]);
    }
    if (h->x) (void)hipFree(h->x);
    if (h->p) (void)hipFree(h->p);
    if (h->Ap) (void)hipFree(h->Ap);
    if (h->partial) (void)hipFree(h->partial);
    if (h->sc) (void)hipFree(h->sc);
    if (h->host_rtr) (void)hipHostFree(h->host_rtr);
    delete h;
    return PFM_OK;
}

int pfm_mg_n_levels(void* hv) { return hv ? ((MGHandle*)hv)->n_levels : 0; }

void* pfm_mg_field(void* hv, int which, int level) {
    MGHandle* h = (MGHandle*)hv;
    if (!h || level < 0 || level >= h->n_levels) return nullptr;
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
    hipLaunchKernelGGL(k_init_rhs, dim3(grid_for(n)), dim3(kBlock), 0, as_stream(stream), n, rhs, k, h->r[0],
                       h->z[0], h->Ap, h->p, h->x);
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
    return vcycle(h, as_stream(stream));
}

int pfm_mg_smooth(void* hv, int level, int phase, pfm_stream_t stream) {
    MGHandle* h = (MGHandle*)hv;
    PFM_CHECK_ARG(h && level >= 0 && level < h->n_levels && (phase == 0 || phase == 1), "bad arguments");
    return smooth(h, level, phase, as_stream(stream));
}

int pfm_mg_restrict(void* hv, int level, pfm_stream_t stream) {
    MGHandle* h = (MGHandle*)hv;
    PFM_CHECK_ARG(h && level >= 0 && level < h->n_levels - 1, "bad level");
    return restrict_level(h, level, as_stream(stream));
}

int pfm_mg_prolongate(void* hv, int level, pfm_stream_t stream) {
    MGHandle* h = (MGHandle*)hv;
    PFM_CHECK_ARG(h && level >= 0 && level < h->n_levels - 1, "bad level");
    return prolongate_level(h, level, as_stream(stream));
}

int pfm_mg_apply_A(void* hv, const float* p, float* Ap, pfm_stream_t stream) {
    MGHandle* h = (MGHandle*)hv;
    PFM_CHECK_ARG(h && p && Ap, "null argument");
    LvlDev L0 = lvl_dev(h, 0);
    hipLaunchKernelGGL(k_apply_A, cell_grid(L0.d, h->dim, false), kCellBlock, 0, as_stream(stream), L0, p, Ap);
    PFM_LAUNCH_CHECK();
    return PFM_OK;
}

}  // extern "C"
