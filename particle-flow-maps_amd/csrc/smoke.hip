// smoke.hip -- particle smoke of the 3D driver: the particles' smoke and its
// gradient from the grid (init_particles_smoke, 3D/init_conditions.py:275-279
// with interp_grad_2_smoke, 3D/taichi_utils.py:378-421), the APIC splat back
// to the cells (P2G_smoke, 3D/run_3D.py:554-593) and the clamp of each cell
// to the initial smoke around its particles' initial cells
// (clamp_smoke_particle, 3D/run_3D.py:595-614).  Smoke is 4 channels, cell
// fields channel-major (4, cells), particle smoke (4, n), the particle
// gradient a 3x4 matrix per particle stored (12, n) row-major (row r, channel
// c at component 4 r + c).
#include "pfm_host.h"

namespace pfm {

// dN_2 (3D/taichi_utils.py:262-272).
__device__ __forceinline__ float dN2(float x) {
    float r = 0.0f;
    const float ax = fabsf(x);
    if (ax < 0.5f)
        r = -2.0f * ax;
    else if (ax < 1.5f)
        r = 0.5f * (2.0f * ax - 3.0f);
    if (x < 0.0f) r *= -1.0f;
    return r;
}

// The 64-point loop of interp_grad_2_smoke, literally (a reinitialisation
// kernel, run at most once per step): interpolated smoke and new_C rows.
__global__ void __launch_bounds__(kBlock) k_g2p_smoke(GridDev g, int64_t n, const float* __restrict__ pos,
                                                      const float* __restrict__ smoke, float* __restrict__ ps,
                                                      float* __restrict__ pg) {
#pragma clang fp contract(off)
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const int64_t C = (int64_t)g.res[0] * g.res[1] * g.res[2];
    float sv[3], fl[3];
    for (int a = 0; a < 3; ++a) {
        const float u = pos[a * n + p] * g.inv_dx - 0.5f;
        sv[a] = fmaxf(1.0f, fminf(u, g.hi_c[a]));
        fl[a] = floorf(sv[a]);
    }
    float val[4] = {0.f, 0.f, 0.f, 0.f}, nc[3][4] = {};
    for (int i = -1; i < 3; ++i)
        for (int j = -1; j < 3; ++j)
            for (int k = -1; k < 3; ++k) {
                const float x = sv[0] - (fl[0] + (float)i), y = sv[1] - (fl[1] + (float)j),
                            z = sv[2] - (fl[2] + (float)k);
                const int ci = min(max((int)(fl[0] + (float)i), 0), g.res[0] - 1);
                const int cj = min(max((int)(fl[1] + (float)j), 0), g.res[1] - 1);
                const int ck = min(max((int)(fl[2] + (float)k), 0), g.res[2] - 1);
                const int64_t q = ((int64_t)ci * g.res[1] + cj) * g.res[2] + ck;
                const float Nx = N2(x), Ny = N2(y), Nz = N2(z);
                const float dw[3] = {g.inv_dx * dN2(x) * Ny * Nz, g.inv_dx * Nx * dN2(y) * Nz,
                                     g.inv_dx * Nx * Ny * dN2(z)};
                for (int c = 0; c < 4; ++c) {
                    const float v = smoke[c * C + q];
                    val[c] = val[c] + v * Nx * Ny * Nz;
                    for (int r = 0; r < 3; ++r) nc[r][c] = nc[r][c] + dw[r] * v;
                }
            }
    for (int c = 0; c < 4; ++c) ps[c * n + p] = val[c];
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 4; ++c) pg[(4 * r + c) * n + p] = nc[r][c];
}

struct SmokeP2GArgs {
    const float* pos;
    const float* ps;     // (4, n)
    const float* grad;   // (12, n)
    const float* Tgm[3];
    const float* Tin[3];
    int apic;
};

// P2G_smoke as a gather: each cell visits the particles binned in the cells
// within two of it (every particle whose 4^3 stencil int(pos - 0.5) - 1 ..
// + 2 covers the cell lies there, including particles outside the domain,
// which the binning clamps to the boundary cells), in cell then particle-id
// order, so the sums are deterministic.  Each contribution is the
// reference's: weight N_2 N_2 N_2 of the cell-centre distances, and with
// APIC the particle smoke plus dpos (T_grad_m^T T_init grad).
__global__ void __launch_bounds__(kBlock) k_p2g_smoke(GridDev g, int64_t n, SmokeP2GArgs a,
                                                      const int32_t* __restrict__ start,
                                                      const int32_t* __restrict__ perm, float* __restrict__ smoke,
                                                      float* __restrict__ weight) {
#pragma clang fp contract(off)
    const int64_t C = (int64_t)g.res[0] * g.res[1] * g.res[2];
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= C) return;
    const int I[3] = {(int)(q / ((int64_t)g.res[1] * g.res[2])), (int)((q / g.res[2]) % g.res[1]),
                      (int)(q % g.res[2])};
    float acc[4] = {0.f, 0.f, 0.f, 0.f}, wsum = 0.f;
    for (int ja = max(I[0] - 2, 0); ja <= min(I[0] + 2, g.res[0] - 1); ++ja)
        for (int jb = max(I[1] - 2, 0); jb <= min(I[1] + 2, g.res[1] - 1); ++jb)
            for (int jc = max(I[2] - 2, 0); jc <= min(I[2] + 2, g.res[2] - 1); ++jc) {
                const int64_t J = ((int64_t)ja * g.res[1] + jb) * g.res[2] + jc;
                for (int32_t s = start[J]; s < start[J + 1]; ++s) {
                    const int64_t p = perm[s];
                    float P[3];
                    bool cover = true;
                    for (int r = 0; r < 3; ++r) {
                        P[r] = a.pos[r * n + p] * g.inv_dx;
                        const int off = I[r] - (int)(P[r] - 0.5f);
                        cover = cover && off >= -1 && off <= 2;
                    }
                    if (!cover) continue;
                    const float w = N2((P[0] - (float)I[0]) - 0.5f) * N2((P[1] - (float)I[1]) - 0.5f) *
                                    N2((P[2] - (float)I[2]) - 0.5f);
                    wsum = wsum + w;
                    float delta[4] = {0.f, 0.f, 0.f, 0.f};
                    if (a.apic) {
                        float dpos[3], M[3][3], MG[3][4];
                        for (int r = 0; r < 3; ++r) dpos[r] = (((float)I[r] + 0.5f) - P[r]) * g.dx;
                        // cols(T_grad_m) @ cols(T_init): M[r][c] = sum_k Tgm_k[r] Tin_c[k]
                        for (int r = 0; r < 3; ++r)
                            for (int c = 0; c < 3; ++c) {
                                float m = a.Tgm[0][r * n + p] * a.Tin[c][0 * n + p];
                                m = m + a.Tgm[1][r * n + p] * a.Tin[c][1 * n + p];
                                m = m + a.Tgm[2][r * n + p] * a.Tin[c][2 * n + p];
                                M[r][c] = m;
                            }
                        for (int r = 0; r < 3; ++r)
                            for (int c = 0; c < 4; ++c) {
                                float m = M[r][0] * a.grad[(0 + c) * n + p];
                                m = m + M[r][1] * a.grad[(4 + c) * n + p];
                                m = m + M[r][2] * a.grad[(8 + c) * n + p];
                                MG[r][c] = m;
                            }
                        for (int c = 0; c < 4; ++c) {
                            float d = dpos[0] * MG[0][c];
                            d = d + dpos[1] * MG[1][c];
                            delta[c] = d + dpos[2] * MG[2][c];
                        }
                    }
                    for (int c = 0; c < 4; ++c) acc[c] = acc[c] + (a.ps[c * n + p] + delta[c]) * w;
                }
            }
    weight[q] = wsum;
    for (int c = 0; c < 4; ++c) {
        float v = acc[c];
        if (wsum > 0.f) {
            v = v / wsum;
            if (v < 0.f) v = 0.0f;
            if (v > 1.f) v = 1.0f;
        }
        smoke[c * C + q] = v;
    }
}

// clamp_smoke_particle: each cell is clamped to the range of the initial
// smoke over the 3^3 cells around the initial cell of each of its (first
// cell_max, by particle id) particles.
__global__ void __launch_bounds__(kBlock) k_clamp_smoke_particle(GridDev g, int64_t n, int cell_max,
                                                                 const int32_t* __restrict__ start,
                                                                 const int32_t* __restrict__ perm,
                                                                 const float* __restrict__ init_pos,
                                                                 const float* __restrict__ init_smoke,
                                                                 float* __restrict__ smoke) {
    const int64_t C = (int64_t)g.res[0] * g.res[1] * g.res[2];
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= C) return;
    float mn[4] = {1000.f, 1000.f, 1000.f, 1000.f}, mx[4] = {-1000.f, -1000.f, -1000.f, -1000.f};
    int valid = 0;
    const int32_t s0 = start[q], m = min(start[q + 1] - s0, cell_max);
    for (int32_t s = s0; s < s0 + m; ++s) {
        const int64_t p = perm[s];
        int b[3];
        for (int r = 0; r < 3; ++r) b[r] = (int)(init_pos[r * n + p] * g.inv_dx);
        for (int oa = -1; oa < 2; ++oa)
            for (int ob = -1; ob < 2; ++ob)
                for (int oc = -1; oc < 2; ++oc) {
                    const int ca = b[0] + oa, cb = b[1] + ob, cc = b[2] + oc;
                    if (ca < 0 || ca >= g.res[0] || cb < 0 || cb >= g.res[1] || cc < 0 || cc >= g.res[2]) continue;
                    const int64_t nb = ((int64_t)ca * g.res[1] + cb) * g.res[2] + cc;
                    for (int c = 0; c < 4; ++c) {
                        const float v = init_smoke[c * C + nb];
                        mn[c] = fminf(mn[c], v);
                        mx[c] = fmaxf(mx[c], v);
                    }
                    ++valid;
                }
    }
    if (valid >= 1)
        for (int c = 0; c < 4; ++c) smoke[c * C + q] = fminf(fmaxf(smoke[c * C + q], mn[c]), mx[c]);
}

}  // namespace pfm

using namespace pfm;

extern "C" int pfm_g2p_smoke(const pfm_grid* gp, int64_t n, const float* pos, const float* smoke, float* p_smoke,
                             float* p_grad, pfm_stream_t stream) {
    GridDev g;
    int rc = make_grid(gp, &g);
    if (rc) return rc;
    PFM_CHECK_ARG(g.dim == 3, "particle smoke is 3D");
    PFM_CHECK_ARG(n >= 0 && (n == 0 || (pos && smoke && p_smoke && p_grad)), "null pointer");
    if (n == 0) return PFM_OK;
    hipLaunchKernelGGL(k_g2p_smoke, dim3(grid_for(n)), dim3(kBlock), 0, as_stream(stream), g, n, pos, smoke, p_smoke,
                       p_grad);
    PFM_LAUNCH_CHECK();
    return PFM_OK;
}

extern "C" int pfm_p2g_smoke(const pfm_grid* gp, int64_t n, const float* pos, const float* p_smoke,
                             const float* init_grad, const float* Tgm_x, const float* Tgm_y, const float* Tgm_z,
                             const float* Tin_x, const float* Tin_y, const float* Tin_z, int use_apic,
                             const int32_t* cell_start, const int32_t* perm, float* smoke, float* weight,
                             pfm_stream_t stream) {
    GridDev g;
    int rc = make_grid(gp, &g);
    if (rc) return rc;
    PFM_CHECK_ARG(g.dim == 3, "particle smoke is 3D");
    PFM_CHECK_ARG(cell_start && perm && smoke && weight && (n == 0 || (pos && p_smoke)), "null pointer");
    PFM_CHECK_ARG(!use_apic || (init_grad && Tgm_x && Tgm_y && Tgm_z && Tin_x && Tin_y && Tin_z),
                  "APIC smoke needs the gradient and both T");
    SmokeP2GArgs a = {pos, p_smoke, init_grad, {Tgm_x, Tgm_y, Tgm_z}, {Tin_x, Tin_y, Tin_z}, use_apic ? 1 : 0};
    const int64_t C = cell_count(g);
    hipLaunchKernelGGL(k_p2g_smoke, dim3(grid_for(C)), dim3(kBlock), 0, as_stream(stream), g, n, a, cell_start, perm,
                       smoke, weight);
    PFM_LAUNCH_CHECK();
    return PFM_OK;
}

extern "C" int pfm_clamp_smoke_particle(const pfm_grid* gp, int64_t n, const int32_t* cell_start,
                                        const int32_t* perm, int32_t cell_max, const float* init_pos,
                                        const float* init_smoke, float* smoke, pfm_stream_t stream) {
    GridDev g;
    int rc = make_grid(gp, &g);
    if (rc) return rc;
    PFM_CHECK_ARG(g.dim == 3, "particle smoke is 3D");
    PFM_CHECK_ARG(cell_start && perm && init_smoke && smoke && cell_max >= 1 && (n == 0 || init_pos),
                  "bad arguments");
    const int64_t C = cell_count(g);
    hipLaunchKernelGGL(k_clamp_smoke_particle, dim3(grid_for(C)), dim3(kBlock), 0, as_stream(stream), g, n,
                       (int)cell_max, cell_start, perm, init_pos, init_smoke, smoke);
    PFM_LAUNCH_CHECK();
    return PFM_OK;
}
