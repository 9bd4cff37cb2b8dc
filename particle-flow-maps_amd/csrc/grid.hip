// grid.hip -- face-centred kernels of the flow-map step: backward RK4 of the
// face flow map, semi-Lagrangian advection with the Jacobian, identity
// resets, CFL max speed, centre<->face averaging, vortex initial conditions
// and the kinetic-energy diagnostic.  One thread per face or cell.
#include "pfm_batch.h"

#include <mutex>

// Device scratch of pfm_kinetic_energy_noalloc: the FP64 per-block partial
// sums of the three face axes (3 x 1024 doubles = 24 KiB), held by every GPU
// the library runs on until the process ends.  Host code never passes this
// name to a kernel: it uses the per-device address from hipGetSymbolAddress.
static __device__ __align__(256) double scratch[3 * 1024];

namespace pfm {

// Decode a flat face index of axis `a` into (i, j, k) (last axis fastest).
template <int DIM>
__device__ __forceinline__ void face_ijk(const GridDev& g, int a, int64_t f, int I[3]) {
    int S[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) S[c] = g.res[c] + (c == a ? 1 : 0);
    if (DIM == 3) {
        I[2] = (int)(f % S[2]);
        int64_t q = f / S[2];
        I[1] = (int)(q % S[1]);
        I[0] = (int)(q / S[1]);
    } else {
        I[2] = 0;
        I[1] = (int)(f % S[1]);
        I[0] = (int)(f / S[1]);
    }
}

// RK4_grid: 3D/run_3D.py:144-173, 2D/run_2D.py:233-261.  dT = G @ T (rows form).
// Face f of one axis; psi and T are (d, nf) with component stride nf.
// rk4_face_reg: the same on values (p0, T0 in; p1, T1 out).
template <int DIM>
__device__ __forceinline__ void rk4_face_reg(const GridDev& g, const float p0[3], const float T0[3],
                                             const float* __restrict__ ux, const float* __restrict__ uy,
                                             const float* __restrict__ uz, float dt, float p1[3], float T1[3],
                                             RowCache<DIM>& rows) {
    const float neg_dt = -1.0f * dt;
    const float hs[3] = {0.5f * neg_dt, 0.5f * neg_dt, 1.0f * neg_dt};
    const float ws[4] = {1.f, 2.f, 2.f, 1.f};
    float pk[3] = {p0[0], p0[1], p0[2]}, Tk[3] = {T0[0], T0[1], T0[2]};
    float us[3] = {0.f, 0.f, 0.f}, dTs[3] = {0.f, 0.f, 0.f};
    // the stages reuse a component's rows where its stencil did not move
    // (interp_mac_cached; bit-identical to interp_mac); `rows` arrives cleared
    // and leaves holding the last stage's rows
#pragma unroll 1
    for (int st = 0; st < 4; ++st) {
        float u[3], G[3][3], dT[3];
        interp_mac_cached<DIM, true>(ux, uy, uz, g, pk, u, G, rows);
#pragma unroll
        for (int r = 0; r < DIM; ++r) {
            float acc = G[r][0] * Tk[0];
#pragma unroll
            for (int c = 1; c < DIM; ++c) acc += G[r][c] * Tk[c];
            dT[r] = acc;
        }
#pragma unroll
        for (int c = 0; c < DIM; ++c) {
            us[c] += ws[st] * u[c];
            dTs[c] += ws[st] * dT[c];
        }
        if (st < 3) {
#pragma unroll
            for (int c = 0; c < DIM; ++c) {
                pk[c] = p0[c] + hs[st] * u[c];
                Tk[c] = T0[c] + hs[st] * dT[c];
            }
        }
    }
    const float sixth = neg_dt * 1.0f / 6.0f;
#pragma unroll
    for (int c = 0; c < DIM; ++c) {
        p1[c] = p0[c] + sixth * us[c];
        T1[c] = T0[c] + sixth * dTs[c];
    }
}

template <int DIM>
__device__ __forceinline__ void rk4_face(const GridDev& g, int64_t f, int64_t nf, float* __restrict__ psi,
                                         float* __restrict__ T, const float* __restrict__ ux,
                                         const float* __restrict__ uy, const float* __restrict__ uz, float dt) {
    float p0[3] = {0.f, 0.f, 0.f}, T0[3] = {0.f, 0.f, 0.f}, p1[3] = {0.f, 0.f, 0.f}, T1[3] = {0.f, 0.f, 0.f};
#pragma unroll
    for (int c = 0; c < DIM; ++c) {
        p0[c] = psi[c * nf + f];
        T0[c] = T[c * nf + f];
    }
    RowCache<DIM> rows;
    rows.clear();
    rk4_face_reg<DIM>(g, p0, T0, ux, uy, uz, dt, p1, T1, rows);
#pragma unroll
    for (int c = 0; c < DIM; ++c) {
        psi[c * nf + f] = p1[c];
        T[c * nf + f] = T1[c];
    }
}

template <int DIM>
__global__ void __launch_bounds__(kBlock, 3) k_rk4_grid(GridDev g, int64_t nf,
                                                     float* __restrict__ psi,
                                                     float* __restrict__ T,
                                                     const float* __restrict__ ux,
                                                     const float* __restrict__ uy,
                                                     const float* __restrict__ uz, float dt) {
    const int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nf) return;
    rk4_face<DIM>(g, f, nf, psi, T, ux, uy, uz, dt);
}

// advect_u: 3D/run_3D.py:296-312 (one launch covers the faces of all axes).
struct AdvectArgs {
    const float* u0[3];
    float* u1[3];
    const float* T[3];
    const float* psi[3];
    int64_t nf[3];
};

// Face f of axis ax: psi_ax / T_ax have component stride `stride`.
template <int DIM>
__device__ __forceinline__ float advect_face_reg(const GridDev& g, const float* const u0[3], const float p[3],
                                                 const float Tv[3]) {
    float u[3], G[3][3];
    interp_mac<DIM, false>(u0[0], u0[1], u0[2], g, p, u, G);
    float acc = Tv[0] * u[0];
#pragma unroll
    for (int c = 1; c < DIM; ++c) acc += Tv[c] * u[c];
    return acc;
}

// 3D advect_face_reg with the rows the RK4 stages left in `rows` (over the
// same field u0) serving the sample where the stencil did not move.
__device__ __forceinline__ float advect_face_reg3_cached(const GridDev& g, const float* const u0[3], const float p[3],
                                                         const float Tv[3], RowCache<3>& rows) {
    float u[3], G[3][3];
    interp_mac_cached<3, false>(u0[0], u0[1], u0[2], g, p, u, G, rows);
    float acc = Tv[0] * u[0];
#pragma unroll
    for (int c = 1; c < 3; ++c) acc += Tv[c] * u[c];
    return acc;
}

template <int DIM>
__device__ __forceinline__ void advect_face(const GridDev& g, const float* const u0[3], const float* __restrict__ psi,
                                            const float* __restrict__ T, int64_t f, int64_t stride,
                                            float* __restrict__ u1) {
    float p[3] = {0.f, 0.f, 0.f}, Tv[3] = {0.f, 0.f, 0.f};
#pragma unroll
    for (int c = 0; c < DIM; ++c) {
        p[c] = psi[c * stride + f];
        Tv[c] = T[c * stride + f];
    }
    u1[f] = advect_face_reg<DIM>(g, u0, p, Tv);
}

template <int DIM>
__global__ void __launch_bounds__(kBlock, 4) k_advect_u(GridDev g, AdvectArgs a) {
    int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int ax = 0;
    while (ax < DIM - 1 && f >= a.nf[ax]) {
        f -= a.nf[ax];
        ++ax;
    }
    if (f >= a.nf[ax]) return;
    advect_face<DIM>(g, a.u0, a.psi[ax], a.T[ax], f, a.nf[ax], a.u1[ax]);
}

// reset_to_identity_grid with x/y/z_coords_func: 3D/run_3D.py:106-120,
// 3D/taichi_utils.py:33-46: psi = (I + 0.5 off-axis) * dx, T = e_axis.
// Face f of the grid's array of `axis`, stored at index `at` of psi / T
// (component stride nf).
template <int DIM>
__device__ __forceinline__ void identity_face(const GridDev& g, int axis, int64_t f, float p[3], float Tv[3]) {
    int I[3];
    face_ijk<DIM>(g, axis, f, I);
#pragma unroll
    for (int c = 0; c < DIM; ++c) {
        float x = c == axis ? (float)I[c] : __fadd_rn((float)I[c], 0.5f);
        p[c] = __fmul_rn(x, g.dx);
        Tv[c] = c == axis ? 1.0f : 0.0f;
    }
}

template <int DIM>
__device__ __forceinline__ void reset_identity_face(const GridDev& g, int axis, int64_t f, int64_t at, int64_t nf,
                                                    float* __restrict__ psi, float* __restrict__ T) {
    float p[3] = {0.f, 0.f, 0.f}, Tv[3] = {0.f, 0.f, 0.f};
    identity_face<DIM>(g, axis, f, p, Tv);
#pragma unroll
    for (int c = 0; c < DIM; ++c) {
        psi[c * nf + at] = p[c];
        T[c * nf + at] = Tv[c];
    }
}

// The midpoint velocity's grid part (3D/run_3D.py:778-784): per face,
// reset_to_identity_grid, RK4_grid over `dt` and advect_u, in registers.
// Every face samples u0 (the copy of u), so u1 can be the live u.  psi and
// T are still written, as the separate kernels leave them.
struct MidOut {
    float* psi[3];
    float* T[3];
};

template <int DIM>
__global__ void __launch_bounds__(kBlock, 3) k_midpoint_advect(GridDev g, AdvectArgs a, MidOut o, float dt) {
    int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int ax = 0;
    while (ax < DIM - 1 && f >= a.nf[ax]) {
        f -= a.nf[ax];
        ++ax;
    }
    if (f >= a.nf[ax]) return;
    float p0[3] = {0.f, 0.f, 0.f}, T0[3] = {0.f, 0.f, 0.f}, p1[3] = {0.f, 0.f, 0.f}, T1[3] = {0.f, 0.f, 0.f};
    identity_face<DIM>(g, ax, f, p0, T0);
    RowCache<DIM> rows;
    rows.clear();
    rk4_face_reg<DIM>(g, p0, T0, a.u0[0], a.u0[1], a.u0[2], dt, p1, T1, rows);
    const int64_t nf = a.nf[ax];
#pragma unroll
    for (int c = 0; c < DIM; ++c) {
        o.psi[ax][c * nf + f] = p1[c];
        o.T[ax][c * nf + f] = T1[c];
    }
    if constexpr (DIM == 3)
        a.u1[ax][f] = advect_face_reg3_cached(g, a.u0, p1, T1, rows);
    else
        a.u1[ax][f] = advect_face_reg<DIM>(g, a.u0, p1, T1);
}

template <int DIM>
__global__ void __launch_bounds__(kBlock) k_reset_identity_grid(GridDev g, int axis, int64_t nf,
                                                                float* __restrict__ psi,
                                                                float* __restrict__ T) {
    const int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nf) return;
    reset_identity_face<DIM>(g, axis, f, f, nf, psi, T);
}

// |u| at the centre of cell (i, j, k) (calc_max_speed, 3D/run_3D.py:95-103).
template <int DIM>
__device__ __forceinline__ float cell_speed(const GridDev& g, const float* __restrict__ ux,
                                            const float* __restrict__ uy, const float* __restrict__ uz, int i, int j,
                                            int k) {
    float s2;
    if (DIM == 3) {
        const int ry = g.res[1], rz = g.res[2];
        float u = 0.5f * (ux[((int64_t)i * ry + j) * rz + k] + ux[((int64_t)(i + 1) * ry + j) * rz + k]);
        float v = 0.5f * (uy[((int64_t)i * (ry + 1) + j) * rz + k] + uy[((int64_t)i * (ry + 1) + j + 1) * rz + k]);
        float w = 0.5f * (uz[((int64_t)i * ry + j) * (rz + 1) + k] + uz[((int64_t)i * ry + j) * (rz + 1) + k + 1]);
        s2 = __fadd_rn(__fadd_rn(__fmul_rn(u, u), __fmul_rn(v, v)), __fmul_rn(w, w));
    } else {
        const int ry = g.res[1];
        float u = 0.5f * (ux[(int64_t)i * ry + j] + ux[(int64_t)(i + 1) * ry + j]);
        float v = 0.5f * (uy[(int64_t)i * (ry + 1) + j] + uy[(int64_t)i * (ry + 1) + j + 1]);
        s2 = __fadd_rn(__fmul_rn(u, u), __fmul_rn(v, v));
    }
    return __fsqrt_rn(s2);
}

// calc_max_speed: 3D/run_3D.py:95-103.  Max over non-negative floats is exact
// and order independent, so unsigned atomicMax on the bit patterns is exact.
template <int DIM>
__global__ void __launch_bounds__(kBlock) k_max_speed(GridDev g, const float* __restrict__ ux,
                                                      const float* __restrict__ uy,
                                                      const float* __restrict__ uz,
                                                      unsigned int* __restrict__ out) {
    const int64_t C = (int64_t)g.res[0] * g.res[1] * g.res[2];
    float m = 0.0f;
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < C;
         q += (int64_t)gridDim.x * blockDim.x) {
        int i, j, k;
        if (DIM == 3) {
            k = (int)(q % g.res[2]);
            j = (int)((q / g.res[2]) % g.res[1]);
            i = (int)(q / ((int64_t)g.res[2] * g.res[1]));
        } else {
            k = 0;
            j = (int)(q % g.res[1]);
            i = (int)(q / g.res[1]);
        }
        m = fmaxf(m, cell_speed<DIM>(g, ux, uy, uz, i, j, k));
    }
    m = wave_max(m);
    if ((threadIdx.x & 63) == 0) atomicMax(out, __float_as_uint(m));
}

__global__ void k_set_u32(unsigned int* p, unsigned int v) { *p = v; }

// ---- batched 2D face and cell kernels ------------------------------------------
// A face unit covers kBlock faces of one scene, numbered over both axes
// (x faces first, as in k_advect_u).
struct BFace {
    int ax;
    int64_t f;  // face within the scene's array of axis ax
};

__device__ __forceinline__ bool face_of_unit(const BScene& S, int32_t chunk, BFace& bf) {
    int64_t f = (int64_t)chunk * kBlock + threadIdx.x;
    bf.ax = 0;
    if (f >= S.nf[0]) {
        f -= S.nf[0];
        bf.ax = 1;
    }
    bf.f = f;
    return f < S.nf[bf.ax];
}

__global__ void __launch_bounds__(kBlock) kb_max_speed(BUnits t, const float* __restrict__ ux,
                                                       const float* __restrict__ uy, BState* __restrict__ st) {
    PFM_FOR_UNITS(t, u) {
        const BUnit w = t.u[u];
        if (!unit_on(t, w.scene)) continue;
        const BScene& S = t.sc[w.scene];
        const int64_t q = (int64_t)w.chunk * kBlock + threadIdx.x;
        float m = 0.0f;
        if (q < S.nc) {
            const int j = (int)(q % S.g.res[1]), i = (int)(q / S.g.res[1]);
            m = cell_speed<2>(S.g, ux + S.f_off[0], uy + S.f_off[1], nullptr, i, j, 0);
        }
        m = wave_max(m);
        if ((threadIdx.x & 63) == 0) atomicMax(&st[w.scene].ms_bits, __float_as_uint(m));
    }
}

__global__ void kb_ms_init(int n, BState* st) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < n) st[s].ms_bits = __float_as_uint(1.e-3f);  // max_speed[None] = 1.e-3 (3D/run_3D.py:97)
}

// curr_dt = CFL * dx / max_speed and the frame clip (2D/run_2D.py:782-800), in
// double as the host driver computes it.
__global__ void kb_dt(int n, const BScene* __restrict__ sc, BState* __restrict__ st, double CFL, double vis_dt) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n || !st[s].enabled) return;
    BState& b = st[s];
    const double dx = 1.0 / (double)sc[s].g.res[1];
    double curr = CFL * dx / (double)__uint_as_float(b.ms_bits);
    if (b.sub_t + curr >= vis_dt) {
        curr = vis_dt - b.sub_t;
        b.sub_t = 0.0;
        b.frame += 1;
    } else {
        b.sub_t += curr;
    }
    b.dt = curr;
}

int b_max_speed(BUnits cells, const float* ux, const float* uy, BState* st, unsigned blocks, hipStream_t s) {
    hipLaunchKernelGGL(kb_max_speed, dim3(blocks), dim3(kBlock), 0, s, cells, ux, uy, st);
    PFM_LAUNCH_CHECK();
    return PFM_OK;
}

int b_ms_init(int n_scenes, BState* st, hipStream_t s) {
    hipLaunchKernelGGL(kb_ms_init, dim3(grid_for(n_scenes)), dim3(kBlock), 0, s, n_scenes, st);
    PFM_LAUNCH_CHECK();
    return PFM_OK;
}

int b_dt(int n_scenes, const BScene* sc, BState* st, double CFL, double visualize_dt, hipStream_t s) {
    hipLaunchKernelGGL(kb_dt, dim3(grid_for(n_scenes)), dim3(kBlock), 0, s, n_scenes, sc, st, CFL, visualize_dt);
    PFM_LAUNCH_CHECK();
    return PFM_OK;
}

// Batched k_midpoint_advect: every face samples u0 (the copy of u) for
// RK4_grid over half its scene's dt and for advect_u, writes psi / T and the
// advected velocity into u1.
__global__ void __launch_bounds__(kBlock, 4) kb_midpoint_advect(BUnits t, int64_t F0, int64_t F1,
                                                             const float* __restrict__ u0x,
                                                             const float* __restrict__ u0y, float* __restrict__ u1x,
                                                             float* __restrict__ u1y, float* __restrict__ psi_x,
                                                             float* __restrict__ psi_y, float* __restrict__ T_x,
                                                             float* __restrict__ T_y, const BState* __restrict__ st) {
    PFM_FOR_UNITS(t, u) {
        const BUnit w = t.u[u];
        if (!unit_on(t, w.scene)) continue;
        const BScene& S = t.sc[w.scene];
        BFace bf;
        if (!face_of_unit(S, w.chunk, bf)) continue;
        const float dt = (float)(0.5 * st[w.scene].dt);
        const float* u0[3] = {u0x + S.f_off[0], u0y + S.f_off[1], nullptr};
        const int64_t gf = S.f_off[bf.ax] + bf.f, nf = bf.ax == 0 ? F0 : F1;
        float p0[3] = {0.f, 0.f, 0.f}, T0[3] = {0.f, 0.f, 0.f}, p1[3] = {0.f, 0.f, 0.f}, T1[3] = {0.f, 0.f, 0.f};
        identity_face<2>(S.g, bf.ax, bf.f, p0, T0);
        RowCache<2> rows;
        rows.clear();
        rk4_face_reg<2>(S.g, p0, T0, u0[0], u0[1], nullptr, dt, p1, T1, rows);
        float* psi = bf.ax == 0 ? psi_x : psi_y;
        float* T = bf.ax == 0 ? T_x : T_y;
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            psi[c * nf + gf] = p1[c];
            T[c * nf + gf] = T1[c];
        }
        (bf.ax == 0 ? u1x : u1y)[gf] = advect_face_reg<2>(S.g, u0, p1, T1);
    }
}

int b_midpoint_advect(BUnits faces, const int64_t F[2], const float* u0x, const float* u0y, float* u1x, float* u1y,
                      float* psi_x, float* psi_y, float* T_x, float* T_y, const BState* st, unsigned blocks,
                      hipStream_t s) {
    hipLaunchKernelGGL(kb_midpoint_advect, dim3(blocks), dim3(kBlock), 0, s, faces, F[0], F[1], u0x, u0y, u1x, u1y,
                       psi_x, psi_y, T_x, T_y, st);
    PFM_LAUNCH_CHECK();
    return PFM_OK;
}

// split_central_vector: 3D/taichi_utils.py:87-100 (clamped samples).
template <int DIM>
__global__ void __launch_bounds__(kBlock) k_split_central(GridDev g, int axis, int64_t nf,
                                                          const float* __restrict__ vc,
                                                          float* __restrict__ u) {
    const int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nf) return;
    int I[3];
    face_ijk<DIM>(g, axis, f, I);
    const int64_t C = (int64_t)g.res[0] * g.res[1] * g.res[2];
    int R[3] = {I[0], I[1], I[2]}, L[3] = {I[0], I[1], I[2]};
    L[axis] -= 1;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        R[c] = max(0, min(R[c], g.res[c] - 1));
        L[c] = max(0, min(L[c], g.res[c] - 1));
    }
    const float* v = vc + axis * C;
    float r = v[((int64_t)R[0] * g.res[1] + R[1]) * g.res[2] + R[2]];
    float l = v[((int64_t)L[0] * g.res[1] + L[1]) * g.res[2] + L[2]];
    u[f] = 0.5f * __fadd_rn(r, l);
}

// get_central_vector: 3D/taichi_utils.py:80-85.
template <int DIM>
__global__ void __launch_bounds__(kBlock) k_get_central(GridDev g, const float* __restrict__ ux,
                                                        const float* __restrict__ uy,
                                                        const float* __restrict__ uz,
                                                        float* __restrict__ vc) {
    const int64_t C = (int64_t)g.res[0] * g.res[1] * g.res[2];
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= C) return;
    const int rx = g.res[0], ry = g.res[1], rz = g.res[2];
    const int k = (int)(q % rz);
    const int j = (int)((q / rz) % ry);
    const int i = (int)(q / ((int64_t)rz * ry));
    (void)rx;
    vc[q] = 0.5f * __fadd_rn(ux[((int64_t)(i + 1) * ry + j) * rz + k], ux[((int64_t)i * ry + j) * rz + k]);
    vc[C + q] = 0.5f * __fadd_rn(uy[((int64_t)i * (ry + 1) + j + 1) * rz + k], uy[((int64_t)i * (ry + 1) + j) * rz + k]);
    if (DIM == 3)
        vc[2 * C + q] = 0.5f * __fadd_rn(uz[((int64_t)i * ry + j) * (rz + 1) + k + 1], uz[((int64_t)i * ry + j) * (rz + 1) + k]);
}

// curl: 3D/taichi_utils.py:56-78 -- central differences of the cell-centred
// vector vc (component-major, (3, cells)) with sample() clamping and
// inv = float(1 / (2 dx)); 2D/taichi_utils.py:97-103 -- the scalar
// (vr.y - vl.y - vt.x + vb.x) / (2 dx).  No contraction: every product and
// difference rounds as in the reference.
template <int DIM>
__global__ void __launch_bounds__(kBlock) k_curl(GridDev g, const float* __restrict__ vc, float inv, float two_dx,
                                                 float* __restrict__ cf) {
#pragma clang fp contract(off)
    const int64_t C = (int64_t)g.res[0] * g.res[1] * g.res[2];
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= C) return;
    const int ry = g.res[1], rz = g.res[2];
    const int k = (int)(q % rz);
    const int j = (int)((q / rz) % ry);
    const int i = (int)(q / ((int64_t)rz * ry));
    auto at = [&](int c, int ii, int jj, int kk) {
        ii = min(max(ii, 0), g.res[0] - 1);
        jj = min(max(jj, 0), g.res[1] - 1);
        kk = min(max(kk, 0), g.res[2] - 1);
        return vc[c * C + ((int64_t)ii * ry + jj) * rz + kk];
    };
    if (DIM == 3) {
        const float d_vx_dz = inv * (at(0, i, j, k + 1) - at(0, i, j, k - 1));
        const float d_vx_dy = inv * (at(0, i, j + 1, k) - at(0, i, j - 1, k));
        const float d_vy_dx = inv * (at(1, i + 1, j, k) - at(1, i - 1, j, k));
        const float d_vy_dz = inv * (at(1, i, j, k + 1) - at(1, i, j, k - 1));
        const float d_vz_dx = inv * (at(2, i + 1, j, k) - at(2, i - 1, j, k));
        const float d_vz_dy = inv * (at(2, i, j + 1, k) - at(2, i, j - 1, k));
        cf[q] = d_vz_dy - d_vy_dz;
        cf[C + q] = d_vx_dz - d_vz_dx;
        cf[2 * C + q] = d_vy_dx - d_vx_dy;
    } else {
        const float vr_y = at(1, i + 1, j, 0), vl_y = at(1, i - 1, j, 0);
        const float vt_x = at(0, i, j + 1, 0), vb_x = at(0, i, j - 1, 0);
        cf[q] = (((vr_y - vl_y) - vt_x) + vb_x) / two_dx;
    }
}

// Smoke transport on the grid: advect_smoke (3D/run_3D.py:315-323,
// 2D/run_2D.py:444-450) and clamp_smoke (3D/run_3D.py:325-334).  The cell's
// backward-map point is the mean of the map on its 2d faces (psi_c), the
// smoke is interpolated there with interp_1 (3D/taichi_utils.py:193-220,
// 2D/taichi_utils.py:119-137: linear hat weights, the clamp s = max(lo,
// min(u, dim-1-eps)) with lo = 1 in 3D and 0 in 2D) and, for clamp_smoke,
// clamped to the range of the 2^d samples around that point
// (sample_min_max_1, 3D/taichi_utils.py:222-247).  Smoke fields are
// channel-major (nch, cells).  mode bit 0: advect into s1; bit 1: clamp s1.
// No contraction: every operation rounds as in the reference.
struct SmokeArgs {
    int nch, mode;
    float hi[3];     // float(extent - 1 - eps) per axis
    float lo;        // 1 (3D) or 0 (2D)
    float third;     // the psi_c weight: float(1/6) in 3D, 0.25 in 2D
};

template <int DIM>
__global__ void __launch_bounds__(kBlock) k_advect_smoke(GridDev g, SmokeArgs a, const float* __restrict__ s0,
                                                         float* __restrict__ s1, const float* __restrict__ px,
                                                         const float* __restrict__ py,
                                                         const float* __restrict__ pz) {
#pragma clang fp contract(off)
    const int64_t C = (int64_t)g.res[0] * g.res[1] * g.res[2];
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= C) return;
    const int rx = g.res[0], ry = g.res[1], rz = g.res[2];
    float pc[3] = {0.f, 0.f, 0.f};
    if (DIM == 3) {
        const int k = (int)(q % rz), j = (int)((q / rz) % ry), i = (int)(q / ((int64_t)rz * ry));
        const int64_t Fx = (int64_t)(rx + 1) * ry * rz, Fy = (int64_t)rx * (ry + 1) * rz, Fz = (int64_t)rx * ry * (rz + 1);
        const int64_t x0 = ((int64_t)i * ry + j) * rz + k, x1 = x0 + (int64_t)ry * rz;
        const int64_t y0 = ((int64_t)i * (ry + 1) + j) * rz + k, y1 = y0 + rz;
        const int64_t z0 = ((int64_t)i * ry + j) * (rz + 1) + k, z1 = z0 + 1;
        for (int c = 0; c < 3; ++c)
            pc[c] = a.third * (((((px[c * Fx + x0] + px[c * Fx + x1]) + py[c * Fy + y0]) + py[c * Fy + y1]) +
                                pz[c * Fz + z0]) + pz[c * Fz + z1]);
    } else {
        // 2D cell (i, j); u_x faces (rx + 1, ry), u_y faces (rx, ry + 1)
        const int j = (int)(q % ry), i = (int)(q / ry);
        const int64_t Fx = (int64_t)(rx + 1) * ry, Fy = (int64_t)rx * (ry + 1);
        const int64_t x0 = (int64_t)i * ry + j, x1 = x0 + ry;
        const int64_t y0 = (int64_t)i * (ry + 1) + j, y1 = y0 + 1;
        for (int c = 0; c < 2; ++c)
            pc[c] = a.third * (((px[c * Fx + x0] + px[c * Fx + x1]) + py[c * Fy + y0]) + py[c * Fy + y1]);
    }
    const int ext[3] = {rx, ry, rz};
    float sv[3], fl[3];
    for (int c = 0; c < DIM; ++c) {
        const float u = pc[c] * g.inv_dx - 0.5f;
        sv[c] = fmaxf(a.lo, fminf(u, a.hi[c]));
        fl[c] = floorf(sv[c]);
    }
    // sample(): int() of the float index, clamped into the field
    auto idx = [&](int c, int o) { return min(max((int)(fl[c] + (float)o), 0), ext[c] - 1); };
    auto hat = [](float x) { return 1.0f - fabsf(x); };
    if (DIM == 3) {
        int64_t off[8];
        for (int i = 0; i < 2; ++i)
            for (int j = 0; j < 2; ++j)
                for (int k = 0; k < 2; ++k) off[(i * 2 + j) * 2 + k] = ((int64_t)idx(0, i) * ry + idx(1, j)) * rz + idx(2, k);
        const float nx[2] = {hat(sv[0] - (fl[0] + 0.0f)), hat(sv[0] - (fl[0] + 1.0f))};
        const float ny[2] = {hat(sv[1] - (fl[1] + 0.0f)), hat(sv[1] - (fl[1] + 1.0f))};
        const float nz[2] = {hat(sv[2] - (fl[2] + 0.0f)), hat(sv[2] - (fl[2] + 1.0f))};
        for (int ch = 0; ch < a.nch; ++ch) {
            const float* f = s0 + (int64_t)ch * C;
            const float v0 = f[off[0]];
            if (a.mode & 1) {
                float acc = 0.0f * v0;
                for (int e = 0; e < 8; ++e) acc = acc + ((f[off[e]] * nx[e >> 2]) * ny[(e >> 1) & 1]) * nz[e & 1];
                s1[(int64_t)ch * C + q] = acc;
            }
            if (a.mode & 2) {
                float mn = v0, mx = v0;
                for (int e = 0; e < 8; ++e) {
                    mn = fminf(mn, f[off[e]]);
                    mx = fmaxf(mx, f[off[e]]);
                }
                const float cur = s1[(int64_t)ch * C + q];
                s1[(int64_t)ch * C + q] = fminf(fmaxf(cur, mn), mx);
            }
        }
    } else {
        const int64_t oa = (int64_t)idx(0, 0) * ry + idx(1, 0), ob = (int64_t)idx(0, 1) * ry + idx(1, 0);
        const int64_t oc = (int64_t)idx(0, 0) * ry + idx(1, 1), od = (int64_t)idx(0, 1) * ry + idx(1, 1);
        const float fu = sv[0] - fl[0], fv = sv[1] - fl[1];
        const float nu0 = hat(fu), nu1 = hat(fu - 1.0f), nv0 = hat(fv), nv1 = hat(fv - 1.0f);
        for (int ch = 0; ch < a.nch; ++ch) {
            const float* f = s0 + (int64_t)ch * C;
            s1[(int64_t)ch * C + q] =
                (((f[oa] * nu0) * nv0 + (f[ob] * nu1) * nv0) + (f[oc] * nu0) * nv1) + (f[od] * nu1) * nv1;
        }
    }
}

// Point vortices: angular_vel_func + leapfrog_vel_func, 2D/init_conditions.py:5-9,29-60.
// The vortices travel by value in the kernel arguments, kVortChunk per
// launch; a launch after the first continues the running sums stored in vc,
// so the summation order (vortex 0, 1, ...) is that of a single loop.
constexpr int kMaxVort = 4096;
constexpr int kVortChunk = 128;
struct VortChunk {
    float x[kVortChunk], y[kVortChunk], w[kVortChunk];
};

__global__ void __launch_bounds__(kBlock) k_point_vortices(GridDev g, VortChunk vt, int nv, bool first, float rad2,
                                                           float strength, float* __restrict__ vc) {
    const int64_t C = (int64_t)g.res[0] * g.res[1];
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= C) return;
    const int j = (int)(q % g.res[1]);
    const int i = (int)(q / g.res[1]);
    const float X0 = __fmul_rn(__fadd_rn((float)i, 0.5f), g.dx);
    const float X1 = __fmul_rn(__fadd_rn((float)j, 0.5f), g.dx);
    float vx = first ? 0.f : vc[q], vy = first ? 0.f : vc[C + q];
    for (int v = 0; v < nv; ++v) {
        float px = __fsub_rn(X0, vt.x[v]);
        float py = __fsub_rn(X1, vt.y[v]);
        float r = __fsqrt_rn(__fadd_rn(__fmul_rn(px, px), __fmul_rn(py, py)));
        r = __fadd_rn(r, 1e-6f);
        float lin = __fmul_rn(__fdiv_rn(__fmul_rn(strength, 1.0f), r),
                              __fsub_rn(1.0f, expf(__fdiv_rn(-__fmul_rn(r, r), rad2))));
        float ang = __fmul_rn(__fdiv_rn(1.0f, r), lin);
        float s = __fmul_rn(ang, vt.w[v]);
        const bool init = first && v == 0;
        vx = init ? __fmul_rn(s, -py) : __fadd_rn(vx, __fmul_rn(s, -py));
        vy = init ? __fmul_rn(s, px) : __fadd_rn(vy, __fmul_rn(s, px));
    }
    vc[q] = vx;
    vc[C + q] = vy;
}

// add_vortex_ring: 3D/init_conditions.py:10-22 (smoothed Biot-Savart of a
// sampled ring, added into the cell-centred velocity).
struct RingArgs {
    float w, rad, delta, curve_length, four_pi, pi;
    float c[3], ex[3], ey[3];
    int num_samples;
};

__global__ void __launch_bounds__(kBlock) k_vortex_ring(GridDev g, RingArgs a, float* __restrict__ vc) {
    const int64_t C = (int64_t)g.res[0] * g.res[1] * g.res[2];
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= C) return;
    const int k = (int)(q % g.res[2]);
    const int j = (int)((q / g.res[2]) % g.res[1]);
    const int i = (int)(q / ((int64_t)g.res[2] * g.res[1]));
    const float P[3] = {__fmul_rn(__fadd_rn((float)i, 0.5f), g.dx),
                        __fmul_rn(__fadd_rn((float)j, 0.5f), g.dx),
                        __fmul_rn(__fadd_rn((float)k, 0.5f), g.dx)};
    float acc[3] = {vc[q], vc[C + q], vc[2 * C + q]};
    for (int l = 0; l < a.num_samples; ++l) {
        float theta = __fmul_rn(__fmul_rn(__fdiv_rn((float)l, (float)a.num_samples), 2.0f), a.pi);
        float ct = cosf(theta), st = sinf(theta);
        float pd[3], wv[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            float ps = __fadd_rn(__fmul_rn(a.rad, __fadd_rn(__fmul_rn(ct, a.ex[c]), __fmul_rn(st, a.ey[c]))), a.c[c]);
            pd[c] = __fsub_rn(P[c], ps);
            wv[c] = __fmul_rn(a.w, __fadd_rn(__fmul_rn(-st, a.ex[c]), __fmul_rn(ct, a.ey[c])));
        }
        float r = __fsqrt_rn(__fadd_rn(__fadd_rn(__fmul_rn(pd[0], pd[0]), __fmul_rn(pd[1], pd[1])), __fmul_rn(pd[2], pd[2])));
        float rd = __fdiv_rn(r, a.delta);
        float r3 = __fmul_rn(__fmul_rn(r, r), r);
        float rd3 = __fmul_rn(__fmul_rn(rd, rd), rd);
        float fac = __fmul_rn(a.curve_length,
                              __fmul_rn(__fdiv_rn(-1.0f, __fmul_rn(a.four_pi, r3)), __fsub_rn(1.0f, expf(-rd3))));
        float cr[3] = {__fsub_rn(__fmul_rn(pd[1], wv[2]), __fmul_rn(pd[2], wv[1])),
                       __fsub_rn(__fmul_rn(pd[2], wv[0]), __fmul_rn(pd[0], wv[2])),
                       __fsub_rn(__fmul_rn(pd[0], wv[1]), __fmul_rn(pd[1], wv[0]))};
#pragma unroll
        for (int c = 0; c < 3; ++c) acc[c] = __fadd_rn(acc[c], __fmul_rn(fac, cr[c]));
    }
    vc[q] = acc[0];
    vc[C + q] = acc[1];
    vc[2 * C + q] = acc[2];
}

// The smoke part of add_vortex_ring_and_smoke (3D/init_conditions.py:31-56):
// channel 3 accumulates curve_length * exp(-(r/delta)^3) over the ring's
// samples (the velocity part is k_vortex_ring); k_smoke_threshold then sets
// every cell to (color, 1) where channel 3 exceeds 0.002 and to 0 elsewhere.
__global__ void __launch_bounds__(kBlock) k_ring_smoke(GridDev g, RingArgs a, float* __restrict__ smoke) {
    const int64_t C = (int64_t)g.res[0] * g.res[1] * g.res[2];
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= C) return;
    const int k = (int)(q % g.res[2]);
    const int j = (int)((q / g.res[2]) % g.res[1]);
    const int i = (int)(q / ((int64_t)g.res[2] * g.res[1]));
    const float P[3] = {__fmul_rn(__fadd_rn((float)i, 0.5f), g.dx),
                        __fmul_rn(__fadd_rn((float)j, 0.5f), g.dx),
                        __fmul_rn(__fadd_rn((float)k, 0.5f), g.dx)};
    float acc = smoke[3 * C + q];
    for (int l = 0; l < a.num_samples; ++l) {
        float theta = __fmul_rn(__fmul_rn(__fdiv_rn((float)l, (float)a.num_samples), 2.0f), a.pi);
        float ct = cosf(theta), st = sinf(theta);
        float pd[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            float ps = __fadd_rn(__fmul_rn(a.rad, __fadd_rn(__fmul_rn(ct, a.ex[c]), __fmul_rn(st, a.ey[c]))), a.c[c]);
            pd[c] = __fsub_rn(P[c], ps);
        }
        float r = __fsqrt_rn(__fadd_rn(__fadd_rn(__fmul_rn(pd[0], pd[0]), __fmul_rn(pd[1], pd[1])), __fmul_rn(pd[2], pd[2])));
        float rd = __fdiv_rn(r, a.delta);
        float rd3 = __fmul_rn(__fmul_rn(rd, rd), rd);
        acc = __fadd_rn(acc, __fmul_rn(a.curve_length, expf(-rd3)));
    }
    smoke[3 * C + q] = acc;
}

__global__ void __launch_bounds__(kBlock) k_smoke_threshold(int64_t C, float c0, float c1, float c2,
                                                            float* __restrict__ smoke) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= C) return;
    const bool on = smoke[3 * C + q] > 0.002f;
    smoke[q] = on ? c0 : 0.f;
    smoke[C + q] = on ? c1 : 0.f;
    smoke[2 * C + q] = on ? c2 : 0.f;
    smoke[3 * C + q] = on ? 1.0f : 0.f;
}

// Kinetic energy (2D/run_2D.py:765-770): per-axis double sums of u^2,
// reduced per block; partials summed in a fixed order by k_ke_finish.
constexpr int kRedBlocks = 1024;

__global__ void __launch_bounds__(kBlock) k_sumsq(const float* __restrict__ u, int64_t n,
                                                  double* __restrict__ partial) {
    __shared__ double sh[kBlock / 64];
    double s = 0.0;
    for (int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; f < n;
         f += (int64_t)gridDim.x * blockDim.x) {
        double v = (double)u[f];
        s += v * v;
    }
    s = block_sum(s, sh);
    if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

__global__ void __launch_bounds__(kBlock) k_ke_finish(const double* __restrict__ partial, int dim,
                                                      double n0, double n1, double n2, double* out) {
    __shared__ double sh[kBlock / 64];
    double E = 0.0;
    const double nf[3] = {n0, n1, n2};
    for (int a = 0; a < dim; ++a) {
        double s = 0.0;
        for (int b = threadIdx.x; b < kRedBlocks; b += blockDim.x) s += partial[a * kRedBlocks + b];
        s = block_sum(s, sh);
        if (threadIdx.x == 0) E += s / nf[a];
    }
    if (threadIdx.x == 0) *out = E;
}

// add_fields (3D/taichi_utils.py:18-21): dest = f1 + multiplier * f2 per
// element, the product rounded before the sum as in the reference's source.
__global__ void __launch_bounds__(kBlock) k_add_fields(int64_t n, const float* __restrict__ f1,
                                                       const float* __restrict__ f2, float* __restrict__ dest,
                                                       float m) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const float t = __fmul_rn(m, f2[i]);
        dest[i] = __fadd_rn(f1[i], t);
    }
}

}  // namespace pfm

using namespace pfm;

extern "C" {

int pfm_add_fields(int64_t n, const float* f1, const float* f2, float* dest, float multiplier,
                   pfm_stream_t stream) {
    PFM_CHECK_ARG(n >= 0, "negative element count");
    if (n == 0) return PFM_OK;
    PFM_CHECK_ARG(f1 && f2 && dest, "null pointer");
    hipLaunchKernelGGL(k_add_fields, dim3(grid_for(n)), dim3(kBlock), 0, as_stream(stream), n, f1, f2, dest,
                       multiplier);
    PFM_LAUNCH_CHECK();
    return PFM_OK;
}

int pfm_rk4_grid(const pfm_grid* gp, int axis, float* psi, float* T, const float* u_x,
                 const float* u_y, const float* u_z, float dt, pfm_stream_t stream) {
    GridDev g;
    int rc = make_grid(gp, &g);
    if (rc) return rc;
    PFM_CHECK_ARG(axis >= 0 && axis < g.dim, "axis out of range");
    PFM_CHECK_ARG(psi && T && u_x && u_y && (g.dim == 2 || u_z), "null pointer");
    const int64_t nf = face_count_h(g, axis);
    if (g.dim == 3)
        hipLaunchKernelGGL(k_rk4_grid<3>, dim3(grid_for(nf)), dim3(kBlock), 0, as_stream(stream), g, nf,
                           psi, T, u_x, u_y, u_z, dt);
    else
        hipLaunchKernelGGL(k_rk4_grid<2>, dim3(grid_for(nf)), dim3(kBlock), 0, as_stream(stream), g, nf,
                           psi, T, u_x, u_y, u_z, dt);
    PFM_LAUNCH_CHECK();
    return PFM_OK;
}

int pfm_advect_u(const pfm_grid* gp, const float* u0_x, const float* u0_y, const float* u0_z,
                 float* u1_x, float* u1_y, float* u1_z, const float* T_x, const float* T_y,
                 const float* T_z, const float* psi_x, const float* psi_y, const float* psi_z,
                 pfm_stream_t stream) {
    GridDev g;
    int rc = make_grid(gp, &g);
    if (rc) return rc;
    AdvectArgs a;
    const float* u0[3] = {u0_x, u0_y, u0_z};
    float* u1[3] = {u1_x, u1_y, u1_z};
    const float* T[3] = {T_x, T_y, T_z};
    const float* psi[3] = {psi_x, psi_y, psi_z};
    int64_t total = 0;
    for (int c = 0; c < 3; ++c) {
        a.u0[c] = u0[c];
        a.u1[c] = u1[c];
        a.T[c] = T[c];
        a.psi[c] = psi[c];
        a.nf[c] = c < g.dim ? face_count_h(g, c) : 0;
        if (c < g.dim) {
            PFM_CHECK_ARG(u0[c] && u1[c] && T[c] && psi[c], "null pointer for axis %d", c);
            for (int d = 0; d < g.dim; ++d)
                PFM_CHECK_ARG(u1[c] != u0[d], "u1 must not alias u0");
        }
        total += a.nf[c];
    }
    if (g.dim == 3)
        hipLaunchKernelGGL(k_advect_u<3>, dim3(grid_for(total)), dim3(kBlock), 0, as_stream(stream), g, a);
    else
        hipLaunchKernelGGL(k_advect_u<2>, dim3(grid_for(total)), dim3(kBlock), 0, as_stream(stream), g, a);
    PFM_LAUNCH_CHECK();
    return PFM_OK;
}

int pfm_midpoint_advect(const pfm_grid* gp, const float* u0_x, const float* u0_y, const float* u0_z,
                        float* u1_x, float* u1_y, float* u1_z, float* psi_x, float* psi_y, float* psi_z,
                        float* T_x, float* T_y, float* T_z, float dt, pfm_stream_t stream) {
    GridDev g;
    int rc = make_grid(gp, &g);
    if (rc) return rc;
    AdvectArgs a;
    MidOut o;
    const float* u0[3] = {u0_x, u0_y, u0_z};
    float* u1[3] = {u1_x, u1_y, u1_z};
    float* psi[3] = {psi_x, psi_y, psi_z};
    float* T[3] = {T_x, T_y, T_z};
    int64_t total = 0;
    for (int c = 0; c < 3; ++c) {
        a.u0[c] = u0[c];
        a.u1[c] = u1[c];
        a.T[c] = T[c];
        a.psi[c] = psi[c];
        o.psi[c] = psi[c];
        o.T[c] = T[c];
        a.nf[c] = c < g.dim ? face_count_h(g, c) : 0;
        if (c < g.dim) {
            PFM_CHECK_ARG(u0[c] && u1[c] && T[c] && psi[c], "null pointer for axis %d", c);
            for (int d = 0; d < g.dim; ++d)
                PFM_CHECK_ARG(u1[c] != u0[d], "u1 must not alias u0");
        }
        total += a.nf[c];
    }
    if (g.dim == 3)
        hipLaunchKernelGGL(k_midpoint_advect<3>, dim3(grid_for(total)), dim3(kBlock), 0, as_stream(stream), g, a, o,
                           dt);
    else
        hipLaunchKernelGGL(k_midpoint_advect<2>, dim3(grid_for(total)), dim3(kBlock), 0, as_stream(stream), g, a, o,
                           dt);
    PFM_LAUNCH_CHECK();
    return PFM_OK;
}

int pfm_reset_identity_grid(const pfm_grid* gp, int axis, float* psi, float* T,
                            pfm_stream_t stream) {
    GridDev g;
    int rc = make_grid(gp, &g);
    if (rc) return rc;
    PFM_CHECK_ARG(axis >= 0 && axis < g.dim, "axis out of range");
    PFM_CHECK_ARG(psi && T, "null pointer");
    const int64_t nf = face_count_h(g, axis);
    if (g.dim == 3)
        hipLaunchKernelGGL(k_reset_identity_grid<3>, dim3(grid_for(nf)), dim3(kBlock), 0, as_stream(stream), g,
                           axis, nf, psi, T);
    else
        hipLaunchKernelGGL(k_reset_identity_grid<2>, dim3(grid_for(nf)), dim3(kBlock), 0, as_stream(stream), g,
                           axis, nf, psi, T);
    PFM_LAUNCH_CHECK();
    return PFM_OK;
}

int pfm_max_speed(const pfm_grid* gp, const float* u_x, const float* u_y, const float* u_z,
                  float* out, pfm_stream_t stream) {
    GridDev g;
    int rc = make_grid(gp, &g);
    if (rc) return rc;
    PFM_CHECK_ARG(u_x && u_y && (g.dim == 2 || u_z) && out, "null pointer");
    hipStream_t s = as_stream(stream);
    unsigned int init_bits;
    float floor_v = 1.e-3f;  // max_speed[None] = 1.e-3 (3D/run_3D.py:97)
    memcpy(&init_bits, &floor_v, 4);
    hipLaunchKernelGGL(k_set_u32, dim3(1), dim3(1), 0, s, reinterpret_cast<unsigned int*>(out), init_bits);
    PFM_LAUNCH_CHECK();
    const int64_t C = cell_count(g);
    unsigned blocks = grid_for(C);
    if (blocks > 2048) blocks = 2048;
    if (g.dim == 3)
        hipLaunchKernelGGL(k_max_speed<3>, dim3(blocks), dim3(kBlock), 0, s, g, u_x, u_y, u_z,
                           reinterpret_cast<unsigned int*>(out));
    else
        hipLaunchKernelGGL(k_max_speed<2>, dim3(blocks), dim3(kBlock), 0, s, g, u_x, u_y, u_z,
                           reinterpret_cast<unsigned int*>(out));
    PFM_LAUNCH_CHECK();
    return PFM_OK;
}

int pfm_split_central(const pfm_grid* gp, const float* vc, float* u_x, float* u_y, float* u_z,
                      pfm_stream_t stream) {
    GridDev g;
    int rc = make_grid(gp, &g);
    if (rc) return rc;
    float* u[3] = {u_x, u_y, u_z};
    PFM_CHECK_ARG(vc && u_x && u_y && (g.dim == 2 || u_z), "null pointer");
    for (int a = 0; a < g.dim; ++a) {
        const int64_t nf = face_count_h(g, a);
        if (g.dim == 3)
            hipLaunchKernelGGL(k_split_central<3>, dim3(grid_for(nf)), dim3(kBlock), 0, as_stream(stream), g, a,
                               nf, vc, u[a]);
        else
            hipLaunchKernelGGL(k_split_central<2>, dim3(grid_for(nf)), dim3(kBlock), 0, as_stream(stream), g, a,
                               nf, vc, u[a]);
        PFM_LAUNCH_CHECK();
    }
    return PFM_OK;
}

int pfm_get_central(const pfm_grid* gp, const float* u_x, const float* u_y, const float* u_z,
                    float* vc, pfm_stream_t stream) {
    GridDev g;
    int rc = make_grid(gp, &g);
    if (rc) return rc;
    PFM_CHECK_ARG(vc && u_x && u_y && (g.dim == 2 || u_z), "null pointer");
    const int64_t C = cell_count(g);
    if (g.dim == 3)
        hipLaunchKernelGGL(k_get_central<3>, dim3(grid_for(C)), dim3(kBlock), 0, as_stream(stream), g, u_x, u_y,
                           u_z, vc);
    else
        hipLaunchKernelGGL(k_get_central<2>, dim3(grid_for(C)), dim3(kBlock), 0, as_stream(stream), g, u_x, u_y,
                           u_z, vc);
    PFM_LAUNCH_CHECK();
    return PFM_OK;
}

int pfm_curl(const pfm_grid* gp, const float* vc, float* out, pfm_stream_t stream) {
    GridDev g;
    int rc = make_grid(gp, &g);
    if (rc) return rc;
    PFM_CHECK_ARG(vc && out, "null pointer");
    const int64_t C = cell_count(g);
    const float two_dx = 2.0f * g.dx;
    const float inv = 1.0f / two_dx;
    if (g.dim == 3)
        hipLaunchKernelGGL(k_curl<3>, dim3(grid_for(C)), dim3(kBlock), 0, as_stream(stream), g, vc, inv, two_dx, out);
    else
        hipLaunchKernelGGL(k_curl<2>, dim3(grid_for(C)), dim3(kBlock), 0, as_stream(stream), g, vc, inv, two_dx, out);
    PFM_LAUNCH_CHECK();
    return PFM_OK;
}

int pfm_point_vortices(const pfm_grid* gp, int nv, const float* cx, const float* cy,
                       const float* w, float rad, float strength, float* vc,
                       pfm_stream_t stream) {
    GridDev g;
    int rc = make_grid(gp, &g);
    if (rc) return rc;
    PFM_CHECK_ARG(g.dim == 2, "point vortices are 2D");
    PFM_CHECK_ARG(nv >= 1 && nv <= kMaxVort, "nv must be in [1, %d]", kMaxVort);
    PFM_CHECK_ARG(cx && cy && w && vc, "null pointer");
    hipStream_t s = as_stream(stream);
    const int64_t C = cell_count(g);
    const float rad2 = rad * rad;
    for (int v0 = 0; v0 < nv; v0 += kVortChunk) {
        const int m = nv - v0 < kVortChunk ? nv - v0 : kVortChunk;
        VortChunk vt = {};
        memcpy(vt.x, cx + v0, sizeof(float) * m);
        memcpy(vt.y, cy + v0, sizeof(float) * m);
        memcpy(vt.w, w + v0, sizeof(float) * m);
        hipLaunchKernelGGL(k_point_vortices, dim3(grid_for(C)), dim3(kBlock), 0, s, g, vt, m, v0 == 0, rad2,
                           strength, vc);
        PFM_LAUNCH_CHECK();
    }
    return PFM_OK;
}

int pfm_vortex_ring(const pfm_grid* gp, float w, float rad, float delta, const float* c,
                    const float* ex, const float* ey, int num_samples, float* vc,
                    pfm_stream_t stream) {
    GridDev g;
    int rc = make_grid(gp, &g);
    if (rc) return rc;
    PFM_CHECK_ARG(g.dim == 3, "vortex rings are 3D");
    PFM_CHECK_ARG(num_samples >= 1 && c && ex && ey && vc, "bad arguments");
    RingArgs a;
    a.w = w;
    a.rad = rad;
    a.delta = delta;
    a.curve_length = (float)((2.0 * M_PI * (double)rad) / num_samples);
    a.four_pi = (float)(4.0 * M_PI);
    a.pi = (float)M_PI;
    for (int q = 0; q < 3; ++q) {
        a.c[q] = c[q];
        a.ex[q] = ex[q];
        a.ey[q] = ey[q];
    }
    a.num_samples = num_samples;
    const int64_t C = cell_count(g);
    hipLaunchKernelGGL(k_vortex_ring, dim3(grid_for(C)), dim3(kBlock), 0, as_stream(stream), g, a, vc);
    PFM_LAUNCH_CHECK();
    return PFM_OK;
}

int pfm_advect_smoke(const pfm_grid* gp, int channels, const float* smoke0, float* smoke1, const float* psi_x,
                     const float* psi_y, const float* psi_z, int mode, pfm_stream_t stream) {
    GridDev g;
    int rc = make_grid(gp, &g);
    if (rc) return rc;
    PFM_CHECK_ARG(channels >= 1 && channels <= 4, "channels must be 1..4");
    PFM_CHECK_ARG(mode >= 1 && mode <= 3, "mode must be 1 (advect), 2 (clamp) or 3 (both)");
    PFM_CHECK_ARG(g.dim == 3 || mode == 1, "clamp_smoke is 3D only");
    PFM_CHECK_ARG(smoke0 && smoke1 && psi_x && psi_y && (g.dim == 2 || psi_z), "null pointer");
    PFM_CHECK_ARG(smoke0 != smoke1 || mode == 2, "advect_smoke needs distinct input and output fields");
    SmokeArgs a;
    a.nch = channels;
    a.mode = mode;
    for (int c = 0; c < 3; ++c) a.hi[c] = (float)((double)g.res[c] - 1.0 - 1e-6);
    a.lo = g.dim == 3 ? 1.0f : 0.0f;
    a.third = g.dim == 3 ? (float)(1.0 / 6.0) : 0.25f;
    const int64_t C = cell_count(g);
    if (g.dim == 3)
        hipLaunchKernelGGL(k_advect_smoke<3>, dim3(grid_for(C)), dim3(kBlock), 0, as_stream(stream), g, a, smoke0,
                           smoke1, psi_x, psi_y, psi_z);
    else
        hipLaunchKernelGGL(k_advect_smoke<2>, dim3(grid_for(C)), dim3(kBlock), 0, as_stream(stream), g, a, smoke0,
                           smoke1, psi_x, psi_y, psi_z);
    PFM_LAUNCH_CHECK();
    return PFM_OK;
}

int pfm_vortex_ring_smoke(const pfm_grid* gp, float rad, float delta, const float* c, const float* ex,
                          const float* ey, int num_samples, const float* color, float* smoke, pfm_stream_t stream) {
    GridDev g;
    int rc = make_grid(gp, &g);
    if (rc) return rc;
    PFM_CHECK_ARG(g.dim == 3, "vortex rings are 3D");
    PFM_CHECK_ARG(num_samples >= 1 && c && ex && ey && color && smoke, "bad arguments");
    RingArgs a = {};
    a.rad = rad;
    a.delta = delta;
    a.curve_length = (float)((2.0 * M_PI * (double)rad) / num_samples);
    a.pi = (float)M_PI;
    for (int q = 0; q < 3; ++q) {
        a.c[q] = c[q];
        a.ex[q] = ex[q];
        a.ey[q] = ey[q];
    }
    a.num_samples = num_samples;
    const int64_t C = cell_count(g);
    hipStream_t s = as_stream(stream);
    hipLaunchKernelGGL(k_ring_smoke, dim3(grid_for(C)), dim3(kBlock), 0, s, g, a, smoke);
    PFM_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_smoke_threshold, dim3(grid_for(C)), dim3(kBlock), 0, s, C, color[0], color[1], color[2],
                       smoke);
    PFM_LAUNCH_CHECK();
    return PFM_OK;
}

int pfm_kinetic_energy(const pfm_grid* gp, const float* u_x, const float* u_y, const float* u_z,
                       double* out, pfm_stream_t stream) {
    GridDev g;
    int rc = make_grid(gp, &g);
    if (rc) return rc;
    PFM_CHECK_ARG(u_x && u_y && (g.dim == 2 || u_z) && out, "null pointer");
    // The partial sums live in the output's trailing scratch: the caller's
    // `out` must hold 1 + 3*1024 doubles.
    double* partial = out + 1;
    const float* u[3] = {u_x, u_y, u_z};
    double nf[3] = {1, 1, 1};
    hipStream_t s = as_stream(stream);
    for (int a = 0; a < g.dim; ++a) {
        nf[a] = (double)face_count_h(g, a);
        hipLaunchKernelGGL(k_sumsq, dim3(kRedBlocks), dim3(kBlock), 0, s, u[a], face_count_h(g, a),
                           partial + a * kRedBlocks);
        PFM_LAUNCH_CHECK();
    }
    hipLaunchKernelGGL(k_ke_finish, dim3(1), dim3(kBlock), 0, s, partial, g.dim, nf[0], nf[1], nf[2], out);
    PFM_LAUNCH_CHECK();
    return PFM_OK;
}

}  // extern "C"

namespace {

// Device address of `scratch` on each GPU, looked up once per device with that
// device current (every GPU holds its own copy of the array).
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};

int scratch_for_current_device(double** out) {
    int dev = -1;
    PFM_HIP_TRY(hipGetDevice(&dev));
    PFM_CHECK_ARG(dev >= 0 && dev < kMaxDevices, "device %d beyond the %d devices the scratch table holds", dev,
                  kMaxDevices);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        PFM_HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        PFM_CHECK_ARG(p != nullptr, "hipGetSymbolAddress returned no address for the scratch array");
        g_scratch_ptr[dev] = static_cast<double*>(p);
    }
    *out = g_scratch_ptr[dev];
    return PFM_OK;
}

}  // namespace

extern "C" int pfm_kinetic_energy_noalloc(const pfm_grid* gp, const float* u_x, const float* u_y,
                                          const float* u_z, double* out, pfm_stream_t stream) {
    GridDev g;
    int rc = make_grid(gp, &g);
    if (rc) return rc;
    PFM_CHECK_ARG(u_x && u_y && (g.dim == 2 || u_z) && out, "null pointer");
    // The partials of every axis fill all kRedBlocks slots of their row
    // (one per block of k_sumsq) before k_ke_finish reads them, so nothing
    // left by an earlier call is read and nothing needs clearing.
    static_assert(sizeof(scratch) / sizeof(double) >= 3 * kRedBlocks, "scratch holds 3 x kRedBlocks partials");
    double* partial = nullptr;
    rc = scratch_for_current_device(&partial);
    if (rc) return rc;
    const float* u[3] = {u_x, u_y, u_z};
    double nf[3] = {1, 1, 1};
    hipStream_t s = as_stream(stream);
    for (int a = 0; a < g.dim; ++a) {
        nf[a] = (double)face_count_h(g, a);
        hipLaunchKernelGGL(k_sumsq, dim3(kRedBlocks), dim3(kBlock), 0, s, u[a], face_count_h(g, a),
                           partial + a * kRedBlocks);
        PFM_LAUNCH_CHECK();
    }
    hipLaunchKernelGGL(k_ke_finish, dim3(1), dim3(kBlock), 0, s, partial, g.dim, nf[0], nf[1], nf[2], out);
    PFM_LAUNCH_CHECK();
    return PFM_OK;
}
