// particles.hip -- per-particle kernels of the flow-map step: forward RK4 of
// position and backward-map Jacobian, G2P at reinitialisation, Jacobian
// composition, impulse/affine updates, identity resets and lattice seeding.
// One thread per particle over SoA arrays (coalesced); the 4^d B-spline
// stencil is evaluated separably from 16-byte row loads (see pfm_device.h).
#include "pfm_batch.h"

namespace pfm {

// RK4_T_forward: 3D/run_3D.py:175-219, 2D/run_2D.py:304-341.
// dT = M @ T with M = cols(grad u_x, grad u_y, grad u_z): dT[r] = sum_c G[c][r] T[c].
// Optional epilogue: the P2G record of the advected particle (the composition
// of 3D/run_3D.py:460-479 with the new T_grad_m), written in particle order.
struct RecArgs {
    const float* Tin[3];
    const float* C[3];
    const float* imp;
    float* rec;
    float* Tfold[3];  // FOLD: where T_init <- cols(T_grad_m) cols(T_init) goes (in place)
    // BIN: the count pass of the particle binning of the advected positions
    int32_t* cell_of;
    int32_t* counts;
    int32_t* slot;
};

// update_T's product (3D/run_3D.py:430-439): O[m][r] = sum_k A[r][k] B[k][m]
// with A = rows of T_grad_m (A[r][k] = T_k[r]) and B[k][m] = T_init_m[k];
// the multiply-adds are written as FMAs so every kernel that forms it
// (k_update_T and the folded RK4 epilogue) rounds identically.
template <int DIM>
__device__ __forceinline__ float t_product(const float A[3][3], const float B[3][3], int r, int m) {
    float acc = A[r][0] * B[0][m];
#pragma unroll
    for (int k = 1; k < DIM; ++k) acc = __builtin_fmaf(A[r][k], B[k][m], acc);
    return acc;
}

// FOLD (3D step with a short reinitialisation every step): T_grad_m enters as
// the identity and is not read; the epilogue writes the record and then
// T_init <- cols(T_grad_m) cols(T_init), the update_T that the next step's
// short reinitialisation would apply (3D/run_3D.py:767-771), leaving
// T_grad_m in memory as the identity that reset would write.
// One particle: i indexes the SoA arrays, whose component stride is n.
// A non-null `lds_rec` receives the record (kRec floats at lds_rec, stride 1)
// instead of global memory; the kernel then copies its block's records out.
template <int DIM, bool REC, bool FOLD, bool BIN = false>
__device__ __forceinline__ void rk4_particle(const GridDev& g, int64_t i, int64_t n, float* __restrict__ pos,
                                             float* __restrict__ Tx, float* __restrict__ Ty,
                                             float* __restrict__ Tz, const float* __restrict__ ux,
                                             const float* __restrict__ uy, const float* __restrict__ uz,
                                             float dt, int advect_psi, const RecArgs& ra,
                                             float* lds_rec = nullptr) {
    float* Ts[3] = {Tx, Ty, Tz};
    float p0[3] = {0.f, 0.f, 0.f};
    float T0[3][3];
#pragma unroll
    for (int c = 0; c < DIM; ++c) p0[c] = pos[c * n + i];
    // FOLD: the identity as opaque values, so the stage arithmetic is
    // compiled exactly as for a T_grad_m read from memory (no constant
    // folding of the 0 / 1 entries, which could change roundings)
    float one = 1.0f, zero = 0.0f;
    if (FOLD) asm volatile("" : "+v"(one), "+v"(zero));
#pragma unroll
    for (int m = 0; m < DIM; ++m)
#pragma unroll
        for (int c = 0; c < DIM; ++c) T0[m][c] = FOLD ? (m == c ? one : zero) : Ts[m][c * n + i];

    const float half_dt = 0.5f * dt;
    const float hstage[3] = {half_dt, half_dt, 1.0f * dt};
    const float wstage[4] = {1.f, 2.f, 2.f, 1.f};
    float pk[3] = {p0[0], p0[1], p0[2]};
    float Tk[3][3];
    float us[3] = {0.f, 0.f, 0.f};
    float dTs[3][3];
#pragma unroll
    for (int m = 0; m < DIM; ++m)
#pragma unroll
        for (int c = 0; c < DIM; ++c) {
            Tk[m][c] = T0[m][c];
            dTs[m][c] = 0.f;
        }

    RowCache<DIM> rows;
    rows.clear();
#pragma unroll 1
    for (int st = 0; st < 4; ++st) {
        float u[3], G[3][3];
        interp_mac_cached<DIM, true>(ux, uy, uz, g, pk, u, G, rows);
        float dT[3][3];
#pragma unroll
        for (int m = 0; m < DIM; ++m)
#pragma unroll
            for (int r = 0; r < DIM; ++r) {
                float acc = G[0][r] * Tk[m][0];
#pragma unroll
                for (int c = 1; c < DIM; ++c) acc += G[c][r] * Tk[m][c];
                dT[m][r] = acc;
            }
        const float ws = wstage[st];
#pragma unroll
        for (int c = 0; c < DIM; ++c) us[c] += ws * u[c];
#pragma unroll
        for (int m = 0; m < DIM; ++m)
#pragma unroll
            for (int c = 0; c < DIM; ++c) dTs[m][c] += ws * dT[m][c];
        if (st < 3) {
            const float h = hstage[st];
#pragma unroll
            for (int c = 0; c < DIM; ++c) pk[c] = p0[c] + h * u[c];
#pragma unroll
            for (int m = 0; m < DIM; ++m)
#pragma unroll
                for (int c = 0; c < DIM; ++c) Tk[m][c] = T0[m][c] - h * dT[m][c];
        }
    }
    const float sixth_dt = dt * 1.0f / 6.0f;
    float p1[3] = {p0[0], p0[1], p0[2]};
    if (advect_psi) {
#pragma unroll
        for (int c = 0; c < DIM; ++c) {
            p1[c] = p0[c] + sixth_dt * us[c];
            pos[c * n + i] = p1[c];
        }
    }
    float T1[3][3];
#pragma unroll
    for (int m = 0; m < DIM; ++m)
#pragma unroll
        for (int c = 0; c < DIM; ++c) {
            T1[m][c] = T0[m][c] - sixth_dt * dTs[m][c];
            if (!FOLD) Ts[m][c * n + i] = T1[m][c];
        }
    if (REC) {
        float A[3][3], B[3][3], C[3][3], m0[3];
#pragma unroll
        for (int r = 0; r < DIM; ++r) {
            m0[r] = ra.imp[r * n + i];
#pragma unroll
            for (int c = 0; c < DIM; ++c) {
                A[r][c] = T1[r][c];               // rows(T_grad_m): A[r][c] = T_r[c]
                B[r][c] = ra.Tin[c][r * n + i];   // cols(T_init)
                C[r][c] = ra.C[r][c * n + i];     // rows(init_C)
            }
        }
        if (lds_rec) {
            float rec[Pay<DIM>::kRec];
            form_p2g_record<DIM>(rec, p1, g.inv_dx, true, A, B, C, m0);
#pragma unroll
            for (int k = 0; k < Pay<DIM>::kRec; ++k) lds_rec[k] = rec[k];
        } else {
            write_p2g_record<DIM>(ra.rec + i * Pay<DIM>::kRec, p1, g.inv_dx, true, A, B, C, m0);
        }
        if (FOLD) {
            // update_T's A is cols(T_grad_m): Au[r][k] = T_k[r] (the record's A transposed)
            float Au[3][3];
#pragma unroll
            for (int r = 0; r < DIM; ++r)
#pragma unroll
                for (int k = 0; k < DIM; ++k) Au[r][k] = T1[k][r];
#pragma unroll
            for (int m = 0; m < DIM; ++m)
#pragma unroll
                for (int r = 0; r < DIM; ++r) ra.Tfold[m][r * n + i] = t_product<DIM>(Au, B, r, m);
        }
    }
    if (BIN) {
        // get_particles_id_in_every_cell on the advected position (the count
        // pass of pfm_bin_particles: cell, then a slot from the run-merged
        // count atomics; inactive lanes of the last wave are its top lanes)
        const int32_t cell = particle_cell<DIM>(g, pos, i, n);
        ra.cell_of[i] = cell;
        count_slots(cell, true, ra.counts, ra.slot + i);
    }
}

template <int DIM, bool REC = false, bool FOLD = false, bool BIN = false>
__global__ void __launch_bounds__(kBlock, 3) k_rk4_particles(GridDev g, int64_t n,
                                                          float* __restrict__ pos,
                                                          float* __restrict__ Tx,
                                                          float* __restrict__ Ty,
                                                          float* __restrict__ Tz,
                                                          const float* __restrict__ ux,
                                                          const float* __restrict__ uy,
                                                          const float* __restrict__ uz, float dt,
                                                          int advect_psi, RecArgs ra = RecArgs{}) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if constexpr (REC && DIM == 3) {
        // the block's records pass through LDS (stride kRec + 1 floats: the
        // per-lane writes hit distinct banks) so that they leave as
        // contiguous 16-byte chunks, one 1 KB span per wave store
        constexpr int kR = Pay<3>::kRec, kS = kR + 1;
        __shared__ float recs[kBlock * kS];
        if (i < n)
            rk4_particle<DIM, REC, FOLD, BIN>(g, i, n, pos, Tx, Ty, Tz, ux, uy, uz, dt, advect_psi, ra,
                                              recs + threadIdx.x * kS);
        __syncthreads();
        const int64_t b0 = (int64_t)blockIdx.x * blockDim.x;
        const int nact = (int)min((int64_t)blockDim.x, n - b0);
        float4* dst = reinterpret_cast<float4*>(ra.rec + b0 * kR);
        for (int c = threadIdx.x; c < nact * (kR / 4); c += blockDim.x) {
            const float* q = recs + (c / (kR / 4)) * kS + (c % (kR / 4)) * 4;
            dst[c] = make_float4(q[0], q[1], q[2], q[3]);
        }
    } else {
        if (i >= n) return;
        rk4_particle<DIM, REC, FOLD, BIN>(g, i, n, pos, Tx, Ty, Tz, ux, uy, uz, dt, advect_psi, ra);
    }
}


// init_particles_imp / init_particles_imp_grad_m: 3D/init_conditions.py:254-273,
// 2D/init_conditions.py:183-210 (interp_u_MAC_grad_imp, 3D/taichi_utils.py:326-376:
// new_C rows equal the velocity gradient rows).
template <int DIM, bool GRAD>
__device__ __forceinline__ void g2p_particle(const GridDev& g, int64_t i, int64_t n, const float* __restrict__ pos,
                                             const float* __restrict__ ux, const float* __restrict__ uy,
                                             const float* __restrict__ uz, float* __restrict__ imp,
                                             float* __restrict__ Cx, float* __restrict__ Cy,
                                             float* __restrict__ Cz) {
    float p[3] = {0.f, 0.f, 0.f};
#pragma unroll
    for (int c = 0; c < DIM; ++c) p[c] = pos[c * n + i];
    float u[3], G[3][3];
    interp_mac<DIM, GRAD>(ux, uy, uz, g, p, u, G);
    if (imp != nullptr) {
#pragma unroll
        for (int c = 0; c < DIM; ++c) imp[c * n + i] = u[c];
    }
    if (GRAD) {
        float* Cs[3] = {Cx, Cy, Cz};
#pragma unroll
        for (int a = 0; a < DIM; ++a)
#pragma unroll
            for (int c = 0; c < DIM; ++c) Cs[a][c * n + i] = G[a][c];
    }
}

template <int DIM, bool GRAD>
__global__ void __launch_bounds__(kBlock, 6) k_g2p(GridDev g, int64_t n, const float* __restrict__ pos,
                                                const float* __restrict__ ux,
                                                const float* __restrict__ uy,
                                                const float* __restrict__ uz,
                                                float* __restrict__ imp, float* __restrict__ Cx,
                                                float* __restrict__ Cy, float* __restrict__ Cz) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if constexpr (DIM == 3) {
        // one component per trip of a rolled loop: the registers of one
        // component's stencil and rows are live at a time
        float p[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) p[c] = pos[c * n + i];
#pragma unroll 1
        for (int a = 0; a < 3; ++a) {
            const float* fld = a == 0 ? ux : (a == 1 ? uy : uz);
            float val, grad[3];
            sample_component_rt<GRAD>(fld, g, a, p, val, grad);
            if (imp != nullptr) imp[a * n + i] = val;
            if (GRAD) {
                float* Ca = a == 0 ? Cx : (a == 1 ? Cy : Cz);
#pragma unroll
                for (int c = 0; c < 3; ++c) Ca[c * n + i] = grad[c];
            }
        }
    } else {
        g2p_particle<DIM, GRAD>(g, i, n, pos, ux, uy, uz, imp, Cx, Cy, Cz);
    }
}

template <bool GRAD>
__global__ void __launch_bounds__(kBlock, 2) kb_g2p(BUnits t, int64_t P, const float* __restrict__ pos,
                                                 const float* __restrict__ ux, const float* __restrict__ uy,
                                                 float* __restrict__ imp, float* __restrict__ Cx,
                                                 float* __restrict__ Cy) {
    PFM_FOR_UNITS(t, u) {
        const BUnit w = t.u[u];
        if (!unit_on(t, w.scene)) continue;
        const BScene& S = t.sc[w.scene];
        const int64_t loc = (int64_t)w.chunk * kBlock + threadIdx.x;
        if (loc >= S.n) continue;
        g2p_particle<2, GRAD>(S.g, S.p_off + loc, P, pos, ux + S.f_off[0], uy + S.f_off[1], nullptr, imp, Cx, Cy,
                              nullptr);
    }
}

// update_T: out = cols(T_gm) @ cols(T_in) (3D/run_3D.py:430-439, 2D/run_2D.py:538-546).
template <int DIM>
__global__ void __launch_bounds__(kBlock) k_update_T(int64_t n, const float* __restrict__ gx,
                                                     const float* __restrict__ gy,
                                                     const float* __restrict__ gz,
                                                     const float* inx, const float* iny,
                                                     const float* inz, float* ox, float* oy,
                                                     float* oz) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float* G[3] = {gx, gy, gz};
    const float* In[3] = {inx, iny, inz};
    float* O[3] = {ox, oy, oz};
    float A[3][3], B[3][3];
#pragma unroll
    for (int m = 0; m < DIM; ++m)
#pragma unroll
        for (int r = 0; r < DIM; ++r) {
            A[r][m] = G[m][r * n + i];
            B[r][m] = In[m][r * n + i];
        }
#pragma unroll
    for (int m = 0; m < DIM; ++m)
#pragma unroll
        for (int r = 0; r < DIM; ++r) O[m][r * n + i] = t_product<DIM>(A, B, r, m);
}

// Row r of cols(T) @ m0 with Tc[c][r] = component r of T_c.
template <int DIM>
__device__ __forceinline__ float imp_row(const float Tc[3][3], const float m0[3], int r) {
    float acc = Tc[0][r] * m0[0];
#pragma unroll
    for (int c = 1; c < DIM; ++c) acc += Tc[c][r] * m0[c];
    return acc;
}

// update_particles_imp: imp = cols(T) @ init_imp (2D/run_2D.py:517-524).
template <int DIM>
__global__ void __launch_bounds__(kBlock) k_update_imp(int64_t n, const float* __restrict__ tx,
                                                       const float* __restrict__ ty,
                                                       const float* __restrict__ tz,
                                                       const float* __restrict__ init_imp,
                                                       float* __restrict__ imp) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float* T[3] = {tx, ty, tz};
    float m0[3], Tc[3][3];
#pragma unroll
    for (int c = 0; c < DIM; ++c) m0[c] = init_imp[c * n + i];
#pragma unroll
    for (int c = 0; c < DIM; ++c)
#pragma unroll
        for (int r = 0; r < DIM; ++r) Tc[c][r] = T[c][r * n + i];
#pragma unroll
    for (int r = 0; r < DIM; ++r) imp[r * n + i] = imp_row<DIM>(Tc, m0, r);
}

// T_init_C_T = cols(T_gm) @ (rows(init_C) @ rows(T_gm)) (2D/run_2D.py:526-536,
// 3D/run_3D.py:462-468).  A[r][c] = T_r[c] (rows), C[r][c] = init_C_r[c].
template <int DIM>
__device__ __forceinline__ void compose_C(const float A[3][3], const float C[3][3], float Q[3][3]) {
    float P[3][3];
#pragma unroll
    for (int r = 0; r < DIM; ++r)
#pragma unroll
        for (int c = 0; c < DIM; ++c) {
            float acc = C[r][0] * A[0][c];
#pragma unroll
            for (int k = 1; k < DIM; ++k) acc += C[r][k] * A[k][c];
            P[r][c] = acc;
        }
#pragma unroll
    for (int r = 0; r < DIM; ++r)
#pragma unroll
        for (int c = 0; c < DIM; ++c) {
            float acc = A[0][r] * P[0][c];
#pragma unroll
            for (int k = 1; k < DIM; ++k) acc += A[k][r] * P[k][c];
            Q[r][c] = acc;
        }
}

template <int DIM>
__global__ void __launch_bounds__(kBlock) k_update_grad_m(
    int64_t n, float* __restrict__ cx, float* __restrict__ cy, float* __restrict__ cz,
    const float* __restrict__ icx, const float* __restrict__ icy, const float* __restrict__ icz,
    const float* __restrict__ gx, const float* __restrict__ gy, const float* __restrict__ gz) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float* Gs[3] = {gx, gy, gz};
    const float* Ic[3] = {icx, icy, icz};
    float* Co[3] = {cx, cy, cz};
    float A[3][3], C[3][3], Q[3][3];
#pragma unroll
    for (int r = 0; r < DIM; ++r)
#pragma unroll
        for (int c = 0; c < DIM; ++c) {
            A[r][c] = Gs[r][c * n + i];
            C[r][c] = Ic[r][c * n + i];
        }
    compose_C<DIM>(A, C, Q);
#pragma unroll
    for (int r = 0; r < DIM; ++r)
#pragma unroll
        for (int c = 0; c < DIM; ++c) Co[r][c * n + i] = Q[r][c];
}

// reset_T_to_identity: 3D/run_3D.py:133-140 (T_m = unit vector m).
template <int DIM>
__global__ void __launch_bounds__(kBlock) k_reset_T(int64_t n, float* tx, float* ty, float* tz) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float* T[3] = {tx, ty, tz};
#pragma unroll
    for (int m = 0; m < DIM; ++m)
#pragma unroll
        for (int c = 0; c < DIM; ++c) T[m][c * n + i] = (m == c) ? 1.0f : 0.0f;
}

// Seeded jitter (extension): counter-based hash of (seed, particle id, component),
// identical to the oracle's so CPU and GPU seed the same positions bit-exactly.
__device__ __forceinline__ uint32_t hash32(uint32_t x) {
    x ^= x >> 16;
    x *= 0x85ebca6bu;
    x ^= x >> 13;
    x *= 0xc2b2ae35u;
    x ^= x >> 16;
    return x;
}

__device__ __forceinline__ float jitter_u(uint32_t seed, uint64_t id, uint32_t comp) {
    uint32_t h = hash32((uint32_t)id ^ hash32(seed * 0x9e3779b9u + comp));
    h = hash32(h ^ (uint32_t)(id >> 32) * 0x27d4eb2fu);
    return (float)(h >> 8) * (1.0f / 16777216.0f);
}

// init_particles_pos_uniform: 3D/init_conditions.py:238-252, 2D/init_conditions.py:139-166.
// order 0: the reference numbering (lattice x fastest).  order 1: cell-major
// (cells with the last axis fastest, the ppc_axis^d sub-lattice inside).
// Particle i of a scene (lattice id and jitter hash use i); its position
// goes to pos[c * stride].
template <int DIM>
__device__ __forceinline__ void seed_particle(const GridDev& g, int ppc_axis, float dist, int64_t i,
                                              float* __restrict__ pos, int64_t stride, int64_t seed, int order) {
    int64_t id[3] = {0, 0, 0};
    if (order == 0) {
        const int64_t nx = (int64_t)ppc_axis * g.res[0];
        const int64_t ny = (int64_t)ppc_axis * g.res[1];
        id[0] = i % nx;
        const int64_t iyz = i / nx;
        if (DIM == 3) {
            id[1] = iyz % ny;
            id[2] = iyz / ny;
        } else {
            id[1] = iyz;
        }
    } else {
        int64_t per = ppc_axis;
        for (int c = 1; c < DIM; ++c) per *= ppc_axis;
        const int64_t cell = i / per;
        int64_t sub = i % per;
        int64_t ci[3] = {0, 0, 0};
        if (DIM == 3) {
            ci[2] = cell % g.res[2];
            ci[1] = (cell / g.res[2]) % g.res[1];
            ci[0] = cell / ((int64_t)g.res[2] * g.res[1]);
        } else {
            ci[1] = cell % g.res[1];
            ci[0] = cell / g.res[1];
        }
        for (int c = DIM - 1; c >= 0; --c) {
            id[c] = ci[c] * ppc_axis + sub % ppc_axis;
            sub /= ppc_axis;
        }
    }
#pragma unroll
    for (int c = 0; c < DIM; ++c) {
        float off = seed >= 0 ? jitter_u((uint32_t)seed, (uint64_t)i, (uint32_t)c) : 0.5f;
        pos[c * stride] = __fmul_rn(__fadd_rn((float)id[c], off), dist);
    }
}

template <int DIM>
__global__ void __launch_bounds__(kBlock) k_seed(GridDev g, int ppc_axis, float dist, int64_t n,
                                                 float* __restrict__ pos, int64_t seed,
                                                 int order) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    seed_particle<DIM>(g, ppc_axis, dist, i, pos + i, n, seed, order);
}

__global__ void __launch_bounds__(kBlock) kb_seed(BUnits t, int ppc_axis, int order, int64_t P,
                                                  float* __restrict__ pos) {
    PFM_FOR_UNITS(t, u) {
        const BUnit w = t.u[u];
        if (!unit_on(t, w.scene)) continue;
        const BScene& S = t.sc[w.scene];
        const int64_t loc = (int64_t)w.chunk * kBlock + threadIdx.x;
        if (loc >= S.n) continue;
        seed_particle<2>(S.g, ppc_axis, S.dist, loc, pos + S.p_off + loc, P, S.seed, order);
    }
}

// Batched 2D RK4_T_forward followed, per particle, by the three updates of
// the 2D step (2D/run_2D.py:874-879: update_particles_grad_m, update_T,
// update_particles_imp, with the arithmetic of k_update_grad_m, k_update_T
// and k_update_imp) and the particle's P2G record in particle order (the
// record of k_p2g_prep<2, false>), so the P2G tile pass reads records through
// the cell permutation and no separate update or record pass runs.
struct Upd2Args {
    const float* Tin[2];   // T_init columns
    const float* iC[2];    // init_C rows
    const float* init_imp;
    float* C[2];           // out: C rows
    float* T[2];           // out: T columns
    float* imp;            // out
    float* rec;            // out: P2G records, particle order
};

__global__ void __launch_bounds__(kBlock, 3) kb_rk4_update(BUnits t, int64_t P, float* __restrict__ pos,
                                                        float* __restrict__ Tx, float* __restrict__ Ty,
                                                        const float* __restrict__ ux, const float* __restrict__ uy,
                                                        const BState* __restrict__ st, Upd2Args a) {
    PFM_FOR_UNITS(t, u) {
        const BUnit w = t.u[u];
        if (!unit_on(t, w.scene)) continue;
        const BScene& S = t.sc[w.scene];
        const int64_t loc = (int64_t)w.chunk * kBlock + threadIdx.x;
        if (loc >= S.n) continue;
        const float dt = (float)st[w.scene].dt;
        const int64_t i = S.p_off + loc;
        rk4_particle<2, false, false>(S.g, i, P, pos, Tx, Ty, nullptr, ux + S.f_off[0], uy + S.f_off[1], nullptr,
                                      dt, 1, RecArgs{});
        const float* G[2] = {Tx, Ty};
        float A[3][3] = {}, Au[3][3] = {}, B[3][3] = {}, Ci[3][3] = {}, Q[3][3] = {}, Tn[3][3] = {};
        float p1[3] = {0.f, 0.f, 0.f}, m0[3] = {0.f, 0.f, 0.f}, mo[3] = {0.f, 0.f, 0.f};
#pragma unroll
        for (int r = 0; r < 2; ++r) {
            p1[r] = pos[r * P + i];
            m0[r] = a.init_imp[r * P + i];
#pragma unroll
            for (int c = 0; c < 2; ++c) {
                A[r][c] = G[r][c * P + i];      // rows of T_grad_m: A[r][c] = T_r[c]
                Ci[r][c] = a.iC[r][c * P + i];  // rows of init_C
                B[r][c] = a.Tin[c][r * P + i];  // B[r][m] = T_init_m[r]
            }
        }
        compose_C<2>(A, Ci, Q);  // C = cols(T_gm) (rows(init_C) rows(T_gm))
#pragma unroll
        for (int r = 0; r < 2; ++r)
#pragma unroll
            for (int m = 0; m < 2; ++m) Au[r][m] = A[m][r];
#pragma unroll
        for (int m = 0; m < 2; ++m)
#pragma unroll
            for (int r = 0; r < 2; ++r) Tn[m][r] = t_product<2>(Au, B, r, m);  // T = cols(T_gm) cols(T_init)
#pragma unroll
        for (int r = 0; r < 2; ++r) mo[r] = imp_row<2>(Tn, m0, r);           // imp = cols(T) init_imp
#pragma unroll
        for (int r = 0; r < 2; ++r) {
            a.imp[r * P + i] = mo[r];
#pragma unroll
            for (int c = 0; c < 2; ++c) {
                a.C[r][c * P + i] = Q[r][c];
                a.T[r][c * P + i] = Tn[r][c];
            }
        }
        write_p2g_record<2>(a.rec + i * Pay<2>::kRec, p1, S.g.inv_dx, false, A, B, Q, mo);
    }
}

// The folded RK4 + P2G record kernel with the count pass of the binning of the
// advected positions in its epilogue (3D).
int launch_rk4_fold_bin(const GridDev& g, int64_t n, float* pos, const float* ux, const float* uy, const float* uz,
                        float dt, float* const Tin[3], const float* const iC[3], const float* init_imp, float* rec,
                        int32_t* cell_of, int32_t* counts, int32_t* slot, hipStream_t s) {
    RecArgs ra{};
    for (int c = 0; c < 3; ++c) {
        ra.Tin[c] = Tin[c];
        ra.Tfold[c] = Tin[c];
        ra.C[c] = iC[c];
    }
    ra.imp = init_imp;
    ra.rec = rec;
    ra.cell_of = cell_of;
    ra.counts = counts;
    ra.slot = slot;
    hipLaunchKernelGGL((k_rk4_particles<3, true, true, true>), dim3(grid_for(n)), dim3(kBlock), 0, s, g, n, pos,
                       nullptr, nullptr, nullptr, ux, uy, uz, dt, 1, ra);
    PFM_LAUNCH_CHECK();
    return PFM_OK;
}

// ---- launchers of the batched kernels ----------------------------------------
int b_seed(BUnits t, int ppc_axis, int order, int64_t P, float* pos, unsigned blocks, hipStream_t s) {
    hipLaunchKernelGGL(kb_seed, dim3(blocks), dim3(kBlock), 0, s, t, ppc_axis, order, P, pos);
    PFM_LAUNCH_CHECK();
    return PFM_OK;
}

int b_g2p(BUnits t, int64_t P, const int64_t F[2], const float* pos, const float* ux, const float* uy, float* imp,
          float* Cx, float* Cy, unsigned blocks, hipStream_t s) {
    (void)F;
    if (Cx)
        hipLaunchKernelGGL(kb_g2p<true>, dim3(blocks), dim3(kBlock), 0, s, t, P, pos, ux, uy, imp, Cx, Cy);
    else
        hipLaunchKernelGGL(kb_g2p<false>, dim3(blocks), dim3(kBlock), 0, s, t, P, pos, ux, uy, imp, Cx, Cy);
    PFM_LAUNCH_CHECK();
    return PFM_OK;
}

int b_rk4_update(BUnits t, int64_t P, float* pos, float* Tx, float* Ty, const float* ux, const float* uy,
                 const BState* st, const float* const* Tin, const float* const* iC, const float* init_imp, float* const* C,
                 float* const* T, float* imp, float* rec, unsigned blocks, hipStream_t s) {
    Upd2Args a;
    for (int k = 0; k < 2; ++k) {
        a.Tin[k] = Tin[k];
        a.iC[k] = iC[k];
        a.C[k] = C[k];
        a.T[k] = T[k];
    }
    a.init_imp = init_imp;
    a.imp = imp;
    a.rec = rec;
    hipLaunchKernelGGL(kb_rk4_update, dim3(blocks), dim3(kBlock), 0, s, t, P, pos, Tx, Ty, ux, uy, st, a);
    PFM_LAUNCH_CHECK();
    return PFM_OK;
}


}  // namespace pfm

using namespace pfm;

extern "C" {

int pfm_rk4_particles(const pfm_grid* gp, int64_t n, float* pos, float* T_x, float* T_y,
                      float* T_z, const float* u_x, const float* u_y, const float* u_z, float dt,
                      int advect_psi, pfm_stream_t stream) {
    GridDev g;
    int rc = make_grid(gp, &g);
    if (rc) return rc;
    PFM_CHECK_ARG(n >= 0, "n must be >= 0");
    if (n == 0) return PFM_OK;
    PFM_CHECK_ARG(pos && T_x && T_y && u_x && u_y, "null pointer");
    PFM_CHECK_ARG(g.dim == 2 || (T_z && u_z), "3D needs T_z and u_z");
    if (g.dim == 3)
        hipLaunchKernelGGL((k_rk4_particles<3, false>), dim3(grid_for(n)), dim3(kBlock), 0, as_stream(stream),
                           g, n, pos, T_x, T_y, T_z, u_x, u_y, u_z, dt, advect_psi);
    else
        hipLaunchKernelGGL((k_rk4_particles<2, false>), dim3(grid_for(n)), dim3(kBlock), 0, as_stream(stream),
                           g, n, pos, T_x, T_y, T_z, u_x, u_y, u_z, dt, advect_psi);
    PFM_LAUNCH_CHECK();
    return PFM_OK;
}

int pfm_rk4_particles_p2g_record(const pfm_grid* gp, int64_t n, float* pos, float* T_x, float* T_y,
                                 float* T_z, const float* u_x, const float* u_y, const float* u_z, float dt,
                                 const float* Tin_x, const float* Tin_y, const float* Tin_z, const float* iC_x,
                                 const float* iC_y, const float* iC_z, const float* init_imp, float* records,
                                 pfm_stream_t stream) {
    GridDev g;
    int rc = make_grid(gp, &g);
    if (rc) return rc;
    PFM_CHECK_ARG(n >= 0, "n must be >= 0");
    if (n == 0) return PFM_OK;
    PFM_CHECK_ARG(pos && T_x && T_y && u_x && u_y && Tin_x && Tin_y && iC_x && iC_y && init_imp && records,
                  "null pointer");
    PFM_CHECK_ARG(g.dim == 2 || (T_z && u_z && Tin_z && iC_z), "3D needs the z columns/rows");
    PFM_CHECK_ARG(((uintptr_t)records & 15) == 0, "records must be 16-byte aligned");
    RecArgs ra;
    ra.Tin[0] = Tin_x; ra.Tin[1] = Tin_y; ra.Tin[2] = Tin_z;
    ra.C[0] = iC_x; ra.C[1] = iC_y; ra.C[2] = iC_z;
    ra.imp = init_imp;
    ra.rec = records;
    if (g.dim == 3)
        hipLaunchKernelGGL((k_rk4_particles<3, true>), dim3(grid_for(n)), dim3(kBlock), 0, as_stream(stream), g,
                           n, pos, T_x, T_y, T_z, u_x, u_y, u_z, dt, 1, ra);
    else
        hipLaunchKernelGGL((k_rk4_particles<2, true>), dim3(grid_for(n)), dim3(kBlock), 0, as_stream(stream), g,
                           n, pos, T_x, T_y, T_z, u_x, u_y, u_z, dt, 1, ra);
    PFM_LAUNCH_CHECK();
    return PFM_OK;
}

int pfm_rk4_particles_p2g_record_fold(const pfm_grid* gp, int64_t n, float* pos, const float* u_x,
                                      const float* u_y, const float* u_z, float dt, float* Tin_x, float* Tin_y,
                                      float* Tin_z, const float* iC_x, const float* iC_y, const float* iC_z,
                                      const float* init_imp, float* records, pfm_stream_t stream) {
    GridDev g;
    int rc = make_grid(gp, &g);
    if (rc) return rc;
    PFM_CHECK_ARG(g.dim == 3, "the folded RK4 + record path is 3D");
    PFM_CHECK_ARG(n >= 0, "n must be >= 0");
    if (n == 0) return PFM_OK;
    PFM_CHECK_ARG(pos && u_x && u_y && u_z && Tin_x && Tin_y && Tin_z && iC_x && iC_y && iC_z && init_imp &&
                      records,
                  "null pointer");
    PFM_CHECK_ARG(((uintptr_t)records & 15) == 0, "records must be 16-byte aligned");
    RecArgs ra;
    ra.Tin[0] = Tin_x; ra.Tin[1] = Tin_y; ra.Tin[2] = Tin_z;
    ra.Tfold[0] = Tin_x; ra.Tfold[1] = Tin_y; ra.Tfold[2] = Tin_z;
    ra.C[0] = iC_x; ra.C[1] = iC_y; ra.C[2] = iC_z;
    ra.imp = init_imp;
    ra.rec = records;
    hipLaunchKernelGGL((k_rk4_particles<3, true, true>), dim3(grid_for(n)), dim3(kBlock), 0, as_stream(stream), g,
                       n, pos, nullptr, nullptr, nullptr, u_x, u_y, u_z, dt, 1, ra);
    PFM_LAUNCH_CHECK();
    return PFM_OK;
}

int pfm_g2p(const pfm_grid* gp, int64_t n, const float* pos, const float* u_x, const float* u_y,
            const float* u_z, float* imp, float* C_x, float* C_y, float* C_z,
            pfm_stream_t stream) {
    GridDev g;
    int rc = make_grid(gp, &g);
    if (rc) return rc;
    PFM_CHECK_ARG(n >= 0, "n must be >= 0");
    if (n == 0) return PFM_OK;
    PFM_CHECK_ARG(pos && u_x && u_y, "null pointer");
    PFM_CHECK_ARG(g.dim == 2 || u_z, "3D needs u_z");
    const bool grad = C_x != nullptr;
    PFM_CHECK_ARG(!grad || (C_y && (g.dim == 2 || C_z)), "C_x given without C_y/C_z");
    PFM_CHECK_ARG(imp || grad, "nothing to write: imp and C_x are both NULL");
    dim3 gr(grid_for(n)), bl(kBlock);
    hipStream_t s = as_stream(stream);
    if (g.dim == 3) {
        if (grad)
            hipLaunchKernelGGL((k_g2p<3, true>), gr, bl, 0, s, g, n, pos, u_x, u_y, u_z, imp, C_x, C_y, C_z);
        else
            hipLaunchKernelGGL((k_g2p<3, false>), gr, bl, 0, s, g, n, pos, u_x, u_y, u_z, imp, C_x, C_y, C_z);
    } else {
        if (grad)
            hipLaunchKernelGGL((k_g2p<2, true>), gr, bl, 0, s, g, n, pos, u_x, u_y, u_z, imp, C_x, C_y, C_z);
        else
            hipLaunchKernelGGL((k_g2p<2, false>), gr, bl, 0, s, g, n, pos, u_x, u_y, u_z, imp, C_x, C_y, C_z);
    }
    PFM_LAUNCH_CHECK();
    return PFM_OK;
}

int pfm_update_T(int dim, int64_t n, const float* Tgm_x, const float* Tgm_y, const float* Tgm_z,
                 const float* Tin_x, const float* Tin_y, const float* Tin_z, float* Tout_x,
                 float* Tout_y, float* Tout_z, pfm_stream_t stream) {
    PFM_CHECK_ARG(dim == 2 || dim == 3, "dim must be 2 or 3");
    PFM_CHECK_ARG(n >= 0, "n must be >= 0");
    if (n == 0) return PFM_OK;
    PFM_CHECK_ARG(Tgm_x && Tgm_y && Tin_x && Tin_y && Tout_x && Tout_y, "null pointer");
    PFM_CHECK_ARG(dim == 2 || (Tgm_z && Tin_z && Tout_z), "3D needs z columns");
    if (dim == 3)
        hipLaunchKernelGGL(k_update_T<3>, dim3(grid_for(n)), dim3(kBlock), 0, as_stream(stream), n,
                           Tgm_x, Tgm_y, Tgm_z, Tin_x, Tin_y, Tin_z, Tout_x, Tout_y, Tout_z);
    else
        hipLaunchKernelGGL(k_update_T<2>, dim3(grid_for(n)), dim3(kBlock), 0, as_stream(stream), n,
                           Tgm_x, Tgm_y, Tgm_z, Tin_x, Tin_y, Tin_z, Tout_x, Tout_y, Tout_z);
    PFM_LAUNCH_CHECK();
    return PFM_OK;
}

int pfm_update_particles_imp(int dim, int64_t n, const float* T_x, const float* T_y,
                             const float* T_z, const float* init_imp, float* imp,
                             pfm_stream_t stream) {
    PFM_CHECK_ARG(dim == 2 || dim == 3, "dim must be 2 or 3");
    PFM_CHECK_ARG(n >= 0, "n must be >= 0");
    if (n == 0) return PFM_OK;
    PFM_CHECK_ARG(T_x && T_y && init_imp && imp && (dim == 2 || T_z), "null pointer");
    if (dim == 3)
        hipLaunchKernelGGL(k_update_imp<3>, dim3(grid_for(n)), dim3(kBlock), 0, as_stream(stream), n,
                           T_x, T_y, T_z, init_imp, imp);
    else
        hipLaunchKernelGGL(k_update_imp<2>, dim3(grid_for(n)), dim3(kBlock), 0, as_stream(stream), n,
                           T_x, T_y, T_z, init_imp, imp);
    PFM_LAUNCH_CHECK();
    return PFM_OK;
}

int pfm_update_particles_grad_m(int dim, int64_t n, float* C_x, float* C_y, float* C_z,
                                const float* iC_x, const float* iC_y, const float* iC_z,
                                const float* Tgm_x, const float* Tgm_y, const float* Tgm_z,
                                pfm_stream_t stream) {
    PFM_CHECK_ARG(dim == 2 || dim == 3, "dim must be 2 or 3");
    PFM_CHECK_ARG(n >= 0, "n must be >= 0");
    if (n == 0) return PFM_OK;
    PFM_CHECK_ARG(C_x && C_y && iC_x && iC_y && Tgm_x && Tgm_y, "null pointer");
    PFM_CHECK_ARG(dim == 2 || (C_z && iC_z && Tgm_z), "3D needs z rows");
    if (dim == 3)
        hipLaunchKernelGGL(k_update_grad_m<3>, dim3(grid_for(n)), dim3(kBlock), 0, as_stream(stream), n,
                           C_x, C_y, C_z, iC_x, iC_y, iC_z, Tgm_x, Tgm_y, Tgm_z);
    else
        hipLaunchKernelGGL(k_update_grad_m<2>, dim3(grid_for(n)), dim3(kBlock), 0, as_stream(stream), n,
                           C_x, C_y, C_z, iC_x, iC_y, iC_z, Tgm_x, Tgm_y, Tgm_z);
    PFM_LAUNCH_CHECK();
    return PFM_OK;
}

int pfm_reset_T_identity(int dim, int64_t n, float* T_x, float* T_y, float* T_z,
                         pfm_stream_t stream) {
    PFM_CHECK_ARG(dim == 2 || dim == 3, "dim must be 2 or 3");
    PFM_CHECK_ARG(n >= 0, "n must be >= 0");
    if (n == 0) return PFM_OK;
    PFM_CHECK_ARG(T_x && T_y && (dim == 2 || T_z), "null pointer");
    if (dim == 3)
        hipLaunchKernelGGL(k_reset_T<3>, dim3(grid_for(n)), dim3(kBlock), 0, as_stream(stream), n, T_x, T_y, T_z);
    else
        hipLaunchKernelGGL(k_reset_T<2>, dim3(grid_for(n)), dim3(kBlock), 0, as_stream(stream), n, T_x, T_y, T_z);
    PFM_LAUNCH_CHECK();
    return PFM_OK;
}

int pfm_seed_particles(const pfm_grid* gp, int ppc_axis, float dist, int64_t n, float* pos,
                       int64_t seed, int order, pfm_stream_t stream) {
    GridDev g;
    int rc = make_grid(gp, &g);
    if (rc) return rc;
    PFM_CHECK_ARG(ppc_axis >= 1, "ppc_axis must be >= 1");
    PFM_CHECK_ARG(order == 0 || order == 1, "order must be 0 or 1");
    PFM_CHECK_ARG(n >= 0 && pos, "bad particle buffer");
    if (n == 0) return PFM_OK;
    if (g.dim == 3)
        hipLaunchKernelGGL(k_seed<3>, dim3(grid_for(n)), dim3(kBlock), 0, as_stream(stream), g, ppc_axis,
                           dist, n, pos, seed, order);
    else
        hipLaunchKernelGGL(k_seed<2>, dim3(grid_for(n)), dim3(kBlock), 0, as_stream(stream), g, ppc_axis,
                           dist, n, pos, seed, order);
    PFM_LAUNCH_CHECK();
    return PFM_OK;
}

}  // extern "C"
