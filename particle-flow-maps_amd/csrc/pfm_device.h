// pfm_device.h -- device-side building blocks shared by the libpfm kernels
// (gfx950, wave64).  Quadratic B-spline weights in closed form, the separable
// MAC interpolation over the 3^d live nodes, and block reductions.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pfm {

constexpr int kBlock = 256;

// Workgroups are dealt to the 8 XCDs round robin by id.  This block index is
// remapped so that XCD b % 8 takes a contiguous range of [0, gridDim.x), so
// that lines read by neighbouring blocks are fetched into one XCD's L2 rather
// than several.  The map is a bijection for any block count.  (The gather
// kernels bound by the address unit -- RK4, G2P, the midpoint pass -- fetch
// less with it but run 1-2% slower, so they keep the plain order.)
__device__ __forceinline__ int64_t xcd_contiguous_block() {
    const int nb = gridDim.x, x = blockIdx.x & 7, k = blockIdx.x >> 3;
    return (int64_t)x * (nb >> 3) + min(x, nb & 7) + k;
}

// Grid description passed by value to kernels.
struct GridDev {
    int dim;
    int res[3];    // cells per axis (res[2] = 1 in 2D)
    float dx;
    float inv_dx;  // 1/dx as a float constant, the reference's `1./dx`
    // Clamp bound of interp_grad_2 for a face field whose extent along an
    // axis is res (hi_c) or res+1 (hi_f): float(extent - 2 - 1e-6)
    // (3D/taichi_utils.py:282-284 with eps = 1e-6 from :4).
    float hi_c[3];
    float hi_f[3];
};

__device__ __forceinline__ int64_t face_count(const GridDev& g, int a) {
    int64_t n = 1;
    for (int c = 0; c < g.dim; ++c) n *= (int64_t)(g.res[c] + (c == a ? 1 : 0));
    return n;
}

// N_2 on an arbitrary argument (3D/taichi_utils.py:252-260), used by P2G.
__device__ __forceinline__ float N2(float x) {
    float ax = fabsf(x);
    float t = 1.5f - ax;
    float r = ax < 1.5f ? 0.5f * (t * t) : 0.0f;
    return ax < 0.5f ? 0.75f - ax * ax : r;
}

// Per-axis stencil of interp_grad_2 (3D/taichi_utils.py:278-307): u = p/dx -
// BL, s = clamp(u, 1, hi), iu = floor(s), f = s - iu.  The reference sums the
// 4 nodes iu-1 .. iu+2, but a quadratic B-spline is zero at one of them:
// N_2(f+1) = 0 when f >= 1/2 and N_2(f-2) = 0 when f <= 1/2.  The stencil
// therefore keeps the 3 live nodes j0 .. j0+2 with j0 = iu-1 (f < 1/2) or iu
// (f >= 1/2); the dropped term adds an exact zero in the reference.
struct AxisStencil {
    int j0;
    float w[3];
    float dw[3];
};

template <bool GRAD>
__device__ __forceinline__ void axis_stencil(float p, float inv_dx, float BL, float hi, AxisStencil& a) {
    const float u = p * inv_dx - BL;
    const float s = fmaxf(1.0f, fminf(u, hi));
    const float fl = floorf(s);
    const float f = s - fl;
    const bool up = f >= 0.5f;
    a.j0 = (int)fl - 1 + (up ? 1 : 0);
    // d = s - (j0 + 1) in [-1/2, 1/2): the distances to the live nodes are
    // d + 1, d, d - 1.  With f in [0, 1) every operation below is exact, so
    // the weights are bit-equal to N_2/dN_2 at those distances.
    const float d = up ? f - 1.0f : f;
    const float a0 = 0.5f - d, a2 = 0.5f + d;
    a.w[0] = 0.5f * (a0 * a0);
    a.w[1] = 0.75f - d * d;
    a.w[2] = 0.5f * (a2 * a2);
    if (GRAD) {
        a.dw[0] = d - 0.5f;
        a.dw[1] = -2.0f * d;
        a.dw[2] = d + 0.5f;
    }
}

// The stencils of all MAC components at one point: along axis c, component c
// uses the face stencil (BL = 0, extent res+1) and the others the cell
// stencil (BL = 0.5, extent res), so 2 per axis serve all components.
template <int DIM, bool GRAD>
struct MacStencil {
    AxisStencil F[3];  // BL = 0, extent res[c] + 1
    AxisStencil C[3];  // BL = 0.5, extent res[c]
    __device__ __forceinline__ void build(const GridDev& g, const float p[3]) {
#pragma unroll
        for (int c = 0; c < DIM; ++c) {
            axis_stencil<GRAD>(p[c], g.inv_dx, 0.0f, g.hi_f[c], F[c]);
            axis_stencil<GRAD>(p[c], g.inv_dx, 0.5f, g.hi_c[c], C[c]);
        }
    }
};

// Raw buffer resource over a face field (word 3 = 0x00020000 on gfx9-family
// parts: raw 32-bit dwords).  Loads through it take a 32-bit byte offset, so a
// row address costs one VALU add; offsets past num_records read zero.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t face_rsrc(const float* fld, uint32_t bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<float*>(fld), (short)0, (int)bytes, 0x00020000);
}

typedef float v3f __attribute__((ext_vector_type(3)));

// Rows of MAC component A under its stencil: the 3^(d-1) rows of 3 nodes
// along the contiguous axis, one 12-byte load each.  The clamp
// s <= extent - 2 - eps keeps the live nodes j0 .. j0+2 inside
// [0, extent-1] (the reference's sample() clamps only the fourth, zero-weight
// node of interp_grad_2), so no row needs folding.
template <int DIM>
struct Rows {
    v3f q[DIM == 3 ? 9 : 3];
};

template <int DIM, int A, bool GRAD>
__device__ __forceinline__ void load_rows(const float* __restrict__ fld, const GridDev& g,
                                          const MacStencil<DIM, GRAD>& st, Rows<DIM>& r) {
    const AxisStencil& s0 = A == 0 ? st.F[0] : st.C[0];
    const AxisStencil& s1 = A == 1 ? st.F[1] : st.C[1];
    const int S0 = g.res[0] + (A == 0 ? 1 : 0);
    const int S1 = g.res[1] + (A == 1 ? 1 : 0);
    if (DIM == 3) {
        const AxisStencil& s2 = A == 2 ? st.F[2] : st.C[2];
        const int S2 = g.res[2] + (A == 2 ? 1 : 0);
        const uint32_t rowb = (uint32_t)S2 * 4u;
        const uint32_t plane = (uint32_t)S1 * rowb;
        const __amdgpu_buffer_rsrc_t rs = face_rsrc(fld, plane * (uint32_t)S0);
        const uint32_t base = (uint32_t)s0.j0 * plane + (uint32_t)s1.j0 * rowb + (uint32_t)s2.j0 * 4u;
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int b = 0; b < 3; ++b)
                r.q[3 * a + b] = __builtin_amdgcn_raw_buffer_load_b96(rs, base + a * plane + b * rowb, 0, 0);
    } else {
        const uint32_t rowb = (uint32_t)S1 * 4u;
        const __amdgpu_buffer_rsrc_t rs = face_rsrc(fld, rowb * (uint32_t)S0);
        const uint32_t base = (uint32_t)s0.j0 * rowb + (uint32_t)s1.j0 * 4u;
#pragma unroll
        for (int a = 0; a < 3; ++a) r.q[a] = __builtin_amdgcn_raw_buffer_load_b96(rs, base + a * rowb, 0, 0);
    }
}

// Value and (if GRAD) gradient of component A from its rows: the separable
// form of interp_grad_2 (3D/taichi_utils.py:274-307; 2D :214-248).
template <int DIM, int A, bool GRAD>
__device__ __forceinline__ void combine_rows(const GridDev& g, const MacStencil<DIM, GRAD>& st, const Rows<DIM>& r,
                                             float& val, float grad[3]) {
    const AxisStencil& s0 = A == 0 ? st.F[0] : st.C[0];
    const AxisStencil& s1 = A == 1 ? st.F[1] : st.C[1];
    if (DIM == 3) {
        const AxisStencil& s2 = A == 2 ? st.F[2] : st.C[2];
        float v0 = 0.f, gx = 0.f, gy = 0.f, gz = 0.f;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            float c_v = 0.f, c_dy = 0.f, c_dz = 0.f;
#pragma unroll
            for (int b = 0; b < 3; ++b) {
                const v3f q = r.q[3 * a + b];
                const float rz = s2.w[0] * q.x + s2.w[1] * q.y + s2.w[2] * q.z;
                c_v += s1.w[b] * rz;
                if (GRAD) {
                    const float rdz = s2.dw[0] * q.x + s2.dw[1] * q.y + s2.dw[2] * q.z;
                    c_dy += s1.dw[b] * rz;
                    c_dz += s1.w[b] * rdz;
                }
            }
            v0 += s0.w[a] * c_v;
            if (GRAD) {
                gx += s0.dw[a] * c_v;
                gy += s0.w[a] * c_dy;
                gz += s0.w[a] * c_dz;
            }
        }
        val = v0;
        if (GRAD) {
            grad[0] = gx * g.inv_dx;
            grad[1] = gy * g.inv_dx;
            grad[2] = gz * g.inv_dx;
        }
    } else {
        float v0 = 0.f, gx = 0.f, gy = 0.f;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const v3f q = r.q[a];
            const float ry = s1.w[0] * q.x + s1.w[1] * q.y + s1.w[2] * q.z;
            v0 += s0.w[a] * ry;
            if (GRAD) {
                const float rdy = s1.dw[0] * q.x + s1.dw[1] * q.y + s1.dw[2] * q.z;
                gx += s0.dw[a] * ry;
                gy += s0.w[a] * rdy;
            }
        }
        val = v0;
        if (GRAD) {
            grad[0] = gx * g.inv_dx;
            grad[1] = gy * g.inv_dx;
            grad[2] = 0.f;
        }
    }
}

// One MAC component `a` (chosen at run time) of a 3D face field at p, value
// and (if GRAD) gradient: the face stencil along a, the cell stencils across,
// the rows and their separable contraction exactly as load_rows and
// combine_rows form them for that component, so the result is bit-identical
// to interp_mac's; only one component's registers are live at a time.
template <bool GRAD>
__device__ __forceinline__ void sample_component_rt(const float* __restrict__ fld, const GridDev& g, int a,
                                                    const float p[3], float& val, float grad[3]) {
    AxisStencil s[3];
    int S[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const bool face = c == a;
        axis_stencil<GRAD>(p[c], g.inv_dx, face ? 0.0f : 0.5f, face ? g.hi_f[c] : g.hi_c[c], s[c]);
        S[c] = g.res[c] + (face ? 1 : 0);
    }
    const uint32_t rowb = (uint32_t)S[2] * 4u;
    const uint32_t plane = (uint32_t)S[1] * rowb;
    const __amdgpu_buffer_rsrc_t rs = face_rsrc(fld, plane * (uint32_t)S[0]);
    const uint32_t base = (uint32_t)s[0].j0 * plane + (uint32_t)s[1].j0 * rowb + (uint32_t)s[2].j0 * 4u;
    float v0 = 0.f, gx = 0.f, gy = 0.f, gz = 0.f;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        float c_v = 0.f, c_dy = 0.f, c_dz = 0.f;
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const v3f q = __builtin_amdgcn_raw_buffer_load_b96(rs, base + i * plane + j * rowb, 0, 0);
            const float rz = s[2].w[0] * q.x + s[2].w[1] * q.y + s[2].w[2] * q.z;
            c_v += s[1].w[j] * rz;
            if (GRAD) {
                const float rdz = s[2].dw[0] * q.x + s[2].dw[1] * q.y + s[2].dw[2] * q.z;
                c_dy += s[1].dw[j] * rz;
                c_dz += s[1].w[j] * rdz;
            }
        }
        v0 += s[0].w[i] * c_v;
        if (GRAD) {
            gx += s[0].dw[i] * c_v;
            gy += s[0].w[i] * c_dy;
            gz += s[0].w[i] * c_dz;
        }
    }
    val = v0;
    if (GRAD) {
        grad[0] = gx * g.inv_dx;
        grad[1] = gy * g.inv_dx;
        grad[2] = gz * g.inv_dx;
    }
}

// interp_u_MAC_grad (3D/taichi_utils.py:312-317): u[a] and G[a][c] = du_a/dx_c.
template <int DIM, bool GRAD>
__device__ __forceinline__ void interp_mac(const float* __restrict__ ux,
                                           const float* __restrict__ uy,
                                           const float* __restrict__ uz, const GridDev& g,
                                           const float p[3], float u[3], float G[3][3]) {
    MacStencil<DIM, GRAD> st;
    st.build(g, p);
    Rows<DIM> r0, r1;
    load_rows<DIM, 0, GRAD>(ux, g, st, r0);
    load_rows<DIM, 1, GRAD>(uy, g, st, r1);
    combine_rows<DIM, 0, GRAD>(g, st, r0, u[0], G[0]);
    combine_rows<DIM, 1, GRAD>(g, st, r1, u[1], G[1]);
    if (DIM == 3) {
        Rows<DIM> r2;
        load_rows<DIM, DIM == 3 ? 2 : 1, GRAD>(uz, g, st, r2);
        combine_rows<DIM, DIM == 3 ? 2 : 1, GRAD>(g, st, r2, u[2], G[2]);
    } else {
        u[2] = 0.f;
        G[2][0] = G[2][1] = G[2][2] = 0.f;
    }
}

// The rows of the three MAC components of one point kept across calls (the
// stages of a particle's RK4), each with the byte offset of its first row.
// A later point whose stencil starts at the same offset reads the same rows,
// so they are not gathered again: with exec masks only the lanes whose
// stencil moved issue the row loads, and the vector-memory address unit's
// work falls with the active lanes.  The values are those the loads would
// return, so the results do not change.
template <int DIM>
struct RowCache {
    Rows<DIM> r[DIM];
    uint32_t key[DIM];
    __device__ __forceinline__ void clear() {
#pragma unroll
        for (int a = 0; a < DIM; ++a) key[a] = 0xFFFFFFFFu;  // no row starts there (offsets < 2^31)
    }
};

template <int DIM, int A, bool GRAD>
__device__ __forceinline__ void load_rows_cached(const float* __restrict__ fld, const GridDev& g,
                                                 const MacStencil<DIM, GRAD>& st, Rows<DIM>& r, uint32_t& key) {
    const AxisStencil& s0 = A == 0 ? st.F[0] : st.C[0];
    const AxisStencil& s1 = A == 1 ? st.F[1] : st.C[1];
    const int S0 = g.res[0] + (A == 0 ? 1 : 0);
    const int S1 = g.res[1] + (A == 1 ? 1 : 0);
    if constexpr (DIM == 3) {
        const AxisStencil& s2 = A == 2 ? st.F[2] : st.C[2];
        const int S2 = g.res[2] + (A == 2 ? 1 : 0);
        const uint32_t rowb = (uint32_t)S2 * 4u;
        const uint32_t plane = (uint32_t)S1 * rowb;
        const uint32_t base = (uint32_t)s0.j0 * plane + (uint32_t)s1.j0 * rowb + (uint32_t)s2.j0 * 4u;
        if (base != key) {
            key = base;
            const __amdgpu_buffer_rsrc_t rs = face_rsrc(fld, plane * (uint32_t)S0);
#pragma unroll
            for (int a = 0; a < 3; ++a)
#pragma unroll
                for (int b = 0; b < 3; ++b)
                    r.q[3 * a + b] = __builtin_amdgcn_raw_buffer_load_b96(rs, base + a * plane + b * rowb, 0, 0);
        }
    } else {
        const uint32_t rowb = (uint32_t)S1 * 4u;
        const uint32_t base = (uint32_t)s0.j0 * rowb + (uint32_t)s1.j0 * 4u;
        if (base != key) {
            key = base;
            const __amdgpu_buffer_rsrc_t rs = face_rsrc(fld, rowb * (uint32_t)S0);
#pragma unroll
            for (int a = 0; a < 3; ++a) r.q[a] = __builtin_amdgcn_raw_buffer_load_b96(rs, base + a * rowb, 0, 0);
        }
    }
}

// interp_mac with the rows taken from `cache` where the stencil did not
// move; the contraction is combine_rows's, so the result is bit-identical.
template <int DIM, bool GRAD>
__device__ __forceinline__ void interp_mac_cached(const float* __restrict__ ux, const float* __restrict__ uy,
                                                  const float* __restrict__ uz, const GridDev& g, const float p[3],
                                                  float u[3], float G[3][3], RowCache<DIM>& cache) {
    MacStencil<DIM, GRAD> st;
    st.build(g, p);
    load_rows_cached<DIM, 0, GRAD>(ux, g, st, cache.r[0], cache.key[0]);
    load_rows_cached<DIM, 1, GRAD>(uy, g, st, cache.r[1], cache.key[1]);
    if constexpr (DIM == 3) load_rows_cached<3, 2, GRAD>(uz, g, st, cache.r[2], cache.key[2]);
    combine_rows<DIM, 0, GRAD>(g, st, cache.r[0], u[0], G[0]);
    combine_rows<DIM, 1, GRAD>(g, st, cache.r[1], u[1], G[1]);
    if constexpr (DIM == 3) {
        combine_rows<3, 2, GRAD>(g, st, cache.r[2], u[2], G[2]);
    } else {
        u[2] = 0.f;
        G[2][0] = G[2][1] = G[2][2] = 0.f;
    }
}

// ---- P2G particle record ---------------------------------------------------
// Record of one particle for the tiled P2G, kRec floats:
//   3D: [0..2] pos/dx, [3] 0, then a quad per face axis A at 4 + 4A holding
//       imp_A and row A of C (the tile pass of axis A reads two aligned quads)
//   2D: [0..1] pos/dx, [2..3] imp, [4..7] C rows (row r at 4 + 2r)
template <int DIM>
struct Pay {
    static constexpr int kPos = 0;
    static constexpr int kRec = DIM == 3 ? 16 : 8;
    static constexpr int imp(int r) { return DIM == 3 ? 4 + 4 * r : DIM + r; }
    static constexpr int C(int r, int c) { return DIM == 3 ? 5 + 4 * r + c : 2 * DIM + r * DIM + c; }
};

// P2G per-particle composition (3D/run_3D.py:460-479) and record write:
// with A = rows(T_grad_m) (A[r][c] = T_r[c]) and B = cols(T_init)
// (B[r][c] = T_c_init[r]), C <- A^T (init_C A) and imp <- (A^T B) init_imp.
// Without `compose`, C and imp are written as given (the 2D driver forms them
// in update_particles_grad_m / update_particles_imp).
template <int DIM>
__device__ __forceinline__ void form_p2g_record(float rec[Pay<DIM>::kRec], const float pos[3], float inv_dx,
                                                bool compose, const float A[3][3], const float B[3][3],
                                                const float C[3][3], const float imp[3]) {
    using P = Pay<DIM>;
#pragma unroll
    for (int k = 0; k < P::kRec; ++k) rec[k] = 0.f;
#pragma unroll
    for (int c = 0; c < DIM; ++c) rec[P::kPos + c] = pos[c] * inv_dx;
    float Co[3][3], mo[3];
    if (compose) {
        // every multiply-add is an explicit FMA, so kernels that form the same
        // products (update_T folded into the RK4 epilogue) round identically
        float P2[3][3], Tt[3][3];
#pragma unroll
        for (int r = 0; r < DIM; ++r)
#pragma unroll
            for (int c = 0; c < DIM; ++c) {
                float acc = C[r][0] * A[0][c];
#pragma unroll
                for (int k = 1; k < DIM; ++k) acc = __builtin_fmaf(C[r][k], A[k][c], acc);
                P2[r][c] = acc;
            }
#pragma unroll
        for (int r = 0; r < DIM; ++r)
#pragma unroll
            for (int c = 0; c < DIM; ++c) {
                float acc = A[0][r] * P2[0][c];
#pragma unroll
                for (int k = 1; k < DIM; ++k) acc = __builtin_fmaf(A[k][r], P2[k][c], acc);
                Co[r][c] = acc;
                float t = A[0][r] * B[0][c];
#pragma unroll
                for (int k = 1; k < DIM; ++k) t = __builtin_fmaf(A[k][r], B[k][c], t);
                Tt[r][c] = t;
            }
#pragma unroll
        for (int r = 0; r < DIM; ++r) {
            float acc = Tt[r][0] * imp[0];
#pragma unroll
            for (int c = 1; c < DIM; ++c) acc = __builtin_fmaf(Tt[r][c], imp[c], acc);
            mo[r] = acc;
        }
    } else {
#pragma unroll
        for (int r = 0; r < DIM; ++r) {
            mo[r] = imp[r];
#pragma unroll
            for (int c = 0; c < DIM; ++c) Co[r][c] = C[r][c];
        }
    }
#pragma unroll
    for (int r = 0; r < DIM; ++r) {
        rec[P::imp(r)] = mo[r];
#pragma unroll
        for (int c = 0; c < DIM; ++c) rec[P::C(r, c)] = Co[r][c];
    }
}

template <int DIM>
__device__ __forceinline__ void write_p2g_record(float* __restrict__ dst, const float pos[3], float inv_dx,
                                                 bool compose, const float A[3][3], const float B[3][3],
                                                 const float C[3][3], const float imp[3]) {
    using P = Pay<DIM>;
    float rec[P::kRec];
    form_p2g_record<DIM>(rec, pos, inv_dx, compose, A, B, C, imp);
    float4* d4 = reinterpret_cast<float4*>(dst);
#pragma unroll
    for (int k = 0; k < P::kRec / 4; ++k)
        d4[k] = make_float4(rec[4 * k], rec[4 * k + 1], rec[4 * k + 2], rec[4 * k + 3]);
}

// ---- binning (get_particles_id_in_every_cell, 3D/run_3D.py:342-351) --------
// cell = int(pos / dx) per axis, clamped to the grid (particle i, component stride n).
template <int DIM>
__device__ __forceinline__ int32_t particle_cell(const GridDev& g, const float* __restrict__ pos, int64_t i,
                                                 int64_t n) {
    int ci[3] = {0, 0, 0};
#pragma unroll
    for (int c = 0; c < DIM; ++c) {
        int v = (int)(pos[c * n + i] * g.inv_dx);
        ci[c] = max(0, min(v, g.res[c] - 1));
    }
    return (int32_t)(((int64_t)ci[0] * g.res[1] + ci[1]) * g.res[2] + ci[2]);
}

// The count atomics of one wave: each run of equal cells among consecutive
// valid lanes takes one atomic; every valid lane gets its slot in the cell.
__device__ __forceinline__ void count_slots(int32_t cell, bool valid, int32_t* __restrict__ counts,
                                            int32_t* __restrict__ slot_i) {
    const int lane = threadIdx.x & 63;
    const int32_t prev = __shfl_up(cell, 1, 64);
    const bool head = valid && (lane == 0 || prev != cell);
    const uint64_t heads = __ballot(head);
    // run start: highest head at or below this lane; run end: next head
    const uint64_t upto = heads & (lane == 63 ? ~0ull : ((2ull << lane) - 1));
    const int first = 63 - __clzll(upto | 1ull);
    const uint64_t above = lane == 63 ? 0ull : (heads & ~((2ull << lane) - 1));
    const uint64_t validm = __ballot(valid);
    int next = above ? __ffsll((long long)above) - 1 : 64;
    const int last_valid = 63 - __clzll(validm | 1ull);
    if (next > last_valid + 1) next = last_valid + 1;
    int32_t base = 0;
    if (head) base = atomicAdd(&counts[cell], next - lane);
    base = __shfl(base, first, 64);
    if (valid) *slot_i = base + (lane - first);
}

// ---- reductions ----------------------------------------------------------
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

__device__ __forceinline__ float wave_max(float v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = fmaxf(v, __shfl_xor(v, off, 64));
    return v;
}

// Block sum of one double (blockDim.x == kBlock); result valid in thread 0.
__device__ __forceinline__ double block_sum(double v, double* sh) {
    v = wave_sum(v);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) sh[wid] = v;
    __syncthreads();
    double r = 0.0;
    if (threadIdx.x == 0) {
        for (int i = 0; i < (int)(blockDim.x >> 6); ++i) r += sh[i];
    }
    __syncthreads();
    return r;
}

}  // namespace pfm
