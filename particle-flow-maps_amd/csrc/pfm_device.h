// pfm_device.h -- device-side building blocks shared by the libpfm kernels
// (gfx950, wave64).  Quadratic B-spline weights in closed form, the separable
// 4^d-point MAC interpolation, and block reductions.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pfm {

constexpr int kBlock = 256;

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

// Quadratic B-spline values and derivatives at the four stencil distances
// f+1, f, f-1, f-2 (f in [0,1)), i.e. N_2/dN_2 of 3D/taichi_utils.py:252-272
// evaluated at x_p - x_i for i = iu-1 .. iu+2.  Every intermediate is exact
// in fp32 (all operands are multiples of 2^-23 below 4), so these closed forms
// give the same bits as evaluating the reference's branchy N_2/dN_2 on the
// four distances.
__device__ __forceinline__ void bspline4(float f, float w[4], float dw[4]) {
    const bool lo = f < 0.5f;
    const bool hi = f > 0.5f;
    const float a = 0.5f - f;
    const float b = 1.5f - f;
    const float c = 0.5f + f;
    const float d = f - 0.5f;
    const float e = 1.0f - f;
    w[0] = lo ? 0.5f * (a * a) : 0.0f;
    w[1] = lo ? 0.75f - f * f : 0.5f * (b * b);
    w[2] = hi ? 0.75f - e * e : 0.5f * (c * c);
    w[3] = hi ? 0.5f * (d * d) : 0.0f;
    const float t = 2.0f * f;
    dw[0] = lo ? 0.5f * (t - 1.0f) : 0.0f;
    dw[1] = lo ? -t : 0.5f * (t - 3.0f);
    dw[2] = hi ? 2.0f * e : 0.5f * (1.0f + t);
    dw[3] = hi ? 0.5f * (t - 1.0f) : 0.0f;
}

__device__ __forceinline__ void bspline4_val(float f, float w[4]) {
    const bool lo = f < 0.5f;
    const bool hi = f > 0.5f;
    const float a = 0.5f - f;
    const float b = 1.5f - f;
    const float c = 0.5f + f;
    const float d = f - 0.5f;
    const float e = 1.0f - f;
    w[0] = lo ? 0.5f * (a * a) : 0.0f;
    w[1] = lo ? 0.75f - f * f : 0.5f * (b * b);
    w[2] = hi ? 0.75f - e * e : 0.5f * (c * c);
    w[3] = hi ? 0.5f * (d * d) : 0.0f;
}

// N_2 on an arbitrary argument (3D/taichi_utils.py:252-260), used by P2G.
__device__ __forceinline__ float N2(float x) {
    float ax = fabsf(x);
    float t = 1.5f - ax;
    float r = ax < 1.5f ? 0.5f * (t * t) : 0.0f;
    return ax < 0.5f ? 0.75f - ax * ax : r;
}

// Unaligned 16-byte row load of 4 consecutive floats (gfx950 global loads
// only need dword alignment).
typedef float f4u __attribute__((ext_vector_type(4), aligned(4)));

__device__ __forceinline__ void load_row4(const float* __restrict__ row, int i0, int ext,
                                          float v[4]) {
    // elements i0-1 .. i0+2 clamped to ext-1 (sample(), 3D/taichi_utils.py:48-54);
    // i0 >= 1 always (s >= 1) and i0 <= ext-2, ext >= 4.
    const int shift = (i0 + 2 > ext - 1) ? 1 : 0;
    const f4u q = *reinterpret_cast<const f4u*>(row + (i0 - 1 - shift));
    if (shift) {
        v[0] = q.y; v[1] = q.z; v[2] = q.w; v[3] = q.w;
    } else {
        v[0] = q.x; v[1] = q.y; v[2] = q.z; v[3] = q.w;
    }
}

// Per-axis stencil set-up of interp_grad_2 (3D/taichi_utils.py:278-287):
// u = p/dx - BL, s = clamp(u, 1, hi), iu = floor(s), f = s - iu.
__device__ __forceinline__ void axis_setup(float p, float inv_dx, float BL, float hi, int& i0,
                                           float& f) {
    float u = p * inv_dx - BL;
    float s = fmaxf(1.0f, fminf(u, hi));
    float fl = floorf(s);
    i0 = (int)fl;
    f = s - fl;
}

// Interpolate MAC component A (a face field of extents S) at p:
// value and (if GRAD) gradient, separable form of interp_grad_2
// (3D/taichi_utils.py:274-307; 2D :214-248).
template <int DIM, int A, bool GRAD>
__device__ __forceinline__ void interp_comp(const float* __restrict__ fld, const GridDev& g,
                                            const float p[3], float& val, float grad[3]) {
    int S[3];
    float hi[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        S[c] = g.res[c] + (c == A ? 1 : 0);
        hi[c] = c == A ? g.hi_f[c] : g.hi_c[c];
    }
    int i0[3];
    float f[3];
    float w[3][4], dw[3][4];
#pragma unroll
    for (int c = 0; c < DIM; ++c) {
        axis_setup(p[c], g.inv_dx, c == A ? 0.0f : 0.5f, hi[c], i0[c], f[c]);
        if (GRAD)
            bspline4(f[c], w[c], dw[c]);
        else
            bspline4_val(f[c], w[c]);
    }
    int ix[4];
#pragma unroll
    for (int m = 0; m < 4; ++m) ix[m] = min(i0[0] + m - 1, S[0] - 1);
    if (DIM == 3) {
        int iy[4];
#pragma unroll
        for (int m = 0; m < 4; ++m) iy[m] = min(i0[1] + m - 1, S[1] - 1);
        float v0 = 0.f, gx = 0.f, gy = 0.f, gz = 0.f;
        // One x-slab (4 rows of 4) per iteration keeps 16 loaded values live
        // instead of 64, which is what bounds the register count here.
#pragma unroll 1
        for (int a = 0; a < 4; ++a) {
            float c_v = 0.f, c_dy = 0.f, c_dz = 0.f;
#pragma unroll
            for (int b = 0; b < 4; ++b) {
                const float* row = fld + ((int64_t)ix[a] * S[1] + iy[b]) * S[2];
                float v[4];
                load_row4(row, i0[2], S[2], v);
                float rz = w[2][0] * v[0] + w[2][1] * v[1] + w[2][2] * v[2] + w[2][3] * v[3];
                c_v += w[1][b] * rz;
                if (GRAD) {
                    float rdz = dw[2][0] * v[0] + dw[2][1] * v[1] + dw[2][2] * v[2] + dw[2][3] * v[3];
                    c_dy += dw[1][b] * rz;
                    c_dz += w[1][b] * rdz;
                }
            }
            v0 += w[0][a] * c_v;
            if (GRAD) {
                gx += dw[0][a] * c_v;
                gy += w[0][a] * c_dy;
                gz += w[0][a] * c_dz;
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
        for (int a = 0; a < 4; ++a) {
            const float* row = fld + (int64_t)ix[a] * S[1];
            float v[4];
            load_row4(row, i0[1], S[1], v);
            float ry = w[1][0] * v[0] + w[1][1] * v[1] + w[1][2] * v[2] + w[1][3] * v[3];
            v0 += w[0][a] * ry;
            if (GRAD) {
                float rdy = dw[1][0] * v[0] + dw[1][1] * v[1] + dw[1][2] * v[2] + dw[1][3] * v[3];
                gx += dw[0][a] * ry;
                gy += w[0][a] * rdy;
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

// interp_u_MAC_grad (3D/taichi_utils.py:312-317): u[a] and G[a][c] = du_a/dx_c.
template <int DIM, bool GRAD>
__device__ __forceinline__ void interp_mac(const float* __restrict__ ux,
                                           const float* __restrict__ uy,
                                           const float* __restrict__ uz, const GridDev& g,
                                           const float p[3], float u[3], float G[3][3]) {
    interp_comp<DIM, 0, GRAD>(ux, g, p, u[0], G[0]);
    interp_comp<DIM, 1, GRAD>(uy, g, p, u[1], G[1]);
    if (DIM == 3) {
        interp_comp<DIM, 2, GRAD>(uz, g, p, u[2], G[2]);
    } else {
        u[2] = 0.f;
        G[2][0] = G[2][1] = G[2][2] = 0.f;
    }
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
