/*
 * pfm_oracle.c -- CPU restatement of the particle-flow-map solver step.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP library:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it.  The product path never calls it.
 *
 * Every function follows the reference file:line cited above it, in fp32,
 * with the reference's operation order (left-to-right products, literal
 * 4^d-point B-spline loops, scatter P2G in particle order).  It is compiled
 * with -ffp-contract=off so results do not depend on FMA contraction.
 * Deliberate, documented differences from the reference:
 *   - dot products, means and the kinetic energy accumulate in double in a
 *     fixed sequential order (the reference uses fp32 atomics whose order is
 *     undefined; BASELINE asks FP64 dots);
 *   - MGPCG restriction sums the 2^d children of a coarse cell in a fixed
 *     lexicographic order (the reference uses fp32 atomics);
 *   - particle binning clamps the cell index to the grid (the reference writes
 *     out of bounds for particles that left the domain);
 *   - optional extensions (jittered seeding, relative tolerance) are off by
 *     default and reduce to the reference when off.
 *
 * Layout: grid fields are row-major with the last axis fastest, exactly the
 * numpy view of the reference's Taichi fields: a 3D face field u_x has shape
 * (rx+1, ry, rz) and element (i,j,k) at (i*ry + j)*rz + k.  2D fields are
 * (nx, ny) with j fastest.  Particle vector fields are SoA: component c of
 * particle p at [c*n + p].  A 2D run uses dim=2 and res[2]=1.
 */
#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define ORACLE_EPS 1e-6 /* 3D/taichi_utils.py:4 */
#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

typedef struct {
    int dim;
    int res[3];    /* cells per axis; res[2] = 1 in 2D */
    float dx;      /* 1/res_y */
    float inv_dx;  /* 1/dx, the reference's `1./dx` constant */
} grid_t;

static inline int64_t idx3(const int shp[3], int i, int j, int k) {
    return ((int64_t)i * shp[1] + j) * shp[2] + k;
}

/* face-array shape of axis a: res with res[a] + 1 */
static void face_shape(const grid_t* g, int a, int shp[3]) {
    shp[0] = g->res[0]; shp[1] = g->res[1]; shp[2] = g->res[2];
    shp[a] += 1;
}

static void make_grid(grid_t* g, int dim, const int* res, float dx) {
    g->dim = dim;
    g->res[0] = res[0]; g->res[1] = res[1]; g->res[2] = dim == 3 ? res[2] : 1;
    g->dx = dx;
    g->inv_dx = (float)(1.0 / (double)dx);
}

/* ---- 1. quadratic B-spline: 3D/taichi_utils.py:252-272 (2D :181-201) ---- */
float oracle_N2(float x) {
    float result = 0.0f;
    float abs_x = fabsf(x);
    if (abs_x < 0.5f) {
        result = 0.75f - abs_x * abs_x;
    } else if (abs_x < 1.5f) {
        float t = 1.5f - abs_x;
        result = 0.5f * (t * t);
    }
    return result;
}

float oracle_dN2(float x) {
    float result = 0.0f;
    float abs_x = fabsf(x);
    if (abs_x < 0.5f) {
        result = -2.0f * abs_x;
    } else if (abs_x < 1.5f) {
        result = 0.5f * (2.0f * abs_x - 3.0f);
    }
    if (x < 0.0f) result *= -1.0f;
    return result;
}

/* ---- 2. clamped fetch: 3D/taichi_utils.py:48-54 (2D :23-28) ---- */
static inline float sample3(const float* f, const int shp[3], int i, int j, int k) {
    if (i > shp[0] - 1) i = shp[0] - 1;
    if (i < 0) i = 0;
    if (j > shp[1] - 1) j = shp[1] - 1;
    if (j < 0) j = 0;
    if (k > shp[2] - 1) k = shp[2] - 1;
    if (k < 0) k = 0;
    return f[idx3(shp, i, j, k)];
}

/* ---- 3/5. interp_grad_2 (3D/taichi_utils.py:274-307, 2D :214-248) and
 *      interp_grad_2_imp (3D :326-368, 2D :263-304).
 * imp_form selects the _imp product order: dw = inv_dx*dN*N*N, partial +=
 * value*dw.  Otherwise partial += inv_dx*(value*dN*N*N).  In 2D the z factor
 * is the exact constant 1.0f so the products equal the 2D reference's. */
static void interp_grad(const grid_t* g, const float* vf, const int shp[3],
                        const float p[3], float BLx, float BLy, float BLz,
                        int imp_form, float* val, float grad[3]) {
    float u = p[0] * g->inv_dx - BLx;
    float v = p[1] * g->inv_dx - BLy;
    float w = g->dim == 3 ? p[2] * g->inv_dx - BLz : 0.0f;
    float hx = (float)((double)shp[0] - 2.0 - ORACLE_EPS);
    float hy = (float)((double)shp[1] - 2.0 - ORACLE_EPS);
    float hz = (float)((double)shp[2] - 2.0 - ORACLE_EPS);
    float s = fmaxf(1.0f, fminf(u, hx));
    float t = fmaxf(1.0f, fminf(v, hy));
    float l = g->dim == 3 ? fmaxf(1.0f, fminf(w, hz)) : 0.0f;
    float iu = floorf(s), iv = floorf(t), iw = floorf(l);
    float px = 0.0f, py = 0.0f, pz = 0.0f, interped = 0.0f;
    int k0 = g->dim == 3 ? -1 : 0, k1 = g->dim == 3 ? 2 : 0;
    /* The spline factors of offset o along an axis depend only on that axis,
     * so they are evaluated once per axis (the same values the point loop of
     * the reference computes for every point, so the products are unchanged). */
    float Nxa[4], dNxa[4], Nya[4], dNya[4], Nza[4], dNza[4];
    int ia[4], ja[4], ka[4];
    for (int o = -1; o <= 2; ++o) {
        float xp = s - (iu + (float)o);
        float yp = t - (iv + (float)o);
        float zp = l - (iw + (float)o);
        Nxa[o + 1] = oracle_N2(xp); dNxa[o + 1] = oracle_dN2(xp);
        Nya[o + 1] = oracle_N2(yp); dNya[o + 1] = oracle_dN2(yp);
        Nza[o + 1] = g->dim == 3 ? oracle_N2(zp) : 1.0f;
        dNza[o + 1] = g->dim == 3 ? oracle_dN2(zp) : 0.0f;
        ia[o + 1] = (int)(iu + (float)o);
        ja[o + 1] = (int)(iv + (float)o);
        ka[o + 1] = (int)(iw + (float)o);
    }
    for (int i = -1; i <= 2; ++i)
        for (int j = -1; j <= 2; ++j)
            for (int k = k0; k <= k1; ++k) {
                float value = sample3(vf, shp, ia[i + 1], ja[j + 1], ka[k + 1]);
                float Nx = Nxa[i + 1], Ny = Nya[j + 1];
                float dNx = dNxa[i + 1], dNy = dNya[j + 1];
                float Nz = Nza[k + 1];
                float dNz = dNza[k + 1];
                if (imp_form) {
                    float dwx = g->inv_dx * dNx * Ny * Nz;
                    float dwy = g->inv_dx * Nx * dNy * Nz;
                    float dwz = g->inv_dx * Nx * Ny * dNz;
                    px += value * dwx;
                    py += value * dwy;
                    pz += value * dwz;
                } else {
                    px += g->inv_dx * (value * dNx * Ny * Nz);
                    py += g->inv_dx * (value * Nx * dNy * Nz);
                    pz += g->inv_dx * (value * Nx * Ny * dNz);
                }
                interped += value * Nx * Ny * Nz;
            }
    *val = interped;
    grad[0] = px; grad[1] = py; grad[2] = pz;
}

/* ---- 4. interp_u_MAC_grad: 3D/taichi_utils.py:312-324 (2D :257-261).
 * Returns u[a] and G[a][c] = d u_a / d x_c (the `rows` form). */
static void interp_mac(const grid_t* g, const float* const* uf, const float p[3],
                       int imp_form, float u[3], float G[3][3]) {
    for (int a = 0; a < 3; ++a) {
        u[a] = 0.0f;
        G[a][0] = G[a][1] = G[a][2] = 0.0f;
    }
    for (int a = 0; a < g->dim; ++a) {
        int shp[3];
        face_shape(g, a, shp);
        float BL[3] = {0.5f, 0.5f, 0.5f};
        BL[a] = 0.0f;
        interp_grad(g, uf[a], shp, p, BL[0], BL[1], BL[2], imp_form, &u[a], G[a]);
    }
}

/* ---- 14. RK4_T_forward: 3D/run_3D.py:175-219, 2D/run_2D.py:304-341.
 * dT = M @ T with M = cols(grad u_x, grad u_y, grad u_z), i.e.
 * dT[r] = sum_c G[c][r] T[c].  pos, Tx, Ty, Tz are SoA (dim x n). */
static void matT_vec(int dim, float G[3][3], const float T[3], float out[3]) {
    for (int r = 0; r < 3; ++r) out[r] = 0.0f;
    for (int r = 0; r < dim; ++r) {
        float acc = G[0][r] * T[0];
        for (int c = 1; c < dim; ++c) acc = acc + G[c][r] * T[c];
        out[r] = acc;
    }
}

void oracle_rk4_particles(int dim, const int* res, float dx, int64_t n, float* pos,
                          float* Tx, float* Ty, float* Tz, const float* ux, const float* uy,
                          const float* uz, float dt, int advect_psi) {
    grid_t g;
    make_grid(&g, dim, res, dx);
    const float* uf[3] = {ux, uy, uz};
    float* Ts[3] = {Tx, Ty, Tz};
    float half_dt = 0.5f * dt;
    float sixth_dt = dt * 1.0f / 6.0f;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
        float p0[3] = {0, 0, 0}, T0[3][3] = {{0}};
        for (int c = 0; c < dim; ++c) p0[c] = pos[c * n + i];
        for (int m = 0; m < dim; ++m)
            for (int c = 0; c < dim; ++c) T0[m][c] = Ts[m][c * n + i];
        float u[4][3], dT[4][3][3], G[3][3];
        float pk[3] = {p0[0], p0[1], p0[2]}, Tk[3][3];
        memcpy(Tk, T0, sizeof(Tk));
        float stage_h[3] = {half_dt, half_dt, 1.0f * dt};
        for (int st = 0; st < 4; ++st) {
            interp_mac(&g, uf, pk, 0, u[st], G);
            for (int m = 0; m < dim; ++m) matT_vec(dim, G, Tk[m], dT[st][m]);
            if (st < 3) {
                float h = stage_h[st];
                for (int c = 0; c < dim; ++c) pk[c] = p0[c] + h * u[st][c];
                for (int m = 0; m < dim; ++m)
                    for (int c = 0; c < dim; ++c) Tk[m][c] = T0[m][c] - h * dT[st][m][c];
            }
        }
        if (advect_psi)
            for (int c = 0; c < dim; ++c) {
                float s = u[0][c] + 2.0f * u[1][c];
                s = s + 2.0f * u[2][c];
                s = s + u[3][c];
                pos[c * n + i] = p0[c] + sixth_dt * s;
            }
        for (int m = 0; m < dim; ++m)
            for (int c = 0; c < dim; ++c) {
                float s = dT[0][m][c] + 2.0f * dT[1][m][c];
                s = s + 2.0f * dT[2][m][c];
                s = s + dT[3][m][c];
                Ts[m][c * n + i] = T0[m][c] - sixth_dt * s;
            }
    }
}

/* ---- 12. RK4_grid: 3D/run_3D.py:144-173, 2D/run_2D.py:233-261.
 * Backward RK4 of the faces of axis `axis`: dT = G @ T (rows form).
 * psi and T are SoA (dim x Fa). */
void oracle_rk4_grid(int dim, const int* res, float dx, int axis, float* psi, float* T,
                     const float* ux, const float* uy, const float* uz, float dt) {
    grid_t g;
    make_grid(&g, dim, res, dx);
    const float* uf[3] = {ux, uy, uz};
    int shp[3];
    face_shape(&g, axis, shp);
    int64_t nf = (int64_t)shp[0] * shp[1] * shp[2];
    float neg_dt = -1.0f * dt;
    float hs[3] = {0.5f * neg_dt, 0.5f * neg_dt, 1.0f * neg_dt};
    float sixth = neg_dt * 1.0f / 6.0f;
#pragma omp parallel for schedule(static)
    for (int64_t f = 0; f < nf; ++f) {
        float p0[3] = {0, 0, 0}, T0[3] = {0, 0, 0};
        for (int c = 0; c < dim; ++c) {
            p0[c] = psi[c * nf + f];
            T0[c] = T[c * nf + f];
        }
        float u[4][3], dT[4][3], G[3][3];
        float pk[3] = {p0[0], p0[1], p0[2]}, Tk[3] = {T0[0], T0[1], T0[2]};
        for (int st = 0; st < 4; ++st) {
            interp_mac(&g, uf, pk, 0, u[st], G);
            for (int r = 0; r < 3; ++r) dT[st][r] = 0.0f;
            for (int r = 0; r < dim; ++r) {
                float acc = G[r][0] * Tk[0];
                for (int c = 1; c < dim; ++c) acc = acc + G[r][c] * Tk[c];
                dT[st][r] = acc;
            }
            if (st < 3)
                for (int c = 0; c < dim; ++c) {
                    pk[c] = p0[c] + hs[st] * u[st][c];
                    Tk[c] = T0[c] + hs[st] * dT[st][c];
                }
        }
        for (int c = 0; c < dim; ++c) {
            float s = u[0][c] + 2.0f * u[1][c];
            s = s + 2.0f * u[2][c];
            s = s + u[3][c];
            psi[c * nf + f] = p0[c] + sixth * s;
            float q = dT[0][c] + 2.0f * dT[1][c];
            q = q + 2.0f * dT[2][c];
            q = q + dT[3][c];
            T[c * nf + f] = T0[c] + sixth * q;
        }
    }
}

/* ---- 13. advect_u: 3D/run_3D.py:296-312, 2D/run_2D.py:429-441.
 * u1_a[f] = T_a[f] . u0(psi_a[f]).  Called per axis. */
void oracle_advect_u(int dim, const int* res, float dx, int axis, const float* ux0,
                     const float* uy0, const float* uz0, float* u1, const float* T,
                     const float* psi) {
    grid_t g;
    make_grid(&g, dim, res, dx);
    const float* uf[3] = {ux0, uy0, uz0};
    int shp[3];
    face_shape(&g, axis, shp);
    int64_t nf = (int64_t)shp[0] * shp[1] * shp[2];
#pragma omp parallel for schedule(static)
    for (int64_t f = 0; f < nf; ++f) {
        float p[3] = {0, 0, 0}, u[3], G[3][3];
        for (int c = 0; c < dim; ++c) p[c] = psi[c * nf + f];
        interp_mac(&g, uf, p, 0, u, G);
        float acc = T[0 * nf + f] * u[0];
        for (int c = 1; c < dim; ++c) acc = acc + T[c * nf + f] * u[c];
        u1[f] = acc;
    }
}

/* ---- 7/8. init_particles_imp / init_particles_imp_grad_m:
 * 3D/init_conditions.py:254-273, 2D/init_conditions.py:183-210.
 * imp[c] = u_c(p); if C != NULL, C holds dim rows, row a = grad u_a (SoA
 * (dim*dim) x n, row-major: C[(a*dim + c)*n + p]). */
void oracle_g2p(int dim, const int* res, float dx, int64_t n, const float* pos,
                const float* ux, const float* uy, const float* uz, float* imp, float* C) {
    grid_t g;
    make_grid(&g, dim, res, dx);
    const float* uf[3] = {ux, uy, uz};
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
        float p[3] = {0, 0, 0}, u[3], G[3][3];
        for (int c = 0; c < dim; ++c) p[c] = pos[c * n + i];
        interp_mac(&g, uf, p, 1, u, G);
        for (int c = 0; c < dim; ++c) imp[c * n + i] = u[c];
        if (C)
            for (int a = 0; a < dim; ++a)
                for (int c = 0; c < dim; ++c) C[(int64_t)(a * dim + c) * n + i] = G[a][c];
    }
}

/* ---- 6. init_particles_pos_uniform: 3D/init_conditions.py:238-252,
 * 2D/init_conditions.py:139-166.  seed >= 0 adds the seeded jitter extension:
 * pos = (id + U) * dist with U from oracle_jitter_u (same hash as the GPU). */
static inline uint32_t pfm_hash32(uint32_t x) {
    /* 32-bit integer finaliser (murmur3 fmix32 constants) */
    x ^= x >> 16; x *= 0x85ebca6bu;
    x ^= x >> 13; x *= 0xc2b2ae35u;
    x ^= x >> 16;
    return x;
}

float oracle_jitter_u(uint32_t seed, uint64_t id, uint32_t comp) {
    uint32_t h = pfm_hash32((uint32_t)id ^ pfm_hash32(seed * 0x9e3779b9u + comp));
    h = pfm_hash32(h ^ (uint32_t)(id >> 32) * 0x27d4eb2fu);
    return (float)(h >> 8) * (1.0f / 16777216.0f);
}

void oracle_seed_particles(int dim, const int* res, int ppc_axis, float dist, int64_t n,
                           float* pos, int64_t seed, int order) {
    int64_t nx = (int64_t)ppc_axis * res[0];
    int64_t ny = (int64_t)ppc_axis * res[1];
    int64_t per = ppc_axis;
    for (int c = 1; c < dim; ++c) per *= ppc_axis;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
        int64_t id[3] = {0, 0, 0};
        if (order == 0) {
            /* reference numbering: lattice index x fastest */
            id[0] = i % nx;
            int64_t iyz = i / nx;
            if (dim == 3) {
                id[1] = iyz % ny;
                id[2] = iyz / ny;
            } else {
                id[1] = iyz;
            }
        } else {
            /* cell-major extension: cells with the last axis fastest, then
             * the ppc_axis^dim sub-lattice (last axis fastest) */
            int64_t cell = i / per, sub = i % per, ci[3] = {0, 0, 0};
            if (dim == 3) {
                ci[2] = cell % res[2];
                ci[1] = (cell / res[2]) % res[1];
                ci[0] = cell / ((int64_t)res[2] * res[1]);
            } else {
                ci[1] = cell % res[1];
                ci[0] = cell / res[1];
            }
            for (int c = dim - 1; c >= 0; --c) {
                id[c] = ci[c] * ppc_axis + sub % ppc_axis;
                sub /= ppc_axis;
            }
        }
        for (int c = 0; c < dim; ++c) {
            float off = seed >= 0 ? oracle_jitter_u((uint32_t)seed, (uint64_t)i, (uint32_t)c) : 0.5f;
            pos[c * n + i] = ((float)id[c] + off) * dist;
        }
    }
}

/* ---- 9. reset_to_identity_grid (3D/run_3D.py:106-120) with the face
 * coordinates of x_coords_func etc. (3D/taichi_utils.py:33-46). */
void oracle_reset_identity_grid(int dim, const int* res, float dx, int axis, float* psi, float* T) {
    grid_t g;
    make_grid(&g, dim, res, dx);
    int shp[3];
    face_shape(&g, axis, shp);
    int64_t nf = (int64_t)shp[0] * shp[1] * shp[2];
    for (int i = 0; i < shp[0]; ++i)
        for (int j = 0; j < shp[1]; ++j)
            for (int k = 0; k < shp[2]; ++k) {
                int64_t f = idx3(shp, i, j, k);
                float c[3] = {(float)i + 0.5f, (float)j + 0.5f, (float)k + 0.5f};
                c[axis] = (float)(axis == 0 ? i : axis == 1 ? j : k);
                for (int a = 0; a < dim; ++a) {
                    psi[a * nf + f] = c[a] * dx;
                    T[a * nf + f] = a == axis ? 1.0f : 0.0f;
                }
            }
}

/* ---- 10. update_T: 3D/run_3D.py:430-439 (T_init <- cols(T_gm) @ cols(T_init)),
 * 2D/run_2D.py:538-546 (T <- cols(T_gm) @ cols(T_init)).
 * Tgm and Tin hold dim column vectors, each SoA dim x n (column m at m*dim*n). */
void oracle_update_T(int dim, int64_t n, const float* Tgm, const float* Tin, float* Tout) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
        float A[3][3], B[3][3];
        for (int m = 0; m < dim; ++m)
            for (int r = 0; r < dim; ++r) {
                A[r][m] = Tgm[((int64_t)m * dim + r) * n + i];
                B[r][m] = Tin[((int64_t)m * dim + r) * n + i];
            }
        for (int m = 0; m < dim; ++m)
            for (int r = 0; r < dim; ++r) {
                float acc = A[r][0] * B[0][m];
                for (int k = 1; k < dim; ++k) acc = acc + A[r][k] * B[k][m];
                Tout[((int64_t)m * dim + r) * n + i] = acc;
            }
    }
}

/* ---- 16/17(prep). update_particles_grad_m (2D/run_2D.py:526-536) and the
 * in-P2G composition of 3D/run_3D.py:460-475:
 *   C_out = cols(T_gm) @ (rows(init_C) @ rows(T_gm))   (rows of C_out stored)
 *   imp   = cols(T) @ init_imp
 * with T = cols(T_gm) @ cols(T_init) in 3D (Tin != NULL) or T = Tfull in 2D.
 * C_in / C_out are row-major SoA (dim*dim) x n. */
void oracle_p2g_prep(int dim, int64_t n, const float* Tgm, const float* Tin, const float* C_in,
                     const float* init_imp, float* C_out, float* imp_out) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
        float A[3][3], Ci[3][3], P[3][3];
        /* A = rows(T_gm): A[r][c] = T_r[c] */
        for (int r = 0; r < dim; ++r)
            for (int c = 0; c < dim; ++c) {
                A[r][c] = Tgm[((int64_t)r * dim + c) * n + i];
                Ci[r][c] = C_in[((int64_t)r * dim + c) * n + i];
            }
        /* P = init_C @ A */
        for (int r = 0; r < dim; ++r)
            for (int c = 0; c < dim; ++c) {
                float acc = Ci[r][0] * A[0][c];
                for (int k = 1; k < dim; ++k) acc = acc + Ci[r][k] * A[k][c];
                P[r][c] = acc;
            }
        /* Q = A^T @ P */
        for (int r = 0; r < dim; ++r)
            for (int c = 0; c < dim; ++c) {
                float acc = A[0][r] * P[0][c];
                for (int k = 1; k < dim; ++k) acc = acc + A[k][r] * P[k][c];
                C_out[((int64_t)r * dim + c) * n + i] = acc;
            }
        if (imp_out) {
            float Tt[3][3];
            if (Tin) {
                /* T_transpose = cols(T_gm) @ cols(T_init) = A^T @ B */
                float B[3][3];
                for (int m = 0; m < dim; ++m)
                    for (int r = 0; r < dim; ++r) B[r][m] = Tin[((int64_t)m * dim + r) * n + i];
                for (int r = 0; r < dim; ++r)
                    for (int c = 0; c < dim; ++c) {
                        float acc = A[0][r] * B[0][c];
                        for (int k = 1; k < dim; ++k) acc = acc + A[k][r] * B[k][c];
                        Tt[r][c] = acc;
                    }
            } else {
                for (int r = 0; r < dim; ++r)
                    for (int c = 0; c < dim; ++c) Tt[r][c] = A[r][c];
            }
            for (int r = 0; r < dim; ++r) {
                float acc = Tt[r][0] * init_imp[0 * n + i];
                for (int c = 1; c < dim; ++c) acc = acc + Tt[r][c] * init_imp[(int64_t)c * n + i];
                imp_out[(int64_t)r * n + i] = acc;
            }
        }
    }
}

/* ---- 16. update_particles_imp: 2D/run_2D.py:517-524 (imp = cols(T) @ init_imp) */
void oracle_update_imp(int dim, int64_t n, const float* T, const float* init_imp, float* imp) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i)
        for (int r = 0; r < dim; ++r) {
            float acc = T[((int64_t)0 * dim + r) * n + i] * init_imp[i];
            for (int c = 1; c < dim; ++c)
                acc = acc + T[((int64_t)c * dim + r) * n + i] * init_imp[(int64_t)c * n + i];
            imp[(int64_t)r * n + i] = acc;
        }
}

/* ---- 15. get_particles_id_in_every_cell: 3D/run_3D.py:342-351.
 * cell = int(pos/dx) (truncation), clamped to the grid (extension).  Fills
 * cell[n] (flat cell index), counts[C] and the stable permutation perm[n]
 * (= numpy.argsort(cell, kind='stable')) and cell_start[C+1]. */
void oracle_bin_particles(int dim, const int* res, float dx, int64_t n, const float* pos,
                          int32_t* cell, int32_t* counts, int32_t* cell_start, int32_t* perm) {
    grid_t g;
    make_grid(&g, dim, res, dx);
    int64_t C = (int64_t)g.res[0] * g.res[1] * g.res[2];
    for (int64_t c = 0; c < C; ++c) counts[c] = 0;
    for (int64_t i = 0; i < n; ++i) {
        int ci[3] = {0, 0, 0};
        for (int a = 0; a < dim; ++a) {
            int v = (int)(pos[a * n + i] * g.inv_dx);
            if (v < 0) v = 0;
            if (v > g.res[a] - 1) v = g.res[a] - 1;
            ci[a] = v;
        }
        int64_t c = idx3(g.res, ci[0], ci[1], ci[2]);
        cell[i] = (int32_t)c;
        counts[c] += 1;
    }
    cell_start[0] = 0;
    for (int64_t c = 0; c < C; ++c) cell_start[c + 1] = cell_start[c] + counts[c];
    int32_t* fill = (int32_t*)malloc(sizeof(int32_t) * (size_t)C);
    for (int64_t c = 0; c < C; ++c) fill[c] = cell_start[c];
    for (int64_t i = 0; i < n; ++i) perm[fill[cell[i]]++] = (int32_t)i;
    free(fill);
}

/* ---- 17. P2G scatter: 3D/run_3D.py:441-552 (2D/run_2D.py:548-611 without the
 * dead psi splat).  Particles in index order, offsets -1..2 per axis with k
 * fastest, then u /= W where W > 0.  C holds rows (row a dotted with dpos for
 * axis a), imp the particle impulse.  W may be NULL. */
void oracle_p2g(int dim, const int* res, float dx, int64_t n, const float* pos, const float* imp,
                const float* C, float* ux, float* uy, float* uz, float* Wx, float* Wy, float* Wz) {
    grid_t g;
    make_grid(&g, dim, res, dx);
    float* uf[3] = {ux, uy, uz};
    float* Wf[3] = {Wx, Wy, Wz};
    float* Wtmp[3] = {NULL, NULL, NULL};
    int shp[3][3];
    int64_t nf[3] = {0, 0, 0};
    for (int a = 0; a < dim; ++a) {
        face_shape(&g, a, shp[a]);
        nf[a] = (int64_t)shp[a][0] * shp[a][1] * shp[a][2];
        if (!Wf[a]) {
            Wtmp[a] = (float*)malloc(sizeof(float) * (size_t)nf[a]);
            Wf[a] = Wtmp[a];
        }
        memset(uf[a], 0, sizeof(float) * (size_t)nf[a]);
        memset(Wf[a], 0, sizeof(float) * (size_t)nf[a]);
    }
    /* Each thread owns a contiguous range of first face indices of every axis
     * and walks all particles in index order, adding only to its own faces:
     * every face still receives its contributions in particle order, so the
     * sums equal the sequential scatter bit for bit. */
#pragma omp parallel
    {
        int nt = omp_get_num_threads(), tid = omp_get_thread_num();
        int lo[3], hi[3];
        for (int a = 0; a < dim; ++a) {
            lo[a] = (int)((int64_t)shp[a][0] * tid / nt);
            hi[a] = (int)((int64_t)shp[a][0] * (tid + 1) / nt);
        }
    for (int64_t i = 0; i < n; ++i) {
        float pos_g[3] = {0, 0, 0};
        for (int c = 0; c < dim; ++c) pos_g[c] = pos[c * n + i] * g.inv_dx;
        for (int a = 0; a < dim; ++a) {
            int base[3] = {0, 0, 0};
            for (int c = 0; c < dim; ++c) base[c] = (int)(c == a ? pos_g[c] : pos_g[c] - 0.5f);
            if (base[0] + 2 < lo[a] || base[0] - 1 >= hi[a]) continue;
            int k0 = dim == 3 ? -1 : 0, k1 = dim == 3 ? 2 : 0;
            for (int oi = -1; oi <= 2; ++oi)
                for (int oj = -1; oj <= 2; ++oj)
                    for (int ok = k0; ok <= k1; ++ok) {
                        int f[3] = {base[0] + oi, base[1] + oj, base[2] + ok};
                        if (f[0] < lo[a] || f[0] >= hi[a]) continue;
                        int inb = 1;
                        for (int c = 0; c < dim; ++c)
                            if (f[c] < 0 || f[c] > shp[a][c] - 1) inb = 0;
                        if (!inb) continue;
                        float weight = 1.0f, dpos[3] = {0, 0, 0};
                        for (int c = 0; c < dim; ++c) {
                            float d = c == a ? pos_g[c] - (float)f[c] : (pos_g[c] - (float)f[c]) - 0.5f;
                            float Nc = oracle_N2(d);
                            weight = c == 0 ? Nc : weight * Nc;
                            float e = c == a ? (float)f[c] - pos_g[c] : ((float)f[c] + 0.5f) - pos_g[c];
                            dpos[c] = e * dx;
                        }
                        float delta = C[((int64_t)a * dim + 0) * n + i] * dpos[0];
                        for (int c = 1; c < dim; ++c)
                            delta = delta + C[((int64_t)a * dim + c) * n + i] * dpos[c];
                        int64_t fi = idx3(shp[a], f[0], f[1], dim == 3 ? f[2] : 0);
                        Wf[a][fi] += weight;
                        uf[a][fi] += (imp[(int64_t)a * n + i] + delta) * weight;
                    }
        }
    }
    }
    for (int a = 0; a < dim; ++a) {
#pragma omp parallel for schedule(static)
        for (int64_t f = 0; f < nf[a]; ++f)
            if (Wf[a][f] > 0.0f) uf[a][f] /= Wf[a][f];
        free(Wtmp[a]);
    }
}

/* ---- 11. calc_max_speed: 3D/run_3D.py:95-103, 2D/run_2D.py:101-108 */
float oracle_max_speed(int dim, const int* res, const float* ux, const float* uy, const float* uz) {
    grid_t g;
    make_grid(&g, dim, res, 1.0f);
    int sx[3], sy[3], sz[3];
    face_shape(&g, 0, sx);
    face_shape(&g, 1, sy);
    face_shape(&g, 2, sz);
    float m = 1.e-3f;
    for (int i = 0; i < g.res[0]; ++i)
        for (int j = 0; j < g.res[1]; ++j)
            for (int k = 0; k < g.res[2]; ++k) {
                float u = 0.5f * (ux[idx3(sx, i, j, k)] + ux[idx3(sx, i + 1, j, k)]);
                float v = 0.5f * (uy[idx3(sy, i, j, k)] + uy[idx3(sy, i, j + 1, k)]);
                float s2 = u * u + v * v;
                if (dim == 3) {
                    float w = 0.5f * (uz[idx3(sz, i, j, k)] + uz[idx3(sz, i, j, k + 1)]);
                    s2 = s2 + w * w;
                }
                float sp = sqrtf(s2);
                if (sp > m) m = sp;
            }
    return m;
}

/* ---- 31. split_central_vector: 3D/taichi_utils.py:87-100 (2D :56-65).
 * vc is SoA dim x C. */
void oracle_split_central(int dim, const int* res, const float* vc, float* ux, float* uy,
                          float* uz) {
    grid_t g;
    make_grid(&g, dim, res, 1.0f);
    float* uf[3] = {ux, uy, uz};
    int64_t C = (int64_t)g.res[0] * g.res[1] * g.res[2];
    for (int a = 0; a < dim; ++a) {
        int shp[3];
        face_shape(&g, a, shp);
        for (int i = 0; i < shp[0]; ++i)
            for (int j = 0; j < shp[1]; ++j)
                for (int k = 0; k < shp[2]; ++k) {
                    int I[3] = {i, j, k}, L[3] = {i, j, k};
                    L[a] -= 1;
                    float r = sample3(vc + a * C, g.res, I[0], I[1], I[2]);
                    float l = sample3(vc + a * C, g.res, L[0], L[1], L[2]);
                    uf[a][idx3(shp, i, j, k)] = 0.5f * (r + l);
                }
    }
}

/* ---- get_central_vector: 3D/taichi_utils.py:80-85 (2D :50-54) */
void oracle_get_central(int dim, const int* res, const float* ux, const float* uy,
                        const float* uz, float* vc) {
    grid_t g;
    make_grid(&g, dim, res, 1.0f);
    const float* uf[3] = {ux, uy, uz};
    int64_t C = (int64_t)g.res[0] * g.res[1] * g.res[2];
    for (int a = 0; a < dim; ++a) {
        int shp[3];
        face_shape(&g, a, shp);
        for (int i = 0; i < g.res[0]; ++i)
            for (int j = 0; j < g.res[1]; ++j)
                for (int k = 0; k < g.res[2]; ++k) {
                    int I[3] = {i, j, k};
                    I[a] += 1;
                    vc[a * C + idx3(g.res, i, j, k)] =
                        0.5f * (uf[a][idx3(shp, I[0], I[1], I[2])] + uf[a][idx3(shp, i, j, k)]);
                }
    }
}

/* ---- curl: 3D/taichi_utils.py:56-78 (sample() clamps the neighbour indices,
 * 3D/taichi_utils.py:48-54), 2D/taichi_utils.py:97-103.  vc is SoA (d x C);
 * out is 3 x C (3D) or C (2D). */
static float vc_at(const grid_t* g, const float* vc, int c, int i, int j, int k) {
    int64_t C = (int64_t)g->res[0] * g->res[1] * g->res[2];
    i = i < 0 ? 0 : (i > g->res[0] - 1 ? g->res[0] - 1 : i);
    j = j < 0 ? 0 : (j > g->res[1] - 1 ? g->res[1] - 1 : j);
    k = k < 0 ? 0 : (k > g->res[2] - 1 ? g->res[2] - 1 : k);
    return vc[c * C + idx3(g->res, i, j, k)];
}

void oracle_curl(int dim, const int* res, float dx, const float* vc, float* out) {
    grid_t g;
    make_grid(&g, dim, res, dx);
    int64_t C = (int64_t)g.res[0] * g.res[1] * g.res[2];
    const float inv_dist = 1.0f / (2.0f * dx); /* 1./(2*dx) in the kernel's f32 */
    for (int i = 0; i < g.res[0]; ++i)
        for (int j = 0; j < g.res[1]; ++j)
            for (int k = 0; k < g.res[2]; ++k) {
                int64_t q = idx3(g.res, i, j, k);
                if (dim == 3) {
                    float d_vx_dz = inv_dist * (vc_at(&g, vc, 0, i, j, k + 1) - vc_at(&g, vc, 0, i, j, k - 1));
                    float d_vx_dy = inv_dist * (vc_at(&g, vc, 0, i, j + 1, k) - vc_at(&g, vc, 0, i, j - 1, k));
                    float d_vy_dx = inv_dist * (vc_at(&g, vc, 1, i + 1, j, k) - vc_at(&g, vc, 1, i - 1, j, k));
                    float d_vy_dz = inv_dist * (vc_at(&g, vc, 1, i, j, k + 1) - vc_at(&g, vc, 1, i, j, k - 1));
                    float d_vz_dx = inv_dist * (vc_at(&g, vc, 2, i + 1, j, k) - vc_at(&g, vc, 2, i - 1, j, k));
                    float d_vz_dy = inv_dist * (vc_at(&g, vc, 2, i, j + 1, k) - vc_at(&g, vc, 2, i, j - 1, k));
                    out[q] = d_vz_dy - d_vy_dz;
                    out[C + q] = d_vx_dz - d_vz_dx;
                    out[2 * C + q] = d_vy_dx - d_vx_dy;
                } else {
                    float vr_y = vc_at(&g, vc, 1, i + 1, j, 0), vl_y = vc_at(&g, vc, 1, i - 1, j, 0);
                    float vt_x = vc_at(&g, vc, 0, i, j + 1, 0), vb_x = vc_at(&g, vc, 0, i, j - 1, 0);
                    out[q] = (vr_y - vl_y - vt_x + vb_x) / (2.0f * dx);
                }
            }
}

/* ---- 31. point vortices: angular_vel_func + leapfrog_vel_func,
 * 2D/init_conditions.py:5-9,29-60, generalised to a list of vortices.
 * Writes the cell-centred velocity vc (SoA 2 x C) at centres (i+0.5)dx. */
void oracle_point_vortices(const int* res, float dx, int nv, const float* cx, const float* cy,
                           const float* w, float rad, float strength, float* vc) {
    int64_t C = (int64_t)res[0] * res[1];
    float rad2 = rad * rad;
    for (int i = 0; i < res[0]; ++i)
        for (int j = 0; j < res[1]; ++j) {
            float X[2] = {((float)i + 0.5f) * dx, ((float)j + 0.5f) * dx};
            float vx = 0.0f, vy = 0.0f;
            for (int q = 0; q < nv; ++q) {
                float px = X[0] - cx[q], py = X[1] - cy[q];
                float r = sqrtf(px * px + py * py);
                r = r + 1e-6f;
                float lin = strength * 1.0f / r * (1.0f - expf(-(r * r) / rad2));
                float ang = 1.0f / r * lin;
                float s = ang * w[q];
                vx = q == 0 ? s * -py : vx + s * -py;
                vy = q == 0 ? s * px : vy + s * px;
            }
            vc[0 * C + (int64_t)i * res[1] + j] = vx;
            vc[1 * C + (int64_t)i * res[1] + j] = vy;
        }
}

/* ---- 31. add_vortex_ring: 3D/init_conditions.py:10-22.  Adds into vc
 * (SoA 3 x C) at cell centres (I+0.5)dx. */
void oracle_vortex_ring(const int* res, float dx, float w, float rad, float delta,
                        const float* c, const float* ex, const float* ey, int num_samples,
                        float* vc) {
    int64_t C = (int64_t)res[0] * res[1] * res[2];
    float curve_length = (float)((2.0 * M_PI * (double)rad) / num_samples);
    float four_pi = (float)(4.0 * M_PI);
    float pi_f = (float)M_PI;
#pragma omp parallel for schedule(static)
    for (int64_t q = 0; q < C; ++q) {
        int k = (int)(q % res[2]);
        int j = (int)((q / res[2]) % res[1]);
        int i = (int)(q / ((int64_t)res[1] * res[2]));
        float P[3] = {((float)i + 0.5f) * dx, ((float)j + 0.5f) * dx, ((float)k + 0.5f) * dx};
        float acc[3] = {vc[q], vc[C + q], vc[2 * C + q]};
        for (int l = 0; l < num_samples; ++l) {
            float theta = (float)l / (float)num_samples * 2.0f * pi_f;
            float ct = cosf(theta), st = sinf(theta);
            float ps[3], pd[3], wv[3];
            for (int a = 0; a < 3; ++a) {
                ps[a] = rad * (ct * ex[a] + st * ey[a]) + c[a];
                pd[a] = P[a] - ps[a];
                wv[a] = w * (-st * ex[a] + ct * ey[a]);
            }
            float r = sqrtf(pd[0] * pd[0] + pd[1] * pd[1] + pd[2] * pd[2]);
            float rd = r / delta;
            float fac = curve_length * (-1.0f / (four_pi * (r * r * r)) * (1.0f - expf(-(rd * rd * rd))));
            float cr[3] = {pd[1] * wv[2] - pd[2] * wv[1], pd[2] * wv[0] - pd[0] * wv[2],
                           pd[0] * wv[1] - pd[1] * wv[0]};
            for (int a = 0; a < 3; ++a) acc[a] += fac * cr[a];
        }
        vc[q] = acc[0];
        vc[C + q] = acc[1];
        vc[2 * C + q] = acc[2];
    }
}

/* ---- (f)4 smoke transport on the grid.
 * advect_smoke: 3D/run_3D.py:315-323 (psi_c = 1./6 * (sum of the 6 face maps,
 * left to right)), 2D/run_2D.py:444-450 (0.25 * sum of 4); interp_1:
 * 3D/taichi_utils.py:193-220 (s = max(1, min(u, dim-1-eps)), 8 samples in
 * (i, j, k) order, value * N_1 * N_1 * N_1 accumulated from 0 * sample) and
 * 2D/taichi_utils.py:119-137 (s = max(0, ...), a N N + b N N + c N N + d N N);
 * clamp_smoke: 3D/run_3D.py:325-334 with sample_min_max_1
 * (3D/taichi_utils.py:222-247).  p / dx is evaluated as p * (1 / dx), as
 * everywhere in this restatement.  mode bit 0 advects, bit 1 clamps. */
static inline float hat1(float x) { return 1.0f - fabsf(x); }

void oracle_advect_smoke(int dim, const int* res, float dx, int nch, const float* s0, float* s1,
                         const float* px, const float* py, const float* pz, int mode) {
    grid_t g;
    make_grid(&g, dim, res, dx);
    const int rx = g.res[0], ry = g.res[1], rz = g.res[2];
    const int64_t C = (int64_t)rx * ry * rz;
    const int shp[3] = {rx, ry, rz};
    float hi[3];
    for (int a = 0; a < 3; ++a) hi[a] = (float)((double)shp[a] - 1.0 - ORACLE_EPS);
#pragma omp parallel for schedule(static)
    for (int64_t q = 0; q < C; ++q) {
        float pc[3] = {0.f, 0.f, 0.f};
        if (dim == 3) {
            int k = (int)(q % rz), j = (int)((q / rz) % ry), i = (int)(q / ((int64_t)rz * ry));
            int64_t Fx = (int64_t)(rx + 1) * ry * rz, Fy = (int64_t)rx * (ry + 1) * rz, Fz = (int64_t)rx * ry * (rz + 1);
            for (int c = 0; c < 3; ++c) {
                float sum = px[c * Fx + ((int64_t)i * ry + j) * rz + k];
                sum = sum + px[c * Fx + ((int64_t)(i + 1) * ry + j) * rz + k];
                sum = sum + py[c * Fy + ((int64_t)i * (ry + 1) + j) * rz + k];
                sum = sum + py[c * Fy + ((int64_t)i * (ry + 1) + j + 1) * rz + k];
                sum = sum + pz[c * Fz + ((int64_t)i * ry + j) * (rz + 1) + k];
                sum = sum + pz[c * Fz + ((int64_t)i * ry + j) * (rz + 1) + k + 1];
                pc[c] = (float)(1.0 / 6.0) * sum;
            }
        } else {
            int j = (int)(q % ry), i = (int)(q / ry);
            int64_t Fx = (int64_t)(rx + 1) * ry, Fy = (int64_t)rx * (ry + 1);
            for (int c = 0; c < 2; ++c) {
                float sum = px[c * Fx + (int64_t)i * ry + j];
                sum = sum + px[c * Fx + (int64_t)(i + 1) * ry + j];
                sum = sum + py[c * Fy + (int64_t)i * (ry + 1) + j];
                sum = sum + py[c * Fy + (int64_t)i * (ry + 1) + j + 1];
                pc[c] = 0.25f * sum;
            }
        }
        float sv[3] = {0.f, 0.f, 0.f}, fl[3] = {0.f, 0.f, 0.f};
        for (int c = 0; c < dim; ++c) {
            float u = pc[c] * g.inv_dx - 0.5f;
            float lo = dim == 3 ? 1.0f : 0.0f;
            float m = u < hi[c] ? u : hi[c];
            sv[c] = m > lo ? m : lo;
            fl[c] = floorf(sv[c]);
        }
        for (int ch = 0; ch < nch; ++ch) {
            const float* f = s0 + (int64_t)ch * C;
            if (dim == 3) {
                const int i0 = (int)fl[0], j0 = (int)fl[1], k0 = (int)fl[2];
                float base = sample3(f, shp, i0, j0, k0);
                if (mode & 1) {
                    float acc = 0.0f * base;
                    for (int i = 0; i < 2; ++i)
                        for (int j = 0; j < 2; ++j)
                            for (int k = 0; k < 2; ++k) {
                                float xi = sv[0] - (fl[0] + (float)i);
                                float yj = sv[1] - (fl[1] + (float)j);
                                float zk = sv[2] - (fl[2] + (float)k);
                                float v = sample3(f, shp, (int)(fl[0] + (float)i), (int)(fl[1] + (float)j),
                                                  (int)(fl[2] + (float)k));
                                acc = acc + v * hat1(xi) * hat1(yj) * hat1(zk);
                            }
                    s1[(int64_t)ch * C + q] = acc;
                }
                if (mode & 2) {
                    float mn = base, mx = base;
                    for (int i = 0; i < 2; ++i)
                        for (int j = 0; j < 2; ++j)
                            for (int k = 0; k < 2; ++k) {
                                float v = sample3(f, shp, (int)(fl[0] + (float)i), (int)(fl[1] + (float)j),
                                                  (int)(fl[2] + (float)k));
                                mn = v < mn ? v : mn;
                                mx = v > mx ? v : mx;
                            }
                    float cur = s1[(int64_t)ch * C + q];
                    cur = cur > mn ? cur : mn;
                    s1[(int64_t)ch * C + q] = cur < mx ? cur : mx;
                }
            } else {
                int shp2[3] = {rx, ry, 1};
                float fu = sv[0] - fl[0], fv = sv[1] - fl[1];
                float a = sample3(f, shp2, (int)fl[0], (int)fl[1], 0);
                float b = sample3(f, shp2, (int)(fl[0] + 1.0f), (int)fl[1], 0);
                float c = sample3(f, shp2, (int)fl[0], (int)(fl[1] + 1.0f), 0);
                float d = sample3(f, shp2, (int)(fl[0] + 1.0f), (int)(fl[1] + 1.0f), 0);
                s1[(int64_t)ch * C + q] = a * hat1(fu) * hat1(fv) + b * hat1(fu - 1.0f) * hat1(fv) +
                                          c * hat1(fu) * hat1(fv - 1.0f) + d * hat1(fu - 1.0f) * hat1(fv - 1.0f);
            }
        }
    }
}

/* ---- the smoke part of add_vortex_ring_and_smoke (3D/init_conditions.py:31-56):
 * channel 3 += curve_length * exp(-(r/delta)^3) per ring sample, then every
 * cell becomes (color, 1) if channel 3 > 0.002, else 0.  smoke is 4 x C. */
void oracle_vortex_ring_smoke(const int* res, float dx, float rad, float delta, const float* c,
                              const float* ex, const float* ey, int num_samples, const float* color,
                              float* smoke) {
    int64_t C = (int64_t)res[0] * res[1] * res[2];
    float curve_length = (float)((2.0 * M_PI * (double)rad) / num_samples);
    float pi_f = (float)M_PI;
#pragma omp parallel for schedule(static)
    for (int64_t q = 0; q < C; ++q) {
        int k = (int)(q % res[2]);
        int j = (int)((q / res[2]) % res[1]);
        int i = (int)(q / ((int64_t)res[1] * res[2]));
        float P[3] = {((float)i + 0.5f) * dx, ((float)j + 0.5f) * dx, ((float)k + 0.5f) * dx};
        float acc = smoke[3 * C + q];
        for (int l = 0; l < num_samples; ++l) {
            float theta = (float)l / (float)num_samples * 2.0f * pi_f;
            float ct = cosf(theta), st = sinf(theta);
            float pd[3];
            for (int a = 0; a < 3; ++a) pd[a] = P[a] - (rad * (ct * ex[a] + st * ey[a]) + c[a]);
            float r = sqrtf(pd[0] * pd[0] + pd[1] * pd[1] + pd[2] * pd[2]);
            float rd = r / delta;
            acc += curve_length * expf(-(rd * rd * rd));
        }
        smoke[3 * C + q] = acc;
    }
    for (int64_t q = 0; q < C; ++q) {
        int on = smoke[3 * C + q] > 0.002f;
        smoke[q] = on ? color[0] : 0.f;
        smoke[C + q] = on ? color[1] : 0.f;
        smoke[2 * C + q] = on ? color[2] : 0.f;
        smoke[3 * C + q] = on ? 1.0f : 0.f;
    }
}

/* ---- (f)4 particle smoke (3D, 4 channels; cell fields (4, C), particle
 * smoke (4, n), gradient (12, n) row-major 3x4).
 * init_particles_smoke: 3D/init_conditions.py:275-279 with
 * interp_grad_2_smoke (3D/taichi_utils.py:378-421): s = max(1, min(u,
 * dim-2-eps)), 64 samples in (i, j, k) order, dw = 1/dx dN N N,
 * interped += value N N N, new_C += outer(dw, value). */
void oracle_g2p_smoke(const int* res, float dx, int64_t n, const float* pos, const float* smoke,
                      float* ps, float* pg) {
    grid_t g;
    make_grid(&g, 3, res, dx);
    const int shp[3] = {g.res[0], g.res[1], g.res[2]};
    const int64_t C = (int64_t)shp[0] * shp[1] * shp[2];
#pragma omp parallel for schedule(static)
    for (int64_t p = 0; p < n; ++p) {
        float sv[3], fl[3];
        for (int a = 0; a < 3; ++a) {
            float u = pos[a * n + p] * g.inv_dx - 0.5f;
            float hi = (float)((double)shp[a] - 2.0 - ORACLE_EPS);
            float m = u < hi ? u : hi;
            sv[a] = m > 1.0f ? m : 1.0f;
            fl[a] = floorf(sv[a]);
        }
        float val[4] = {0.f, 0.f, 0.f, 0.f}, nc[3][4];
        memset(nc, 0, sizeof(nc));
        for (int i = -1; i < 3; ++i)
            for (int j = -1; j < 3; ++j)
                for (int k = -1; k < 3; ++k) {
                    float x = sv[0] - (fl[0] + (float)i), y = sv[1] - (fl[1] + (float)j), z = sv[2] - (fl[2] + (float)k);
                    int ci = (int)(fl[0] + (float)i), cj = (int)(fl[1] + (float)j), ck = (int)(fl[2] + (float)k);
                    float dw[3] = {g.inv_dx * oracle_dN2(x) * oracle_N2(y) * oracle_N2(z),
                                   g.inv_dx * oracle_N2(x) * oracle_dN2(y) * oracle_N2(z),
                                   g.inv_dx * oracle_N2(x) * oracle_N2(y) * oracle_dN2(z)};
                    for (int c = 0; c < 4; ++c) {
                        float v = sample3(smoke + c * C, shp, ci, cj, ck);
                        val[c] = val[c] + v * oracle_N2(x) * oracle_N2(y) * oracle_N2(z);
                        for (int r = 0; r < 3; ++r) nc[r][c] = nc[r][c] + dw[r] * v;
                    }
                }
        for (int c = 0; c < 4; ++c) ps[c * n + p] = val[c];
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 4; ++c) pg[(4 * r + c) * n + p] = nc[r][c];
    }
}

/* P2G_smoke: 3D/run_3D.py:554-593, particles in id order, offsets -1..2 in
 * (i, j, k) order around int(pos - 0.5), weights N_2 N_2 N_2, APIC term
 * dpos @ (cols(T_grad_m) @ cols(T_init) @ grad); then smoke /= weight where
 * positive and each channel clamped to [0, 1]. */
void oracle_p2g_smoke(const int* res, float dx, int64_t n, const float* pos, const float* ps,
                      const float* grad, const float* tgm, const float* tin, int apic, float* smoke,
                      float* weight) {
    grid_t g;
    make_grid(&g, 3, res, dx);
    const int shp[3] = {g.res[0], g.res[1], g.res[2]};
    const int64_t C = (int64_t)shp[0] * shp[1] * shp[2];
    const float* Tg[3] = {tgm, tgm + 3 * n, tgm + 6 * n};
    const float* Ti[3] = {tin, tin + 3 * n, tin + 6 * n};
    for (int64_t q = 0; q < 4 * C; ++q) smoke[q] = 0.f;
    for (int64_t q = 0; q < C; ++q) weight[q] = 0.f;
    for (int64_t p = 0; p < n; ++p) {
        float P[3];
        int base[3];
        for (int r = 0; r < 3; ++r) {
            P[r] = pos[r * n + p] * g.inv_dx;
            base[r] = (int)(P[r] - 0.5f);
        }
        float MG[3][4];
        if (apic) {
            float M[3][3];
            for (int r = 0; r < 3; ++r)
                for (int c = 0; c < 3; ++c) {
                    float m = Tg[0][r * n + p] * Ti[c][0 * n + p];
                    m = m + Tg[1][r * n + p] * Ti[c][1 * n + p];
                    m = m + Tg[2][r * n + p] * Ti[c][2 * n + p];
                    M[r][c] = m;
                }
            for (int r = 0; r < 3; ++r)
                for (int c = 0; c < 4; ++c) {
                    float m = M[r][0] * grad[(0 + c) * n + p];
                    m = m + M[r][1] * grad[(4 + c) * n + p];
                    m = m + M[r][2] * grad[(8 + c) * n + p];
                    MG[r][c] = m;
                }
        }
        for (int oi = -1; oi < 3; ++oi)
            for (int oj = -1; oj < 3; ++oj)
                for (int ok = -1; ok < 3; ++ok) {
                    int F[3] = {base[0] + oi, base[1] + oj, base[2] + ok};
                    if (F[0] < 0 || F[0] >= shp[0] || F[1] < 0 || F[1] >= shp[1] || F[2] < 0 || F[2] >= shp[2])
                        continue;
                    float w = oracle_N2(P[0] - (float)F[0] - 0.5f) * oracle_N2(P[1] - (float)F[1] - 0.5f) *
                              oracle_N2(P[2] - (float)F[2] - 0.5f);
                    int64_t q = ((int64_t)F[0] * shp[1] + F[1]) * shp[2] + F[2];
                    weight[q] += w;
                    for (int c = 0; c < 4; ++c) {
                        float delta = 0.f;
                        if (apic) {
                            float dpos[3];
                            for (int r = 0; r < 3; ++r) dpos[r] = ((float)F[r] + 0.5f - P[r]) * dx;
                            delta = dpos[0] * MG[0][c];
                            delta = delta + dpos[1] * MG[1][c];
                            delta = delta + dpos[2] * MG[2][c];
                        }
                        smoke[c * C + q] += (ps[c * n + p] + delta) * w;
                    }
                }
    }
    for (int64_t q = 0; q < C; ++q)
        if (weight[q] > 0.f)
            for (int c = 0; c < 4; ++c) {
                float v = smoke[c * C + q] / weight[q];
                if (v < 0.f) v = 0.0f;
                if (v > 1.f) v = 1.0f;
                smoke[c * C + q] = v;
            }
}

/* clamp_smoke_particle: 3D/run_3D.py:595-614 over cell lists (ids ascending,
 * at most cell_max per cell) given as start / ids arrays (CSR). */
void oracle_clamp_smoke_particle(const int* res, float dx, int64_t n, const int32_t* start,
                                 const int32_t* ids, int cell_max, const float* init_pos,
                                 const float* init_smoke, float* smoke) {
    grid_t g;
    make_grid(&g, 3, res, dx);
    const int shp[3] = {g.res[0], g.res[1], g.res[2]};
    const int64_t C = (int64_t)shp[0] * shp[1] * shp[2];
    for (int64_t q = 0; q < C; ++q) {
        float mn[4] = {1000.f, 1000.f, 1000.f, 1000.f}, mx[4] = {-1000.f, -1000.f, -1000.f, -1000.f};
        int valid = 0;
        int m = start[q + 1] - start[q];
        if (m > cell_max) m = cell_max;
        for (int s = 0; s < m; ++s) {
            int64_t p = ids[start[q] + s];
            int b[3];
            for (int r = 0; r < 3; ++r) b[r] = (int)(init_pos[r * n + p] * g.inv_dx);
            for (int oa = -1; oa < 2; ++oa)
                for (int ob = -1; ob < 2; ++ob)
                    for (int oc = -1; oc < 2; ++oc) {
                        int ca = b[0] + oa, cb = b[1] + ob, cc = b[2] + oc;
                        if (ca < 0 || ca >= shp[0] || cb < 0 || cb >= shp[1] || cc < 0 || cc >= shp[2]) continue;
                        int64_t nb = ((int64_t)ca * shp[1] + cb) * shp[2] + cc;
                        for (int c = 0; c < 4; ++c) {
                            float v = init_smoke[c * C + nb];
                            mn[c] = v < mn[c] ? v : mn[c];
                            mx[c] = v > mx[c] ? v : mx[c];
                        }
                        valid += 1;
                    }
        }
        if (valid >= 1)
            for (int c = 0; c < 4; ++c) {
                float v = smoke[c * C + q];
                v = v > mn[c] ? v : mn[c];
                smoke[c * C + q] = v < mx[c] ? v : mx[c];
            }
    }
}

/* ---- 32. kinetic energy (2D/run_2D.py:765-770): sum u_a^2 / |F_a|, in double */
double oracle_kinetic_energy(int dim, const int* res, const float* ux, const float* uy,
                             const float* uz) {
    grid_t g;
    make_grid(&g, dim, res, 1.0f);
    const float* uf[3] = {ux, uy, uz};
    double E = 0.0;
    for (int a = 0; a < dim; ++a) {
        int shp[3];
        face_shape(&g, a, shp);
        int64_t nf = (int64_t)shp[0] * shp[1] * shp[2];
        double s = 0.0;
        for (int64_t f = 0; f < nf; ++f) s += (double)uf[a][f] * (double)uf[a][f];
        E += s / (double)nf;
    }
    return E;
}

/* =====================================================================
 * MGPCG: 3D/mgpcg.py:8-280 (base), 350-430 (MGPCG_3), 2D/mgpcg.py (MGPCG_2).
 * Precision is selected per build of this block: real = float (the
 * reference) via the oracle_mg_* functions, and double via oracle_mg64_*
 * (the FP64 Poisson oracle).
 * ===================================================================== */
#define MG_REAL float
#define MG_PREFIX(name) oracle_mg_##name
#include "pfm_oracle_mg.inc"
#undef MG_REAL
#undef MG_PREFIX
#define MG_REAL double
#define MG_PREFIX(name) oracle_mg64_##name
#include "pfm_oracle_mg.inc"
#undef MG_REAL
#undef MG_PREFIX
