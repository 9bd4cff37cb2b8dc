/*
 * pfm.h -- C ABI of libpfm.so, the MI355X (gfx950) particle-flow-map step.
 *
 * Every entry point replaces one Taichi kernel (or solver method) of the
 * reference zjw49246/particle-flow-maps; the reference file:line is given
 * beside each declaration.  All entry points are asynchronous on the given
 * HIP stream (pass NULL for the default stream) and return PFM_OK (0) or a
 * negative error code; pfm_last_error() returns a thread-local description.
 *
 * Memory: every pointer argument is a device pointer owned by the caller
 * (e.g. a torch tensor's data_ptr()).  Layouts:
 *   - grid fields are row-major with the last axis fastest (the numpy view of
 *     the reference's fields): a 3D x-face field has shape (rx+1, ry, rz) and
 *     (i,j,k) lives at (i*ry + j)*rz + k; 2D fields are (nx, ny), j fastest;
 *   - particle vector fields are SoA (d, n): component c of particle p at
 *     c*n + p; one column/row vector per argument (T_x, T_y, T_z / C_x ...);
 *   - grid-attached vector fields (psi_x_grid, T_x_grid ...) are SoA
 *     (d, faces of that axis).
 * The library allocates device memory only in pfm_mg_create (the multigrid
 * hierarchy); step entry points take caller workspaces and never allocate.
 */
#ifndef PFM_H
#define PFM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PFM_ABI_VERSION 1

#define PFM_OK 0
#define PFM_ERR_ARG (-1)      /* bad shape, null pointer, unsupported option */
#define PFM_ERR_HIP (-2)      /* a HIP runtime call failed */
#define PFM_ERR_OOM (-3)      /* allocation failed (pfm_mg_create, pfm_mg_set_solid, pfm_scene_create) */
#define PFM_WARN_NOT_CONVERGED 1 /* MGPCG hit max_iters (the reference prints and returns) */

typedef void* pfm_stream_t; /* hipStream_t */

/* Grid descriptor: the reference's module globals res_x/res_y/res_z and
 * dx = 1/res_y (3D/run_3D.py:16, 2D/run_2D.py:14).  res[2] is ignored in 2D. */
typedef struct {
    int32_t dim;     /* 2 or 3 */
    int32_t res[3];  /* cells per axis, each >= 4 */
    float dx;        /* cell size */
} pfm_grid;

int pfm_abi_version(void);
const char* pfm_last_error(void);

/* ---- flow-map advection ------------------------------------------------ */

/* RK4_T_forward: 3D/run_3D.py:175-219 (via stretch_T_and_advect_particles
 * :336-337); 2D/run_2D.py:304-341 (advect_psi flag).  pos (d,n) and the d
 * column vectors T_x..T_z (each (d,n)) are updated in place; u_* are the face
 * velocities u_x0..u_z0.  T_z/u_z are ignored in 2D. */
int pfm_rk4_particles(const pfm_grid* g, int64_t n, float* pos, float* T_x, float* T_y,
                      float* T_z, const float* u_x, const float* u_y, const float* u_z,
                      float dt, int advect_psi, pfm_stream_t stream);

/* RK4_grid: 3D/run_3D.py:144-173, 2D/run_2D.py:233-261.  Backward RK4 of the
 * faces of `axis`: psi (d, F_axis) and T (d, F_axis) in place. */
int pfm_rk4_grid(const pfm_grid* g, int axis, float* psi, float* T, const float* u_x,
                 const float* u_y, const float* u_z, float dt, pfm_stream_t stream);

/* advect_u: 3D/run_3D.py:296-312, 2D/run_2D.py:429-441.  For each axis a:
 * u1_a[f] = T_a[f] . u0(psi_a[f]).  u1_* must not alias u0_*. */
int pfm_advect_u(const pfm_grid* g, const float* u0_x, const float* u0_y, const float* u0_z,
                 float* u1_x, float* u1_y, float* u1_z, const float* T_x, const float* T_y,
                 const float* T_z, const float* psi_x, const float* psi_y, const float* psi_z,
                 pfm_stream_t stream);

/* The grid part of the midpoint velocity, 3D/run_3D.py:778-784 (2D:
 * 2D/run_2D.py:865-870) in one launch: for every axis a,
 * reset_to_identity_grid(psi_a, T_a), RK4_grid(psi_a, T_a, u0, dt) and
 * advect_u(u0 -> u1_a), per face in registers.  u0 is the copy of the
 * velocity that the reference's RK4_grid samples as u and advect_u as u0,
 * so u1 (not aliasing u0) can be the live velocity; psi and T end as the
 * three separate calls leave them. */
int pfm_midpoint_advect(const pfm_grid* g, const float* u0_x, const float* u0_y, const float* u0_z,
                        float* u1_x, float* u1_y, float* u1_z, float* psi_x, float* psi_y, float* psi_z,
                        float* T_x, float* T_y, float* T_z, float dt, pfm_stream_t stream);

/* reset_to_identity_grid: 3D/run_3D.py:106-120 with the face coordinates of
 * x/y/z_coords_func (3D/taichi_utils.py:33-46); 2D/run_2D.py:120-129. */
int pfm_reset_identity_grid(const pfm_grid* g, int axis, float* psi, float* T,
                            pfm_stream_t stream);

/* reset_T_to_identity: 3D/run_3D.py:133-140, 2D/run_2D.py:141-146. */
int pfm_reset_T_identity(int dim, int64_t n, float* T_x, float* T_y, float* T_z,
                         pfm_stream_t stream);

/* ---- particle <-> grid transfer ---------------------------------------- */

/* init_particles_pos_uniform: 3D/init_conditions.py:238-252,
 * 2D/init_conditions.py:139-166 (lattice, reference particle order).
 * seed >= 0 adds seeded per-particle jitter inside each lattice sub-cell
 * (extension; seed < 0 gives the reference lattice bit-exactly).  order 0 is
 * the reference order (x fastest); order 1 numbers particles cell by cell
 * (cells z fastest, ppc_axis^d per cell) for coalesced transfers. */
int pfm_seed_particles(const pfm_grid* g, int ppc_axis, float dist, int64_t n, float* pos,
                       int64_t seed, int order, pfm_stream_t stream);

/* init_particles_imp (3D/init_conditions.py:254-263, 2D :183-191) when C_x is
 * NULL; init_particles_imp_grad_m (3D :265-273, 2D :203-210) otherwise: imp
 * (d,n) = u(pos) and rows C_a (d,n) = grad u_a(pos).  imp may be NULL when C_x
 * is given (the 3D driver never reads particles_init_imp_grad_m). */
int pfm_g2p(const pfm_grid* g, int64_t n, const float* pos, const float* u_x, const float* u_y,
            const float* u_z, float* imp, float* C_x, float* C_y, float* C_z,
            pfm_stream_t stream);

/* update_T: 3D/run_3D.py:430-439 (out = in_init, in place) and
 * 2D/run_2D.py:538-546 (out = T): out = cols(T_gm) @ cols(T_in). */
int pfm_update_T(int dim, int64_t n, const float* Tgm_x, const float* Tgm_y, const float* Tgm_z,
                 const float* Tin_x, const float* Tin_y, const float* Tin_z, float* Tout_x,
                 float* Tout_y, float* Tout_z, pfm_stream_t stream);

/* update_particles_imp: 2D/run_2D.py:517-524: imp = cols(T) @ init_imp. */
int pfm_update_particles_imp(int dim, int64_t n, const float* T_x, const float* T_y,
                             const float* T_z, const float* init_imp, float* imp,
                             pfm_stream_t stream);

/* update_particles_grad_m: 2D/run_2D.py:526-536: rows of
 * cols(T_gm) @ (rows(init_C) @ rows(T_gm)) into C_x..C_z. */
int pfm_update_particles_grad_m(int dim, int64_t n, float* C_x, float* C_y, float* C_z,
                                const float* iC_x, const float* iC_y, const float* iC_z,
                                const float* Tgm_x, const float* Tgm_y, const float* Tgm_z,
                                pfm_stream_t stream);

/* Workspace bytes needed by pfm_bin_particles and pfm_p2g for n particles. */
size_t pfm_p2g_workspace_size(const pfm_grid* g, int64_t n);

/* get_particles_id_in_every_cell: 3D/run_3D.py:342-351, as a deterministic
 * stable counting sort.  cell_of (n, int32, nullable) = flat cell
 * (i*ry + j)*rz + k of int(pos/dx) clamped to the grid; cell_start (C+1)
 * exclusive offsets; perm (n) = particle ids sorted by cell, ascending id
 * inside a cell (= numpy.argsort(cell_of, kind='stable')). */
int pfm_bin_particles(const pfm_grid* g, int64_t n, const float* pos, int32_t* cell_of,
                      int32_t* cell_start, int32_t* perm, void* workspace, size_t ws_bytes,
                      pfm_stream_t stream);

/* The reference's output layout of get_particles_id_in_every_cell: ids
 * (C x cell_max, -1 filled, lowest ids first) and counts (C, full count). */
int pfm_cell_lists(const pfm_grid* g, const int32_t* cell_start, const int32_t* perm,
                   int32_t cell_max, int32_t* ids, int32_t* counts, pfm_stream_t stream);

/* P2G: 3D/run_3D.py:441-552 (Tgm_x != NULL: imp = cols(T_gm)@cols(T_init)@
 * init_imp and C = cols(T_gm)@rows(init_C)@rows(T_gm) are formed per particle,
 * with `imp` = init_imp and C_* = init_C_*), 2D/run_2D.py:548-611 (Tgm_x ==
 * NULL: `imp` and C_* are used as given).  Deterministic gather over the
 * cell-sorted particles of pfm_bin_particles (cell_start, perm must describe
 * pos).  Writes u_* and, when non-NULL, the weights W_* (p2g_weight_*). */
int pfm_p2g(const pfm_grid* g, int64_t n, const float* pos, const float* imp, const float* C_x,
            const float* C_y, const float* C_z, const float* Tgm_x, const float* Tgm_y,
            const float* Tgm_z, const float* Tin_x, const float* Tin_y, const float* Tin_z,
            const int32_t* cell_start, const int32_t* perm, float* u_x, float* u_y, float* u_z,
            float* W_x, float* W_y, float* W_z, void* workspace, size_t ws_bytes,
            pfm_stream_t stream);

/* Fused forms used by the step driver (no separate counterpart in the
 * reference; together they compute exactly RK4_T_forward followed by P2G).
 * pfm_rk4_particles_p2g_record = pfm_rk4_particles plus the per-particle P2G
 * composition of 3D/run_3D.py:460-479 with the advected T_grad_m, written as
 * 64-byte records in particle order into `records` (2D: the composition of
 * update_particles_grad_m / update_T / update_particles_imp is not fused; use
 * pfm_p2g).  pfm_p2g_records = the P2G tile and combine passes over those
 * records, read through `perm` (after pfm_bin_particles on the new
 * positions).  pfm_p2g_record_buffer returns the record area inside a P2G
 * workspace (NULL on bad arguments). */
int pfm_rk4_particles_p2g_record(const pfm_grid* g, int64_t n, float* pos, float* T_x, float* T_y,
                                 float* T_z, const float* u_x, const float* u_y, const float* u_z,
                                 float dt, const float* Tin_x, const float* Tin_y, const float* Tin_z,
                                 const float* iC_x, const float* iC_y, const float* iC_z,
                                 const float* init_imp, float* records, pfm_stream_t stream);
/* pfm_rk4_particles_p2g_record for a 3D step whose short reinitialisation
 * runs every step (reinit_every_grad_m = 1, 3D/hyperparameters.py:33): T_grad_m
 * enters as the identity (not read), and T_init is replaced in place by
 * cols(T_grad_m) cols(T_init) -- the update_T of the next step's short
 * reinitialisation (3D/run_3D.py:767-771) -- after the record is formed, so
 * that update_T and reset_T_to_identity(T_grad_m) can be skipped there. */
int pfm_rk4_particles_p2g_record_fold(const pfm_grid* g, int64_t n, float* pos, const float* u_x,
                                      const float* u_y, const float* u_z, float dt, float* Tin_x,
                                      float* Tin_y, float* Tin_z, const float* iC_x, const float* iC_y,
                                      const float* iC_z, const float* init_imp, float* records,
                                      pfm_stream_t stream);
/* pfm_rk4_particles_p2g_record_fold with the count pass of the binning of
 * the advected positions (the cell of each particle and its slot from the
 * per-cell count atomics, get_particles_id_in_every_cell 3D/run_3D.py:342-351)
 * done in the RK4 kernel's epilogue; the records, counts and slots go to the
 * workspace, the cells to cell_of (NULL: the workspace).  pfm_bin_finish then
 * gives the cell_start and perm of pfm_bin_particles on the new positions. */
int pfm_rk4_particles_p2g_record_fold_count(const pfm_grid* g, int64_t n, float* pos, const float* u_x,
                                            const float* u_y, const float* u_z, float dt, float* Tin_x,
                                            float* Tin_y, float* Tin_z, const float* iC_x, const float* iC_y,
                                            const float* iC_z, const float* init_imp, int32_t* cell_of,
                                            void* workspace, size_t ws_bytes, pfm_stream_t stream);
/* The rest of pfm_bin_particles after a fused count pass: exclusive scan of
 * the counts into cell_start (C + 1), scatter by (cell, slot) into perm and the
 * per-cell sort to ascending ids.  cell_of as given to the count pass. */
int pfm_bin_finish(const pfm_grid* g, int64_t n, const int32_t* cell_of, int32_t* cell_start, int32_t* perm,
                   void* workspace, size_t ws_bytes, pfm_stream_t stream);
int pfm_p2g_records(const pfm_grid* g, int64_t n, const int32_t* cell_start, const int32_t* perm,
                    float* u_x, float* u_y, float* u_z, float* W_x, float* W_y, float* W_z,
                    void* workspace, size_t ws_bytes, pfm_stream_t stream);
void* pfm_p2g_record_buffer(const pfm_grid* g, int64_t n, void* workspace);

/* ---- grid utilities ---------------------------------------------------- */

/* calc_max_speed: 3D/run_3D.py:95-103, 2D/run_2D.py:101-108.  Writes the max
 * cell-centre speed (floored at 1e-3) to out[0] (device float). */
int pfm_max_speed(const pfm_grid* g, const float* u_x, const float* u_y, const float* u_z,
                  float* out, pfm_stream_t stream);

/* split_central_vector: 3D/taichi_utils.py:87-100, 2D :56-65.  vc (d, cells). */
int pfm_split_central(const pfm_grid* g, const float* vc, float* u_x, float* u_y, float* u_z,
                      pfm_stream_t stream);

/* get_central_vector: 3D/taichi_utils.py:80-85, 2D :50-54. */
int pfm_get_central(const pfm_grid* g, const float* u_x, const float* u_y, const float* u_z,
                    float* vc, pfm_stream_t stream);

/* add_fields: 3D/taichi_utils.py:18-21 (2D/taichi_utils.py has the same
 * kernel): dest[i] = f1[i] + multiplier * f2[i] over n floats (any field,
 * e.g. the 4-channel smoke; dest may alias f1 or f2). */
int pfm_add_fields(int64_t n, const float* f1, const float* f2, float* dest, float multiplier,
                   pfm_stream_t stream);

/* curl: 3D/taichi_utils.py:56-78 (vc (3, cells) -> out (3, cells)) and
 * 2D/taichi_utils.py:97-103 (vc (2, cells) -> out (cells)); central
 * differences with sample() clamping at the walls. */
int pfm_curl(const pfm_grid* g, const float* vc, float* out, pfm_stream_t stream);

/* leapfrog_vel_func generalised to nv point vortices (2D/init_conditions.py:
 * 5-9,29-60): vc (2, cells) = sum_q w_q * angular_vel_func(r, rad, strength)
 * * (-y, x).  cx, cy, w are host arrays of nv floats (nv <= 4096). */
int pfm_point_vortices(const pfm_grid* g, int nv, const float* cx, const float* cy,
                       const float* w, float rad, float strength, float* vc,
                       pfm_stream_t stream);

/* add_vortex_ring: 3D/init_conditions.py:10-22.  Adds the smoothed
 * Biot-Savart velocity of one ring to vc (3, cells).  c, ex, ey: host float[3]. */
int pfm_vortex_ring(const pfm_grid* g, float w, float rad, float delta, const float* c,
                    const float* ex, const float* ey, int num_samples, float* vc,
                    pfm_stream_t stream);

/* Kinetic energy (2D/run_2D.py:765-770, commented there): sum_a sum u_a^2 /
 * |F_a| accumulated in double into out[0].  `out` is a device buffer of
 * 1 + 3*1024 doubles; out[1..] holds the per-block partial sums. */
/* advect_smoke (3D/run_3D.py:315-323, 2D/run_2D.py:444-450) and clamp_smoke
 * (3D/run_3D.py:325-334): cell-centred smoke (channels x cells, channel-major)
 * sampled with interp_1 at the mean backward map of each cell's faces
 * (psi_x/y/z as pfm_rk4_grid's psi, (dim, face shape) each).  mode 1 writes
 * smoke1 = interp_1(smoke0, psi_c); 2 clamps smoke1 in place to the range of
 * smoke0's 2^d samples around psi_c (3D only); 3 does both. */
int pfm_advect_smoke(const pfm_grid* g, int channels, const float* smoke0, float* smoke1,
                     const float* psi_x, const float* psi_y, const float* psi_z, int mode,
                     pfm_stream_t stream);
/* The smoke part of add_vortex_ring_and_smoke (3D/init_conditions.py:31-56):
 * adds the ring's density into channel 3 of smoke (4 x cells), then sets every
 * cell to (color, 1) where channel 3 > 0.002 and to 0 elsewhere.  The velocity
 * part is pfm_vortex_ring. */
int pfm_vortex_ring_smoke(const pfm_grid* g, float rad, float delta, const float* c, const float* ex,
                          const float* ey, int num_samples, const float* color, float* smoke,
                          pfm_stream_t stream);
/* Particle smoke (3D; 4 channels).  init_particles_smoke
 * (3D/init_conditions.py:275-279, interp_grad_2_smoke 3D/taichi_utils.py:378-421):
 * p_smoke (4, n) and p_grad (12, n: a 3x4 matrix per particle, row-major) from
 * the cell smoke (4, cells). */
int pfm_g2p_smoke(const pfm_grid* g, int64_t n, const float* pos, const float* smoke, float* p_smoke,
                  float* p_grad, pfm_stream_t stream);
/* P2G_smoke (3D/run_3D.py:554-593) as a deterministic gather over the cell
 * sort of pfm_bin_particles: smoke (4, cells) and the weights (cells) are
 * overwritten; with use_apic the particle term is p_smoke + dpos
 * (cols(T_grad_m) cols(T_init) init_grad). */
int pfm_p2g_smoke(const pfm_grid* g, int64_t n, const float* pos, const float* p_smoke,
                  const float* init_grad, const float* Tgm_x, const float* Tgm_y, const float* Tgm_z,
                  const float* Tin_x, const float* Tin_y, const float* Tin_z, int use_apic,
                  const int32_t* cell_start, const int32_t* perm, float* smoke, float* weight,
                  pfm_stream_t stream);
/* clamp_smoke_particle (3D/run_3D.py:595-614) with the cell lists of
 * pfm_bin_particles (the first cell_max particles of a cell by id, as
 * get_particles_id_in_every_cell keeps them). */
int pfm_clamp_smoke_particle(const pfm_grid* g, int64_t n, const int32_t* cell_start,
                             const int32_t* perm, int32_t cell_max, const float* init_pos,
                             const float* init_smoke, float* smoke, pfm_stream_t stream);
int pfm_kinetic_energy(const pfm_grid* g, const float* u_x, const float* u_y, const float* u_z,
                       double* out, pfm_stream_t stream);
/* pfm_kinetic_energy without a caller workspace and without any allocation:
 * no counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, per-step serving loops).  `out` is a
 * device double[1]; the per-block partial sums go to a fixed 24 KiB device
 * array of the library (3 axes x 1024 doubles, one per GPU, kept until the
 * process ends), so any grid size is accepted.  Calls that overlap on one GPU
 * (different streams) race on that array and must be serialised. */
int pfm_kinetic_energy_noalloc(const pfm_grid* g, const float* u_x, const float* u_y, const float* u_z,
                               double* out, pfm_stream_t stream);

/* ---- MGPCG pressure solver --------------------------------------------- */

/* MGPCG.__init__ + MGPCG_2/MGPCG_3.__init__: 3D/mgpcg.py:17-60,350-357,
 * 2D/mgpcg.py.  bc[2*dim] = boundary types (1 Dirichlet, 2 Neumann) per axis
 * (low, high).  precision 0: fp32 levels (the reference); 1: fp16-stored
 * V-cycle levels (fp32 compute).  Allocates the hierarchy. */
int pfm_mg_create(int dim, const int32_t* N, int base_level, const int32_t* bc, int precision,
                  void** handle);
int pfm_mg_destroy(void* handle);
int pfm_mg_n_levels(void* handle);
/* 1: replay each PCG iteration of pfm_mg_solve / pfm_mg_poisson from a HIP
 * graph captured once per handle, enqueuing the iterations the previous
 * solves needed before the first host read; 0 (default): issue the launches
 * directly, one host read per iteration.  Both run the same kernels in the
 * same order (equal results).  No reference counterpart. */
int pfm_mg_set_graphs(void* handle, int enable);

/* Device pointer of an internal field (0 r, 1 z, 2 x, 3 p, 4 Ap, 5 u_div,
 * 6 pressure) at `level` (r and z only), for tests and get_result. */
void* pfm_mg_field(void* handle, int which, int level);

/* Solid obstacle (an extension: the reference solver has none).  `solid`
 * marks solid cells of the finest level (device bytes, nonzero = solid, the
 * cell array's shape); NULL removes the obstacle.  Solid cells carry no
 * pressure; their fluid neighbours see them as walls; the faces touching them
 * are zeroed by apply_bc / poisson (free slip); the recentring mean is over
 * the fluid cells.  A coarse cell is solid when all the cells it owns are. */
int pfm_mg_set_solid(void* h, const uint8_t* solid, pfm_stream_t stream);

/* MGPCG_3.apply_bc (3D/mgpcg.py:359-378), MGPCG_2.apply_bc (2D/mgpcg.py:
 * 292-324). */
int pfm_mg_apply_bc(void* handle, float* u_x, float* u_y, float* u_z, pfm_stream_t stream);

/* divergence (3D/mgpcg.py:380-390) fused with init(u_div, k) (3D/mgpcg.py:
 * 71-79): r = k * div u, z = Ap = p = x = 0. */
int pfm_mg_divergence_init(void* handle, const float* u_x, const float* u_y, const float* u_z,
                           float k, pfm_stream_t stream);

/* init (3D/mgpcg.py:71-79) from an explicit right-hand side field. */
int pfm_mg_init(void* handle, const float* rhs, float k, pfm_stream_t stream);

/* solve (3D/mgpcg.py:202-280).  tol_mode 0: stop when rTr < tol (reference,
 * absolute); 1: stop when sqrt(rTr/r0Tr0) < tol (relative extension).
 * The convergence test runs on the device and the stream is synchronised
 * once per group of iterations (the reference does three host reads per
 * iteration).  A group is one iteration, or with pfm_mg_set_graphs(h, 1) the
 * first group is as long as the shorter of the handle's previous two solves
 * (iterations past convergence change nothing) and each iteration is replayed
 * from a HIP graph captured on the handle's own stream.  Returns PFM_OK or
 * PFM_WARN_NOT_CONVERGED; iters_out / rTr_out may be NULL. */
int pfm_mg_solve(void* handle, int max_iters, double eps, double tol, int tol_mode,
                 int* iters_out, double* rTr_out, pfm_stream_t stream);

/* subtract_grad_p: 3D/mgpcg.py:392-418 using the solution x (get_result). */
int pfm_mg_subtract_grad_p(void* handle, float* u_x, float* u_y, float* u_z,
                           pfm_stream_t stream);

/* Poisson: 3D/mgpcg.py:425-430 (solve_pressure_MGPCG :420-423): apply_bc,
 * divergence, init(-1), solve, get_result, subtract_grad_p, apply_bc. */
int pfm_mg_poisson(void* handle, float* u_x, float* u_y, float* u_z, int max_iters, double tol,
                   int tol_mode, int* iters_out, double* rTr_out, pfm_stream_t stream);

/* Building blocks exposed for parity tests: the V-cycle z0 = M^-1 r0
 * (apply_preconditioner, 3D/mgpcg.py:182-200), one colour of RB-GS on a level
 * (smooth, :165-172), restrict (:152-158), prolongate (:160-163) and the
 * operator Ap = A p on level 0 (get_Ap, :124-129). */
int pfm_mg_vcycle(void* handle, pfm_stream_t stream);
int pfm_mg_smooth(void* handle, int level, int phase, pfm_stream_t stream);
int pfm_mg_restrict(void* handle, int level, pfm_stream_t stream);
int pfm_mg_prolongate(void* handle, int level, pfm_stream_t stream);
int pfm_mg_apply_A(void* handle, const float* p, float* Ap, pfm_stream_t stream);

/* ---- native step driver ------------------------------------------------
 * One scene of the solver stepped by the library: the per-step phase
 * sequence of the reference main loops (3D/run_3D.py:716-831,
 * 2D/run_2D.py:772-909 -- reinit schedule, CFL dt with the frame clip,
 * midpoint velocity, particle RK4, T/imp/C updates, binning, P2G,
 * projections) issued from C++ with exactly the arguments of the per-kernel
 * entry points above.  Every field stays owned by the caller. */
typedef struct {
    int32_t dim;                  /* 2 or 3 */
    int32_t res[3];               /* cells per axis; dx = 1 / res[1] */
    /* hyperparameters.py knobs */
    int32_t reinit_every, reinit_every_grad_m;
    int32_t use_midpoint_vel, reinit_particle_pos;
    double CFL, visualize_dt;
    int32_t ppc_axis;             /* particles per cell = ppc_axis^dim */
    int32_t particle_order;       /* pfm_seed_particles order */
    int64_t seed;                 /* pfm_seed_particles seed (< 0: the reference lattice) */
    /* projection (pfm_mg_poisson) */
    int32_t max_iters, tol_mode;
    double tol;
    int32_t fused_p2g;            /* 3D: pfm_rk4_particles_p2g_record + pfm_p2g_records */
    int32_t fold_T;               /* 3D, fused, reinit_every_grad_m == 1: the _fold variant */
    int64_t n;                    /* particles */
    /* grid fields (face arrays; unused axes NULL) */
    float* u[3];
    float* p2g_weight[3];         /* also the copy of u for advect_u (3D/run_3D.py:783-787) */
    float* psi_grid[3];           /* (dim, face shape) per axis */
    float* T_grid[3];
    /* particle fields, (dim, n) each; T and C are 2D only; imp_grad_m may be NULL */
    float* pos;
    float* init_imp;
    float* imp_grad_m;
    float* T_init[3];
    float* T_grad_m[3];
    float* init_C[3];
    float* imp;
    float* T[3];
    float* C[3];
    /* binning (pfm_bin_particles) and P2G workspace */
    int32_t* cell_of;
    int32_t* cell_start;
    int32_t* perm;
    void* workspace;
    size_t ws_bytes;
    float* max_speed;             /* device float[1] */
    void* mg;                     /* pfm_mg_create handle of the scene's grid */
} pfm_scene_desc;

/* Schedule state (3D/run_3D.py:708-713), read and advanced by each step. */
typedef struct {
    int64_t i;                    /* step index, -1 before the first step */
    int64_t ik;                   /* index of the last reinit that moved ik */
    int64_t frame_idx;
    double sub_t;
    double last_dt;
    int32_t n_solves;             /* projections of the last step (2 with the midpoint) */
    int32_t iters[2];             /* their PCG iteration counts */
    int32_t converged[2];
} pfm_scene_state;

/* Validates the descriptor and keeps a copy of it. */
int pfm_scene_create(const pfm_scene_desc* desc, void** scene);
/* One step.  dt < 0: calc_max_speed + CFL + frame clip (3D/run_3D.py:725-743);
 * else the given dt (sub_t unchanged).  Synchronises `stream` for the max
 * speed and for each projection's convergence tests. */
int pfm_scene_step(void* scene, pfm_scene_state* state, double dt, pfm_stream_t stream);
int pfm_scene_destroy(void* scene);

/* ---- batched 2D scenes (extension) --------------------------------------
 * Many independent 2D scenes of different sizes advanced together: every
 * phase of the 2D step (2D/run_2D.py:772-909) runs as one persistent launch
 * over all scenes, the per-phase launch sequences are captured once into HIP
 * graphs and replayed, and the persistent grid sizes (and the block size of
 * the per-scene coarse multigrid kernel) are picked by pfm_batch2d_tune.
 * The reference runs one scene per process; this has no reference
 * counterpart beyond the per-scene step it repeats.
 *
 * Layout: every field is the concatenation of the scenes' fields in scene
 * order.  Scene s (res_x, res_y = res[2s], res[2s+1], dx = 1/res_y) owns
 * n_s = res_x res_y ppc_axis^2 particles, (res_x+1) res_y x-faces,
 * res_x (res_y+1) y-faces and res_x res_y cells; particle vector fields are
 * (2, P) and grid-attached vector fields (2, F_axis) over the totals
 * (pfm_batch2d_sizes).  Walls are free slip (Neumann) on every side. */
typedef struct {
    int32_t n_scenes;
    const int32_t* res;           /* host: res_x, res_y per scene (>= 4) */
    const int64_t* seed;          /* host: per-scene jitter seed (< 0: the reference lattice) */
    int32_t ppc_axis;             /* sqrt(particles_per_cell) */
    int32_t particle_order;       /* pfm_seed_particles order */
    int32_t reinit_every, reinit_every_grad_m;
    int32_t use_midpoint_vel, reinit_particle_pos;
    double CFL, visualize_dt;
    int32_t base_level;           /* MGPCG levels: int(log2(min res)) - base_level + 1 */
    int32_t max_iters;            /* >= 0 */
    int32_t tol_mode;             /* 0: absolute r.r < tol, 1: relative |r| < tol |r0| */
    double tol;
    float* u[2];                  /* face velocities (x, y) */
    float* p2g_weight[2];         /* P2G weights, also the copy of u before advect_u */
    float* psi_grid[2];           /* (2, F_axis) */
    float* T_grid[2];             /* (2, F_axis) */
    float* pos;                   /* (2, P) */
    float* T[2];
    float* T_init[2];
    float* T_grad_m[2];
    float* C[2];
    float* init_C[2];
    float* imp;
    float* init_imp;
    float* imp_grad_m;            /* may be NULL (written, never read) */
    int32_t* cell_of;             /* P */
    int32_t* cell_start;          /* cells + 1 */
    int32_t* perm;                /* P */
} pfm_batch2d_desc;

typedef struct {
    int64_t i;                    /* step index, -1 before the first step */
    int64_t ik;
    int32_t n_solves;             /* projections of the last step */
    int32_t lockstep_iters[2];    /* PCG iterations the batch ran per projection */
} pfm_batch2d_state;

/* Totals over the scenes: out = {particles, x-faces, y-faces, cells}. */
int pfm_batch2d_sizes(int32_t n_scenes, const int32_t* res, int32_t ppc_axis, int64_t* out);
/* Validates the descriptor, builds the scene and unit tables and allocates
 * the multigrid levels and step workspaces (device bytes: *bytes_out). */
int pfm_batch2d_create(const pfm_batch2d_desc* desc, void** batch, int64_t* bytes_out);
/* Poisson projection of every scene's u (MGPCG_2.Poisson, 2D/mgpcg.py). */
int pfm_batch2d_project(void* batch, pfm_stream_t stream);
/* initialize_2d_particles (2D/run_2D.py:731-738) for every scene. */
int pfm_batch2d_init_particles(void* batch, pfm_stream_t stream);
/* One step of every scene (2D/run_2D.py:772-909; dt per scene from its own
 * max speed, CFL and frame clip).  Synchronises on the batch's own stream
 * once per group of PCG iterations; ordered after prior work on `stream`. */
int pfm_batch2d_step(void* batch, pfm_batch2d_state* state, pfm_stream_t stream);
/* Per-scene results of the last step (host arrays of n_scenes, iters and
 * converged of 2 per scene; any may be NULL).  Synchronises. */
int pfm_batch2d_scene_info(void* batch, double* dt, double* sub_t, int32_t* frame, int32_t* iters,
                           int32_t* converged);
/* Autotuner: runs `rounds` steps per candidate launch shape, timing every
 * kernel kind, and keeps the fastest shape per kind.  The fields are
 * stepped, so the caller restores its state afterwards. */
int pfm_batch2d_tune(void* batch, pfm_batch2d_state* state, pfm_stream_t stream);
/* The chosen shapes: caps[k] = persistent grid cap of kernel kind k (0: one
 * workgroup per unit), then the coarse kernel's block size; returns the
 * number of values written (at most n). */
int pfm_batch2d_tuning(void* batch, int32_t* caps, int32_t n);
/* Scene `scene` takes part in the following steps (1, the default) or is
 * left exactly as it is (0: a finished served job; its units are skipped by
 * every launch of the step except the particle binning).  Synchronises. */
int pfm_batch2d_set_enabled(void* batch, int32_t scene, int32_t enabled);
/* 0: issue every launch directly; 1 (default): capture and replay graphs. */
int pfm_batch2d_set_graphs(void* batch, int32_t enable);
int pfm_batch2d_destroy(void* batch);

#ifdef __cplusplus
}
#endif
#endif /* PFM_H */
