// scene.cpp -- native step driver of one scene: the reference main loop's
// per-step phase sequence (3D/run_3D.py:716-831, 2D/run_2D.py:772-909) issued
// from C++ as one library call.
//
// The Python driver (pfm.simulation.Simulation) issues the same entry points
// one by one; this driver runs the identical sequence with the identical
// arguments, so the two produce the same bits.  One call per step keeps the
// host side of a step out of the Python interpreter: every synchronisation of
// the step (the CFL max-speed read, the MGPCG convergence tests) happens with
// the GIL released, so host threads driving different scenes on different
// streams overlap fully.
#include <new>

#include "pfm_host.h"

namespace pfm {
namespace {

struct Scene {
    pfm_scene_desc d;
    pfm_grid grid;       // dx = 1 / res[1] (3D/run_3D.py:16, 2D/run_2D.py:14)
    double dx;
    float* ms_host;      // pinned: the max speed read back each step
    int64_t faces[3];
};

// The reinitialisation decisions of step i (3D/run_3D.py:748-768,
// 2D/run_2D.py:804-823): long reinit when i % reinit_every == 0; the 3D driver
// moves ik only when it reseeds, the 2D driver at every long reinit; short
// reinit when (i - ik) % reinit_every_grad_m == 0.
struct Decision {
    bool long_r, reseed, short_r;
};

Decision decide(const pfm_scene_desc& d, int64_t i, int64_t* ik) {
    Decision r;
    r.long_r = i % d.reinit_every == 0;
    r.reseed = r.long_r && (d.reinit_particle_pos || i == 0);
    if (r.long_r && (r.reseed || d.dim == 2)) *ik = i;
    r.short_r = (i - *ik) % d.reinit_every_grad_m == 0;
    return r;
}

#define PFM_STEP_TRY(expr)          \
    do {                            \
        int _rc = (expr);           \
        if (_rc < 0) return _rc;    \
    } while (0)

// calc_max_speed + CFL + frame clip (3D/run_3D.py:725-743) in host double.
int compute_dt(Scene* sc, pfm_scene_state* st, hipStream_t s, double* dt) {
    const pfm_scene_desc& d = sc->d;
    PFM_STEP_TRY(pfm_max_speed(&sc->grid, d.u[0], d.u[1], d.u[2], d.max_speed, s));
    PFM_HIP_TRY(hipMemcpyAsync(sc->ms_host, d.max_speed, sizeof(float), hipMemcpyDeviceToHost, s));
    PFM_HIP_TRY(hipStreamSynchronize(s));
    double curr = d.CFL * sc->dx / (double)sc->ms_host[0];
    if (st->sub_t + curr >= d.visualize_dt) {
        curr = d.visualize_dt - st->sub_t;
        st->sub_t = 0.0;
        st->frame_idx += 1;
    } else {
        st->sub_t += curr;
    }
    *dt = curr;
    return PFM_OK;
}

int project(Scene* sc, pfm_scene_state* st, hipStream_t s) {
    const pfm_scene_desc& d = sc->d;
    int it = 0;
    double rtr = 0.0;
    const int rc = pfm_mg_poisson(d.mg, d.u[0], d.u[1], d.u[2], d.max_iters, d.tol, d.tol_mode, &it, &rtr, s);
    if (rc < 0) return rc;
    if (st->n_solves < 2) {
        st->iters[st->n_solves] = it;
        st->converged[st->n_solves] = rc == PFM_OK;
    }
    st->n_solves += 1;
    return PFM_OK;
}

int copy(void* dst, const void* src, int64_t nfloat, hipStream_t s) {
    PFM_HIP_TRY(hipMemcpyAsync(dst, src, sizeof(float) * (size_t)nfloat, hipMemcpyDeviceToDevice, s));
    return PFM_OK;
}

int seed(Scene* sc, hipStream_t s) {
    const pfm_scene_desc& d = sc->d;
    const double dist = sc->dx / d.ppc_axis;
    return pfm_seed_particles(&sc->grid, d.ppc_axis, (float)dist, d.n, d.pos, d.seed, d.particle_order, s);
}

// The midpoint velocity (3D/run_3D.py:778-789, 2D/run_2D.py:865-872):
// backward RK4 of the identity map over dt/2 on every face, advection of u
// through it (the p2g weight arrays hold the copy of u), projection.
int midpoint(Scene* sc, pfm_scene_state* st, double dt, hipStream_t s) {
    const pfm_scene_desc& d = sc->d;
    // the copy of u first: the fused launch samples it for RK4_grid and
    // advect_u alike and writes the advected velocity into u
    for (int a = 0; a < d.dim; ++a) PFM_STEP_TRY(copy(d.p2g_weight[a], d.u[a], sc->faces[a], s));
    PFM_STEP_TRY(pfm_midpoint_advect(&sc->grid, d.p2g_weight[0], d.p2g_weight[1], d.p2g_weight[2], d.u[0], d.u[1],
                                     d.u[2], d.psi_grid[0], d.psi_grid[1], d.psi_grid[2], d.T_grid[0], d.T_grid[1],
                                     d.T_grid[2], (float)(0.5 * dt), s));
    return project(sc, st, s);
}

int step3(Scene* sc, pfm_scene_state* st, double dt, hipStream_t s) {
    const pfm_scene_desc& d = sc->d;
    const pfm_grid* g = &sc->grid;
    const Decision r = decide(d, st->i, &st->ik);
    if (r.long_r) {
        if (r.reseed) PFM_STEP_TRY(seed(sc, s));
        PFM_STEP_TRY(pfm_g2p(g, d.n, d.pos, d.u[0], d.u[1], d.u[2], d.init_imp, nullptr, nullptr, nullptr, s));
        PFM_STEP_TRY(pfm_reset_T_identity(3, d.n, d.T_init[0], d.T_init[1], d.T_init[2], s));
        PFM_STEP_TRY(pfm_reset_T_identity(3, d.n, d.T_grad_m[0], d.T_grad_m[1], d.T_grad_m[2], s));
    }
    if (r.short_r) {
        PFM_STEP_TRY(pfm_g2p(g, d.n, d.pos, d.u[0], d.u[1], d.u[2], d.imp_grad_m, d.init_C[0], d.init_C[1],
                             d.init_C[2], s));
        // with fold_T the previous step's RK4 already applied update_T and
        // left T_grad_m at the identity
        if (!d.fold_T) {
            PFM_STEP_TRY(pfm_update_T(3, d.n, d.T_grad_m[0], d.T_grad_m[1], d.T_grad_m[2], d.T_init[0],
                                      d.T_init[1], d.T_init[2], d.T_init[0], d.T_init[1], d.T_init[2], s));
            PFM_STEP_TRY(pfm_reset_T_identity(3, d.n, d.T_grad_m[0], d.T_grad_m[1], d.T_grad_m[2], s));
        }
    }
    if (d.use_midpoint_vel) PFM_STEP_TRY(midpoint(sc, st, dt, s));
    if (d.fused_p2g) {
        // RK4_T_forward with the P2G composition fused in, binning, P2G
        // (3D/run_3D.py:795-812)
        float* rec = (float*)pfm_p2g_record_buffer(g, d.n, d.workspace);
        PFM_CHECK_ARG(rec != nullptr, "pfm_p2g_record_buffer failed");
        if (d.fold_T) {
            // the binning's count pass rides in the RK4 kernel
            PFM_STEP_TRY(pfm_rk4_particles_p2g_record_fold_count(g, d.n, d.pos, d.u[0], d.u[1], d.u[2], (float)dt,
                                                                 d.T_init[0], d.T_init[1], d.T_init[2], d.init_C[0],
                                                                 d.init_C[1], d.init_C[2], d.init_imp, d.cell_of,
                                                                 d.workspace, d.ws_bytes, s));
            PFM_STEP_TRY(pfm_bin_finish(g, d.n, d.cell_of, d.cell_start, d.perm, d.workspace, d.ws_bytes, s));
        } else {
            PFM_STEP_TRY(pfm_rk4_particles_p2g_record(g, d.n, d.pos, d.T_grad_m[0], d.T_grad_m[1], d.T_grad_m[2],
                                                      d.u[0], d.u[1], d.u[2], (float)dt, d.T_init[0], d.T_init[1],
                                                      d.T_init[2], d.init_C[0], d.init_C[1], d.init_C[2],
                                                      d.init_imp, rec, s));
            PFM_STEP_TRY(pfm_bin_particles(g, d.n, d.pos, d.cell_of, d.cell_start, d.perm, d.workspace, d.ws_bytes,
                                           s));
        }
        PFM_STEP_TRY(pfm_p2g_records(g, d.n, d.cell_start, d.perm, d.u[0], d.u[1], d.u[2], d.p2g_weight[0],
                                     d.p2g_weight[1], d.p2g_weight[2], d.workspace, d.ws_bytes, s));
    } else {
        PFM_STEP_TRY(pfm_rk4_particles(g, d.n, d.pos, d.T_grad_m[0], d.T_grad_m[1], d.T_grad_m[2], d.u[0], d.u[1],
                                       d.u[2], (float)dt, 1, s));
        PFM_STEP_TRY(pfm_bin_particles(g, d.n, d.pos, d.cell_of, d.cell_start, d.perm, d.workspace, d.ws_bytes, s));
        PFM_STEP_TRY(pfm_p2g(g, d.n, d.pos, d.init_imp, d.init_C[0], d.init_C[1], d.init_C[2], d.T_grad_m[0],
                             d.T_grad_m[1], d.T_grad_m[2], d.T_init[0], d.T_init[1], d.T_init[2], d.cell_start,
                             d.perm, d.u[0], d.u[1], d.u[2], d.p2g_weight[0], d.p2g_weight[1], d.p2g_weight[2],
                             d.workspace, d.ws_bytes, s));
    }
    return project(sc, st, s);
}

int step2(Scene* sc, pfm_scene_state* st, double dt, hipStream_t s) {
    const pfm_scene_desc& d = sc->d;
    const pfm_grid* g = &sc->grid;
    const int64_t nv = 2 * d.n;
    const Decision r = decide(d, st->i, &st->ik);
    if (r.long_r) {
        if (r.reseed) PFM_STEP_TRY(seed(sc, s));
        PFM_STEP_TRY(pfm_g2p(g, d.n, d.pos, d.u[0], d.u[1], nullptr, d.imp, nullptr, nullptr, nullptr, s));
        PFM_STEP_TRY(copy(d.init_imp, d.imp, nv, s));
        PFM_STEP_TRY(pfm_reset_T_identity(2, d.n, d.T[0], d.T[1], nullptr, s));
        PFM_STEP_TRY(pfm_reset_T_identity(2, d.n, d.T_init[0], d.T_init[1], nullptr, s));
    }
    if (r.short_r) {
        PFM_STEP_TRY(pfm_g2p(g, d.n, d.pos, d.u[0], d.u[1], nullptr, d.imp_grad_m, d.init_C[0], d.init_C[1],
                             nullptr, s));
        PFM_STEP_TRY(pfm_reset_T_identity(2, d.n, d.T_grad_m[0], d.T_grad_m[1], nullptr, s));
        for (int a = 0; a < 2; ++a) PFM_STEP_TRY(copy(d.T_init[a], d.T[a], nv, s));
    }
    if (d.use_midpoint_vel) PFM_STEP_TRY(midpoint(sc, st, dt, s));
    PFM_STEP_TRY(pfm_rk4_particles(g, d.n, d.pos, d.T_grad_m[0], d.T_grad_m[1], nullptr, d.u[0], d.u[1], nullptr,
                                   (float)dt, 1, s));
    PFM_STEP_TRY(pfm_update_particles_grad_m(2, d.n, d.C[0], d.C[1], nullptr, d.init_C[0], d.init_C[1], nullptr,
                                             d.T_grad_m[0], d.T_grad_m[1], nullptr, s));
    PFM_STEP_TRY(pfm_update_T(2, d.n, d.T_grad_m[0], d.T_grad_m[1], nullptr, d.T_init[0], d.T_init[1], nullptr,
                              d.T[0], d.T[1], nullptr, s));
    PFM_STEP_TRY(pfm_update_particles_imp(2, d.n, d.T[0], d.T[1], nullptr, d.init_imp, d.imp, s));
    PFM_STEP_TRY(pfm_bin_particles(g, d.n, d.pos, d.cell_of, d.cell_start, d.perm, d.workspace, d.ws_bytes, s));
    PFM_STEP_TRY(pfm_p2g(g, d.n, d.pos, d.imp, d.C[0], d.C[1], nullptr, nullptr, nullptr, nullptr, nullptr, nullptr,
                         nullptr, d.cell_start, d.perm, d.u[0], d.u[1], nullptr, d.p2g_weight[0], d.p2g_weight[1],
                         nullptr, d.workspace, d.ws_bytes, s));
    return project(sc, st, s);
}

}  // namespace
}  // namespace pfm

using namespace pfm;

extern "C" int pfm_scene_create(const pfm_scene_desc* desc, void** scene) {
    PFM_CHECK_ARG(scene != nullptr, "null scene out-pointer");
    *scene = nullptr;
    PFM_CHECK_ARG(desc != nullptr, "null scene descriptor");
    const pfm_scene_desc& d = *desc;
    PFM_CHECK_ARG(d.dim == 2 || d.dim == 3, "dim must be 2 or 3");
    PFM_CHECK_ARG(d.reinit_every >= 1 && d.reinit_every_grad_m >= 1, "reinit intervals must be >= 1");
    PFM_CHECK_ARG(d.ppc_axis >= 1 && d.n >= 1, "bad particle counts");
    PFM_CHECK_ARG(d.CFL > 0.0 && d.visualize_dt > 0.0, "CFL and visualize_dt must be positive");
    PFM_CHECK_ARG(!d.fold_T || (d.fused_p2g && d.reinit_every_grad_m == 1),
                  "fold_T needs the fused path and a short reinitialisation every step");
    PFM_CHECK_ARG(d.mg && d.max_speed && d.pos && d.init_imp && d.cell_start && d.perm && d.workspace,
                  "null field pointer");
    for (int a = 0; a < d.dim; ++a)
        PFM_CHECK_ARG(d.u[a] && d.p2g_weight[a] && d.psi_grid[a] && d.T_grid[a] && d.T_init[a] && d.T_grad_m[a] &&
                          d.init_C[a],
                      "null field pointer (axis %d)", a);
    if (d.dim == 2) {
        PFM_CHECK_ARG(d.imp && d.T[0] && d.T[1] && d.C[0] && d.C[1], "2D scenes need imp, T and C");
        PFM_CHECK_ARG(!d.fused_p2g && !d.fold_T, "the fused RK4 + P2G record path is 3D only");
    }
    Scene* sc = new (std::nothrow) Scene();
    if (!sc) {
        set_error("host allocation failed");
        return PFM_ERR_OOM;
    }
    sc->d = d;
    sc->grid.dim = d.dim;
    for (int a = 0; a < 3; ++a) sc->grid.res[a] = a < d.dim ? d.res[a] : 1;
    sc->dx = 1.0 / d.res[1];
    sc->grid.dx = (float)sc->dx;
    GridDev gd;
    int rc = make_grid(&sc->grid, &gd);
    if (rc) {
        delete sc;
        return rc;
    }
    PFM_CHECK_ARG(d.n == cell_count(gd) * (int64_t)(d.dim == 3 ? d.ppc_axis * d.ppc_axis * d.ppc_axis
                                                               : d.ppc_axis * d.ppc_axis),
                  "n must be cells x ppc_axis^dim");
    for (int a = 0; a < 3; ++a) sc->faces[a] = a < d.dim ? face_count_h(gd, a) : 0;
    const size_t need = pfm_p2g_workspace_size(&sc->grid, d.n);
    if (d.ws_bytes < need) {
        delete sc;
        PFM_CHECK_ARG(false, "workspace too small: %zu < %zu bytes", d.ws_bytes, need);
    }
    if (hipHostMalloc((void**)&sc->ms_host, sizeof(float), hipHostMallocDefault) != hipSuccess) {
        delete sc;
        set_error("hipHostMalloc failed");
        return PFM_ERR_OOM;
    }
    *scene = sc;
    return PFM_OK;
}

extern "C" int pfm_scene_step(void* scene, pfm_scene_state* st, double dt, pfm_stream_t stream) {
    PFM_CHECK_ARG(scene && st, "null scene or state");
    Scene* sc = (Scene*)scene;
    hipStream_t s = as_stream(stream);
    st->i += 1;
    st->n_solves = 0;
    if (dt < 0.0) PFM_STEP_TRY(compute_dt(sc, st, s, &dt));
    st->last_dt = dt;
    return sc->d.dim == 3 ? step3(sc, st, dt, s) : step2(sc, st, dt, s);
}

extern "C" int pfm_scene_destroy(void* scene) {
    if (!scene) return PFM_OK;
    Scene* sc = (Scene*)scene;
    if (sc->ms_host) (void)hipHostFree(sc->ms_host);
    delete sc;
    return PFM_OK;
}
