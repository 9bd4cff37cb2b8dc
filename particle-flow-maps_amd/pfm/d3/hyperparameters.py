"""Knobs of 3D/hyperparameters.py, same names and values.

``case`` selects the scene as in the reference: 0 is the leapfrogging rings
(3D/hyperparameters.py:8-44, ``init_vorts_leapfrog``), 1 the eight vortex
rings (:45-80, ``init_vorts_eight``).  The module-level names are those of
``case``; ``knobs(case)`` returns either block as a dict.  The reference
star-imports this module into its driver; ``pfm.simulation.StepConfig``
takes the same names.  The encoder and neural-buffer knobs of the reference
belong to a component outside the solver step and are not carried."""

case = 0  # leapfrogs
# case = 1  # eight vortices

dim = 3

_COMMON = dict(
    use_midpoint_vel=True,
    forward_update_T=True,
    use_APIC=True,
    use_APIC_smoke=True,
    save_particle_pos_numpy=False,
    save_frame_each_step=False,
    reinit_particle_pos=True,
    visualize_dt=0.1,
    ckpt_every=1,
    CFL=0.5,
    from_frame=0,
    use_total_steps=False,
    total_steps=100,
    particles_per_cell=8,
    total_particles_num_ratio=1,
    cell_max_particle_num_ratio=1.6,
)

_CASES = {
    0: dict(_COMMON, save_particle_imp_numpy=False, res_x=256, res_y=128, res_z=128, reinit_every=20,
            reinit_every_grad_m=1, total_frames=600, exp_name="3D_leapfrog_reinit-20-1"),
    1: dict(_COMMON, res_x=128, res_y=128, res_z=128, reinit_every=12, reinit_every_grad_m=4,
            total_frames=1200, exp_name="3D_eight_vorts_reinit-12-4"),
}


def knobs(which: int) -> dict:
    """The knob block of scene ``which`` (0: leapfrog, 1: eight vortices)."""
    return dict(_CASES[which])


globals().update(_CASES[case])
