"""Reference-compatible 3D call surface (mirrors 3D/run_3D.py, 3D/init_conditions.py,
3D/taichi_utils.py, 3D/mgpcg.py and 3D/hyperparameters.py)."""
