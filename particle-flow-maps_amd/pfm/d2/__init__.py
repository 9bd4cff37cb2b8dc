"""Reference-compatible 2D call surface (mirrors 2D/run_2D.py, 2D/init_conditions.py,
2D/taichi_utils.py, 2D/mgpcg.py and 2D/hyperparameters.py)."""
