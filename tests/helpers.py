"""Shared test helpers (scene construction and comparisons)."""
import numpy as np


def assembled_system(scene, perturb=0.0, seed=0, steps=0):
    """Oracle world for the scene; optionally perturb positions so the elastic
    Hessian is away from rest.  Returns (world, grad, bsr)."""
    import pyoracle as O
    w = O.World(scene)
    if perturb:
        rng = np.random.default_rng(seed)
        x = scene.x + perturb * rng.normal(size=scene.x.shape)
        w.set_positions(x)
    for _ in range(steps):
        w.time_step()
    g, A = w.assemble()
    return w, g, A


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))
