"""Initial conformations (chromo/util/poly_paths.py:269-295, 506-545).

Host-side numpy set-up of the starting chain, drawn from numpy's global stream
exactly as the reference draws it, so a seeded script starts from the same
conformation.
"""
from __future__ import annotations

from typing import Tuple

import numpy as np


def gaussian_walk(num_steps: int, step_sizes: np.ndarray) -> np.ndarray:
    """Random walk with unit-Gaussian directions scaled to ``step_sizes``."""
    steps = np.random.standard_normal((num_steps, 3))
    magnitude_steps = np.linalg.norm(steps, axis=1)
    positions = np.cumsum(
        np.divide(steps, magnitude_steps[:, None]) * step_sizes[:, np.newaxis], axis=0)
    return np.vstack((np.zeros(3), positions))


def estimate_tangents_from_coordinates(coordinates: np.ndarray) -> Tuple[np.ndarray, np.ndarray]:
    """t3 from (central) differences, t2 = normalised t3 x e_x (or t3 x e_y)."""
    num_beads = coordinates.shape[0]
    t3 = np.zeros((num_beads, 3))
    t2 = np.zeros((num_beads, 3))
    first_diff = coordinates[1, :] - coordinates[0, :]
    last_diff = coordinates[num_beads - 1, :] - coordinates[num_beads - 2, :]
    t3[0, :] = first_diff / np.linalg.norm(first_diff)
    t3[num_beads - 1, :] = last_diff / np.linalg.norm(last_diff)
    for i in range(1, num_beads - 1):
        d = coordinates[i + 1, :] - coordinates[i - 1, :]
        t3[i, :] = d / np.linalg.norm(d)
    e1 = np.array([1, 0, 0])
    e2 = np.array([0, 1, 0])
    for i in range(num_beads):
        trial = np.cross(t3[i, :], e1)
        if np.all(trial == 0):
            trial = np.cross(t3[i, :], e2)
        t2[i, :] = trial / np.linalg.norm(trial)
    return t3, t2
