"""Test helper: the light curves of adiabatic.hpp:228-333 in numpy."""
import numpy as np


def light_curves(U, fc, area, gam):
    """adiabatic.hpp:228-333 restated with numpy, face by face."""
    d = 9.4 * 3 * 1e19 / 15000
    sq = np.sqrt(d)
    obs = np.array([[d, 0, 0], [sq, 0, sq], [0, 0, d], [0, 0, -d]])
    n = np.linalg.norm(fc, axis=1)
    r = fc / n[:, None]
    phi = np.arctan2(fc[:, 1] / n, fc[:, 0] / n)
    theta = np.arccos(fc[:, 2] / n)
    vel = np.cross(r, U[:, 1:4]) / (-U[:, 0])[:, None]
    PI = np.maximum(1e-16, (U[:, 4] - U[:, 0] * np.sum(vel * vel, axis=1) / 2) * (gam - 1))
    cos = np.abs(r @ obs.T) / np.linalg.norm(obs, axis=1)[None, :]
    masks = [(phi < np.pi / 2) & (phi > -np.pi / 2), r @ obs[1] > 0, theta < np.pi / 2, theta > np.pi / 2]
    return np.array([np.sum((PI * cos[:, k] * area)[masks[k]]) for k in range(4)])
