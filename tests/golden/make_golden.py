"""Writes the golden vectors under tests/golden/ (run from the repo root).

se3_expm.npz  - twists (omega, v) and the 4x4 matrix exponential of the
                se(3) element [[omega]x, v; 0 0] from scipy.linalg.expm, the
                textbook map that exp_se3 (include/se3_math.cuh:40-98) is a
                closed form of.  Includes |omega| < 1e-6 (first-order branch)
                and rotations near pi about each axis (all four Shepperd
                branches of mat4_to_dq, se3_math.cuh:252-283).
"""
import os

import numpy as np
from scipy.linalg import expm

HERE = os.path.dirname(os.path.abspath(__file__))


def hat(w):
    return np.array([[0, -w[2], w[1]], [w[2], 0, -w[0]], [-w[1], w[0], 0]])


def main():
    rng = np.random.default_rng(20261203)
    tw = [rng.normal(0, 0.3, 6) for _ in range(40)]
    tw += [np.concatenate([rng.normal(0, 1e-8, 3), rng.normal(0, 0.05, 3)]) for _ in range(6)]
    tw += [np.array([0, 0, 0, 0.1, -0.2, 0.3])]
    for ax in range(3):  # near-pi rotations exercise the trace <= 0 branches
        w = np.zeros(3)
        w[ax] = np.pi * 0.97
        tw.append(np.concatenate([w, rng.normal(0, 0.1, 3)]))
        w2 = rng.normal(0, 0.05, 3)
        w2[ax] += np.pi * 0.9
        tw.append(np.concatenate([w2, rng.normal(0, 0.1, 3)]))
    tw = np.array(tw)
    mats = []
    for t in tw:
        m = np.zeros((4, 4))
        m[:3, :3] = hat(t[:3])
        m[:3, 3] = t[3:]
        mats.append(expm(m))
    np.savez(os.path.join(HERE, "se3_expm.npz"), twists=tw, expm=np.array(mats))


if __name__ == "__main__":
    main()
