"""Writes the small golden fixtures of tests/golden/ (run in the build container).

The reference's own tests store no golden vectors (they recompute direct
sums at run time, tests/nuft_direct_kernels.cpp of the reference), so these
fixtures are produced here by brute force in numpy FP64: explicit
exp(i sign k.x) matrices, independent of the C oracle. Inputs follow the
reference tests' layout: coordinates far outside [-pi, pi) (x in [20,30],
y in [-20,20], z in [-20,-15], tests/test_type_1_2.cpp:210-225) and
strengths uniform in [-4,4]^2 (tests/test_type_1_2.cpp:124).

Usage: python tests/golden/gen_golden.py
"""
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def modes_1d(N):
    return np.arange(N) - N // 2


def kgrid(modes):
    axes = [modes_1d(N) for N in modes]
    mesh = np.meshgrid(*axes, indexing="ij")
    # contiguous CMCL order with x fastest: flatten in Fortran order
    return np.stack([m.ravel(order="F") for m in mesh], axis=1)


def coords(rng, dim, m):
    lo_hi = [(20.0, 30.0), (-20.0, 20.0), (-20.0, -15.0)]
    return np.stack([rng.uniform(*lo_hi[d], m) for d in range(dim)], axis=0)


def cplx(rng, m):
    return rng.uniform(-4, 4, m) + 1j * rng.uniform(-4, 4, m)


def main():
    rng = np.random.default_rng(20240601)
    cases = {}
    for dim, modes, m in [(1, (23,), 37), (2, (12, 7), 41), (3, (6, 5, 4), 29)]:
        for sign in (-1, 1):
            x = coords(rng, dim, m)
            c = cplx(rng, m)
            k = kgrid(modes)                         # (K, dim)
            E = np.exp(1j * sign * (k @ x))          # (K, m)
            f1 = E @ c                               # type 1
            f = cplx(rng, k.shape[0])
            c2 = E.T @ f                             # type 2 (same exponent matrix)
            s = coords(rng, dim, 17) * 0.7 - 3.0
            c3 = np.exp(1j * sign * (s.T @ x)) @ c   # type 3
            key = f"d{dim}_s{'p' if sign > 0 else 'm'}"
            cases[key + "_x"] = x
            cases[key + "_c"] = c
            cases[key + "_modes"] = np.array(modes)
            cases[key + "_t1"] = f1
            cases[key + "_f"] = f
            cases[key + "_t2"] = c2
            cases[key + "_s"] = s
            cases[key + "_t3"] = c3
    np.savez_compressed(os.path.join(HERE, "direct_small.npz"), **cases)
    print("wrote", os.path.join(HERE, "direct_small.npz"), len(cases), "arrays")


if __name__ == "__main__":
    main()
