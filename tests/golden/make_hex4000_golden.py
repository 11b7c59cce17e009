"""Writes hex4000_resonant.npz: oracle solutions of three near-resonant systems of the
hex-monolayer frequency sweep at its full size (N = 4000, grid 210 x 114 x 61, the
9-plane slab on the GPU path), for the GPU parity test in tests/test_gpu_configs.py.

Geometry: 80 x 50 hexagonal monolayer, radius 5 nm, spacing 11 nm, Gaussian jitter
0.2 nm (numpy default_rng(2)), box 880 x 476.31 x 250 nm; Drude eps_inf 4, wp 1.7 eV,
gamma 0.08 eV at 10 frequencies in 0.2-0.8 eV; errortol = tol = 1e-6, xi = 0.5; Jacobi.
The oracle solves each system with GMRES(50) from the mean-field guess to tol 1e-11
(the reference's preconditioned criterion), which leaves the solution error far below the
1e-6 parity bar; the test compares GPU solves at tol 1e-6 and 1e-10 with it.

Run from the repository root: python tests/golden/make_hex4000_golden.py (about 10
minutes on 8 cores).
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))
from oracle import oracle as O  # noqa: E402
from gpu_helpers import hex_monolayer  # noqa: E402

SYSTEMS = [(6, 0), (7, 0), (7, 2)]  # (frequency index, polarisation)


def main():
    pos, box = hex_monolayer(80, 50, 11.0, jitter=0.2, seed=2)
    box = np.array([box[0], box[1], 250.0])
    a = 5.0
    omega = np.linspace(0.2, 0.8, 10)
    eps = 4.0 - 1.7 ** 2 / (omega ** 2 + 1j * omega * 0.08)
    L, P = box / a, pos / a
    N = len(P)
    op = O.Operator(L, 1e-6, 0.5)
    op.set_positions(P)
    vol_frac = 4.0 / 3.0 * np.pi * N / np.prod(L)
    xs, iters = [], []
    for w, k in SYSTEMS:
        sc = np.full(N, O.self_coefficient(eps[w]))
        b = np.zeros((N, 3), np.complex128)
        b[:, k] = 1.0
        beta = (eps[w] - 1.0) / (eps[w] + 2.0)
        x0 = 4.0 * np.pi * beta / (1.0 - beta * vol_frac) * b
        t = time.time()
        x, rep = op.solve(b, sc, x0, tol=1e-11, restart=50, maxiter=5000)
        print(f"system {(w, k)}: {rep} in {time.time() - t:.1f} s", flush=True)
        assert rep["converged"]
        xs.append(x.reshape(N, 3))
        iters.append(rep["iters"])
    np.savez_compressed(os.path.join(HERE, "hex4000_resonant.npz"), positions=P, box=L, eps=eps,
                        systems=np.array(SYSTEMS), x=np.stack(xs), iters=np.array(iters))


if __name__ == "__main__":
    main()
