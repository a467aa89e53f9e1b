"""Write golden vectors of the reference's L0 numerics.

Run where the reference tree is present (after ``make -C oracle ref``):

    python tests/golden/make_golden.py

The outputs come from oracle/_ref/ref_numerics, a driver that compiles the
reference's own kernels/_interpolation/linear.pxd and
kernels/roms/_vertical_coordinate.pxd unmodified.  The file written,
``l0_vectors.npz``, holds inputs and expected outputs only.
"""

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle", "_ref"))
sys.path.insert(0, ROOT)


def inputs(seed=2024):
    rng = np.random.default_rng(seed)
    a1 = rng.standard_normal(7)
    a2 = rng.standard_normal((5, 6))
    a3 = rng.standard_normal((4, 5, 6))
    edge = np.array([-3.7, -1.0, -0.999, -0.5, -1e-17, 0.0, 1e-300, 0.25, 0.5, 0.999999, 1.0, 2.5,
                     3.0, 3.999999999, 4.0, 4.5, 5.0, 5.5, 6.0, 7.25, 100.0, 1e19, -1e19, np.inf, -np.inf, np.nan])
    q1 = np.concatenate([edge, rng.uniform(-2, 9, 200)])
    q2 = [np.concatenate([edge, rng.uniform(-2, 7, 200)]) for _ in range(2)]
    q2[1] = rng.permutation(q2[1])
    q3 = [np.concatenate([edge, rng.uniform(-2, 7, 200)]) for _ in range(3)]
    q3[1] = rng.permutation(q3[1])
    q3[2] = rng.permutation(q3[2])
    return a1, a2, a3, q1, q2, q3


def vcoord_inputs(seed=7):
    from offline_particles_amd.workloads import stretching_cs_w

    rng = np.random.default_rng(seed)
    NZ = 8
    C = stretching_cs_w(NZ)
    n = 300
    h = rng.uniform(5.0, 4000.0, n)
    zeta = rng.uniform(-1.0, 1.0, n)
    zidx = np.concatenate([rng.uniform(-1.0, NZ, n - 6), [-0.5, 0.0, NZ - 1.0, NZ - 0.5, -2.0, NZ + 1.0]])
    Cz = rng.uniform(-1.0, 0.0, n)
    # z from the surface to below the bed, including exact nodes
    z = np.concatenate([-h[: n - 10] * rng.uniform(-0.2, 1.2, n - 10), -h[n - 10:] * 0.5])
    return NZ, C, h, zeta, zidx, Cz, z


def main():
    import ref_numerics as R

    a1, a2, a3, q1, q2, q3 = inputs()
    NZ, C, h, zeta, zidx, Cz, z = vcoord_inputs()
    hc = 10.0
    out = dict(
        a1=a1, a2=a2, a3=a3, q1=q1, q2a=q2[0], q2b=q2[1], q3a=q3[0], q3b=q3[1], q3c=q3[2],
        trunc=R.ref_truncate_index(q1, 5),
        lin=R.ref_linear(a1, q1),
        bil=R.ref_bilinear(a2, q2[0], q2[1]),
        tril=R.ref_trilinear(a3, q3[0], q3[1], q3[2]),
        NZ=np.int64(NZ), C=C, h=h, zeta=zeta, zidx=zidx, Cz=Cz, z=z, hc=np.float64(hc),
        compute_z=R.ref_compute_z(zidx, NZ, hc, h, Cz, zeta),
        compute_zidx=R.ref_compute_zidx(z, h, zeta, hc, NZ, C, 0.5),
    )
    path = os.path.join(HERE, "l0_vectors.npz")
    np.savez_compressed(path, **out)
    print("wrote", path)


if __name__ == "__main__":
    main()
