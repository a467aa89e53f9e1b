# cython: language_level=3
"""Thin driver over the reference's header-only numeric core.  TEST INFRASTRUCTURE ONLY.

Built by `make -C oracle ref` into oracle/_ref/ from the reference's own .pxd
files, which are cimported in place (nothing is copied):

- kernels/_interpolation/linear.pxd: truncate_index, linear/bilinear/trilinear
- kernels/roms/_vertical_coordinate.pxd: compute_z, compute_zidx

Every wrapper loops over sample points so that one call produces a whole
vector of outputs.  These outputs pin the C oracle's L0 functions.
"""

import numpy as np

from offline_particles.kernels._interpolation.linear cimport (
    bilinear_interpolation,
    linear_interpolation,
    trilinear_interpolation,
    truncate_index,
)
from offline_particles.kernels.roms._vertical_coordinate cimport compute_z, compute_zidx


def ref_truncate_index(double[::1] idx, Py_ssize_t max_idx):
    cdef Py_ssize_t i
    out = np.empty(idx.shape[0], dtype=np.int64)
    cdef long long[::1] o = out
    for i in range(idx.shape[0]):
        o[i] = truncate_index(idx[i], max_idx)
    return out


def ref_linear(const double[::1] a, double[::1] idx):
    cdef Py_ssize_t i
    out = np.empty(idx.shape[0])
    cdef double[::1] o = out
    for i in range(idx.shape[0]):
        o[i] = linear_interpolation(a, idx[i])
    return out


def ref_bilinear(const double[:, ::1] a, double[::1] i0, double[::1] i1):
    cdef Py_ssize_t i
    out = np.empty(i0.shape[0])
    cdef double[::1] o = out
    for i in range(i0.shape[0]):
        o[i] = bilinear_interpolation(a, i0[i], i1[i])
    return out


def ref_trilinear(const double[:, :, ::1] a, double[::1] i0, double[::1] i1, double[::1] i2):
    cdef Py_ssize_t i
    out = np.empty(i0.shape[0])
    cdef double[::1] o = out
    for i in range(i0.shape[0]):
        o[i] = trilinear_interpolation(a, i0[i], i1[i], i2[i])
    return out


def ref_compute_z(double[::1] zidx, int NZ, double hc, double[::1] h, double[::1] C, double[::1] zeta):
    cdef Py_ssize_t i
    out = np.empty(zidx.shape[0])
    cdef double[::1] o = out
    for i in range(zidx.shape[0]):
        o[i] = compute_z(zidx[i], NZ, hc, h[i], C[i], zeta[i])
    return out


def ref_compute_zidx(double[::1] z, double[::1] h, double[::1] zeta, double hc, int NZ,
                     const double[::1] Carr, double C_offset):
    cdef Py_ssize_t i
    out = np.empty(z.shape[0])
    cdef double[::1] o = out
    for i in range(z.shape[0]):
        o[i] = compute_zidx(z[i], h[i], zeta[i], hc, NZ, Carr, C_offset)
    return out
