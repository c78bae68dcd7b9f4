#!/usr/bin/env python3
"""Writes the golden fixtures of tests/golden/ from independent numpy
statements of the reference formulas (no oracle, no HIP code involved).

The reference ships no golden vectors for this path (its test baseline is an
LFS pointer), so each fixture is computed here from the mathematics the
reference states, vectorised differently from both the CPU oracle and the
GPU kernels:

* ``decode.npz``   -- every int8 and int16 compressed coordinate through the
  reference's own formula ``(int(xp+ishift, izipx) + rshift) * x_resolution``
  (parameters.f90:91-92, 131-133; particle_mesh.f90:36) with modular
  conversion to the izipx-byte integer.
* ``green.npz``    -- Green_3D (Green/Green.f90:22-71) in FP64 for the three
  mesh kinds of a 64^3 run: coarse (n=16, soft 3.5, range 0, gridsize 4),
  tile (n=48, soft 3.5, range 14, gridsize 1) and sub-tile at ratio_sf=2
  (n=24, soft 3.5, range 7, gridsize 1/2).
* ``f_ra.npz``     -- the PP force law F_ra (basic_functions.f08:143-157)
  in FP64 at 257 radii for the PP softening 0.06 and the PP range 1.75.

Run from the repository root:  python tests/golden/make_golden.py
"""
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def decode_ref(xp, izipx):
    B = 8 * izipx
    ishift = -(2 ** (B - 1))
    rshift = 0.5 - ishift
    t = xp.astype(np.int64) + ishift
    t = (t + 2 ** (B - 1)) % (2 ** B) - 2 ** (B - 1)  # int(., izipx): two's-complement wrap
    return (t + rshift) * (1.0 / 2 ** B)


def green_ref(n, apm_soft, apm_range, gridsize, n_int=2, p=2):
    apm_soft = np.float32(apm_soft)
    apm_range = np.float32(apm_range)
    gridsize = np.float32(gridsize)
    alpha = np.float64(np.float32(4.0) / np.float32(3.0))
    c215 = np.float64(np.float32(2.0) / np.float32(15.0))
    inv_n = np.float64(np.float32(1.0) / np.float32(n))
    kk = np.arange(1, n + 1)
    fftspace = 2 * np.pi * inv_n * (((kk + n // 2 - 1) % n) - n // 2)
    k1 = fftspace[: n // 2 + 1][None, None, :]
    k2 = fftspace[None, :, None]
    k3 = fftspace[:, None, None]
    shape = (n, n, n // 2 + 1)
    kv = [np.broadcast_to(k1, shape), np.broadcast_to(k2, shape), np.broadcast_to(k3, shape)]
    D = [alpha * np.sin(k) + (1 - alpha) * np.sin(2 * k) / 2 for k in kv]
    ur = [np.zeros(shape) for _ in range(3)]
    rng = range(-n_int, n_int + 1)
    with np.errstate(divide="ignore", invalid="ignore"):
        for i3 in rng:
            for i2 in rng:
                for i1 in rng:
                    kn = [kv[0] + 2 * np.pi * i1, kv[1] + 2 * np.pi * i2, kv[2] + 2 * np.pi * i3]
                    kmag = np.sqrt(kn[0] ** 2 + kn[1] ** 2 + kn[2] ** 2)
                    U = np.ones(shape)
                    for k in kn:
                        U = U * np.where(k == 0, 1.0, np.sin(k / 2) / (k / 2))
                    U = U ** (2 * p + 2)
                    y = np.float64(apm_soft) * kmag / 2
                    Ss = np.where(kmag == 0, 1.0, 12 * y ** -4.0 * (2 - 2 * np.cos(y) - y * np.sin(y)))
                    if apm_range == 0:
                        Sr = np.zeros(shape)
                    else:
                        y = np.float64(apm_range) * kmag / 2
                        Sr = np.where(kmag == 0, 1.0, 12 * y ** -4.0 * (2 - 2 * np.cos(y) - y * np.sin(y)))
                    for d in range(3):
                        ur[d] += np.where(kmag == 0, 0.0, U * kn[d] * (Ss ** 2 - Sr ** 2) / kmag ** 2)
        s = [np.sin(k / 2) for k in kv]
        u2n = np.prod([1 - si ** 2 + c215 * si ** 4 for si in s], axis=0)
        G = 4 * np.pi * (D[0] * ur[0] + D[1] * ur[1] + D[2] * ur[2]) / (D[0] ** 2 + D[1] ** 2 + D[2] ** 2)
        G = G / u2n ** 2 / np.float64(gridsize * gridsize)
    for a in (0, n // 2):
        for b in (0, n // 2):
            for c in (0, n // 2):
                G[c, b, a] = 0
    return G.astype(np.float32)


def f_ra_ref(r, apm):
    r = np.asarray(r, np.float64)
    ep = 2 * r / apm
    c = 1.0 / 35.0 / apm ** 2
    out = np.zeros_like(r)
    with np.errstate(divide="ignore"):
        far = r ** -2.0
    mid = c * (12 * ep ** -2.0 - 224 + 896 * ep - 840 * ep ** 2 + 224 * ep ** 3 + 70 * ep ** 4 - 48 * ep ** 5 + 7 * ep ** 6)
    near = c * (224 * ep - 224 * ep ** 3 + 70 * ep ** 4 + 48 * ep ** 5 - 21 * ep ** 6)
    out = np.where(ep > 2, far, np.where(ep > 1, mid, np.where(ep > 0, near, 0.0)))
    return out


def main():
    x8 = np.arange(-128, 128, dtype=np.int64).astype(np.int8)
    x16 = np.arange(-32768, 32768, dtype=np.int64).astype(np.int16)
    np.savez_compressed(os.path.join(HERE, "decode.npz"), xp8=x8, frac8=decode_ref(x8, 1), xp16=x16,
                        frac16=decode_ref(x16, 2))
    cases = {"coarse16": (16, 3.5, 0.0, 4.0), "tile48": (48, 3.5, 14.0, 1.0), "subtile24": (24, 3.5, 7.0, 0.5)}
    out = {}
    for name, args in cases.items():
        out[name] = green_ref(*args)
        out[name + "_args"] = np.array(args, np.float64)
    np.savez_compressed(os.path.join(HERE, "green.npz"), **out)
    r = np.linspace(0.0, 2.0, 257)[1:].astype(np.float32)
    np.savez_compressed(os.path.join(HERE, "f_ra.npz"), r=r, app=f_ra_ref(r, np.float32(0.06)),
                        range175=f_ra_ref(r, np.float32(1.75)))
    for f in ("decode.npz", "green.npz", "f_ra.npz"):
        print(f, os.path.getsize(os.path.join(HERE, f)))


if __name__ == "__main__":
    main()
