"""Shared set-up of layered simulation cases for the oracle and the GPU.

Follows the reference's set-up (AtmosphericTurbulence.c:1947-2363): Cn2
normalised to 1, super-layer binning, masters synthesised from injected
seeded noise, r0 measured on every master and all masters rescaled to the
target r0, then layer l weighted by sqrt(cn2_l / cos z); initial positions
0.05 M plus the direction offset (angle * altitude / dx), velocities
speed (cos, sin)(dir) / dx.  Masters are stored as float32 (the reference's
WFprecision = 0 storage) and handed to the oracle as the same values.
"""
import math

import numpy as np


class Case:
    pass


def build_case(oracle, N, Nout, M, alts, cn2, speeds, dirs, lambdas, lam_ref, dx, dt, r0_m, L0, l0,
               bicubic=True, seed=0xA75, bin_m=1000.0, directions=((0.0, 0.0),), unwrap_mode=None,
               zangle=0.0, scoeff=None):
    c = Case()
    L = len(alts)
    alts = np.asarray(alts, np.float64)
    cn2 = np.asarray(cn2, np.float64)
    cn2 = cn2 / cn2.sum()
    masters = []
    for l in range(L):
        _, z32 = oracle.noise_c64(seed, l, M * M, f32=True)
        o1, _, _, _ = oracle.make_master(z32.reshape(M, M).astype(np.complex128), L0 / dx, l0 / dx)
        masters.append(o1.astype(np.float32))
    dpix = min(50, M // 4)
    r0s = [oracle.r0_measure(m.astype(np.float64), dpix) for m in masters]
    r0pix_goal = r0_m / dx
    tmp1 = (np.mean(r0s) / r0pix_goal) ** (5.0 / 6.0)
    for l in range(L):
        m = (masters[l].astype(np.float64) * tmp1).astype(np.float32)
        m = (m.astype(np.float64) * math.sqrt(cn2[l] / math.cos(zangle))).astype(np.float32)
        masters[l] = m
    c.masters = np.stack(masters)
    c.slayer_alt, c.sli = oracle.superlayers(alts, cn2, bin_m)
    D = len(directions)
    c.xpos0 = np.array([[0.05 * M + dxr * alts[l] / dx for l in range(L)] for dxr, _ in directions])
    c.ypos0 = np.array([[0.05 * M + dyr * alts[l] / dx for l in range(L)] for _, dyr in directions])
    c.vx = np.array([speeds[l] * math.cos(dirs[l]) / dx for l in range(L)])
    c.vy = np.array([speeds[l] * math.sin(dirs[l]) / dx for l in range(L)])
    c.lambdas = np.asarray(lambdas, np.float64)
    W = c.lambdas.size
    if scoeff is None:
        scoeff = np.array([[lam_ref / lam] * L for lam in c.lambdas])
    c.scoeff = np.asarray(scoeff, np.float64).reshape(W, L)
    c.unwrap = np.asarray(unwrap_mode if unwrap_mode is not None else [0] * W, np.int32)
    c.N, c.Nout, c.M, c.L, c.W, c.D = N, Nout, M, L, W, D
    c.dx, c.dt, c.zangle, c.bicubic = dx, dt, zangle, bicubic
    return c


def oracle_sims(oracle, c, **kw):
    """one oracle.Sim per direction (the reference runs one direction per process)"""
    return [oracle.Sim(c.N, c.Nout, c.masters.astype(np.float64), c.slayer_alt, c.sli, c.lambdas, c.scoeff,
                       c.unwrap, c.xpos0[d], c.ypos0[d], c.vx, c.vy, c.dx, c.dt, c.zangle, c.bicubic, **kw)
            for d in range(c.D)]


def gpu_sim(ctx, c, **kw):
    from atmosphericturbulencesimul_amd.api import Simulation
    s = Simulation(ctx, c.N, c.Nout, c.M, c.slayer_alt, c.sli, c.lambdas, c.scoeff, c.unwrap, c.xpos0, c.ypos0,
                   c.vx, c.vy, c.dx, c.dt, c.zangle, c.bicubic, **kw)
    for l in range(c.L):
        s.set_master(l, c.masters[l])
    return s
