"""quad6 genstep builders (SEvt genstep arrays, shape (n,6,4) float32).

Layouts follow the reference structs, field for field:

* torch       sysrap/storch.h:43-69
* Cerenkov    sysrap/scerenkov.h:29-58
* scintillation sysrap/sscint.h:34-70

gentype values are OpticksGenstep_* (sysrap/OpticksGenstep.h:21-44).
"""
from __future__ import annotations

import numpy as np

GS_CERENKOV_1042 = 1
GS_SCINT_1042 = 2
GS_TORCH = 6
GS_CERENKOV = 15
GS_SCINTILLATION = 16
GS_INPUT_PHOTON = 19

# storch types (sysrap/storchtype.h:10-17); T_POINT_ISOTROPIC is an extension of
# this engine: isotropic directions from gs.pos with a random linear polarisation
# perpendicular to the direction (three uniforms: phi, cos(theta), pol phase).
T_UNDEF, T_DISC, T_LINE, T_POINT, T_CIRCLE, T_RECTANGLE, T_SPHERE_MARSAGLIA, T_SPHERE = range(8)
T_POINT_ISOTROPIC = 8


def _blank(n):
    gs = np.zeros((n, 6, 4), dtype=np.float32)
    return gs, gs.view(np.uint32), gs.view(np.int32)


def torch(numphoton, pos=(0, 0, 0), mom=(0, 0, 1), pol=(1, 0, 0), wavelength=420.0,
          type=T_POINT_ISOTROPIC, radius=0.0, distance=0.0, zenith=(0.0, 1.0),
          azimuth=(0.0, 1.0), time=0.0, matline=0, trackid=0, weight=1.0, mode=255):
    """One torch genstep per entry of numphoton (scalar or array); pos may be (n,3)."""
    numphoton = np.atleast_1d(np.asarray(numphoton, dtype=np.int64))
    n = len(numphoton)
    gs, u, _ = _blank(n)
    u[:, 0, 0] = GS_TORCH
    u[:, 0, 1] = trackid
    u[:, 0, 2] = matline
    u[:, 0, 3] = numphoton
    gs[:, 1, :3] = np.broadcast_to(np.asarray(pos, dtype=np.float32), (n, 3))
    gs[:, 1, 3] = time
    gs[:, 2, :3] = np.broadcast_to(np.asarray(mom, dtype=np.float32), (n, 3))
    gs[:, 2, 3] = weight
    gs[:, 3, :3] = np.broadcast_to(np.asarray(pol, dtype=np.float32), (n, 3))
    gs[:, 3, 3] = wavelength
    gs[:, 4, 0:2] = zenith
    gs[:, 4, 2:4] = azimuth
    gs[:, 5, 0] = radius
    gs[:, 5, 1] = distance
    u[:, 5, 2] = mode
    u[:, 5, 3] = type
    return gs


def cerenkov(numphoton, pos, delta_pos, step_length, beta, matline, wmin, wmax, nmax,
             mean_photons, time=0.0, charge=-1.0):
    """Cerenkov gensteps as U4::CollectGenstep_G4Cerenkov_modified (u4/U4.cc:159)."""
    numphoton = np.atleast_1d(np.asarray(numphoton, dtype=np.int64))
    n = len(numphoton)
    gs, u, i = _blank(n)
    beta = np.broadcast_to(np.asarray(beta, dtype=np.float64), (n,))
    bi = 1.0 / beta
    maxcos = bi / nmax
    maxsin2 = (1.0 - maxcos) * (1.0 + maxcos)
    v = beta * 299.792458
    u[:, 0, 0] = GS_CERENKOV
    u[:, 0, 2] = matline
    u[:, 0, 3] = numphoton
    gs[:, 1, :3] = pos
    gs[:, 1, 3] = time
    gs[:, 2, :3] = delta_pos
    gs[:, 2, 3] = step_length
    i[:, 3, 0] = 11
    gs[:, 3, 1] = charge
    gs[:, 3, 2] = 1.0
    gs[:, 3, 3] = v
    gs[:, 4, 0] = bi
    gs[:, 4, 1] = wmin
    gs[:, 4, 2] = wmax
    gs[:, 4, 3] = maxcos
    gs[:, 5, 0] = maxsin2
    gs[:, 5, 1] = mean_photons
    gs[:, 5, 2] = mean_photons
    gs[:, 5, 3] = v
    return gs


def scintillation(numphoton, pos, delta_pos, step_length, mean_velocity, scint_time,
                  matline=0, time=0.0, charge=-1.0):
    """Scintillation gensteps as U4::CollectGenstep_DsG4Scintillation_r4695 (u4/U4.cc:278)."""
    numphoton = np.atleast_1d(np.asarray(numphoton, dtype=np.int64))
    n = len(numphoton)
    gs, u, i = _blank(n)
    u[:, 0, 0] = GS_SCINTILLATION
    u[:, 0, 2] = matline
    u[:, 0, 3] = numphoton
    gs[:, 1, :3] = pos
    gs[:, 1, 3] = time
    gs[:, 2, :3] = delta_pos
    gs[:, 2, 3] = step_length
    i[:, 3, 0] = 11
    gs[:, 3, 1] = charge
    gs[:, 3, 2] = 1.0
    gs[:, 3, 3] = mean_velocity
    gs[:, 5, 0] = scint_time
    return gs


def input_photon(numphoton):
    gs, u, _ = _blank(1)
    u[0, 0, 0] = GS_INPUT_PHOTON
    u[0, 0, 3] = numphoton
    return gs


def num_photons(gs):
    return int(np.asarray(gs).view(np.uint32)[:, 0, 3].astype(np.int64).sum())
