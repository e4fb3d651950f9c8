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
GS_FRAME = 17
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


# grid axes of FRAME gensteps (sysrap/sxyz.h)
AX_XYZ, AX_YZ, AX_XZ, AX_XY = range(4)


def grid_axes(nx, ny, nz):
    """SGenstep::GridAxes: a planar grid (one dimension zero) names its plane."""
    if nx == 0 and ny > 0 and nz > 0:
        return AX_YZ
    if nx > 0 and ny == 0 and nz > 0:
        return AX_XZ
    if nx > 0 and ny > 0 and nz == 0:
        return AX_XY
    return AX_XYZ


def genstep_id(ix, iy, iz, iw=0):
    """SGenstep::GenstepID: four signed chars packed little-endian into 32 bits."""
    return int(np.array([ix, iy, iz, iw], dtype=np.int8).view(np.uint32)[0])


def frame(ce, cegs, gridscale=1.0, geotran=None, offsets=((0.0, 0.0, 0.0),), ce_scale=True):
    """FRAME gensteps on a centre-extent grid, SFrameGenstep::MakeCenterExtentGenstep
    (sysrap/SFrameGenstep.cc:604-700). ce = (cx, cy, cz, extent); cegs = (nx, ny, nz,
    photons) for the symmetric grid -n..n, or the standardised
    (ix0, ix1, iy0, iy1, iz0, iz1, photons, high). Each genstep carries the transform
    geotran x translate(grid point * extent * gridscale / high) as a qat4 in q2..q5;
    geotran defaults to a translation to the centre (row-convention 4x4)."""
    cegs = list(cegs)
    if len(cegs) == 4:
        nx, ny, nz, ph = cegs
        cegs = [-nx, nx, -ny, ny, -nz, nz, ph, 1]
    if len(cegs) != 8:
        raise ValueError('cegs has 4 or 8 entries')
    ix0, ix1, iy0, iy1, iz0, iz1, ph, high = cegs
    if not 1 <= high <= 8:
        raise ValueError('cegs high must be 1..8')
    scale = float(gridscale) / float(high)
    ix0, ix1, iy0, iy1, iz0, iz1 = (v * high for v in (ix0, ix1, iy0, iy1, iz0, iz1))
    axes = grid_axes((ix1 - ix0) // 2, (iy1 - iy0) // 2, (iz1 - iz0) // 2)
    local_scale = scale * float(ce[3]) if ce_scale else scale
    if geotran is None:
        geotran = np.eye(4)
        geotran[3, :3] = ce[:3]
    geotran = np.asarray(geotran, dtype=np.float64)
    out = []
    for ip, off in enumerate(offsets):
        for ix in range(ix0, ix1 + 1):
            for iy in range(iy0, iy1 + 1):
                for iz in range(iz0, iz1 + 1):
                    shift = np.eye(4)
                    shift[3, :3] = (ix * local_scale, iy * local_scale, iz * local_scale)
                    m = shift @ geotran                      # grid shift first, then geotran
                    gs = np.zeros((6, 4), dtype=np.float32)
                    u = gs.view(np.uint32)
                    u[0, 0] = GS_FRAME
                    u[0, 1] = axes
                    u[0, 2] = genstep_id(ix, iy, iz, ip)
                    u[0, 3] = ph
                    gs[1, :3] = off
                    gs[1, 3] = 1.0
                    gs[2:6] = m.astype(np.float32)
                    out.append(gs)
    return np.asarray(out, dtype=np.float32).reshape(-1, 6, 4)
