"""Photon and genstep sources of the reference's example applications.

* load_photons_txt / write_photons_txt: the photon text file of GPUPhotonFileSource
  (src/GPUPhotonFileSource.h:47-88): one photon per line, 11 numbers
  "pos_x pos_y pos_z time mom_x mom_y mom_z pol_x pol_y pol_z wavelength"; lines that
  are empty or start with '#' are skipped, malformed lines are skipped with a warning.
* torch_config: the JSON torch configuration of gphox::Config::ReadConfig
  (src/config.cpp:110-146) turned into one quad6 torch genstep (sysrap/storch.h:43-69);
  mom is normalised, type names are storchtype's ("disc", "sphere", ...).
* write_hits_txt: the hit listing GPUPhotonFileSource writes after a launch
  (src/GPUPhotonFileSource.h:176-194).
"""
from __future__ import annotations

import json
import sys
from dataclasses import dataclass
from typing import List, Optional, Tuple

import numpy as np

from . import gensteps as G

# storchtype names (sysrap/storchtype.h:28-35)
TORCH_TYPES = {'undef': G.T_UNDEF, 'disc': G.T_DISC, 'line': G.T_LINE, 'point': G.T_POINT,
               'circle': G.T_CIRCLE, 'rectangle': G.T_RECTANGLE,
               'sphere_marsaglia': G.T_SPHERE_MARSAGLIA, 'sphere': G.T_SPHERE}


def load_photons_txt(path, warn=True) -> np.ndarray:
    """Photons of a text file as sphoton (n,4,4) float32, every other field zero."""
    rows: List[List[float]] = []
    with open(path) as fh:
        for lineno, line in enumerate(fh, 1):
            line = line.rstrip('\n')
            if not line or line[0] == '#':
                continue
            parts = line.split()
            vals = None
            if len(parts) >= 11:
                try:
                    vals = [float(np.float32(x)) for x in parts[:11]]
                except ValueError:
                    vals = None
            if vals is None:
                if warn:
                    print(f'WARNING: skipping malformed line {lineno}: {line}', file=sys.stderr)
                continue
            rows.append(vals)
    p = np.zeros((len(rows), 4, 4), dtype=np.float32)
    if rows:
        a = np.asarray(rows, dtype=np.float32)
        p[:, 0, :3] = a[:, 0:3]
        p[:, 0, 3] = a[:, 3]
        p[:, 1, :3] = a[:, 4:7]
        p[:, 2, :3] = a[:, 7:10]
        p[:, 2, 3] = a[:, 10]
    return p


def write_photons_txt(path, photons, header=True):
    p = np.asarray(photons, dtype=np.float32).reshape(-1, 4, 4)
    with open(path, 'w') as fh:
        if header:
            fh.write('# pos_x pos_y pos_z time mom_x mom_y mom_z pol_x pol_y pol_z wavelength\n')
        for q in p:
            vals = list(q[0, :3]) + [q[0, 3]] + list(q[1, :3]) + list(q[2, :3]) + [q[2, 3]]
            fh.write(' '.join(repr(float(v)) for v in vals) + '\n')


def _g(v):
    """Shortest round-trip decimal of a float32, like an ostream of a float at its
    default precision for the values these files hold."""
    f = float(np.float32(v))
    s = f'{f:.6g}'
    return s


def write_hits_txt(path, hits):
    """'time wavelength  (x, y, z)  (mx, my, mz)  (px, py, pz)' per hit."""
    h = np.asarray(hits, dtype=np.float32).reshape(-1, 4, 4)
    with open(path, 'w') as fh:
        for q in h:
            fh.write(f'{_g(q[0, 3])} {_g(q[2, 3])}  ({_g(q[0, 0])}, {_g(q[0, 1])}, {_g(q[0, 2])})  '
                     f'({_g(q[1, 0])}, {_g(q[1, 1])}, {_g(q[1, 2])})  '
                     f'({_g(q[2, 0])}, {_g(q[2, 1])}, {_g(q[2, 2])})\n')


@dataclass
class TorchConfig:
    genstep: np.ndarray            # (1,6,4) float32
    numphoton: int
    event_mode: Optional[str]
    max_slot: Optional[int]


def torch_config(path_or_dict) -> TorchConfig:
    """A gphox JSON config ({"torch": {...}, "event": {...}}) as one torch genstep."""
    if isinstance(path_or_dict, dict):
        cfg = path_or_dict
    else:
        with open(path_or_dict) as fh:
            cfg = json.load(fh)
    t = cfg['torch']
    if t.get('gentype', 'TORCH') != 'TORCH':
        raise ValueError(f"torch gentype must be TORCH, got {t.get('gentype')!r}")
    ttype = TORCH_TYPES.get(t['type'], G.T_UNDEF)
    if ttype == G.T_UNDEF:
        raise ValueError(f"unknown torch type {t['type']!r}")
    mom = np.asarray(t['mom'], dtype=np.float32)
    mom = (mom / np.float32(np.sqrt(np.float32(np.dot(mom, mom))))).astype(np.float32)
    gs = G.torch(int(t['numphoton']), pos=t['pos'], mom=mom, pol=t['pol'],
                 wavelength=float(t['wavelength']), type=ttype, radius=float(t['radius']),
                 distance=float(t['distance']), zenith=tuple(t['zenith']), azimuth=tuple(t['azimuth']),
                 time=float(t['time']), matline=int(t['matline']), trackid=int(t['trackid']),
                 weight=float(t['weight']), mode=int(t['mode']))
    ev = cfg.get('event', {})
    return TorchConfig(gs, int(t['numphoton']), ev.get('mode'),
                       None if ev.get('maxslot') is None else int(ev['maxslot']))
