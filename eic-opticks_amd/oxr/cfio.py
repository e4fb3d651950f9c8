"""CSGFoundry folders in the reference's on-disk layout.

save_foundry writes a scene's arrays the way CSGFoundry::save_ (CSG/CSGFoundry.cc:2768-2800)
and SSim::save lay them out; load_foundry reads such a folder through the library's own
reader (oxr_foundry_load), the path a reference geometry takes into the engine:

    <cfbase>/CSGFoundry/{solid,prim,node,plan,tran,itra,inst}.npy
    <cfbase>/CSGFoundry/SSim/stree/standard/{bnd,optical,icdf}.npy, bnd_meta.txt
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import abi


def save_foundry(cfbase, geometry, physics):
    """Write geometry (solid, prim, node, plan, tran, itra, inst) and physics (bnd,
    optical, icdf, wl_min, wl_step) arrays under cfbase/CSGFoundry."""
    d = os.path.join(cfbase, 'CSGFoundry')
    s = os.path.join(d, 'SSim', 'stree', 'standard')
    os.makedirs(s, exist_ok=True)
    kinds = dict(solid=np.int32, prim=np.float32, node=np.float32, plan=np.float32, tran=np.float32,
                 itra=np.float32, inst=np.float32)
    for k, dt in kinds.items():
        a = geometry.get(k)
        if a is not None and len(a):
            a = np.ascontiguousarray(a, dtype=dt)
            if k == 'plan':
                a = a.reshape(len(a), 1, 4)          # NP::Write(dir, "plan.npy", ..., 1, 4)
            np.save(os.path.join(d, k + '.npy'), a)
    bnd = np.ascontiguousarray(physics['bnd'], dtype=np.float32)
    np.save(os.path.join(s, 'bnd.npy'), bnd)
    np.save(os.path.join(s, 'optical.npy'), np.ascontiguousarray(physics['optical'], dtype=np.uint32))
    lo, step = float(physics.get('wl_min', 60.0)), float(physics.get('wl_step', 1.0))
    hi = lo + step * (bnd.shape[3] - 1)
    with open(os.path.join(s, 'bnd_meta.txt'), 'w') as f:
        f.write(f'domain_low:{lo}\ndomain_high:{hi}\ndomain_step:{step}\ndomain_range:{hi - lo}\n')
    icdf = physics.get('icdf')
    if icdf is not None:
        icdf = np.asarray(icdf, dtype=np.float32)
        np.save(os.path.join(s, 'icdf.npy'), icdf.reshape(icdf.shape[0], -1, 1))
    return d


class LoadedFoundry:
    """A CSGFoundry folder read by the library; geometry/physics are numpy copies."""

    def __init__(self, cfbase):
        from .engine import lib, _check
        h = C.c_void_p()
        _check(lib().oxr_foundry_load(os.fsencode(cfbase), C.byref(h)))
        try:
            g, p = abi.Geometry(), abi.Physics()
            _check(lib().oxr_foundry_arrays(h, C.byref(g), C.byref(p)))

            def grab(ptr, n, shape, dt):
                if not ptr or n == 0:
                    return np.zeros((0,) + shape, dtype=dt)
                ct = C.c_int32 if dt == np.int32 else (C.c_uint32 if dt == np.uint32 else C.c_float)
                count = int(n * np.prod(shape))
                return np.ctypeslib.as_array(C.cast(ptr, C.POINTER(ct)), shape=(count,)).copy().view(dt).reshape((n,) + shape)
            self.geometry = dict(solid=grab(g.solid, g.num_solid, (3, 4), np.int32),
                                 prim=grab(g.prim, g.num_prim, (4, 4), np.float32),
                                 node=grab(g.node, g.num_node, (4, 4), np.float32),
                                 plan=grab(g.plan, g.num_plan, (4,), np.float32),
                                 itra=grab(g.itra, g.num_itra, (4, 4), np.float32),
                                 inst=grab(g.inst, g.num_inst, (4, 4), np.float32))
            phys = dict(bnd=grab(p.bnd, p.num_bnd, (4, 2, p.num_wl, 4), np.float32),
                        optical=grab(p.optical, 4 * p.num_bnd, (4,), np.uint32),
                        wl_min=p.wl_min, wl_step=p.wl_step, hd_factor=p.hd_factor)
            if p.icdf:
                phys['icdf'] = grab(p.icdf, p.icdf_rows, (p.icdf_width,), np.float32)
            self.physics = phys
        finally:
            lib().oxr_foundry_free(h)


def load_foundry(cfbase):
    f = LoadedFoundry(cfbase)
    return f.geometry, f.physics
