"""Autotuner driver for the persistent form.

Runs the C-ABI autotuner (coordinate search over workgroup size, workgroups per CU,
top-level BVH nodes staged in LDS, refill granularity, the traversal stack form, the BVH
form with or without radial culling and the CSG evaluation stack depth, timed with HIP
events) on a
workload and stores the winning settings in tuning_gfx950.json, which bench.py and the
serving loop apply.

python -m oxr.autotune csi_calorimeter instanced_csg
"""
from __future__ import annotations

import json
import os
import sys

from . import abi, scenes
from .engine import Engine

HERE = os.path.dirname(os.path.abspath(__file__))
TABLE = os.path.join(HERE, 'tuning_gfx950.json')


def tune(name, repeat=2, scale=1.0, out=TABLE):
    s = scenes.SCENES[name]()
    table = {}
    if os.path.exists(out):
        with open(out) as f:
            table = json.load(f)
    with Engine.from_scene(s) as e:
        if name in table:
            e.set_tuning(**table[name])   # the search starts from the recorded best
        e.set_gensteps(s.gensteps(scale_photons=scale))
        res, best = e.autotune(repeat=repeat)
    table[name] = best
    table.setdefault('_trials', {})[name] = [dict(block_size=int(r[0]), blocks_per_cu=int(r[1]),
                                                  lds_top_nodes=int(r[2]), refill_batch=int(r[3]),
                                                  stack_mode=int(r[4]), bvh_mode=int(r[5]),
                                                  csg_stack=int(r[6]), bvh_radial=int(r[7]),
                                                  photon_order=int(r[8]), generate_inplace=int(r[9]),
                                                  kernel_ms=round(float(r[10]), 4))
                                             for r in res]
    with open(out, 'w') as f:
        json.dump(table, f, indent=1, sort_keys=True)
    return best, res


if __name__ == '__main__':
    for name in sys.argv[1:] or ['csi_calorimeter']:
        best, res = tune(name)
        print(name, 'best', best, 'of', len(res), 'trials', flush=True)
