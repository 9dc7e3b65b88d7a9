"""Train, validate and save a surrogate model for one set of march parameters, on the GPU.

    python -m blackholerender_amd.train_surrogate --max-steps 450 --disk 6 20 \
        --rays 16777216 --steps 20000 --out blackholerender_amd/models/lawS_450_disk6-20.bhsmlp

The training set is generated by the FP64 integrator itself (bhs_mlp_make_dataset), training and
validation run in the library's kernels; this script only sequences the calls and prints progress.
"""
from __future__ import annotations

import argparse
import json
import math
import time

import numpy as np


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--max-steps", type=int, default=450, help="quality preset (200/300/450/700)")
    ap.add_argument("--disk", type=float, nargs=2, default=(6.0, 20.0), metavar=("INNER", "OUTER"))
    ap.add_argument("--rays", type=int, default=1 << 24, help="training rays")
    ap.add_argument("--steps", type=int, default=20000, help="Adam steps")
    ap.add_argument("--batch", type=int, default=8192)
    ap.add_argument("--lr", type=float, nargs=2, default=(2e-3, 1e-5), metavar=("START", "END"))
    ap.add_argument("--validate", type=int, default=1 << 22, help="validation rays")
    ap.add_argument("--tol", type=float, default=1e-4, help="escape-direction tolerance (rad)")
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--out", default=None, help="model file to write")
    ap.add_argument("--chunk", type=int, default=1000, help="steps per progress line")
    ap.add_argument("--revalidate", default=None, metavar="MODEL",
                    help="no training: load MODEL, rebuild its error map and margins by the "
                         "validation pass (after a change of the decode), save to --out")
    ap.add_argument("--restep", action="store_true",
                    help="with --revalidate: also rebuild the step table from --rays fresh "
                         "integrator rays (after a change of the sampled domain)")
    a = ap.parse_args(argv)

    from . import Tracer, march_params
    from .surrogate import Surrogate

    t0 = time.time()
    tr = Tracer(0)
    if a.revalidate:
        sur = Surrogate.load(tr, a.revalidate)
        if a.restep:
            sur.make_dataset(a.rays, seed=a.seed)
        v = sur.validate(a.validate, seed=a.seed + 1000, tol_phi=a.tol)
        print("validation", json.dumps(v), flush=True)
        if a.out:
            sur.save(a.out)
            print(f"saved {a.out}", flush=True)
        return v
    params = march_params(max_steps=a.max_steps, disk_inner=a.disk[0], disk_outer=a.disk[1])
    sur = Surrogate(tr, params, seed=a.seed)
    sur.make_dataset(a.rays, seed=a.seed)
    print(f"dataset: {a.rays} rays marched in FP64 ({time.time() - t0:.1f} s)", flush=True)
    done = 0
    hist = []
    lr0, lr1 = a.lr
    def lr_at(t):
        return lr1 + 0.5 * (lr0 - lr1) * (1.0 + math.cos(math.pi * t / max(a.steps, 1)))
    while done < a.steps:
        n = min(a.chunk, a.steps - done)
        loss = sur.train(n, a.batch, lr_at(done), lr_at(done + n))
        done += n
        hist.append(float(np.mean(loss[-min(100, n):])))
        print(f"step {done:6d}  loss {hist[-1]:.4e}  lr {lr_at(done):.2e}  "
              f"({time.time() - t0:.1f} s)", flush=True)
    v = sur.validate(a.validate, seed=a.seed + 1000, tol_phi=a.tol)
    print("validation", json.dumps(v), flush=True)
    info = sur.info()
    print("info", json.dumps({k: info[k] for k in ("served_cells", "radius_margin", "steps_max",
                                                  "adam_steps")}), flush=True)
    if a.out:
        sur.save(a.out)
        print(f"saved {a.out}", flush=True)
    return v


if __name__ == "__main__":
    main()
