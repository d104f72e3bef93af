"""Train the surrogate on the GPU and write its parameters and held-out report.

    python -m black_simulation_amd.train_surrogate --out PATH [--net wide|fast] [--n-train N] [--epochs E] ...

Writes PATH.f32 (kt_surrogate_param_count() little-endian float32 values, the layout
kt_surrogate_load takes) and PATH.json (the options and the kt_surrogate_report).  The
default options are the full training set: 2^24 rays labelled by the mixed-precision
integrator, 2^20 held-out rays judged by the FP64 integrator, Adam lr 1e-3 cosine,
batch 65536, 192 epochs, sky regression weighted 30x.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

from . import api
from . import _lib as K

DATA = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data")
DEFAULT_WEIGHTS = os.path.join(DATA, "surrogate")        # the wide network (5-256x4-8, SiLU)
FAST_WEIGHTS = os.path.join(DATA, "surrogate_fast")      # the fast one (5-128x4-8, hard SiLU)


def load_params(path: str = DEFAULT_WEIGHTS) -> np.ndarray:
    """Parameters written by this script (PATH.f32); raises if absent or if the count is
    not one of the networks' (the count names the network)."""
    p = np.fromfile(path + ".f32", dtype="<f4")
    sizes = {K.lib().kt_surrogate_net_param_count(n) for n in (K.NET_WIDE, K.NET_FAST)}
    if p.size not in sizes:
        raise ValueError(f"{path}.f32 holds {p.size} values; the networks have {sorted(sizes)}")
    return p


def load_into(tracer, path: str = DEFAULT_WEIGHTS):
    """Loads PATH.f32 into `tracer` and sets the network's domain to the one PATH.json
    records (the scene it was trained on); a network without that record keeps the
    default training's domain."""
    tracer.surrogate_load(load_params(path))
    meta = path + ".json"
    if os.path.exists(meta):
        with open(meta) as f:
            rec = json.load(f)
        dom = rec.get("domain")
        if dom is None:  # written before the domain was recorded: the scene from its options
            o = rec.get("opts", {})
            dom = {k: o[k] for k in ("disk_outer", "escape_radius", "capture_factor") if k in o}
        tracer.surrogate_set_domain(**dom)


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", required=True)
    ap.add_argument("--n-train", type=int, default=0)
    ap.add_argument("--n-holdout", type=int, default=0)
    ap.add_argument("--epochs", type=int, default=0)
    ap.add_argument("--batch", type=int, default=0)
    ap.add_argument("--lr", type=float, default=0.0)
    ap.add_argument("--w-sky", type=float, default=0.0)
    ap.add_argument("--w-disk", type=float, default=0.0)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--init-seed", type=int, default=1)
    ap.add_argument("--net", choices=["wide", "fast"], default="wide")
    args = ap.parse_args(argv)
    kw = {"verbose": 1, "net": K.NET_FAST if args.net == "fast" else K.NET_WIDE}
    for k in ("n_train", "n_holdout", "epochs", "batch", "lr", "seed", "w_sky", "w_disk"):
        v = getattr(args, k)
        if v:
            kw[k] = v
    tr = api.Tracer(0)
    tr.surrogate_set_net(kw["net"])
    tr.surrogate_init(args.init_seed)
    t0 = time.perf_counter()
    rep = tr.surrogate_train(**kw)
    rep["wall_seconds"] = time.perf_counter() - t0
    params = tr.surrogate_save()
    domain = tr.surrogate_domain()
    tr.close()
    os.makedirs(os.path.dirname(os.path.abspath(args.out)), exist_ok=True)
    params.astype("<f4").tofile(args.out + ".f32")
    o = K.KtSurrogateOpts()
    K.lib().kt_surrogate_opts_default(o)
    for k, v in kw.items():
        setattr(o, k, v)
    with open(args.out + ".json", "w") as f:
        json.dump({"opts": {n: getattr(o, n) for n, _ in K.KtSurrogateOpts._fields_},
                   "init_seed": args.init_seed, "domain": domain, "report": rep}, f, indent=1)
    print(json.dumps(rep), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
