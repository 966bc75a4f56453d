"""Train the learned preconditioner on the GPU and save its weights.

    python -m parallelpoissonsolver_amd.train_nn --out weights.npy

Training set (pps_nn_generate): plain FP64 CG on an n^3 Dirichlet unit cube;
per right-hand side the converged u* and `snapshots` CG iterates x_k give the
pairs (r_k / rms(r_k), energy norm of u* - x_k).  Right-hand sides follow a
k^-2 power-law spectrum, a `white_fraction` of them white noise.  Training
(pps_nn_train): Adam on 1 - cos^2(y, e) in the energy norm, cosine schedule.
The weights are saved with numpy (no pickle) together with the recipe.
"""
import argparse
import json
import time

import numpy as np

from . import _lib as L

DEFAULT = dict(n=64, nrhs=256, snapshots=16, kmax=32, seed=1, white_fraction=0.5, batch=16, steps=2000, lr=2e-3,
               lr_final=5e-5, init_seed=7)


def train(**kw):
    cfg = dict(DEFAULT, **kw)
    net = L.Network(seed=cfg["init_seed"])
    t0 = time.perf_counter()
    pairs = net.generate(n=cfg["n"], nrhs=cfg["nrhs"], snapshots=cfg["snapshots"], kmax=cfg["kmax"], seed=cfg["seed"],
                         white_fraction=cfg["white_fraction"])
    t_gen = time.perf_counter() - t0
    st = net.train(n=cfg["n"], batch=cfg["batch"], steps=cfg["steps"], lr=cfg["lr"], lr_final=cfg["lr_final"],
                   seed=cfg["seed"] + 1)
    info = dict(cfg, pairs=pairs, generate_seconds=t_gen, train_seconds=st["seconds"], loss_first=st["loss_first"],
                loss_last=st["loss_last"], params_effective=net.n_effective)
    return net, info


def save(net, info, path):
    np.save(path, net.get_params(), allow_pickle=False)
    with open(path + ".json", "w") as f:
        json.dump(info, f, indent=1)


def load(net, path):
    p = np.load(path, allow_pickle=False)
    net.set_params(p)
    try:
        with open(path + ".json") as f:
            return json.load(f)
    except OSError:
        return {}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", required=True)
    for k, v in DEFAULT.items():
        ap.add_argument("--" + k.replace("_", "-"), type=type(v), default=v)
    a = vars(ap.parse_args())
    out = a.pop("out")
    net, info = train(**a)
    save(net, info, out)
    print(json.dumps(info), flush=True)


if __name__ == "__main__":
    main()
