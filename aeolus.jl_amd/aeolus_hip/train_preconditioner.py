"""Train the learned preconditioner on the GPU and save its weights.

Pairs (source, solution) come from the multigrid solver itself: seeded
building layouts (cut-cell operator, Δ = 2 m), the divergence of a seeded
k^-5/3 Gaussian velocity as the source, MGPreconditionedCG (FP64 V-cycle) to
1e-10 as the solution. The solution is taken over the fluid, its mean
removed (the operator's null space), and 0 in solid cells (whose values the
equations do not fix). The network is fully convolutional, so weights
trained on one grid serve any grid with the same spacing.

    python -m aeolus_hip.train_preconditioner --shape 128 128 64 --pairs 64 \
        --epochs 20 --out weights.npy
"""
import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_WEIGHTS = os.path.join(HERE, "weights", "unet_cutcell_dx2.npy")


def make_pair(N, seed, spacing=2.0, n_buildings=16, height=(10.0, 80.0), tol=1e-10):
    """One (source, solution, mask, coefficients) sample, FP64 numpy."""
    from . import workloads as W
    from .pressure import MGPreconditionedCG, solve_interior
    w = W.cutcell_workload(N, spacing, n_buildings=n_buildings, seed=seed, height=height)
    mg = MGPreconditionedCG(N, (spacing,) * 3, coeff=w["coeff"], tol=tol, maxit=500)
    x = solve_interior(mg, w["source"])
    fluid = w["mask"] > 0
    x[~fluid] = 0.0
    x[fluid] -= x[fluid].mean()
    return w["source"], np.asfortranarray(x), w["mask"], w["coeff"]


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--shape", type=int, nargs=3, default=[128, 128, 64])
    ap.add_argument("--pairs", type=int, default=64)
    ap.add_argument("--epochs", type=int, default=20)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--buildings", type=int, default=16)
    ap.add_argument("--seed", type=int, default=7000)
    ap.add_argument("--out", default=DEFAULT_WEIGHTS)
    args = ap.parse_args(argv)

    import aeolus_hip as hip
    from aeolus_hip.fields import DeviceArray
    from aeolus_hip.learned import UNet, train

    hip.init(0)
    N = tuple(args.shape)
    t0 = time.time()
    samples = []
    for s in range(args.pairs):
        src, x, mask, _ = make_pair(N, args.seed + s, n_buildings=args.buildings)
        samples.append((DeviceArray.from_numpy(src), DeviceArray.from_numpy(x), DeviceArray.from_numpy(mask)))
        if (s + 1) % 16 == 0:
            print(f"pairs: {s + 1}/{args.pairs} ({time.time() - t0:.0f} s)", flush=True)
    t1 = time.time()
    net = UNet(N, in_channels=2, seed=args.seed)
    losses = train(net, samples, epochs=args.epochs, lr=args.lr, log=lambda m: print(m, flush=True))
    hip.sync()
    t2 = time.time()
    os.makedirs(os.path.dirname(os.path.abspath(args.out)), exist_ok=True)
    np.save(args.out, net.params(), allow_pickle=False)
    print(json.dumps({"shape": N, "pairs": args.pairs, "epochs": args.epochs, "lr": args.lr,
                      "first_loss": losses[0], "last_loss": losses[-1], "pairs_s": t1 - t0, "train_s": t2 - t1,
                      "steps": args.pairs * args.epochs, "out": args.out}), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
