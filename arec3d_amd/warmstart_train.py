"""Train the SIRT warm-start network on pairs the method itself generates.

    python -m arec3d_amd.warmstart_train --samples 2000 --epochs 3 --out unet_sq128.npz

Pairs: seeded ellipsoid phantoms at 128^3 (support radius 62.5, 40
ellipsoids, semi-axes U[2, 30], density U[0.05, 1]), 360 angles on [0, pi),
SQ projector, Gaussian noise (--noise: a fixed level or a range drawn per
stack of 16 phantoms); input = SIRT after 5 iterations, target =
SIRT after 200 iterations (lam = 0.9 / sigma_max^2).  Training: Adam on the
mean squared error, batch of one cube, cosine learning-rate decay, all on
the GPU through the U-Net kernels.  A held-out set (other seeds) reports the
network's error beside the identity's (the 5-iteration input itself).
"""
import argparse
import json
import sys
import time

import numpy as np

from . import Context
from .unet import UNet, make_pairs, train


def held_out(ctx, params, scale, xin, xout):
    """Mean squared error to the target of the identity and of the network."""
    n = xin.shape[1]
    net = UNet(ctx, n, n, n, training=False, params=params, scale=scale)
    dx, dy = ctx.empty(xin.shape[1:]), ctx.empty(xin.shape[1:])
    e_id, e_net = [], []
    for a, b in zip(xin, xout):
        dx.upload(a)
        net.forward(dx, dy)
        y = dy.download()
        e_id.append(float(np.mean((a - b) ** 2)))
        e_net.append(float(np.mean((y - b) ** 2)))
    net.close()
    return float(np.mean(e_id)), float(np.mean(e_net))


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--samples", type=int, default=2000)
    ap.add_argument("--heldout", type=int, default=32)
    ap.add_argument("--epochs", type=int, default=3)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--n", type=int, default=128)
    ap.add_argument("--angles", type=int, default=360)
    ap.add_argument("--k-in", type=int, default=5)
    ap.add_argument("--k-out", type=int, default=200)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--noise", type=float, nargs="+", default=[0.01],
                    help="Gaussian noise sigma relative to the sinogram maximum, or a range lo hi "
                         "drawn log-uniformly per stack of phantoms")
    ap.add_argument("--out", required=True)
    args = ap.parse_args(argv)

    def log(*a):
        print(*a, flush=True)

    ctx = Context(0)
    noise = args.noise[0] if len(args.noise) == 1 else (args.noise[0], args.noise[1])
    t0 = time.time()
    xin, xout, lam = make_pairs(ctx, args.samples, n=args.n, na=args.angles, k_in=args.k_in,
                                k_out=args.k_out, seed0=40000, noise=noise, log=log)
    t_pairs = time.time() - t0
    hin, hout, _ = make_pairs(ctx, args.heldout, n=args.n, na=args.angles, k_in=args.k_in,
                              k_out=args.k_out, seed0=48000, noise=noise, log=None)
    log(f"pairs: {args.samples} in {t_pairs:.1f}s, lam {lam:.4e}")
    t0 = time.time()
    params, scale, losses = train(ctx, xin, xout, epochs=args.epochs, lr=args.lr, seed=args.seed, log=log)
    t_train = time.time() - t0
    e_id, e_net = held_out(ctx, params, scale, hin, hout)
    log(f"trained in {t_train:.1f}s; held-out MSE identity {e_id:.4e}, network {e_net:.4e} "
        f"({e_id / max(e_net, 1e-30):.2f}x lower)")
    meta = {
        "samples": args.samples, "epochs": args.epochs, "lr": args.lr, "n": args.n, "angles": args.angles,
        "k_in": args.k_in, "k_out": args.k_out, "lam": lam, "projector": "sq", "noise": args.noise,
        "pairs_s": round(t_pairs, 1), "train_s": round(t_train, 1),
        "heldout_mse_identity": e_id, "heldout_mse_network": e_net,
    }
    curve = np.array([np.mean(losses[i:i + 100]) for i in range(0, len(losses), 100)], np.float32)
    np.savez(args.out, params=params, scale=np.float32(scale), loss_curve=curve, meta=json.dumps(meta))
    log(json.dumps(meta))
    return 0


if __name__ == "__main__":
    sys.exit(main())
