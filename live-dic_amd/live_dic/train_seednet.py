"""Train the learned seed network on pairs the engine renders itself.

    python -m live_dic.train_seednet --pairs 50000 --out live-dic_amd/live_dic/weights/seednet.npz

Each step renders ``batch`` crops of a 3840x2160 speckle frame and their
random affine deformations (translation U[-64, 64] px, rotation U[-10, 10] deg,
normal strains U[-3 %, 3 %] about the frame centre, noise 2 grey levels) with
``ldic_render_speckle_batch``, reduces them twice with the integer pyramid and
takes one Adam step on the supervised flow loss.  Progress lines go to stdout.
"""

from __future__ import annotations

import argparse
import sys
import time

from . import seednet


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--pairs", type=int, default=50_000)
    ap.add_argument("--batch", type=int, default=8)
    ap.add_argument("--crop", type=int, default=1024)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--init", default=None, help="start from these weights")
    ap.add_argument("--out", default=str(seednet.WEIGHTS_PATH))
    args = ap.parse_args(argv)
    net = seednet.SeedNet(seed=args.seed)
    if args.init:
        net.load(args.init)
    print(f"parameters {net.arch.parameter_count()}, pairs {args.pairs}, batch {args.batch}, crop {args.crop}",
          flush=True)
    started = time.perf_counter()

    def log(msg):
        print(f"[{time.perf_counter() - started:7.1f}s] {msg}", flush=True)

    history = seednet.train(net, pairs=args.pairs, batch=args.batch, crop=args.crop, lr=args.lr, log=log,
                            seed=args.seed)
    net.save(args.out)
    log(f"saved {args.out}; first/last loss {history[0][1]:.4f} / {history[-1][1]:.4f}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
