"""Train the learned seed network on pairs the engine renders itself.

    python -m live_dic.train_seednet --out live-dic_amd/live_dic/weights/seednet.npz

Each step renders ``batch`` crops of a 3840x2160 speckle frame and their
random affine deformations (translation U[-64, 64] px, rotation U[-10, 10] deg,
normal strains U[-3 %, 3 %] about the frame centre, noise 2 grey levels) with
``ldic_render_speckle_batch``, reduces them twice with the integer pyramid and
takes one Adam step on the supervised flow loss.  Progress lines go to stdout.

The default schedule is the one the committed weights come from: two stages
of 2048 x 2048 crops in batches of 4, each with its own cosine learning-rate
decay and its own stream of rendered pairs (data seeds 1 and 2):
  stage 1: 50,000 pairs at lr 1e-3 from the seeded initialisation;
  stage 2: 100,000 pairs at lr 5e-4, continuing from stage 1.
``--stages`` replaces it (pairs:crop:batch:lr[,...]).
"""

from __future__ import annotations

import argparse
import sys
import time

from . import seednet

DEFAULT_STAGES = "50000:2048:4:1e-3,100000:2048:4:5e-4"


def parse_stages(text: str) -> list[tuple[int, int, int, float]]:
    stages = []
    for part in text.split(","):
        pairs, crop, batch, lr = part.split(":")
        stages.append((int(pairs), int(crop), int(batch), float(lr)))
    return stages


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--stages", default=DEFAULT_STAGES, help="pairs:crop:batch:lr for each stage, comma separated")
    ap.add_argument("--seed", type=int, default=1, help="initialisation seed; stage k renders with seed + k")
    ap.add_argument("--init", default=None, help="start from these weights")
    ap.add_argument("--out", default=str(seednet.WEIGHTS_PATH))
    args = ap.parse_args(argv)
    stages = parse_stages(args.stages)
    net = seednet.SeedNet(seed=args.seed)
    if args.init:
        net.load(args.init)
    print(f"parameters {net.arch.parameter_count()}, stages {stages}", flush=True)
    started = time.perf_counter()

    def log(msg):
        print(f"[{time.perf_counter() - started:7.1f}s] {msg}", flush=True)

    for k, (pairs, crop, batch, lr) in enumerate(stages):
        log(f"stage {k + 1}: {pairs} pairs, crop {crop}, batch {batch}, lr {lr}, data seed {args.seed + k}")
        history = seednet.train(net, pairs=pairs, batch=batch, crop=crop, lr=lr, log=log, seed=args.seed + k)
        log(f"stage {k + 1} done; first/last loss {history[0][1]:.4f} / {history[-1][1]:.4f}")
    net.save(args.out)
    log(f"saved {args.out} after {net.step_count} steps")
    return 0


if __name__ == "__main__":
    sys.exit(main())
