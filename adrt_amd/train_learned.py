"""Train the learned CG initial guess / preconditioner on the GPU.

    python -m adrt_amd.train_learned --steps 12000 --out adrt_amd/weights/learned_256.npz

(the committed weights: 12000 steps of batch 64 over the 50k pairs, 342 s on
one MI355X).  Fits both preconditioners (least squares), then trains the U-Net
with Adam (cosine learning-rate decay) on seeded phantom pairs generated on
the device: phantom i of seed ``--seed``, noisy ADRT data at ``--sigma`` of
max|adrt(x)|, input = preconditioned normal-equation right-hand side,
target = the clean phantom (default), the normal equations' solution
(``--target solution``, computed by preconditioned CG per batch) or no target
(``--target residual``: the relative residual of the guess itself).
Progress lines go to stdout; the weights, preconditioner coefficients and
loss history go to ``--out`` (NumPy .npz, no pickles).
"""

import argparse
import json
import math
import sys
import time


def main(argv=None):
    p = argparse.ArgumentParser()
    p.add_argument("--n", type=int, default=256)
    p.add_argument("--batch", type=int, default=64)
    p.add_argument("--pairs", type=int, default=50000)
    p.add_argument("--steps", type=int, default=2000)
    p.add_argument("--lr", type=float, default=1e-3)
    p.add_argument("--sigma", type=float, default=0.005)
    p.add_argument("--seed", type=int, default=1000)
    p.add_argument("--time-limit", type=float, default=0.0,
                   help="stop after this many seconds of training (0: no limit)")
    p.add_argument("--log-every", type=int, default=100)
    p.add_argument("--target", default="phantom", choices=["phantom", "solution", "residual"])
    p.add_argument("--input", default="pyramid", choices=["pyramid", "dct"],
                   help="preconditioner forming the network's input z = P b")
    p.add_argument("--out", required=True)
    args = p.parse_args(argv)

    import torch

    from adrt_amd import learned

    torch.cuda.init()
    solver = learned.LearnedSolver(n=args.n, sigma_frac=args.sigma)
    solver.input_prec = args.input
    t0 = time.perf_counter()
    solver.fit_preconditioner()
    torch.cuda.synchronize()
    print(json.dumps({"event": "preconditioner", "seconds": round(time.perf_counter() - t0, 2),
                      "theta_level0": [round(v, 5) for v in solver.pyramid.theta[0]]}), flush=True)
    losses = []
    t1 = time.perf_counter()
    steps_done = 0

    def lr_fn(step):
        return args.lr * (0.02 + 0.98 * 0.5 * (1 + math.cos(math.pi * step / args.steps)))

    def log(step, loss):
        losses.append((step, loss))
        print(json.dumps({"step": step, "loss": loss, "lr": lr_fn(step),
                          "seconds": round(time.perf_counter() - t1, 2)}), flush=True)

    chunk = max(args.log_every, 1)
    while steps_done < args.steps:
        k = min(chunk, args.steps - steps_done)
        # one logged step per chunk keeps the progress lines coming
        offset = steps_done
        solver.train(k, batch=args.batch, seed=args.seed, pairs=args.pairs, log_every=k,
                     log=lambda s, l: log(offset + s, l),
                     lr_fn=lambda s: lr_fn(offset + s), start=offset, target=args.target)
        steps_done += k
        if args.time_limit and time.perf_counter() - t1 > args.time_limit:
            break
    torch.cuda.synchronize()
    secs = time.perf_counter() - t1
    solver.save(args.out, losses=losses, steps=steps_done, train_seconds=secs,
                batch=args.batch, pairs=args.pairs, target=args.target)
    print(json.dumps({"event": "done", "steps": steps_done, "seconds": round(secs, 2),
                      "ms_per_step": round(1e3 * secs / max(steps_done, 1), 2),
                      "out": args.out}), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
