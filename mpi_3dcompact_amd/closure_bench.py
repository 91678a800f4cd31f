"""Throughput of the learned closure's training and inference kernels on
synthetic pairs, with the batch of the 256^3 study (8192 pairs).

    python -m mpi_3dcompact_amd.closure_bench [--samples 1048576] [--epochs 2] [--batch 8192]

Prints the time per training batch (gradient partials, their fixed-order sum
and the Adam update), the training rate in pairs/s and in TFLOP/s
(52,608 FLOP per pair: the forward pass and two backward products of
7-64-64-64-2), and the inference rate over the same pairs.
"""
import argparse
import sys
import time

import numpy as np

FLOP_PER_PAIR = 2 * 3 * (7 * 64 + 64 * 64 + 64 * 64 + 64 * 2)


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--samples", type=int, default=1 << 20)
    ap.add_argument("--epochs", type=int, default=2)
    ap.add_argument("--batch", type=int, default=8192)
    args = ap.parse_args(argv)
    import torch
    from .closure import Closure
    if not torch.cuda.is_available():
        raise SystemExit("closure_bench needs a GPU")
    r = np.random.default_rng(0)
    x = r.standard_normal((args.samples, 7)).astype(np.float32)
    y = np.stack([np.sin(x[:, 0]) + x[:, 1] * x[:, 2], np.tanh(x[:, 3] - x[:, 4])], axis=1).astype(np.float32)
    X, T = torch.from_numpy(x).cuda(), torch.from_numpy(y).cuda()
    c = Closure(3)
    c.train(X, T, epochs=1, batch=args.batch, lr=1e-3, seed=1)  # warm-up
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    losses = c.train(X, T, epochs=args.epochs, batch=args.batch, lr=1e-3, seed=1)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    batches = args.epochs * ((args.samples + args.batch - 1) // args.batch)
    pairs = args.epochs * args.samples
    print(f"training: {args.epochs} epochs of {args.samples} pairs, batch {args.batch}: {dt * 1e3:.2f} ms, "
          f"{dt * 1e6 / batches:.1f} us per batch, {pairs / dt / 1e6:.1f} M pairs/s, "
          f"{pairs * FLOP_PER_PAIR / dt / 1e12:.2f} TFLOP/s; losses {', '.join(f'{v:.4e}' for v in losses)}",
          flush=True)
    c.infer(X)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(5):
        c.infer(X)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / 5
    print(f"inference: {args.samples} pairs in {dt * 1e3:.3f} ms, {args.samples / dt / 1e9:.2f} G pairs/s", flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
