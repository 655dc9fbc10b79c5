#!/usr/bin/env python3
"""Step time of the batched engine on a mix of water systems.

nsys systems (default 512) of round(exp(U(ln 100, ln 10000))) waters each
(seed 4), boxes at 0.0334 waters/A^3, rcut 9 A, tol 2e-5 (at which every
such box is within 4e-6 relative force error of the direct sum), packed
into one GLSTBatch; inputs resident on the device.  Prints the mean step
time (zero_ef + calc_ener_force) over `steps` steps after one warm-up step.

  python3 examples/batch_bench.py [--nsys 512] [--ct float] [--steps 5]
                                  [--overlap 1] [--tol 2e-5]
"""
import argparse
import os
import sys
import time

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--nsys", type=int, default=512)
    ap.add_argument("--ct", default="float", choices=("float", "double"))
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--overlap", type=int, default=1,
                    help="overlap_near option of the batch (0: serial)")
    ap.add_argument("--tol", type=float, default=2e-5)
    ap.add_argument("--sf-cheb", type=int, default=1, choices=(0, 1),
                    help="sf_cheb option of the batch (0: direct S(k) sums)")
    args = ap.parse_args()
    import numpy as np
    import torch
    import glst_amd as g
    rng = np.random.default_rng(4)
    nw = np.round(np.exp(rng.uniform(np.log(100), np.log(10000),
                                     args.nsys))).astype(int)
    boxes = (nw / 0.0334) ** (1 / 3)
    hs = [g.make_water_box(int(w), float(b), seed=100 + i)
          for i, (w, b) in enumerate(zip(nw, boxes))]
    natoms = [h[0].size for h in hs]
    cols = [torch.as_tensor(np.concatenate([h[i] for h in hs]), device="cuda")
            for i in range(4)]
    b = g.GLSTBatch(natoms, boxes, 9.0, args.tol, ct=args.ct)
    b.set_option("overlap_near", args.overlap)
    if args.sf_cheb != 1:  # 1 is the default; setting it packs the batch again
        b.set_option("sf_cheb", args.sf_cheb)
    b.zero_ef()
    b.calc_ener_force(*cols)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        b.zero_ef()
        b.calc_ener_force(*cols)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / args.steps
    b.close()
    print(f"batch: {args.nsys} systems, {sum(natoms)} atoms, ct {args.ct}, "
          f"tol {args.tol}: "
          f"{dt * 1e3:.2f} ms per step, {sum(natoms) / dt:.3e} atom-steps/s",
          flush=True)


if __name__ == "__main__":
    main()
