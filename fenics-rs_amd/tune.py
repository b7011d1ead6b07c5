#!/usr/bin/env python3
"""Tile tuner for the batched block-CG solve (fh_batched_poisson_block_cg).

Sweeps the two launch parameters of the batched solve on this device:

* ``batched_wgs_per_cu``: workgroups per compute unit in the fixed grid of
  the product and update kernels, which stride over the work list of active
  256-row tiles;
* ``batched_chunk``: block iterations per HIP-graph chunk (the work list is
  compacted and the host looks at it once per chunk);

times the kernel (HIP events inside the library) on the batched workload of
256 unit_square(N, N) instances, N ~ U{32..512} (seed 2024), four nodal loads
each, tol 1e-10, in the given precision, and writes the fastest pair to
``fh_tune.json`` beside libfenics_hip.so (or to --out), which the library
reads at every batched solve.  FH_BATCH_WGS_PER_CU / FH_BATCH_CHUNK override
the file.

Usage: python fenics-rs_amd/tune.py [--precision mixed|fp64] [--reps 2]
                                    [--instances 256] [--out PATH]
"""
import argparse
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--precision", choices=("mixed", "fp64"), default="mixed")
    ap.add_argument("--reps", type=int, default=2)
    ap.add_argument("--instances", type=int, default=256)
    ap.add_argument("--wgs", default="4,8,16,32")
    ap.add_argument("--chunks", default="8,16,32")
    ap.add_argument("--out", default=os.path.join(HERE, "fh_tune.json"))
    args = ap.parse_args()

    import numpy as np
    import fenics_hip as fh

    if not fh.discover_gpus():
        raise SystemExit("tune.py: no HIP device")
    prec = fh.MIXED if args.precision == "mixed" else fh.FP64
    rng = np.random.default_rng(2024)
    sizes = rng.integers(32, 513, args.instances)
    fs = [rng.uniform(-1, 1, ((int(N) + 1) ** 2, 4)) for N in sizes]
    results = []
    for chunk in [int(t) for t in args.chunks.split(",")]:
        for wgs in [int(w) for w in args.wgs.split(",")]:
            os.environ["FH_BATCH_CHUNK"] = str(chunk)
            os.environ["FH_BATCH_WGS_PER_CU"] = str(wgs)
            best = None
            for _ in range(args.reps):
                _, its, st = fh.batched_poisson_block_cg(sizes, fs, tol=1e-10, max_iter=40000,
                                                         precision=prec)
                best = st.kernel_ms if best is None else min(best, st.kernel_ms)
            results.append({"batched_chunk": chunk, "batched_wgs_per_cu": wgs, "kernel_ms": best,
                            "gdof_iter_per_s": st.gflops})
            print(f"chunk {chunk:3d} wgs/CU {wgs:2d}: {best:8.1f} ms", file=sys.stderr, flush=True)
    os.environ.pop("FH_BATCH_CHUNK", None)
    os.environ.pop("FH_BATCH_WGS_PER_CU", None)
    win = min(results, key=lambda r: r["kernel_ms"])
    dev = fh.discover_gpus()[0]
    out = {"batched_chunk": win["batched_chunk"], "batched_wgs_per_cu": win["batched_wgs_per_cu"],
           "device": dev.name, "arch": dev.arch, "precision": args.precision,
           "workload": f"{args.instances} unit_square(N,N), N~U[32,512] seed 2024, 4 RHS, tol 1e-10",
           "sweep": results}
    with open(args.out, "w") as f:
        json.dump(out, f, indent=1)
    print(json.dumps({"winner": win, "out": args.out}))


if __name__ == "__main__":
    main()
