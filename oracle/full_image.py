"""TEST INFRASTRUCTURE: the CPU oracle's whole reconstruction of a configured
scan -- the reference's main() pipeline (preprocess: intersections, binning,
cuts, FBP, MSC hull, x0, endpoints) then DROP (.cu:8023-8055 semantics) --
saving the image after chosen iterations as float32 .npy files, so that GPU
images of the same scan can be compared with it (bench.py --oracle-images).

  python -m oracle.full_image --config cfg2 --iterations 1,6,12 --threads 8 --out DIR
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--config", default="cfg2")
    ap.add_argument("--histories", type=int, default=None)
    ap.add_argument("--iterations", default="1,6,12")
    ap.add_argument("--threads", type=int, default=os.cpu_count())
    ap.add_argument("--out", required=True)
    args = ap.parse_args()
    import oracle
    import pct_amd
    from pct_amd import configs
    its = sorted(int(i) for i in args.iterations.split(","))
    os.makedirs(args.out, exist_ok=True)
    p = configs.params(args.config)
    cfg = configs.synth_cfg(args.config, n_histories=args.histories)
    t0 = time.time()
    raw = pct_amd.generate_host(cfg).numpy()
    h, x0, hull, fbp = oracle.preprocess(p, raw)
    del raw
    t_pre = time.time() - t0
    print(f"preprocess {t_pre:.1f} s, {h.n} reconstruction histories", flush=True)
    np.save(os.path.join(args.out, "x0.npy"), x0)
    x = x0
    done = 0
    times = {}
    for it in its:
        t1 = time.time()
        x = oracle.drop(p, h, x, iterations=it - done, threads=args.threads)
        times[it] = time.time() - t1
        done = it
        np.save(os.path.join(args.out, f"x{it}.npy"), x)
        print(f"iteration {it}: {times[it]:.1f} s", flush=True)
    meta = {"config": args.config, "histories": int(cfg.n_histories), "reconstruction_histories": int(h.n),
            "iterations": its, "threads": args.threads, "preprocess_s": round(t_pre, 1),
            "drop_s": {str(k): round(v, 1) for k, v in times.items()},
            "command": " ".join(["python -m oracle.full_image"] + sys.argv[1:])}
    json.dump(meta, open(os.path.join(args.out, "meta.json"), "w"), indent=1)
    print(json.dumps(meta), flush=True)


if __name__ == "__main__":
    main()
