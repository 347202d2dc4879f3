"""Block-size autotuner of the batched form.

The batched DROP iteration runs block step j of every scan in one pair of
launches, so its block loop takes as many steps as the scan with the most
blocks, and each step costs a launch latency plus a little per history. The
MLP pass (FP64-bound, ~proportional to the total number of MLP steps) runs
beside it. The block size of each scan therefore trades steps against work
per step; it also changes the reconstruction (DROP's relaxation per block),
so the candidates are restricted to the sizes the configuration allows.

The tuner measures, on the real scans, one timed iteration (after a warm-up
that builds the plans and graphs) for each candidate assignment:
  - every scan at the same size (one candidate per allowed size);
  - step-balanced assignments: for a target number of steps T, each scan gets
    the smallest allowed size whose block count is at most T (the largest
    size when none is), for T between the smallest and largest block counts.
It returns the fastest assignment and the table of measurements.
"""
import math
import time


def candidate_assignments(n_histories, sizes):
    sizes = sorted(sizes)
    n = len(n_histories)
    cands = [tuple([b] * n) for b in sizes]
    counts = sorted({math.ceil(h / b) for h in n_histories for b in sizes})
    for t in counts:
        assign = []
        for h in n_histories:
            fit = [b for b in sizes if math.ceil(h / b) <= t]
            assign.append(fit[0] if fit else sizes[-1])
        cands.append(tuple(assign))
    out, seen = [], set()
    for c in cands:
        if c not in seen:
            seen.add(c)
            out.append(c)
    return out


def tune(batch, hists, x0s, sizes, lambdas=None, mode=0, iterations=1, max_candidates=8, log=None):
    """Times each candidate assignment (`iterations` timed iterations after one
    warm-up, images restarted from x0s) and returns (best, table)."""
    import torch
    n_hist = [h.n for h in hists]
    cands = candidate_assignments(n_hist, sizes)
    if len(cands) > max_candidates:
        # keep the uniform ones and an even spread of the balanced ones
        uni = cands[:len(sizes)]
        rest = cands[len(sizes):]
        k = max_candidates - len(uni)
        step = max(1, len(rest) // max(1, k))
        cands = uni + rest[::step][:k]
    stream = batch.contexts[0].stream
    table = []
    for assign in cands:
        t0 = time.time()
        batch.prepare(hists, list(assign), mode & 0xff)
        xs = [x.clone() for x in x0s]
        batch.iteration(hists, xs, list(assign), lambdas, mode)         # warm-up: graphs
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(iterations):
            batch.iteration(hists, xs, list(assign), lambdas, mode)
        e1.record(stream)
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / iterations
        steps = max(math.ceil(h / b) for h, b in zip(n_hist, assign))
        table.append({"block_sizes": list(assign), "ms_per_iteration": round(ms, 3), "block_steps": steps,
                      "setup_s": round(time.time() - t0, 2)})
        if log:
            log(f"tuner: {list(assign)} steps {steps}: {ms:.2f} ms/iteration")
        del xs
    best = min(table, key=lambda r: r["ms_per_iteration"])
    return best, table
