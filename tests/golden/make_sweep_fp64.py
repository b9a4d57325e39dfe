"""Writes tests/golden/sweep_fp64_rates.json: FP64-oracle population rates of the sampled
points of the 256-point Brunel sweep, each under several realisations of the external
Poisson drive (same connectivity; drive seed = point seed + 7919 k, k = 0 is the point's
own seed).  The sweep's rate-error gate compares the mean over these realisations with
the mean over the same realisations of the batched FP32 GPU runs.

    python tests/golden/make_sweep_fp64.py [--realisations 16] [--every 8] [--procs 8]

Single-threaded oracle runs in a process pool; deterministic (the oracle's results do
not depend on its thread count).
"""
import argparse
import json
import multiprocessing as mp
import os
import sys
import time

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

OUT = os.path.join(ROOT, "tests", "golden", "sweep_fp64_rates.json")


def drive_seed(point_seed, k):
    from spike_amd import sweep
    return int(point_seed) + sweep.DRIVE_SEED_STEP * int(k)


def _one(job):
    from spike_amd import sweep
    from sweep_replay import oracle_rate
    k, r, seconds, n_points = job
    p = sweep.sample_points(n_points)[k]
    rate, el = oracle_rate(p, seconds, precision=64, threads=1, drive_seed=drive_seed(p.seed, r))
    return k, r, rate, el


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--points", type=int, default=256)
    ap.add_argument("--every", type=int, default=8)
    ap.add_argument("--realisations", type=int, default=16)
    ap.add_argument("--seconds", type=float, default=1.0)
    ap.add_argument("--procs", type=int, default=min(8, os.cpu_count() or 1))
    args = ap.parse_args()
    from spike_amd import sweep
    pts = sweep.sample_points(args.points)
    idx = list(range(0, args.points, args.every))
    rates = {k: [None] * args.realisations for k in idx}
    # realisations already in the fixture (same sweep, seconds and seeds) are kept
    if os.path.exists(OUT):
        old = json.load(open(OUT))
        if old.get("points_in_sweep") == args.points and old.get("seconds") == args.seconds:
            for e in old["points"]:
                if e["index"] in rates:
                    for r, (sd, v) in enumerate(zip(e["drive_seeds"], e["rate_fp64"])):
                        if r < args.realisations and sd == drive_seed(e["seed"], r):
                            rates[e["index"]][r] = v
    jobs = [(k, r, args.seconds, args.points) for k in idx for r in range(args.realisations)
            if rates[k][r] is None]
    # longest first (cost grows with N^2)
    jobs.sort(key=lambda j: -pts[j[0]].n)
    t0 = time.time()
    with mp.get_context("spawn").Pool(args.procs) as pool:
        for n_done, (k, r, rate, el) in enumerate(pool.imap_unordered(_one, jobs), 1):
            rates[k][r] = rate
            print(f"[{n_done}/{len(jobs)}] point {k} N={pts[k].n} realisation {r}: {rate:.4f} Hz "
                  f"({el:.1f} s)", flush=True)
    out = {
        "description": "FP64 oracle population rates (Hz) of sampled points of the Brunel sweep "
                       "(spike_amd.sweep.sample_points), one entry per external-drive realisation",
        "generator": "tests/golden/make_sweep_fp64.py",
        "points_in_sweep": args.points, "seconds": args.seconds, "drive_seed_step": 7919,
        "realisations": args.realisations,
        "points": [dict(index=k, n=pts[k].n, g=pts[k].g, nu_ratio=pts[k].nu_ratio, J=pts[k].J,
                        delay_steps=pts[k].delay_steps, seed=pts[k].seed,
                        drive_seeds=[drive_seed(pts[k].seed, r) for r in range(args.realisations)],
                        rate_fp64=rates[k]) for k in idx],
        "wall_s": round(time.time() - t0, 1),
    }
    with open(OUT, "w") as f:
        json.dump(out, f, indent=1)
    print("wrote", OUT)


if __name__ == "__main__":
    main()
