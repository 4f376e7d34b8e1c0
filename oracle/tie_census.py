"""Near-tie census of the FP32 index decisions, per benchmark configuration.

Test infrastructure (the oracle's own C scan, go_tie_census): for every
projection of a configuration's tilt series, counts the grid points whose
distance to the projection plane lies within k ulps of the cutoff
(fillInFourierGrid_DFT.m:178, D < cutoff) and the points with D < cutoff whose
in-plane |u| or |v| lies within k ulps of n/2 (:190).  Those are the only
points where a 1-ulp difference in the reference's own norm() or mldivide
(whose FP32 operation order MATLAB does not publish) could move a point in or
out of the master list; everywhere else the index contract is exact.

Usage: python -m oracle.tie_census [--configs tilt64,tilt256,...] [--ulps 1,4]
Prints one markdown row per configuration and ulp bound.
"""
import argparse
import time

import numpy as np

from genfire_amd import phantom
from oracle import oracle as O


def workload_angles(w: phantom.Workload) -> np.ndarray:
    """The tilt series of phantom.make_inputs: its generator draws the phantom's
    blobs before the angles, so they are drawn again here."""
    rng = np.random.Generator(np.random.PCG64(w.seed))
    phantom.blob_phantom(w.dim, w.n_blobs, w.sigma, w.radius, rng)
    return phantom.tilt_angles(w, rng)


def series(name: str):
    """[(workload, angles)] for one configuration: a named workload, or 'mix' (the
    16-request serving mix of the benchmark, seed 7)."""
    if name == "mix":
        out = []
        for r in phantom.serving_mix(16, 4.0, 7):
            w = phantom.mix_workload(r)
            out.append((w, phantom.tilt_angles(w, np.random.Generator(np.random.PCG64(r.seed)))))
        return out
    w = phantom.WORKLOADS[name]
    return [(w, workload_angles(w))]


def census(name: str, ulps: int):
    nd = nuv = kept = exact = nproj = 0
    grids = set()
    for w, ang in series(name):
        n = O.oversampled_size(w.dim, w.oversampling)
        grids.add(n)
        for a in ang:
            R, _ = O.rotation(a[0], a[1], a[2], *w.vectors)
            d, uv, k, e = O.tie_census(n, n, True, R, w.cutoff, ulps)
            nd += d
            nuv += uv
            kept += k
            exact += e
            nproj += 1
    return dict(config=name, grids=sorted(grids), projections=nproj, ulps=ulps, kept=kept, near_cutoff=nd,
                near_edge=nuv, exact_edge=exact)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--configs", default="tilt64,tilt256,tilt512,mix,tilt512x121")
    ap.add_argument("--ulps", default="1,4")
    ap.add_argument("--threads", type=int, default=0)
    args = ap.parse_args()
    if args.threads > 0:
        O.set_threads(args.threads)
    print("| config | grid n | projections | ulps | points kept | D within ulps of cutoff | |u| or |v| within "
          "ulps of n/2 (of them exactly n/2) | near-ties, fraction of kept (exact ones excluded) |")
    print("|---|---|---|---|---|---|---|---|")
    for name in args.configs.split(","):
        for u in (int(x) for x in args.ulps.split(",")):
            t = time.perf_counter()
            c = census(name, u)
            near = c["near_cutoff"] + c["near_edge"] - c["exact_edge"]
            grids = ",".join(str(g) for g in c["grids"])
            print(f"| {c['config']} | {grids} | {c['projections']} | {u} | {c['kept']} | {c['near_cutoff']} | "
                  f"{c['near_edge']} ({c['exact_edge']}) | {near} ({near / max(c['kept'], 1):.1e}) |  "
                  f"<!-- {time.perf_counter() - t:.0f} s -->", flush=True)


if __name__ == "__main__":
    main()
