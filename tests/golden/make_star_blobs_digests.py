"""Writes tests/golden/star_blobs_digests.json: the oracle's output for every
path of the large-fills workload (200 star blobs of log-uniform 1k-100k
cubics, radius 200-2000 px with 15% noise, centred on a 4096x4096 canvas),
as per-path SHA-256 digests of the tile/span sets and of the alpha bytes,
with a digest of each path's input.  The oracle takes about a minute on one
core for the whole workload, so the GPU tests and the bench compare against
these digests instead of re-running it.

    python tests/golden/make_star_blobs_digests.py [--jobs 8]
"""
import argparse
import json
import multiprocessing as mp
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

OUT = os.path.join(HERE, "star_blobs_digests.json")


def _one(i):
    import oracle_ref
    from ochre_amd import workloads as W
    from parity import path_digests
    b = W.star_blobs()
    sub = b.subset([i])
    r = oracle_ref.run_batch(sub)
    assert r.status == 0, (i, r.status)
    return i, path_digests(r.tile_xy, r.tile_alpha, r.span_xy, r.span_w, r.path_tile_off, r.path_span_off)[0], \
        r.stats


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=min(8, os.cpu_count() or 1))
    args = ap.parse_args()
    from ochre_amd import workloads as W
    from parity import input_digests
    b = W.star_blobs()
    t0 = time.time()
    # largest paths first so the pool's tail is short
    order = sorted(range(b.n_paths), key=lambda i: -int(b.verb_off[i + 1] - b.verb_off[i]))
    with mp.get_context("fork").Pool(args.jobs) as pool:
        res = dict((i, (d, st)) for i, d, st in pool.imap_unordered(_one, order))
    paths = []
    inp = input_digests(b)
    for i in range(b.n_paths):
        (nt, ns, sets, alpha), st = res[i]
        paths.append({"input": inp[i], "tiles": nt, "spans": ns, "sets": sets, "alpha": alpha,
                      "lines": st["lines"], "increments": st["increments"]})
    doc = {"workload": b.name, "generator": "ochre_amd.workloads.star_blobs() (defaults)",
           "what": "per path: SHA-256 of the input bytes (verbs, points, weights, transform), and of the oracle's "
                   "output in TileBuilder order: sets = tile xy | span xy | span widths, alpha = tile alpha bytes",
           "tiles": sum(p["tiles"] for p in paths), "spans": sum(p["spans"] for p in paths), "paths": paths}
    with open(OUT, "w") as f:
        json.dump(doc, f, indent=0)
        f.write("\n")
    print(f"{OUT}: {b.n_paths} paths, {doc['tiles']} tiles, {doc['spans']} spans in {time.time() - t0:.1f} s")


if __name__ == "__main__":
    main()
