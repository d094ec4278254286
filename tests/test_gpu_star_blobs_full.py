"""The large self-intersecting fills at full size: all 200 star blobs
(log-uniform 1k-100k cubics, radius 200-2000 px, 14.8 M tiles, winding above
1, lines of hundreds of pixels) through the plain pipeline and both batched
modes, against the oracle's per-path digests (tests/golden/
star_blobs_digests.json, made by make_star_blobs_digests.py in that folder).
Plain and the ordered batched call (which hands every blob to the plain
pipeline) must match every digest; the dense mode (which hands them to the
unordered tile stage) must match every tile/span-set digest and stay within
1 LSB of the plain alpha, bit-exact for at least 99.9% of the bytes."""
import json
import os

import numpy as np
import pytest

from ochre_amd import workloads as W
from parity import input_digests, path_digests

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "star_blobs_digests.json")

_STATE = {}


def _blobs():
    if "b" not in _STATE:
        b = W.star_blobs()
        doc = json.load(open(GOLDEN))
        assert input_digests(b) == [p["input"] for p in doc["paths"]], \
            "the star-blob generator produced other paths on this host than the fixture was made from"
        _STATE["b"], _STATE["doc"] = b, doc
    return _STATE["b"], _STATE["doc"]


def _run(ctx, mode):
    from ochre_amd.device import run_to_host
    b, doc = _blobs()
    print(f"[{mode}] star_blobs: 200 paths, {doc['tiles']} tiles expected", flush=True)
    return run_to_host(ctx, b, mode, tile_capacity=doc["tiles"] + 64, span_capacity=doc["spans"] + 64)


@pytest.mark.parametrize("mode", ["plain", "batched"])
def test_star_blobs_full_bit_exact(gpu_ctx, mode):
    b, doc = _blobs()
    got = _run(gpu_ctx, mode)
    dig = path_digests(*got)
    bad = [i for i, (d, p) in enumerate(zip(dig, doc["paths"]))
           if d != (p["tiles"], p["spans"], p["sets"], p["alpha"])]
    print(f"[{mode}] star_blobs: {len(dig) - len(bad)}/{len(dig)} paths match the oracle digests", flush=True)
    assert not bad, f"paths differing from the oracle: {bad[:10]}"
    if mode == "plain":
        _STATE["plain"] = got


def test_star_blobs_full_dense(gpu_ctx):
    b, doc = _blobs()
    got = _run(gpu_ctx, "dense")
    dig = path_digests(*got)
    bad = [i for i, (d, p) in enumerate(zip(dig, doc["paths"])) if d[:3] != (p["tiles"], p["spans"], p["sets"])]
    assert not bad, f"tile/span sets differing from the oracle: {bad[:10]}"
    ref = _STATE.get("plain") or _run(gpu_ctx, "plain")
    d = np.abs(got[1].astype(np.int16) - ref[1].astype(np.int16))
    exact = float((d == 0).mean())
    print(f"[dense] star_blobs: sets identical; alpha max {int(d.max())} LSB, bit-exact {exact:.7f}", flush=True)
    assert int(d.max()) <= 1 and exact >= 0.999
