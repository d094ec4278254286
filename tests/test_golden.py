"""The committed fixtures under tests/golden/ against the generators and the
oracle on this host (CPU only):

* readme_path_basic.txt is the oracle's README-path output in the reference's
  examples/basic.rs print format (its 100 tiles and 23 spans are pinned
  independently by tests/test_oracle.py::test_readme_tiles_and_spans_from_geometry);
* star_blobs_digests.json was made from the same star-blob paths the
  generator produces here (input digests), and the oracle reproduces the
  digests of a sample of its paths (the smallest five and one large one)."""
import json
import os

import numpy as np
import pytest

import oracle_ref as O
from ochre_amd import workloads as W
from parity import basic_rs_text, input_digests, path_digests

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_readme_fixture_is_the_oracle_output_in_basic_rs_format():
    v, p, w, xf = W.readme_path().path(0)
    r = O.Rasterizer()
    r.fill(v, p, w, xf)
    text = basic_rs_text(r.finish())
    assert text == open(os.path.join(GOLDEN, "readme_path_basic.txt")).read()
    assert text.count("tile at (") == 100 and text.count("span at (") == 23


@pytest.fixture(scope="module")
def blobs():
    return W.star_blobs(), json.load(open(os.path.join(GOLDEN, "star_blobs_digests.json")))


def test_star_blob_fixture_inputs_match_the_generator(blobs):
    b, doc = blobs
    assert len(doc["paths"]) == b.n_paths == 200
    assert input_digests(b) == [p["input"] for p in doc["paths"]]
    assert doc["tiles"] == sum(p["tiles"] for p in doc["paths"])


def test_star_blob_fixture_reproduced_by_the_oracle(blobs):
    b, doc = blobs
    sizes = np.diff(b.verb_off)
    sample = list(np.argsort(sizes)[:5]) + [int(np.argsort(sizes)[100])]
    r = O.run_batch(b.subset(sample))
    got = path_digests(r.tile_xy, r.tile_alpha, r.span_xy, r.span_w, r.path_tile_off, r.path_span_off)
    for k, i in enumerate(sample):
        p = doc["paths"][int(i)]
        assert got[k] == (p["tiles"], p["spans"], p["sets"], p["alpha"]), int(i)
