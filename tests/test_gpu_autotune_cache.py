"""The library's autotune cache (ochre_autotune_batched_cached): the first
call on a batch tunes and appends a line to the cache file; a later call with
a batch of the same signature applies the stored parameters without timing;
another process reads them from the file; a batch of another signature is
tuned again."""
import os
import subprocess
import sys

import pytest

from ochre_amd import workloads as W

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _params(ctx):
    from ochre_amd.device import batched_params
    return {k: v for k, v in batched_params(ctx).items() if k != "graph"}


def test_autotune_cache_round_trip(tmp_path):
    import torch
    from ochre_amd.device import Context, DeviceOutput, autotune_cached, set_batched_params, upload
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    # a context of its own: tuning changes the launch shapes other tests use
    gpu_ctx = Context(0)
    path = str(tmp_path / "tune.txt")
    b = W.glyph_run(600, seed=41)
    db = upload(b)
    out = DeviceOutput(b.n_paths, b.n_paths * 64, b.n_paths * 8)
    print("[autotune-cache] tuning 600 glyphs", flush=True)
    assert autotune_cached(gpu_ctx, db, out, reps=1, cache_path=path)
    tuned = _params(gpu_ctx)
    lines = open(path).read().splitlines()
    assert len(lines) == 1 and " = " in lines[0]
    # the same signature (the same glyphs in another order): applied, not timed
    set_batched_params(gpu_ctx, 1, 1, 16)
    b2 = b.subset(list(range(b.n_paths))[::-1])
    assert not autotune_cached(gpu_ctx, upload(b2), DeviceOutput(b2.n_paths, b2.n_paths * 64, b2.n_paths * 8),
                               reps=1, cache_path=path)
    assert _params(gpu_ctx) == tuned
    # another process reads the file
    code = ("import sys; sys.path.insert(0, %r)\n"
            "from ochre_amd import workloads as W\n"
            "from ochre_amd.device import Context, DeviceOutput, autotune_cached, upload\n"
            "b = W.glyph_run(600, seed=41); ctx = Context(0)\n"
            "t = autotune_cached(ctx, upload(b), DeviceOutput(b.n_paths, b.n_paths * 64, b.n_paths * 8), reps=1,"
            " cache_path=%r)\n"
            "print('TUNED' if t else 'CACHED'); ctx.close()\n") % (ROOT, path)
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stderr[-2000:]
    assert "CACHED" in r.stdout
    # another signature: tuned again, a second line
    b3 = W.glyph_run(40, seed=44)
    assert autotune_cached(gpu_ctx, upload(b3), DeviceOutput(b3.n_paths, b3.n_paths * 64, b3.n_paths * 8),
                           reps=1, cache_path=path)
    assert len(open(path).read().splitlines()) == 2
    gpu_ctx.close()
