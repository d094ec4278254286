"""The device coverage adjoint against the f64 oracle: loss, analytic
gradients (f64 CPU adjoint of the same coverage formula) and central finite
differences of the f64 forward (h = 1e-3 px) on sampled parameters."""
import numpy as np
import pytest

import oracle_ref as O
from ochre_amd import workloads as W

pytestmark = pytest.mark.gpu


def perturbed(batch, sigma_px, seed):
    rng = np.random.default_rng(seed)
    t = W.PathBatch(batch.verbs, batch.verb_off, batch.points.copy(), batch.point_off, batch.weights,
                    batch.weight_off, batch.xforms, name=batch.name + "[target]")
    for i in range(batch.n_paths):
        a, b = batch.point_off[i], batch.point_off[i + 1]
        s = float(batch.xforms[i, 0])
        t.points[a:b] += (rng.normal(0, sigma_px, (b - a, 2)) / s).astype(np.float32)
    return t


def render_target(ctx, batch):
    import torch
    from ochre_amd.device import DeviceOutput, rasterize_batched, upload
    db = upload(batch)
    out = DeviceOutput(batch.n_paths, batch.n_paths * 128, batch.n_paths * 16, coverage=True)
    rasterize_batched(ctx, db, out)
    torch.cuda.synchronize()
    return out


def host_target(out, i):
    ts = int(out.path_tile_start[i]); tc = int(out.path_tile_count[i])
    ss = int(out.path_span_start[i]); sc = int(out.path_span_count[i])
    return O.Target(out.tile_xy[ts:ts + tc].cpu().numpy(), out.tile_cov[ts:ts + tc].cpu().numpy(),
                    out.span_xy[ss:ss + sc].cpu().numpy(), out.span_w[ss:ss + sc].cpu().numpy().view(np.uint16))


def test_forward_coverage_matches_alpha(gpu_ctx):
    b = W.adjoint_glyphs(n=64)
    out = render_target(gpu_ctx, b)
    cov = out.tile_cov[:out.n_tiles].cpu().numpy()
    alpha = out.tile_alpha[:out.n_tiles].cpu().numpy()
    assert np.array_equal(np.minimum(np.floor(cov * 256.0), 255).astype(np.uint8), alpha)


@pytest.mark.parametrize("n,seed", [(96, 0), (96, 1)])
def test_adjoint_matches_f64_oracle_and_finite_differences(gpu_ctx, n, seed):
    import torch
    from ochre_amd.device import coverage_adjoint, upload
    b = W.adjoint_glyphs(n=n, seed=100 + seed)
    tb = perturbed(b, 0.5, 900 + seed)
    tgt = render_target(gpu_ctx, tb)
    loss, dp, dx = coverage_adjoint(gpu_ctx, upload(b), tgt)
    torch.cuda.synchronize()
    loss, dp, dx = loss.cpu().numpy(), dp.cpu().numpy(), dx.cpu().numpy()
    rng = np.random.default_rng(seed)
    worst_g = 0.0
    checked_fd = 0
    for i in range(b.n_paths):
        v, p, w, xf = b.path(i)
        t = host_target(tgt, i)
        p64, xf64 = p.astype(np.float64), xf.astype(np.float64)
        st, L, gp, gx, _ = O.path_loss64(v, p64, w, xf64, t, grad=True)
        assert st == 0
        # f32 coverage against f64: at canvas coordinates up to 1024 px an f32
        # point carries ~6e-5 px of rounding, so an edge pixel's coverage
        # differs by up to ~1e-4 and the loss (hundreds of edge pixels, each
        # d(cov - t)^2 = 2 (cov - t) dcov) by ~1e-3 relative at most
        assert abs(loss[i] - L) <= 1e-3 * max(L, 1.0), (i, loss[i], L)
        a, e = b.point_off[i], b.point_off[i + 1]
        scale = max(np.abs(gp).max(), np.abs(gx).max(), 1.0)
        err = max(np.abs(dp[a:e] - gp).max(), np.abs(dx[i] - gx).max()) / scale
        worst_g = max(worst_g, err)
        assert err <= 1e-3, (i, err)
        # central finite differences of the f64 forward (h = 1e-3 px) on
        # sampled parameters, up to 3 tries per path for one without a kink
        # in [x - h, x + h] (a kink: the one-sided quotients disagree by more
        # than the curvature of a smooth loss explains at this h)
        for _ in range(3):
            if checked_fd >= 64:
                break
            k = int(rng.integers(0, p64.size + 6))
            if k < p64.size:
                h = 1e-3 / float(xf64[0])
                f = lambda z: O.path_loss64(v, z.reshape(-1, 2), w, xf64, t)[1]
                x0, gk = p64.reshape(-1).copy(), float(dp[a:e].reshape(-1)[k])
            else:
                k -= p64.size
                h = 1e-3 / 32.0 if k < 4 else 1e-3
                f = lambda z: O.path_loss64(v, p64, w, z, t)[1]
                x0, gk = xf64.copy(), float(dx[i, k])
            xp, xm = x0.copy(), x0.copy()
            xp[k] += h
            xm[k] -= h
            lp, lm, l0 = f(xp), f(xm), f(x0)
            c, r, l = (lp - lm) / (2 * h), (lp - l0) / h, (l0 - lm) / h
            if abs(r - l) <= 2e-2 * max(abs(c), 1.0):
                assert abs(gk - c) <= 1e-3 * max(abs(c), 1.0), (i, k, gk, c)
                checked_fd += 1
                break
    print(f"adjoint n={n}: worst gradient rel err vs f64 adjoint {worst_g:.2e}, fd checks {checked_fd}", flush=True)
    assert checked_fd >= 32
