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


def _check_adjoint(ctx, b, tb, fd_samples, seed):
    """Loss and every gradient element of the device adjoint against the f64
    oracle adjoint, and `fd_samples` seeded parameters against central finite
    differences of the f64 forward."""
    tgt = render_target(ctx, tb)
    _check_adjoint_against(ctx, b, tgt, [host_target(tgt, i) for i in range(b.n_paths)], fd_samples, seed)


def oracle_target(ctx, tb):
    """A target render from the oracles: tiles with the f64 forward's coverage
    and the f32 oracle's spans (the two tile sets are identical), as a device
    output with per-path ranges, and the same per path on the host."""
    import torch
    from ochre_amd.device import DeviceOutput
    ref = O.run_batch(tb)
    txys, covs, hosts = [], [], []
    for i in range(tb.n_paths):
        v, p, w, xf = tb.path(i)
        n = int(ref.path_tile_off[i + 1] - ref.path_tile_off[i])
        st, _, _, _, (txy, cov) = O.path_loss64(v, p.astype(np.float64), w, xf.astype(np.float64), want_cov=True,
                                                cov_cap=max(n, 1))
        assert st == 0 and np.array_equal(txy, ref.tile_xy[ref.path_tile_off[i]:ref.path_tile_off[i + 1]])
        s0, s1 = ref.path_span_off[i], ref.path_span_off[i + 1]
        txys.append(txy)
        covs.append(cov.astype(np.float32))
        hosts.append(O.Target(txy, cov.astype(np.float32), ref.span_xy[s0:s1], ref.span_w[s0:s1]))
    nt, ns = ref.tile_xy.shape[0], ref.span_xy.shape[0]
    out = DeviceOutput(tb.n_paths, nt, ns, coverage=True)
    out.tile_xy[:nt] = torch.from_numpy(np.concatenate(txys).astype(np.int16))
    out.tile_cov[:nt] = torch.from_numpy(np.concatenate(covs))
    out.span_xy[:ns] = torch.from_numpy(ref.span_xy.astype(np.int16))
    out.span_w[:ns] = torch.from_numpy(ref.span_w.astype(np.uint16).view(np.int16))
    out.path_tile_start[:] = torch.from_numpy(ref.path_tile_off[:-1].astype(np.int64))
    out.path_tile_count[:] = torch.from_numpy(np.diff(ref.path_tile_off).astype(np.int32))
    out.path_span_start[:] = torch.from_numpy(ref.path_span_off[:-1].astype(np.int64))
    out.path_span_count[:] = torch.from_numpy(np.diff(ref.path_span_off).astype(np.int32))
    out.n_tiles, out.n_spans = nt, ns
    return out, hosts


def _check_adjoint_against(ctx, b, tgt, targets, fd_samples, seed):
    import torch
    from ochre_amd.device import coverage_adjoint, upload
    loss, dp, dx = coverage_adjoint(ctx, upload(b), tgt)
    torch.cuda.synchronize()
    loss, dp, dx = loss.cpu().numpy(), dp.cpu().numpy(), dx.cpu().numpy()
    errs = []
    worst = (0.0, None)
    beyond = []  # elements off the f64 gradient by more than 1e-3
    for i in range(b.n_paths):
        v, p, w, xf = b.path(i)
        st, L, gp, gx, _ = O.path_loss64(v, p.astype(np.float64), w, xf.astype(np.float64), targets[i], grad=True)
        assert st == 0
        # f32 coverage against f64: at canvas coordinates up to 1024 px an f32
        # point carries ~6e-5 px of rounding, so an edge pixel's coverage
        # differs by up to ~1e-4 and the loss (hundreds of edge pixels, each
        # d(cov - t)^2 = 2 (cov - t) dcov) by ~1e-3 relative at most
        assert abs(loss[i] - L) <= 1e-3 * max(L, 1.0), (i, loss[i], L)
        a, e = b.point_off[i], b.point_off[i + 1]
        # per parameter: |g - g64| <= 1e-3 * max(|g64|, 1)
        got = np.concatenate([dp[a:e].reshape(-1), dx[i]])
        ref = np.concatenate([gp.reshape(-1), gx])
        rel = np.abs(got - ref) / np.maximum(np.abs(ref), 1.0)
        errs.append(rel)
        k = int(np.argmax(rel))
        if rel[k] > worst[0]:
            worst = (float(rel[k]), (i, k, float(got[k]), float(ref[k])))
        beyond += [(i, int(k), float(got[k]), float(ref[k])) for k in np.nonzero(rel > 1e-3)[0]]
    errs = np.concatenate(errs)
    print(f"adjoint {b.name}: {errs.size} gradient elements, per-parameter rel err vs f64 adjoint: "
          f"median {np.median(errs):.1e}, p99.9 {np.quantile(errs, 0.999):.1e}, max {worst[0]:.1e} at {worst[1]}",
          flush=True)
    # The device takes the forward's discrete decisions in f32 (which pixels a
    # line visits, in the local frame) on end points that lie within ~1e-4 px
    # of the f64 ones.  Where a kink of the loss lies that close (a y-extremal
    # vertex on a pixel-row boundary, a sliver taking |v| to 0 or 1), the
    # device's gradient is the one-sided derivative of the other side.  Such an
    # element must lie between the f64 forward's one-sided difference
    # quotients at h = 1e-4 px, which must disagree (a kink), and there may be
    # only a handful of them.
    for i, k, g_dev, g_ref in beyond:
        v, p, w, xf = b.path(i)
        p64, xf64, t = p.astype(np.float64), xf.astype(np.float64), targets[i]
        if k < p64.size:
            h = 1e-4 / float(xf64[0])
            x0 = p64.reshape(-1).copy()
            f = lambda z: O.path_loss64(v, z.reshape(-1, 2), w, xf64, t)[1]  # noqa: E731
        else:
            h = 1e-4 / 32.0 if k - p64.size < 4 else 1e-4
            x0 = xf64.copy()
            f = lambda z: O.path_loss64(v, p64, w, z, t)[1]  # noqa: E731
        kk = k if k < p64.size else k - p64.size
        xp, xm = x0.copy(), x0.copy()
        xp[kk] += h
        xm[kk] -= h
        lp, lm, l0 = f(xp), f(xm), f(x0)
        r, l = (lp - l0) / h, (l0 - lm) / h
        tol = 1e-3 * max(abs(g_ref), 1.0)
        print(f"  path {i} element {k}: device {g_dev:.6g}, f64 {g_ref:.6g}, one-sided f64 at 1e-4 px "
              f"{l:.6g} / {r:.6g}", flush=True)
        assert abs(r - l) > 2 * tol, (i, k, g_dev, g_ref, l, r)
        assert min(l, r) - tol <= g_dev <= max(l, r) + tol, (i, k, g_dev, g_ref, l, r)
    assert len(beyond) <= max(4, errs.size // 100000), len(beyond)
    # central finite differences of the f64 forward (h = 1e-3 px) on seeded
    # (path, parameter) samples; a sample whose [x - h, x + h] holds a kink
    # (the one-sided quotients disagree by more than the curvature of a smooth
    # loss explains at this h) is replaced by another draw
    rng = np.random.default_rng(seed)
    checked, draws = 0, 0
    while checked < fd_samples and draws < 8 * fd_samples:
        draws += 1
        i = int(rng.integers(0, b.n_paths))
        v, p, w, xf = b.path(i)
        p64, xf64, t = p.astype(np.float64), xf.astype(np.float64), targets[i]
        a = b.point_off[i]
        k = int(rng.integers(0, p64.size + 6))
        if k < p64.size:
            h = 1e-3 / float(xf64[0])
            f = lambda z: O.path_loss64(v, z.reshape(-1, 2), w, xf64, t)[1]  # noqa: E731
            x0, gk = p64.reshape(-1).copy(), float(dp[a:a + p64.shape[0]].reshape(-1)[k])
        else:
            k -= p64.size
            h = 1e-3 / 32.0 if k < 4 else 1e-3
            f = lambda z: O.path_loss64(v, p64, w, z, t)[1]  # noqa: E731
            x0, gk = xf64.copy(), float(dx[i, k])
        xp, xm = x0.copy(), x0.copy()
        xp[k] += h
        xm[k] -= h
        lp, lm, l0 = f(xp), f(xm), f(x0)
        c, r, l = (lp - lm) / (2 * h), (lp - l0) / h, (l0 - lm) / h
        if abs(r - l) <= 2e-2 * max(abs(c), 1.0):
            assert abs(gk - c) <= 1e-3 * max(abs(c), 1.0), (i, k, gk, c)
            checked += 1
    print(f"adjoint {b.name}: {checked} finite-difference checks in {draws} draws", flush=True)
    assert checked >= fd_samples


def test_adjoint_configured_batch(gpu_ctx):
    """The differentiable configuration itself: 4,096 paths of 16 cubics
    (32-64 px) on 1024x1024 against a target with control points perturbed by
    N(0, 0.5 px); every gradient element against the f64 oracle adjoint and 64
    finite-difference samples."""
    b = W.adjoint_glyphs()
    assert b.n_paths == 4096
    _check_adjoint(gpu_ctx, b, perturbed(b, 0.5, 900), fd_samples=64, seed=4)


def test_adjoint_side_batch(gpu_ctx):
    """Another seed of the same generator (96 paths)."""
    b = W.adjoint_glyphs(n=96, seed=100)
    _check_adjoint(gpu_ctx, b, perturbed(b, 0.5, 901), fd_samples=32, seed=5)


def test_adjoint_large_paths(gpu_ctx):
    """Paths over one wave's limits (star blobs of 1,000-1,500 cubics, a few
    thousand tiles each) go through the large-path adjoint: every element
    against the f64 oracle adjoint and finite differences on seeded samples."""
    b = W.star_blobs(n=3, seed=71, nmin=1e3, nmax=1.5e3, rmin=120.0, rmax=250.0)
    b.name = "star_blobs(n=3, 1k-1.5k cubics)"
    tb = perturbed(b, 0.5, 902)
    tgt, hosts = oracle_target(gpu_ctx, tb)
    print(f"[adjoint/large] {b.n_paths} blobs, {int(np.diff(b.verb_off).max())} verbs max, "
          f"{tgt.n_tiles} target tiles", flush=True)
    _check_adjoint_against(gpu_ctx, b, tgt, hosts, fd_samples=12, seed=6)
