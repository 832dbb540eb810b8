"""The sizing MLP: the numpy checker pinned by finite differences (CPU), the
GPU forward / gradient / Adam kernels against it, training on logged exact
counts, and learned sizing end to end (identical pixels, fewer host round
trips and no counting pass, fewer fallbacks than history sizing)."""
import numpy as np
import pytest

import checkers
import cost_model_ref as R


def _data(n, seed):
    rng = np.random.default_rng(seed)
    X = rng.uniform(0, 8, (n, R.IN)).astype(np.float32)
    T = np.stack([X[:, :4].sum(1) * 0.5 + np.sin(X[:, 4]), X[:, 5] * 0.3 + 1.0], 1).astype(np.float32)
    return X, T


def test_reference_gradient_matches_finite_differences():
    X, T = _data(32, 1)
    p = R.init_params(3).astype(np.float64)
    _, g = R.loss_and_grad(X, T, p)
    rng = np.random.default_rng(5)
    idx = rng.choice(R.N_PARAMS, 60, replace=False)
    eps = 1e-6
    for i in idx:
        dp = np.zeros_like(p)
        dp[i] = eps
        num = (R.loss_and_grad(X, T, p + dp)[0] - R.loss_and_grad(X, T, p - dp)[0]) / (2 * eps)
        assert abs(num - g[i]) <= 1e-6 + 1e-5 * abs(num), (i, num, g[i])


def test_sizing_modes_are_accepted_and_others_rejected():
    import blend2d_amd
    try:
        for mode in (blend2d_amd.SIZING_LEARNED, blend2d_amd.SIZING_HISTORY, blend2d_amd.SIZING_EXACT):
            blend2d_amd.set_sizing(mode)
        with pytest.raises(blend2d_amd.BLError) as e:
            blend2d_amd.set_sizing(3)
        assert e.value.code == 0x10001      # BL_ERROR_INVALID_VALUE
    finally:
        blend2d_amd.set_sizing(blend2d_amd.SIZING_EXACT)


def test_reference_adam_matches_closed_form_first_step():
    # the first Adam step moves every parameter by lr * sign(g) (bias-corrected)
    p = np.zeros(4)
    g = np.array([0.5, -2.0, 1e-3, -1e-4])
    p1, _, _ = R.adam(p, g, np.zeros(4), np.zeros(4), 1, 0.01)
    np.testing.assert_allclose(p1, -0.01 * np.sign(g), rtol=1e-4)


@pytest.fixture(scope="module")
def amd():
    import blend2d_amd
    if blend2d_amd.lib().bl_amd_device_count() < 1:
        pytest.skip("no HIP device")
    return blend2d_amd


@pytest.mark.gpu
def test_gpu_forward_matches_reference(amd):
    X, _ = _data(1000, 2)
    p = R.init_params(4)
    got = amd.cost_model_forward(X, p)
    want, _ = R.forward(X, p)
    np.testing.assert_allclose(got, want, rtol=1e-4, atol=1e-4)


@pytest.mark.gpu
@pytest.mark.parametrize("n", [1, 63, 64, 65, 1000])
def test_gpu_gradient_matches_reference(amd, n):
    X, T = _data(n, 3 + n)
    p = R.init_params(6)
    g, loss = amd.cost_model_grad(X, T, p)
    want_loss, want = R.loss_and_grad(X, T, p)
    assert abs(loss - want_loss) <= 1e-4 * max(1.0, abs(want_loss))
    np.testing.assert_allclose(g, want, rtol=2e-3, atol=1e-5 * max(1.0, np.abs(want).max()))


@pytest.mark.gpu
def test_gpu_adam_training_matches_reference_steps(amd):
    # three full-batch steps on the GPU against the numpy Adam
    X, T = _data(256, 9)
    p0 = R.init_params(7)
    got, losses = amd.cost_model_train(X, T, p0, epochs=3, batch=256, lr=1e-3)
    p = p0.astype(np.float64)
    m1 = np.zeros_like(p)
    m2 = np.zeros_like(p)
    small = np.zeros(len(p), dtype=bool)
    for step in range(1, 4):
        _, g = R.loss_and_grad(X, T, p)
        # Adam's update is ~lr * g / |g|: where |g| is at rounding level the
        # sign of a float32 gradient is arbitrary, so those are not compared
        small |= np.abs(g) < 1e-4 * np.abs(g).max()
        p, m1, m2 = R.adam(p, g, m1, m2, step, 1e-3)
    keep = ~small
    assert keep.mean() > 0.5
    np.testing.assert_allclose(got[keep], p[keep], rtol=1e-3, atol=1e-5)
    assert losses[2] < losses[0]


def _render(amd, sc):
    images = []
    for fmt, px in sc.images:
        im = amd.BLImage(px.shape[1], px.shape[0], fmt)
        im.upload(px)
        images.append(im)
    img = amd.BLImage(sc.width, sc.height, sc.format)
    img.fill_raw(sc.clear_value)
    ctx = amd.BLContext(img)
    amd.replay_into(ctx, sc, images)
    ctx.flush()
    st = ctx.last_stats()
    out = img.download()
    return ctx, out, st


@pytest.fixture(scope="module")
def trained(amd):
    from blend2d_amd import scene as S, train_cost_model as TC
    X, T = TC.collect([11, 12], 24, 128, 768, 64, 3000)
    Xv, Tv = TC.collect([13], 12, 128, 768, 64, 3000)
    p0 = TC.init_params(0)
    params, losses = TC.train(X, T, epochs=30, batch=1024, lr=3e-3)
    return dict(X=X, T=T, Xv=Xv, Tv=Tv, p0=p0, params=params, losses=losses)


@pytest.mark.gpu
def test_training_on_logged_counts_predicts_held_out_counts(amd, trained):
    assert len(trained["X"]) > 500
    losses = trained["losses"]
    assert losses[-1] < 0.2 * losses[0]
    Yv = amd.cost_model_forward(trained["Xv"], trained["params"])
    err = np.abs(Yv - trained["Tv"])
    # the untrained network against the same held-out rows
    err0 = np.abs(amd.cost_model_forward(trained["Xv"], trained["p0"]) - trained["Tv"])
    assert np.median(err[:, 0]) < 0.35 and np.median(err[:, 1]) < 0.35      # within ~27 %
    assert np.median(err[:, 0]) < 0.25 * np.median(err0[:, 0])


@pytest.mark.gpu
def test_learned_sizing_keeps_pixels_and_skips_the_counting_pass(amd, trained):
    from blend2d_amd import scene as S
    amd.cost_model_set(trained["params"])
    try:
        scenes = S.cost_canvases(21, 8, 128, 768, 64, 3000)
        stats = {}
        for mode in (amd.SIZING_EXACT, amd.SIZING_LEARNED, amd.SIZING_HISTORY):
            amd.set_sizing(mode)
            outs, sts = [], []
            for sc in scenes:
                ctx, out, st = _render(amd, sc)
                ctx.end()
                outs.append(out)
                sts.append(st)
            stats[mode] = (outs, sts)
        exact_outs, exact_sts = stats[amd.SIZING_EXACT]
        for mode in (amd.SIZING_LEARNED, amd.SIZING_HISTORY):
            outs, sts = stats[mode]
            for a, b in zip(outs, exact_outs):
                assert (a == b).all()
        learned = stats[amd.SIZING_LEARNED][1]
        history = stats[amd.SIZING_HISTORY][1]
        # 3 host round trips per frame either way (exact: segment count,
        # binned count, end; learned: predicted totals, render-grid counts,
        # end), but learned sizing runs no counting pass of the edge builder
        assert all(s["syncs"] == 3 for s in exact_sts)
        ok = [s for s in learned if not s["fallback"]]
        assert len(ok) >= len(learned) - 1
        assert all(s["syncs"] == 3 and s["sizing"] == amd.SIZING_LEARNED for s in ok)
        # history sizing (no model) reuses the last frame's totals and falls
        # back whenever a frame is larger than its predecessor
        assert sum(s["fallback"] for s in learned) < max(1, sum(s["fallback"] for s in history))
    finally:
        amd.set_sizing(amd.SIZING_EXACT)
        amd.cost_model_set(None)


@pytest.mark.gpu
def test_untrained_model_falls_back_to_exact_counting(amd):
    from blend2d_amd import scene as S
    import cost_model_ref as Rf
    sc = S.cost_canvases(31, 1, 256, 256, 128, 3000)[0]   # 8287 edge segments
    amd.set_sizing(amd.SIZING_EXACT)
    _, want, _ = _render(amd, sc)
    # a model that predicts nothing: every pool overflows.  On a fresh
    # context the pool buffers hold stale device memory, so this also checks
    # that the slots an overflowing item owns below the capacity are marked
    # unused before binning reads them.
    zero = np.zeros(Rf.N_PARAMS, dtype=np.float32)
    zero[-2:] = -10.0
    amd.cost_model_set(zero)
    amd.set_sizing(amd.SIZING_LEARNED)
    try:
        ctx, got, st = _render(amd, sc)
        ctx.end()
    finally:
        amd.set_sizing(amd.SIZING_EXACT)
        amd.cost_model_set(None)
    assert st["fallback"] == 1
    assert (got == want).all()


@pytest.mark.gpu
def test_learned_sizing_of_a_batched_frame_matches_the_oracle(amd):
    # the shipped model sizes one batched frame of glyph and cost canvases
    from blend2d_amd import scene as S, train_cost_model as TC
    params = TC.load()
    if params is None:
        pytest.skip("no trained model parameters in the package")
    scenes = S.glyph_canvases(9, 24, 60) + S.cost_canvases(9, 4, 128, 512, 32, 3000)
    amd.cost_model_set(params)
    amd.set_sizing(amd.SIZING_LEARNED)
    try:
        outs, errs = amd.render_scenes_batched(scenes)
    finally:
        amd.set_sizing(amd.SIZING_EXACT)
        amd.cost_model_set(None)
    for sc, out, err in zip(scenes, outs, errs):
        want, oerr = checkers.render_oracle(sc)
        assert err == oerr
        assert checkers.diff_info(out, want)[0] == 0
