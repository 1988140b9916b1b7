"""The learned denoiser (include/pt_denoise.h): the CPU oracle pinned by
finite differences, then the device network's forward pass, loss and
gradients against the oracle, training, and the gain of the trained network
over the plain render at the same sample count."""
import os

import numpy as np
import pytest

import denoise_oracle as D
import ptrace
from conftest import ROOT


def _params(rng, last_scale=0.3):
    p = np.zeros(D.parameter_count())
    o = 0
    for k, (a, b, ca, cb, co, *_rest) in enumerate(D.LAYERS):
        nw = co * 9 * (ca + cb)
        sd = np.sqrt(2 / (9 * (ca + cb))) * (last_scale if k == len(D.LAYERS) - 1 else 1)
        p[o:o + nw] = rng.normal(0, sd, nw)
        p[o + nw:o + nw + co] = rng.normal(0, 0.05, co)
        o += nw + co
    return p


def _data(rng, n, h, w):
    feat = np.concatenate([rng.gamma(1.0, 0.5, (n, h, w, 3)), rng.uniform(0, 1, (n, h, w, 3)),
                           rng.normal(0, 1, (n, h, w, 3))], -1)
    return feat, rng.gamma(1.0, 0.5, (n, h, w, 3))


def test_oracle_gradients_match_finite_differences():
    rng = np.random.default_rng(0)
    params = _params(rng)
    feat, tgt = _data(rng, 1, 16, 16)
    _, g = D.loss_and_gradients(params, feat, tgt, rounding=False)
    offs = np.cumsum([0] + [co * 9 * (ca + cb) + co for _, _, ca, cb, co, *_ in D.LAYERS])
    worst = 0.0
    for k in range(len(D.LAYERS)):
        for i in rng.integers(offs[k], offs[k + 1], 3):
            e = 1e-6
            p1, p2 = params.copy(), params.copy()
            p1[i] += e
            p2[i] -= e
            fd = (D.loss_and_gradients(p1, feat, tgt, False)[0] - D.loss_and_gradients(p2, feat, tgt, False)[0]) / (2 * e)
            worst = max(worst, abs(fd - g[i]) / max(1e-9, abs(fd) + abs(g[i])))
    assert worst < 1e-4


def test_oracle_detects_a_flipped_tap():
    # A plausible mistake (taps applied mirrored) must change the output.
    rng = np.random.default_rng(1)
    params = _params(rng)
    feat, _ = _data(rng, 1, 16, 16)
    out, _ = D.forward(params, feat, rounding=False)
    mirrored = params.copy()
    o = 0
    for a, b, ca, cb, co, *_ in D.LAYERS:
        nw = co * 9 * (ca + cb)
        w = mirrored[o:o + nw].reshape(co, 9, ca + cb)
        mirrored[o:o + nw] = w[:, ::-1, :].reshape(-1)
        o += nw + co
    out2, _ = D.forward(mirrored, feat, rounding=False)
    assert np.abs(out - out2).max() > 1e-3


def test_bf16_rounding_is_nearest_even():
    x = np.array([1.0, 1.00390625, 1.01171875, 1.0078125 + 2 ** -16, -3.5, 0.0, np.inf], np.float32)
    # 1 + 2^-8 is a tie -> even (1.0); 1 + 3*2^-8 a tie -> 1 + 2^-6; 1 + 2^-7 + 2^-16 -> 1 + 2^-7.
    assert list(D.bf16(x)) == [1.0, 1.0, 1.015625, 1.0078125, -3.5, 0.0, np.inf]


def test_library_declares_the_denoiser():
    import ctypes
    import re
    text = open(os.path.join(ROOT, "include", "pt_denoise.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = sorted(set(re.findall(r"\b(ptd_[a-z_]+)\s*\(", text)))
    lib = ctypes.CDLL(ptrace.LIBRARY_PATH)
    assert names and all(hasattr(lib, n) for n in names)
    import denoise
    assert set(names) == set(denoise.PTD_FUNCTIONS)
    assert denoise.Denoiser.parameter_count() == D.parameter_count()


# ---------------------------------------------------------------- device
def _rel_l2(a, b):
    return float(np.linalg.norm(a - b) / max(1e-30, np.linalg.norm(b)))


@pytest.mark.gpu
def test_device_forward_matches_oracle():
    import denoise
    rng = np.random.default_rng(2)
    params = _params(rng, last_scale=1.0).astype(np.float32)
    feat, _ = _data(rng, 2, 32, 48)
    net = denoise.Denoiser(seed=5)
    net.set_parameters(params)
    assert np.array_equal(net.parameters(), params)
    got = net.forward(feat)
    want = D.denoise(params.astype(np.float64), feat.astype(np.float32), rounding=True)
    # bf16 activations: the two differ where a value sits on a rounding boundary.
    assert _rel_l2(got, want) < 2e-2
    assert np.mean(np.abs(got - want) <= 2e-2 * np.abs(want) + 1e-3) > 0.98


@pytest.mark.gpu
def test_device_gradients_match_oracle():
    import denoise
    rng = np.random.default_rng(3)
    params = _params(rng).astype(np.float32)
    feat, tgt = _data(rng, 2, 32, 32)
    net = denoise.Denoiser(seed=5)
    net.set_parameters(params)
    loss, g = net.loss_and_gradients(feat, tgt)
    want_loss, want_g = D.loss_and_gradients(params.astype(np.float64), feat.astype(np.float32), tgt, rounding=True)
    assert abs(loss - want_loss) <= 1e-2 * abs(want_loss)
    o = 0
    for k, (a, b, ca, cb, co, *_r) in enumerate(D.LAYERS):
        n = co * 9 * (ca + cb) + co
        assert _rel_l2(g[o:o + n], want_g[o:o + n]) < 5e-2, k
        o += n


@pytest.mark.gpu
def test_training_reduces_the_loss():
    import denoise
    rng = np.random.default_rng(4)
    # A learnable map: targets are a smoothed version of the noisy input.
    feat, _ = _data(rng, 16, 32, 32)
    rad = feat[..., :3]
    smooth = (rad + np.roll(rad, 1, 1) + np.roll(rad, -1, 1) + np.roll(rad, 1, 2) + np.roll(rad, -1, 2)) / 5
    net = denoise.Denoiser(seed=9)
    net.set_training_data(feat, smooth)
    ids = rng.integers(0, 16, 60 * 4)
    losses, ms = net.train(ids, 4, 1e-3)
    assert np.all(np.isfinite(losses)) and ms > 0
    assert losses[-10:].mean() < 0.7 * losses[:5].mean()


@pytest.mark.gpu
def test_trained_network_beats_the_plain_render():
    # Held-out view (a room seed not used in training): the shipped weights
    # at 4 spp against the plain 4-spp image, both scored against 512 spp.
    import denoise
    if not os.path.exists(denoise.WEIGHTS_PATH):
        pytest.skip("no trained weights shipped")
    net = denoise.Denoiser(weights=denoise.WEIGHTS_PATH)
    scene, cam = ptrace.cornell(424242, 128, 128)
    ref, _, _ = denoise.render_features(scene, cam, 128, 128, 77, 512)
    r = ptrace.Renderer(scene, cam, 128, 128, 5, False, ptrace.DEBUG_SHADED, True, 8)
    r.enable_aovs()
    r.add_samples(4)
    out, ms = net.denoise_renderer(r)
    plain = denoise.relmse(r.image(), ref)
    den = denoise.relmse(out, ref)
    assert den < 0.5 * plain, (den, plain)
    # the same result delivered into a page-locked buffer
    buf = ptrace.pinned((128 * 128, 3))
    got, _ = net.denoise_renderer(r, out=buf)
    assert got is buf and np.array_equal(buf, out)


@pytest.mark.gpu
def test_feature_buffers_leave_radiance_unchanged():
    scene, cam = ptrace.cornell(3, 64, 48)
    a = ptrace.Renderer(scene, cam, 64, 48, 9, False, ptrace.DEBUG_SHADED, True, 8)
    b = ptrace.Renderer(scene, cam, 64, 48, 9, False, ptrace.DEBUG_SHADED, True, 8)
    b.enable_aovs()
    a.add_samples(3)
    b.add_samples(3)
    assert np.array_equal(a.image(), b.image()) and a.rays_traced() == b.rays_traced()
    albedo, normal = b.aovs()
    assert np.all(albedo >= 0) and np.all(albedo <= 1 + 1e-6)
    lens = np.linalg.norm(normal, axis=1)
    # Unit normals averaged over a pixel's samples; zero where every camera ray escapes.
    assert np.all(lens <= 1 + 1e-5) and np.mean(lens > 0.99) > 0.3
    assert np.all(albedo[lens == 0] == 0)
    with pytest.raises(ptrace.InvalidArgument):
        b.enable_aovs()  # samples already added
    b.reset(cam)
    b.enable_aovs(False)
    with pytest.raises(ptrace.InvalidArgument):
        b.aovs()
