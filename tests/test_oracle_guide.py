"""Pins the oracle's guiding network (CPU only).

The reference has no path guiding, so the oracle restates the network the C
header specifies.  What fixes it:
  * the analytic gradient of the FP64 loss against central finite
    differences, parameter by parameter, in every layer;
  * the mixture pdf integrates to one over the sphere, and a single dominant
    lobe has the closed-form vMF peak kappa / (2 pi (1 - exp(-2 kappa)));
  * sampled directions are unit vectors whose moments match the pdf's
    (quadrature), so sample_vmf and the pdf describe the same density;
  * features of axis-aligned inputs have closed forms.
"""
import numpy as np
import pytest

from oracle import pyoracle as po

NP = po.GUIDE_NUM_PARAMS
LAYERS = [(160, 64), (64, 64), (64, 64), (64, 64), (64, 40)]


def offsets():
    o, out = 0, []
    for ni, no in LAYERS:
        out.append((o, o + ni * no))
        o += ni * no + no
    return out


def random_params(seed, out_scale=0.3):
    rng = np.random.default_rng(seed)
    p = np.zeros(NP)
    for l, ((ni, no), (ow, ob)) in enumerate(zip(LAYERS, offsets())):
        a = np.sqrt(6 / ni) * (out_scale if l == 4 else 1.0)
        p[ow:ob] = rng.uniform(-a, a, ni * no)
        p[ob:ob + no] = rng.uniform(-0.1, 0.1, no)
    # keep log kappa inside the clamp range so its gradient is live
    ob = offsets()[4][1]
    for j in range(8):
        p[ob + 5 * j + 3] = rng.uniform(0.0, 2.0)
    return p


def random_batch(seed, n):
    rng = np.random.default_rng(seed)
    feats = np.zeros((n, 8), np.float32)
    feats[:, :7] = rng.uniform(0, 1, (n, 7))
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    dirs = np.zeros((n, 4), np.float32)
    dirs[:, :3] = d
    dirs[:, 3] = rng.uniform(0.1, 2.0, n)
    return feats, dirs


def test_gradient_matches_finite_differences():
    p = random_params(1)
    feats, dirs = random_batch(2, 12)
    loss, g = po.guide_loss_grad(p, feats, dirs, emulate_bf16=False)
    assert np.isfinite(loss)
    rng = np.random.default_rng(3)
    idx = []
    for (ni, no), (ow, ob) in zip(LAYERS, offsets()):
        idx += list(rng.integers(ow, ob, 12)) + list(rng.integers(ob, ob + no, 4))
    ob4 = offsets()[4][1]
    idx += [ob4 + 5 * j + k for j in (0, 5) for k in range(5)]  # every head parameter kind
    worst = 0.0
    for k in idx:
        h = 1e-5 * max(1.0, abs(p[k]))
        pp, pm = p.copy(), p.copy()
        pp[k] += h
        pm[k] -= h
        fd = (po.guide_loss_grad(pp, feats, dirs)[0] - po.guide_loss_grad(pm, feats, dirs)[0]) / (2 * h)
        err = abs(fd - g[k]) / max(abs(fd), abs(g[k]), 1e-6)
        worst = max(worst, err)
        assert err < 2e-4, (k, fd, g[k])
    print("worst relative FD error", worst)


def test_gradient_of_weighted_batch_is_normalised():
    """Doubling every weight leaves the normalised loss and gradient unchanged."""
    p = random_params(4)
    feats, dirs = random_batch(5, 8)
    l1, g1 = po.guide_loss_grad(p, feats, dirs)
    d2 = dirs.copy()
    d2[:, 3] *= 2
    l2, g2 = po.guide_loss_grad(p, feats, d2)
    assert abs(l1 - l2) < 1e-12 * abs(l1) + 1e-12
    np.testing.assert_allclose(g1, g2, rtol=1e-10, atol=1e-14)


def fib_sphere(n):
    i = np.arange(n) + 0.5
    z = 1 - 2 * i / n
    r = np.sqrt(1 - z * z)
    ph = np.pi * (1 + 5 ** 0.5) * i
    return np.stack([r * np.cos(ph), r * np.sin(ph), z], 1)


def random_raw(seed, kmax=2.5):
    rng = np.random.default_rng(seed)
    raw = np.zeros(40)
    for j in range(8):
        raw[5 * j:5 * j + 3] = rng.normal(size=3)
        raw[5 * j + 3] = rng.uniform(-1, kmax)
        raw[5 * j + 4] = rng.normal()
    return raw


@pytest.mark.parametrize("seed", [0, 1])
def test_mixture_pdf_integrates_to_one(seed):
    raw = random_raw(seed)
    n = 200_000
    dirs = fib_sphere(n)
    pdf = po.guide_pdf(raw, dirs)
    integral = pdf.mean() * 4 * np.pi
    assert abs(integral - 1) < 2e-3, integral


def test_dominant_lobe_peak_closed_form():
    raw = np.zeros(40)
    for j in range(8):
        raw[5 * j:5 * j + 3] = [0, 0, 1] if j == 3 else [1, 0, 0]
        raw[5 * j + 3] = np.log(5.0)
        raw[5 * j + 4] = 30.0 if j == 3 else 0.0
    k = 5.0
    peak = po.guide_pdf(raw, np.array([[0, 0, 1.0]]))[0]
    np.testing.assert_allclose(peak, k / (2 * np.pi * (1 - np.exp(-2 * k))), rtol=1e-9)
    # kappa clamp: log kappa beyond 8 behaves as 8
    raw[5 * 3 + 3] = 12.0
    k = np.exp(8.0)
    peak = po.guide_pdf(raw, np.array([[0, 0, 1.0]]))[0]
    np.testing.assert_allclose(peak, k / (2 * np.pi), rtol=1e-9)


def test_samples_follow_the_pdf():
    raw = random_raw(7)
    rng = np.random.default_rng(8)
    n = 100_000
    d = po.guide_sample(raw, rng.uniform(0, 1, (n, 3)))
    np.testing.assert_allclose(np.linalg.norm(d, axis=1), 1, atol=1e-9)
    # moments against quadrature
    q = fib_sphere(400_000)
    pdf = po.guide_pdf(raw, q) * (4 * np.pi / len(q))
    for f in (lambda v: v[:, 0], lambda v: v[:, 1], lambda v: v[:, 2], lambda v: v[:, 0] * v[:, 2],
              lambda v: v[:, 1] ** 2):
        mc = f(d)
        expect = (f(q) * pdf).sum()
        se = mc.std() / np.sqrt(n)
        assert abs(mc.mean() - expect) < 5 * se + 1e-4, (mc.mean(), expect, se)


def test_features_closed_forms():
    lo, hi = np.array([-1.0, 0, 2]), np.array([1.0, 4, 6])
    f = po.guide_features([0, 1, 6], [0, 0, 1], [-1, 0, 0], lo, hi)
    np.testing.assert_allclose(f[:3], [0.5, 0.25, 1.0])
    np.testing.assert_allclose(f[3], 0.0, atol=1e-7)      # normal +z: theta 0
    np.testing.assert_allclose(f[5:7], [0.5, 0.5])        # -x: theta pi/2, phi pi
    f = po.guide_features([5, -3, 0], [0, -1, 0], [0, 0, -1], lo, hi)
    np.testing.assert_allclose(f[:3], [1.0, 0.0, 0.0])    # clamped to the box
    np.testing.assert_allclose(f[3:5], [0.5, 0.75])       # -y: phi 3 pi / 2
    np.testing.assert_allclose(f[5], 1.0)


def test_bf16_emulation_stays_close():
    p = random_params(9)
    feats, _ = random_batch(10, 32)
    a = po.guide_forward(p, feats, emulate_bf16=False)
    b = po.guide_forward(p, feats, emulate_bf16=True)
    # BF16 carries 8 significand bits: the outputs agree to a few 1e-2
    assert np.abs(a - b).max() < 0.05 * (np.abs(a).max() + 1)
    assert np.abs(a - b).max() > 0  # the emulation does round
