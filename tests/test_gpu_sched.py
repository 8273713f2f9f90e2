"""The learned tile-cost model's device kernels against numpy.

* inference (the per-bin MLP k_bin_predict runs, through srg_sched_eval):
  4 -> 32 -> 32 -> 1 with ReLU, against a float64 forward pass;
* training (k_mlp_train, through srg_sched_train): mini-batch Adam on the
  squared log2-cost error, batch 256, each thread drawing its sample with its
  own linear congruential generator, restated in float64 here.

Floating-point tolerances are stated in each test: the device sums in fp32
in a fixed order, the reference in float64."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FEAT, HID, BATCH = 4, 32, 256
NPARAMS = HID * FEAT + HID + HID * HID + HID + HID + 1  # 1249


def unpack(p):
    p = np.asarray(p, np.float64)
    o = 0
    w1 = p[o:o + HID * FEAT].reshape(HID, FEAT); o += HID * FEAT
    b1 = p[o:o + HID]; o += HID
    w2 = p[o:o + HID * HID].reshape(HID, HID); o += HID * HID
    b2 = p[o:o + HID]; o += HID
    w3 = p[o:o + HID]; o += HID
    return w1, b1, w2, b2, w3, p[o]


def forward(p, x):
    w1, b1, w2, b2, w3, b3 = unpack(p)
    h1 = np.maximum(x @ w1.T + b1, 0.0)
    h2 = np.maximum(h1 @ w2.T + b2, 0.0)
    return h2 @ w3 + b3, h1, h2


def draws(seed, n, steps):
    """Sample index of each thread at each step: rng = seed*747796405 +
    t*2891336453 + 1, then rng = rng*1664525 + 1013904223 per step, index
    (rng >> 8) % n (uint32 arithmetic)."""
    t = np.arange(BATCH, dtype=np.uint64)
    rng = (np.uint64(seed) * np.uint64(747796405) + t * np.uint64(2891336453) + np.uint64(1)) & np.uint64(0xFFFFFFFF)
    out = []
    for _ in range(steps):
        rng = (rng * np.uint64(1664525) + np.uint64(1013904223)) & np.uint64(0xFFFFFFFF)
        out.append(((rng >> np.uint64(8)) % np.uint64(n)).astype(np.int64))
    return out


def adam_reference(p0, samples, steps, lr, seed):
    """Returns the parameters after `steps` Adam steps from a cleared state,
    and the smallest |gradient| each parameter saw."""
    p = np.asarray(p0, np.float64).copy()
    m = np.zeros_like(p)
    v = np.zeros_like(p)
    gmin = np.full_like(p, np.inf)
    for step, idx in enumerate(draws(seed, len(samples), steps)):
        x, y = samples[idx, :4].astype(np.float64), samples[idx, 4].astype(np.float64)
        w1, b1, w2, b2, w3, b3 = unpack(p)
        o, h1, h2 = forward(p, x)
        dout = 2.0 * (o - y) / BATCH
        d2 = np.where(h2 > 0, dout[:, None] * w3[None, :], 0.0)
        d1 = np.where(h1 > 0, d2 @ w2, 0.0)
        g = np.concatenate([(d1.T @ x).reshape(-1), d1.sum(0), (d2.T @ h1).reshape(-1), d2.sum(0),
                            (dout[:, None] * h2).sum(0), [dout.sum()]])
        gmin = np.minimum(gmin, np.abs(g))
        m = 0.9 * m + 0.1 * g
        v = 0.999 * v + 0.001 * g * g
        tt = step + 1
        p = p - lr * (m / (1 - 0.9 ** tt)) / (np.sqrt(v / (1 - 0.999 ** tt)) + 1e-8)
    return p, gmin


def random_model(rng, scale=0.35):
    p = (scale * rng.standard_normal(NPARAMS)).astype(np.float32)
    p[-1] = 14.0
    return p


def test_inference_matches_numpy():
    """|device - float64| <= 1e-4 * (1 + |out|) on 4,000 rows (fp32
    accumulation of 4 + 32 + 32 terms of magnitude <= ~30)."""
    import softrast_amd as sr
    rng = np.random.default_rng(11)
    ctx = sr.RenderContext()
    p = random_model(rng)
    ctx.set_sched_model(p)
    x = np.concatenate([rng.uniform(0, 16, (4000, 1)), rng.uniform(0, 8, (4000, 2)), rng.uniform(0, 12, (4000, 1))],
                       1).astype(np.float32)
    got = ctx.sched_eval(x)
    want, h1, _ = forward(p, x.astype(np.float64))
    assert np.count_nonzero(h1 > 0) > 1000  # the ReLUs are exercised on both sides
    err = np.abs(got - want)
    assert np.all(err <= 1e-4 * (1.0 + np.abs(want))), float(err.max())
    # a planted mistake fails: swapping two layer-1 input weights changes the output
    q = p.copy()
    q[[0, 1]] = q[[1, 0]]
    assert np.max(np.abs(forward(q, x.astype(np.float64))[0] - got)) > 1e-2
    ctx.Shutdown()


@pytest.mark.parametrize("steps", [1, 6])
def test_adam_steps_match_numpy(steps):
    """Parameters after `steps` device Adam steps equal the float64 restatement
    within 1e-6 + 1e-4 |p|, except where some step's gradient was below 1e-5
    in magnitude (there the update's sign is decided by rounding): those may
    differ by at most 2 * lr * steps."""
    import softrast_amd as sr
    rng = np.random.default_rng(12 + steps)
    ctx = sr.RenderContext()
    p0 = random_model(rng, 0.25)
    n = 3000
    x = np.concatenate([rng.uniform(0, 14, (n, 1)), rng.uniform(0, 6, (n, 2)), rng.uniform(0, 10, (n, 1))], 1)
    y = 9.0 + 0.8 * x[:, 0] + 0.3 * x[:, 1] - 0.2 * x[:, 3] + rng.normal(0, 0.3, n)
    samples = np.concatenate([x, y[:, None]], 1).astype(np.float32)
    lr, seed = 3e-3, 977
    ctx.set_sched_model(p0)
    ctx.sched_train(samples, steps, lr, seed)
    got = ctx.sched_model().astype(np.float64)
    want, gmin = adam_reference(p0, samples, steps, lr, seed)
    firm = gmin >= 1e-5
    assert np.count_nonzero(firm) > NPARAMS // 2
    moved = np.abs(want - p0) > 0.5 * lr
    assert np.count_nonzero(moved & firm) > NPARAMS // 4  # the step did something on both sides
    err = np.abs(got - want)
    assert np.all(err[firm] <= 1e-6 + 1e-4 * np.abs(want[firm])), float(err[firm].max())
    assert np.all(err[~firm] <= 2 * lr * steps + 1e-6)
    # the training lowers the batch loss it optimises
    l0 = np.mean((forward(p0, samples[:, :4].astype(np.float64))[0] - samples[:, 4]) ** 2)
    l1 = np.mean((forward(got, samples[:, :4].astype(np.float64))[0] - samples[:, 4]) ** 2)
    assert l1 < l0
    ctx.Shutdown()
