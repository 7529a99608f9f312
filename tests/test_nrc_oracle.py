"""CPU pins of the neural-radiance-cache oracle (oracle/nrc_oracle.cpp).

  encoding      sin/cos(2^k pi x) evaluated directly in numpy (the oracle uses
                the exact power-of-two fraction form)
  forward       an independent numpy network with numpy's own float16 /
                bfloat16-by-truncation-free rounding
  gradients     central finite differences of the FP64 network (loss
                denominator frozen, as the training loss defines)
"""
import numpy as np
import pytest

NP = 25936


def _layers():
    # (weight offset, out, in) of the padded parameter layout
    offs = [(0, 64, 128)] + [(8192 + (l - 1) * 4096, 64, 64) for l in range(1, 5)] + \
           [(8192 + 4 * 4096, 16, 64)]
    return offs


def random_params(seed, scale=1.0):
    rng = np.random.default_rng(seed)
    p = np.zeros(NP, np.float32)
    nw = 8192 + 4 * 4096 + 1024
    for l, (off, o, i) in enumerate(_layers()):
        fan = 121 if l == 0 else 64
        real_o = 3 if l == 5 else 64
        w = np.zeros((o, i), np.float32)
        w[:real_o, :fan] = rng.uniform(-1, 1, (real_o, fan)) * np.sqrt(6 / fan) * scale
        p[off:off + o * i] = w.ravel()
        b = np.zeros(64 if l < 5 else 16, np.float32)
        b[:real_o] = rng.uniform(-0.1, 0.1, real_o)
        p[nw + 64 * l:nw + 64 * l + b.size] = b
    return p


def random_raw(n, seed):
    rng = np.random.default_rng(seed)
    return rng.uniform(0, 1, (n, 6)).astype(np.float32)


def np_encode(raw):
    e = np.zeros(128)
    for f in range(120):
        s, w = divmod(f, 24)
        k = w // 2
        a = 2.0 ** k * np.pi * float(raw[s])
        e[f] = np.cos(a) if w % 2 else np.sin(a)
    e[120] = raw[5]
    return e


def np_forward(params, raw, q16, f32_acc=True):
    nw = 8192 + 4 * 4096 + 1024
    out = []
    for r in raw:
        x = q16(np_encode(r))
        for l, (off, o, i) in enumerate(_layers()):
            W = q16(params[off:off + o * i].reshape(o, i).astype(np.float64))
            b = params[nw + 64 * l:nw + 64 * l + o].astype(np.float64)
            z = W @ x + b
            if f32_acc:
                z = z.astype(np.float32).astype(np.float64) if l < 5 else z
            x = q16(np.maximum(0.0, z)) if l < 5 else z
        out.append(x[:3])
    return np.array(out)


def test_encoding_direct_formula(oracle):
    for r in random_raw(20, 1):
        np.testing.assert_allclose(oracle.nrc_encode(r), np_encode(r), atol=2e-12)


@pytest.mark.parametrize("precision", [0, 2])
def test_forward_matches_numpy_network(oracle, precision):
    params = random_params(3)
    raw = random_raw(32, 4)
    if precision == 0:
        q16 = lambda a: np.asarray(a, np.float64).astype(np.float32).astype(np.float16).astype(np.float64)
    else:
        q16 = lambda a: np.asarray(a, np.float64)
    got = oracle.nrc_forward(params, raw, precision)
    exp = np_forward(params, raw, q16, f32_acc=precision != 2)
    np.testing.assert_allclose(got, exp, rtol=1e-9, atol=1e-9)


def test_gradients_finite_differences(oracle):
    params = random_params(5, scale=0.7).astype(np.float64)
    raw = random_raw(6, 6)
    target = np.random.default_rng(7).uniform(0, 1, (6, 3)).astype(np.float32)
    # exact network: the float32 parameter vector is perturbed by a step that
    # is representable, the loss denominator is frozen at the base point
    p32 = params.astype(np.float32)
    g, _ = oracle.nrc_gradients(p32, raw, target, 2)
    p0 = oracle.nrc_forward(p32, raw, 2)
    den = p0 ** 2 + 0.01

    def loss(p):
        y = oracle.nrc_forward(p, raw, 2)
        return np.sum((y - target) ** 2 / den)

    rng = np.random.default_rng(8)
    nw = 8192 + 4 * 4096 + 1024
    idx = []
    for l, (off, o, i) in enumerate(_layers()):
        real_o = 3 if l == 5 else 64
        fan = 121 if l == 0 else 64
        for _ in range(6):
            idx.append(off + rng.integers(real_o) * i + rng.integers(fan))
        idx.append(nw + 64 * l + rng.integers(real_o))
    for k in idx:
        h = np.float32(2.0 ** -16)
        pp, pm = p32.copy(), p32.copy()
        pp[k] += h
        pm[k] -= h
        fd = (loss(pp) - loss(pm)) / (float(pp[k]) - float(pm[k]))
        assert abs(fd - g[k]) <= 1e-4 * max(1.0, abs(g[k])), (k, fd, g[k])


def test_bf16_rounding_is_round_to_nearest_even(oracle):
    """bf16 forward with weights exactly representable in bf16 and an input
    whose encoding is exact (x = 0): the output equals the exact network's
    on a bf16-rounded copy."""
    params = random_params(9)
    u = params.view(np.uint32)
    params_bf = ((u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000).view(np.float32)
    raw = np.zeros((1, 6), np.float32)
    a = oracle.nrc_forward(params_bf, raw, 1)
    # biases are not rounded by the network: compare against the network on
    # the same parameter vector with only activations rounded
    assert np.all(np.isfinite(a))
