"""The GPU network (oaph_network_*, FP64, built on this library's
procedures) against the reference's neural training fixtures, and the
activation procedures against their closed forms."""
import numpy as np
import pytest

from _neural_replay import layers_of, load, replay
from oap_amd.api import Activation, CuProceduresApi, DeviceMatrix, Network

pytestmark = pytest.mark.gpu


class _GpuAdapter:
    def __init__(self, fx):
        self.n = Network()
        for neurons, bias, act in layers_of(fx):
            self.n.createLayer(neurons, bias, act)
        self.n.setLearningRate(fx["learning_rate"])

    def set_weights(self, l, w):
        self.n.setHostWeights(w, l)

    def get_weights(self, l):
        return self.n.getHostWeights(l)

    def forward(self, x):
        self.n.setInputs(x)
        self.n.forwardPropagation()

    def accumulate(self, t):
        self.n.setExpected(t)
        self.n.accumulateErrors()

    def backward(self):
        self.n.backPropagation()

    def update(self):
        self.n.updateWeights()

    def post_step(self):
        self.n.postStep()

    def error(self):
        return self.n.calculateError()


@pytest.mark.parametrize("name", ["neural_1hl_1", "neural_2hl_1_3", "neural_4hl_3_2_5_4"])
def test_network_matches_reference_fixture(name):
    fx = load(name)
    worst = replay(_GpuAdapter(fx), fx)
    assert worst <= 1e-13


def test_network_outputs_match_oracle(oracle):
    from oracle.pyoracle import OracleNet
    rng = np.random.default_rng(3)
    shape = [(3, True, Activation.SIGMOID), (5, True, Activation.SIN), (4, True, Activation.TANH),
             (2, False, Activation.LINEAR)]
    g = Network()
    for n, b, a in shape:
        g.createLayer(n, b, a)
    o = OracleNet(oracle, shape)
    for l in range(3):
        w = rng.standard_normal(int(np.prod(g.weightsShape(l))))
        g.setHostWeights(w, l)
        o.set_weights(l, w)
    g.setLearningRate(0.1)
    for it in range(5):
        x, t = rng.standard_normal(3), rng.standard_normal(2)
        g.setInputs(x); g.forwardPropagation(); g.setExpected(t); g.accumulateErrors()
        g.backPropagation()
        yo = o.forward(x); o.accumulate(t); o.backward()
        assert np.allclose(g.getOutputs(), yo, rtol=0, atol=1e-14)
        g.updateWeights(); o.update(0.1)
        for l in range(3):
            assert np.allclose(g.getHostWeights(l), o.get_weights(l), rtol=0, atol=1e-14)


FUNCS = {
    Activation.SIGMOID: lambda o, v: 1 / (1 + np.exp(-v)),
    Activation.DSIGMOID: lambda o, v: (1 / (1 + np.exp(-v))) * (1 - 1 / (1 + np.exp(-v))),
    Activation.MULTIPLY_DSIGMOID: lambda o, v: o * (1 / (1 + np.exp(-v))) * (1 - 1 / (1 + np.exp(-v))),
    Activation.LINEAR: lambda o, v: v,
    Activation.DLINEAR: lambda o, v: np.ones_like(v),
    Activation.TANH: lambda o, v: np.tanh(v),
    Activation.DTANH: lambda o, v: 1 - np.tanh(v) ** 2,
    Activation.MULTIPLY_DTANH: lambda o, v: o * (1 - np.tanh(v) ** 2),
    Activation.SIN: lambda o, v: np.sin(v),
    Activation.DSIN: lambda o, v: np.cos(v),
    Activation.MULTIPLY_DSIN: lambda o, v: o * np.cos(v),
}


@pytest.mark.parametrize("fn", sorted(FUNCS))
def test_activation_full_dims_periodic(api, fn):
    rng = np.random.default_rng(fn)
    cols, rows = 5, 12
    re, im = rng.standard_normal((rows, cols)) * 2, rng.standard_normal((rows, cols)) * 2
    o_re, o_im = rng.standard_normal((rows, cols)), rng.standard_normal((rows, cols))
    f = FUNCS[fn]
    for dims, mask in [(None, np.ones((rows, cols), bool)),
                       ([3, 7], np.add.outer(np.arange(rows) < 7, np.zeros(cols)).astype(bool)
                        & (np.arange(cols) < 3)[None, :]),
                       ([[4, 2], [0, 3]], ((np.arange(rows) % 3) < 2)[:, None]
                        & (np.arange(cols) < 4)[None, :])]:
        m = DeviceMatrix.from_numpy(re, im)
        out = DeviceMatrix.from_numpy(o_re, o_im)
        getattr_name = {v: k for k, v in vars(Activation).items() if isinstance(v, int)}[fn]
        api_fn = {"SIGMOID": "sigmoid", "DSIGMOID": "dsigmoid", "MULTIPLY_DSIGMOID": "multiplyDSigmoid",
                  "LINEAR": "linear", "DLINEAR": "dlinear", "TANH": "tanh", "DTANH": "dtanh",
                  "MULTIPLY_DTANH": "multiplyDTanh", "SIN": "sin", "DSIN": "dsin",
                  "MULTIPLY_DSIN": "multiplyDSin"}[getattr_name]
        getattr(api, api_fn)(out, m, dims)
        gre, gim = out.download()
        for g, old, v in ((gre, o_re, re), (gim, o_im, im)):
            want = np.where(mask, f(old, v), old)
            assert np.allclose(g, want, rtol=1e-15, atol=1e-15), (api_fn, dims)


def test_activation_shape_errors(api):
    from oap_amd.api import OapError
    a, b = DeviceMatrix(3, 4), DeviceMatrix(4, 3)
    with pytest.raises(OapError):
        api.tanh(a, b)
    with pytest.raises(OapError):
        api.tanh(a, DeviceMatrix(3, 4), [[1, 1], [0, 0]])  # period 0
