"""The oracle's FP64 network (orc_net_*) against the reference's neural
training fixtures (every weight checkpoint and loss to the reference's own
1e-13), and its back propagation against finite differences."""
import numpy as np
import pytest

from _neural_replay import layers_of, load, replay
from oracle.pyoracle import OracleNet

FIXTURES = ["neural_1hl_1", "neural_2hl_1_3", "neural_4hl_3_2_5_4"]


class _OracleAdapter:
    def __init__(self, net):
        self.n = net

    def set_weights(self, l, w):
        self.n.set_weights(l, w)

    def get_weights(self, l):
        return self.n.get_weights(l)

    def forward(self, x):
        self.n.forward(np.asarray(x, dtype=np.float64))

    def accumulate(self, t):
        self.n.accumulate(t)

    def backward(self):
        self.n.backward()

    def update(self):
        self.n.update(0.03)

    def post_step(self):
        self.n.post_step()

    def error(self):
        return self.n.error()


@pytest.mark.parametrize("name", FIXTURES)
def test_oracle_network_matches_reference_fixture(oracle, name):
    fx = load(name)
    net = OracleNet(oracle, layers_of(fx))
    worst = replay(_OracleAdapter(net), fx)
    assert worst <= 1e-13


@pytest.mark.parametrize("act", [0, 3, 5, 8])
def test_oracle_backprop_finite_differences(oracle, act):
    rng = np.random.default_rng(act)
    shape = [(3, True, act), (4, True, act), (2, True, act), (2, False, act)]
    net = OracleNet(oracle, shape)
    ws = [rng.standard_normal(np.prod(net.wshape(l))) * 0.5 for l in range(3)]
    for l, w in enumerate(ws):
        net.set_weights(l, w)
    x, t = rng.standard_normal(3), rng.standard_normal(2)

    def loss():
        net.post_step()
        net.forward(x)
        net.accumulate(t)
        return net.error() * 2  # sum of 0.5 e^2 over the 2 outputs

    net.post_step()
    net.forward(x)
    net.accumulate(t)
    net.backward()
    grads = [net.gradient(l) for l in range(3)]
    h = 1e-6
    for l in range(3):
        for i in range(ws[l].size):
            wp = ws[l].copy(); wp[i] += h
            wm = ws[l].copy(); wm[i] -= h
            net.set_weights(l, wp); fp = loss()
            net.set_weights(l, wm); fm = loss()
            net.set_weights(l, ws[l])
            fd = (fp - fm) / (2 * h)
            assert abs(fd - grads[l][i]) <= 1e-7 * max(1.0, abs(fd)), (l, i, fd, grads[l][i])
