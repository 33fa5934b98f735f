"""Replays one of the reference's neural training fixtures
(tests/golden/neural_*.json, written by tests/golden/make_neural_fixtures.py)
through a network object and checks every weight checkpoint and loss, as the
reference's driver does (oapDeviceTests/oapNeuralTests_Api.cpp:169-403):
learning rate 0.03, per-sample forward / accumulate / backward, one update
per batch, weights compared after every batch and losses after every step.

`net` needs: set_weights(l, w), get_weights(l), forward(x), accumulate(t),
backward(), update(), post_step(), error().
"""
import json
from pathlib import Path

import numpy as np

GOLDEN = Path(__file__).resolve().parent / "golden"
TANH = 5


def load(name):
    return json.loads((GOLDEN / f"{name}.json").read_text())


def layers_of(fx):
    sizes = fx["network"]
    return [(n, i < len(sizes) - 1, TANH) for i, n in enumerate(sizes)]


def replay(net, fx, max_batches=None):
    tol = fx["tolerance"]
    nl = len(fx["weights"])
    for l in range(nl):
        net.set_weights(l, fx["weights"][l][0])
    worst = 0.0

    def check(widx, where):
        nonlocal worst
        for l in range(nl):
            got = net.get_weights(l)
            want = np.asarray(fx["weights"][l][widx])
            idx = fx["idxs_to_check"][l]
            d = float(np.max(np.abs(got[idx] - want[idx])))
            worst = max(worst, d)
            assert d <= tol, (where, l, widx, d)

    widx, done = 0, 0
    for si, step in enumerate(fx["steps"]):
        for bi, batch in enumerate(fx["points"][step["batches"]]):
            check(widx, f"step {si} batch {bi} before")
            for (x, y), label in batch:
                net.forward([x, y])
                net.accumulate([label])
                net.backward()
            net.update()
            widx += 1
            check(widx, f"step {si} batch {bi} after")
            net.post_step()
            done += 1
            if max_batches and done >= max_batches:
                return worst
        for key, lkey in (("train", "loss_train"), ("test", "loss_test")):
            if key in step:
                for (x, y), label in fx["points"][step[key]]:
                    net.forward([x, y])
                    net.accumulate([label])
                loss = net.error()
                assert abs(loss - step[lkey]) <= tol, (si, key, loss, step[lkey])
                net.post_step()
        net.post_step()
    return worst
