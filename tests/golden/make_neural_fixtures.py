"""Writes tests/golden/neural_*.json from the reference's neural training
fixtures (oapDeviceTestsData/oapNeuralTests_Data_*.h): network shape,
initial weights, the sample batches, the expected weights after every batch
update and the expected train/test losses after every step.  The reference
test driver that consumes them is oapDeviceTests/oapNeuralTests_Api.cpp:90-403
(learning rate 0.03, tanh layers, per-sample forward + backward, one weight
update per batch, checked to 1e-13).

Run in a container that has the reference checkout at /root/reference:
    python tests/golden/make_neural_fixtures.py
Only data values are extracted (numbers and the names that tie them
together); nothing of the reference's code is stored.
"""
import json
import re
from pathlib import Path

SRC = Path("/root/reference/oapDeviceTestsData")
OUT = Path(__file__).resolve().parent
DATASETS = {
    "neural_1hl_1": ("oapNeuralTests_Data_1HL_1.h", "Test_1"),
    "neural_2hl_1_3": ("oapNeuralTests_Data_2HL_1_3_Test_1.h", "Test_1"),
    "neural_4hl_3_2_5_4": ("oapNeuralTests_Data_4HL_3_2_5_4.h", None),
}


def braces_to_json(text: str):
    t = re.sub(r"//[^\n]*", "", text)
    t = t.replace("{", "[").replace("}", "]")
    t = re.sub(r",\s*]", "]", t)
    return json.loads(t)


def block(src: str, decl: str) -> str:
    """The brace-balanced initializer following `decl =`."""
    i = src.index(decl)
    j = src.index("{", i)
    depth = 0
    for k in range(j, len(src)):
        if src[k] == "{":
            depth += 1
        elif src[k] == "}":
            depth -= 1
            if depth == 0:
                return src[j:k + 1]
    raise ValueError(decl)


def extract(fname: str, ns):
    src = (SRC / fname).read_text()
    if ns is not None:
        src = src[src.index(f"namespace {ns}"):]
    hidden = braces_to_json(block(src, "g_networkInfo"))
    idxs = braces_to_json(block(src, "g_idxsToCheck"))
    vecs = {m.group(1): [float(x) for x in m.group(2).split(",") if x.strip()]
            for m in re.finditer(r"const std::vector<floatt> (\w+) = \{([^}]*)\};", src)}
    scalars = {m.group(1): float(m.group(2))
               for m in re.finditer(r"const floatt (\w+) = ([-0-9.eE+]+);", src)}
    points = {}
    for m in re.finditer(r"const (Points|Batches) (\w+) =", src):
        points[m.group(2)] = braces_to_json(block(src, m.group(0)[6:-2].split(" ", 1)[1] + " ="))
    steps = []
    for m in re.finditer(r"createStep\s*\(([^)]*)\)", block(src, "const Steps g_steps")):
        args = [a.strip() for a in m.group(1).split(",")]
        st = {"batches": args[0]}
        if len(args) == 5:
            st.update(train=args[1], loss_train=scalars[args[2]], test=args[3],
                      loss_test=scalars[args[4]])
        steps.append(st)
    layers = []
    li = 1
    while f"g_weights{li}to{li + 1}Vec" in src:
        names = re.findall(r"\w+", block(src, f"g_weights{li}to{li + 1}Vec"))
        layers.append([vecs[n] for n in names])
        li += 1
    used = {s["batches"] for s in steps} | {s.get("train") for s in steps} | {s.get("test") for s in steps}
    return {"network": [2] + list(hidden) + [1], "bias": True, "activation": "tanh",
            "learning_rate": 0.03, "tolerance": 1e-13, "idxs_to_check": idxs,
            "weights": layers, "steps": steps,
            "points": {k: v for k, v in points.items() if k in used}}


def main():
    for name, (fname, ns) in DATASETS.items():
        d = extract(fname, ns)
        nb = sum(len(d["points"][s["batches"]]) for s in d["steps"])
        assert all(len(w) == nb + 1 for w in d["weights"]), name
        (OUT / f"{name}.json").write_text(json.dumps(d, separators=(",", ":")))
        print(name, "layers", len(d["weights"]), "batches", nb, "steps", len(d["steps"]))


if __name__ == "__main__":
    main()
