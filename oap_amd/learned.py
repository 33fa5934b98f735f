"""Learned start for the implicitly restarted Arnoldi solver on a
parameterised operator family: the 2-D convection-diffusion-reaction
operators

    -(alpha u_xx + u_yy) + pe (u_x + u_y) + c u   on a 64 x 64 interior grid,

pe ~ U[0, 50], c ~ U[0, 10], alpha ~ U[0.5, 2], assembled dense (n = 4096,
FP32 coefficients), 6 eigenpairs of smallest real part.

A small network (3 -> 256 -> 256 -> 70, tanh, BF16 operands / FP32
accumulation, trained on the GPU by oaph_mlp_*) maps the normalised
parameters to
  * 64 coefficients of the start vector in a coarse basis: the indicator
    functions of the 8 x 8 blocks of the grid (each block 8 x 8 points)
    multiplied by the convection profile D(ix, iy) = rho_x^ix rho_y^iy,
    rho_x = sqrt((alpha + q) / (alpha - q)), rho_y = sqrt((1 + q) / (1 - q)),
    q = pe h / 2 (the diagonal similarity that makes the constant-coefficient
    operator symmetric, so the block coefficients describe a smooth shape
    rather than an exponential boundary layer), and
  * 6 estimates of the wanted eigenvalues (as log(Re lambda) / 10).
The targets come from converged solves of the plain solver (random start):
the wanted Ritz vectors, each rotated so that its largest entry is real and
positive, summed, divided by D, averaged over every block and normalised;
and the Rayleigh quotients.

solve() starts the solver from the predicted vector and falls back to the
plain random start when that solve does not converge; every answer is
judged by its explicit FP64 residual ||A x - lambda x||_2 <= tol ||A||_F.
The eigenvalue estimates are returned with the solution and can place the
initial shifts: solve(shifts=LearnedStart.shifts(...)) filters the start
vector by the Chebyshev polynomial of the unwanted interval above them (the
shifts are its roots).  Measured on held-out members (14 well-conditioned, 8
strongly convective; degrees 8-32, margins 0.15-0.6) it changes the matvec
count by -1 % to +4 % against the learned start alone: the wanted eigenvalues
sit at the low end of a spectrum ~1e4 times wider than their gap to the next,
where a polynomial of degree 16 amplifies them by ~1.1 against the unwanted
part, and the restarts' exact shifts do that filtering adaptively.  The
filter is therefore not a default.
"""
from __future__ import annotations

import time
from dataclasses import dataclass
from pathlib import Path

import numpy as np

from ._lib import check, load
from .api import (MLP, ArnUtils, CuHArnoldiCallback, DeviceMatrix, MatrixInfo, Precision, QRType)

GRID = 64
N = GRID * GRID
K = 6
M = 24
TOL = 1e-10
BLOCKS = 8                   # coarse basis: BLOCKS x BLOCKS blocks
NCOEF = BLOCKS * BLOCKS
DIMS = [3, 256, 256, NCOEF + K]
PARAM_LO = np.array([0.0, 0.0, 0.5])
PARAM_HI = np.array([50.0, 10.0, 2.0])
DEFAULT_MODEL = Path(__file__).resolve().parent / "data" / "learned_start_cdr64.npz"


def sample_params(count: int, seed: int) -> np.ndarray:
    """count x 3 parameters (pe, c, alpha), uniform on the family's box."""
    rng = np.random.default_rng(seed)
    return PARAM_LO + (PARAM_HI - PARAM_LO) * rng.random((count, 3))


def features(params: np.ndarray) -> np.ndarray:
    """Parameters scaled to [-1, 1]."""
    p = np.atleast_2d(params)
    return (2.0 * (p - PARAM_LO) / (PARAM_HI - PARAM_LO) - 1.0).astype(np.float32)


def eig_encode(ev):
    return np.log(np.maximum(np.real(ev), 1e-3)) / 10.0


def eig_decode(y):
    return np.exp(10.0 * np.asarray(y, dtype=np.float64))


def closed_form_smallest(pe: float, c: float, alpha: float, k: int = K,
                         grid: int = GRID) -> np.ndarray:
    """The k smallest eigenvalues of a family member in closed form.  The
    assembled operator (oaph_synth_cdr) has five coefficients, each rounded
    to FP32: the diagonal d, the x neighbours (xm, xp) and the y neighbours
    (ym, yp).  It is the Kronecker sum of two tridiagonal Toeplitz matrices,
    so its eigenvalues are d + 2 sqrt(xm xp) cos(i pi h) + 2 sqrt(ym yp)
    cos(j pi h), i, j = 1..grid (all real: the cell Peclet number pe h / 2
    stays below alpha on this box)."""
    h = 1.0 / (grid + 1)
    ih2, adv = 1.0 / (h * h), pe / (2.0 * h)
    f32 = lambda v: float(np.float32(v))
    d = f32(2.0 * alpha * ih2 + 2.0 * ih2 + c)
    xm, xp = f32(-alpha * ih2 - adv), f32(-alpha * ih2 + adv)
    ym, yp = f32(-ih2 - adv), f32(-ih2 + adv)
    cs = np.cos(np.arange(1, grid + 1) * np.pi * h)
    ex, ey = 2.0 * np.sqrt(xm * xp) * cs, 2.0 * np.sqrt(ym * yp) * cs
    return np.sort((ex[:, None] + ey[None, :]).ravel() + d)[:k]


def condition_estimate(pe: float, alpha: float) -> float:
    """Growth of the eigenvector condition number with convection: the
    diagonal similarity D that symmetrises the operator has entries ratio
    ((1 + r_x) / (1 - r_x))^(GRID / 2) ((1 + r_y) / (1 - r_y))^(GRID / 2),
    r_x = pe h / (2 alpha), r_y = pe h / 2 (cond(D) up to the grid-size
    factor).  The eigenvalue error of a pair whose residual meets
    1e-10 ||A||_F grows with it: measured on the MI355X, 1.6e-6 relative at
    8.5e5 (pe 9.2, alpha 0.5), and at pe 50 (~1e39) the returned values come
    from the pseudospectrum."""
    h = 1.0 / (GRID + 1)
    rx, ry = pe * h / (2 * alpha), pe * h / 2
    return float(((1 + rx) / (1 - rx)) ** (GRID / 2) * ((1 + ry) / (1 - ry)) ** (GRID / 2))


# condition_estimate bound of the members whose eigenvalues an FP64 solve at
# 1e-10 ||A||_F resolves to 1e-6 relative (pe up to 6.2, 9.3, 12.4 at alpha
# 0.5, 1, 2; about 20 % of the parameter box)
WELL_CONDITIONED = 1e4

# Initial shifts from the predicted eigenvalues: the start vector is filtered
# by the degree-SHIFT_DEGREE Chebyshev polynomial of the unwanted interval
# [(1 + SHIFT_MARGIN) max(predicted), Gershgorin bound], whose roots are the
# shifts (see CdrProblem.filter_start).
SHIFT_DEGREE = 16
SHIFT_MARGIN = 0.3


def gershgorin_bound(pe: float, c: float, alpha: float) -> float:
    """An upper bound of the spectrum of a family member: the largest row sum
    of |entries| with the generator's FP32-rounded coefficients (every
    eigenvalue is real here)."""
    h = 1.0 / (GRID + 1)
    ih2, adv = 1.0 / (h * h), pe / (2.0 * h)
    f32 = lambda v: float(np.float32(v))
    d = f32(2.0 * alpha * ih2 + 2.0 * ih2 + c)
    return d + sum(abs(f32(v)) for v in (-alpha * ih2 - adv, -alpha * ih2 + adv, -ih2 - adv,
                                         -ih2 + adv))


def log_profile(pe: float, alpha: float) -> np.ndarray:
    """log D on the grid (index ix + GRID iy), shifted so that its maximum is 0."""
    h = 1.0 / (GRID + 1)
    q = pe * h / 2.0
    if not (q < alpha and q < 1.0):
        raise ValueError("pe h / 2 must stay below alpha and 1 (real spectrum)")
    lx = 0.5 * np.log((alpha + q) / (alpha - q))
    ly = 0.5 * np.log((1.0 + q) / (1.0 - q))
    i = np.arange(GRID)
    ld = (ly * i)[:, None] + (lx * i)[None, :]
    return (ld - ld.max()).ravel()


def block_average(v: np.ndarray) -> np.ndarray:
    """64 block means of a grid vector (index ix + GRID iy)."""
    s = GRID // BLOCKS
    g = np.asarray(v, dtype=np.float64).reshape(GRID, GRID)  # [iy, ix]
    return g.reshape(BLOCKS, s, BLOCKS, s).mean(axis=(1, 3)).ravel()


def expand(coef: np.ndarray) -> np.ndarray:
    """Grid vector that is constant on each block."""
    s = GRID // BLOCKS
    c = np.asarray(coef, dtype=np.float64).reshape(BLOCKS, BLOCKS)
    return np.kron(c, np.ones((s, s))).ravel()


def start_target(vectors: np.ndarray, pe: float, alpha: float) -> np.ndarray:
    """Coarse coefficients (D-weighted basis) of the sum of the Ritz vectors,
    each rotated so that its largest entry is real and positive (the solver
    works in complex arithmetic, so a real eigenvector comes back with an
    arbitrary phase)."""
    t = np.zeros(N)
    for x in np.asarray(vectors):
        j = np.argmax(np.abs(x))
        x = x * (np.conj(x[j]) / abs(x[j]))
        t += np.real(x) / np.linalg.norm(x)
    c = block_average(t * np.exp(-log_profile(pe, alpha)))
    return c / np.linalg.norm(c)


def start_from(coef: np.ndarray, pe: float, alpha: float) -> np.ndarray:
    """Start vector sum_b coef_b D chi_b, normalised."""
    v = expand(coef) * np.exp(log_profile(pe, alpha))
    return v / np.linalg.norm(v)


@dataclass
class Solution:
    evals: np.ndarray        # wanted eigenvalues (complex), wanted order
    resid: np.ndarray        # explicit ||A x - lambda x||_2
    converged: int           # pairs within tol ||A||_F
    matvecs: int             # operator applications, fallback included
    restarts: int
    seconds: float
    fallback: bool           # the random start was needed
    predicted: np.ndarray | None = None  # network's eigenvalue estimates
    vectors: np.ndarray | None = None
    fro: float = 0.0         # ||A||_F, the residual scale


class CdrProblem:
    """One operator of the family, resident in HBM (FP64 re plane)."""

    def __init__(self):
        self.lib = load()
        self.A = DeviceMatrix(N, N, True, False)
        self.solver = CuHArnoldiCallback()
        s = self.solver
        s.setQRType(QRType.QRGR)
        s.setSortType(ArnUtils.SortSmallestReValues)
        s.setCheckTolerance(TOL)
        s.setOutputType(ArnUtils.DEVICE)
        s.setMaxRestarts(3000)
        s.setMatrix(self.A)
        self.re, self.im = np.zeros(K), np.zeros(K)
        s.setOutputsEigenvalues(self.re, self.im)
        self.vecs = [DeviceMatrix(1, N, True, True) for _ in range(K)]
        s.setOutputsEigenvectors(self.vecs)
        self.fro = 0.0
        self._api = None

    def set_params(self, pe: float, c: float, alpha: float) -> None:
        check(self.lib.oaph_synth_cdr(self.A.ptr, GRID, pe, c, alpha), "synth_cdr")
        # new values: the solver's FP32 copy of the operator is remade
        self.solver.setMatrix(self.A)
        from .api import CuProceduresApi
        self.fro = CuProceduresApi().magnitude(self.A)

    def _run(self, start, seed, want_vectors):
        s = self.solver
        s.setSeed(seed)
        s.setStartVector(start)
        s.execute(M, K, MatrixInfo(True, False, N, N), raise_on_not_converged=False)
        st = s.stats()
        resid = np.array([s.testOutcome(t) for t in range(K)])
        ok = int(np.sum(resid <= TOL * self.fro))
        vecs = np.array([v.to_complex()[:, 0] for v in self.vecs]) if want_vectors else None
        return self.re + 1j * self.im, resid, ok, st, vecs

    def filter_start(self, v0: np.ndarray, lo: float, hi: float, a0: float,
                     degree: int) -> np.ndarray:
        """p(A) v0 on the device, p the degree-`degree` Chebyshev polynomial
        of [lo, hi] scaled to p(a0) = 1 (a0 < lo: the wanted end), by the
        scaled three-term recurrence of Chebyshev-filtered subspace
        iteration (Zhou & Saad, 2007): the components of eigenvalues in
        [lo, hi] are damped by 1 / |T_degree| at a0 against those below lo.
        Its roots, the Chebyshev nodes of [lo, hi], are the initial shifts."""
        from .api import CuProceduresApi
        api = self._api
        if api is None:
            api = self._api = CuProceduresApi()
            self._vec = [DeviceMatrix(1, N, True, False) for _ in range(4)]
        X, Y, W, T = self._vec
        X.upload(np.asarray(v0, dtype=np.float64).real)
        e, c = (hi - lo) / 2.0, (hi + lo) / 2.0
        sigma = e / (a0 - c)
        sigma1 = sigma
        # Y = (A - c) X sigma1 / e
        api.dotProduct(W, self.A, X)
        api.multiplyReConstant(T, X, c)
        api.substract(W, W, T)
        api.multiplyReConstant(Y, W, sigma1 / e)
        for _ in range(2, degree + 1):
            sigma_new = 1.0 / (2.0 / sigma1 - sigma)
            # X <- 2 sigma_new / e (A - c) Y - sigma sigma_new X, then swap
            api.dotProduct(W, self.A, Y)
            api.multiplyReConstant(T, Y, c)
            api.substract(W, W, T)
            api.multiplyReConstant(W, W, 2.0 * sigma_new / e)
            api.multiplyReConstant(T, X, sigma * sigma_new)
            api.substract(X, W, T)
            X, Y = Y, X
            sigma = sigma_new
        v = Y.download()[0][:, 0]
        return v / np.linalg.norm(v)

    def solve(self, start=None, seed: int = 1, want_vectors: bool = False,
              shifts=None) -> Solution:
        """From `start` (None: uniform random start); with `shifts` = (lo, hi,
        a0, degree) the start vector is first filtered by the Chebyshev
        polynomial of [lo, hi] (filter_start; its matvecs are counted and
        timed).  Falls back to the random start when the solve from `start`
        does not converge."""
        t0 = time.perf_counter()
        extra = 0
        if shifts is not None and start is not None:
            lo, hi, a0, degree = shifts
            start = self.filter_start(start, lo, hi, a0, degree)
            extra = degree
        ev, resid, ok, st, vecs = self._run(start, seed, want_vectors)
        mv, rs, fb = st["matvecs"] + extra, st["restarts"], False
        if ok < K and start is not None:
            ev, resid, ok, st, vecs = self._run(None, seed, want_vectors)
            mv += st["matvecs"]
            rs += st["restarts"]
            fb = True
        return Solution(ev.copy(), resid, ok, mv, rs, time.perf_counter() - t0, fb, None, vecs,
                        self.fro)


class LearnedStart:
    """The network and its use as a start-vector predictor."""

    def __init__(self, max_batch: int = 256):
        self.mlp = MLP(DIMS, max_batch=max_batch)
        self.mlp.init_params(seed=0)
        w = np.ones(DIMS[-1], dtype=np.float32)
        w[NCOEF:] = 0.5
        self.mlp.set_output_weights(w)

    def fit(self, params: np.ndarray, targets: np.ndarray, epochs: int = 400,
            batch: int = 64, lr: float = 2e-3, seed: int = 1) -> np.ndarray:
        self.mlp.set_dataset(features(params), targets.astype(np.float32))
        return self.mlp.train(batch=batch, epochs=epochs, lr=lr, seed=seed)

    def predict(self, params: np.ndarray):
        y = self.mlp.predict(features(params)).astype(np.float64)
        coef = y[:, :NCOEF]
        coef /= np.maximum(np.linalg.norm(coef, axis=1, keepdims=True), 1e-30)
        return coef, eig_decode(y[:, NCOEF:])

    def start_vector(self, pe: float, c: float, alpha: float):
        coef, ev = self.predict(np.array([[pe, c, alpha]]))
        return start_from(coef[0], pe, alpha), ev[0]

    @staticmethod
    def shifts(ev_hat, pe: float, c: float, alpha: float, degree: int = SHIFT_DEGREE,
               margin: float = SHIFT_MARGIN):
        """(lo, hi, a0, degree) of the start-vector filter from the predicted
        eigenvalues: the unwanted interval starts `margin` above the largest
        predicted one and ends at the Gershgorin bound; the filter is 1 at
        the smallest predicted one."""
        ev = np.sort(np.real(ev_hat))
        return ((1.0 + margin) * ev[-1], gershgorin_bound(pe, c, alpha), ev[0], degree)

    def save(self, path: str | Path) -> None:
        np.savez(path, params=self.mlp.get_params(), dims=np.array(DIMS), blocks=BLOCKS,
                 grid=GRID, k=K)

    @classmethod
    def load(cls, path: str | Path = DEFAULT_MODEL) -> "LearnedStart":
        d = np.load(path, allow_pickle=False)
        if list(d["dims"]) != DIMS or int(d["blocks"]) != BLOCKS or int(d["grid"]) != GRID:
            raise ValueError(f"{path}: model shape does not match this module")
        ls = cls()
        ls.mlp.set_params(d["params"])
        return ls


def targets_from(sol: Solution, pe: float, alpha: float) -> np.ndarray:
    """Training target row: 64 coarse coefficients and 6 encoded eigenvalues."""
    return np.concatenate([start_target(sol.vectors, pe, alpha),
                           eig_encode(np.sort(sol.evals.real))])


def generate(params: np.ndarray, seed: int = 1, log=None):
    """Plain-solver solves over `params`; returns (kept params, targets,
    matvecs of each solve).  Solves that do not converge are dropped."""
    prob = CdrProblem()
    keep, rows, mvs = [], [], []
    for i, (pe, c, a) in enumerate(params):
        prob.set_params(pe, c, a)
        sol = prob.solve(None, seed=seed + i, want_vectors=True)
        mvs.append(sol.matvecs)
        if sol.converged == K:
            keep.append((pe, c, a))
            rows.append(targets_from(sol, pe, a))
        if log and (i + 1) % 64 == 0:
            log(f"  {i + 1}/{len(params)} solves, mean matvecs {np.mean(mvs):.0f}")
    return np.array(keep), np.array(rows, dtype=np.float32), np.array(mvs)


def compare(model: LearnedStart, params: np.ndarray, seed: int = 7, mixed: bool = False,
            shifts: bool = False):
    """Each operator solved by the plain FP64 solver from the random start and
    from the learned start (same random seed for the fallback); with `mixed`
    also from the learned start with the FP32 operand (the CDR operator's
    entries are FP32 values, so the FP32 copy is the operator itself); with
    `shifts` (and `mixed`) also from the learned start filtered by the
    predicted shifts (LearnedStart.shifts), FP32 operand."""
    prob = CdrProblem()
    out = []
    for i, (pe, c, a) in enumerate(params):
        prob.set_params(pe, c, a)
        prob.solver.setPrecision(Precision.FP64)
        plain = prob.solve(None, seed=seed + i)
        v0, ev_hat = model.start_vector(pe, c, a)
        learned = prob.solve(v0, seed=seed + i)
        learned.predicted = ev_hat
        row = [plain, learned]
        if mixed:
            prob.solver.setPrecision(Precision.MIXED_FP32)
            lm = prob.solve(v0, seed=seed + i)
            lm.predicted = ev_hat
            row.append(lm)
            if shifts:
                ls = prob.solve(v0, seed=seed + i, shifts=model.shifts(ev_hat, pe, c, a))
                ls.predicted = ev_hat
                row.append(ls)
            prob.solver.setPrecision(Precision.FP64)
        out.append(tuple(row))
    return out


def main(argv=None) -> int:
    """python -m oap_amd.learned train --solves 1024 --out model.npz
    python -m oap_amd.learned generate --total 20000 --chunk 0 --chunk-size 5000 --out d0.npz
    python -m oap_amd.learned train --data d0.npz d1.npz ... --out model.npz

    `generate` solves one slice of the `total` operators drawn from the
    sampling box (the same draw for every chunk, so chunks partition it) and
    stores (params, targets, matvecs); `train --data` fits on the union of
    such files instead of generating."""
    import argparse
    import sys
    ap = argparse.ArgumentParser()
    ap.add_argument("cmd", choices=["train", "generate"])
    ap.add_argument("--solves", type=int, default=1024)
    ap.add_argument("--total", type=int, default=20000)
    ap.add_argument("--chunk", type=int, default=0)
    ap.add_argument("--chunk-size", type=int, default=5000)
    ap.add_argument("--data", nargs="*", default=None)
    ap.add_argument("--epochs", type=int, default=600)
    ap.add_argument("--batch", type=int, default=64)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--seed", type=int, default=0x5EED0004)
    ap.add_argument("--out", default=str(DEFAULT_MODEL))
    ap.add_argument("--eval", type=int, default=32)
    a = ap.parse_args(argv)
    log = lambda s: print(s, file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    if a.cmd == "generate":
        lo = a.chunk * a.chunk_size
        params = sample_params(a.total, a.seed)[lo:lo + a.chunk_size]
        P, T, mv = generate(params, seed=1 + lo, log=log)
        np.savez(a.out, params=P, targets=T, matvecs=mv)
        log(f"chunk {a.chunk}: {len(P)}/{len(params)} converged solves in "
            f"{time.perf_counter() - t0:.1f} s, mean matvecs {mv.mean():.0f}")
        return 0
    if a.data:
        parts = [np.load(f, allow_pickle=False) for f in a.data]
        P = np.concatenate([d["params"] for d in parts])
        T = np.concatenate([d["targets"] for d in parts])
        mv = np.concatenate([d["matvecs"] for d in parts])
        log(f"loaded {len(P)} converged solves from {len(a.data)} files, mean matvecs "
            f"{mv.mean():.0f}")
    else:
        params = sample_params(a.solves, a.seed)
        P, T, mv = generate(params, log=log)
        log(f"generated {len(P)}/{a.solves} converged solves in {time.perf_counter() - t0:.1f} s, "
            f"mean matvecs {mv.mean():.0f}")
    model = LearnedStart()
    t1 = time.perf_counter()
    losses = model.fit(P, T, epochs=a.epochs, batch=a.batch, lr=a.lr)
    log(f"trained {a.epochs} epochs in {time.perf_counter() - t1:.1f} s, loss "
        f"{losses[0]:.4g} -> {losses[-1]:.4g}")
    Path(a.out).parent.mkdir(parents=True, exist_ok=True)
    model.save(a.out)
    ev = compare(model, sample_params(a.eval, a.seed + 1))
    mp = np.mean([p.matvecs for p, _ in ev])
    ml = np.mean([l.matvecs for _, l in ev])
    log(f"held-out {a.eval}: mean matvecs random {mp:.0f}, learned {ml:.0f} "
        f"(fallbacks {sum(l.fallback for _, l in ev)})")
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
