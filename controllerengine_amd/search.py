"""Gain search: the caller side of the evaluation path.

The reference evaluates blocks of candidate gain rows [K, L, N] that its caller generates, and
leaves the ranking by the total score sc[0] to that caller as well (global search and Monte Carlo
use, reference README.md:31-35,105-106; score layout src/config.cu:51-56). GainSearch is such a
caller: a cross-entropy search over the log-magnitudes of the state-feedback gains K and the
observer gains L, signs held at those of the problem's own candidates. Each iteration

  1. draws a block of candidates on the device,
  2. forms each candidate's reference gain N = 1 / (C (I - A + B K^T)^-1 B) on the observer model
     (the rule the problem generators use), as one batched solve,
  3. evaluates the block through the engine (one ce_evaluate on the caller's stream),
  4. keeps the elite by total score and moves the sampling distribution towards it.

The candidate block and its result rows stay in device memory; only the best row of each
iteration is copied to the host.
"""
from dataclasses import dataclass, field

import numpy as np

from ._abi import PRECISION_MIXED


def reference_gain(A, B, C, K):
    """N = 1 / (C (I - A + B K^T)^-1 B) for every row of K (torch tensors, batched)."""
    import torch
    n, ns = K.shape
    eye = torch.eye(ns, dtype=K.dtype, device=K.device)
    M = (eye - A).unsqueeze(0) + B.view(1, ns, 1) * K.view(n, 1, ns)
    x = torch.linalg.solve(M, B.view(1, ns, 1).expand(n, ns, 1)).squeeze(-1)
    return 1.0 / (x @ C)


def cem_update(z, score, mu, sigma, elite, smoothing, min_sigma):
    """One cross-entropy step: the `elite` rows of z with the lowest finite score give the new mean
    and spread, blended with the old ones by `smoothing`.  Returns (mu, sigma, elite indices)."""
    import torch
    s = torch.where(torch.isfinite(score), score, torch.full_like(score, float("inf")))
    idx = torch.topk(s, elite, largest=False).indices
    ez = z[idx]
    mu = smoothing * ez.mean(0) + (1.0 - smoothing) * mu
    sigma = smoothing * ez.std(0) + (1.0 - smoothing) * sigma
    return mu, torch.clamp(sigma, min=min_sigma), idx


@dataclass
class SearchResult:
    best_gains: np.ndarray        # (2ns+1,) float32 gain row [K, L, N]
    best_score: float             # its total score sc[0]
    best_row: np.ndarray          # its INFO row
    history: list = field(default_factory=list)


class GainSearch:
    """Cross-entropy gain search for one problem (plant, instructions, config) on one device.

    n: candidates per iteration; elite: fraction kept; seed: sampling seed; precision: chain mode.
    The initial distribution is centred on the median log-magnitude of the problem's own candidate
    rows, with their spread."""

    def __init__(self, problem, engine=None, precision=PRECISION_MIXED, n=16384, elite=0.02, seed=0,
                 smoothing=0.7, min_sigma=0.02):
        import torch
        from .engine import Engine
        self.torch = torch
        self.problem = problem
        self.engine = engine if engine is not None else Engine(0)
        self.own_engine = engine is None
        self.precision = precision
        self.n = int(n)
        if self.n < 2:
            raise ValueError("GainSearch needs at least 2 candidates per iteration")
        self.elite = max(2, int(round(elite * self.n)))
        self.smoothing = float(smoothing)
        self.min_sigma = float(min_sigma)
        self.dev = f"cuda:{self.engine.device}"
        d = problem.derived()
        self.sc = int(d.info_scoring)
        self.info_w = int(d.info_mx_width)
        self.ns = problem.ns
        self.cfg = problem.cfg_struct()
        self.plant = problem.plant_struct(0)
        f64 = dict(dtype=torch.float64, device=self.dev)
        self.A = torch.tensor(np.asarray(problem.ss_a, np.float64), **f64)
        self.B = torch.tensor(np.asarray(problem.ss_b, np.float64), **f64)
        self.C = torch.tensor(np.asarray(problem.ss_c, np.float64), **f64)
        g = np.asarray(problem.gains[:, : 2 * self.ns], np.float64)
        sign = np.sign(np.median(g, axis=0))
        sign[sign == 0] = 1.0
        lg = np.log(np.maximum(np.abs(g), 1e-30))
        self.sign = torch.tensor(sign, **f64)
        self.mu = torch.tensor(np.median(lg, axis=0), **f64)
        self.sigma = torch.clamp(torch.tensor(lg.std(axis=0), **f64), min=0.1)
        self.gen = torch.Generator(device=self.dev)
        self.gen.manual_seed(int(seed))
        self.info_in = torch.zeros((self.n, self.info_w), dtype=torch.float32, device=self.dev)
        self.info_out = torch.empty_like(self.info_in)

    def close(self):
        if self.own_engine:
            self.engine.close()

    def step(self):
        """One iteration; returns (score, gain row, INFO row) of the iteration's best candidate."""
        torch = self.torch
        z = self.mu + self.sigma * torch.randn((self.n, 2 * self.ns), generator=self.gen, dtype=torch.float64,
                                                device=self.dev)
        g = self.sign * torch.exp(z)
        N = reference_gain(self.A, self.B, self.C, g[:, : self.ns])
        rows = torch.cat([g, N.unsqueeze(1)], dim=1).to(torch.float32).contiguous()
        self.engine.evaluate_device(self.cfg, self.plant, self.precision, rows, self.info_in, self.info_out)
        score = self.info_out[:, self.sc].to(torch.float64)
        self.mu, self.sigma, idx = cem_update(z, score, self.mu, self.sigma, self.elite, self.smoothing,
                                              self.min_sigma)
        b = int(idx[0])
        return float(score[b]), rows[b].cpu().numpy(), self.info_out[b].cpu().numpy()

    def run(self, iters=10):
        best = None
        hist = []
        for it in range(int(iters)):
            s, row, info = self.step()
            if best is None or s < best.best_score:
                best = SearchResult(best_gains=row, best_score=s, best_row=info)
            hist.append(dict(iter=it, iter_best=s, best=best.best_score,
                             sigma_mean=float(self.sigma.mean())))
        best.history = hist
        return best
