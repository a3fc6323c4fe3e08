"""Host logic of the gain search (controllerengine_amd/search.py), on CPU tensors."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from controllerengine_amd import workloads
from controllerengine_amd.search import cem_update, reference_gain


@pytest.mark.parametrize("name", ["c1", "c2"])
def test_reference_gain_gives_unit_dc_gain(name):
    """N is the gain that makes the closed loop x+ = A x + B (N r - K x), y = C x settle at y = r:
    checked by running the loop to steady state (state feedback placed by scipy at fast poles), not
    by restating the formula."""
    from scipy.signal import place_poles
    p = workloads.CONFIGS[name](n=8)
    A = np.asarray(p.ss_a, np.float64)
    B = np.asarray(p.ss_b, np.float64)
    C = np.asarray(p.ss_c, np.float64)
    rng = np.random.default_rng(0)
    Ks = []
    for _ in range(6):
        poles = np.sort(rng.uniform(0.2, 0.7, size=p.ns))
        Ks.append(place_poles(A, B.reshape(-1, 1), poles).gain_matrix.ravel())
    K = torch.tensor(np.array(Ks))
    N = reference_gain(torch.tensor(A), torch.tensor(B), torch.tensor(C), K).numpy()
    for k, nk in zip(np.array(Ks), N):
        x = np.zeros(p.ns)
        for _ in range(400):                     # |poles| <= 0.7: 0.7^400 ~ 1e-62
            x = A @ x + B * (nk - k @ x)
        assert abs(C @ x - 1.0) < 1e-9


def test_cem_update_moves_towards_the_minimum_and_skips_non_finite():
    g = torch.Generator().manual_seed(0)
    target = torch.tensor([1.5, -2.0, 0.25], dtype=torch.float64)
    mu = torch.zeros(3, dtype=torch.float64)
    sigma = torch.full((3,), 2.0, dtype=torch.float64)
    for it in range(30):
        z = mu + sigma * torch.randn((4096, 3), generator=g, dtype=torch.float64)
        score = ((z - target) ** 2).sum(1)
        score[:64] = float("nan")          # failed candidates
        score[64:128] = float("inf")
        mu, sigma, idx = cem_update(z, score, mu, sigma, 64, 0.7, 1e-3)
        assert (idx >= 128).all()
        s = torch.where(torch.isfinite(score), score, torch.full_like(score, float("inf")))
        assert s[idx].max() <= torch.sort(s).values[63]
    assert torch.allclose(mu, target, atol=0.05)
