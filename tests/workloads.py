"""Synthetic inputs of the benchmark configurations, generated on the device (seeded), and
the device-randomness provider that injects the GPU's sketch draws into the oracle."""
import numpy as np


class DeviceRand:
    """Oracle randomness provider returning exactly what the GPU sketches draw for round
    `rnd` (Philox stream base + rnd under `seed`): Gaussian Omega, SRFT signs and samples,
    subset rows, sparse-Gaussian permutation and coefficients (lra_sketch_draw)."""

    def __init__(self, lra, seed, base=0):
        self.lra, self.seed, self.base = lra, seed, base

    def _h(self, t):
        return t.detach().cpu().numpy()

    def gaussian(self, rnd, rows, cols):
        return np.asfortranarray(self._h(self.lra.sketch_draw("randn", cols, rows, self.seed,
                                                              self.base + rnd)))

    def srft(self, rnd, m, k):
        d, idx = self.lra.sketch_draw("srft", m, k, self.seed, self.base + rnd)
        return self._h(d), self._h(idx)

    def subset(self, rnd, m, k):
        return self._h(self.lra.sketch_draw("sub", m, k, self.seed, self.base + rnd))

    def sprn(self, rnd, m, k):
        r, s = self.lra.sketch_draw("sprn", m, k, self.seed, self.base + rnd)
        return self._h(r), self._h(s)


def usv(torch, m, n, sig, seed_u, seed_v):
    """A = U diag(sig) V' (m x n, column-major on the device) with U, V the Q factors of
    seeded Gaussian m x r and n x r matrices (r = len(sig))."""
    r = len(sig)
    g = torch.Generator(device="cuda")
    g.manual_seed(seed_u)
    U = torch.randn(m, r, dtype=torch.float64, device="cuda", generator=g)
    g.manual_seed(seed_v)
    V = torch.randn(n, r, dtype=torch.float64, device="cuda", generator=g)
    U, _ = torch.linalg.qr(U)
    V, _ = torch.linalg.qr(V)
    s = torch.as_tensor(np.asarray(sig, dtype=np.float64), device="cuda")
    A = (U * s) @ V.T
    return A.t().contiguous().t()


def config2(torch, n=8192):
    """8192 x 8192, sigma_i = 10^(-12(i-1)/511) for i <= 512 and 0 beyond, seeds 1 and 2 (the
    idfact throughput workload)."""
    sig = 10.0 ** (-12.0 * np.arange(512) / 511.0)
    return usv(torch, n, n, sig, 1, 2), sig


def config3(torch, m=65536, n=4096, r=1400):
    """65536 x 4096, sigma_i = exp(-i/40) (i = 1..1400, below 1e-15 sigma_1 beyond), seeds 3
    and 4 (the SRFT-vs-Gaussian workload)."""
    sig = np.exp(-np.arange(1, r + 1) / 40.0)
    return usv(torch, m, n, sig, 3, 4), sig


def config5_requests(count, seed=6, smax=8192):
    """Requests of the served stream: m log-uniform on [256, smax]; square or tall 50/50
    (tall: n log-uniform on [256, m]); idfact / psvdfact 50/50; rtol 1e-6 / 1e-10 50/50;
    sigma_i = exp(-r (i-1)) with r uniform on [0.01, 0.2]."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        m = int(round(np.exp(rng.uniform(np.log(256), np.log(smax)))))
        n = m if rng.random() < 0.5 else int(round(np.exp(rng.uniform(np.log(256), np.log(m)))))
        kind = "idfact" if rng.random() < 0.5 else "psvdfact"
        rtol = 1e-6 if rng.random() < 0.5 else 1e-10
        r = float(rng.uniform(0.01, 0.2))
        out.append(dict(m=m, n=n, kind=kind, rtol=rtol, decay=r, seed=100 + i))
    return out


def decay_matrix(torch, m, n, decay, seed):
    """U diag(exp(-decay (i-1))) V' with the spectrum cut where it drops below 1e-17."""
    r = min(m, n, int(np.ceil(17 * np.log(10) / decay)) + 1)
    sig = np.exp(-decay * np.arange(r))
    return usv(torch, m, n, sig, 2 * seed, 2 * seed + 1), sig
