"""Synthetic catalogs for tests and the benchmark (workload generation, not
part of the estimator).

uniform(Np, L, seed)
    Np positions uniform in [0, L)^3 (the plain small-size case).

lognormal(Np, L, seed, Ngs=256, ...)
    Clustered mock from a Gaussian random field with a power-law spectrum
    P(k) = A (k/k0)^n on an Ngs^3 seed grid:
      1. white noise w (N(0,1) per cell) -> delta_G(k) = DFT(w) sqrt(P(k)/H^3),
         H = L/Ngs, so that |delta_k|^2/V = P(k) in the reference's Fourier
         convention (delta_k = H^3 DFT(delta_x), P = |delta_k|^2/V); DC = 0;
      2. rho = exp(delta_G - sigma^2/2), normalised to mean 1;
      3. exactly Np particles: multinomial counts over rho, each particle
         uniform inside its seed cell;
      4. linear redshift-space shift along z from the Zel'dovich
         displacement psi_z(k) = i k_z / k^2 delta_G(k):  z += f psi_z[cell];
      5. wrapped into [0, L), returned in a random order, float32 by default.
    Deterministic for a given seed (numpy PCG64 and pocketfft).

lognormal_device(Np, L, seed, ...)
    The same recipe drawn on the GPU with torch (catalogs of 2^30).

batch_specs(nmocks)
    Parameters of the batched uneven workload: per seed s, L ~ U[250, 2000],
    Ng from {128, 192, 256, 384}, Np log-uniform in [5e5, 2e7].
"""
import numpy as np


def uniform(Np, L, seed, dtype=np.float64):
    rng = np.random.default_rng(seed)
    pos = rng.random((3, Np)) * L
    return pos[0].astype(dtype), pos[1].astype(dtype), pos[2].astype(dtype)


def gaussian_field(Ngs, L, seed, A=2e4, k0=0.1, n=-1.5):
    """delta_G(k) (rfft layout) of a Gaussian field with P(k) = A (k/k0)^n."""
    rng = np.random.default_rng(seed)
    H = L / Ngs
    w = rng.standard_normal((Ngs, Ngs, Ngs))
    W = np.fft.rfftn(w)
    del w
    kf = 2 * np.pi / L
    ki = np.fft.fftfreq(Ngs, d=1.0 / Ngs) * kf
    kk = np.fft.rfftfreq(Ngs, d=1.0 / Ngs) * kf
    k2 = ki[:, None, None] ** 2 + ki[None, :, None] ** 2 + kk[None, None, :] ** 2
    k2[0, 0, 0] = 1.0
    amp = np.sqrt(A * (np.sqrt(k2) / k0) ** n / H ** 3)
    amp[0, 0, 0] = 0.0
    W *= amp
    return W, ki, kk, k2


def lognormal(Np, L, seed, Ngs=256, A=2e4, k0=0.1, n=-1.5, f=0.76, rsd=True,
              dtype=np.float32, log=None):
    """log: optional callable, called with a short message after each step
    (large catalogs take minutes)"""
    say = log if log is not None else (lambda msg: None)
    rng = np.random.default_rng([seed, 1])
    W, ki, kk, k2 = gaussian_field(Ngs, L, seed, A, k0, n)
    delta = np.fft.irfftn(W, s=(Ngs, Ngs, Ngs), axes=(0, 1, 2))
    sigma2 = float(delta.var())
    rho = np.exp(delta - 0.5 * sigma2)
    del delta
    rho /= rho.sum()
    counts = rng.multinomial(int(Np), rho.ravel())
    say("mock: counts drawn")
    del rho
    cells = np.repeat(np.arange(Ngs ** 3, dtype=np.int64), counts)
    del counts
    H = L / Ngs
    ci = cells // (Ngs * Ngs)
    cj = (cells // Ngs) % Ngs
    ck = cells % Ngs
    x = (ci + rng.random(Np)) * H
    y = (cj + rng.random(Np)) * H
    z = (ck + rng.random(Np)) * H
    say("mock: positions placed")
    if rsd and f != 0:
        psi = np.fft.irfftn(1j * kk[None, None, :] / k2 * W, s=(Ngs, Ngs, Ngs), axes=(0, 1, 2))
        z += f * psi.ravel()[cells]
        del psi
        say("mock: RSD applied")
    del cells, ci, cj, ck, W
    x %= L
    y %= L
    z %= L
    perm = rng.permutation(Np)
    say("mock: permutation drawn")
    out = tuple(np.ascontiguousarray(a[perm]).astype(dtype) for a in (x, y, z))
    # float32 rounding can land exactly on L: wrap again
    for a in out:
        a[a >= L] -= L
    return out, sigma2


def batch_specs(nmocks):
    """(seed, L, Ng, Np) of mocks 0..nmocks-1 of the batched workload: per
    seed s (numpy PCG64 stream [s, 3]) L ~ U[250, 2000], Ng uniform from
    {128, 192, 256, 384}, Np = round(exp U[ln 5e5, ln 2e7])."""
    out = []
    for s in range(nmocks):
        r = np.random.default_rng([s, 3])
        L = float(r.uniform(250.0, 2000.0))
        Ng = int((128, 192, 256, 384)[r.integers(4)])
        Np = int(round(np.exp(r.uniform(np.log(5e5), np.log(2e7)))))
        out.append((s, L, Ng, Np))
    return out


def lognormal_device(Np, L, seed, Ngs=256, A=2e4, k0=0.1, n=-1.5, f=0.76, rsd=True,
                     device="cuda", chunk=1 << 26):
    """The lognormal recipe of lognormal() drawn on the GPU with torch (for
    catalogs of 2^30 particles, which numpy draws in minutes): the same
    field, density and displacement, each particle's cell drawn from the
    density (inverse CDF, i.e. the multinomial of lognormal() particle by
    particle) and its position uniform in the cell.  Deterministic for a
    given seed and device, but not the same catalog as lognormal() with
    that seed.  Returns float32 x, y, z tensors on the device, in random
    order, and sigma^2 of the field."""
    import torch
    gen = torch.Generator(device=device)
    gen.manual_seed(int(seed))
    H = L / Ngs
    w = torch.randn((Ngs, Ngs, Ngs), generator=gen, device=device, dtype=torch.float64)
    W = torch.fft.rfftn(w)
    del w
    kf = 2 * np.pi / L
    ki = torch.fft.fftfreq(Ngs, d=1.0 / Ngs, device=device, dtype=torch.float64) * kf
    kk = torch.fft.rfftfreq(Ngs, d=1.0 / Ngs, device=device, dtype=torch.float64) * kf
    k2 = ki[:, None, None] ** 2 + ki[None, :, None] ** 2 + kk[None, None, :] ** 2
    k2[0, 0, 0] = 1.0
    amp = torch.sqrt(A * (torch.sqrt(k2) / k0) ** n / H ** 3)
    amp[0, 0, 0] = 0.0
    W *= amp
    del amp
    delta = torch.fft.irfftn(W, s=(Ngs, Ngs, Ngs))
    sigma2 = float(delta.var())
    cdf = torch.cumsum(torch.exp(delta - 0.5 * sigma2).reshape(-1), 0)
    del delta
    psi = None
    if rsd and f != 0:
        psi = torch.fft.irfftn(1j * kk[None, None, :] / k2 * W, s=(Ngs, Ngs, Ngs)).reshape(-1)
    del W, k2
    out = [torch.empty(Np, dtype=torch.float32, device=device) for _ in range(3)]
    for s0 in range(0, Np, chunk):
        m = min(chunk, Np - s0)
        u = torch.rand(m, generator=gen, device=device, dtype=torch.float64) * cdf[-1]
        cell = torch.searchsorted(cdf, u).clamp_(max=Ngs ** 3 - 1)
        del u
        ci = torch.div(cell, Ngs * Ngs, rounding_mode="floor")
        cj = torch.div(cell, Ngs, rounding_mode="floor") % Ngs
        ck = cell % Ngs
        for a, c in zip(out, (ci, cj, ck)):
            v = (c + torch.rand(m, generator=gen, device=device, dtype=torch.float64)) * H
            if a is out[2] and psi is not None:
                v += f * psi[cell]
            v = torch.remainder(v, L).to(torch.float32)
            v[v >= L] -= L          # float32 rounding can land exactly on L
            a[s0:s0 + m] = v
        del cell, ci, cj, ck
    return tuple(out), sigma2
