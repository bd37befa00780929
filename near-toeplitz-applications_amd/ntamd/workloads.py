"""Seeded synthetic inputs of the measured workloads.

* compact field: f = sin(2 pi x) sin(2 pi y) sin(2 pi z) + N(0, sigma=1e-3) noise
  on a uniform grid over [0,1]^3 with endpoints (h = 1/(N-1), the reference's
  linspace convention, experiments/run-compact.py:34-39), C order (z, y, x).
* ADI initial condition: exp(-((x-.5)^2+(y-.5)^2)/0.01) + U[0, 1e-3) noise on
  [0,1]^2 with zero Dirichlet boundary.
Host versions use numpy.random.default_rng(seed) and are bit-reproducible;
device versions draw the noise with a seeded torch generator on the GPU (used
for the large benchmark fields, whose parity is checked by properties and by
lines copied back from the device).
"""
import numpy as np
import torch

__all__ = ["compact_field", "compact_field_device", "gaussian_ic", "grid_spacing"]


def grid_spacing(n):
    return 1.0 / (n - 1)


def compact_field(shape, seed=0, sigma=1e-3, dtype=np.float64):
    nz, ny, nx = (shape,) * 3 if isinstance(shape, int) else shape
    z = np.linspace(0.0, 1.0, nz)[:, None, None]
    y = np.linspace(0.0, 1.0, ny)[None, :, None]
    x = np.linspace(0.0, 1.0, nx)[None, None, :]
    f = np.sin(2 * np.pi * x) * np.sin(2 * np.pi * y) * np.sin(2 * np.pi * z)
    f = f + np.random.default_rng(seed).normal(0.0, sigma, size=(nz, ny, nx))
    return f.astype(dtype, copy=False)


def compact_field_device(shape, seed=0, sigma=1e-3, dtype=torch.float64, device="cuda"):
    nz, ny, nx = (shape,) * 3 if isinstance(shape, int) else shape
    dev = torch.device(device)
    lin = lambda n: torch.linspace(0.0, 1.0, n, dtype=torch.float64, device=dev)  # noqa: E731
    sz = torch.sin(2 * np.pi * lin(nz))
    sy = torch.sin(2 * np.pi * lin(ny))
    sx = torch.sin(2 * np.pi * lin(nx))
    g = torch.Generator(device=dev)
    g.manual_seed(seed)
    f = torch.empty((nz, ny, nx), dtype=dtype, device=dev)
    step = max(1, (1 << 26) // max(1, ny * nx))   # ~64 M points (512 MiB FP64) per chunk
    for k0 in range(0, nz, step):
        k1 = min(nz, k0 + step)
        blk = sz[k0:k1, None, None] * sy[None, :, None] * sx[None, None, :]
        blk += sigma * torch.randn((k1 - k0, ny, nx), generator=g, dtype=torch.float64, device=dev)
        f[k0:k1].copy_(blk)
    return f


def gaussian_ic(N, seed=0, noise=1e-3):
    x = np.linspace(0.0, 1.0, N)
    y, x2 = np.meshgrid(x, x, indexing="ij")
    u = np.exp(-((x2 - 0.5) ** 2 + (y - 0.5) ** 2) / 0.01)
    u += np.random.default_rng(seed).uniform(0.0, noise, size=(N, N))
    u[0, :] = u[-1, :] = 0.0
    u[:, 0] = u[:, -1] = 0.0
    return u
