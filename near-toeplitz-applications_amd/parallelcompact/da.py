"""One-device stand-in for ``parallelcompact/da.py`` (reference lines 9-488).

The reference's ``DA`` is an MPI Cartesian decomposition.
* Each rank owns one (nz, ny, nx) block of an (npz*nz, npy*ny, npx*nx) grid.
* Rank ``mz*npy*npx + my*npx + mx`` holds block (mz, my, mx).
* ``global_to_local`` fills a width-``sw`` ghosted copy of the block. It
  exchanges the six faces with the neighbouring ranks (da.py:61-132). Ghost
  cells on the physical boundary, and the edges and corners, keep what they
  held.

Here every block lives on one device, stacked block-major:
* a "global vector" is ``(npz, npy, npx, nz, ny, nx)``;
* a "local vector" is ``(npz, npy, npx, nz+2sw, ny+2sw, nx+2sw)``.
The MPI halo swaps become seven strided copies on the stacked arrays. The
blocks of one x-line of subdomains, ``[mz, my]``, are a contiguous
(npx, nz, ny, nx) slab: the layout ``reduced.get_params`` takes.

``DA_arange``, ``DA_scatter_blocks`` and ``DA_gather_blocks`` (da.py:402-488)
work on the stacked layout too. The helpers are torch tensor operations, so
they run on the device of their arguments, or on the CPU.
"""
import numpy as np
import torch

__all__ = ["DA", "DA_arange", "DA_scatter_blocks", "DA_gather_blocks"]


class DA:
    """All (npz, npy, npx) blocks of a decomposition, on one device."""

    def __init__(self, local_dims, proc_sizes, stencil_width=1, device="cuda"):
        self.local_dims = tuple(int(v) for v in local_dims)
        self.proc_sizes = tuple(int(v) for v in proc_sizes)
        if len(self.local_dims) != 3 or len(self.proc_sizes) != 3 or min(self.local_dims + self.proc_sizes) < 1:
            raise ValueError("local_dims and proc_sizes are (z, y, x) triples of positive integers")
        self.stencil_width = int(stencil_width)
        if self.stencil_width < 0 or self.stencil_width > min(self.local_dims):
            raise ValueError("stencil width must be in [0, min(local_dims)]")
        self.nz, self.ny, self.nx = self.local_dims
        self.npz, self.npy, self.npx = self.proc_sizes
        self.size = self.npz * self.npy * self.npx
        self.device = torch.device(device)

    # ---- vectors ------------------------------------------------------------------
    def create_global_vector(self):
        """Zeros, (npz, npy, npx, nz, ny, nx): every block's own values."""
        return torch.zeros(self.proc_sizes + self.local_dims, dtype=torch.float64, device=self.device)

    def create_local_vector(self):
        """Zeros, (npz, npy, npx, nz+2sw, ny+2sw, nx+2sw): every block with ghosts."""
        sw = self.stencil_width
        dims = (self.nz + 2 * sw, self.ny + 2 * sw, self.nx + 2 * sw)
        return torch.zeros(self.proc_sizes + dims, dtype=torch.float64, device=self.device)

    def _check(self, t, local):
        sw = self.stencil_width if local else 0
        want = self.proc_sizes + (self.nz + 2 * sw, self.ny + 2 * sw, self.nx + 2 * sw)
        if tuple(t.shape) != want:
            raise ValueError(f"expected shape {want}, got {tuple(t.shape)}")

    def global_to_local(self, global_array, local_array):
        """Interior of every block plus the six faces its neighbours own
        (da.py:61-132); ghosts without a neighbour are left as they are."""
        self._check(global_array, False)
        self._check(local_array, True)
        G, L, sw = global_array, local_array, self.stencil_width
        nz, ny, nx = self.local_dims
        iz, iy, ix = slice(sw, sw + nz), slice(sw, sw + ny), slice(sw, sw + nx)
        L[..., iz, iy, ix] = G
        if sw == 0:
            return
        # x faces: left ghosts from the left neighbour's last sw columns, and back
        L[:, :, 1:, iz, iy, :sw] = G[:, :, :-1, :, :, nx - sw:]
        L[:, :, :-1, iz, iy, sw + nx:] = G[:, :, 1:, :, :, :sw]
        # y faces
        L[:, 1:, :, iz, :sw, ix] = G[:, :-1, :, :, ny - sw:, :]
        L[:, :-1, :, iz, sw + ny:, ix] = G[:, 1:, :, :, :sw, :]
        # z faces
        L[1:, :, :, :sw, iy, ix] = G[:-1, :, :, nz - sw:, :, :]
        L[:-1, :, :, sw + nz:, iy, ix] = G[1:, :, :, :sw, :, :]

    def local_to_global(self, local_array, global_array):
        """Interior of every ghosted block back to the global vector (da.py:134-140)."""
        self._check(global_array, False)
        self._check(local_array, True)
        sw = self.stencil_width
        global_array.copy_(local_array[..., sw:sw + self.nz, sw:sw + self.ny, sw:sw + self.nx])

    def get_line_DA(self, direction):
        """The decomposition of the lines along x (0), y (1) or z (2): proc sizes
        (1, 1, P) and the line axis last, as da.py:154-180 builds it."""
        if direction == 0:
            return DA((self.nz, self.ny, self.nx), (1, 1, self.npx), self.stencil_width, self.device)
        if direction == 1:
            return DA((self.nz, self.nx, self.ny), (1, 1, self.npy), self.stencil_width, self.device)
        if direction == 2:
            return DA((self.ny, self.nx, self.nz), (1, 1, self.npz), self.stencil_width, self.device)
        raise ValueError("direction must be 0 (x), 1 (y) or 2 (z)")


def DA_arange(da, x_range, y_range, z_range):
    """Coordinates x, y, z of every point of every block, each stacked as
    (npz, npy, npx, nz, ny, nx) host arrays (da.py:402-432): spacing
    (max - min)/(N - 1) over the whole grid, and block (mz, my, mx) starts at
    min + m * n * spacing in each direction."""
    nz, ny, nx = da.nz, da.ny, da.nx
    dx = float(x_range[-1] - x_range[0]) / (nx * da.npx - 1)
    dy = float(y_range[-1] - y_range[0]) / (ny * da.npy - 1)
    dz = float(z_range[-1] - z_range[0]) / (nz * da.npz - 1)
    shape = da.proc_sizes + da.local_dims
    x, y, z = np.empty(shape), np.empty(shape), np.empty(shape)
    for mz in range(da.npz):
        for my in range(da.npy):
            for mx in range(da.npx):
                xs, ys, zs = x_range[0] + mx * nx * dx, y_range[0] + my * ny * dy, z_range[0] + mz * nz * dz
                zz, yy, xx = np.meshgrid(np.linspace(zs, zs + (nz - 1) * dz, nz),
                                         np.linspace(ys, ys + (ny - 1) * dy, ny),
                                         np.linspace(xs, xs + (nx - 1) * dx, nx), indexing="ij")
                x[mz, my, mx], y[mz, my, mx], z[mz, my, mx] = xx, yy, zz
    return x, y, z


def _blocked_view(x_global, da):
    """(NZ, NY, NX) -> a (npz, npy, npx, nz, ny, nx) view of the same storage."""
    return x_global.reshape(da.npz, da.nz, da.npy, da.ny, da.npx, da.nx).permute(0, 2, 4, 1, 3, 5)


def DA_scatter_blocks(da, x_global, x_local):
    """Cut an (NZ, NY, NX) array into the stacked blocks ``x_local`` (da.py:434-460)."""
    want = (da.npz * da.nz, da.npy * da.ny, da.npx * da.nx)
    if tuple(x_global.shape) != want:
        raise ValueError(f"x_global must be {want}, got {tuple(x_global.shape)}")
    da._check(x_local, False)
    if isinstance(x_global, np.ndarray):
        x_global = torch.from_numpy(x_global)
    x_local.copy_(_blocked_view(x_global.to(x_local.device), da))


def DA_gather_blocks(da, x_local, x_global):
    """Assemble the stacked blocks ``x_local`` into an (NZ, NY, NX) array (da.py:462-488)."""
    want = (da.npz * da.nz, da.npy * da.ny, da.npx * da.nx)
    if tuple(x_global.shape) != want:
        raise ValueError(f"x_global must be {want}, got {tuple(x_global.shape)}")
    da._check(x_local, False)
    whole = x_local.permute(0, 3, 1, 4, 2, 5).reshape(want)
    if isinstance(x_global, np.ndarray):
        x_global[...] = whole.cpu().numpy()
    else:
        x_global.copy_(whole)
