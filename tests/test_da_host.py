"""The one-device domain decomposition (parallelcompact/da.py drop-in) on the CPU.

Block layout, halo faces and coordinates are restated from the reference's
definitions (da.py:61-132 faces only, da.py:402-432 coordinates, rank order
mz*npy*npx + my*npx + mx) and checked element by element."""
import numpy as np
import pytest
import torch

from parallelcompact.da import DA, DA_arange, DA_gather_blocks, DA_scatter_blocks


def _da():
    return DA((3, 4, 5), (2, 3, 2), 1, device="cpu")


def test_scatter_gather_round_trip_and_block_placement():
    da = _da()
    g = torch.arange(6 * 12 * 10, dtype=torch.float64).reshape(6, 12, 10)
    blocks = da.create_global_vector()
    DA_scatter_blocks(da, g, blocks)
    for mz in range(2):
        for my in range(3):
            for mx in range(2):
                assert torch.equal(blocks[mz, my, mx], g[3 * mz:3 * mz + 3, 4 * my:4 * my + 4, 5 * mx:5 * mx + 5])
    back = torch.empty_like(g)
    DA_gather_blocks(da, blocks, back)
    assert torch.equal(back, g)
    arr = np.empty((6, 12, 10))
    DA_gather_blocks(da, blocks, arr)
    assert np.array_equal(arr, g.numpy())


def test_global_to_local_fills_faces_from_neighbours_only():
    da = _da()
    g = torch.arange(6 * 12 * 10, dtype=torch.float64).reshape(6, 12, 10) + 1.0
    blocks = da.create_global_vector()
    DA_scatter_blocks(da, g, blocks)
    L = da.create_local_vector() - 7.0          # untouched ghosts keep -7
    da.global_to_local(blocks, L)
    gp = torch.nn.functional.pad(g, (1, 1, 1, 1, 1, 1), value=-7.0)
    for mz in range(2):
        for my in range(3):
            for mx in range(2):
                win = gp[3 * mz:3 * mz + 5, 4 * my:4 * my + 6, 5 * mx:5 * mx + 7]
                loc = L[mz, my, mx]
                assert torch.equal(loc[1:-1, 1:-1, 1:-1], win[1:-1, 1:-1, 1:-1])
                assert torch.equal(loc[1:-1, 1:-1, 0], win[1:-1, 1:-1, 0])       # x faces
                assert torch.equal(loc[1:-1, 1:-1, -1], win[1:-1, 1:-1, -1])
                assert torch.equal(loc[1:-1, 0, 1:-1], win[1:-1, 0, 1:-1])       # y faces
                assert torch.equal(loc[1:-1, -1, 1:-1], win[1:-1, -1, 1:-1])
                assert torch.equal(loc[0, 1:-1, 1:-1], win[0, 1:-1, 1:-1])       # z faces
                assert torch.equal(loc[-1, 1:-1, 1:-1], win[-1, 1:-1, 1:-1])
                assert float(loc[0, 0, 0]) == -7.0                              # corners untouched
    back = da.create_global_vector()
    da.local_to_global(L, back)
    assert torch.equal(back, blocks)


def test_arange_matches_reference_formula():
    da = _da()
    x, y, z = DA_arange(da, (0.0, 1.0), (2.0, 5.0), (-1.0, 1.0))
    dx, dy, dz = 1.0 / 9, 3.0 / 11, 2.0 / 5
    assert x.shape == (2, 3, 2, 3, 4, 5)
    # block (1, 2, 1) starts at min + m n d (da.py:422-425)
    assert np.array_equal(x[1, 2, 1, 0, 0, :], np.linspace(5 * dx, 5 * dx + 4 * dx, 5))
    assert np.array_equal(y[1, 2, 1, 0, :, 0], np.linspace(2.0 + 8 * dy, 2.0 + 8 * dy + 3 * dy, 4))
    assert np.array_equal(z[1, 2, 1, :, 0, 0], np.linspace(-1.0 + 3 * dz, -1.0 + 3 * dz + 2 * dz, 3))


def test_line_da_shapes():
    da = _da()
    assert da.get_line_DA(0).proc_sizes == (1, 1, 2) and da.get_line_DA(0).local_dims == (3, 4, 5)
    assert da.get_line_DA(1).proc_sizes == (1, 1, 3) and da.get_line_DA(1).local_dims == (3, 5, 4)
    assert da.get_line_DA(2).proc_sizes == (1, 1, 2) and da.get_line_DA(2).local_dims == (4, 5, 3)
    with pytest.raises(ValueError):
        da.get_line_DA(3)


def test_dropins_validate_before_launch():
    from ntamd import NtInvalidArgument
    from parallelcompact.reduced import get_params
    from parallelcompact.rhs import compute_RHS
    from parallelcompact.sum import sum_solutions
    x = torch.zeros((2, 3, 4), dtype=torch.float64)
    with pytest.raises(NtInvalidArgument):
        compute_RHS(torch.zeros((4, 5, 6), dtype=torch.float64), x, 0.1, (2, 3, 5), 0, 1)
    with pytest.raises(NtInvalidArgument):
        compute_RHS(torch.zeros((4, 5, 6), dtype=torch.float64), x, 0.1, (2, 3, 4), 0, 1)   # host tensors
    with pytest.raises(NtInvalidArgument):
        sum_solutions(x[0, 0], x[0, 0], x, x[:, :, 0], x[:, :, 0], (2, 3, 5))
    with pytest.raises(NtInvalidArgument):
        get_params(DA((3, 4, 5), (1, 1, 2), device="cpu"), x, x, torch.zeros((3, 3, 4, 5), dtype=torch.float64))
