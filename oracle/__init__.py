"""CPU oracle for the sparse-voxel SDF rasterizer path.

TEST INFRASTRUCTURE ONLY.  Nothing in the product (`svraster_amd`, `svraster_cuda`) imports
this package.  It may be imported by `tests/`, by `__graft_entry__.smoke()` (as the checker)
and by the `cpu_baseline` leg of `bench.py` -- never as the thing measured or shipped.

Modules
  binning   numpy FP32 restatement of the integer stage: preprocess (bbox, octant bitset,
            n_dup), inclusive scan, (tile, rank) key duplication, stable sort, tile ranges.
  raster    torch FP64 restatement of the forward compositor; its autograd graph (with the
            reference's two documented gradient quirks) is the backward oracle.  A scalar
            per-pixel loop in the reference's exact control flow pins the vectorised version.
  sh        torch FP64 spherical-harmonics evaluation (degree <= 3) and its backward.
  adam      FP64 closed form of the reference's Adam kernels.
  grid_loss numpy FP64 restatement of the grid eikonal / Laplacian smoothness backward.
  octree    numpy integer restatement of ijk <-> octpath and the valid-gradient table.

How each function is pinned (closed forms, library routines, finite differences, brute force)
is stated in the header of each module and exercised by tests/test_oracle_*.py.
"""
