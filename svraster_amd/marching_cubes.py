"""Import-compatibility module for the reference's mesh extraction helpers
(cuda/svraster_cuda/marching_cubes.py:271-380).  Mesh extraction is outside this rasterizer
build; the names exist so that `import svraster_cuda` succeeds, and calling them raises."""


def _unavailable(name):
    def fn(*args, **kwargs):
        raise NotImplementedError(f"{name}: mesh extraction is not provided by this library")
    fn.__name__ = name
    return fn


vert_interp = _unavailable("vert_interp")
torch_marching_cubes_grid = _unavailable("torch_marching_cubes_grid")
torch_marching_cubes_voxels = _unavailable("torch_marching_cubes_voxels")
