"""Import-compatibility module for the reference's NeuS ray-marching placeholder
(cuda/svraster_cuda/render_neus.py), which is unimplemented there as well."""


def _unavailable(name):
    def fn(*args, **kwargs):
        raise NotImplementedError(f"{name} is not implemented")
    fn.__name__ = name
    return fn


build_rays = _unavailable("build_rays")
render_neus = _unavailable("render_neus")
build_voxinfo_for_neus = _unavailable("build_voxinfo_for_neus")
