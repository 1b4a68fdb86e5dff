"""Drop-in name for the reference's extension package: `import svraster_cuda` resolves to the
MI355X build in `svraster_amd` (same submodules and `_C` functions)."""
import importlib
import sys

for _name in ("_C", "meta", "utils", "renderer", "sparse_adam", "marching_cubes", "grid_loss_bw",
              "render_neus"):
    _mod = importlib.import_module(f"svraster_amd.{_name}")
    sys.modules[f"{__name__}.{_name}"] = _mod
    globals()[_name] = _mod
del _name, _mod
