"""The C-ABI library loads and exports every symbol its header declares (no GPU needed)."""
from svraster_amd import _lib


def test_library_exports_every_header_symbol():
    lib = _lib.load()
    names = _lib.header_symbols()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert set(names) == set(_lib._SIGS), "ctypes signature table out of sync with the header"


def test_abi_version_and_state_sizes():
    lib = _lib.load()
    assert lib.svr_abi_version() == 1
    g1 = lib.svr_geom_state_bytes(1, 1000)
    g2 = lib.svr_geom_state_bytes(1, 2000)
    assert g2 > g1 > 17 * 1000
    assert lib.svr_binning_state_bytes(10_000) >= 24 * 10_000
    assert lib.svr_image_state_bytes(1, 128 * 128, 64) >= 4 * 128 * 128 + 12 * 64
    for n in (0, 1, 255, 256, 4097):
        assert lib.svr_geom_state_bytes(1, n) % 256 == 0
