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
    assert lib.svr_abi_version() == 5
    g1 = lib.svr_geom_state_bytes(1, 1000)
    g2 = lib.svr_geom_state_bytes(1, 2000)
    assert g2 > g1 > 17 * 1000
    assert lib.svr_binning_state_bytes(10_000) >= 24 * 10_000
    assert lib.svr_image_state_bytes(1, 128 * 128, 64) >= 4 * 128 * 128 + 12 * 64
    for n in (0, 1, 255, 256, 4097):
        assert lib.svr_geom_state_bytes(1, n) % 256 == 0


def test_adam_bias_table_matches_double_expressions():
    """Host-side bias table: float32 of the double expressions the host-driven unbiased step uses
    (lr / (1 - b1^t), 1 / sqrt(1 - b2^t)); past the table length both corrections are exactly 1."""
    import ctypes
    import math

    import numpy as np

    from svraster_amd.batched import adam_table_length
    lib = _lib.load()
    for lr, betas in ((1e-2, (0.9, 0.99)), (3e-3, (0.1, 0.99)), (1e-3, (0.9, 0.999))):
        n = adam_table_length(betas)
        out = (ctypes.c_float * (2 * n))()
        _lib.check(lib.svr_adam_bias_table(ctypes.c_double(lr), ctypes.c_double(betas[0]),
                                           ctypes.c_double(betas[1]), n, out))
        tab = np.frombuffer(out, dtype=np.float32).reshape(n, 2)
        t = np.arange(1, n + 1, dtype=np.float64)
        ref0 = np.float32(lr / (1.0 - np.power(betas[0], t)))
        ref1 = np.float32(1.0 / np.sqrt(1.0 - np.power(betas[1], t)))
        assert np.array_equal(tab[:, 0], ref0) and np.array_equal(tab[:, 1], ref1)
        assert 1.0 - betas[0] ** n == 1.0 and 1.0 - betas[1] ** n == 1.0
        assert tab[-1, 0] == np.float32(lr) and tab[-1, 1] == 1.0
        # t = 1 closed form: lr / (1 - b1), 1 / sqrt(1 - b2)
        assert tab[0, 0] == np.float32(lr / (1.0 - betas[0]))
        assert tab[0, 1] == np.float32(1.0 / math.sqrt(1.0 - betas[1]))
    assert lib.svr_adam_bias_table(ctypes.c_double(1e-2), ctypes.c_double(0.9), ctypes.c_double(0.99), 0, None) != 0
