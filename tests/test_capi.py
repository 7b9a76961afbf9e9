"""Host-side checks of the C ABI library: it loads, exports every symbol
include/sopgpu.h declares, and reports errors without a GPU."""
import ctypes
import os
import re

import numpy as np

import sopgpu as S

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    txt = open(os.path.join(ROOT, "include", "sopgpu.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(sop_[a-z0-9_]+)\s*\(", txt)))


def test_every_declared_symbol_is_exported():
    lib = ctypes.CDLL(S.LIB_PATH)
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing
    assert set(declared_symbols()) <= set(S.EXPORTED)


def test_params_struct_layout_matches_oracle():
    import oracle as O

    assert ctypes.sizeof(S.Params) == ctypes.sizeof(O.OrParams)
    assert [f[0] for f in S.Params._fields_] == [f[0] for f in O.OrParams._fields_]


def test_seed_table_matches_oracle():
    import oracle as O

    for seed in (1, 99, 1119727):
        np.testing.assert_array_equal(S.generate_seeds(seed, 300), O.generate_seeds(seed, 300))


def test_job_bytes_and_limits():
    sysm = S.synth_chain(300, seed=5)
    b = S.job_bytes(S.make_params(), sysm)
    assert b > 300 * 16 * 4
    try:
        S.job_bytes(S.make_params(max_native=1), sysm)
    except S.SopError as e:
        assert e.code == -4 and "native contacts" in str(e)
    else:
        raise AssertionError("limit not enforced")


def test_synthetic_chain_statistics():
    sysm = S.synth_chain(1000, seed=7)
    c = sysm.coords[:, :3].astype(np.float64)
    b = np.linalg.norm(c[1:] - c[:-1], axis=1)
    assert np.allclose(b, 3.8, atol=1e-4)
    d = np.linalg.norm(c[:, None] - c[None], axis=2)
    iu = np.triu_indices(1000, 2)
    assert d[iu].min() >= 3.99  # self-avoiding
    # contacts: |i-j|>2 and r<8, each listed once
    assert np.all(sysm.natives[:, 1] - sysm.natives[:, 0] > 2)
    assert np.all(sysm.nat_r0 < 8.0)
    assert 2500 < len(sysm.nat_r0) < 5000, len(sysm.nat_r0)
    assert sysm.nat_eh.min() >= 0.9 and sysm.nat_eh.max() <= 1.5


def test_engine_create_without_gpu_reports_error():
    from conftest import gpu_present

    if gpu_present():
        return
    h = ctypes.c_void_p()
    rc = S.lib.sop_engine_create(0, 0, ctypes.byref(h))
    assert rc == -2
    assert b"device" in S.lib.sop_last_error()
