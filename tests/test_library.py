"""CPU-side checks of the HIP library: it loads, exports every symbol the
header declares, its host recipe matches the oracle, and errors come back as
codes (never a crash) without a GPU."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import emeans_hip as E
import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _header_functions():
    txt = open(os.path.join(ROOT, "include", "emeans_hip.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(emh_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    L = E.lib()
    declared = _header_functions()
    assert len(declared) >= 17
    for name in declared:
        assert hasattr(L, name), name
    assert sorted(E.EXPORTS) == declared


@pytest.mark.parametrize("n,d,k,nep", [(200_000, 96, 1024, 5.0), (1_000_000, 8, 100, 2.5), (4000, 4, 8, 5.0),
                                       (1000, 2, 10, 5.0), (2_000_000, 128, 16384, 5.0),
                                       (1_000_000, 512, 4096, 5.0), (10_000_000, 256, 65536, 2.0),
                                       (5000, 64, 64, 3.0), (500_000, 256, 4096, 3.0), (777, 3, 5, 0.3)])
def test_host_recipe_matches_oracle(n, d, k, nep):
    a = E.resolve_recipe(n, d, k, E.Config(nepochs=nep))
    b = O.resolve_recipe(n, d, k, O.Config(nepochs=nep))
    assert a.batch_size == b.batch_size and a.alpha0 == b.alpha0
    assert a.epoch_rows == b.epoch_rows and a.total_steps == b.total_steps
    assert a.unused_threshold == b.unused_threshold


def test_many_epochs_up_to_the_documented_limit():
    """EMH_MAX_EPOCHS = 1024 table entries: nepochs up to 1022.x resolves as
    the oracle does; 1023 and above are refused (the reference's Vec has no
    bound)."""
    for nep in (63.0, 100.5, 1022.25):
        a = E.resolve_recipe(5000, 16, 10, E.Config(nepochs=nep))
        b = O.resolve_recipe(5000, 16, 10, O.Config(nepochs=nep))
        assert a.epoch_rows == b.epoch_rows and a.total_steps == b.total_steps and a.batch_size == b.batch_size
        assert len(a.epoch_rows) == int(np.ceil(nep))
    with pytest.raises(E.EmeansError) as ei:
        E.resolve_recipe(5000, 16, 10, E.Config(nepochs=1023.0))
    assert ei.value.code == E.E_ARG


def test_host_recipe_golden():
    # src/lib.rs:526-540 (faiss 1.16.0)
    r = E.resolve_recipe(200_000, 96, 1024)
    assert r.batch_size == 16384 and abs(r.alpha0 - 0.728183) < 1e-5
    r = E.resolve_recipe(1_000_000, 8, 100, E.Config(nepochs=2.5))
    assert r.epoch_rows == [1_000_000, 500_000, 1_000_000] and r.batch_size == 4096
    assert abs(r.alpha0 - 0.319215) < 1e-5


def test_bad_arguments_return_codes():
    L = E.lib()
    with pytest.raises(E.EmeansError) as ei:
        E.resolve_recipe(100, 4, 2, E.Config(nepochs=0.0))
    assert ei.value.code == E.E_ARG
    assert L.emh_create(0, 0, 0, None) == E.E_ARG
    h = C.c_void_p()
    assert L.emh_create(0, 99, 0, C.byref(h)) == E.E_ARG
    assert L.emh_set_centroids(None, None, 1, 1) == E.E_ARG
    assert L.emh_assign_f32(None, None, 0, None, None) == E.E_ARG
    assert L.emh_last_error()


def test_create_without_gpu_fails_cleanly():
    import torch

    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    h = C.c_void_p()
    rc = E.lib().emh_create(0, E.FP32, 0, C.byref(h))
    assert rc == E.E_DEVICE and not h.value
    assert b"device" in E.lib().emh_last_error()
