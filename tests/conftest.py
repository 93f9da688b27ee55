import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X); run with -m gpu")


def _have_gpu() -> bool:
    try:
        from starsmasher_amd import _lib
        import ctypes as C
        n = C.c_int(0)
        return _lib.load().sg_device_count(C.byref(n)) == 0 and n.value > 0
    except FileNotFoundError:
        # on a machine with a GPU a missing library is an error, not a skip
        return os.path.exists("/dev/kfd")


HAVE_GPU = None


def pytest_collection_modifyitems(config, items):
    global HAVE_GPU
    gpu_items = [it for it in items if "gpu" in it.keywords]
    if not gpu_items:
        return
    HAVE_GPU = _have_gpu()
    if HAVE_GPU:
        return
    skip = pytest.mark.skip(reason="no HIP device visible")
    for it in gpu_items:
        it.add_marker(skip)


def rel_errors(acc, pot, acc_ref, pot_ref):
    """Per-particle relative errors: |a - a_ref| / |a_ref| and |phi - phi_ref| / |phi_ref|."""
    acc = np.asarray(acc, dtype=np.float64)
    acc_ref = np.asarray(acc_ref, dtype=np.float64)
    ea = np.linalg.norm(acc - acc_ref, axis=-1) / np.linalg.norm(acc_ref, axis=-1)
    ep = np.abs(np.asarray(pot) - np.asarray(pot_ref)) / np.abs(np.asarray(pot_ref))
    return ea, ep


@pytest.fixture(scope="session")
def golden_dumps():
    d = np.load(os.path.join(ROOT, "tests", "golden", "dumps.npz"))
    names = sorted({k.split("/")[0] for k in d.files})
    out = {}
    for nm in names:
        rec = {k.split("/")[1]: d[k] for k in d.files if k.startswith(nm + "/")}
        r2 = 4.0 * rec["hp"] ** 2
        rec["range2"] = np.where(rec["u"] == 0.0, -r2, r2)
        out[nm] = rec
    return out


@pytest.fixture(scope="session")
def twin_vectors():
    d = np.load(os.path.join(ROOT, "tests", "golden", "twin_vectors.npz"))
    names = sorted({k.split("/")[0] for k in d.files})
    return {nm: {k.split("/")[1]: d[k] for k in d.files if k.startswith(nm + "/")} for nm in names}
