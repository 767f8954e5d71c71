import json
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))

GOLDEN = os.path.join(ROOT, "tests", "golden")
PARAMS = os.path.join(ROOT, "joint_bp_plus_pp_amd", "codes",
                      "H_P768_J3_L12_dmax3_nc0-3_1-2_seed11579811919164041.txt")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X)")


_HIP_DEVICE = []


def hip_device_available():
    """Whether a HIP device is visible, probed once per session (through torch,
    before any test's decoder opens the device: probed afterwards the answer
    can come back false)."""
    if not _HIP_DEVICE:
        import torch  # device probe only
        _HIP_DEVICE.append(torch.cuda.is_available())
    return _HIP_DEVICE[0]


def pytest_collection_finish(session):
    if any(item.get_closest_marker("gpu") for item in session.items):
        hip_device_available()


def golden_npz(name):
    with np.load(os.path.join(GOLDEN, name), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def golden_json(name):
    with open(os.path.join(GOLDEN, name)) as f:
        return json.load(f)


def unpack2(packed):
    """Inverse of make_golden.pack2: four qubit states per byte."""
    b = np.asarray(packed, dtype=np.uint8)
    out = np.stack([(b >> s) & 3 for s in (0, 2, 4, 6)], axis=-1)
    return out.reshape(b.shape[:-1] + (-1,)).astype(np.int32)


def fnv1a(arr):
    h = 1469598103934665603
    for v in np.asarray(arr, dtype=np.int64).tolist():
        h ^= v & 0xFFFFFFFFFFFFFFFF
        h = (h * 1099511628211) & 0xFFFFFFFFFFFFFFFF
    return h


@pytest.fixture(scope="session")
def oracle():
    from _oracle import Oracle
    return Oracle()


@pytest.fixture(scope="session")
def main_code_oracle(oracle):
    return oracle.code_from_params(PARAMS)


@pytest.fixture(scope="session")
def gpu_available():
    return hip_device_available()


@pytest.fixture(scope="session")
def main_decoder():
    if not hip_device_available():
        pytest.skip("no HIP device")
    import joint_bp_plus_pp_amd as jbp
    code = jbp.Code.from_params(PARAMS)
    dec = jbp.Decoder(code, 0)
    yield dec
    dec.destroy()
