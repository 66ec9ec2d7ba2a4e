import hashlib
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
PKG = os.path.join(ROOT, "ea-lcd-mc_amd")
for p in (os.path.join(ROOT, "oracle"), PKG, ROOT):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line(
        "markers", "gpu: needs an MI355X (gfx950) and the built HIP engine")


def gpu_available():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


def golden_files():
    return sorted(f for f in os.listdir(GOLDEN)
                  if f.startswith("ref_") and f.endswith(".npz"))


class Golden:
    """One fixture file: verbatim small blobs, SHA-256 of larger ones."""

    def __init__(self, fname):
        import json
        self.name = fname
        self.z = np.load(os.path.join(GOLDEN, fname), allow_pickle=False)
        self.meta = json.loads(bytes(self.z["meta"]).decode())
        self.betas = self.z["betas"]

    def has(self, key):
        return ("raw:" + key) in self.z or ("sha:" + key) in self.z

    def raw(self, key):
        return bytes(self.z["raw:" + key])

    def check(self, key, data):
        data = bytes(np.ascontiguousarray(data).view(np.uint8)) \
            if isinstance(data, np.ndarray) else bytes(data)
        if ("raw:" + key) in self.z:
            want = bytes(self.z["raw:" + key])
            if data != want:
                a = np.frombuffer(data, np.uint8)
                b = np.frombuffer(want, np.uint8)
                if a.size != b.size:
                    raise AssertionError("%s/%s: %d bytes, fixture has %d" %
                                         (self.name, key, a.size, b.size))
                i = int(np.nonzero(a != b)[0][0])
                raise AssertionError("%s/%s differs at byte %d" %
                                     (self.name, key, i))
        else:
            want = bytes(self.z["sha:" + key])
            n = int(self.z["len:" + key])
            assert len(data) == n, "%s/%s: %d bytes, fixture has %d" % (
                self.name, key, len(data), n)
            assert hashlib.sha256(data).digest() == want, \
                "%s/%s: digest differs" % (self.name, key)


@pytest.fixture(params=golden_files())
def golden(request):
    return Golden(request.param)
