import json
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP library)")


@pytest.fixture(scope="session")
def kats():
    with open(os.path.join(GOLDEN, "reference_kats.json")) as fh:
        return json.load(fh)


@pytest.fixture(scope="session")
def orc():
    import oracle
    oracle.build()
    return oracle


def np_col(spec):
    """Fixture column spec -> numpy array (or ('string', blob, idx))."""
    t = spec["type"]
    if t == "STRING":
        return ("string", spec["blob"].encode(), np.array(spec["idx"], dtype=np.int32))
    dt = {"LONG": np.int64, "INT": np.int32, "DOUBLE": np.float64, "FLOAT": np.float32}[t]
    return np.array(spec["values"], dtype=dt)


@pytest.fixture(scope="session")
def ctx():
    """One HIP context for the GPU tests (session-wide; tests run in one process)."""
    import torch
    from ghive_amd import ops
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    c = ops.Context(0)
    yield c
    c.close()


def to_dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def to_np(t):
    return t.cpu().numpy()


def str_col(values):
    """bytes / None values -> a ('string', blob, idx) host column (the oracle's form)."""
    blob, idx, o = bytearray(), [], 0
    for v in values:
        if v is None:
            idx += [-1, -1]
            continue
        blob += v
        idx += [o, o + len(v)]
        o += len(v)
    return ("string", bytes(blob), np.array(idx, dtype=np.int32))


def to_dev_col(c):
    """numpy column or ('string', blob, idx) -> a device column (tensor or StringColumn)."""
    if isinstance(c, tuple) and c[0] == "string":
        import torch
        from ghive_amd import ops
        blob = torch.frombuffer(bytearray(c[1]), dtype=torch.uint8) if len(c[1]) else torch.zeros(0, dtype=torch.uint8)
        return ops.StringColumn(blob.cuda(), to_dev(np.asarray(c[2], dtype=np.int32)))
    return to_dev(c)
