import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")
    config.addinivalue_line("markers", "ref: needs the compiled reference (oracle/_ref)")


def gpu_count():
    try:
        import torch
        return torch.cuda.device_count()
    except Exception:
        return 0


def pytest_collection_modifyitems(config, items):
    have_gpu = gpu_count() > 0
    import refbind
    have_ref = refbind.available()
    for it in items:
        if "gpu" in it.keywords and not have_gpu:
            it.add_marker(pytest.mark.skip(reason="no GPU in this environment"))
        # a GPU parity test whose checker is missing fails (pytest_runtest_setup)
        # rather than passing as skipped
        if "ref" in it.keywords and not have_ref and not ("gpu" in it.keywords and have_gpu):
            it.add_marker(pytest.mark.skip(reason="oracle/_ref not built"))


def pytest_runtest_setup(item):
    if "ref" in item.keywords and "gpu" in item.keywords:
        import refbind
        if not refbind.available():
            pytest.fail("oracle/_ref (the compiled reference) is missing on a GPU machine: "
                        "build it with __graft_entry__.build() before the GPU tests")
