import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "langevinmedium.jl_amd")
for p in (PKG, os.path.join(ROOT, "tests"), os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")
    # build what is missing (nothing is rebuilt when the libraries exist)
    import build_oracle
    if not os.path.exists(build_oracle.LIB_PATH):
        build_oracle.build()
    from lmhip import build as hipbuild
    if not os.path.exists(hipbuild.LIB_PATH):
        hipbuild.build()


@pytest.fixture(scope="session")
def gpu_ok():
    import torch
    if not torch.cuda.is_available():
        pytest.fail("a gpu-marked test ran without a GPU")
    return True
