import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP library)")


@pytest.fixture(scope="session")
def orc():
    from oracle import oracle
    if not os.path.exists(oracle.ORC_PATH):
        oracle.build(ref=False)
    return oracle.Orc()


@pytest.fixture(scope="session")
def ref():
    from oracle import oracle
    if not os.path.exists(oracle.REF_PATH):
        pytest.skip("reference library not built (needs the DelFEM tree at build time)")
    return oracle.Ref()


@pytest.fixture(scope="session")
def gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from delfem_amd import linsys
    linsys.init(0)
    return linsys
