import os
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
sys.path.insert(0, HERE)
sys.path.insert(0, os.path.join(REPO, "echo-qgp-mhd-devel_amd"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP C-ABI)")


@pytest.fixture(scope="session")
def oracle():
    from libs import Oracle, ORACLE_SO
    if not os.path.exists(ORACLE_SO):
        pytest.skip("oracle/liboracle.so not built (run __graft_entry__.build())")
    return Oracle()


@pytest.fixture(scope="session")
def ref():
    from libs import Ref, ref_available
    if not ref_available():
        pytest.skip("oracle/_ref/libechoref.so not built (needs the reference sources)")
    return Ref()
