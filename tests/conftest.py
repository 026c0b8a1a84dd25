import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path)")


@pytest.fixture(scope="session")
def oracle_mod():
    from oracle import oracle
    if not os.path.exists(oracle.ORACLE_LIB):
        oracle.build(ref=False)
    return oracle


@pytest.fixture(scope="session")
def ref_mod(oracle_mod):
    if not oracle_mod.have_ref():
        if os.path.isdir(oracle_mod.REF_SRC):
            oracle_mod.build(ref=True)
        else:
            pytest.skip("compiled reference not available (oracle/_ref not built)")
    return oracle_mod
