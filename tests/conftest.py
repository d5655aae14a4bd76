import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built HIP library")


@pytest.fixture(scope="session")
def gpu_ctx():
    import threepoint_amd as tp
    if tp.device_count() < 1:
        pytest.fail("no HIP device visible for a gpu-marked test")
    return tp.default_context()


@pytest.fixture(scope="session")
def fid_nz():
    from threepoint_amd import workloads as W
    return W.fiducial_nz()


@pytest.fixture(scope="session")
def fid_tables(gpu_ctx, fid_nz):
    import threepoint_amd as tp
    from threepoint_amd import workloads as W
    return tp.CosmologyTables([W.fiducial()], fid_nz, 100, 3.0, gpu_ctx)


@pytest.fixture(scope="session")
def fid_oracle(fid_nz):
    from _oracle import Oracle
    from threepoint_amd import workloads as W
    return Oracle(W.fiducial(), fid_nz)
