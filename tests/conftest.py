import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "ib3d-gpu_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs on the GPU box)")


@pytest.fixture(scope="session")
def case_dir(tmp_path_factory):
    """Generated case directories, one per configuration name."""
    from ib3d_amd import cases
    made = {}

    def get(name):
        if name not in made:
            d = tmp_path_factory.mktemp("case_" + name)
            cases.write_case(cases.CONFIGS[name], str(d))
            made[name] = str(d)
        return made[name]
    return get
