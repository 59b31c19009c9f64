"""Shared test configuration: import paths, the `gpu` marker, data directory."""

import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(ROOT, 'spinor-gpe_amd'), ROOT):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP GPU (MI355X) and the built libsgpe")


@pytest.fixture(autouse=True)
def _data_dir(tmp_path, monkeypatch):
    """PSpinor writes trial data; keep it inside the test's tmp dir."""
    monkeypatch.setenv('SPINOR_GPE_DATA', str(tmp_path / 'data'))


@pytest.fixture
def gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch.device('cuda', 0)
