import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) to run")


@pytest.fixture
def in_tmp(tmp_path, monkeypatch):
    """Run the test from an empty scratch directory (GPF paths are relative)."""
    monkeypatch.chdir(tmp_path)
    return tmp_path
