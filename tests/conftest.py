"""Shared pytest setup: the `gpu` marker, repo-root imports, oracle build."""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libclv.so)")
    config.addinivalue_line("markers", "slow: long statistical runs")


@pytest.fixture(scope="session")
def oracle_lib():
    from oracle import pyoracle
    if not os.path.exists(pyoracle.LIB):
        pyoracle.build()
    return pyoracle
