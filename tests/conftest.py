"""Shared fixtures: the product package (directory taichi-bem_amd/, loaded by
path) and the CPU oracle (oracle/, test infrastructure)."""
import importlib
import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if REPO not in sys.path:
    sys.path.insert(0, REPO)
ORACLE_DIR = os.path.join(REPO, "oracle")
if ORACLE_DIR not in sys.path:
    sys.path.insert(0, ORACLE_DIR)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


def load_package():
    return importlib.import_module("taichi-bem_amd")


@pytest.fixture(scope="session")
def bem():
    return load_package()


@pytest.fixture(scope="session")
def oracle():
    import oracle as O

    O.build()
    return O
