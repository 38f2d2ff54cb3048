"""Shared test set-up.

`gpu` marks tests that need the MI355X and the built HIP library; the rest run
on any CPU. The oracle (oracle/libkestrel_oracle.so) is built on first use.
"""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X and the built libkestrel_hip.so")


@pytest.fixture(scope="session")
def oracle_mod():
    from oracle import oracle
    oracle.lib()
    return oracle
