"""Shared pytest setup: import paths for the package and the oracle, the
`gpu` marker, and session fixtures for graphs and the device context."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(ROOT, "radar-graph_amd"), os.path.join(ROOT, "oracle"), ROOT):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP library)")


@pytest.fixture(scope="session")
def rg():
    import radar_graph
    return radar_graph


@pytest.fixture(scope="session")
def orc():
    import pyoracle
    return pyoracle


@pytest.fixture(scope="session")
def ctx(rg):
    # The gpu-marked tests need device 0: a missing device is an error, not a skip.
    c = rg.Context(0)
    yield c
    c.close()


_GRAPHS = {}


@pytest.fixture(scope="session")
def graphs(rg):
    """Seeded test graphs (GAP generator + reader dedup), built once."""
    def get(name):
        if name not in _GRAPHS:
            kind, scale, deg, sym = {
                "rmat10s": ("rmat", 10, 16, True),
                "rmat12s": ("rmat", 12, 16, True),
                "rmat14s": ("rmat", 14, 16, True),
                "rmat16s": ("rmat", 16, 16, True),
                "rmat12d": ("rmat", 12, 16, False),
                "rmat14d": ("rmat", 14, 16, False),
                "er12s": ("er", 12, 16, True),
                "er12d": ("er", 12, 8, False),
            }[name]
            _GRAPHS[name] = rg.HostGraph.generate(scale, deg, uniform=(kind == "er"), symmetric=sym)
        return _GRAPHS[name]
    return get
