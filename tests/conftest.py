import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libchs on the GPU)")
    config.addinivalue_line("markers", "slow: long-running CPU oracle comparison")


def spinor(rs, shape):
    return (rs.randn(*shape) + 1j * rs.randn(*shape)) / np.sqrt(2.0)


def phi_field(rs, kvol, beta=0.475):
    """phi_a ~ N(0, 1/beta)"""
    return rs.randn(3, kvol) / np.sqrt(beta)
