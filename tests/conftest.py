import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


def lcg(state):
    """The seeded LCG of the reference tests (hal/cpu.rs:1132-1135, hal/ntt.rs:674-677):
    returns (new_state, output)."""
    state = (state * 6364136223846793005 + 1442695040888963407) & ((1 << 64) - 1)
    return state, state >> 33


class Lcg:
    def __init__(self, seed):
        self.state = seed

    def __call__(self):
        self.state, out = lcg(self.state)
        return out


def splitmix64_stream(seed):
    """splitmix64 (bfv.rs:879-886)."""
    state = seed

    def nxt():
        nonlocal state
        state = (state + 0x9E3779B97F4A7C15) & ((1 << 64) - 1)
        z = state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & ((1 << 64) - 1)
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & ((1 << 64) - 1)
        return z ^ (z >> 31)

    return nxt


@pytest.fixture(scope="session")
def cpu():
    from oracle import cpu_backend

    cpu_backend.build()
    return cpu_backend.CpuBackend()
