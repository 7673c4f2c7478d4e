import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def oracle_lib():
    from oracle import oracle as O
    return O


def pytest_sessionfinish(session, exitstatus):
    """Writes the divergent-path fractions the GPU parity tests measured (name, fraction, bar,
    pixels) to gpurun_out/parity_divergence.txt."""
    from tests.helpers import DIVERGENCE
    if not DIVERGENCE:
        return
    out = os.path.join(ROOT, "gpurun_out")
    os.makedirs(out, exist_ok=True)
    with open(os.path.join(out, "parity_divergence.txt"), "w") as f:
        f.write("# case  divergent_fraction  bar  pixels  (pixels whose radiance differs from the oracle's by "
                "more than 1e-3 relative)\n")
        for name, frac, bar, px in DIVERGENCE:
            f.write(f"{name:40s} {frac:.6f} {bar:.6f} {px}\n")


def pytest_terminal_summary(terminalreporter, exitstatus, config):
    """The measured divergent-path fractions, printed with the session's result."""
    from tests.helpers import DIVERGENCE
    if DIVERGENCE:
        terminalreporter.write_line("divergent-path fractions (case: measured / bar):")
        terminalreporter.write_line("  " + ", ".join(f"{n} {f:.2e}/{b:.1e}" for n, f, b, _ in DIVERGENCE))
