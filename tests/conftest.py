import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)
# SPIKE_ORACLE_PYDIR: a directory whose Python oracle modules replace oracle/'s (the
# mutation test runs the pins against copies with planted mistakes this way)
if os.environ.get("SPIKE_ORACLE_PYDIR"):
    sys.path.insert(0, os.environ["SPIKE_ORACLE_PYDIR"])


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")
