import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
TESTS = os.path.dirname(os.path.abspath(__file__))
if TESTS not in sys.path:
    sys.path.insert(0, TESTS)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (run with -m gpu)")


def pytest_sessionfinish(session, exitstatus):
    """Keep the parity distances the GPU suites measured (tests/_parity.py)
    in $HGDG_PARITY_DIR/parity_distances.json when that folder is given."""
    mod = sys.modules.get("_parity")
    out = os.environ.get("HGDG_PARITY_DIR", "")
    if mod is None or not mod.DISTANCES or not out or not os.path.isdir(out):
        return
    import json

    with open(os.path.join(out, "parity_distances.json"), "w") as f:
        json.dump(mod.DISTANCES, f, indent=1)
