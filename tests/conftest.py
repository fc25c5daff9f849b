import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path through the C ABI)")


# The autotuned transform plans bench.py and pi_compute load: the GPU tests run
# under them, so the configuration that is measured is the one that is checked.
PLAN_CACHE = os.path.join(ROOT, "pi_amd", "plans_gfx950.json")


def tuned_plans():
    """{log_n: split} of the committed plan cache."""
    import json

    with open(PLAN_CACHE) as f:
        return {int(k): v for k, v in json.load(f)["plans"].items()}


@pytest.fixture(scope="session")
def ctx():
    """One device context for all GPU tests (fails loudly without the HIP
    library), with the tuned plan cache installed as gnt_init installs it."""
    from pi_amd import gnt

    c = gnt.Context(0, plan_cache=PLAN_CACHE)
    yield c
    c.close()


@pytest.fixture
def builtin_plans(ctx):
    """The library's built-in splits for every length for the test's duration;
    the tuned cache is installed again afterwards."""
    from pi_amd import gnt

    for n in range(1, 31):
        gnt.set_plan(n)
    try:
        yield
    finally:
        gnt.load_plans(PLAN_CACHE)


@pytest.fixture(scope="session")
def orc():
    from oracle import oracle

    oracle.lib()
    return oracle
