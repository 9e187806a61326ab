"""Shared fixtures.  GPU tests carry @pytest.mark.gpu; everything else runs on CPU."""
import gzip
import json
import os
import shutil
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")
    config.addinivalue_line("markers", "slow: large inputs (seconds to minutes)")


@pytest.fixture(scope="session")
def known():
    with open(os.path.join(GOLDEN, "known_answers.json")) as f:
        return json.load(f)


@pytest.fixture(scope="session")
def fixture_path(tmp_path_factory):
    """Path of a golden file; .gz fixtures are decompressed once per session."""
    tmp = tmp_path_factory.mktemp("golden")

    def get(name: str) -> str:
        p = os.path.join(GOLDEN, name)
        if not name.endswith(".gz"):
            return p
        out = os.path.join(str(tmp), name[:-3])
        if not os.path.exists(out):
            with gzip.open(p, "rb") as fi, open(out, "wb") as fo:
                shutil.copyfileobj(fi, fo)
        return out

    return get
