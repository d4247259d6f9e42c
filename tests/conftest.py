"""Shared fixtures.  Tests marked ``gpu`` need an MI355X; everything else runs on CPU.

The CPU oracle (oracle/) is imported here only to check results."""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent
GOLDEN = ROOT / "tests" / "golden"
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device 0)")


@pytest.fixture(scope="session")
def oracle():
    import oracle as O  # noqa: F401  (oracle/oracle.py)
    O.lib()
    return O


@pytest.fixture(scope="session")
def golden_scene():
    import rayfactorcl_amd as rf

    def get(name: str):
        return rf.read_rfscene(GOLDEN / "scenes" / f"{name}.rfscene")
    return get


SCENE_NAMES = ["config1", "input-C109", "input-C122", "input-C14", "input-C14-BV", "input-C47", "input-C52",
               "input-C52-BV", "input-C77", "input-W100", "input-furnace", "input-uni"]


@pytest.fixture(scope="session")
def ctx():
    import rayfactorcl_amd as rf
    c = rf.Context(0)
    yield c
    c.close()


def batch_mean_sigma(hits_per_block: np.ndarray, rays_per_block: int) -> float:
    """Standard error of a VF from per-block estimates (blocks of whole work items)."""
    est = hits_per_block / rays_per_block
    return float(est.std(ddof=1) / np.sqrt(len(est)))


def oracle_trace_mt(O, prims, bvs, emitter: int, n_wi: int, chunks: int = 64, threads: int = 16, **kw):
    """Per-ray oracle trace of work items [0, n_wi) of one emitter, split over
    host threads (the ctypes call releases the GIL)."""
    from concurrent.futures import ThreadPoolExecutor
    step = -(-n_wi // chunks)
    spans = [(w, min(step, n_wi - w)) for w in range(0, n_wi, step)]
    with ThreadPoolExecutor(max_workers=threads) as ex:
        parts = list(ex.map(lambda s: O.trace(prims, bvs, emitter, s[0], s[1], **kw), spans))
    return np.concatenate(parts)
