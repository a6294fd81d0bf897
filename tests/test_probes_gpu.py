"""Tier-2 per-function parity: every probed device function (ixg_probe, the
same __forceinline__ code the film kernels run) against the oracle's function
on the same records, held to the reference's per-function ULP budgets and
absolute floors (renderer/gpu/ulp.rs within_tolerance; budgets as recorded in
crates/indicatrix/docs/gpu_harness_last_run.json:24-73).  Integer outputs
(the path seed, the facet hit) must match bit for bit.

With IXG_PROBE_REPORT set to a file path, each case appends its summary
(the harness's record: comparisons, max raw ULP, exempted, max genuine ULP)
as one JSON line."""
import json
import os

import numpy as np
import pytest

import oracle_lib
import probe_cases
from ulp import summarize

pytestmark = pytest.mark.gpu

PARAMS = [(ci, si) for ci, case in enumerate(probe_cases.CASES) for si in range(len(case["scenes"]))]


@pytest.mark.parametrize("ci,si", PARAMS,
                         ids=[f"{probe_cases.CASES[c]['fn']}-{s}" for c, s in PARAMS])
def test_function_within_budget(backend, ci, si):
    case = probe_cases.CASES[ci]
    sc, rec = probe_cases.case_inputs(case, si)
    gpu = backend.probe(sc, case["fn"], rec)
    cpu = oracle_lib.probe(sc, case["fn"], rec)
    assert gpu.shape == cpu.shape
    for col in case.get("exact", []):
        assert np.array_equal(gpu[:, col].view(np.uint32), cpu[:, col].view(np.uint32)), (case["label"], col)
    report = []
    for cols, budget, floor in case["checks"]:
        cols = probe_cases.columns(cols, cpu.shape[1])
        s = summarize(cpu[:, cols], gpu[:, cols], budget, floor)
        s.update(label=case["label"], scene=si, columns=[min(cols), max(cols)], abs_floor=floor)
        report.append(s)
        assert s["passed"], s
    path = os.environ.get("IXG_PROBE_REPORT")
    if path:
        with open(path, "a") as f:
            for s in report:
                f.write(json.dumps(s) + "\n")
