"""The library's own timer scopes around real products (spla_amd_timer_set_library_scopes):
pgemm_ssb_gpu / pgemm_sbs_gpu (SPLA's SCOPED_TIMING names, pgemm_ssb_gpu.cpp:243,
pgemm_sbs_gpu.cpp:236) enclose gemm_gpu, which encloses split_plan for a split-precision
product; nothing is recorded while the scopes are off."""

from __future__ import annotations

import json

import numpy as np
import pytest

import spla_amd as spla
from helpers import make_buffer, to_device

pytestmark = pytest.mark.gpu


def test_library_scopes_nest_around_products(ctx, device, tmp_path):
    m, n, k = 96, 80, 512
    A = to_device(make_buffer(5, 1, k, m, k, np.complex128), device)
    B = to_device(make_buffer(5, 2, k, n, k, np.complex128), device)
    C = to_device(make_buffer(5, 3, m, n, m, np.complex128), device)
    As = to_device(make_buffer(5, 4, m, n, m, np.complex128), device)
    Bs = to_device(make_buffer(5, 5, n, n, n, np.complex128), device)
    dist = spla.MatrixDistribution.create_mirror(spla.MPI_COMM_SELF)
    split = spla.Context(spla.ProcessingUnit.GPU)
    split.set_gemm_mode(spla.GemmMode.SPLIT_INT8)

    spla.timer.clear()
    try:
        # off: nothing recorded
        spla.pgemm_ssb(m, n, k, spla.Operation.CONJ_TRANSPOSE, 1.0, A, k, B, k, 0.0, C, m, 0, 0,
                       dist, ctx)
        path = tmp_path / "off.json"
        spla.timer.export_json(str(path))
        assert json.loads(path.read_text()) == {}

        spla.timer.set_library_scopes(True)
        spla.timer.start("user")
        for _ in range(2):
            spla.pgemm_ssb(m, n, k, spla.Operation.CONJ_TRANSPOSE, 1.0, A, k, B, k, 0.0, C, m, 0,
                           0, dist, split)
        spla.pgemm_sbs(m, n, n, 1.0, As, m, Bs, n, 0, 0, dist, 0.0, C, m, ctx)
        spla.timer.stop("user")
        path = tmp_path / "on.json"
        spla.timer.export_json(str(path))
        t = json.loads(path.read_text())
    finally:
        spla.timer.set_library_scopes(False)
        spla.timer.clear()

    user = t["user"]
    assert list(user["sub-timings"]) == ["pgemm_ssb_gpu", "pgemm_sbs_gpu"]
    ssb = user["sub-timings"]["pgemm_ssb_gpu"]
    assert len(ssb["timings"]) == 2
    local = ssb["sub-timings"]["gemm_gpu"]
    assert len(local["timings"]) == 2
    assert len(local["sub-timings"]["split_plan"]["timings"]) == 2
    sbs = user["sub-timings"]["pgemm_sbs_gpu"]
    assert len(sbs["sub-timings"]["gemm_gpu"]["timings"]) == 1
    assert "split_plan" not in sbs["sub-timings"]["gemm_gpu"]["sub-timings"]
    assert sum(ssb["timings"]) + sum(sbs["timings"]) <= user["timings"][0]
