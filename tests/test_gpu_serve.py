"""Served form: continuous batching of Redi step requests (MI355X).

Bar: every request's result from the served loop -- admitted while others
are on the GPU, stepped in a changing batch, under an HBM budget that makes
later requests wait -- is bit-identical to the same domain stepped alone by
the single-domain FP32-store entry for the same number of steps.
"""
from __future__ import annotations

import numpy as np
import pytest

from more_benchmarks_dc_ocean_amd import inputs

pytestmark = pytest.mark.gpu


def test_served_results_same_alone_and_in_the_mix(gpu_device):
    import torch
    from more_benchmarks_dc_ocean_amd import serve
    from more_benchmarks_dc_ocean_amd.batch import domain_bytes
    from more_benchmarks_dc_ocean_amd.redi import DeviceDomain, make_redi

    print("[gpu] served loop: 12 requests over 3 pool domains", flush=True)
    cases = [inputs.case_domain(4, 201, 30, 22, 25), inputs.case_domain(4, 202, 18, 40, 50),
             inputs.case_domain(4, 203, 26, 26, 75)]
    pool = [DeviceDomain(c, gpu_device, "fp32_store") for c in cases]
    pool_bytes = sum(serve._field_bytes(d) for d in pool)
    need = max(domain_bytes(c.nx, c.ny, c.nz, "fp32_store", 4) for c in cases)
    # room for about three requests at a time (arena and tracer budget): of
    # the 12 requests, arriving in two bursts, most wait for the budget
    budget = pool_bytes + int(3.3 * need / 0.94)
    srv = serve.RediServer(pool, budget, "fp32_store", max_slots=8, device=gpu_device,
                           keep_results=True)
    rng = np.random.default_rng(5)
    reqs = [serve.Request(rid=i, arrival=0.0 if i < 6 else 0.01, nsteps=int(rng.integers(1, 5)),
                          domain=int(rng.integers(0, 3))) for i in range(12)]
    stats = srv.run(reqs)
    srv.close()
    assert stats["requests"] == 12
    assert srv.max_active >= 2, "requests never shared a batched step"
    assert any(r.admitted > r.arrival + 1e-4 for r in reqs), "the budget never made a request wait"
    assert all(r.done >= r.admitted >= r.arrival for r in reqs)
    alone = {}
    for r in reqs:
        key = (r.domain, r.nsteps)
        if key not in alone:
            ro, d = make_redi(cases[r.domain], gpu_device, "fp32_store")
            ro.step(d, r.nsteps)
            torch.cuda.synchronize()
            alone[key] = d.tracers_host()
            ro.close()
        assert np.array_equal(r.result, alone[key]), f"request {r.rid} differs from its solo run"
