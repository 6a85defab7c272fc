"""The algorithmic byte model used for the roofline figure (CPU)."""
from __future__ import annotations

import pytest

from more_benchmarks_dc_ocean_amd import byte_model as bm


class _Sizes:
    """A domain's sizes only (nghost = 3 on each side)."""

    def __init__(self, nxp, nyp, nz, nt):
        self.nx, self.ny, self.nz, self.nt = nxp + 6, nyp + 6, nz, nt
        self.interior_cells = nxp * nyp * nz


@pytest.mark.parametrize("nxp,nyp,nz,nt,fp64,fp32", [
    (64, 64, 25, 2, 372.8, 188.2),
    (473, 297, 75, 4, 343.7, 172.3),
    (1440, 1080, 75, 4, 335.4, 168.2),
])
def test_bytes_per_cell_step(nxp, nyp, nz, nt, fp64, fp32):
    """Bytes per cell-step of the three single-domain workloads, as tabulated
    from the two-call contract's compulsory traffic."""
    c = _Sizes(nxp, nyp, nz, nt)
    assert round(bm.bytes_per_cell_step(c), 1) == fp64
    assert round(bm.bytes_per_cell_step(c, "fp32_store"), 1) == fp32


def test_hand_count_tiny_domain():
    """1x1 interior, nghost 0, nz 1, one tracer: h 8 + tracer 16 +
    coefficients 2*(2+2)*(4*24+3*8) + metrics 8*(4*2+4*2+1), counted by hand."""
    c = _Sizes(1, 1, 1, 1)
    c.nx, c.ny = 1, 1
    assert bm.step_bytes(c) == 8 + 16 + 2 * 4 * (4 * 24 + 3 * 8) + 8 * 17


@pytest.mark.parametrize("merged", [True, False])
@pytest.mark.parametrize("form", ["fp64", "fp32_store"])
@pytest.mark.parametrize("group", [1, 2, 3, 4])
@pytest.mark.parametrize("nt", [1, 2, 3, 4])
def test_launch_split_adds_up(group, nt, form, merged):
    c = _Sizes(100, 70, 30, nt)
    launches = bm.step_launches(c, group, form, merged)
    assert ("k_neutral_sweep_uv" in launches) == merged
    assert ("k_neutral_sweep_u" in launches) == (not merged)
    total = sum(sum(v) for v in launches.values())
    assert total == pytest.approx(bm.step_bytes(c, form), rel=1e-12)
    ngroups = -(-nt // max(1, min(group, nt)))
    assert len(launches["k_tracer_recon"]) == ngroups
    assert len(launches["k_tracer_update"]) == ngroups
    assert len(launches["k_flux_x"]) == ngroups
    assert len(launches["k_flux_y"]) == ngroups


def test_ceiling():
    c = _Sizes(1440, 1080, 75, 4)
    assert bm.ceiling_cell_steps(c) / 1e9 == pytest.approx(23.9, abs=0.05)
