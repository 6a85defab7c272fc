"""Algorithmic HBM bytes of one Redi step and their split over kernel launches.

The algorithmic figure counts the compulsory bytes of the public two-call
contract (redi_calc_coeffs then redi_apply_flux, redi/ocean_redi.F90:572-607,
927-1002), per step:

    h read once                               s * N3
    every tracer read once and written once   2 * nt * s * N3
    the Phase-A coefficients written once and read back once
        2 * (Nu + Nv) * (ns * (2 s_Po + 2 s_Ko) + (ns - 1) * s_hEff)
    the 2-D metrics read once (dy_cu, idxCu, wet_u, khtr_u per u face; the
    same per v face; areaT per column)         8 * (4 Nu + 4 Nv + N2)

with N3 = nx*ny*nz and N2 = nx*ny over the whole array (ghosts included, as
the reference computes them), Nu = (nx+1)*ny, Nv = nx*(ny+1), ns = 2nz+2.
Dividing by the interior cells nx_phys*ny_phys*nz gives bytes per cell-step.

Each term is charged to the kernel that moves it:

    k_column_state      h
    k_neutral_sweep_u/v coefficient write of its face grid, wet of that grid
                        (k_neutral_sweep_uv, which sweeps both grids in one
                        launch: both)
    k_face_khtr         khtr_u / khtr_v
    k_tracer_recon      the group's tracer reads (one launch per tracer group)
    k_flux_x / k_flux_y coefficient read-back and dy_cu/idxCu (dx_cv/idyCv)
                        of its face grid, shared over the tracer groups
    k_tracer_update     the group's tracer writes, areaT shared over groups

so the launches of one step add up to the step total exactly.
"""
from __future__ import annotations

from typing import Dict

ELEMENT_BYTES = {
    # form: (s, s_Po, s_Ko, s_hEff)
    "fp64": (8, 8, 4, 8),
    "fp32_store": (4, 4, 2, 4),
}


def _dims(case):
    nx, ny, nz = case.nx, case.ny, case.nz
    return dict(nx=nx, ny=ny, nz=nz, ns=2 * nz + 2, N3=nx * ny * nz, N2=nx * ny,
                Nu=(nx + 1) * ny, Nv=nx * (ny + 1))


def coeff_bytes(nfaces: int, ns: int, form: str = "fp64") -> float:
    """One pass over the Po/Ko/hEff coefficients of ``nfaces`` faces."""
    _, sp, sk, sh = ELEMENT_BYTES[form]
    return nfaces * (ns * (2 * sp + 2 * sk) + (ns - 1) * sh)


def step_bytes(case, form: str = "fp64") -> float:
    d = _dims(case)
    s = ELEMENT_BYTES[form][0]
    nt = case.nt
    return (s * d["N3"] + 2 * nt * s * d["N3"]
            + 2 * coeff_bytes(d["Nu"] + d["Nv"], d["ns"], form)
            + 8 * (4 * d["Nu"] + 4 * d["Nv"] + d["N2"]))


def bytes_per_cell_step(case, form: str = "fp64") -> float:
    return step_bytes(case, form) / case.interior_cells


def step_launches(case, group: int, form: str = "fp64", merged_sweep: bool = True) -> Dict[str, list]:
    """Algorithmic bytes of every launch of one step, per kernel name, for
    tracer groups of ``group`` tracers, with fields and coefficients of
    ``form``'s element sizes (the 2-D metrics are FP64 in every form).
    ``merged_sweep``: both face grids are swept by one launch (64-lane
    sweep tiles, nx > 1 and ny > 1)."""
    d = _dims(case)
    nt = case.nt
    s = float(ELEMENT_BYTES[form][0])
    g = max(1, min(group, 4, nt))
    groups = [min(g, nt - g0) for g0 in range(0, nt, g)]
    out: Dict[str, list] = {"k_column_state": [s * d["N3"]], "k_face_khtr": [8.0 * (d["Nu"] + d["Nv"])]}
    if d["nx"] > 1:
        out["k_neutral_sweep_u"] = [coeff_bytes(d["Nu"], d["ns"], form) + 8.0 * d["Nu"]]
    if d["ny"] > 1:
        out["k_neutral_sweep_v"] = [coeff_bytes(d["Nv"], d["ns"], form) + 8.0 * d["Nv"]]
    if merged_sweep and d["nx"] > 1 and d["ny"] > 1:
        out["k_neutral_sweep_uv"] = [out.pop("k_neutral_sweep_u")[0] + out.pop("k_neutral_sweep_v")[0]]
    fx, fy, cf, rc = [], [], [], []
    for gn in groups:
        share = gn / nt
        rc.append(s * d["N3"] * gn)
        fx.append(share * (coeff_bytes(d["Nu"], d["ns"], form) + 16.0 * d["Nu"]) if d["nx"] > 1 else 0.0)
        fy.append(share * (coeff_bytes(d["Nv"], d["ns"], form) + 16.0 * d["Nv"]) if d["ny"] > 1 else 0.0)
        cf.append(s * d["N3"] * gn + share * 8.0 * d["N2"])
    out["k_tracer_recon"] = rc
    out["k_flux_x"] = fx
    out["k_flux_y"] = fy
    out["k_tracer_update"] = cf
    return out


def launch_bytes(case, kernel: str, group: int, form: str = "fp64") -> float:
    """Mean algorithmic bytes per launch of ``kernel`` in one step."""
    v = step_launches(case, group, form).get(kernel, [])
    return sum(v) / len(v) if v else 0.0


def ceiling_cell_steps(case, peak_bytes_per_s: float = 8e12, form: str = "fp64") -> float:
    """Cell-steps/s at which the algorithmic bytes alone fill HBM."""
    return peak_bytes_per_s / bytes_per_cell_step(case, form)


__all__ = ["ELEMENT_BYTES", "coeff_bytes", "step_bytes", "bytes_per_cell_step", "step_launches",
           "launch_bytes", "ceiling_cell_steps"]
