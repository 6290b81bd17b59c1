"""Case drivers: one case of the reference's case tables, end to end on the device.

``prescribed_field_case`` follows scripts/run_prescribed_field.py:78-157 and
``coupled_case`` follows scripts/run_coupled.py:78-159: the same ray sets
regenerated from the stored seeds, the same operators, solves, acceptance
checks and output columns, so their rows can be compared with
results/prescribed_field.csv and results/coupled.csv column for column.  The
ray blocks of a set are traced and pooled in one device pass
(``first_hit_edges``, equal to ``pool_counts`` over per-seed
``first_hit_counts``) and the operator is assembled from them on the device
(``ExchangeOperator.from_edges``, the same operator bit for bit).

Command line, with the reference's case tables and microstructure files::

    python -m porerad.cases prescribed pf_OF_core_phi0.10_s2_x --ab \\
        --cases data/cases_prescribed_field.csv --microstructure-dir data/microstructures --out rerun.csv
    python -m porerad.cases coupled --all --cases data/cases_coupled.csv \\
        --microstructure-dir data/microstructures --out coupled_rerun.csv
"""

from __future__ import annotations

import argparse
import csv
import hashlib
import sys
import time
from pathlib import Path
from typing import Callable

import numpy as np

from . import conduction, coupled, exchange, view_factors
from .constants import K_GAS, K_SOLID

DELTA_T_K = 20.0
PRESCRIBED_COLUMNS = [
    "case_id", "group", "structure", "microstructure", "transfer_axis", "T_mean_K", "delta_T_K", "epsilon",
    "q_h_W_m2", "q_c_W_m2", "q_i_W_m2", "ab_wall_channel_rel_difference",
]
COUPLED_COLUMNS = [
    "case_id", "group", "structure", "microstructure", "transfer_axis", "H_m", "N_axis", "T_mean_K", "delta_T_K",
    "epsilon", "q_cond_hot_W_m2", "q_rad_hot_W_m2", "q_total_hot_W_m2", "q_cond_cold_W_m2", "q_rad_cold_W_m2",
    "q_total_cold_W_m2", "radiative_fraction_hot", "k_eff_apparent_W_mK", "wall_flux_mismatch_rel",
    "energy_residual_rel", "newton_updates", "T_min_K", "T_max_K",
]

Log = Callable[[str], None]


def _quiet(message: str) -> None:
    pass


def _seeds(value) -> list[int]:
    if isinstance(value, str):
        return [int(s) for s in value.split(";") if s]
    return [int(s) for s in value]


def _conditions(value) -> list[tuple[float, float]]:
    if isinstance(value, str):
        return [(float(T), float(eps)) for T, eps in (item.split("/") for item in value.split(";"))]
    return [(float(T), float(eps)) for T, eps in value]


def _check_phases(phases_x: np.ndarray, expected_sha256: str) -> None:
    if expected_sha256 and hashlib.sha256(phases_x.tobytes()).hexdigest() != expected_sha256:
        raise ValueError("phase array differs from the one listed for this case")


def build_operator(patches, rays_per_patch: int, seeds, log: Log = _quiet) -> exchange.ExchangeOperator:
    """Projected operator of the pooled ray blocks of ``seeds`` (run_prescribed_field.py:78-92)."""

    edges = view_factors.first_hit_edges(patches, rays_per_patch, seeds)
    rays = np.full(patches.n_patches, rays_per_patch * len(seeds), dtype=np.int64)
    try:
        operator = exchange.ExchangeOperator.from_edges(edges, rays, patches.patch_area_m2)
        n_pairs = edges.nnz
    finally:
        edges.close()
    log(f"operator of {len(seeds)} blocks: {n_pairs} sampled patch pairs, balancing {operator.balance_method} "
        f"({operator.balance_iterations} iterations), largest row-sum deviation {operator.row_sum_error:.2e}")
    return operator


def _channels(operator, patches, conditions):
    return [coupled.prescribed_field_channels(operator, patches, T, DELTA_T_K, eps) for T, eps in conditions]


def prescribed_field_case(case: dict, phases: np.ndarray, *, ab: bool = False, log: Log = _quiet) -> list[dict]:
    """Rows of one prescribed-field case (run_prescribed_field.py:98-157).

    ``case`` holds the columns of data/cases_prescribed_field.csv (strings as
    in the CSV, or parsed values); ``phases`` is the microstructure in its
    stored orientation.  With ``ab`` the sets A and B are also evaluated on
    their own, the append rule decides whether the appended blocks join the
    final set, and the A/B wall-channel difference fills its column.
    """

    rpp = int(case["rays_per_patch_per_block"])
    seeds = {key: _seeds(case[key]) for key in ("seeds_A", "seeds_B", "append_seeds_A", "append_seeds_B")}
    append = case["append_used"] in (True, "true")
    conditions = _conditions(case["conditions"])
    phases_x = view_factors.orient_along_x(np.asarray(phases, dtype=np.uint8), case["transfer_axis"])
    _check_phases(phases_x, case.get("phase_sha256_x", ""))
    patches = view_factors.build_surface_patches(phases_x, float(case["H_m"]))
    log(f"{case['case_id']}: {phases_x.shape} cells, {patches.n_patches} surface patches")

    ab_difference = ""
    if ab:
        set_a = _channels(build_operator(patches, rpp, seeds["seeds_A"], log), patches, conditions)
        set_b = _channels(build_operator(patches, rpp, seeds["seeds_B"], log), patches, conditions)
        channel, source = coupled.ab_differences(set_a, set_b)
        append_rule = coupled.append_required(channel, source)
        log(f"sets A and B: wall-channel difference {channel:.3e}, coarse-source L1 difference {source:.3e}, "
            f"append {'required' if append_rule else 'not required'}")
        if append_rule != append:
            log("the append decision differs from the case table; the decision of the rule is used")
            append = append_rule
        if append:
            set_a = _channels(build_operator(patches, rpp, seeds["seeds_A"] + seeds["append_seeds_A"], log),
                              patches, conditions)
            set_b = _channels(build_operator(patches, rpp, seeds["seeds_B"] + seeds["append_seeds_B"], log),
                              patches, conditions)
            channel, source = coupled.ab_differences(set_a, set_b)
            log(f"appended sets: wall-channel difference {channel:.3e}, coarse-source L1 difference {source:.3e}")
        ab_difference = channel

    final_seeds = seeds["seeds_A"] + seeds["seeds_B"]
    if append:
        final_seeds += seeds["append_seeds_A"] + seeds["append_seeds_B"]
    final = build_operator(patches, rpp, final_seeds, log)
    rows = []
    for result in _channels(final, patches, conditions):
        if result.closure_rel > 1.0e-4:
            raise RuntimeError(f"channel balance not closed: {result.closure_rel:.2e}")
        rows.append({
            **{key: case[key] for key in ("case_id", "group", "structure", "microstructure", "transfer_axis")},
            "T_mean_K": f"{result.T_mean_K:g}", "delta_T_K": f"{DELTA_T_K:g}",
            "epsilon": result.epsilon, "q_h_W_m2": result.q_h, "q_c_W_m2": result.q_c, "q_i_W_m2": result.q_i,
            "ab_wall_channel_rel_difference": ab_difference,
        })
        log(f"T_m {result.T_mean_K:g} K, eps {result.epsilon:g}: q_h {result.q_h:.6f}, q_c {result.q_c:.6f}, "
            f"q_i {result.q_i:.6f} W m^-2, balance {result.closure_rel:.1e}")
    return rows


def _solve(system, operator, patches, T_hot, T_cold, epsilon, T_initial, log: Log):
    result = coupled.solve_coupled(system, operator, patches, T_hot, T_cold, epsilon, T_initial)
    metrics = coupled.wall_flux_metrics(system, patches, result, T_hot, T_cold)
    log(f"coupled solve: converged {result.converged}, {result.newton_iterations} Newton updates, "
        f"final residual {result.final_residual_inf_W:.2e} W")
    failures = coupled.acceptance_failures(result, metrics)
    if result.failure:
        failures.insert(0, result.failure)
    if failures:
        raise RuntimeError("solution not accepted: " + "; ".join(failures))
    return result, metrics


def coupled_case(case: dict, phases: np.ndarray, *, log: Log = _quiet) -> dict:
    """The row of one coupled case (run_coupled.py:111-155).

    ``case`` holds the columns of data/cases_coupled.csv; ``phases`` is the
    microstructure in its stored orientation (repeated along the transfer
    axis as the case asks).  The pooled counts of sets A and B are checked
    against the stored hashes where listed and the outcome logged.
    """

    axis, H, eps = case["transfer_axis"], float(case["H_m"]), float(case["epsilon"])
    T_mean, delta_T = float(case["T_mean_K"]), float(case["delta_T_K"])
    T_hot, T_cold = T_mean + delta_T / 2.0, T_mean - delta_T / 2.0
    rpp = int(case["rays_per_patch_per_block"])
    seeds = {name: _seeds(case[f"seeds_{name}"]) for name in ("A", "B")}
    phases_x = view_factors.orient_along_x(np.asarray(phases, dtype=np.uint8), axis)
    phases_x = np.ascontiguousarray(np.concatenate([phases_x] * int(case["repeats_along_axis"]), axis=0))
    if phases_x.shape[0] != int(case["N_axis"]):
        raise ValueError(f"microstructure has {phases_x.shape[0]} cells along the transfer axis")
    _check_phases(phases_x, case.get("phase_sha256_x", ""))
    patches = view_factors.build_surface_patches(phases_x, H)
    log(f"{case['case_id']}: {phases_x.shape} cells, {patches.n_patches} surface patches")

    for name in "AB":
        expected = case.get(f"raw_count_sha256_{name}", "")
        if not expected:
            continue
        o, r, c = view_factors.first_hit_counts_pooled(patches, rpp, seeds[name])
        rays = np.full(patches.n_patches, rpp * len(seeds[name]), dtype=np.int64)
        match = view_factors.count_sha256(o, r, c, rays) == expected
        log(f"first-hit counts of set {name}: {'identical to' if match else 'DIFFERENT FROM'} the stored hash")
    final = build_operator(patches, rpp, seeds["A"] + seeds["B"], log)

    system = conduction.assemble_conduction(np.where(phases_x == 1, K_SOLID, K_GAS), H)
    T_initial = conduction.solve_conduction(system, T_hot, T_cold)
    result, metrics = _solve(system, final, patches, T_hot, T_cold, eps, T_initial, log)
    row = {key: case[key] for key in ("case_id", "group", "structure", "microstructure", "transfer_axis", "H_m",
                                      "N_axis", "T_mean_K", "delta_T_K", "epsilon")}
    row.update({key: metrics[key] for key in COUPLED_COLUMNS if key in metrics})
    row["newton_updates"] = result.newton_iterations
    return row


def read_cases(path: Path) -> list[dict]:
    with open(path, newline="") as handle:
        return list(csv.DictReader(handle))


def load_phases(directory: Path, microstructure: str) -> np.ndarray:
    path = Path(directory) / f"{microstructure}.npz"
    if not path.is_file():
        raise FileNotFoundError(f"{path} not found")
    with np.load(path, allow_pickle=False) as stored:
        return np.asarray(stored["phases"], dtype=np.uint8)


def write_rows(path: Path | None, rows: list[dict], columns: list[str]) -> None:
    """Write rows to stdout, or merge them into a CSV file replacing rows of the same case."""

    if path is None:
        writer = csv.DictWriter(sys.stdout, fieldnames=columns, lineterminator="\n")
        writer.writeheader()
        writer.writerows(rows)
        return
    kept = []
    if path.is_file():
        with open(path, newline="") as handle:
            kept = [r for r in csv.DictReader(handle) if r["case_id"] not in {row["case_id"] for row in rows}]
    with open(path, "w", newline="") as handle:
        writer = csv.DictWriter(handle, fieldnames=columns, lineterminator="\n")
        writer.writeheader()
        writer.writerows(kept + rows)


def main(argv=None) -> None:
    parser = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    parser.add_argument("kind", choices=("prescribed", "coupled"))
    parser.add_argument("case_ids", nargs="*")
    parser.add_argument("--all", action="store_true", help="every case of the table whose microstructure exists")
    parser.add_argument("--ab", action="store_true", help="prescribed cases: also evaluate sets A and B")
    parser.add_argument("--cases", type=Path, required=True, help="the case table (CSV)")
    parser.add_argument("--microstructure-dir", type=Path, required=True)
    parser.add_argument("--out", type=Path, help="result CSV (rows of the same case are replaced)")
    args = parser.parse_args(argv)
    start = time.perf_counter()

    def log(message: str) -> None:
        print(f"[{time.perf_counter() - start:8.1f} s] {message}", file=sys.stderr, flush=True)

    table = read_cases(args.cases)
    if args.all:
        chosen = [c for c in table if (args.microstructure_dir / f"{c['microstructure']}.npz").is_file()]
    else:
        by_id = {c["case_id"]: c for c in table}
        missing = [c for c in args.case_ids if c not in by_id]
        if missing:
            raise SystemExit(f"unknown cases {missing}")
        chosen = [by_id[c] for c in args.case_ids]
    rows = []
    for case in chosen:
        phases = load_phases(args.microstructure_dir, case["microstructure"])
        if args.kind == "prescribed":
            rows += prescribed_field_case(case, phases, ab=args.ab, log=log)
        else:
            rows.append(coupled_case(case, phases, log=log))
        if args.out is not None:  # results so far survive an interrupted run
            write_rows(args.out, rows, PRESCRIBED_COLUMNS if args.kind == "prescribed" else COUPLED_COLUMNS)
    if args.out is None:
        write_rows(None, rows, PRESCRIBED_COLUMNS if args.kind == "prescribed" else COUPLED_COLUMNS)
    log("done")


if __name__ == "__main__":
    main()
