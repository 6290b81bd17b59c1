"""Write the golden fixtures of the test suite from the reference's own data files.

Reads (never imports or runs) the reference repository's shipped data:
  data/cases_coupled.csv          seeds + archived count hashes (raw_count_sha256_A/B)
  data/cases_prescribed_field.csv seeds of the prescribed-field ray sets
  data/microstructures/*.npz      phase arrays (loaded with allow_pickle=False)
  results/prescribed_field.csv    archived FP64 channels q_h, q_c, q_i
  results/coupled.csv             archived coupled-solve outputs
  results/fullpath_crosscheck.csv archived first-hit radiosity channels
  results/radiative_properties.csv, results/penetration_response.csv and
  their results/hit_point_restart/ twins: archived Monte Carlo statistics
  data/microstructures.csv        build parameters, scalar descriptors and
                                  phase SHA-256 of every microstructure

and writes, next to this script:
  geometries.npz  bit-packed phase arrays (np.packbits of the C-order uint8
                  array) of the microstructures used by the tests, with shapes
                  and the SHA-256 of the original bytes
  cases.json      hash cases, prescribed-field cases with their archived rows,
                  canonical coupled rows and the full-path cross-check rows;
                  every prescribed-field case and every coupled case with its
                  archived result rows (the 8 resolution cases use
                  microstructures that are not shipped: they are
                  regenerated from the microstructure table);
                  the Monte Carlo rows (both restart rules) of those
                  microstructures; every row of the microstructure table

Run once in a container that holds the reference:
    python tests/golden/make_fixtures.py /path/to/reference
"""

from __future__ import annotations

import csv
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))

PRESCRIBED_CASES = [
    "pf_OF_core_phi0.05_s3_x",
    "pf_FA_core_phi0.10_s2_z",
    "pf_FI_core_phi0.05_s3_y",
    "pf_TL_core_phi0.10_s0_y",  # the one case whose final set includes the appended blocks
]


def read_csv(path):
    with open(path, newline="") as handle:
        return list(csv.DictReader(handle))


def main(reference: str) -> None:
    cases_c = read_csv(os.path.join(reference, "data", "cases_coupled.csv"))
    cases_p = read_csv(os.path.join(reference, "data", "cases_prescribed_field.csv"))
    res_p = read_csv(os.path.join(reference, "results", "prescribed_field.csv"))
    res_c = read_csv(os.path.join(reference, "results", "coupled.csv"))
    res_x = read_csv(os.path.join(reference, "results", "fullpath_crosscheck.csv"))

    hash_cases, seen = [], set()
    for row in cases_c:
        if row["group"] != "canonical" or not row["raw_count_sha256_A"]:
            continue
        key = (row["structure"], row["transfer_axis"])
        if key in seen:
            continue
        seen.add(key)
        hash_cases.append({
            "structure": row["structure"],
            "axis": row["transfer_axis"],
            "microstructure": row["microstructure"],
            "H_m": float(row["H_m"]),
            "rays_per_patch_per_block": int(row["rays_per_patch_per_block"]),
            "seeds_A": [int(s) for s in row["seeds_A"].split(";")],
            "seeds_B": [int(s) for s in row["seeds_B"].split(";")],
            "phase_sha256_x": row["phase_sha256_x"],
            "raw_count_sha256_A": row["raw_count_sha256_A"],
            "raw_count_sha256_B": row["raw_count_sha256_B"],
        })

    prescribed = []
    for case_id in PRESCRIBED_CASES:
        row = next(r for r in cases_p if r["case_id"] == case_id)
        seeds = [int(row["seeds_A"]), int(row["seeds_B"])]
        if row["append_used"] == "true":
            seeds += [int(row["append_seeds_A"]), int(row["append_seeds_B"])]
        prescribed.append({
            "case_id": case_id,
            "microstructure": row["microstructure"],
            "axis": row["transfer_axis"],
            "H_m": float(row["H_m"]),
            "rays_per_patch_per_block": int(row["rays_per_patch_per_block"]),
            "final_seeds": seeds,
            "phase_sha256_x": row["phase_sha256_x"],
            "results": [
                {k: float(r[k]) for k in ("T_mean_K", "delta_T_K", "epsilon", "q_h_W_m2", "q_c_W_m2", "q_i_W_m2")}
                for r in res_p if r["case_id"] == case_id
            ],
        })

    coupled = [
        {k: (float(v) if k not in ("case_id", "group", "structure", "microstructure", "transfer_axis") else v)
         for k, v in r.items()}
        for r in res_c if r["group"] == "canonical"
    ]
    crosscheck = [
        {k: r[k] for k in ("structure", "microstructure", "transfer_axis", "T_mean_K", "delta_T_K", "epsilon",
                           "method", "q_h_W_m2", "q_c_W_m2", "q_i_W_m2")}
        for r in res_x if r["method"] == "first_hit_radiosity"
    ]

    # every case of the two case tables whose microstructure is shipped, with its archived rows
    shipped = {f[:-4] for f in os.listdir(os.path.join(reference, "data", "microstructures")) if f.endswith(".npz")}
    prescribed_cases = []
    for row in cases_p:
        prescribed_cases.append({
            **{k: row[k] for k in ("case_id", "group", "structure", "microstructure", "transfer_axis",
                                   "phase_sha256_x", "conditions")},
            "H_m": float(row["H_m"]),
            "rays_per_patch_per_block": int(row["rays_per_patch_per_block"]),
            **{k: [int(s) for s in row[k].split(";")]
               for k in ("seeds_A", "seeds_B", "append_seeds_A", "append_seeds_B")},
            "append_used": row["append_used"] == "true",
            "results": [
                {k: float(r[k]) for k in ("T_mean_K", "delta_T_K", "epsilon", "q_h_W_m2", "q_c_W_m2", "q_i_W_m2",
                                          "ab_wall_channel_rel_difference")}
                for r in res_p if r["case_id"] == row["case_id"]
            ],
        })
    archived_c = {r["case_id"]: r for r in res_c}
    coupled_cases = []
    for row in cases_c:
        if row["case_id"] not in archived_c:
            continue
        coupled_cases.append({
            # microstructures that are not shipped are regenerated from the microstructure table
            "shipped": row["microstructure"] in shipped,
            **{k: row[k] for k in ("case_id", "group", "structure", "microstructure", "transfer_axis",
                                   "phase_sha256_x", "raw_count_sha256_A", "raw_count_sha256_B")},
            **{k: float(row[k]) for k in ("H_m", "T_mean_K", "delta_T_K", "epsilon")},
            **{k: int(row[k]) for k in ("repeats_along_axis", "N_axis", "rays_per_patch_per_block")},
            **{k: [int(s) for s in row[k].split(";")] for k in ("seeds_A", "seeds_B")},
            "archived": {k: (v if k in ("case_id", "group", "structure", "microstructure", "transfer_axis") or v == ""
                             else float(v)) for k, v in archived_c[row["case_id"]].items()},
        })

    names = sorted({c["microstructure"] for c in hash_cases} | {c["microstructure"] for c in prescribed}
                   | {c["microstructure"] for c in prescribed_cases}
                   | {c["microstructure"] for c in coupled_cases if c["shipped"]})
    arrays = {}
    for name in names:
        path = os.path.join(reference, "data", "microstructures", f"{name}.npz")
        with np.load(path, allow_pickle=False) as stored:
            phases = np.asarray(stored["phases"], dtype=np.uint8)
        if phases.max() > 1:
            raise ValueError(f"{name}: phases are not 0/1")
        arrays[f"{name}__bits"] = np.packbits(phases.ravel())
        arrays[f"{name}__shape"] = np.asarray(phases.shape, dtype=np.int64)
        arrays[f"{name}__sha256"] = np.frombuffer(hashlib.sha256(phases.tobytes()).digest(), dtype=np.uint8)
    mc = {"mc_properties": [], "mc_penetration": []}
    for restart, sub in (("article", ""), ("hit_point", "hit_point_restart")):
        for key, table, fields in (
                ("mc_properties", "radiative_properties.csv", ("beta_per_cell", "beta_se_per_cell", "omega", "omega_se",
                                                                "g", "g_se")),
                ("mc_penetration", "penetration_response.csv", ("reflectance", "transmittance", "absorptance"))):
            for r in read_csv(os.path.join(reference, "results", sub, table)):
                if r["microstructure"] not in names:
                    continue
                mc[key].append({"restart": restart, "microstructure": r["microstructure"],
                                "transfer_axis": r["transfer_axis"], "epsilon": float(r["epsilon"]),
                                "n_rays": int(r["n_rays"]), **{f: float(r[f]) for f in fields}})
    float_cols = ("target_phi", "realized_phi", "interfacial_area_density_per_cell", "mean_beam_length_cells",
                  "mean_pore_chord_x_cells", "mean_pore_chord_y_cells", "mean_pore_chord_z_cells")
    micro = [{"id": r["id"], "structure": r["structure"], "set": r["set"], "N": int(r["N"]),
              "seed": int(r["seed"]) if r["seed"] != "" else None, "medoid": r["medoid"],
              "phase_sha256": r["phase_sha256"], **{k: float(r[k]) for k in float_cols}}
             for r in read_csv(os.path.join(reference, "data", "microstructures.csv"))]
    np.savez_compressed(os.path.join(HERE, "geometries.npz"), **arrays)
    with open(os.path.join(HERE, "cases.json"), "w") as handle:
        json.dump({"hash_cases": hash_cases, "prescribed_field": prescribed, "coupled_canonical": coupled,
                   "fullpath_crosscheck": crosscheck, "prescribed_cases": prescribed_cases,
                   "coupled_cases": coupled_cases, **mc, "microstructures": micro}, handle, indent=1)
    print(f"wrote {len(names)} geometries, {len(hash_cases)} hash cases, {len(prescribed)} prescribed cases")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
