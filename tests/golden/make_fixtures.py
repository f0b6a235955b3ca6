"""Writes tests/golden/ fixtures from the reference's committed data files
(run here, where /root/reference is mounted; the GPU box only reads the
fixtures).  Only data is stored: mesh arrays read from the reference's
.msh files and the expected numbers its result CSVs hold.

  naca0012_coarse.npz  nodes/elements of cases/meshes/naca0012_2.5d/coarse.msh
  naca0012_{coarse,medium}_bc.npz  nodes, elements and the wall / farfield
                       boundary faces of the same meshes (level-set wake)
  golden.json          expected values with their reference sources
"""

import json
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parents[1]
REF = Path("/root/reference")
sys.path.insert(0, str(ROOT))

from up3d_amd.mesh.reader import read_mesh  # noqa: E402


def main():
    m = read_mesh(REF / "cases/meshes/naca0012_2.5d/coarse.msh")
    np.savez_compressed(HERE / "naca0012_coarse.npz", nodes=m.nodes,
                        elements=m.elements)
    for lv in ("coarse", "medium"):
        m = read_mesh(REF / f"cases/meshes/naca0012_2.5d/{lv}.msh")
        np.savez_compressed(HERE / f"naca0012_{lv}_bc.npz", nodes=m.nodes,
                            elements=m.elements,
                            wall=m.boundary_faces["wall"],
                            farfield=m.boundary_faces["farfield"],
                            wake=m.boundary_faces["wake"],
                            symmetry=m.boundary_faces["symmetry"])
    g73 = {}
    for ln in (REF / "cases/demo/p7_diff_flux/results/g73_constructed.csv").read_text().splitlines()[1:]:
        k, v = ln.split(",")
        g73[k] = float(v)
    stats = {}
    for ln in (REF / "cases/meshes/naca0012_2.5d/coarse_stats.csv").read_text().splitlines()[1:]:
        k, v = ln.split(",")
        stats[k] = float(v)
    col = {}
    for ln in (REF / "cases/demo/p0_infrastructure/results/g03_coloring.csv").read_text().splitlines()[1:]:
        p = ln.split(",")
        col[p[0]] = {"n_elements": int(p[1]), "n_colors": int(p[2]),
                     "min_class": int(p[4]), "max_class": int(p[5])}
    out = {
        "g73_constructed": {"source": "cases/demo/p7_diff_flux/results/g73_constructed.csv",
                            "values": g73},
        "naca_coarse_stats": {"source": "cases/meshes/naca0012_2.5d/coarse_stats.csv",
                              "values": stats},
        "g03_coloring": {"source": "cases/demo/p0_infrastructure/results/g03_coloring.csv",
                         "values": col},
        "naca_medium_colors": {"source": "cases/demo/p3_subsonic/results/assembly_debt.csv:6",
                               "values": {"n_colors": 27}},
        # level-set wake, NACA0012 2.5D, alpha 2 deg, incompressible, vortex
        # far field, wall-adjacent pressure-equality Kutta (the reference's
        # measured table, coarse / medium / fine)
        "ls_gamma_naca_a2": {"source": "tests/test_b4_te_control_volume.py:186-189",
                             "values": {"levelset": [0.1177, 0.1191, 0.1197],
                                        "conforming": [0.1175, 0.1200, 0.1202],
                                        "b3_mass_row": [0.2074, 0.1760, 0.1704],
                                        "levels": ["coarse", "medium", "fine"]}},
        # level-set wake, NACA0012 2.5D, alpha 1.25 deg, per-side upwinding
        # C = 1.5, vortex far field, converged circulation (the reference's
        # measured table: M0.60 upwind on/off, then the M0.65-0.75 coarse and
        # M0.70 medium LS column)
        # conforming explicit-circulation wake, coupled (phi, Gamma) Newton,
        # alpha 1.25 deg, C = 1.5, vortex far field (the reference's
        # same-mesh "conf Newton" column)
        "conforming_newton_naca_a125": {"source": "docs/design_track_b.md:873-879",
                                        "values": {"coarse": {"0.65": 0.1038, "0.70": 0.1151,
                                                              "0.75": 0.1377},
                                                   "medium": {"0.70": 0.1190}}},
        "ls_gamma_naca_a125": {"source": "docs/design_track_b.md:799,873-878",
                               "values": {"coarse": {"0.60": 0.0968, "0.65": 0.1043,
                                                     "0.70": 0.1158, "0.75": 0.1391},
                                          "medium": {"0.70": 0.1193}}},
    }
    (HERE / "golden.json").write_text(json.dumps(out, indent=1))


if __name__ == "__main__":
    main()
