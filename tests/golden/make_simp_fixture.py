"""Write simp100_oracle.json: the FP64 oracle's compliance and gradient for SIMP
shell designs 0 and 1 of the 100x100 MITC-4 batch over 15 m (10,201 nodes,
10,000 quads, 61,206 DOF; corners and mid-edge nodes pinned, -1.5 kN/node,
E = rho^3 * 30 GPa).  The density gradient follows from the E gradient by the
chain rule dC/drho = dC/dE * 3 rho^2 * 30 GPa (the reference applies the same
chain rule outside the library, Examples/shells_topo_shape.ipynb).

The oracle takes ~1 minute per design here, so the test compares against this
fixture: C, full-vector norms and seeded samples of u, dC/dz and dC/drho.  Run
from the repository root:
    python tests/golden/make_simp_fixture.py
"""
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from jaxsso_amd import workloads as wl  # noqa: E402
from oracle import gradient as og  # noqa: E402
from oracle import solve as osolve  # noqa: E402
from oracle import system  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "simp100_oracle.json")
N_SAMPLE = 256


def one(design):
    st = wl.simp_shell(design)
    K = system.assemble_K(st.crds, st.cnct_beamcols, st.prop_beamcols, st.cnct_quads, st.prop_quads)
    u, hist = osolve.solve_reduced(K, st.known_id, st.nodal_loads, return_info=True)
    eg = og.element_gradients(st.crds, st.cnct_beamcols, st.prop_beamcols, st.cnct_quads, st.prop_quads,
                              u, 0.5 * u)
    dc = og.node_gradient(st.n_node, st.cnct_beamcols, st.cnct_quads, eg)
    drho = eg["quad_props"][:, 1] * 3.0 * st.rho ** 2 * 30e9
    rng = np.random.default_rng(1000 + design)
    iu = np.sort(rng.choice(st.ndof, N_SAMPLE, replace=False))
    iz = np.sort(rng.choice(st.design_nodes, N_SAMPLE, replace=False))
    ir = np.sort(rng.choice(st.n_quad, N_SAMPLE, replace=False))
    return {
        "design": design,
        "refinement_residuals": [float(h) for h in hist],
        "C": float(0.5 * st.nodal_loads @ u),
        "norm_u": float(np.linalg.norm(u)),
        "norm_dz_free": float(np.linalg.norm(dc[st.design_nodes, 2])),
        "norm_drho": float(np.linalg.norm(drho)),
        "u_index": iu.tolist(), "u": u[iu].tolist(),
        "dz_node": iz.tolist(), "dz": dc[iz, 2].tolist(),
        "drho_elem": ir.tolist(), "drho": drho[ir].tolist(),
    }


def main():
    t0 = time.time()
    out = {"workload": "100x100 SIMP MITC-4 shell, 15 m, corners + mid-edges pinned, -1500 N per free node",
           "oracle": "oracle.solve.solve_reduced + oracle.gradient (complex-step element derivatives)",
           "designs": [one(0), one(1)]}
    out["seconds"] = time.time() - t0
    with open(OUT, "w") as fh:
        json.dump(out, fh, indent=0)
    print("wrote", OUT, [d["C"] for d in out["designs"]], "in", round(out["seconds"], 1), "s")


if __name__ == "__main__":
    main()
