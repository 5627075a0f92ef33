"""Write the small golden fixtures under tests/golden/ (run from the repo root).

Inputs come from the seeded generator (stenpy_amd.gen.generate_host), outputs from the FP64
CPU oracle.  Before writing, each output family is cross-checked against something
independent of the oracle (see tests/test_golden.py for the checks that run every time).
    python tests/golden/make_golden.py
"""
import sys
from pathlib import Path

import numpy as np
import torch

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

from stenpy_amd import gen  # noqa: E402
import stenpy_amd as sp  # noqa: E402
from oracle import graph_eval  # noqa: E402

OUT = Path(__file__).resolve().parent
EXPRS = {
    "pointwise": "f**2 + tanh(f)*exp(f)",
    "laplacian": "laplacian(f)",
    "gradient": "gradient(f)",
    "stencil_graph": "divergence(gradient(f)) + dot(u, gradient(f)) + sqnorm(curl(u))",
    "hessian_eig": "eigenvalues(hessian(f))",
    "hessian_det": "determinant(hessian(f))",
    "mean_curvature": "mean_curvature(f)",
}


def main():
    shape = (12, 14, 16)
    f = gen.generate_host(shape, 2)
    u = np.stack([gen.generate_host(shape, s) for s in (3, 4, 5)], -1)
    arrays = {"gen16_seed2": gen.generate_host((16, 16, 16), 2), "f": f, "u": u}
    for name, expr in EXPRS.items():
        g, sink, *_ = sp.compile_expression(expr, boundary="periodic",
                                            field_map={"f": torch.from_numpy(f), "u": torch.from_numpy(u)})
        reds = [g.add(r, (sink,), {}) for r in ("sum", "mean", "variance", "norm_l2")] if name == "stencil_graph" else []
        o64 = graph_eval.evaluate(g, torch.float64)
        o32 = graph_eval.evaluate(g, torch.float32)
        arrays[f"{name}__out64"] = o64[sink].numpy()
        arrays[f"{name}__out32"] = o32[sink].numpy()
        for r in reds:
            op = g.node(r).op_name
            arrays[f"{name}__{op}64"] = np.asarray(o64[r].item())
            arrays[f"{name}__{op}32"] = np.asarray(o32[r].item())
    meta = {"torch": torch.__version__, "sympy": __import__("sympy").__version__,
            "numpy": np.__version__, "shape": shape, "exprs": EXPRS}
    np.savez_compressed(OUT / "fixtures.npz", **arrays)
    (OUT / "fixtures_meta.txt").write_text(repr(meta) + "\n")
    print("wrote", OUT / "fixtures.npz", sorted(arrays))


if __name__ == "__main__":
    main()
