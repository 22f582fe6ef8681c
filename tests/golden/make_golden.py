"""Write the golden fixtures of tests/golden/.

These are RESTATEMENT-GENERATED, not reference output: the reference (Rust + wgpu)
cannot run in this environment and its repository holds no golden vectors for this
path.  They freeze the oracle's output (after its pins in tests/test_oracle_*.py) so
that later changes to either side are caught, and they carry the workload inputs as
the reference CLI would read them.

    python tests/golden/make_golden.py
"""
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent.parent))

import oracle  # noqa: E402
from sdf2mesh_amd import tape as T  # noqa: E402

CASES = {
    # name: (program, res, bounds, lattice, lag)
    "torus_r16_b2_ref": (T.torus_example(), 16, 2.0, False, True),
    "csg64_r24_b2_ref": (T.csg64(), 24, 2.0, False, True),
    "csg64_r24_b2_lattice_nolag": (T.csg64(), 24, 2.0, True, False),
    "random_csg3_r20_b2_ref": (T.random_csg(3), 20, 2.0, False, True),
    "mandelbulb4_r16_b2.5_ref": (T.mandelbulb(4), 16, 2.5, False, True),
}


def main():
    for name, (node, res, b, lattice, lag) in CASES.items():
        h = np.float32(b) * np.float32(0.5)
        m = oracle.mesh(node, [-h] * 3, [h] * 3, [res] * 3, lattice=lattice, lag=lag)
        np.savez_compressed(HERE / f"{name}.npz", tape=node.tape(), res=res, bounds=b,
                            lattice=lattice, lag=lag, vertices=m.vertices, keys=m.keys,
                            signs=m.signs, triangles=m.triangles,
                            invalid_quads=m.invalid_quads)
    m = oracle.mesh(T.torus_example(), [-1.0] * 3, [1.0] * 3, [12] * 3)
    oracle.write_stl(m.vertices, m.triangles, HERE / "torus_r12_b2_ref.stl")
    (HERE / "torus.sdf3d").write_text(T.TORUS_SDF3D)
    (HERE / "csg64.sdf3d").write_text(T.csg64().sdf3d_source())
    for s in (1, 2, 3):
        (HERE / f"random_csg{s}.sdf3d").write_text(T.random_csg(s).sdf3d_source())
    (HERE / "mandelbulb8.frag").write_text(T.mandelbulb_glsl(8))


if __name__ == "__main__":
    main()
