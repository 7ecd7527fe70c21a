"""Writes the golden fixtures in this directory from the reference's own
solvers (oracle/_ref, compiled from the reference sources by oracle/Makefile).

One fixture per solver level: the reference main()'s scene at a small grid,
the solver state after a few steps and the captured stdout lines. Run from the
repository root: python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import oracle_lib as ol  # noqa: E402

CASES = {3: (40, 32, 10), 4: (40, 32, 8), 5: (40, 32, 8), 6: (40, 32, 8), 7: (40, 32, 8), 8: (32, 24, 6)}


def main():
    for level, (w, h, steps) in CASES.items():
        sc = ol.REF_SCENES[level]
        ref = ol.RefSolver(level, w, h, sc["params"], sc["bodies"], fresh=True)
        ol.run_scene(ref, level, steps)
        out = {"w": w, "h": h, "steps": steps, "log": np.frombuffer(ref.log().encode(), dtype=np.uint8)}
        for f in ol.state_fields(level):
            v = ref.field(f)
            if v is not None:
                if f in (ol.F_POSX, ol.F_POSY) or ol.F_PROP_D <= f <= ol.F_PROP_V:
                    v = v[:ref.particle_count()]
                out[f"f{f}"] = v
        if level == 8:
            out["particle_count"] = np.array([ref.particle_count()])
        path = os.path.join(HERE, f"solver{level}_{w}x{h}_{steps}steps.npz")
        np.savez_compressed(path, **out)
        print(path, os.path.getsize(path))
        ref.close()


if __name__ == "__main__":
    main()
