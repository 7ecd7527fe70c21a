"""The source-level C++ drop-in (include/ifl_fluid.hpp).

tests/cxx/shim_main.cpp is a reference-style main() that uses the reference's
class API. It must compile against the header at every level (CPU test), and
on the GPU it must print exactly the reference's per-update() lines and end
with the oracle's density field bit for bit (sequential dot order).
"""
import os
import subprocess

import numpy as np
import pytest

import oracle_lib as ol

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
PKG = os.path.join(ROOT, "incremental-fluids_amd")
LIB = os.path.join(PKG, "libifl.so")


def build_shim(level, outdir):
    exe = os.path.join(outdir, f"shim_main{level}")
    cmd = ["g++", "-O2", "-std=c++11", "-Wall", "-Wextra", "-Werror", f"-DIFL_SOLVER_LEVEL={level}",
           "-I", os.path.join(ROOT, "include"), os.path.join(HERE, "cxx", "shim_main.cpp"),
           "-L", PKG, "-lifl", f"-Wl,-rpath,{PKG}", "-o", exe]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    return exe


@pytest.mark.skipif(not os.path.exists(LIB), reason="libifl.so not built")
@pytest.mark.parametrize("level", [3, 4, 5, 6, 7, 8])
def test_shim_compiles(level, tmp_path):
    assert os.path.exists(build_shim(level, str(tmp_path)))


SHIM_SCENE = {3: ol.REF_SCENES[3], 4: ol.REF_SCENES[4], 5: ol.REF_SCENES[5],
              6: dict(ol.REF_SCENES[7]), 7: ol.REF_SCENES[7], 8: ol.REF_SCENES[8]}
# shim_main's user-defined body: a sphere (static: moving bodies make the
# pressure solves of levels 5-7 diverge in the reference, 6:990)
USER_SPHERE = (ol.BODY_SPHERE, 0.3, 0.35, 0.15, 0.15, 0.3, 0.0, 0.0, 0.0)


@pytest.mark.gpu
@pytest.mark.parametrize("level,user", [(3, False), (4, False), (5, False), (6, False), (7, False), (8, False),
                                        (4, True), (5, True), (6, True), (7, True)])
def test_shim_matches_oracle(level, user, tmp_path):
    """user=True adds a SolidBody subclass defined in the caller's program,
    evaluated on the host through ifl_set_host_body_ops; its geometry is a
    sphere's, so the oracle runs the same scene with a SolidSphere."""
    w, h, steps = 64, 48, 6
    exe = build_shim(level, str(tmp_path))
    out = str(tmp_path / "d.bin")
    env = dict(os.environ, IFL_DOT_ORDER="sequential")
    print(f"shim level {level} {w}x{h} {steps} steps{' with a user-defined body' if user else ''}", flush=True)
    run = subprocess.run([exe, str(w), str(h), str(steps), out] + (["user"] if user else []), env=env,
                         capture_output=True, text=True, timeout=120)
    assert run.returncode == 0, run.stderr
    sc = SHIM_SCENE[level]
    bodies = list(sc["bodies"]) + ([USER_SPHERE] if user else [])
    orc = ol.OracleSolver(level, w, h, sc["params"], bodies, fresh=True)
    for k in range(steps):
        if sc["inflow"] is not None:
            orc.add_inflow(*sc["inflow"])
        orc.update(sc["dt"])
        if (k + 1) % 4 == 0:
            orc.bodies_update(sc["dt"])
    assert run.stdout == orc.log()
    d = np.fromfile(out, dtype=np.float64)
    want = orc.field(ol.qfield(ol.Q_D, ol.QF_SRC))
    assert d.tobytes() == want.tobytes()
    orc.close()
