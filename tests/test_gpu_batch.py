"""Independent solver instances stepped concurrently from several host
threads (each handle has its own HIP stream; the library releases the GIL in
every call): every instance must still match its own oracle bit for bit.
Instances follow the batched workload's draws (solver 4, sides 64..192 here
so the oracle stays quick)."""
import threading

import pytest

import oracle_lib as ol
from gpu_util import compare_fields

pytestmark = pytest.mark.gpu


def test_concurrent_instances_bit_exact():
    import ifl_fluids as F
    import ifl_scenes as S
    insts = [i for i in S.config4_instances(64) if max(i["w"], i["h"]) <= 192][:6]
    assert len(insts) == 6
    steps = 4
    gpus = [F.FluidSolver(4, i["w"], i["h"], i["params"], [F.SolidBox(*b[1:]) for b in i["bodies"]], dot_order=1)
            for i in insts]
    errors = []

    def run(ks):
        try:
            for _ in range(steps):
                for k in ks:
                    for inf in insts[k]["inflows"]:
                        gpus[k].add_inflow(*inf)
                    gpus[k].update(insts[k]["dt"])
        except Exception as e:
            errors.append(repr(e))

    th = [threading.Thread(target=run, args=([k for k in range(6) if k % 3 == t],)) for t in range(3)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    assert not errors, errors
    for k, inst in enumerate(insts):
        orc = ol.OracleSolver(4, inst["w"], inst["h"], inst["params"], inst["bodies"])
        for _ in range(steps):
            for inf in inst["inflows"]:
                orc.add_inflow(*inf)
            orc.update(inst["dt"])
        bad = compare_fields(orc, gpus[k], ol.state_fields(4), where=f"instance {k} {inst['w']}x{inst['h']}")
        assert not bad, "\n".join(bad[:6])
        print(f"instance {k} {inst['w']}x{inst['h']}: bit-exact after {steps} steps", flush=True)
        orc.close()
        gpus[k].close()
