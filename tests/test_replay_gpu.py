"""The file-driver replay on the GPU (socrates_amd.replay.run through the C
ABI): each example run of the reference's driver reproduced to its printed
digits in the FP64 form, the FP64 form against the oracle to the parity
tolerances, and the mixed form within its stated bounds (0.05 W m-2, 0.01
K/day) of the FP64 oracle; the output files are written and read back."""
import os

import numpy as np
import pytest

from oracle import oracle
from replay_cases import RUNS, kgo_tolerance, run_args
from socrates_amd import replay
from test_gpu_parity import check_fp64, check_mixed

pytestmark = pytest.mark.gpu


def _fluxes(res, case):
    """runes-style outputs from the replay's fields, for the parity checks."""
    up, down = res["uflx"], res["vflx"]
    net = down - up
    return {"flux_up": up, "flux_down": down,
            "flux_direct": res.get("sflx", np.zeros_like(up)),
            "heating_rate": (net[:-1] - net[1:]) / case.layer_heat_capacity}


@pytest.mark.parametrize("name", sorted(RUNS))
def test_replay_reproduces_driver_kgo(name, tmp_path):
    base, argv, surf, kgo = run_args(name)
    print("replay", name, flush=True)
    out = str(tmp_path / "run")
    res = replay.run(base, argv, out_base=out, precision="fp64", surface=surf)
    field, vals, ulp = replay.read_kgo(kgo)
    err = np.abs(res[field] - vals)
    print("%s %s: max %.3f of the last printed digit" % (name, field, (err / ulp).max()))
    assert (err <= kgo_tolerance(vals, ulp)).all(), (err / ulp).max()
    # the written file holds the same field
    f2, back, ulp2 = replay.read_kgo(out + "." + field)
    assert f2 == field and (np.abs(back - res[field]) <= 0.5 * ulp2 * (1 + 1e-9)).all()
    # FP64 against the oracle, mixed within its bounds
    case, spec, mcica, grid = replay.build(base, argv, surface=surf)
    ref = oracle.runes(case, spec, mcica)
    gpu = _fluxes(res, case)
    src = "sw" if case.i_source == 1 else "lw"
    check_fp64(gpu, ref, name, src, case.layer_heat_capacity)
    mixed = replay.run(base, argv, precision="mixed", surface=surf)
    check_mixed(_fluxes(mixed, case), ref, name + " mixed")
