"""The file-driver replay (socrates_amd.replay) on the CPU: its readers and
writer, its option checks, and the oracle on the replayed calls against the
reference driver's own known-good outputs (examples/netcdf: CRM and CIRC
columns, GA7 and GA9 files, McICA, mixed column and clear sky), to the
printed digits."""
import os

import numpy as np
import pytest

from oracle import oracle
from replay_cases import RUNS, kgo_tolerance, run_args
from socrates_amd import replay


@pytest.mark.parametrize("name", sorted(RUNS))
def test_oracle_reproduces_driver_kgo(name):
    base, argv, surf, kgo = run_args(name)
    case, spec, mcica, grid = replay.build(base, argv, surface=surf)
    out = oracle.runes(case, spec, mcica)
    res = replay.diagnostics(case, out, grid, case.i_source)
    field, vals, ulp = replay.read_kgo(kgo)
    err = np.abs(res[field] - vals)
    print("%s %s: max %.3f of the last printed digit" % (name, field, (err / ulp).max()))
    assert (err <= kgo_tolerance(vals, ulp)).all(), (err / ulp).max()


def test_cdl_round_trip(tmp_path):
    rng = np.random.default_rng(3)
    vals = rng.normal(0.0, 300.0, (11, 3))
    vals[2, 1] = 0.0
    p = str(tmp_path / "x.hrts")
    replay.write_cdl(p, "hrts", "k.d-1", "heating rates", np.linspace(1.0, 1e5, 11), np.zeros(1),
                     np.zeros(3), vals)
    name, back, ulp = replay.read_kgo(p)
    assert name == "hrts" and back.shape == vals.shape
    assert (np.abs(back - vals) <= 0.5 * ulp * (1 + 1e-9)).all()
    f = replay.read_field(p)
    assert f.data.shape == (11, 3) and np.allclose(f.plev, np.linspace(1.0, 1e5, 11), rtol=1e-6)


def test_reads_netcdf_and_cdl_alike(tmp_path):
    """A netCDF example field written out as CDL reads back to the same
    values (to the 6 printed digits)."""
    base, argv, surf, kgo = run_args("mcica_lw_ga7")
    f = replay.read_field(base + ".t")
    p = str(tmp_path / "t.cdl")
    replay.write_cdl(p, "t", "K", "temperature", f.plev, f.lat, f.lon, f.data)
    g = replay.read_field(p)
    assert np.allclose(g.data, f.data, rtol=5e-6, atol=0) and np.allclose(g.plev, f.plev, rtol=5e-6, atol=0)


@pytest.mark.parametrize("opts,msg", [
    ("-g 2", "gas overlap"),
    ("-C 6 -K 3", "cloud scheme"),
    ("-t 2", "two-stream"),
    ("-a", "not on the two-stream"),
    ("-R 2 9", "band range"),
])
def test_refuses_options_off_the_path(opts, msg):
    base, argv, surf, kgo = run_args("circ6_lw_ga7_clr")
    argv = [a for a in argv if a not in ("-g", "4")] + opts.split() + ([] if "-g" in opts else ["-g", "4"])
    if opts.startswith("-R"):
        argv = [a for a in argv] + ["-R", "2", "9"]
    with pytest.raises(replay.ReplayError, match=msg):
        replay.build(base, argv, surface=surf)


def test_requires_rescaling():
    base, argv, surf, kgo = run_args("circ6_lw_ga7_clr")
    with pytest.raises(replay.ReplayError, match=r"\+R"):
        replay.build(base, [a for a in argv if a != "+R"], surface=surf)
