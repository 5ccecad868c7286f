"""Oracle runs of many independent calls in worker processes (the
restatement keeps its spectral tables in file-scope state, so it is run in
separate processes, not threads). Test infrastructure only.

Workers are started with the "spawn" method: the parent may hold a HIP
context, which a forked child must not inherit.
"""
import multiprocessing as mp
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _work(job):
    if ROOT not in sys.path:
        sys.path.insert(0, ROOT)
    from oracle import oracle
    case, spec, mcica, variant = job
    return oracle.runes(case, spec, mcica, variant=variant)


def run_many(jobs, workers=None):
    """jobs: list of (case, spectral_file, mcica_file or None, variant or
    None). Returns the oracle outputs in order."""
    if workers is None:
        workers = max(1, min(8, (os.cpu_count() or 1)))
    if len(jobs) <= 1 or workers == 1:
        return [_work(j) for j in jobs]
    ctx = mp.get_context("spawn")
    with ctx.Pool(min(workers, len(jobs))) as pool:
        return pool.map(_work, jobs, chunksize=1)
