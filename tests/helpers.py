"""Shared helpers of the tests: parity metric and small runs."""
import numpy as np


def rel_err(a, b):
    """Norm-wise relative error per component: max|a_c - b_c| / max|a_c|."""
    a = np.asarray(a)
    b = np.asarray(b)
    out = []
    for c in range(a.shape[0]):
        den = np.max(np.abs(a[c]))
        num = np.max(np.abs(a[c] - b[c]))
        out.append(num / den if den > 0 else num)
    return np.array(out)


def run_oracle(case, nsteps, state=None):
    from oracle.pyoracle import OracleLevel
    L = OracleLevel(case)
    L.set_state(case.make_state() if state is None else state)
    L.step(nsteps)
    return L.get_state(), L.info()


def run_gpu(case, nsteps, mode=1, state=None, tile=None, boxes=None):
    from mesoflow_amd import runtime
    ctx = runtime.Context(case)
    L = runtime.Level(ctx, boxes)
    if tile:
        L.set_tile(*tile)
    L.set_state(case.make_state() if state is None else state)
    L.step(nsteps, mode)
    out = L.get_state(), L.info()
    L.close()
    ctx.close()
    return out
