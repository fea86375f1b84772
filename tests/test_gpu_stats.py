"""Statistics of the dimer in the FP32-mixed and the FP64 forms: 48 independent replicas
of a passive dimer (sigma 2, d 4.5, no catalytic conversion) in L=16 boxes, 40 collision
cycles of equilibration, then 120 sampled cycles (tau 0.5, N_MD 50).

* Equipartition: each component of the dimer's centre-of-mass velocity has variance
  T / (m1 + m2).  Checked in both forms to 20 % (about four standard errors for this
  sample).
* The two forms agree within error bars: the velocity variance and the mean squared
  displacement of the centre of mass at lags of 5, 10 and 20 cycles, compared with four
  standard errors of their difference (the error of each form from the spread over
  replicas, whose trajectories are independent).
"""
import ctypes as C

import numpy as np
import pytest

from rmpcdmd_amd import capi
from rmpcdmd_amd.sim import SingleDimer

pytestmark = pytest.mark.gpu

R, L, EQ, SAMPLES = 48, 16, 40, 120


def colloids(sim, r):
    """unfolded positions and velocities of the two colloids of replica r"""
    sol, col, box, cells, st = sim.state(r)

    def get(ptr, dtype):
        a = np.empty(6, dtype=dtype)
        capi.call("rmpcd_memcpy", a.ctypes.data_as(C.c_void_p), C.c_void_p(ptr), a.nbytes, 1)
        return a.reshape(3, 2).T

    pos, vel, img = get(col.pos, np.float64), get(col.vel, np.float64), get(col.image, np.int32)
    edges = np.array(box.edges[:])
    return pos + img * edges, vel, np.array(st.mass[:])


def trajectory(form):
    sim = SingleDimer(replica_L=[(L, L, L)] * R, prob=0.0, form=form, seed=4242, n_md=50, tau=0.5)
    sim.run(EQ)
    X = np.empty((SAMPLES, R, 3))
    V = np.empty((SAMPLES, R, 3))
    for k in range(SAMPLES):
        sim.run(1)
        for r in range(R):
            x, v, m = colloids(sim, r)
            X[k, r] = (m[0] * x[0] + m[1] * x[1]) / (m[0] + m[1])
            V[k, r] = (m[0] * v[0] + m[1] * v[1]) / (m[0] + m[1])
    sim.close()
    return X, V, m[0] + m[1]


@pytest.fixture(scope="module")
def runs(gpu):
    return {f: trajectory(f) for f in (capi.FP64, capi.FP32)}


def per_replica_var(V):
    return (V ** 2).mean(axis=(0, 2))  # mean over samples and components, one per replica


def msd(X, lag):
    d = X[lag:] - X[:-lag]
    return (d ** 2).sum(axis=2).mean(axis=0)  # one per replica


def test_equipartition(runs):
    for form, (X, V, M) in runs.items():
        v = per_replica_var(V)
        ratio = v.mean() * M  # T = 1
        sem = v.std(ddof=1) / np.sqrt(R) * M
        print("form %d: <V^2> M / T = %.3f +- %.3f" % (form, ratio, sem))
        assert abs(ratio - 1) < 0.2


def test_fp32_and_fp64_agree_within_error_bars(runs):
    (X64, V64, M), (X32, V32, _) = runs[capi.FP64], runs[capi.FP32]
    a, b = per_replica_var(V64), per_replica_var(V32)
    err = np.sqrt(a.var(ddof=1) / R + b.var(ddof=1) / R)
    print("velocity variance: FP64 %.4e FP32 %.4e (error of the difference %.1e)" % (a.mean(), b.mean(), err))
    assert abs(a.mean() - b.mean()) < 4 * err
    for lag in (5, 10, 20):
        a, b = msd(X64, lag), msd(X32, lag)
        err = np.sqrt(a.var(ddof=1) / R + b.var(ddof=1) / R)
        print("MSD at %d cycles: FP64 %.4f FP32 %.4f (error of the difference %.4f)" % (lag, a.mean(), b.mean(), err))
        assert a.mean() > 0 and abs(a.mean() - b.mean()) < 4 * err
    assert msd(X64, 20).mean() > msd(X64, 5).mean()
