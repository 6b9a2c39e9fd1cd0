"""The hand-written reverse mode of the step arithmetic (geoswe_amd/csrc/gs_adjoint.h),
built for the host in float64, against central finite differences of the forward
functions (gs_device.h, same build): directional derivatives <obar, J d> vs <J^T obar, d>
on random face and update states in wet, shallow, dry and velocity-capped regimes.

A sample whose forward is not smooth within +-eps (a max/min switch, the dry cut-off)
shows up as a disagreement between the eps and eps/4 differences and is left out; at
least 90 % of the samples of each regime must be smooth and all of those must agree.
The host face is also checked against the NumPy oracle (the same function)."""
import ctypes as C
import subprocess
from pathlib import Path

import numpy as np
import pytest

from oracle import swe_oracle as O

NATIVE = Path(__file__).resolve().parent / "native"
LIB = NATIVE / "_build" / "libgs_vjp_host.so"
G, HMIN, HALF = 9.81, 1e-6, 1.5


@pytest.fixture(scope="module")
def lib():
    if not LIB.exists():
        subprocess.run(["make", "-C", str(NATIVE)], check=True, capture_output=True)
    h = C.CDLL(str(LIB))
    dp = C.POINTER(C.c_double)
    h.gsh_face.argtypes = [dp, C.c_int, C.c_double, C.c_double, C.c_double, dp]
    h.gsh_face_vjp.argtypes = [dp, dp, C.c_int, C.c_double, C.c_double, C.c_double, dp]
    h.gsh_update.argtypes = [dp, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double, C.c_int, dp]
    h.gsh_update_vjp.argtypes = [dp, dp, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double, C.c_int, dp]
    return h


def ptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def face(lib, x):
    x = np.ascontiguousarray(x, np.float64)
    out = np.zeros((len(x), 6))
    lib.gsh_face(ptr(x), len(x), HALF, G, HMIN, ptr(out))
    return out


def face_vjp(lib, x, ob):
    x = np.ascontiguousarray(x, np.float64)
    ob = np.ascontiguousarray(ob, np.float64)
    gx = np.zeros((len(x), 10))
    lib.gsh_face_vjp(ptr(x), ptr(ob), len(x), HALF, G, HMIN, ptr(gx))
    return gx


def face_states(rng, n, regime):
    bL = rng.uniform(-1, 1, n)
    bR = bL + rng.uniform(-0.8, 0.8, n)
    gz = rng.uniform(-0.3, 0.3, (n, 2))
    if regime == "wet":
        h = rng.uniform(0.05, 2.0, (n, 2))
    elif regime == "shallow":
        h = 10 ** rng.uniform(-4, -1, (n, 2))
    else:                                           # one side dry
        h = np.stack([rng.uniform(0.05, 1.0, n), np.zeros(n)], 1)
        flip = rng.random(n) < 0.5
        h[flip] = h[flip][:, ::-1]
    u = rng.uniform(-1.5, 1.5, (n, 4))
    u[h[:, 0] <= HMIN, :2] = 0.0
    u[h[:, 1] <= HMIN, 2:] = 0.0
    return np.column_stack([bL, bR, gz, h, u])


@pytest.mark.parametrize("regime", ["wet", "shallow", "dry_side"])
def test_face_vjp_matches_central_differences(lib, regime):
    rng = np.random.default_rng({"wet": 1, "shallow": 2, "dry_side": 3}[regime])
    n = 4000
    x = face_states(rng, n, regime)
    ob = rng.normal(size=(n, 5))
    d = rng.normal(size=(n, 10))
    if regime == "dry_side":                         # keep the dry side dry (its depth is a kink)
        d[:, 4][x[:, 4] == 0] = 0.0
        d[:, 5][x[:, 5] == 0] = 0.0
        d[:, 6:8][x[:, 4] == 0] = 0.0
        d[:, 8:10][x[:, 5] == 0] = 0.0
    scale = np.maximum(np.abs(x[:, 4:6]).min(axis=1), 1e-3) if regime == "shallow" else np.ones(n)
    eps = 1e-6 * scale[:, None]

    def dd(e):
        fp = face(lib, x + e * d)[:, :5]
        fm = face(lib, x - e * d)[:, :5]
        return np.sum(ob * (fp - fm), axis=1) / (2 * e[:, 0])

    fd1, fd2 = dd(eps), dd(eps / 4)
    gate = face(lib, x)[:, 5] > 0
    g = np.sum(face_vjp(lib, x, ob) * d, axis=1)
    tol = 1e-5 * (1 + np.abs(fd2))
    smooth = gate & (np.abs(fd1 - fd2) < 1e-6 * (1 + np.abs(fd2)))
    assert smooth.mean() > 0.9, smooth.mean()
    bad = smooth & (np.abs(g - fd2) > tol)
    assert not bad.any(), (int(bad.sum()), x[bad][:3], g[bad][:3], fd2[bad][:3])


def test_face_vjp_each_input(lib):
    """Unit directions: every one of the 10 inputs has its own correct cotangent."""
    rng = np.random.default_rng(4)
    x = face_states(rng, 2000, "wet")
    ob = rng.normal(size=(2000, 5))
    gx = face_vjp(lib, x, ob)
    for k in range(10):
        d = np.zeros_like(x)
        d[:, k] = 1.0
        e = 1e-6
        fd1 = np.sum(ob * (face(lib, x + e * d)[:, :5] - face(lib, x - e * d)[:, :5]), 1) / (2 * e)
        fd2 = np.sum(ob * (face(lib, x + e / 4 * d)[:, :5] - face(lib, x - e / 4 * d)[:, :5]), 1) / (e / 2)
        smooth = np.abs(fd1 - fd2) < 1e-6 * (1 + np.abs(fd2))
        assert smooth.mean() > 0.9
        assert np.all(np.abs(gx[smooth, k] - fd2[smooth]) < 1e-5 * (1 + np.abs(fd2[smooth]))), k


def test_host_face_is_the_oracle_face(lib):
    rng = np.random.default_rng(5)
    x = np.concatenate([face_states(rng, 500, r) for r in ("wet", "shallow", "dry_side")])
    got = face(lib, x)
    T = np.float64
    hLf, hRf, gate, srcL, srcR = O.srm_face(x[:, 0], x[:, 1], x[:, 2], x[:, 3], x[:, 4], x[:, 5], T(HALF), T(G),
                                            T(HMIN))
    Fh, FN, FT = O.hllc(hLf, x[:, 6], x[:, 7], hRf, x[:, 8], x[:, 9], T(G), T(HMIN))
    assert np.array_equal(got[:, 5] > 0, gate)
    for k, want in enumerate((Fh, FN, FT, srcL, srcR)):
        assert np.array_equal(got[:, k], want), k


@pytest.mark.parametrize("case", ["plain", "capped", "linear"])
def test_update_vjp_matches_central_differences(lib, case):
    rng = np.random.default_rng({"plain": 6, "capped": 7, "linear": 8}[case])
    n = 4000
    h1 = 10 ** rng.uniform(-3, 0.3, n)
    vel = rng.uniform(-3, 3, (n, 2)) if case != "capped" else rng.uniform(-60, 60, (n, 2))
    x = np.column_stack([h1, h1 * vel[:, 0], h1 * vel[:, 1], rng.uniform(0.01, 0.12, n)])
    dt, vcap, quad = 0.7, 15.0, 0 if case == "linear" else 1
    ob = rng.normal(size=(n, 3))
    d = rng.normal(size=(n, 4)) * np.column_stack([h1, h1, h1, np.full(n, 0.01)])

    def upd(xx):
        xx = np.ascontiguousarray(xx)
        out = np.zeros((n, 3))
        lib.gsh_update(ptr(xx), n, dt, G, HMIN, vcap, quad, ptr(out))
        return out

    gx = np.zeros((n, 4))
    xc = np.ascontiguousarray(x)
    obc = np.ascontiguousarray(ob)
    lib.gsh_update_vjp(ptr(xc), ptr(obc), n, dt, G, HMIN, vcap, quad, ptr(gx))
    e = 1e-6
    fd1 = np.sum(ob * (upd(x + e * d) - upd(x - e * d)), 1) / (2 * e)
    fd2 = np.sum(ob * (upd(x + e / 4 * d) - upd(x - e / 4 * d)), 1) / (e / 2)
    g = np.sum(gx * d, 1)
    smooth = np.abs(fd1 - fd2) < 1e-6 * (1 + np.abs(fd2))
    assert smooth.mean() > 0.9
    assert np.all(np.abs(g[smooth] - fd2[smooth]) < 1e-5 * (1 + np.abs(fd2[smooth])))
    if case == "capped":
        modU = np.hypot(x[:, 1], x[:, 2]) / x[:, 0]
        assert (modU > vcap).mean() > 0.5
