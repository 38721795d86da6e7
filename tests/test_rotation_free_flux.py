"""The mixed-precision kernel evaluates the rotated hybrid HLLC-HLL flux
without rotation matrices (hydro_math.hpp, riemann_f): solving along a unit
normal n in the original frame with u_n = v.n, and n2 = (d - n1_d n1) / s.
This test restates that formulation in float64 numpy and checks it against
the oracle's literal rotate-solve-rotate-back solver
(Riemann.cpp:154-244) on the reference-derived KAT inputs of
tests/golden/riemann.npz: the two must agree to rounding (the identity is
exact in real arithmetic), so a wrong sign of n2, a dropped p n term in the
momentum flux or a wrong star-state velocity fails it."""
import os

import numpy as np
import pytest

import oracle

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
DEN, PRE, HII, ADV, VEL = 0, 1, 2, 3, 4


def _side(q, g):
    rho, p, v = q[:, DEN], q[:, PRE], q[:, VEL:VEL + 3]
    ke = 0.5 * rho * (v * v).sum(1)
    return dict(rho=rho, p=p, x=q[:, HII], a=q[:, ADV], v=v, sr=np.sqrt(rho), a2=g * p / rho, ke=ke,
                E=p / (g - 1.0) + ke)


def _flux(s, un, n, g2):
    m = s["rho"] * un
    F = np.empty((un.shape[0], 7))
    F[:, DEN] = m
    F[:, PRE] = un * (g2 * s["p"] + s["ke"])
    F[:, HII] = m * s["x"]
    F[:, ADV] = m * s["a"]
    F[:, VEL:] = m[:, None] * s["v"] + s["p"][:, None] * n
    return F


def _cons(s):
    U = np.empty((s["rho"].shape[0], 7))
    U[:, DEN] = s["rho"]
    U[:, PRE] = s["E"]
    U[:, HII] = s["rho"] * s["x"]
    U[:, ADV] = s["rho"] * s["a"]
    U[:, VEL:] = s["rho"][:, None] * s["v"]
    return U


def _speeds(L, R, unl, unr):
    is_ = 1.0 / (L["sr"] + R["sr"])
    ut = (L["sr"] * unl + R["sr"] * unr) * is_
    d = np.sqrt((L["sr"] * L["a2"] + R["sr"] * R["a2"]) * is_ + 0.5 * L["sr"] * R["sr"] * is_ * is_ * (unr - unl) ** 2)
    return ut - d, ut + d


def _hll(L, R, n, g2):
    unl, unr = (L["v"] * n).sum(1), (R["v"] * n).sum(1)
    Sl, Sr = _speeds(L, R, unl, unr)
    Fl, Fr = _flux(L, unl, n, g2), _flux(R, unr, n, g2)
    Ul, Ur = _cons(L), _cons(R)
    with np.errstate(all="ignore"):
        Fh = (Sr[:, None] * Fl - Sl[:, None] * Fr + (Sl * Sr)[:, None] * (Ur - Ul)) / (Sr - Sl)[:, None]
    return np.where((Sr <= 0)[:, None], Fr, np.where((Sl >= 0)[:, None], Fl, Fh))


def _hllc(L, R, n, g2):
    unl, unr = (L["v"] * n).sum(1), (R["v"] * n).sum(1)
    Sl, Sr = _speeds(L, R, unl, unr)
    wl, wr = L["rho"] * (Sl - unl), R["rho"] * (Sr - unr)
    with np.errstate(all="ignore"):
        Sc = (R["p"] - L["p"] + wl * unl - wr * unr) / (wl - wr)
    sub = ~(Sl >= 0) & ~(Sr <= 0)
    left = (Sl >= 0) | (sub & (Sc >= 0))
    K = {k: np.where(left[:, None] if L[k].ndim == 2 else left, L[k], R[k]) for k in L}
    un = np.where(left, unl, unr)
    S = np.where(left, Sl, Sr)
    F = _flux(K, un, n, g2)
    with np.errstate(all="ignore"):
        w = S - un
        A = K["rho"] * w / (S - Sc)
        dsc = Sc - un
        dU = np.empty_like(F)
        dU[:, DEN] = A - K["rho"]
        dU[:, PRE] = A * (K["E"] / K["rho"] + dsc * (Sc + K["p"] / (K["rho"] * w))) - K["E"]
        dU[:, HII] = (A - K["rho"]) * K["x"]
        dU[:, ADV] = (A - K["rho"]) * K["a"]
        dU[:, VEL:] = A[:, None] * (K["v"] + dsc[:, None] * n) - K["rho"][:, None] * K["v"]
        Fs = F + S[:, None] * dU
    return np.where(sub[:, None], Fs, F)


def rotation_free(Ql, Qr, dims, g):
    """Rotated hybrid flux along axis dims[i] of each face, rotation-free."""
    L, R = _side(Ql, g), _side(Qr, g)
    g2 = g / (g - 1.0)
    n = Ql.shape[0]
    d = np.zeros((n, 3))
    d[np.arange(n), dims] = 1.0
    n1 = R["v"] - L["v"]
    n1d = n1[np.arange(n), dims]
    n1 = np.where((n1d < 0)[:, None], -n1, n1)
    nn = (n1 * n1).sum(1)
    pure_hllc = np.sqrt(nn) < 1e-6
    with np.errstate(all="ignore"):
        n1 = n1 / np.sqrt(nn)[:, None]
        n1d = n1[np.arange(n), dims]
        s2 = (n1 * n1).sum(1) - n1d * n1d
        s2 = np.maximum(s2, 0.0)
        pure_hll = ~pure_hllc & (s2 == 0)
        s = np.sqrt(s2)
        n2 = (d - n1d[:, None] * n1) / s[:, None]
    hyb = ~pure_hllc & ~pure_hll
    n1s = np.where(hyb[:, None], n1, d)
    n2s = np.where(hyb[:, None], n2, d)
    F1 = _hll(L, R, n1s, g2)
    F2 = _hllc(L, R, n2s, g2)
    Fh = n1d[:, None] * F1 + s[:, None] * F2
    out = np.where(pure_hllc[:, None], _hllc(L, R, d, g2), np.where(pure_hll[:, None], _hll(L, R, d, g2), Fh))
    return out


@pytest.mark.parametrize("solver", ["RotatedHLLC", "HLL", "HLLC"])
def test_rotation_free_matches_rotated_solver(solver):
    f = np.load(os.path.join(GOLD, "riemann.npz"))
    Ql, Qr, dims, g = f["Ql"], f["Qr"], f["dims"].astype(int), float(f["gamma"])
    ref, st = oracle.riemann(solver, Ql, Qr, dims, g)
    ok = st == 0
    L, R = _side(Ql, g), _side(Qr, g)
    n = Ql.shape[0]
    d = np.zeros((n, 3))
    d[np.arange(n), dims] = 1.0
    if solver == "RotatedHLLC":
        got = rotation_free(Ql, Qr, dims, g)
    elif solver == "HLL":
        got = _hll(L, R, d, g / (g - 1.0))
    else:
        got = _hllc(L, R, d, g / (g - 1.0))
    # relative to the face's flux scale (momentum flux ~ p + rho v^2)
    scale = np.abs(ref[ok]).max(1, keepdims=True) + 1e-300
    err = np.abs(got[ok] - ref[ok]) / scale
    assert np.isfinite(got[ok]).all()
    assert np.median(err) < 1e-14, np.median(err)
    assert err.max() < 1e-9, (err.max(), np.argmax(err.max(1)))
