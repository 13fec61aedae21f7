"""The matrix-core M2L engine (fastmultipole_amd/csrc/m2l_mx.hpp), emulated lane by lane on the CPU
(tests/m2l_mx_emul.cpp: per-lane values as arrays of 64 doubles, the MFMA as v_mfma_f64_16x16x4f64's
product with its lane layout), against the oracle's multipole_to_local (src/translate.jl:519-1412).

Each case translates the multipoles of a set of source boxes (P2M of random bodies, the oracle's) to one
target box, in batches of 4 (Lamb-Helmholtz) or 8 interactions, and compares the summed local expansion
and every interaction's expansion order with the oracle's. The same source is what runs on the GPU; the
GPU parity tests check the device build end to end."""
import ctypes as C
import os
import subprocess
import tempfile

import numpy as np
import pytest

from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def emul():
    d = tempfile.mkdtemp(prefix="mxemul")
    so = os.path.join(d, "libmxemul.so")
    subprocess.run(["g++", "-O1", "-std=c++17", "-ffp-contract=off", "-fPIC", "-shared", "-o", so,
                    os.path.join(ROOT, "tests", "m2l_mx_emul.cpp")], check=True)
    L = C.CDLL(so)
    dp = np.ctypeslib.ndpointer(np.float64, flags="C")
    ip = np.ctypeslib.ndpointer(np.int32, flags="C")
    L.mx_emul_m2l.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, C.c_int, dp, C.c_int, dp, dp,
                              C.c_int64, dp, ip, C.POINTER(C.c_int), C.POINTER(C.c_int)]
    L.mx_emul_m2l.restype = C.c_int
    return L


def nh(p):
    return (p + 1) * (p + 2) // 2


def run_case(L, LH, cap, P, kind, rel, eps, BE, nsrc=11, seed=0):
    rng = np.random.default_rng(seed)
    tc = np.array([0.3, -0.2, 0.1])
    tbox = np.array([0.25, 0.25, 0.25])
    tg = O.geom(tc, radius=0.43, box=tbox, max_influence=2.5)
    sgs, exps = [], []
    for _ in range(nsrc):
        d = rng.normal(size=3)
        sc = tc + d / np.linalg.norm(d) * rng.uniform(1.5, 3.0)
        e = np.zeros(4 * nh(P))
        for _ in range(30):
            x = sc + rng.uniform(-0.25, 0.25, 3)
            if LH:
                O.p2m(O.KIND_VORTEX, x - sc, rng.normal(size=3), P, out=e)
            else:
                O.p2m(O.KIND_SOURCE, x - sc, rng.uniform(0, 1, 1), P, out=e)
        sgs.append(O.geom(sc, radius=0.43, box=np.array([0.25] * 3)))
        exps.append(e)
    ref = np.zeros(4 * nh(P))
    orders_ref = []
    for g, e in zip(sgs, exps):
        out, Pi, _ = O.m2l(tg, g, e, P, LH=LH, err_kind=kind, relative=rel, eps=eps, BE=BE)
        ref += out
        orders_ref.append(Pi)
    out = np.zeros(4 * nh(P))
    orders = np.zeros(nsrc, np.int32)
    pm, ok = C.c_int(0), C.c_int(0)
    rc = L.mx_emul_m2l(int(LH), cap, P, kind, int(rel), eps, int(BE), tg, nsrc, np.ascontiguousarray(sgs),
                       np.ascontiguousarray(exps), 4 * nh(P), out, orders, C.byref(pm), C.byref(ok))
    assert rc == 0
    return out, ref, orders, np.array(orders_ref), pm.value


CASES = [
    # LH, cap, P, kind, relative, eps, BE
    (False, 8, 8, O.ERR_NONE, False, 0.0, True),
    (True, 8, 8, O.ERR_NONE, False, 0.0, True),
    (True, 20, 20, O.ERR_NONE, False, 0.0, True),
    (False, 20, 16, O.ERR_NONE, False, 0.0, True),
    (True, 20, 5, O.ERR_NONE, False, 0.0, True),
] + [(LH, 20, 20, kind, rel, eps, BE)
     for kind in (O.ERR_ROTATED_GRADIENT, O.ERR_POWER_POTENTIAL, O.ERR_POWER_GRADIENT)
     for LH in (False, True)
     for rel, eps, BE in ((True, 1e-7, True), (False, 1e-9, False))]


@pytest.mark.slow
@pytest.mark.parametrize("LH,cap,P,kind,rel,eps,BE", CASES)
def test_emulated_engine_matches_oracle(emul, LH, cap, P, kind, rel, eps, BE):
    out, ref, orders, orders_ref, pmax = run_case(emul, LH, cap, P, kind, rel, eps, BE)
    np.testing.assert_array_equal(orders, orders_ref)
    assert pmax == orders_ref.max()
    err = np.abs(out - ref).max() / np.abs(ref).max()
    assert err < 1e-12, err


@pytest.mark.slow
def test_emulated_engine_partial_batch_and_order_cap(emul):
    """Batches with empty slots (nsrc not a multiple of the slot count) and interactions that reach the
    order cap without meeting the tolerance (the reference's P = p, error_success = false)."""
    out, ref, orders, orders_ref, _ = run_case(emul, True, 20, 12, O.ERR_POWER_GRADIENT, False, 1e-14, False,
                                                nsrc=6, seed=3)
    np.testing.assert_array_equal(orders, orders_ref)
    assert (orders_ref == 12).any()
    assert np.abs(out - ref).max() / np.abs(ref).max() < 1e-12
