"""The JNI shim (cartogenesis_amd/jni/cgocean_jni.c) compiled against the test-only jni.h and driven
through a fake JVM environment (tests/jni/fake_jvm.c), whose Get<Type>ArrayElements hand out copies
that Release copies back unless the mode is JNI_ABORT. No JDK exists here, so the Kotlin side
(HipOcean.kt) is not compiled; what these tests pin is the C shim a JVM would load:

- every native releases every element array it takes (no pinned copy left), on success and on
  every refusal;
- outputs reach the Java arrays (released with mode 0) and inputs are never written back;
- on the CPU, nOpen returns 0 and hands the library's reason back in the String[1], and every
  native given a null handle returns the library's CG_ERR_ARG (or 0 for a handle);
- on the GPU, the natives give what the C ABI gives the Python mirror on the same inputs, to the
  bit: the relaxation seam against the oracle, the whole solve, a resident round's plan (forcing,
  solve, read, velocities), a set cancel flag, the device name.
"""
import ctypes
from pathlib import Path

import numpy as np
import pytest

import cartogenesis_amd as cg

HARNESS = Path(__file__).resolve().parent / "jni" / "libjni_harness.so"
K = {"B": np.int8, "S": np.int16, "I": np.int32, "J": np.int64, "F": np.float32, "D": np.float64}


@pytest.fixture(scope="module")
def jni():
    if not HARNESS.exists():
        pytest.fail("tests/jni/libjni_harness.so is not built (__graft_entry__.build builds it)")
    h = ctypes.CDLL(str(HARNESS))
    vp = ctypes.c_void_p
    for name, res, args in (("fj_env", vp, []), ("fj_array", vp, [ctypes.c_char, ctypes.c_int]),
                            ("fj_data", vp, [vp]), ("fj_len", ctypes.c_int, [vp]),
                            ("fj_object_at", vp, [vp, ctypes.c_int]), ("fj_string", ctypes.c_char_p, [vp]),
                            ("fj_direct_buffer", vp, [vp, ctypes.c_int]), ("fj_free", None, [vp]),
                            ("fj_outstanding", ctypes.c_int, []), ("fj_bad_releases", ctypes.c_int, [])):
        f = getattr(h, name)
        f.restype, f.argtypes = res, args
    return h


class Java:
    """Java arrays and strings in the fake JVM, and the shim's natives with their JNI signatures."""

    def __init__(self, h):
        self.h = h
        self.env = h.fj_env()
        self.owned = []

    # Object handles are ctypes.c_void_p (never plain ints, which native() passes as jint).
    def array(self, kind, values):
        values = np.ascontiguousarray(values, dtype=K[kind])
        o = ctypes.c_void_p(self.h.fj_array(kind.encode(), values.size))
        ctypes.memmove(self.h.fj_data(o), values.ctypes.data, values.nbytes)
        self.owned.append(o)
        return o

    def objects(self, n):
        o = ctypes.c_void_p(self.h.fj_array(b"L", n))
        self.owned.append(o)
        return o

    def read(self, o, kind):
        n = self.h.fj_len(o)
        buf = (ctypes.c_char * (n * np.dtype(K[kind]).itemsize)).from_address(self.h.fj_data(o))
        return np.frombuffer(bytes(buf), dtype=K[kind]).copy()

    def native(self, name, res, *args):
        f = getattr(self.h, "Java_com_cartogenesis_desktop_HipOcean_" + name)
        f.restype = res
        types = [ctypes.c_void_p, ctypes.c_void_p]
        vals = [self.env, None]
        for a in args:
            if isinstance(a, float):
                types.append(ctypes.c_double)
            elif isinstance(a, bool):
                types.append(ctypes.c_ubyte)
            elif isinstance(a, tuple):  # ("j", v): a jlong
                types.append(ctypes.c_int64)
                a = a[1]
            elif isinstance(a, int):
                types.append(ctypes.c_int)
            else:
                types.append(ctypes.c_void_p)
            vals.append(a)
        f.argtypes = types
        out = f(*vals)
        assert self.h.fj_outstanding() == 0, f"{name} left element copies unreleased"
        assert self.h.fj_bad_releases() == 0, f"{name} released an array it did not take"
        return out

    def close(self):
        for o in self.owned:
            self.h.fj_free(o)
        self.owned.clear()


def j(v):
    return ("j", int(v))


def test_natives_refuse_null_handles_and_release_every_array(jni):
    java = Java(jni)
    try:
        reason = java.objects(1)
        if java.native("nOpen", ctypes.c_int64, 0, reason) == 0:
            msg = jni.fj_string(jni.fj_object_at(reason, 0))
            assert msg and len(msg) > 0
        n = 8 * 16
        f, s, o = java.array("F", np.ones(n)), java.array("F", np.ones(n)), java.array("F", np.full(n, 7.0))
        assert java.native("nRelax", ctypes.c_int, j(0), j(1), f, s, o, 4, None) == cg.CG_ERR_ARG
        assert np.all(java.read(o, "F") == 7.0)  # refused: the output keeps its values
        water = java.array("B", np.ones(n))
        ws = [java.array("F", np.full(n, 0.25)) for _ in range(4)]
        assert java.native("nUpload", ctypes.c_int64, j(0), 16, 8, water, *ws, None) == 0
        levels = java.array("J", [])
        info = java.array("D", np.full(3, -1.0))
        assert java.native("nSolve", ctypes.c_int, j(0), levels, f, None, 1e-8, True, 3, o, info, None) == cg.CG_ERR_ARG
        assert java.native("nPlanCreate", ctypes.c_int64, j(0), java.array("J", [1, 2]), True) == 0
        assert java.native("nPlanCreate", ctypes.c_int64, j(0), levels, True) == 0
        assert java.native("nPlanSetForcing", ctypes.c_int, j(0), f) == cg.CG_ERR_ARG
        d = java.array("D", np.zeros(n))
        assert java.native("nPlanSetStress", ctypes.c_int, j(0), 1e4, 1e4, d, d, None) == cg.CG_ERR_ARG
        assert java.native("nPlanSetBodies", ctypes.c_int, j(0), java.array("I", np.zeros(n)), 1) == cg.CG_ERR_ARG
        assert java.native("nPlanSetStart", ctypes.c_int, j(0), s) == cg.CG_ERR_ARG
        assert java.native("nPlanSolve", ctypes.c_int, j(0), 1e-8, 3, False, info) == cg.CG_ERR_ARG
        assert np.array_equal(java.read(info, "D"), [0.0, 0.0, 0.0])
        assert java.native("nPlanRead", ctypes.c_int, j(0), o) == cg.CG_ERR_ARG
        assert java.native("nPlanVelocities", ctypes.c_int, j(0), 1e4, 1e4, o, o) == cg.CG_ERR_ARG
        wts = java.array("S", np.zeros(16))
        assert java.native("nPlanSetCoarseNet", ctypes.c_int, j(0), 4, 32, wts, None, 1.0) == cg.CG_ERR_ARG
        assert java.native("nPlanSetCoarseNet", ctypes.c_int, j(0), 4, 32, None, None, 1.0) == cg.CG_ERR_ARG
        err = java.native("nLastError", ctypes.c_void_p, j(0))
        assert jni.fj_string(err) == b"no context"
        java.native("nClose", None, j(0))
        java.native("nRelease", None, j(0), j(5))
        java.native("nPlanDestroy", None, j(0))
    finally:
        java.close()


@pytest.mark.gpu
def test_natives_on_the_device_match_the_c_abi(jni, acc):
    import oracle
    from conftest import to_oracle
    from cartogenesis_amd import CG_METHOD_REFINED
    from cartogenesis_amd.ocean import Plan
    from cartogenesis_amd.problem import circulation_problem

    java = Java(jni)
    p = circulation_problem(256, seed=9)  # the seam declines grids under 65536 cells
    st = p.levels[0]
    n = st.cells
    reason = java.objects(1)
    ctx = java.native("nOpen", ctypes.c_int64, 0, reason)
    assert ctx != 0
    try:
        name = jni.fj_string(java.native("nName", ctypes.c_void_p, j(ctx)))
        assert b"gfx950" in name
        handles = []
        for s in p.levels:
            h = java.native("nUpload", ctypes.c_int64, j(ctx), s.cells_across, s.cells_down,
                            java.array("B", s.is_water), java.array("F", s.east_weight),
                            java.array("F", s.west_weight), java.array("F", s.north_weight),
                            java.array("F", s.south_weight), java.array("D", s.centre_weight))
            assert h != 0
            handles.append(h)
        # The seam: relaxation passes, against the oracle to the bit; start is not written.
        start = np.where(st.is_water != 0, 1.0, 0.0).astype(np.float32)
        js, jo = java.array("F", start), java.array("F", np.zeros(n))
        print("OCEAN jni relax", flush=True)
        assert java.native("nRelax", ctypes.c_int, j(ctx), j(handles[0]), java.array("F", st.forcing), js, jo, 10,
                           None) == cg.CG_OK
        want = oracle.relax(to_oracle(st), start, 10)
        assert np.array_equal(java.read(jo, "F").view(np.uint32), want.view(np.uint32))
        assert np.array_equal(java.read(js, "F"), start)
        # A set cancel flag: the batch is refused and the output left as it was.
        flag = (ctypes.c_int * 1)(1)
        buf = ctypes.c_void_p(jni.fj_direct_buffer(ctypes.addressof(flag), 4))
        jo2 = java.array("F", np.full(n, 3.0))
        rc = java.native("nRelax", ctypes.c_int, j(ctx), j(handles[0]), java.array("F", st.forcing), js, jo2, 50, buf)
        jni.fj_free(buf)
        assert rc == cg.CG_CANCELLED
        assert np.all(java.read(jo2, "F") == 3.0)
        # The whole solve, as the C ABI gives it to the Python mirror.
        levels = java.array("J", handles)
        out, info = java.array("F", np.zeros(n)), java.array("D", np.zeros(3))
        print("OCEAN jni solve", flush=True)
        assert java.native("nSolve", ctypes.c_int, j(ctx), levels, java.array("F", st.forcing), None, 1e-10, True,
                           CG_METHOD_REFINED, out, info, None) == cg.CG_OK
        ref = acc.solve_levels(p.levels, None, 1e-10, True, CG_METHOD_REFINED)
        assert np.array_equal(java.read(out, "F").view(np.uint32), ref.values.view(np.uint32))
        assert java.read(info, "D")[0] == ref.cycles
        # A resident round's plan: forcing, solve, read, velocities, against Plan on the same ladder.
        plan = java.native("nPlanCreate", ctypes.c_int64, j(ctx), levels, True)
        assert plan != 0
        try:
            assert java.native("nPlanSetForcing", ctypes.c_int, j(plan), java.array("F", st.forcing)) == cg.CG_OK
            pinfo = java.array("D", np.zeros(3))
            print("OCEAN jni plan", flush=True)
            assert java.native("nPlanSolve", ctypes.c_int, j(plan), 1e-10, CG_METHOD_REFINED, False, pinfo) == cg.CG_OK
            got = java.array("F", np.zeros(n))
            assert java.native("nPlanRead", ctypes.c_int, j(plan), got) == cg.CG_OK
            ve, vn = java.array("F", np.zeros(n)), java.array("F", np.zeros(n))
            assert java.native("nPlanVelocities", ctypes.c_int, j(plan), float(p.dx), float(p.dy), ve, vn) == cg.CG_OK
            py = Plan(acc, p.levels)
            try:
                pi = py.solve(1e-10, CG_METHOD_REFINED)
                _, x32 = py.read()
                e2, n2 = py.velocities(p.dx, p.dy)
            finally:
                py.close()
            assert java.read(pinfo, "D")[0] == pi.cycles
            assert np.array_equal(java.read(got, "F").view(np.uint32), x32.view(np.uint32))
            assert np.array_equal(java.read(ve, "F").view(np.uint32), np.asarray(e2, np.float32).view(np.uint32))
            assert np.array_equal(java.read(vn, "F").view(np.uint32), np.asarray(n2, np.float32).view(np.uint32))
        finally:
            java.native("nPlanDestroy", None, j(plan))
        for h in handles:
            java.native("nRelease", None, j(ctx), j(h))
    finally:
        java.native("nClose", None, j(ctx))
        java.close()
