"""ctypes binding of the CPU oracle (test infrastructure only).

The oracle is a FP64 C++ restatement of FastMultipole's FMM path
(src/tree.jl, src/interaction_list.jl, src/bodytomultipole.jl, src/rotate.jl,
src/translate.jl, src/evaluate_expansions.jl, src/fmm.jl and the user kernels
of test/gravitational.jl and test/vortex.jl). Only tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg may import this
module, and only as the checker; the product path never touches it.

Parity status of each restated function is recorded in the header of
tests/test_oracle_golden.py (which fixture or closed form pins it).
"""
import ctypes as C
import os
import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

KIND_SOURCE = 0   # Point{Source}, test/gravitational.jl
KIND_VORTEX = 1   # Point{Vortex}, test/vortex.jl
KIND_DIPOLE = 2   # Point{Dipole}: P2M src/bodytomultipole.jl:431-463, direct kernel stated in oracle_fmm.cpp
KIND_VORTEX_GAUSS = 3   # Gaussian-core vortex (sigma = buffer row 4), stated in oracle_fmm.cpp
# constant-strength panels (test/panels.jl; P2M src/bodytomultipole.jl:643-831);
# rows [centroid(3), radius, strength, v1(3), v2(3), v3(3), v4(3) for quadrilaterals]
KIND_PANEL_SOURCE_TRI, KIND_PANEL_DIPOLE_TRI, KIND_PANEL_SOURCE_QUAD, KIND_PANEL_DIPOLE_QUAD = 4, 5, 6, 7
# Point{SourceVortex} (src/bodytomultipole.jl:491-496): rows [x, y, z, sigma, q, Gx, Gy, Gz]; nearfield stated
# in oracle_fmm.cpp (the point source's terms plus the singular vortex's)
KIND_SOURCE_VORTEX = 8
# Panel{SourceDipole} (test/panels.jl:323-429): rows [centroid(3), radius, source strength, dipole strength, vertices]
KIND_PANEL_SOURCE_DIPOLE_TRI, KIND_PANEL_SOURCE_DIPOLE_QUAD = 9, 10
METHOD_BARBA, METHOD_SELFTUNING, METHOD_SELFTUNING_TREESTOP = 0, 1, 2
ERR_NONE, ERR_ROTATED_GRADIENT, ERR_POWER_POTENTIAL, ERR_POWER_GRADIENT = 0, 1, 2, 3

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip64 = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
_ip32 = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_ic = np.ctypeslib.ndpointer(dtype=np.intc, flags="C_CONTIGUOUS")


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built; run `make -C oracle`")
        L = C.CDLL(path)
        L.orc_fgs_create.restype = C.c_void_p
        L.orc_fgs_create.argtypes = [C.c_int, C.c_int64, _dp, _ic, _dp]
        L.orc_fgs_free.argtypes = [C.c_void_p]
        L.orc_fgs_free.restype = None
        L.orc_fgs_n_leaves.argtypes = [C.c_void_p]
        L.orc_fgs_n_leaves.restype = C.c_int64
        L.orc_fgs_leaf_info.argtypes = [C.c_void_p, _ip64]
        for f in ("orc_fgs_self", "orc_fgs_nonself"):
            getattr(L, f).argtypes = [C.c_void_p, C.c_int64, _dp]
        for f in ("orc_fgs_nonself_rows", "orc_fgs_leaf_bodies"):
            getattr(L, f).argtypes = [C.c_void_p, C.c_int64, _ip64]
        L.orc_fgs_solve.argtypes = [C.c_void_p, _dp, C.c_int, C.c_double, _dp, _dp, _dp]
        L.orc_fgs_solve.restype = C.c_int
        L.orc_threads.argtypes = [C.c_int]
        L.orc_threads.restype = C.c_int
        L.orc_run_create.restype = C.c_void_p
        L.orc_run_create.argtypes = [C.c_int, C.c_int64, _dp, C.c_int64, _dp, _ic, _dp]
        L.orc_run_create_tuple.restype = C.c_void_p
        L.orc_run_create_tuple.argtypes = [C.c_int, _ic, _ip64, C.c_void_p, _dp, _dp, _ic, _dp]
        for f in ("orc_run_free", "orc_run_fmm", "orc_run_trees", "orc_run_lists"):
            getattr(L, f).argtypes = [C.c_void_p]
            getattr(L, f).restype = None
        for f in ("orc_n_branches", "orc_n_levels", "orc_n_leaves", "orc_n_list"):
            getattr(L, f).argtypes = [C.c_void_p, C.c_int]
            getattr(L, f).restype = C.c_int64
        L.orc_get_branches.argtypes = [C.c_void_p, C.c_int, _ip64, _dp]
        L.orc_get_levels.argtypes = [C.c_void_p, C.c_int, _ip64]
        L.orc_get_leaf_index.argtypes = [C.c_void_p, C.c_int, _ip64]
        L.orc_get_sort_index.argtypes = [C.c_void_p, C.c_int, _ip64]
        L.orc_get_buffer.argtypes = [C.c_void_p, C.c_int, _dp]
        L.orc_get_expansions.argtypes = [C.c_void_p, C.c_int, _dp]
        L.orc_get_list.argtypes = [C.c_void_p, C.c_int, _ip32]
        L.orc_get_m2l_orders.argtypes = [C.c_void_p, _ip32]
        L.orc_get_m2l_margins.argtypes = [C.c_void_p, _dp]
        L.orc_get_targets.argtypes = [C.c_void_p, _dp]
        L.orc_pmax.argtypes = [C.c_void_p]
        L.orc_pmax.restype = C.c_int
        L.orc_error_success.argtypes = [C.c_void_p]
        L.orc_error_success.restype = C.c_int
        L.orc_direct.argtypes = [C.c_int, C.c_int64, _dp, C.c_int64, _dp, C.c_int, C.c_int, C.c_int]
        L.orc_direct_sampled.argtypes = [C.c_int, C.c_int64, _ip64, _dp, _dp, C.c_int64, _dp, C.c_int, C.c_int, C.c_int]
        L.orc_tables.argtypes = [C.c_int, _dp, _dp, _dp, _dp, _dp]
        L.orc_hs_fresh.argtypes = [C.c_int, _dp]
        L.orc_pnms.argtypes = [C.c_int, C.c_double, C.c_double, _dp]
        L.orc_update_Ts.argtypes = [C.c_double, C.c_int, _dp]
        L.orc_eimphis.argtypes = [C.c_double, C.c_int, _dp]
        L.orc_zeta_sign.argtypes = [C.c_double, C.c_int, C.c_int, _dp]
        L.orc_eta_sign.argtypes = [C.c_double, C.c_int, C.c_int, _dp]
        L.orc_regular_harmonics.argtypes = [C.c_double, C.c_double, C.c_double, C.c_int, _dp]
        L.orc_irregular_harmonics.argtypes = [C.c_double, C.c_double, C.c_double, C.c_int, _dp]
        L.orc_c2s.argtypes = [C.c_double, C.c_double, C.c_double, _dp]
        L.orc_factorial.argtypes = [C.c_int]
        L.orc_factorial.restype = C.c_double
        for f in ("orc_rotate_multipole_y", "orc_rotate_local_y", "orc_rotate_z", "orc_translate_multipole_z",
                  "orc_translate_m2l_z", "orc_translate_local_z"):
            getattr(L, f).argtypes = [_dp, C.c_double, C.c_int, C.c_int, _dp]
        L.orc_transform_lh_multipole.argtypes = [_dp, C.c_double, C.c_int]
        L.orc_transform_lh_local.argtypes = [_dp, C.c_double, C.c_int]
        L.orc_m2m.argtypes = [_dp, _dp, _dp, C.c_int, C.c_int, _dp]
        L.orc_l2l.argtypes = [_dp, _dp, _dp, C.c_int, C.c_int, _dp]
        L.orc_m2l.argtypes = [_dp, _dp, _dp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, C.c_int, _dp,
                              C.POINTER(C.c_int)]
        L.orc_m2l.restype = C.c_int
        L.orc_l2p.argtypes = [_dp, _dp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, _dp]
        L.orc_p2m.argtypes = [C.c_int, _dp, _dp, C.c_int, _dp]
        L.orc_p2m_panel.argtypes = [C.c_int, _dp, _dp, C.c_int, _dp]
        _LIB = L
    return _LIB


def threads(n=0):
    """Set the oracle's OpenMP thread count (n > 0) and return the count in effect."""
    return lib().orc_threads(int(n))


def nh(p):
    return (p + 1) * (p + 2) // 2


def length_Ts(p):
    return (p + 1) * (p + 2) * (2 * p + 3) // 6


def length_Hs(p):
    n = p + 1
    return n * (n + 1) * (n + 2) // 6


def data_per_body(kind):
    if kind == KIND_SOURCE_VORTEX:
        return 8
    if kind in (KIND_PANEL_SOURCE_DIPOLE_TRI, KIND_PANEL_SOURCE_DIPOLE_QUAD):
        return 6 + 3 * (4 if kind == KIND_PANEL_SOURCE_DIPOLE_QUAD else 3)
    if kind >= KIND_PANEL_SOURCE_TRI:
        return 5 + 3 * (4 if kind >= KIND_PANEL_SOURCE_QUAD else 3)
    return 5 if kind == KIND_SOURCE else 7


def _options(p, leaf_size_source, leaf_size_target, method, shrink_recenter, error_kind, relative, eps, BE,
             scalar_potential, gradient, hessian, farfield, nearfield, self_induced, multipole_acceptance):
    if leaf_size_target is None:
        leaf_size_target = leaf_size_source
    oi = np.array([p, leaf_size_source, leaf_size_target, method, int(shrink_recenter), error_kind, int(relative),
                   int(BE), int(scalar_potential), int(gradient), int(hessian), int(farfield), int(nearfield),
                   int(self_induced)], dtype=np.intc)
    od = np.array([multipole_acceptance, eps], dtype=np.float64)
    return oi, od


def pack_tuple(kinds, datas):
    """Concatenate systems system-major into one buffer of R rows per body, R the largest data_per_body of
    the kinds and at least 7: each system's rows [x, y, z, radius, strength rows...] first, zeros after."""
    R = max([7] + [data_per_body(k) for k in kinds])
    rows = []
    for k, d in zip(kinds, datas):
        d = np.asarray(d, dtype=np.float64)
        assert d.shape[1] == data_per_body(k)
        r = np.zeros((len(d), R))
        r[:, :d.shape[1]] = d
        rows.append(r)
    return np.ascontiguousarray(np.concatenate(rows, axis=0))


class Run:
    """One oracle FMM evaluation (trees, lists, passes) on one target and one source system.

    positions: (N,3); source_data: (N, data_per_body) in the reference's source
    buffer row order [x,y,z,radius,strength...] (src/compatibility.jl:5-20).
    """

    def __init__(self, kind, target_xyz, source_data, p=5, leaf_size_source=20, leaf_size_target=None,
                 method=METHOD_SELFTUNING_TREESTOP, shrink_recenter=True, error_kind=ERR_NONE, relative=False,
                 eps=0.0, BE=True, scalar_potential=False, gradient=True, hessian=False, farfield=True,
                 nearfield=True, self_induced=True, multipole_acceptance=0.4):
        L = lib()
        self.kind = kind
        t = np.zeros((len(target_xyz), 16))
        t[:, :3] = target_xyz
        s = np.ascontiguousarray(source_data, dtype=np.float64)
        assert s.shape[1] == data_per_body(kind)
        self.nt, self.ns = len(t), len(s)
        if leaf_size_target is None:
            leaf_size_target = leaf_size_source
        oi, od = _options(p, leaf_size_source, leaf_size_target, method, shrink_recenter, error_kind, relative, eps, BE,
                          scalar_potential, gradient, hessian, farfield, nearfield, self_induced, multipole_acceptance)
        self._h = L.orc_run_create(kind, self.nt, np.ascontiguousarray(t.ravel()), self.ns,
                                   np.ascontiguousarray(s.ravel()), oi, od)
        self.p = p

    @classmethod
    def tuple(cls, kinds, datas, p=5, leaf_size_source=20, leaf_size_target=None, leaf_sizes=None,
              method=METHOD_SELFTUNING_TREESTOP,
              shrink_recenter=True, error_kind=ERR_NONE, relative=False, eps=0.0, BE=True, scalar_potential=False,
              gradient=True, hessian=False, farfield=True, nearfield=True, self_induced=True, multipole_acceptance=0.4):
        """fmm! on a tuple of systems (each system both target and source): one tree over all bodies,
        each source system's nearfield and P2M with its own kernel. Targets come back concatenated in the
        order of `datas`. leaf_sizes: one leaf size per system (the reference's SVector leaf_size), else
        leaf_size_source for each."""
        L = lib()
        self = cls.__new__(cls)
        s = pack_tuple(kinds, datas)
        counts = np.array([len(d) for d in datas], dtype=np.int64)
        t = np.zeros((len(s), 16))
        t[:, :3] = s[:, :3]
        self.kind = KIND_VORTEX if any(k in (KIND_VORTEX, KIND_VORTEX_GAUSS, KIND_SOURCE_VORTEX) for k in kinds) else KIND_DIPOLE
        self.nt = self.ns = len(s)
        oi, od = _options(p, leaf_size_source, leaf_size_target, method, shrink_recenter, error_kind, relative, eps, BE,
                          scalar_potential, gradient, hessian, farfield, nearfield, self_induced, multipole_acceptance)
        self._ls = None if leaf_sizes is None else np.ascontiguousarray(leaf_sizes, dtype=np.int64)
        assert self._ls is None or len(self._ls) == len(kinds)
        self._h = L.orc_run_create_tuple(len(kinds), np.ascontiguousarray(kinds, dtype=np.intc), counts,
                                         None if self._ls is None else self._ls.ctypes.data,
                                         np.ascontiguousarray(t.ravel()), np.ascontiguousarray(s.ravel()), oi, od)
        self.p = p
        return self

    def __del__(self):
        if getattr(self, "_h", None):
            lib().orc_run_free(self._h)
            self._h = None

    def fmm(self):
        lib().orc_run_fmm(self._h)
        return self

    def trees(self):
        lib().orc_run_trees(self._h)
        return self

    def lists(self):
        lib().orc_run_lists(self._h)
        return self

    def branches(self, which):
        L = lib()
        nb = L.orc_n_branches(self._h, which)
        ints = np.zeros(nb * 6, dtype=np.int64)
        reals = np.zeros(nb * 17)
        L.orc_get_branches(self._h, which, ints, reals)
        return ints.reshape(nb, 6), reals.reshape(nb, 17)

    def levels(self, which):
        L = lib()
        n = L.orc_n_levels(self._h, which)
        out = np.zeros(2 * n, dtype=np.int64)
        L.orc_get_levels(self._h, which, out)
        return out.reshape(n, 2)

    def leaf_index(self, which):
        L = lib()
        out = np.zeros(L.orc_n_leaves(self._h, which), dtype=np.int64)
        L.orc_get_leaf_index(self._h, which, out)
        return out

    def sort_index(self, which):
        out = np.zeros(self.nt if which == 0 else self.ns, dtype=np.int64)
        lib().orc_get_sort_index(self._h, which, out)
        return out

    def buffer(self, which):
        rows = 16 if which == 0 else data_per_body(self.kind)
        n = self.nt if which == 0 else self.ns
        out = np.zeros(rows * n)
        lib().orc_get_buffer(self._h, which, out)
        return out.reshape(n, rows)

    def expansions(self, which):
        nb = lib().orc_n_branches(self._h, which)
        out = np.zeros(nb * 4 * nh(self.p))
        lib().orc_get_expansions(self._h, which, out)
        return out.reshape(nb, nh(self.p), 2, 2)   # [branch, i, phi/chi, re/im]

    def list(self, which):
        L = lib()
        n = L.orc_n_list(self._h, which)
        out = np.zeros(2 * n, dtype=np.int32)
        L.orc_get_list(self._h, which, out)
        return out.reshape(n, 2)

    def m2l_orders(self):
        n = lib().orc_n_list(self._h, 0)
        out = np.zeros(n, dtype=np.int32)
        lib().orc_get_m2l_orders(self._h, out)
        return out

    def m2l_margins(self):
        """Per M2L interaction: how close its order decision came to the tolerance, the smallest
        |estimate/limit - 1| over the comparisons made (inf at fixed order)."""
        n = lib().orc_n_list(self._h, 0)
        out = np.zeros(n)
        lib().orc_get_m2l_margins(self._h, out)
        return out

    def targets(self):
        out = np.zeros(self.nt * 16)
        lib().orc_get_targets(self._h, out)
        return out.reshape(self.nt, 16)

    @property
    def pmax(self):
        return lib().orc_pmax(self._h)

    @property
    def error_success(self):
        return bool(lib().orc_error_success(self._h))


def direct(kind, target_xyz, source_data, scalar_potential=False, gradient=True, hessian=False):
    """direct!(targets, sources) of src/direct.jl:43-75; returns the (N,16) target buffer."""
    t = np.zeros((len(target_xyz), 16))
    t[:, :3] = target_xyz
    tb = np.ascontiguousarray(t.ravel())
    s = np.ascontiguousarray(np.asarray(source_data, dtype=np.float64).ravel())
    lib().orc_direct(kind, len(t), tb, len(source_data), s, int(scalar_potential), int(gradient), int(hessian))
    return tb.reshape(-1, 16)


def direct_sampled(kind, target_xyz, idx, source_data, scalar_potential=False, gradient=True, hessian=False):
    idx = np.ascontiguousarray(idx, dtype=np.int64)
    out = np.zeros(len(idx) * 16)
    lib().orc_direct_sampled(kind, len(idx), idx, np.ascontiguousarray(np.asarray(target_xyz, dtype=np.float64).ravel()),
                             out, len(source_data), np.ascontiguousarray(np.asarray(source_data, np.float64).ravel()),
                             int(scalar_potential), int(gradient), int(hessian))
    return out.reshape(-1, 16)


# ---------------------------------------------------------------- operators
def _exp(p):
    return np.zeros(4 * nh(p))


def tables(p):
    Hs, z, e, M, Lt = np.zeros(length_Hs(p)), np.zeros(length_Ts(p)), np.zeros(length_Ts(p)), np.zeros(nh(p)), np.zeros(nh(p))
    lib().orc_tables(p, Hs, z, e, M, Lt)
    return Hs, z, e, M, Lt


def update_Ts(beta, p):
    out = np.zeros(length_Ts(p))
    lib().orc_update_Ts(beta, p, out)
    return out


def pnms(p, s, c):
    out = np.zeros(nh(p))
    lib().orc_pnms(p, s, c, out)
    return out


def eimphis(phi, P):
    out = np.zeros(2 * (P + 1))
    lib().orc_eimphis(phi, P, out)
    return out.reshape(P + 1, 2)


def zeta_sign(mag, mp, m):
    out = np.zeros(2)
    lib().orc_zeta_sign(mag, mp, m, out)
    return out


def eta_sign(mag, mp, m):
    out = np.zeros(2)
    lib().orc_eta_sign(mag, mp, m, out)
    return out


def regular_harmonics(rho, theta, phi, p):
    out = _exp(p)
    lib().orc_regular_harmonics(rho, theta, phi, p, out)
    return out.reshape(nh(p), 2, 2)[:, 0, 0] + 1j * out.reshape(nh(p), 2, 2)[:, 0, 1]


def irregular_harmonics(rho, theta, phi, p):
    out = _exp(p)
    lib().orc_irregular_harmonics(rho, theta, phi, p, out)
    return out.reshape(nh(p), 2, 2)[:, 0, 0] + 1j * out.reshape(nh(p), 2, 2)[:, 0, 1]


def cartesian_to_spherical(x, y, z):
    out = np.zeros(3)
    lib().orc_c2s(x, y, z, out)
    return tuple(out)


def factorial(k):
    return lib().orc_factorial(k)


def _op(name, e, arg, p, LH):
    out = _exp(p)
    getattr(lib(), name)(np.ascontiguousarray(e, dtype=np.float64), arg, p, int(LH), out)
    return out


def rotate_multipole_y(e, theta, p, LH=False):
    return _op("orc_rotate_multipole_y", e, theta, p, LH)


def rotate_local_y(e, theta, p, LH=False):
    return _op("orc_rotate_local_y", e, theta, p, LH)


def rotate_z(e, phi, p, LH=False):
    return _op("orc_rotate_z", e, phi, p, LH)


def translate_multipole_z(e, t, p, LH=False):
    return _op("orc_translate_multipole_z", e, t, p, LH)


def translate_m2l_z(e, t, p, LH=False):
    return _op("orc_translate_m2l_z", e, t, p, LH)


def translate_local_z(e, t, p, LH=False):
    return _op("orc_translate_local_z", e, t, p, LH)


def transform_lh_multipole(e, r, p):
    out = np.array(e, dtype=np.float64)
    lib().orc_transform_lh_multipole(out, r, p)
    return out


def transform_lh_local(e, r, p):
    out = np.array(e, dtype=np.float64)
    lib().orc_transform_lh_local(out, r, p)
    return out


def m2m(target_center, source_center, e, p, LH=False, out=None):
    out = _exp(p) if out is None else out
    lib().orc_m2m(np.asarray(target_center, float), np.asarray(source_center, float), np.ascontiguousarray(e), p, int(LH), out)
    return out


def l2l(target_center, source_center, e, p, LH=False, out=None):
    out = _exp(p) if out is None else out
    lib().orc_l2l(np.asarray(target_center, float), np.asarray(source_center, float), np.ascontiguousarray(e), p, int(LH), out)
    return out


def geom(center, radius=0.0, box=(0.0, 0.0, 0.0), max_influence=0.0):
    g = np.zeros(15)
    g[0:3] = center
    g[3:6] = center
    g[6] = radius
    g[7] = radius
    g[8:11] = box
    g[11:14] = box
    g[14] = max_influence
    return g


def m2l(tgeom, sgeom, e, p, LH=False, err_kind=ERR_NONE, relative=False, eps=0.0, BE=True, out=None):
    out = _exp(p) if out is None else out
    ok = C.c_int(0)
    P = lib().orc_m2l(np.ascontiguousarray(tgeom, float), np.ascontiguousarray(sgeom, float), np.ascontiguousarray(e),
                      p, int(LH), err_kind, int(relative), eps, int(BE), out, C.byref(ok))
    return out, P, bool(ok.value)


def l2p(dx, local, p, LH=False, PS=True, GS=True, HS=True):
    out = np.zeros(13)
    lib().orc_l2p(np.asarray(dx, float), np.ascontiguousarray(local), p, int(LH), int(PS), int(GS), int(HS), out)
    return out[0], out[1:4], out[4:13]


def p2m_panel(kind, panel_row, center, p, out=None):
    """body_to_multipole! of one panel (row in the buffer layout above) about `center`."""
    out = _exp(p) if out is None else out
    lib().orc_p2m_panel(kind, np.ascontiguousarray(panel_row, dtype=np.float64),
                        np.ascontiguousarray(center, dtype=np.float64), p, out)
    return out


def p2m(kind, dx, strength, p, out=None):
    out = _exp(p) if out is None else out
    lib().orc_p2m(kind, np.asarray(dx, float), np.atleast_1d(np.asarray(strength, float)), p, out)
    return out


def exp_to_complex(e, p, index=0):
    """[re/im, phi/chi, H] flat -> complex coefficients of potential `index` (m >= 0 layout)."""
    a = np.asarray(e).reshape(nh(p), 2, 2)
    return a[:, index, 0] + 1j * a[:, index, 1]


def complex_to_exp(c_phi, p, c_chi=None):
    a = np.zeros((nh(p), 2, 2))
    a[:, 0, 0] = np.real(c_phi)
    a[:, 0, 1] = np.imag(c_phi)
    if c_chi is not None:
        a[:, 1, 0] = np.real(c_chi)
        a[:, 1, 1] = np.imag(c_chi)
    return a.ravel()


def full_to_compressed(v, p):
    """Reference test layout (n, m = -n..n) -> (n, m >= 0) as in test/runtests.jl:23-37."""
    out = []
    i = 0
    for n in range(p + 1):
        for m in range(-n, n + 1):
            if m >= 0:
                out.append(v[i])
            i += 1
    return np.array(out)


INFLUENCE_POTENTIAL, INFLUENCE_NORMAL_GRADIENT = 0, 1


class GaussSeidel:
    """FastGaussSeidel + solve! (src/solve.jl:538-946) for one system that is both
    target and source (point sources or source panels). influence: the scalar
    potential (test/solve_test.jl:11-13) or, for panels, the gradient along the
    panel normal. Matrices are returned column-major as in the reference."""

    def __init__(self, kind, source_data, expansion_order=4, multipole_acceptance=0.5, leaf_size=30,
                 shrink_recenter=True, influence=INFLUENCE_POTENTIAL):
        s = np.ascontiguousarray(source_data, dtype=np.float64)
        assert s.shape[1] == data_per_body(kind)
        self.kind, self.n = kind, len(s)
        oi = np.array([expansion_order, leaf_size, int(shrink_recenter), influence], dtype=np.intc)
        od = np.array([multipole_acceptance], dtype=np.float64)
        self._h = lib().orc_fgs_create(kind, self.n, np.ascontiguousarray(s.ravel()), oi, od)

    def __del__(self):
        if getattr(self, "_h", None):
            lib().orc_fgs_free(self._h)
            self._h = None

    def leaves(self):
        """(n_leaves, 4): [branch, n_bodies, nonself rows, first sorted body] per leaf, leaf_index order."""
        n = lib().orc_fgs_n_leaves(self._h)
        out = np.zeros(4 * n, dtype=np.int64)
        lib().orc_fgs_leaf_info(self._h, out)
        return out.reshape(n, 4)

    def self_matrix(self, k):
        nb = int(self.leaves()[k, 1])
        out = np.zeros(nb * nb)
        lib().orc_fgs_self(self._h, k, out)
        return out.reshape(nb, nb).T

    def nonself_matrix(self, k):
        info = self.leaves()[k]
        out = np.zeros(int(info[1] * info[2]))
        lib().orc_fgs_nonself(self._h, k, out)
        return out.reshape(int(info[1]), int(info[2])).T

    def nonself_rows(self, k):
        out = np.zeros(int(self.leaves()[k, 2]), dtype=np.int64)
        lib().orc_fgs_nonself_rows(self._h, k, out)
        return out

    def leaf_bodies(self, k):
        out = np.zeros(int(self.leaves()[k, 1]), dtype=np.int64)
        lib().orc_fgs_leaf_bodies(self._h, k, out)
        return out

    def solve(self, external, max_iterations=10, tolerance=1e-3):
        """Returns (strengths, targets (n,16), sweeps, mse history); original body order."""
        ext = np.ascontiguousarray(external, dtype=np.float64)
        strengths = np.zeros(self.n)
        targets = np.zeros(self.n * 16)
        hist = np.zeros(max(1, max_iterations))
        sweeps = lib().orc_fgs_solve(self._h, ext, max_iterations, tolerance, strengths, targets, hist)
        return strengths, targets.reshape(-1, 16), sweeps, hist[:max(1, max_iterations)]
