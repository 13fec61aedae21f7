"""Host-side mirror of FastMultipole's fmm!/direct! entry points over libfmmhip.

Mirrors src/fmm.jl:807-1176 (fmm! and its keyword arguments and 8-tuple
return), src/direct.jl:20-75 (direct!), the error-method and interaction-list
types of src/containers.jl:64-123, and the user-system interface of
src/compatibility.jl (source_system_to_buffer!, data_per_body, get_position,
strength_dims, get_n_bodies, body_to_multipole! kernel tag, direct!,
buffer_to_target_system!, has_vector_potential).

All evaluation runs on the GPU through libfmmhip; there is no CPU path.
"""
import atexit
import math
import time
import ctypes as C
import weakref
from dataclasses import dataclass

import numpy as np

from . import _lib as L


# ------------------------------------------------------------ error methods
@dataclass(frozen=True)
class ErrorMethod:
    eps: float
    BE: bool = True
    kind: int = L.ERR_NONE
    relative: bool = False


def PowerAbsolutePotential(eps, BE=True):            # src/containers.jl:85-86
    return ErrorMethod(float(eps), bool(BE), L.ERR_POWER_POTENTIAL, False)


def PowerAbsoluteGradient(eps, BE=True):             # :88-89
    return ErrorMethod(float(eps), bool(BE), L.ERR_POWER_GRADIENT, False)


def RotatedCoefficientsAbsoluteGradient(eps, BE=True):   # :91-92
    return ErrorMethod(float(eps), bool(BE), L.ERR_ROTATED_COEFFICIENTS_GRADIENT, False)


def PowerRelativePotential(eps, BE=True):            # :97-98
    return ErrorMethod(float(eps), bool(BE), L.ERR_POWER_POTENTIAL, True)


def PowerRelativeGradient(eps, BE=True):             # :100-101
    return ErrorMethod(float(eps), bool(BE), L.ERR_POWER_GRADIENT, True)


def RotatedCoefficientsRelativeGradient(eps, BE=True):   # :103-104
    return ErrorMethod(float(eps), bool(BE), L.ERR_ROTATED_COEFFICIENTS_GRADIENT, True)


# ------------------------------------------------- interaction list methods
@dataclass(frozen=True)
class InteractionListMethod:
    code: int


def Barba():                 # src/containers.jl:110-112
    return InteractionListMethod(L.BARBA)


def SelfTuning():            # :118-120
    return InteractionListMethod(L.SELFTUNING)


def SelfTuningTreeStop():    # :122-123
    return InteractionListMethod(L.SELFTUNING_TREESTOP)


@dataclass(frozen=True)
class DerivativesSwitch:     # src/containers.jl:50
    scalar_potential: bool = True
    gradient: bool = True
    hessian: bool = True


# ----------------------------------------------------------------- context
class Context:
    """One HIP device + stream (fmmhip_ctx). Not thread-safe."""

    def __init__(self, device=-1):
        lib = L.load()
        h = C.c_void_p()
        st = lib.fmmhip_ctx_create(int(device), C.byref(h))
        if st != L.FMMHIP_OK:
            raise L.FmmHipError(st, "fmmhip_ctx_create failed (no usable GPU?)")
        self._h = h
        self.lib = lib
        _LIVE_CTXS.add(self)

    def check(self, st):
        if st != L.FMMHIP_OK:
            raise L.FmmHipError(st, self.lib.fmmhip_last_error(self._h).decode())

    def set_p2p_free_slots(self, slots):
        """Workgroup slots per compute unit the nearfield leaves to the far-field stream (0..3)."""
        self.check(self.lib.fmmhip_ctx_set_p2p_free_slots(self._h, int(slots)))

    def close(self):
        if getattr(self, "_h", None):
            self.lib.fmmhip_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_DEFAULT = None
_LIVE_EVALS = weakref.WeakSet()
_LIVE_CTXS = weakref.WeakSet()
_LIVE_SERVERS = weakref.WeakSet()


@atexit.register
def _release_all():
    """Release device state before interpreter teardown: servers, evaluations, then contexts."""
    for sv in list(_LIVE_SERVERS):
        sv.close()
    for ev in list(_LIVE_EVALS):
        ev.close()
    for ctx in list(_LIVE_CTXS):
        ctx.close()


def default_context():
    global _DEFAULT
    if _DEFAULT is None:
        _DEFAULT = Context()
    return _DEFAULT


def make_options(expansion_order=5, leaf_size_source=20, leaf_size_target=None, multipole_acceptance=0.4,
                 interaction_list_method=None, shrink_recenter=True, error_tolerance=None, scalar_potential=False,
                 gradient=True, hessian=False, farfield=True, nearfield=True, self_induced=True, precision="fp64"):
    o = L.default_options()
    o.expansion_order = int(expansion_order)
    o.leaf_size_source = int(leaf_size_source)
    o.leaf_size_target = int(leaf_size_source if leaf_size_target is None else leaf_size_target)
    o.multipole_acceptance = float(multipole_acceptance)
    o.list_method = (interaction_list_method or SelfTuningTreeStop()).code
    o.shrink_recenter = int(bool(shrink_recenter))
    if error_tolerance is not None:
        o.error_kind = error_tolerance.kind
        o.error_relative = int(error_tolerance.relative)
        o.error_tolerance = error_tolerance.eps
        o.error_bonus = int(error_tolerance.BE)
    o.scalar_potential, o.gradient, o.hessian = int(scalar_potential), int(gradient), int(hessian)
    o.farfield, o.nearfield, o.self_induced = int(farfield), int(nearfield), int(self_induced)
    o.precision = {"fp64": L.FP64, "mixed": L.MIXED}[precision] if isinstance(precision, str) else int(precision)
    return o


class TreeView:
    """Host view of a device tree (Tree, src/containers.jl:182-195); indices are 0-based."""

    def __init__(self, ev, which):
        lib, h = ev.ctx.lib, ev._h
        nb = lib.fmmhip_eval_n_branches(h, which)
        self.n_branches = nb
        ints = np.zeros((max(nb, 1), 6), dtype=np.int64)
        reals = np.zeros((max(nb, 1), 17))
        if nb:
            ev.ctx.check(lib.fmmhip_eval_get_branches(h, which, ints.ctypes.data, reals.ctypes.data))
        self.ints, self.reals = ints[:nb], reals[:nb]
        nl = lib.fmmhip_eval_n_levels(h, which)
        lv = np.zeros((max(nl, 1), 2), dtype=np.int64)
        if nl:
            ev.ctx.check(lib.fmmhip_eval_get_levels(h, which, lv.ctypes.data))
        self.levels_index = lv[:nl]
        nleaf = lib.fmmhip_eval_n_leaves(h, which)
        li = np.zeros(max(nleaf, 1), dtype=np.int64)
        if nleaf:
            ev.ctx.check(lib.fmmhip_eval_get_leaf_index(h, which, li.ctypes.data))
        self.leaf_index = li[:nleaf]
        n = ev.nt if which == 0 else ev.ns
        si = np.zeros(max(n, 1), dtype=np.int64)
        if n and nb:
            ev.ctx.check(lib.fmmhip_eval_get_sort_index(h, which, si.ctypes.data))
        self.sort_index = si[:n]
        self.inverse_sort_index = np.empty_like(self.sort_index)
        self.inverse_sort_index[self.sort_index] = np.arange(n)
        self._ev, self._which = ev, which

    def expansions(self, p):
        H = (p + 1) * (p + 2) // 2
        out = np.zeros(self.n_branches * 4 * H)
        self._ev.ctx.check(self._ev.ctx.lib.fmmhip_eval_get_expansions(self._ev._h, self._which, out.ctypes.data))
        return out.reshape(self.n_branches, H, 2, 2)


class Evaluation:
    """Device-resident trees, lists and expansions of one fmm! call (the reference's Cache + trees)."""

    def __init__(self, ctx=None):
        self.ctx = ctx or default_context()
        h = C.c_void_p()
        self.ctx.check(self.ctx.lib.fmmhip_eval_create(self.ctx._h, C.byref(h)))
        self._h = h
        self.nt = self.ns = 0
        self._keep = None
        _LIVE_EVALS.add(self)

    def close(self):
        if getattr(self, "_h", None):
            self.ctx.lib.fmmhip_eval_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_systems(self, kernel, target_xyz, source_buffer, on_device=False):
        """target_xyz: (nt,3) float64 or None (targets = sources); source_buffer: (ns, data_per_body).

        With on_device=True both are device pointers (ints) or objects with data_ptr() (torch tensors)."""
        if on_device:
            tp = 0 if target_xyz is None else (target_xyz.data_ptr() if hasattr(target_xyz, "data_ptr") else int(target_xyz))
            sp = source_buffer.data_ptr() if hasattr(source_buffer, "data_ptr") else int(source_buffer)
            ns = source_buffer.shape[0]
            nt = ns if target_xyz is None else target_xyz.shape[0]
            self._keep = (target_xyz, source_buffer)
        else:
            sb = np.ascontiguousarray(source_buffer, dtype=np.float64)
            ns = sb.shape[0]
            tx = None if target_xyz is None else np.ascontiguousarray(target_xyz, dtype=np.float64)
            nt = ns if tx is None else tx.shape[0]
            self._keep = (tx, sb)
            tp = 0 if tx is None else tx.ctypes.data
            sp = sb.ctypes.data
        self.kernel, self.nt, self.ns = kernel, nt, ns
        self.counts = None
        self.ctx.check(self.ctx.lib.fmmhip_eval_set_systems(self._h, kernel, nt, tp or None, ns, sp or None, int(on_device)))

    def set_batch(self, kernel, source_buffers, on_device=False):
        """Several independent systems, each its own target and source (fmmhip_eval_set_batch).

        source_buffers: a list of (n_i, data_per_body) arrays, or with on_device=True one device buffer of the
        concatenated columns (an object with data_ptr() and shape) followed by the list of counts."""
        if on_device:
            buf, counts = source_buffers
            counts = np.ascontiguousarray(counts, dtype=np.int64)
            sp = buf.data_ptr() if hasattr(buf, "data_ptr") else int(buf)
            self._keep = (buf, counts)
        else:
            counts = np.array([b.shape[0] for b in source_buffers], dtype=np.int64)
            sb = np.ascontiguousarray(np.concatenate([np.asarray(b, dtype=np.float64) for b in source_buffers]))
            self._keep = (sb, counts)
            sp = sb.ctypes.data
        n = int(counts.sum())
        self.kernel, self.nt, self.ns = kernel, n, n
        self.counts = counts
        self.ctx.check(self.ctx.lib.fmmhip_eval_set_batch(self._h, kernel, len(counts), counts.ctypes.data, sp,
                                                          int(on_device)))

    def set_tuple(self, kernels, source_buffers, on_device=False):
        """Several systems in one tree, each a target and a source of all of them (fmmhip_eval_set_tuple).

        kernels: one point or panel kernel per system; source_buffers: one (n_i, data_per_body) array per
        system, or with on_device=True one device buffer of R x sum(n_i) rows followed by the list of counts,
        R the largest data_per_body of the kernels (each system's rows first, zeros after them)."""
        kernels = np.ascontiguousarray(kernels, dtype=np.int32)
        R = max([7] + [L.data_per_body(int(k)) for k in kernels])
        if on_device:
            buf, counts = source_buffers
            counts = np.ascontiguousarray(counts, dtype=np.int64)
            sp = buf.data_ptr() if hasattr(buf, "data_ptr") else int(buf)
            self._keep = (buf, counts, kernels)
        else:
            counts = np.array([b.shape[0] for b in source_buffers], dtype=np.int64)
            rows = []
            for k, b in zip(kernels, source_buffers):
                b = np.asarray(b, dtype=np.float64)
                if b.ndim != 2 or b.shape[1] != L.data_per_body(int(k)):
                    raise ValueError(f"a system of kernel {int(k)} takes {L.data_per_body(int(k))} rows per body")
                r = np.zeros((b.shape[0], R))
                r[:, :b.shape[1]] = b
                rows.append(r)
            sb = np.ascontiguousarray(np.concatenate(rows) if rows else np.zeros((0, R)))
            self._keep = (sb, counts, kernels)
            sp = sb.ctypes.data
        n = int(counts.sum())
        self.kernel, self.nt, self.ns = None, n, n
        self.counts = None
        self.tuple_counts = counts
        self.ctx.check(self.ctx.lib.fmmhip_eval_set_tuple(self._h, len(counts), kernels.ctypes.data, counts.ctypes.data,
                                                          sp, int(on_device)))

    def set_tuple_leaf_sizes(self, leaf_sizes):
        """One leaf size per system of the tuple set last (the reference's SVector leaf_size)."""
        ls = np.ascontiguousarray(leaf_sizes, dtype=np.int64)
        self.ctx.check(self.ctx.lib.fmmhip_eval_set_tuple_leaf_sizes(self._h, len(ls), ls.ctypes.data))

    def batch_stats(self):
        """Per system of the last batch evaluation (one entry for a single system): (pmax, error_success)
        arrays, each system's values as its own fmm! call reports them."""
        k = len(getattr(self, "counts", [1])) if getattr(self, "counts", None) is not None else 1
        pm = np.zeros(k, dtype=np.int32)
        ok = np.zeros(k, dtype=np.int32)
        self.ctx.check(self.ctx.lib.fmmhip_eval_get_batch_stats(self._h, pm.ctypes.data, ok.ctypes.data))
        return pm, ok.astype(bool)

    def build_trees(self, opt):
        self.ctx.check(self.ctx.lib.fmmhip_eval_build_trees(self._h, C.byref(opt)))

    def build_lists(self, opt):
        self.ctx.check(self.ctx.lib.fmmhip_eval_build_lists(self._h, C.byref(opt)))

    def run(self, opt):
        st = L.Stats()
        self.ctx.check(self.ctx.lib.fmmhip_eval_run(self._h, C.byref(opt), C.byref(st)))
        return st

    def fmm(self, opt):
        st = L.Stats()
        self.ctx.check(self.ctx.lib.fmmhip_fmm(self._h, C.byref(opt), C.byref(st)))
        return st

    def targets(self, out=None):
        """16 x nt target buffer rows in original order, as an (nt, 16) array (or into a device pointer)."""
        if out is not None:
            ptr = out.data_ptr() if hasattr(out, "data_ptr") else int(out)
            self.ctx.check(self.ctx.lib.fmmhip_eval_get_targets(self._h, ptr, 1))
            return out
        a = np.zeros((self.nt, 16))
        if self.nt:
            self.ctx.check(self.ctx.lib.fmmhip_eval_get_targets(self._h, a.ctypes.data, 0))
        return a

    def tree(self, which):
        return TreeView(self, which)

    def list(self, which):
        n = self.ctx.lib.fmmhip_eval_n_list(self._h, which)
        out = np.zeros((max(n, 1), 2), dtype=np.int32)
        if n:
            self.ctx.check(self.ctx.lib.fmmhip_eval_get_list(self._h, which, out.ctypes.data))
        return out[:n]

    def m2l_orders(self):
        n = self.ctx.lib.fmmhip_eval_n_list(self._h, 0)
        out = np.zeros(max(n, 1), dtype=np.int32)
        if n:
            self.ctx.check(self.ctx.lib.fmmhip_eval_get_m2l_orders(self._h, out.ctypes.data))
        return out[:n]


# ------------------------------------------------------------ FastGaussSeidel
class FastGaussSeidel:
    """FastGaussSeidel(system; expansion_order=4, multipole_acceptance=0.5, leaf_size=30,
    shrink_recenter=true) of src/solve.jl:538-645 for one system that is both target and source
    (point sources or source panels). `influence` selects the value matched per body: the scalar
    potential (the reference test's influence!) or, for panels, the normal component of the gradient.
    Influence matrices and solver state stay in HBM."""

    def __init__(self, system, expansion_order=4, multipole_acceptance=0.5, leaf_size=30, shrink_recenter=True,
                 influence=L.INFLUENCE_POTENTIAL, ctx=None):
        kernel = _kernel_of(system)
        self.system = system
        self.ev = Evaluation(ctx)
        self.ev.set_systems(kernel, None, system.source_buffer())
        self.opt = make_options(expansion_order=expansion_order, leaf_size_source=leaf_size,
                                multipole_acceptance=multipole_acceptance, shrink_recenter=shrink_recenter,
                                interaction_list_method=Barba(), scalar_potential=True)
        self.ev.ctx.check(self.ev.ctx.lib.fmmhip_fgs_setup(self.ev._h, C.byref(self.opt), int(influence)))

    def leaves(self):
        n = self.ev.ctx.lib.fmmhip_fgs_n_leaves(self.ev._h)
        out = np.zeros((max(n, 1), 4), dtype=np.int64)
        if n:
            self.ev.ctx.check(self.ev.ctx.lib.fmmhip_fgs_leaf_info(self.ev._h, out.ctypes.data))
        return out[:n]

    def matrix(self, which, leaf):
        """which 0: self matrix (n x n), 1: non-self matrix (rows x n), as numpy arrays."""
        info = self.leaves()[leaf]
        rows = int(info[1] if which == 0 else info[2])
        buf = np.zeros(max(1, rows * int(info[1])))
        self.ev.ctx.check(self.ev.ctx.lib.fmmhip_fgs_get_matrix(self.ev._h, which, leaf, buf.ctypes.data))
        return buf[:rows * int(info[1])].reshape(int(info[1]), rows).T

    def nonself_rows(self, leaf):
        rows = int(self.leaves()[leaf, 2])
        out = np.zeros(max(1, rows), dtype=np.int64)
        self.ev.ctx.check(self.ev.ctx.lib.fmmhip_fgs_nonself_rows(self.ev._h, leaf, out.ctypes.data))
        return out[:rows]

    def solve(self, external, max_iterations=10, tolerance=1e-3):
        """solve! (src/solve.jl:797-946). external: per-body external influence (original order).
        Returns (strengths, targets (n,16), sweeps, mse history) and writes the strengths and the
        potential/gradient back into the system."""
        ext = np.ascontiguousarray(external, dtype=np.float64)
        n = self.ev.ns
        strengths = np.zeros(n)
        hist = np.zeros(max(1, max_iterations))
        sweeps = C.c_int32(0)
        self.ev.ctx.check(self.ev.ctx.lib.fmmhip_fgs_solve(self.ev._h, ext.ctypes.data, 0, int(max_iterations),
                                                           float(tolerance), strengths.ctypes.data, C.byref(sweeps),
                                                           hist.ctypes.data))
        targets = self.ev.targets()
        if hasattr(self.system, "set_strengths"):
            self.system.set_strengths(strengths)
        if hasattr(self.system, "buffer_to_target_system"):
            self.system.buffer_to_target_system(targets, DerivativesSwitch(True, True, False))
        return strengths, targets, int(sweeps.value), hist[:max(1, max_iterations)]


def solve(system, solver, max_iterations=10, tolerance=1e-3, external=None):
    """solve!(system, solver; max_iterations, tolerance): the external influence defaults to the
    system's current potential (target_influence_to_buffer! of test/solve_test.jl:15-17)."""
    if external is None:
        pot = np.asarray(system.potential, dtype=np.float64)
        external = pot[:, 0] if pot.ndim == 2 else pot
    return solver.solve(external, max_iterations=max_iterations, tolerance=tolerance)


# ------------------------------------------------------------------ request server
class Server:
    """Concurrent request loop on one device (fmmhip_server): fmm!(system) evaluations and
    FastGaussSeidel solves of uneven sizes, each the same computation as the single call.

    n_slots workers each own a stream and device-resident buffers reused between requests; a
    request is admitted when its estimated HBM footprint fits within hbm_budget_bytes beside the
    running ones. FMM requests with the same kernel and options that wait together run as one batch
    evaluation (up to batch_requests requests and batch_bodies bodies; batch_requests=1: each alone).
    submit_* copy the inputs and return a ticket at once; wait(ticket) blocks."""

    def __init__(self, n_slots=4, hbm_budget_bytes=64 * 2**30, device=-1, batch_requests=64, batch_bodies=2 << 20):
        self.lib = L.load()
        h = C.c_void_p()
        st = self.lib.fmmhip_server_create(int(device), int(n_slots), float(hbm_budget_bytes), C.byref(h))
        if st != L.FMMHIP_OK:
            raise L.FmmHipError(st, "fmmhip_server_create failed (bad arguments or no usable GPU)")
        self._h = h
        self._sizes = {}
        _LIVE_SERVERS.add(self)
        st = self.lib.fmmhip_server_set_batching(h, int(batch_requests), int(batch_bodies))
        if st != L.FMMHIP_OK:
            self.close()
            raise L.FmmHipError(st, "fmmhip_server_set_batching rejected its arguments")

    def _submit(self, req, n, kind):
        t = C.c_int64()
        st = self.lib.fmmhip_server_submit(self._h, C.byref(req), C.byref(t))
        if st != L.FMMHIP_OK:
            raise L.FmmHipError(st, "fmmhip_server_submit rejected the request")
        self._sizes[t.value] = (n, kind)
        return t.value

    def submit_fmm(self, system, *, leaf_size=20, expansion_order=5, multipole_acceptance=0.4, error_tolerance=None,
                   shrink_recenter=True, interaction_list_method=None, scalar_potential=False, gradient=True,
                   hessian=False, farfield=True, nearfield=True, self_induced=True, precision="fp64"):
        """Queue fmm!(system; kwargs...) (src/fmm.jl:807-884) for one system as target and source."""
        sbuf = np.ascontiguousarray(system.source_buffer(), dtype=np.float64)
        req = L.Request()
        req.kind, req.kernel, req.n = L.REQUEST_FMM, _kernel_of(system), sbuf.shape[0]
        req.source_buffer = sbuf.ctypes.data
        req.options = make_options(expansion_order, leaf_size, leaf_size, multipole_acceptance,
                                   interaction_list_method, shrink_recenter, error_tolerance, scalar_potential,
                                   gradient, hessian, farfield, nearfield, self_induced, precision)
        return self._submit(req, sbuf.shape[0], L.REQUEST_FMM)

    def submit_solve(self, system, external, *, expansion_order=4, multipole_acceptance=0.5, leaf_size=30,
                     shrink_recenter=True, influence=L.INFLUENCE_POTENTIAL, max_iterations=10, tolerance=1e-3):
        """Queue FastGaussSeidel(system; ...) followed by solve!(system, solver; max_iterations, tolerance)
        (src/solve.jl:538-645, 797-946)."""
        sbuf = np.ascontiguousarray(system.source_buffer(), dtype=np.float64)
        ext = np.ascontiguousarray(external, dtype=np.float64)
        if ext.shape != (sbuf.shape[0],):
            raise ValueError("external must hold one value per body")
        req = L.Request()
        req.kind, req.kernel, req.n = L.REQUEST_SOLVE, _kernel_of(system), sbuf.shape[0]
        req.source_buffer = sbuf.ctypes.data
        req.options = make_options(expansion_order=expansion_order, leaf_size_source=leaf_size,
                                   multipole_acceptance=multipole_acceptance, shrink_recenter=shrink_recenter,
                                   interaction_list_method=Barba(), scalar_potential=True)
        req.influence, req.max_iterations, req.tolerance = int(influence), int(max_iterations), float(tolerance)
        req.external = ext.ctypes.data
        return self._submit(req, sbuf.shape[0], L.REQUEST_SOLVE)

    def wait(self, ticket):
        """Block until the request ends. Returns (targets (n,16) in body order, strengths or None,
        stats dict). A request whose tolerance was missed still returns results, with
        stats["status"] == FMMHIP_ETOL."""
        n, kind = self._sizes.pop(ticket)
        out = np.zeros((max(n, 1), 16))
        strengths = np.zeros(max(n, 1)) if kind == L.REQUEST_SOLVE else None
        rs = L.RequestStats()
        st = self.lib.fmmhip_server_wait(self._h, int(ticket), out.ctypes.data,
                                         None if strengths is None else strengths.ctypes.data, C.byref(rs))
        if st not in (L.FMMHIP_OK, L.FMMHIP_ETOL):
            raise L.FmmHipError(st, f"request {ticket} failed")
        return out[:n], (None if strengths is None else strengths[:n]), rs.as_dict()

    def stats(self):
        s = L.ServerStats()
        st = self.lib.fmmhip_server_stats(self._h, C.byref(s))
        if st != L.FMMHIP_OK:
            raise L.FmmHipError(st, "fmmhip_server_stats failed")
        return s.as_dict()

    def close(self):
        if getattr(self, "_h", None):
            self.lib.fmmhip_server_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ------------------------------------------------------------------ fmm!
def _as_tuple(x):
    return x if isinstance(x, tuple) else (x,)


def _kernel_of(system):
    tag = getattr(system, "fmm_kernel", None)
    if tag is None:
        raise TypeError(f"{type(system).__name__} does not declare a built-in kernel (fmm_kernel); "
                        "only Point{Source} and Point{Vortex} systems run on libfmmhip")
    return tag


def fmm(target_systems, source_systems=None, *, leaf_size=None, leaf_size_source=None, leaf_size_target=None,
        expansion_order=5, multipole_acceptance=None, error_tolerance=None, shrink_recenter=True,
        interaction_list_method=None, scalar_potential=False, gradient=True, hessian=False, farfield=True,
        nearfield=True, self_induced=True, update_target_systems=True, precision="fp64", evaluation=None,
        tune=False):
    """fmm!(systems) / fmm!(targets, sources) of src/fmm.jl:807-915.

    Returns the reference's 8-tuple (optimized_args, cache, target_tree, source_tree,
    m2l_list, direct_list, derivatives_switches, error_success); `cache` is the
    device-resident Evaluation, the trees are host views with 0-based indices.
    """
    single = source_systems is None
    if single:
        source_systems = target_systems
        ls = 20 if leaf_size is None else leaf_size        # fmm!(system; leaf_size=20) (src/fmm.jl:807)
        leaf_size_source = ls if leaf_size_source is None else leaf_size_source
        leaf_size_target = ls if leaf_size_target is None else leaf_size_target
    if leaf_size_source is None:
        leaf_size_source = 20
    if multipole_acceptance is None:
        multipole_acceptance = 0.4
    targets, sources = _as_tuple(target_systems), _as_tuple(source_systems)
    if len(targets) != 1 or len(sources) != 1:
        if len(targets) != len(sources) or any(t is not s for t, s in zip(targets, sources)):
            raise NotImplementedError("several systems run as one tuple that is both the targets and the sources")
        return _fmm_tuple(sources, leaf_size_source, leaf_size_target, expansion_order, multipole_acceptance,
                          error_tolerance, shrink_recenter, interaction_list_method, scalar_potential, gradient, hessian,
                          farfield, nearfield, self_induced, update_target_systems, precision, evaluation)
    tsys, ssys = targets[0], sources[0]
    kernel = _kernel_of(ssys)
    ev = evaluation or Evaluation()
    sbuf = ssys.source_buffer()
    txyz = None if (single or tsys is ssys) else tsys.positions()
    ev.set_systems(kernel, txyz, sbuf)
    opt = make_options(expansion_order, leaf_size_source, leaf_size_target, multipole_acceptance,
                       interaction_list_method, shrink_recenter, error_tolerance, scalar_potential, gradient, hessian,
                       farfield, nearfield, self_induced, precision)
    stats = ev.fmm(opt)
    out = ev.targets()
    switch = DerivativesSwitch(bool(scalar_potential), bool(gradient), bool(hessian))
    if update_target_systems:
        tsys.buffer_to_target_system(out, switch)
    leaf, order = opt.leaf_size_source, opt.expansion_order
    if tune:
        leaf, order = _tuned_leaf_size(stats, leaf), stats.pmax
    optimized_args = {"leaf_size_source": leaf, "expansion_order": max(order, 1),
                      "multipole_acceptance": opt.multipole_acceptance}
    ev.last_stats = stats
    return (optimized_args, ev, ev.tree(0), ev.tree(1), ev.list(0), ev.list(1), (switch,), bool(stats.error_success))


def _fmm_tuple(systems, leaf_size_source, leaf_size_target, expansion_order, multipole_acceptance, error_tolerance,
               shrink_recenter, interaction_list_method, scalar_potential, gradient, hessian, farfield, nearfield,
               self_induced, update_target_systems, precision, evaluation):
    """fmm!((sys1, sys2, ...)) (src/fmm.jl:807-884): one tree over the bodies of every system (src/tree.jl:3-122),
    each system a target of all and a source with its own kernel (point or panel kernels). leaf_size: one
    for every system, or one per system (the reference's SVector); the same for targets and sources.
    SelfTuningTreeStop (the default) or Barba lists."""
    per_system = np.ndim(leaf_size_source) > 0
    if (np.ndim(leaf_size_target) > 0) != per_system or not np.array_equal(leaf_size_target, leaf_size_source):
        raise ValueError("a tuple of systems takes the same leaf sizes for targets and sources")
    if per_system and len(leaf_size_source) != len(systems):
        raise ValueError("one leaf size per system")
    if interaction_list_method is None:
        interaction_list_method = SelfTuningTreeStop()
    ev = evaluation or Evaluation()
    ev.set_tuple([_kernel_of(s) for s in systems], [s.source_buffer() for s in systems])
    if per_system:
        ev.set_tuple_leaf_sizes(leaf_size_source)
        leaf_size_source = leaf_size_target = int(min(leaf_size_source))
    opt = make_options(expansion_order, leaf_size_source, leaf_size_target, multipole_acceptance,
                       interaction_list_method, shrink_recenter, error_tolerance, scalar_potential, gradient, hessian,
                       farfield, nearfield, self_induced, precision)
    stats = ev.fmm(opt)
    out = ev.targets()
    switch = DerivativesSwitch(bool(scalar_potential), bool(gradient), bool(hessian))
    offs = np.concatenate([[0], np.cumsum(ev.tuple_counts)])
    outputs = [out[offs[i]:offs[i + 1]] for i in range(len(systems))]
    if update_target_systems:
        for s, o in zip(systems, outputs):
            s.buffer_to_target_system(o, switch)
    ev.last_stats = stats
    ev.tuple_outputs = outputs
    optimized_args = {"leaf_size_source": opt.leaf_size_source, "expansion_order": max(opt.expansion_order, 1),
                      "multipole_acceptance": opt.multipole_acceptance}
    return (optimized_args, ev, ev.tree(0), ev.tree(1), ev.list(0), ev.list(1), (switch,) * len(systems),
            bool(stats.error_success))


def fmm_batch(systems, *, leaf_size=20, expansion_order=5, multipole_acceptance=0.4, error_tolerance=None,
              shrink_recenter=True, interaction_list_method=None, scalar_potential=False, gradient=True, hessian=False,
              update_target_systems=True, precision="fp64", evaluation=None):
    """fmm!(system) (src/fmm.jl:807-915) for each of several independent systems of one kernel, in one
    evaluation: one tree per system, lists within each system, every pass of all systems in the same
    launches. Each system's results are those of its own fmm! call. Returns (outputs, evaluation): outputs[i]
    is system i's (n_i, 16) target buffer; evaluation.last_stats covers the whole batch (error_success is
    false if any system missed the tolerance), evaluation.batch_stats() gives each system's pmax and
    error_success as its own fmm! call reports them."""
    systems = list(systems)
    if not systems:
        return [], None
    kernel = _kernel_of(systems[0])
    if any(_kernel_of(s) != kernel for s in systems):
        raise ValueError("the systems of one batch share one kernel")
    ev = evaluation or Evaluation()
    ev.set_batch(kernel, [s.source_buffer() for s in systems])
    opt = make_options(expansion_order, leaf_size, leaf_size, multipole_acceptance, interaction_list_method,
                       shrink_recenter, error_tolerance, scalar_potential, gradient, hessian, True, True, True, precision)
    ev.last_stats = ev.fmm(opt)
    out = ev.targets()
    switch = DerivativesSwitch(bool(scalar_potential), bool(gradient), bool(hessian))
    offs = np.concatenate([[0], np.cumsum(ev.counts)])
    outputs = [out[offs[i]:offs[i + 1]] for i in range(len(systems))]
    if update_target_systems:
        for s, o in zip(systems, outputs):
            s.buffer_to_target_system(o, switch)
    return outputs, ev


def _tuned_leaf_size(stats, leaf_size_source):
    """fmm!(...; tune=true)'s leaf-size update (src/fmm.jl:1073-1089): with per-M2L time t_m2l and
    per-body-pair nearfield time t_direct, the leaf size balancing both is ceil(sqrt(t_m2l / t_direct));
    with an empty M2L list the leaf size halves (at least 1). Times are the device phase times of the
    evaluation just run. When no body pair was evaluated the leaf size is kept (the reference's ratio
    is undefined there)."""
    if stats.n_m2l == 0:
        return max(leaf_size_source >> 1, 1)
    if stats.n_p2p_pairs == 0 or stats.ms_p2p <= 0.0:
        return leaf_size_source
    t_m2l = stats.ms_m2l / stats.n_m2l
    t_direct = stats.ms_p2p / stats.n_p2p_pairs
    return max(int(math.ceil(math.sqrt(t_m2l / t_direct))), 1)


def tune_fmm(target_systems, source_systems=None, *, error_tolerance=None, expansion_order=4, leaf_size_source=20,
             max_expansion_order=20, multipole_acceptances=None, verbose=True, **kwargs):
    """tune_fmm(target_systems, source_systems; ...) of src/autotune.jl:53-232.

    For each multipole acceptance (default 0.3, 0.4, ..., 0.8): one fmm!(tune=true) at the maximum
    expansion order gives a leaf size; a second at that leaf size with the maximum lowered to the
    order used + 2 (rerun at the full maximum if the tolerance is then missed) gives the expansion
    order; a final timed fmm! prices the combination. Tolerances missed at the maximum order skip the
    acceptance. Returns (tuned_params, cache): the fastest combination's leaf_size_source,
    expansion_order and multipole_acceptance, and the device-resident Evaluation for later calls."""
    if source_systems is None:
        source_systems = target_systems
    if multipole_acceptances is None:
        multipole_acceptances = [round(0.3 + 0.1 * k, 10) for k in range(6)]
    if verbose:
        print("\n#======= Begin FastMultipole.tune_fmm() =======#")
    kw = dict(kwargs)
    kw.pop("tune", None)
    kw.pop("update_target_systems", None)
    kw.pop("evaluation", None)
    single = source_systems is target_systems
    cache = Evaluation()

    def call(order, leaf, mac):
        t0 = time.perf_counter()
        res = fmm(target_systems, None if single else source_systems, leaf_size_source=leaf,
                  leaf_size_target=kw.get("leaf_size_target", leaf), expansion_order=order, multipole_acceptance=mac,
                  error_tolerance=error_tolerance, tune=True, update_target_systems=False, evaluation=cache,
                  **{k: v for k, v in kw.items() if k != "leaf_size_target"})
        return time.perf_counter() - t0, res[0], res[7]

    fixed = error_tolerance is None
    best = (math.inf, max_expansion_order, leaf_size_source, multipole_acceptances[0])
    call(1, leaf_size_source, multipole_acceptances[0])        # first call: allocations and code paths warm
    original_max = max_expansion_order
    for mac in multipole_acceptances:
        if verbose:
            print(f"\nmultipole_acceptance = {mac}...")
        t, optargs, ok = call(expansion_order if fixed else max_expansion_order, leaf_size_source, mac)
        if not ok:
            print(f"\terror tolerance not satisfied for max expansion order P={max_expansion_order};")
            print("\tskipping this multipole_acceptance...")
            continue
        leaf_size_source = optargs["leaf_size_source"]
        this_max = optargs["expansion_order"] + 2
        t, optargs, ok = call(expansion_order if fixed else this_max, leaf_size_source, mac)
        if not ok:
            max_expansion_order = original_max
            t, optargs, ok = call(expansion_order if fixed else max_expansion_order, leaf_size_source, mac)
        expansion_order = optargs["expansion_order"]
        t, optargs, ok = call(expansion_order, leaf_size_source, mac)
        if t < best[0]:
            best = (t, expansion_order, leaf_size_source, mac)
        if verbose:
            print(f"\n\tBest Parameters: \n\t\tleaf_size_source:    {leaf_size_source}\n"
                  f"\t\texpansion_order:     {expansion_order}\n\t\tmultipole_acceptance: {mac}\n"
                  f"\t\tcost:                {t} seconds")
    # tuned is False when every acceptance was skipped (tolerance missed at the maximum order): the
    # parameters are then the ones passed in, not a measured choice
    tuned = {"leaf_size_source": best[2], "expansion_order": best[1], "multipole_acceptance": best[3],
             "tuned": math.isfinite(best[0])}
    if verbose:
        print(f"\nFinished autotune!\n\nParameters: \n\tleaf_size_source:    {best[2]}\n"
              f"\texpansion_order:     {best[1]}\n\tmultipole_acceptance: {best[3]}\n\tcost:                {best[0]} seconds"
              "\n\n#===============================================#\n")
    return tuned, cache


def autotune(system, *, error_tolerance=None, expansion_order=20, leaf_sizes=(16, 24, 32, 48, 64, 96, 128),
             multipole_acceptances=(0.4, 0.45, 0.5, 0.55, 0.6), free_slots=(0, 1, 2), start=None, rounds=2,
             evaluation=None, **kwargs):
    """The project's own tuner for this GPU (no counterpart in the reference; tune_fmm above is the
    reference's heuristic). Every candidate is a complete evaluation of `system` (trees, lists,
    estimate_influence, all passes), priced by the library's device phase times (ms_tree + ms_lists +
    ms_total of the second of two runs). Coordinate descent over the leaf size, the multipole acceptance
    (error-controlled orders only: with a fixed order the acceptance sets the accuracy and is kept) and
    the nearfield's launch shape (workgroup slots per compute unit left to the far-field stream), for
    `rounds` rounds from `start` (default: the middle candidates). Settings whose evaluation misses the
    error tolerance (error_success false) are never chosen.

    Returns (params, evaluation, history): params has leaf_size, multipole_acceptance, p2p_free_slots,
    ms (device time) and expansion_order (the largest order used); evaluation is left at the chosen
    setting; history lists every measured candidate."""
    kernel = _kernel_of(system)
    ev = evaluation or Evaluation()
    ev.set_systems(kernel, None, system.source_buffer())
    fixed = error_tolerance is None
    macs = (kwargs.pop("multipole_acceptance", 0.5),) if fixed else tuple(multipole_acceptances)
    cur = dict(start or {})
    cur.setdefault("leaf_size", leaf_sizes[len(leaf_sizes) // 2])
    cur.setdefault("multipole_acceptance", macs[len(macs) // 2])
    cur.setdefault("p2p_free_slots", 1)
    seen, history = {}, []

    def price(leaf, mac, slots):
        key = (leaf, mac, slots)
        if key in seen:
            return seen[key]
        ev.ctx.set_p2p_free_slots(slots)
        opt = make_options(expansion_order=expansion_order, leaf_size_source=leaf, multipole_acceptance=mac,
                           error_tolerance=error_tolerance, **kwargs)
        ev.fmm(opt)
        st = ev.fmm(opt)
        ms = st.ms_tree + st.ms_lists + st.ms_total
        ok = bool(st.error_success)
        seen[key] = (ms, ok, int(st.pmax))
        history.append({"leaf_size": leaf, "multipole_acceptance": mac, "p2p_free_slots": slots, "ms": ms,
                        "error_success": ok, "pmax": int(st.pmax)})
        return seen[key]

    def better(a, b):   # a, b: (ms, ok, pmax); feasible beats infeasible, then faster
        return (a[1] and not b[1]) or (a[1] == b[1] and a[0] < b[0])

    best = price(cur["leaf_size"], cur["multipole_acceptance"], cur["p2p_free_slots"])
    for _ in range(rounds):
        for name, cands in (("leaf_size", leaf_sizes), ("multipole_acceptance", macs), ("p2p_free_slots", free_slots)):
            for v in cands:
                trial = dict(cur, **{name: v})
                r = price(trial["leaf_size"], trial["multipole_acceptance"], trial["p2p_free_slots"])
                if better(r, best):
                    best, cur = r, trial
    # leave the evaluation and its context at the chosen setting
    price_key = (cur["leaf_size"], cur["multipole_acceptance"], cur["p2p_free_slots"])
    ev.ctx.set_p2p_free_slots(cur["p2p_free_slots"])
    ev.fmm(make_options(expansion_order=expansion_order, leaf_size_source=cur["leaf_size"],
                        multipole_acceptance=cur["multipole_acceptance"], error_tolerance=error_tolerance, **kwargs))
    params = dict(cur, ms=seen[price_key][0], error_success=seen[price_key][1], expansion_order=seen[price_key][2])
    return params, ev, history


def direct(target_systems, source_systems=None, *, scalar_potential=False, gradient=True, hessian=False, precision="fp64",
           ctx=None):
    """direct!(targets, sources; scalar_potential, gradient, hessian) of src/direct.jl:20-75, on the GPU.

    Tuples of systems: every source system acts on every target system, sources outermost as in
    src/direct.jl:64-69; the switches may be given per target system. Returns the (n, 16) target buffer, or
    one per target system for a tuple."""
    if source_systems is None:
        source_systems = target_systems
    targets, sources = _as_tuple(target_systems), _as_tuple(source_systems)
    ctx = ctx or default_context()
    prec = {"fp64": L.FP64, "mixed": L.MIXED}[precision]

    def per_target(v):
        v = list(v) if np.ndim(v) > 0 else [v] * len(targets)
        if len(v) != len(targets):
            raise ValueError("one derivative switch per target system")
        return [bool(x) for x in v]

    ps, gs, hs = per_target(scalar_potential), per_target(gradient), per_target(hessian)
    txyz = [np.ascontiguousarray(t.positions(), dtype=np.float64) for t in targets]
    outs = []
    for x in txyz:
        o = np.zeros((x.shape[0], 16))
        o[:, :3] = x
        outs.append(o)
    for ssys in sources:
        kernel = _kernel_of(ssys)
        sbuf = np.ascontiguousarray(ssys.source_buffer(), dtype=np.float64)
        for i, x in enumerate(txyz):
            ctx.check(ctx.lib.fmmhip_direct(ctx._h, kernel, x.shape[0], x.ctypes.data, sbuf.shape[0], sbuf.ctypes.data,
                                            int(ps[i]), int(gs[i]), int(hs[i]), prec, outs[i].ctypes.data, 0))
    for i, t in enumerate(targets):
        t.buffer_to_target_system(outs[i], DerivativesSwitch(ps[i], gs[i], hs[i]))
    return outs[0] if len(targets) == 1 and not isinstance(target_systems, tuple) else outs
