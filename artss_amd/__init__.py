"""MI355X-native ARTSS solver hot path: Python bindings of the C ABI.

The product is artss_amd/libartss_hip.so (HIP kernels for gfx950, C++ host
code, C ABI declared in include/artss_hip.h).  This module is a thin ctypes
layer over that ABI for tests, benchmarks and scripting; it performs no
computation of its own and raises if the library has not been built
(`python -m artss_amd.build`).
"""
import ctypes as C
import os
import weakref

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libartss_hip.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "artss_hip.h")
CONFIG_DIR = os.path.join(HERE, "configs")

FIELD_TYPES = {"rho": 0, "u": 1, "v": 2, "w": 3, "p": 4, "T": 5, "nu": 6}
RHO, U, V, W, P, T, NU = range(7)

_lib = None


class ArtssError(RuntimeError):
    pass


class NotConverged(ArtssError):
    """ARTSS_ERR_NOT_CONVERGED: the step completed (fields advanced, counters
    valid) but its residual-target pressure solve stopped above the target."""

ERR_NOT_CONVERGED = 5


class LevelInfo(C.Structure):
    _fields_ = [("N", C.c_int * 3), ("n", C.c_int * 3), ("size", C.c_longlong), ("h", C.c_double * 3),
                ("inner_cells", C.c_longlong), ("n_obstacles", C.c_int)]


def lib():
    """Loads libartss_hip.so; raises if it is missing (no fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ArtssError("artss_amd/libartss_hip.so is not built; run `python -m artss_amd.build`")
    L = C.CDLL(LIB_PATH)
    vp, dp, ip = C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int)
    sig = {
        "artss_last_error": (C.c_char_p, []),
        "artss_version": (C.c_int, []),
        "artss_domain_create": (C.c_int, [C.c_char_p, C.c_int, C.POINTER(vp)]),
        "artss_domain_destroy": (C.c_int, [vp]),
        "artss_domain_levels": (C.c_int, [vp, ip]),
        "artss_domain_level_info": (C.c_int, [vp, C.c_int, C.POINTER(LevelInfo)]),
        "artss_domain_list": (C.c_int, [vp, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_uint64), C.c_size_t,
                                        C.POINTER(C.c_size_t)]),
        "artss_domain_synchronize": (C.c_int, [vp]),
        "artss_field_alloc": (C.c_int, [vp, C.c_int, C.POINTER(vp)]),
        "artss_field_free": (C.c_int, [vp, vp]),
        "artss_field_upload": (C.c_int, [vp, vp, dp, C.c_size_t]),
        "artss_field_download": (C.c_int, [vp, dp, vp, C.c_size_t]),
        "artss_apply_boundary": (C.c_int, [vp, vp, C.c_int, C.c_int]),
        "artss_sl_advect": (C.c_int, [vp, vp, vp, vp, vp, vp, C.c_int]),
        "artss_diffuse": (C.c_int, [vp, C.c_char_p, vp, vp, vp, C.c_double, vp, C.c_int, C.c_int, C.c_double,
                                    C.c_double, ip, dp]),
        "artss_jacobi_step": (C.c_int, [vp, C.c_int, vp, vp, vp] + [C.c_double] * 6),
        "artss_cgs_step": (C.c_int, [vp, C.c_int, vp, vp] + [C.c_double] * 6),
        "artss_divergence": (C.c_int, [vp, vp, vp, vp, vp]),
        "artss_projection": (C.c_int, [vp] + [vp] * 7),
        "artss_smagorinsky": (C.c_int, [vp, vp, vp, vp, vp, C.c_double, C.c_int]),
        "artss_add_source": (C.c_int, [vp, vp, vp, C.c_int]),
        "artss_buoyancy_force": (C.c_int, [vp, vp, vp, vp]),
        "artss_gauss_source": (C.c_int, [vp, vp, C.c_double]),
        "artss_mg_create": (C.c_int, [vp, C.POINTER(vp)]),
        "artss_mg_destroy": (C.c_int, [vp]),
        "artss_mg_pressure": (C.c_int, [vp, vp, vp, C.c_double, ip]),
        "artss_mg_vcycle": (C.c_int, [vp, vp, vp]),
        "artss_mg_residual": (C.c_int, [vp, vp, vp, dp]),
        "artss_mg_residual_static": (C.c_int, [vp, vp, vp, dp]),
        "artss_mg_set_schedule": (C.c_int, [vp, C.c_int, C.c_longlong]),
        "artss_mg_tail_info": (C.c_int, [vp, ip, ip]),
        "artss_solver_create": (C.c_int, [C.c_char_p, C.c_int, C.POINTER(vp)]),
        "artss_solver_destroy": (C.c_int, [vp]),
        "artss_solver_domain": (C.c_int, [vp, C.POINTER(vp)]),
        "artss_solver_step": (C.c_int, [vp, C.c_double]),
        "artss_solver_run": (C.c_int, [vp, C.c_long, C.c_long, dp]),
        "artss_solver_field": (C.c_int, [vp, C.c_char_p, C.POINTER(vp), C.POINTER(C.c_size_t)]),
        "artss_solver_counters": (C.c_int, [vp, ip, ip, dp]),
        "artss_solver_set_pressure_mode": (C.c_int, [vp, C.c_int, C.c_double, C.c_int]),
        "artss_solver_set_mg_schedule": (C.c_int, [vp, C.c_int, C.c_longlong]),
        "artss_solver_set_pressure_method": (C.c_int, [vp, C.c_int, C.c_int]),
        "artss_solver_write_dat": (C.c_int, [vp, C.c_char_p, ip]),
        "artss_solver_set_warm_start": (C.c_int, [vp, C.c_int, vp, C.c_int]),
        "artss_solver_set_warm_start_depth": (C.c_int, [vp, C.c_int, C.c_int]),
        "artss_solver_set_step_graph": (C.c_int, [vp, C.c_int, C.c_int]),
        "artss_solver_step_graph_stats": (C.c_int, [vp, C.POINTER(C.c_longlong), C.POINTER(C.c_longlong),
                                                    C.POINTER(C.c_longlong), C.POINTER(C.c_int),
                                                    C.POINTER(C.c_char_p)]),
        "artss_solver_set_mixed_sweeps": (C.c_int, [vp, C.c_int, C.c_int, C.c_double]),
        "artss_solver_set_coarse_solve": (C.c_int, [vp, C.c_int]),
        "artss_solver_set_mixed_tail": (C.c_int, [vp, C.c_int]),
        "artss_solver_set_coarse_tolerance": (C.c_int, [vp, C.c_double]),
        "artss_solver_fallbacks": (C.c_int, [vp, C.POINTER(C.c_longlong)]),
        "artss_tuned_cells_per_thread": (C.c_int, [C.c_longlong]),
        "artss_solver_snapshot": (C.c_int, [vp, C.c_double, ip]),
        "artss_solver_restore": (C.c_int, [vp, C.c_int, dp]),
        "artss_solver_drop_snapshot": (C.c_int, [vp, C.c_int]),
        "artss_solver_save": (C.c_int, [vp, C.c_char_p, C.c_double]),
        "artss_solver_load": (C.c_int, [vp, C.c_char_p, dp]),
        "artss_ensemble_create": (C.c_int, [C.c_int, C.POINTER(vp)]),
        "artss_ensemble_destroy": (C.c_int, [vp]),
        "artss_ensemble_add": (C.c_int, [vp, vp]),
        "artss_ensemble_step": (C.c_int, [vp, C.c_long]),
        "artss_solver_warm_sample": (C.c_int, [vp, vp, vp, ip]),
        "artss_solver_initial_residual": (C.c_int, [vp, dp]),
        "artss_buffer_alloc": (C.c_int, [vp, C.c_size_t, C.POINTER(vp)]),
        "artss_buffer_free": (C.c_int, [vp, vp]),
        "artss_buffer_upload": (C.c_int, [vp, vp, vp, C.c_size_t]),
        "artss_buffer_download": (C.c_int, [vp, vp, vp, C.c_size_t]),
        "artss_unet_create": (C.c_int, [vp, C.c_ulonglong, C.POINTER(vp)]),
        "artss_unet_destroy": (C.c_int, [vp]),
        "artss_unet_param_count": (C.c_int, [C.POINTER(C.c_size_t)]),
        "artss_unet_set_params": (C.c_int, [vp, vp]),
        "artss_unet_get_params": (C.c_int, [vp, vp]),
        "artss_unet_forward": (C.c_int, [vp, C.c_int, C.c_int, C.c_int, vp, vp]),
        "artss_unet_gradients": (C.c_int, [vp, C.c_int, C.c_int, C.c_int, vp, vp, vp, C.POINTER(C.c_float)]),
        "artss_unet_adam_step": (C.c_int, [vp, C.c_float]),
        "artss_profile_arm": (C.c_int, [vp, C.c_char_p]),
        "artss_profile_read": (C.c_int, [vp, C.POINTER(C.c_longlong), dp]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def tuned_cells_per_thread(level_size):
    """cells per thread along z the autotuner picked for levels of this array size (0: not tuned)"""
    return int(lib().artss_tuned_cells_per_thread(int(level_size)))


def _check(rc):
    if rc == ERR_NOT_CONVERGED:
        raise NotConverged("artss error %d: %s" % (rc, lib().artss_last_error().decode()))
    if rc != 0:
        raise ArtssError("artss error %d: %s" % (rc, lib().artss_last_error().decode()))


def _as_text(xml):
    if not xml.lstrip().startswith("<"):
        with open(xml) as fh:
            xml = fh.read()
    return xml.encode()


class Domain:
    """DomainData + DomainController for one XML input.  device=-1: host-only
    (geometry and index lists without a GPU)."""

    def __init__(self, xml=None, device=0, _handle=None, _owner=None):
        self.L = lib()
        self._owner = _owner
        if _handle is not None:
            self.h = _handle
            self._owned = False
        else:
            h = C.c_void_p()
            _check(self.L.artss_domain_create(_as_text(xml), device, C.byref(h)))
            self.h = h
            self._owned = True
        n = C.c_int()
        _check(self.L.artss_domain_levels(self.h, C.byref(n)))
        self.levels = n.value
        self._dependents = weakref.WeakSet()

    def _register(self, obj):
        """obj (field, buffer, multigrid, network) uses this domain's device
        memory or stream; it is released before the domain goes away."""
        self._dependents.add(obj)

    def _release_dependents(self):
        for obj in list(getattr(self, "_dependents", ())):
            obj._release()

    def close(self):
        if self.h:
            self._release_dependents()
        if getattr(self, "_owned", False) and self.h:
            _check(self.L.artss_domain_destroy(self.h))
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def info(self, level=0):
        i = LevelInfo()
        _check(self.L.artss_domain_level_info(self.h, level, C.byref(i)))
        return dict(N=tuple(i.N), n=tuple(i.n), size=i.size, h=tuple(i.h), inner_cells=i.inner_cells,
                    n_obstacles=i.n_obstacles)

    def size(self, level=0):
        return self.info(level)["size"]

    def list(self, which, level=0, obstacle=0):
        n = C.c_size_t()
        _check(self.L.artss_domain_list(self.h, which, level, obstacle, None, 0, C.byref(n)))
        out = np.zeros(n.value, np.uint64)
        if n.value:
            _check(self.L.artss_domain_list(self.h, which, level, obstacle,
                                            out.ctypes.data_as(C.POINTER(C.c_uint64)), n.value, C.byref(n)))
        return out

    def sync(self):
        _check(self.L.artss_domain_synchronize(self.h))

    def profile_arm(self, tag):
        """Time every launch of one kernel ("<kernel>@<level>") with HIP events."""
        _check(self.L.artss_profile_arm(self.h, tag.encode()))

    def profile_read(self):
        n = C.c_longlong()
        ms = C.c_double()
        _check(self.L.artss_profile_read(self.h, C.byref(n), C.byref(ms)))
        return n.value, ms.value

    # ---- device fields
    def alloc(self, level=0):
        return DeviceField(self, level)

    def field(self, host, level=0):
        f = DeviceField(self, level)
        f.upload(host)
        return f

    # ---- operators (all synchronous)
    def apply_boundary(self, f, ft, level=0):
        _check(self.L.artss_apply_boundary(self.h, f.ptr, ft, level))

    def sl_advect(self, out, inp, u, v, w, ft):
        _check(self.L.artss_sl_advect(self.h, out.ptr, inp.ptr, u.ptr, v.ptr, w.ptr, ft))

    def diffuse(self, method, out, inp, b, D, ev, ft, max_iter, tol, w):
        it = C.c_int()
        res = C.c_double()
        _check(self.L.artss_diffuse(self.h, method.encode(), out.ptr, inp.ptr, b.ptr, D, ev.ptr if ev else None, ft,
                                    max_iter, tol, w, C.byref(it), C.byref(res)))
        return it.value, res.value

    def jacobi_step(self, level, out, inp, b, ax, ay, az, beta, dsign, w):
        _check(self.L.artss_jacobi_step(self.h, level, out.ptr, inp.ptr, b.ptr, ax, ay, az, beta, dsign, w))

    def cgs_step(self, level, out, b, ax, ay, az, beta, dsign, w):
        _check(self.L.artss_cgs_step(self.h, level, out.ptr, b.ptr, ax, ay, az, beta, dsign, w))

    def divergence(self, out, x, y, z):
        _check(self.L.artss_divergence(self.h, out.ptr, x.ptr, y.ptr, z.ptr))

    def projection(self, ou, ov, ow, iu, iv, iw, ip):
        _check(self.L.artss_projection(self.h, ou.ptr, ov.ptr, ow.ptr, iu.ptr, iv.ptr, iw.ptr, ip.ptr))

    def smagorinsky(self, ev, u, v, w, cs, ft=NU):
        _check(self.L.artss_smagorinsky(self.h, ev.ptr, u.ptr, v.ptr, w.ptr, cs, ft))

    def add_source(self, out, s, ft):
        _check(self.L.artss_add_source(self.h, out.ptr, s.ptr, ft))

    def buoyancy_force(self, out, T_, Ta):
        _check(self.L.artss_buoyancy_force(self.h, out.ptr, T_.ptr, Ta.ptr))

    def gauss_source(self, out, t):
        _check(self.L.artss_gauss_source(self.h, out.ptr, t))

    def multigrid(self):
        return Multigrid(self)


class DeviceField:
    """A float64 field in HBM (reference layout) owned by a domain."""

    def __init__(self, dom, level=0):
        self.dom = dom
        self.level = level
        self.n = dom.size(level)
        p = C.c_void_p()
        _check(dom.L.artss_field_alloc(dom.h, level, C.byref(p)))
        self.ptr = p
        self._owned = True
        dom._register(self)

    @classmethod
    def view(cls, dom, ptr, n, level=0):
        f = cls.__new__(cls)
        f.dom, f.level, f.n, f.ptr, f._owned = dom, level, n, C.c_void_p(ptr), False
        return f

    def upload(self, host):
        a = np.ascontiguousarray(host, dtype=np.float64)
        if a.size != self.n:
            raise ValueError("size mismatch: %d != %d" % (a.size, self.n))
        _check(self.dom.L.artss_field_upload(self.dom.h, self.ptr, a.ctypes.data_as(C.POINTER(C.c_double)), a.size))

    def download(self):
        out = np.empty(self.n, np.float64)
        _check(self.dom.L.artss_field_download(self.dom.h, out.ctypes.data_as(C.POINTER(C.c_double)), self.ptr,
                                               self.n))
        return out

    def free(self):
        if self._owned and self.ptr and self.dom.h:
            _check(self.dom.L.artss_field_free(self.dom.h, self.ptr))
        self.ptr = None

    def _release(self):
        self.free()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Multigrid:
    """IPressure::pressure -> VCycleMG with the XML pressure settings."""

    def __init__(self, dom):
        self.dom = dom
        h = C.c_void_p()
        _check(dom.L.artss_mg_create(dom.h, C.byref(h)))
        self.h = h
        dom._register(self)

    def _release(self):
        if self.h:
            self.dom.L.artss_mg_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self._release()
        except Exception:
            pass

    def pressure(self, p, rhs, t):
        c = C.c_int()
        _check(self.dom.L.artss_mg_pressure(self.h, p.ptr, rhs.ptr, t, C.byref(c)))
        return c.value

    def vcycle(self, p, rhs):
        _check(self.dom.L.artss_mg_vcycle(self.h, p.ptr, rhs.ptr))

    def residual(self, p, rhs):
        r = C.c_double()
        _check(self.dom.L.artss_mg_residual(self.h, p.ptr, rhs.ptr, C.byref(r)))
        return r.value

    def residual_static(self, p, rhs):
        """residual() with its partial sums in the library's fixed device
        scratch array: no device allocation (calls overlapping on one GPU race)"""
        r = C.c_double()
        _check(self.dom.L.artss_mg_residual_static(self.h, p.ptr, rhs.ptr, C.byref(r)))
        return r.value

    def set_schedule(self, mode=2, tail_cells=4096):
        """mode 0: one launch per reference operator; 1: folded launches and a
        single-workgroup kernel for levels of <= tail_cells cells; 2: as 1,
        replayed from a captured HIP graph.  Results are bit-identical."""
        _check(self.dom.L.artss_mg_set_schedule(self.h, int(mode), int(tail_cells)))

    def tail_level(self):
        return self.tail_info()[0]

    def tail_info(self):
        """(first level of the single-workgroup kernel, LDS bytes it keeps level arrays in)"""
        f, b = C.c_int(), C.c_int()
        _check(self.dom.L.artss_mg_tail_info(self.h, C.byref(f), C.byref(b)))
        return f.value, b.value


class Solver:
    """SolverController + TimeIntegration for one XML input, resident on one GPU."""

    def __init__(self, xml, device=0):
        self.L = lib()
        h = C.c_void_p()
        _check(self.L.artss_solver_create(_as_text(xml), device, C.byref(h)))
        self.h = h
        d = C.c_void_p()
        _check(self.L.artss_solver_domain(self.h, C.byref(d)))
        self.domain = Domain(_handle=d, _owner=self)

    def close(self):
        if self.h:
            # fields, buffers and networks on the solver's domain go first:
            # they use its stream and device
            self.domain._release_dependents()
            self.domain.h = None
            _check(self.L.artss_solver_destroy(self.h))
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def step(self, t):
        _check(self.L.artss_solver_step(self.h, t))

    def run(self, first_step, count):
        s = C.c_double()
        _check(self.L.artss_solver_run(self.h, first_step, count, C.byref(s)))
        return s.value

    def device_field(self, name):
        p = C.c_void_p()
        n = C.c_size_t()
        _check(self.L.artss_solver_field(self.h, name.encode(), C.byref(p), C.byref(n)))
        return DeviceField.view(self.domain, p.value, n.value)

    def field(self, name):
        return self.device_field(name).download()

    def set_field(self, name, host):
        self.device_field(name).upload(host)

    def counters(self):
        d = (C.c_int * 4)()
        vc = C.c_int()
        r = C.c_double()
        _check(self.L.artss_solver_counters(self.h, d, C.byref(vc), C.byref(r)))
        r0 = C.c_double()
        _check(self.L.artss_solver_initial_residual(self.h, C.byref(r0)))
        return dict(diffusion=list(d), vcycles=vc.value, residual=r.value, initial_residual=r0.value)

    def set_pressure_mode(self, to_tolerance, tolerance=1e-8, max_cycles=200):
        _check(self.L.artss_solver_set_pressure_mode(self.h, 1 if to_tolerance else 0, tolerance, max_cycles))

    def write_dat(self, directory):
        """The reference's final output: <name>.dat and <name>_obstacle.dat per used field."""
        n = C.c_int()
        _check(self.L.artss_solver_write_dat(self.h, str(directory).encode(), C.byref(n)))
        return n.value

    def set_pressure_method(self, method, coarse_max_iter=0):
        """0: V-cycles on their own; 1: V-cycle-preconditioned GCR; 2: GCR with an FP32/FP64
        mixed-precision V-cycle (residual target only); coarse_max_iter > 0 sets the
        sweeps of the GCR preconditioner's coarse solve (0: five)."""
        _check(self.L.artss_solver_set_pressure_method(self.h, int(method), int(coarse_max_iter)))

    def set_step_graph(self, on=True, margin=2):
        """Capture the whole step in one HIP graph after an eager step that records each
        residual-target loop's iterations (budget: that count + margin); a replay that would
        need more is rolled back and the step runs eagerly.  Fields are bit-identical."""
        _check(self.L.artss_solver_set_step_graph(self.h, 1 if on else 0, int(margin)))

    def step_graph_stats(self):
        r, b, c, a, n = C.c_longlong(), C.c_longlong(), C.c_longlong(), C.c_int(), C.c_char_p()
        _check(self.L.artss_solver_step_graph_stats(self.h, C.byref(r), C.byref(b), C.byref(c), C.byref(a),
                                                    C.byref(n)))
        return dict(replays=r.value, rollbacks=b.value, captures=c.value, active=bool(a.value),
                    note=(n.value or b"").decode())

    def set_warm_start(self, mode, net=None, record=False, depth=None):
        """Start of the GCR pressure iteration: 0 the previous p, 1 linear extrapolation in
        time, 2 extrapolation plus the correction of the warm-start UNet `net`, 3 the
        combination of the last `depth` (default 4) increments of the solution closest to
        the solution in the energy norm (Galerkin projection); `record` keeps each step's
        network input and target for warm_sample()."""
        if mode == 3:
            if net is not None or record:
                raise ArtssError("the projection start takes no network and records nothing")
            _check(self.L.artss_solver_set_warm_start_depth(self.h, 3, 4 if depth is None else int(depth)))
            self._warm_net = None
            return
        _check(self.L.artss_solver_set_warm_start(self.h, int(mode), net.h if net is not None else None,
                                                  1 if record else 0))
        self._warm_net = net  # the solver holds a pointer to it
        if net is not None:
            net._users.add(self)

    def warm_sample(self, x_buf, t_buf):
        """Copies the last recorded (input, target) into Buffers of shapes (Z, Y, X, 2) and
        (Z, Y, X, 1), float32; False when nothing is recorded yet."""
        ok = C.c_int()
        _check(self.L.artss_solver_warm_sample(self.h, x_buf.ptr, t_buf.ptr, C.byref(ok)))
        return bool(ok.value)

    def snapshot(self, t):
        """Keeps the state between two steps in HBM; returns its slot."""
        slot = C.c_int()
        _check(self.L.artss_solver_snapshot(self.h, t, C.byref(slot)))
        return slot.value

    def restore(self, slot):
        """Rolls back to a snapshot; returns its time."""
        t = C.c_double()
        _check(self.L.artss_solver_restore(self.h, int(slot), C.byref(t)))
        return t.value

    def drop_snapshot(self, slot):
        _check(self.L.artss_solver_drop_snapshot(self.h, int(slot)))

    def save(self, path, t):
        _check(self.L.artss_solver_save(self.h, os.fspath(path).encode(), t))

    def load(self, path):
        t = C.c_double()
        _check(self.L.artss_solver_load(self.h, os.fspath(path).encode(), C.byref(t)))
        return t.value

    def set_mixed_sweeps(self, pre=0, post=0, weight=0.0):
        """method 2: smoothing sweeps of the FP32 levels and their Jacobi weight (0: the XML's)"""
        _check(self.L.artss_solver_set_mixed_sweeps(self.h, int(pre), int(post), float(weight)))

    def set_coarse_solve(self, direct=True):
        """method 2: dense exact solve of the preconditioner's coarse levels (True) or their V-cycle"""
        _check(self.L.artss_solver_set_coarse_solve(self.h, 1 if direct else 0))

    def set_mixed_tail(self, n_relax=0):
        """method 2: n_relax of the preconditioner's coarse-level kernel (0: the XML's)"""
        _check(self.L.artss_solver_set_mixed_tail(self.h, int(n_relax)))

    def fallbacks(self):
        """steps whose GCR pressure solve stalled and was finished by V-cycles"""
        n = C.c_longlong()
        _check(self.L.artss_solver_fallbacks(self.h, C.byref(n)))
        return n.value

    def set_coarse_tolerance(self, coarse_tol=-1.0):
        """methods 1, 2: stopping test of the preconditioner's coarse solve: 0 the
        XML's absolute tolerance, 0 < coarse_tol < 1 a reduction relative to its
        start, < 0 (default) a fixed number of sweeps (the coarse_max_iter of
        set_pressure_method, 0: five)"""
        _check(self.L.artss_solver_set_coarse_tolerance(self.h, float(coarse_tol)))

    def set_mg_schedule(self, mode=2, tail_cells=4096):
        _check(self.L.artss_solver_set_mg_schedule(self.h, int(mode), int(tail_cells)))


class Buffer:
    """Raw device memory on a domain's device holding a numpy array's bytes."""

    def __init__(self, dom, array):
        self.dom = dom
        a = np.ascontiguousarray(array)
        self.dtype, self.shape, self.nbytes = a.dtype, a.shape, a.nbytes
        h = C.c_void_p()
        _check(dom.L.artss_buffer_alloc(dom.h, a.nbytes, C.byref(h)))
        self.ptr = h
        dom._register(self)
        self.upload(a)

    def upload(self, array):
        a = np.ascontiguousarray(array, dtype=self.dtype)
        assert a.nbytes == self.nbytes
        _check(self.dom.L.artss_buffer_upload(self.dom.h, self.ptr, a.ctypes.data_as(C.c_void_p), a.nbytes))

    def download(self):
        a = np.empty(self.shape, self.dtype)
        _check(self.dom.L.artss_buffer_download(self.dom.h, a.ctypes.data_as(C.c_void_p), self.ptr, a.nbytes))
        return a

    def _release(self):
        if self.ptr and self.dom.h:
            self.dom.L.artss_buffer_free(self.dom.h, self.ptr)
        self.ptr = None

    def __del__(self):
        try:
            self._release()
        except Exception:
            pass


class UNet:
    """Warm-start U-Net (BF16 matrix-core convolutions) on a domain's stream.
    Volumes are (Z, Y, X, C) float32 arrays, sides multiples of 8."""

    def __init__(self, dom, seed=0):
        self.dom = dom
        h = C.c_void_p()
        _check(dom.L.artss_unet_create(dom.h, seed, C.byref(h)))
        self.h = h
        self._users = weakref.WeakSet()  # solvers warm-started by this network
        dom._register(self)

    def _release(self):
        if self.h:
            for s in list(self._users):
                if s.h and getattr(s, "_warm_net", None) is self:
                    s.set_warm_start(0)
            self.dom.L.artss_unet_destroy(self.h)
        self.h = None

    def close(self):
        self._release()

    def __del__(self):
        try:
            self._release()
        except Exception:
            pass

    @staticmethod
    def param_count():
        n = C.c_size_t()
        _check(lib().artss_unet_param_count(C.byref(n)))
        return n.value

    def set_params(self, flat):
        a = np.ascontiguousarray(flat, dtype=np.float32)
        assert a.size == self.param_count()
        _check(self.dom.L.artss_unet_set_params(self.h, a.ctypes.data_as(C.c_void_p)))

    def get_params(self):
        a = np.empty(self.param_count(), np.float32)
        _check(self.dom.L.artss_unet_get_params(self.h, a.ctypes.data_as(C.c_void_p)))
        return a

    def forward(self, x_buf, y_buf):
        Z, Y, X = x_buf.shape[:3]
        _check(self.dom.L.artss_unet_forward(self.h, X, Y, Z, x_buf.ptr, y_buf.ptr))

    def gradients(self, x_buf, t_buf, want_grads=True):
        Z, Y, X = x_buf.shape[:3]
        loss = C.c_float()
        g = np.empty(self.param_count(), np.float32) if want_grads else None
        _check(self.dom.L.artss_unet_gradients(self.h, X, Y, Z, x_buf.ptr, t_buf.ptr,
                                               g.ctypes.data_as(C.c_void_p) if want_grads else None,
                                               C.byref(loss)))
        return loss.value, g

    def adam_step(self, lr):
        _check(self.dom.L.artss_unet_adam_step(self.h, lr))


class Ensemble:
    """Independent solvers of any sizes advanced together: step(n) runs every
    member's step n (at its own n * dt) on `threads` host threads, each
    member on its own stream, so their kernels overlap on the GPU."""

    def __init__(self, members, threads=8):
        self.L = lib()
        self.members = list(members)
        h = C.c_void_p()
        _check(self.L.artss_ensemble_create(int(threads), C.byref(h)))
        self.h = h
        for m in self.members:
            _check(self.L.artss_ensemble_add(self.h, m.h))

    def step(self, n):
        if any(not m.h for m in self.members):
            raise ArtssError("an ensemble member was closed")
        _check(self.L.artss_ensemble_step(self.h, int(n)))

    def close(self):
        if self.h:
            _check(self.L.artss_ensemble_destroy(self.h))
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
