"""ctypes binding of libmdr_hip.so (the C ABI declared in include/mertens_hip.h).

The binding is a thin host-side mirror of the reference entry points
(mertens_dr::DelégliseRivatEngine::compute_mertens / compute_compressed,
mertens_dr.hpp:968 and :1442) for tests and the benchmark.  It never computes
anything itself: when the HIP library is missing it raises, it does not fall
back to a CPU path.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmdr_hip.so")

STRATEGY_AUTO, STRATEGY_DENSE, STRATEGY_STREAMING, STRATEGY_FORCE_STREAMING = 0, 1, 2, 3
PARTITION_STATIC, PARTITION_LEARNED, PARTITION_TERMS = 0, 1, 2
PARAMS_REFERENCE, PARAMS_TUNED = 0, 1
SCHED_DYNAMIC, SCHED_STATIC = 0, 1

STATUS = {0: "MDR_OK", 1: "MDR_ERR_INVALID", 2: "MDR_ERR_HIP", 3: "MDR_ERR_NOMEM",
          4: "MDR_ERR_NODEVICE", 5: "MDR_ERR_RANGE", 6: "MDR_ERR_NOTABLE"}

# Every symbol include/mertens_hip.h declares.
EXPORTS = ["mdr_default_opts", "mdr_init", "mdr_shutdown", "mdr_last_error", "mdr_mertens",
           "mdr_mertens_batch", "mdr_mertens_dirichlet", "mdr_plan", "mdr_table_build", "mdr_table_read_M",
           "mdr_table_read_mu", "mdr_s2_interval", "mdr_s1_sum", "mdr_g_eval",
           "mdr_cost_model_set", "mdr_cost_model_get", "mdr_cost_model_load", "mdr_cost_model_save",
           "mdr_cycle_log", "mdr_mlp_features", "mdr_mlp_predict", "mdr_mlp_grad", "mdr_mlp_train",
           "mdr_quot_waves", "mdr_table_create", "mdr_table_free", "mdr_table_info", "mdr_table_copy_M",
           "mdr_table_copy_mu", "mdr_table_gather_M", "mdr_table_copy_chunks", "mdr_mertens_noalloc",
           "mdr_param_cx", "mdr_param_dense_cutoff", "mdr_param_streaming_cutoff"]

MLP_NPARAM = 1281
MLP_NFEAT = 5


class Opts(ctypes.Structure):
    _fields_ = [("strategy", ctypes.c_int), ("cx", ctypes.c_double), ("u_cap", ctypes.c_longlong),
                ("fx", ctypes.c_double), ("u_cap_dense", ctypes.c_longlong), ("fx_set", ctypes.c_int),
                ("u_cap_dense_set", ctypes.c_int), ("partition", ctypes.c_int), ("params", ctypes.c_int),
                ("sched", ctypes.c_int), ("host_threads", ctypes.c_int), ("chunk", ctypes.c_longlong),
                ("record_cycles", ctypes.c_int), ("no_table_reuse", ctypes.c_int), ("tune_u", ctypes.c_longlong),
                ("tune_ud", ctypes.c_longlong), ("tune_cx", ctypes.c_double),
                ("use_graph", ctypes.c_int), ("tune_batch", ctypes.c_double), ("sieve_shape", ctypes.c_int),
                ("window", ctypes.c_longlong), ("wave_span", ctypes.c_int)]


class Stats(ctypes.Structure):
    _fields_ = [("host_sieve_time", ctypes.c_double), ("h2d_copy_time", ctypes.c_double),
                ("kernel_time", ctypes.c_double), ("d2h_copy_time", ctypes.c_double),
                ("total_time", ctypes.c_double), ("table_ms", ctypes.c_double),
                ("partition_ms", ctypes.c_double), ("quotient_ms", ctypes.c_double),
                ("table_reused", ctypes.c_int), ("U", ctypes.c_longlong), ("u_dense", ctypes.c_longlong),
                ("mu_limit", ctypes.c_longlong), ("W", ctypes.c_longlong), ("u", ctypes.c_longlong),
                ("N", ctypes.c_longlong), ("slots", ctypes.c_longlong), ("chunks", ctypes.c_longlong),
                ("cx", ctypes.c_double), ("s1_terms", ctypes.c_longlong), ("s2_terms", ctypes.c_longlong),
                ("table_bytes", ctypes.c_double), ("s1_comp_terms", ctypes.c_longlong),
                ("partition_used", ctypes.c_int), ("graph_used", ctypes.c_int),
                ("graph_instantiated", ctypes.c_int), ("graph_ms", ctypes.c_double),
                ("windows", ctypes.c_longlong), ("predicted_cycles", ctypes.c_double),
                ("measured_cycles", ctypes.c_double), ("quotient_bytes", ctypes.c_double),
                ("wave_span_ms", ctypes.c_double), ("wave_busy_ms", ctypes.c_double),
                ("wave_busy_frac", ctypes.c_double)]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class Stamp(ctypes.Structure):
    _fields_ = [("y", ctypes.c_longlong), ("start", ctypes.c_longlong), ("end", ctypes.c_longlong),
                ("kind", ctypes.c_uint), ("cycles", ctypes.c_uint), ("count", ctypes.c_ulonglong)]


class MdrError(RuntimeError):
    pass


_lib = None


def load():
    """Load libmdr_hip.so; raise if it is absent (no CPU fallback exists)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise MdrError(f"HIP library not built: {LIB_PATH} (run __graft_entry__.build())")
    lib = ctypes.CDLL(LIB_PATH)
    P = ctypes.POINTER
    ll, i32, dbl = ctypes.c_longlong, ctypes.c_int, ctypes.c_double
    vp = ctypes.c_void_p
    lib.mdr_default_opts.argtypes = [P(Opts)]
    lib.mdr_default_opts.restype = None
    lib.mdr_init.argtypes = [P(vp), i32]
    lib.mdr_init.restype = i32
    lib.mdr_shutdown.argtypes = [vp]
    lib.mdr_shutdown.restype = None
    lib.mdr_last_error.argtypes = [vp]
    lib.mdr_last_error.restype = ctypes.c_char_p
    lib.mdr_mertens.argtypes = [vp, ll, P(Opts), P(ll), P(Stats)]
    lib.mdr_mertens.restype = i32
    lib.mdr_mertens_batch.argtypes = [vp, P(ll), ctypes.c_size_t, P(Opts), P(ll), P(Stats)]
    lib.mdr_mertens_batch.restype = i32
    lib.mdr_mertens_noalloc.argtypes = [vp, ll, P(ll), P(Stats)]
    lib.mdr_mertens_noalloc.restype = i32
    lib.mdr_mertens_dirichlet.argtypes = [vp, ll, P(ll), P(Stats)]
    lib.mdr_mertens_dirichlet.restype = i32
    lib.mdr_plan.argtypes = [vp, ll, P(Opts), P(i32), P(ll), P(ll), P(ll), P(ll), P(dbl)]
    lib.mdr_plan.restype = i32
    lib.mdr_param_cx.argtypes = [ll, P(Opts)]
    lib.mdr_param_cx.restype = dbl
    lib.mdr_param_dense_cutoff.argtypes = [ll, P(Opts)]
    lib.mdr_param_dense_cutoff.restype = ll
    lib.mdr_param_streaming_cutoff.argtypes = [ll, P(Opts)]
    lib.mdr_param_streaming_cutoff.restype = ll
    lib.mdr_table_build.argtypes = [vp, ll, ll, ll]
    lib.mdr_table_build.restype = i32
    lib.mdr_table_read_M.argtypes = [vp, ll, ll, P(ctypes.c_int32)]
    lib.mdr_table_read_M.restype = i32
    lib.mdr_table_read_mu.argtypes = [vp, ll, ll, P(ctypes.c_int8)]
    lib.mdr_table_read_mu.restype = i32
    lib.mdr_s2_interval.argtypes = [vp, i32, ll, ll, ll, P(ll), P(ll)]
    lib.mdr_s2_interval.restype = i32
    lib.mdr_s1_sum.argtypes = [vp, ll, ll, ll, i32, P(ll), P(ll)]
    lib.mdr_s1_sum.restype = i32
    lib.mdr_g_eval.argtypes = [vp, ll, i32, ll, dbl, P(ll), P(ll)]
    lib.mdr_g_eval.restype = i32
    lib.mdr_quot_waves.argtypes, lib.mdr_quot_waves.restype = [vp], ll
    lib.mdr_table_create.argtypes, lib.mdr_table_create.restype = [vp, ll, ll, ll, P(vp)], i32
    lib.mdr_table_free.argtypes, lib.mdr_table_free.restype = [vp], None
    lib.mdr_table_info.argtypes, lib.mdr_table_info.restype = [vp, P(ll), P(ll), P(ll)], i32
    lib.mdr_table_copy_M.argtypes, lib.mdr_table_copy_M.restype = [vp, ll, ll, P(ctypes.c_int32)], i32
    lib.mdr_table_copy_mu.argtypes, lib.mdr_table_copy_mu.restype = [vp, ll, ll, P(ctypes.c_int8)], i32
    lib.mdr_table_gather_M.argtypes = [vp, P(ll), ctypes.c_size_t, P(ctypes.c_int32)]
    lib.mdr_table_gather_M.restype = i32
    lib.mdr_table_copy_chunks.argtypes = [vp, ll, ll, ll, vp, P(ctypes.c_int32)]
    lib.mdr_table_copy_chunks.restype = i32
    fp = P(ctypes.c_float)
    sz = ctypes.c_size_t
    lib.mdr_cost_model_set.argtypes, lib.mdr_cost_model_set.restype = [vp, fp, i32], i32
    lib.mdr_cost_model_get.argtypes, lib.mdr_cost_model_get.restype = [vp, fp, i32], i32
    lib.mdr_cost_model_load.argtypes, lib.mdr_cost_model_load.restype = [vp, ctypes.c_char_p], i32
    lib.mdr_cost_model_save.argtypes, lib.mdr_cost_model_save.restype = [vp, ctypes.c_char_p], i32
    lib.mdr_cycle_log.argtypes, lib.mdr_cycle_log.restype = [vp, P(Stamp), sz, P(sz)], i32
    lib.mdr_mlp_features.argtypes, lib.mdr_mlp_features.restype = [P(Stamp), sz, fp, fp], i32
    lib.mdr_mlp_predict.argtypes, lib.mdr_mlp_predict.restype = [vp, fp, fp, sz, fp], i32
    lib.mdr_mlp_grad.argtypes, lib.mdr_mlp_grad.restype = [vp, fp, fp, fp, sz, fp, fp], i32
    lib.mdr_mlp_train.argtypes = [vp, fp, fp, sz, i32, i32, ctypes.c_float, ctypes.c_uint, fp]
    lib.mdr_mlp_train.restype = i32
    _lib = lib
    return lib


def _i128(lo, hi):
    return (hi.value << 64) | (lo.value & ((1 << 64) - 1))


def make_opts(strategy=STRATEGY_AUTO, cx=0.0, u_cap=0, fx=0.0, u_cap_dense=0, fx_set=0,
              u_cap_dense_set=0, partition=PARTITION_STATIC, params=PARAMS_REFERENCE,
              sched=SCHED_DYNAMIC, chunk=0, record_cycles=0, no_table_reuse=0, tune_u=0, tune_ud=0,
              tune_cx=0.0, use_graph=0, tune_batch=0.0, sieve_shape=0, window=0, wave_span=0):
    o = Opts()
    load().mdr_default_opts(ctypes.byref(o))
    o.strategy, o.cx, o.u_cap, o.fx, o.u_cap_dense = strategy, cx, u_cap, fx, u_cap_dense
    o.fx_set, o.u_cap_dense_set, o.partition, o.params = fx_set, u_cap_dense_set, partition, params
    o.sched, o.chunk, o.record_cycles, o.no_table_reuse = sched, chunk, record_cycles, no_table_reuse
    o.tune_u, o.tune_ud, o.tune_cx = tune_u, tune_ud, tune_cx
    o.use_graph, o.tune_batch, o.sieve_shape, o.window = use_graph, tune_batch, sieve_shape, window
    o.wave_span = wave_span
    return o


def opts_from_env(env=None):
    """Parse the reference's strategy controls exactly as mertens_dr.hpp does
    (MERTENS_USE_CHMT :1027/:1430-1435, MERTENS_CX :915-918,
    MERTENS_CHMT_U_CAP :1412-1415, MERTENS_FX :944-947, MERTENS_U_CAP :957-960)."""
    env = os.environ if env is None else env
    kw = {}
    if "MERTENS_USE_CHMT" in env:
        v = env["MERTENS_USE_CHMT"]
        kw["strategy"] = STRATEGY_STREAMING if v[:1] in ("1", "y", "Y", "t", "T") else STRATEGY_DENSE
    if "MERTENS_CX" in env:
        v = _atof(env["MERTENS_CX"])
        if 0.1 < v < 5.0:
            kw["cx"] = v
    if "MERTENS_CHMT_U_CAP" in env:
        v = _atoll(env["MERTENS_CHMT_U_CAP"])
        if 100000000 <= v <= 10000000000:
            kw["u_cap"] = v
    if "MERTENS_FX" in env:
        kw["fx_set"] = 1
        v = _atof(env["MERTENS_FX"])
        if 0.1 <= v <= 3.0:
            kw["fx"] = v
    if "MERTENS_U_CAP" in env:
        kw["u_cap_dense_set"] = 1
        v = _atoll(env["MERTENS_U_CAP"])
        if 100000000 <= v <= 4000000000:
            kw["u_cap_dense"] = v
    return kw


def _atof(s):
    """C atof: longest valid leading float prefix, 0.0 when none."""
    import re
    m = re.match(r"\s*[+-]?(\d+\.?\d*([eE][+-]?\d+)?|\.\d+([eE][+-]?\d+)?)", s)
    return float(m.group(0)) if m else 0.0


def _atoll(s):
    """C atoll: optional sign and leading digits, 0 when none."""
    import re
    m = re.match(r"\s*[+-]?\d+", s)
    return int(m.group(0)) if m else 0


class Engine:
    """One GPU context (mdr_init / mdr_shutdown)."""

    def __init__(self, device=0):
        self.lib = load()
        self.ctx = ctypes.c_void_p()
        st = self.lib.mdr_init(ctypes.byref(self.ctx), device)
        if st != 0:
            raise MdrError(f"mdr_init failed: {STATUS.get(st, st)}")

    def waves(self):
        """Resident waves of the quotient kernel on this GPU."""
        return self.lib.mdr_quot_waves(self.ctx)

    def close(self):
        if self.ctx:
            self.lib.mdr_shutdown(self.ctx)
            self.ctx = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _check(self, st, what):
        if st != 0:
            msg = self.lib.mdr_last_error(self.ctx).decode(errors="replace")
            raise MdrError(f"{what}: {STATUS.get(st, st)}: {msg}")

    def mertens(self, X, opts=None, **kw):
        o = opts if opts is not None else make_opts(**kw)
        out = ctypes.c_longlong()
        stats = Stats()
        self._check(self.lib.mdr_mertens(self.ctx, X, ctypes.byref(o), ctypes.byref(out), ctypes.byref(stats)),
                    f"mdr_mertens({X})")
        self.last_stats = stats
        return out.value

    def mertens_noalloc(self, X):
        """M(X) through the allocation-free entry point (mdr_mertens_noalloc)."""
        out = ctypes.c_longlong()
        stats = Stats()
        self._check(self.lib.mdr_mertens_noalloc(self.ctx, X, ctypes.byref(out), ctypes.byref(stats)),
                    f"mdr_mertens_noalloc({X})")
        self.last_stats = stats
        return out.value

    def dirichlet(self, X):
        """M(X) from the Dirichlet cross-check engine (independent of the
        Deleglise-Rivat split)."""
        out = ctypes.c_longlong()
        stats = Stats()
        self._check(self.lib.mdr_mertens_dirichlet(self.ctx, X, ctypes.byref(out), ctypes.byref(stats)),
                    f"mdr_mertens_dirichlet({X})")
        self.last_stats = stats
        return out.value

    def batch(self, xs, opts=None, **kw):
        o = opts if opts is not None else make_opts(**kw)
        n = len(xs)
        arr = (ctypes.c_longlong * n)(*xs)
        out = (ctypes.c_longlong * n)()
        stats = Stats()
        self._check(self.lib.mdr_mertens_batch(self.ctx, arr, n, ctypes.byref(o), out, ctypes.byref(stats)),
                    "mdr_mertens_batch")
        self.last_stats = stats
        return list(out)

    def plan(self, X, opts=None, **kw):
        o = opts if opts is not None else make_opts(**kw)
        s, u, ud, mu, N, cx = (ctypes.c_int(), ctypes.c_longlong(), ctypes.c_longlong(), ctypes.c_longlong(),
                               ctypes.c_longlong(), ctypes.c_double())
        self._check(self.lib.mdr_plan(self.ctx, X, ctypes.byref(o), *(ctypes.byref(v) for v in (s, u, ud, mu, N, cx))),
                    "mdr_plan")
        return dict(streaming=bool(s.value), u=u.value, u_dense=ud.value, mu_limit=mu.value, N=N.value, cx=cx.value)

    def table_build(self, U, u_dense, mu_limit):
        self._check(self.lib.mdr_table_build(self.ctx, U, u_dense, mu_limit), "mdr_table_build")

    def read_M(self, first, count):
        import numpy as np
        out = np.empty(count, dtype=np.int32)
        self._check(self.lib.mdr_table_read_M(self.ctx, first, count,
                                              out.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))), "read_M")
        return out

    def read_mu(self, first, count):
        import numpy as np
        out = np.empty(count, dtype=np.int8)
        self._check(self.lib.mdr_table_read_mu(self.ctx, first, count,
                                               out.ctypes.data_as(ctypes.POINTER(ctypes.c_int8))), "read_mu")
        return out

    def s2_interval(self, piece, first, last, y):
        lo, hi = ctypes.c_longlong(), ctypes.c_longlong()
        self._check(self.lib.mdr_s2_interval(self.ctx, piece, first, last, y, ctypes.byref(lo), ctypes.byref(hi)),
                    "s2_interval")
        return _i128(lo, hi)

    def s1_sum(self, y, first, last, stride):
        lo, hi = ctypes.c_longlong(), ctypes.c_longlong()
        self._check(self.lib.mdr_s1_sum(self.ctx, y, first, last, stride, ctypes.byref(lo), ctypes.byref(hi)),
                    "s1_sum")
        return _i128(lo, hi)

    def g_eval(self, y, comb2, u, cx):
        lo, hi = ctypes.c_longlong(), ctypes.c_longlong()
        self._check(self.lib.mdr_g_eval(self.ctx, y, 1 if comb2 else 0, u, cx, ctypes.byref(lo), ctypes.byref(hi)),
                    "g_eval")
        return _i128(lo, hi)


class Table:
    """A device-resident table object (mdr_table_create): M(n) for n <= U,
    mu(n) for n <= mu_limit, built on the engine's GPU."""

    def __init__(self, engine, U, u_dense, mu_limit):
        self.engine = engine
        self.h = ctypes.c_void_p()
        engine._check(engine.lib.mdr_table_create(engine.ctx, U, u_dense, mu_limit, ctypes.byref(self.h)),
                      "mdr_table_create")

    def close(self):
        if self.h:
            self.engine.lib.mdr_table_free(self.h)
            self.h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def info(self):
        U, ud, mu = ctypes.c_longlong(), ctypes.c_longlong(), ctypes.c_longlong()
        self.engine._check(self.engine.lib.mdr_table_info(self.h, ctypes.byref(U), ctypes.byref(ud), ctypes.byref(mu)),
                           "mdr_table_info")
        return U.value, ud.value, mu.value

    def M(self, first, count):
        import numpy as np
        out = np.empty(count, dtype=np.int32)
        self.engine._check(self.engine.lib.mdr_table_copy_M(self.h, first, count,
                                                            out.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))),
                           "mdr_table_copy_M")
        return out

    def mu(self, first, count):
        import numpy as np
        out = np.empty(count, dtype=np.int8)
        self.engine._check(self.engine.lib.mdr_table_copy_mu(self.h, first, count,
                                                             out.ctypes.data_as(ctypes.POINTER(ctypes.c_int8))),
                           "mdr_table_copy_mu")
        return out

    def gather_M(self, qs):
        import numpy as np
        q = np.ascontiguousarray(qs, dtype=np.int64)
        out = np.empty(len(q), dtype=np.int32)
        self.engine._check(self.engine.lib.mdr_table_gather_M(
            self.h, q.ctypes.data_as(ctypes.POINTER(ctypes.c_longlong)), len(q),
            out.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))), "mdr_table_gather_M")
        return out

    def chunks(self, u_dense_ref, first_chunk, count):
        """(uint8[count, 64] chunks {nonzero[32], sign[32]}, int32[count * 4] checkpoints)."""
        import numpy as np
        ch = np.empty((count, 64), dtype=np.uint8)
        ck = np.empty(count * 4, dtype=np.int32)
        self.engine._check(self.engine.lib.mdr_table_copy_chunks(
            self.h, u_dense_ref, first_chunk, count, ch.ctypes.data_as(ctypes.c_void_p),
            ck.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))), "mdr_table_copy_chunks")
        return ch, ck


def _fptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def mlp_features(stamps):
    """Features (n x 5) and targets log2(cycles/count) of cycle-log records."""
    import numpy as np
    n = len(stamps)
    f = np.empty((n, MLP_NFEAT), dtype=np.float32)
    t = np.empty(n, dtype=np.float32)
    arr = (Stamp * n)(*stamps) if not isinstance(stamps, ctypes.Array) else stamps
    st = load().mdr_mlp_features(arr, n, _fptr(f), _fptr(t))
    if st != 0:
        raise MdrError(f"mdr_mlp_features: {STATUS.get(st, st)}")
    return f, t


class _CostModel:
    """Cost-model methods of Engine."""

    def cost_model_set(self, w):
        import numpy as np
        if w is None:
            self._check(self.lib.mdr_cost_model_set(self.ctx, None, 0), "cost_model_set")
            return
        w = np.ascontiguousarray(w, dtype=np.float32)
        self._check(self.lib.mdr_cost_model_set(self.ctx, _fptr(w), len(w)), "cost_model_set")

    def cost_model_get(self):
        import numpy as np
        w = np.empty(MLP_NPARAM, dtype=np.float32)
        self._check(self.lib.mdr_cost_model_get(self.ctx, _fptr(w), MLP_NPARAM), "cost_model_get")
        return w

    def cost_model_load(self, path):
        self._check(self.lib.mdr_cost_model_load(self.ctx, path.encode()), "cost_model_load")

    def cost_model_save(self, path):
        self._check(self.lib.mdr_cost_model_save(self.ctx, path.encode()), "cost_model_save")

    def cycle_log(self):
        n = ctypes.c_size_t()
        self._check(self.lib.mdr_cycle_log(self.ctx, None, 0, ctypes.byref(n)), "cycle_log")
        arr = (Stamp * max(1, n.value))()
        self._check(self.lib.mdr_cycle_log(self.ctx, arr, n.value, ctypes.byref(n)), "cycle_log")
        return arr[:n.value]

    def mlp_predict(self, w, feats):
        import numpy as np
        w = np.ascontiguousarray(w, dtype=np.float32)
        feats = np.ascontiguousarray(feats, dtype=np.float32)
        out = np.empty(len(feats), dtype=np.float32)
        self._check(self.lib.mdr_mlp_predict(self.ctx, _fptr(w), _fptr(feats), len(feats), _fptr(out)), "mlp_predict")
        return out

    def mlp_grad(self, w, feats, targets):
        import numpy as np
        w = np.ascontiguousarray(w, dtype=np.float32)
        feats = np.ascontiguousarray(feats, dtype=np.float32)
        targets = np.ascontiguousarray(targets, dtype=np.float32)
        g = np.empty(MLP_NPARAM, dtype=np.float32)
        loss = ctypes.c_float()
        self._check(self.lib.mdr_mlp_grad(self.ctx, _fptr(w), _fptr(feats), _fptr(targets), len(feats), _fptr(g),
                                          ctypes.byref(loss)), "mlp_grad")
        return g, loss.value

    def mlp_train(self, feats, targets, epochs=30, batch=1024, lr=1e-3, seed=1):
        import numpy as np
        feats = np.ascontiguousarray(feats, dtype=np.float32)
        targets = np.ascontiguousarray(targets, dtype=np.float32)
        loss = ctypes.c_float()
        self._check(self.lib.mdr_mlp_train(self.ctx, _fptr(feats), _fptr(targets), len(feats), epochs, batch, lr,
                                           seed, ctypes.byref(loss)), "mlp_train")
        return loss.value


for _name in dir(_CostModel):
    if not _name.startswith("_"):
        setattr(Engine, _name, getattr(_CostModel, _name))
