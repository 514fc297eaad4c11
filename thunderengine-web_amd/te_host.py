"""Python host of the MI355X large step (ctypes over include/te.h).

Mirrors the reference's host objects for this path:

* ``Replayer`` - web/replay.js ``Replayer``: executes a planned command tape
  (flat word memory + per-dispatch 64-word argument slots) on the GPU;
* ``Run`` - web/core/driver.js ``buildRun``: builds the step program, the
  FieldArena and the t=0 state from a configuration and steps it.

The library is ``lib/libte.so`` next to this file (built by ``make`` here or
by ``__graft_entry__.build()``).  Nothing here falls back to the CPU: every
call fails loudly when the library or a GPU is missing.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libte.so")

TE_OK = 0
TE_ERR_ARG = -1
TE_ERR_HIP = -2
TE_ERR_UNSUPPORTED = -3
TE_ERR_STATE = -4


class TeError(RuntimeError):
    pass


class Config(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int32), ("ny", ctypes.c_int32), ("nz", ctypes.c_int32),
        ("dx", ctypes.c_float), ("dy", ctypes.c_float), ("dz", ctypes.c_float),
        ("zf", ctypes.POINTER(ctypes.c_float)),
        ("dt", ctypes.c_float), ("nsound", ctypes.c_int32), ("moist", ctypes.c_int32),
        ("open_x", ctypes.c_int32), ("open_y", ctypes.c_int32),
        ("weno_scalars", ctypes.c_int32), ("weno_momentum", ctypes.c_int32),
        ("tke", ctypes.c_int32), ("explicit_vdiff", ctypes.c_int32), ("dissipative_heating", ctypes.c_int32),
        ("rayleigh", ctypes.c_int32), ("rayleigh_zd", ctypes.c_float), ("rayleigh_alpha", ctypes.c_float),
        ("alph", ctypes.c_float), ("kdiv", ctypes.c_float), ("smeps", ctypes.c_float),
        ("umove", ctypes.c_float), ("vmove", ctypes.c_float),
        ("th_sfc", ctypes.c_float), ("th_trop", ctypes.c_float), ("t_trop", ctypes.c_float),
        ("z_trop", ctypes.c_float), ("qv_pbl", ctypes.c_float),
        ("hodo_radius", ctypes.c_float), ("hodo_depth", ctypes.c_float), ("shear_umax", ctypes.c_float),
        ("shear_depth", ctypes.c_float),
        ("bubble_amp", ctypes.c_float), ("bubble_rh", ctypes.c_float), ("bubble_rv", ctypes.c_float),
        ("bubble_zc", ctypes.c_float),
        ("noise_sigma", ctypes.c_float), ("seed", ctypes.c_uint64), ("initial_tke", ctypes.c_int32),
        ("fuse", ctypes.c_int32), ("sounding_seed", ctypes.c_uint64), ("th_pert_sigma", ctypes.c_float),
        ("qv_pert_sigma", ctypes.c_float),
        ("species", ctypes.c_int32), ("microphysics", ctypes.c_int32), ("mass_conservation", ctypes.c_int32),
        ("cache_moisture_turbulence", ctypes.c_int32), ("n_arena_fields", ctypes.c_int32),
        ("arena_fields", ctypes.POINTER(ctypes.c_char_p)),
        ("stretch_x", ctypes.c_int32), ("stretch_y", ctypes.c_int32),
        ("dx_inner", ctypes.c_double), ("dx_outer", ctypes.c_double), ("nos_x_len", ctypes.c_double),
        ("tot_x_len", ctypes.c_double),
        ("dy_inner", ctypes.c_double), ("dy_outer", ctypes.c_double), ("nos_y_len", ctypes.c_double),
        ("tot_y_len", ctypes.c_double),
        ("adaptive_dt", ctypes.c_int32), ("tapfrq", ctypes.c_float), ("statfrq", ctypes.c_float),
        ("rstfrq", ctypes.c_float), ("input_sounding", ctypes.c_char_p), ("sounding_winds", ctypes.c_int32),
    ]


class ClockState(ctypes.Structure):
    """te_clock_state: the adaptive clock after its last call."""
    _fields_ = [("mtime", ctypes.c_double), ("dbldt", ctypes.c_double), ("dt", ctypes.c_float),
                ("dtlast", ctypes.c_float), ("cflmax", ctypes.c_float), ("ksmax", ctypes.c_float),
                ("nsound", ctypes.c_int32), ("steps", ctypes.c_int32), ("fired", ctypes.c_uint32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_lib = None


def lib():
    """Loads libte.so (raises TeError when it is not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise TeError("libte.so is not built (run __graft_entry__.build() or make -C thunderengine-web_amd)")
    L = ctypes.CDLL(LIB_PATH)
    P = ctypes.POINTER
    u32p, u64 = P(ctypes.c_uint32), ctypes.c_uint64
    ctx = ctypes.c_void_p
    sig = {
        "te_program_create": ([u64, ctypes.c_int32, P(ctypes.c_char_p), u32p, P(ctx)], ctypes.c_int),
        "te_upload_words": ([ctx, u64, u32p, u64], ctypes.c_int),
        "te_download_words": ([ctx, u64, u32p, u64], ctypes.c_int),
        "te_run": ([ctx, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32], ctypes.c_int),
        "te_sync": ([ctx], ctypes.c_int),
        "te_profile_ops": ([ctx, ctypes.c_int32, P(ctypes.c_float)], ctypes.c_int),
        "te_region_hash": ([ctx, u64, u64, u32p], ctypes.c_int),
        "te_fnv1a": ([ctx, u64, u64, u32p], ctypes.c_int),
        "te_region_hashes_noalloc": ([ctypes.c_void_p, P(u64), P(u64), ctypes.c_int32, u32p], ctypes.c_int),
        "te_memory_words": ([ctx], u64),
        "te_op_count": ([ctx], ctypes.c_int32),
        "te_op_get": ([ctx, ctypes.c_int32, P(ctypes.c_char_p), u32p], ctypes.c_int),
        "te_last_run_ms": ([ctx, P(ctypes.c_float)], ctypes.c_int),
        "te_last_error": ([ctx], ctypes.c_char_p),
        "te_autotune": ([ctx, ctypes.c_int32, ctypes.c_int32], ctypes.c_int),
        "te_tuned_ops": ([ctx], ctypes.c_int32),
        "te_autotune_ms": ([ctx, P(ctypes.c_float), P(ctypes.c_float)], ctypes.c_int),
        "te_kernel_count": ([], ctypes.c_int32),
        "te_kernel_name": ([ctypes.c_int32], ctypes.c_char_p),
        "te_destroy": ([ctx], None),
        "te_config_defaults": ([P(Config)], None),
        "te_encode": ([P(Config), P(ctx)], ctypes.c_int),
        "te_create": ([P(Config), P(ctx)], ctypes.c_int),
        "te_set_time_step": ([ctx, ctypes.c_float, ctypes.c_int32], ctypes.c_int),
        "te_selftest_math": ([P(ctypes.c_uint64), P(ctypes.c_uint32), P(ctypes.c_uint64)], ctypes.c_int),
        "te_adaptive_start": ([ctx], ctypes.c_int),
        "te_adaptive_step": ([ctx, ctypes.c_int32], ctypes.c_int),
        "te_adaptive_clock": ([ctx, P(ClockState)], ctypes.c_int),
        "te_clock_create": ([ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_int32, ctypes.c_int32,
                             P(ctypes.c_float)], ctypes.c_void_p),
        "te_clock_start": ([ctypes.c_void_p, ctypes.c_float, ctypes.c_float], None),
        "te_clock_begin_step": ([ctypes.c_void_p], None),
        "te_clock_end_step": ([ctypes.c_void_p, ctypes.c_float, ctypes.c_float], None),
        "te_clock_get": ([ctypes.c_void_p, P(ClockState)], None),
        "te_clock_destroy": ([ctypes.c_void_p], None),
        "te_init_state": ([ctx], ctypes.c_int),
        "te_step": ([ctx, ctypes.c_int32], ctypes.c_int),
        "te_field_offset": ([ctx, ctypes.c_char_p, P(u64)], ctypes.c_int),
        "te_field_stride_words": ([ctx], u64),
        "te_region_offset": ([ctx, ctypes.c_char_p, P(u64)], ctypes.c_int),
        "te_field_count": ([ctx], ctypes.c_int32),
        "te_field_name": ([ctx, ctypes.c_int32], ctypes.c_char_p),
        "te_grid": ([ctx, P(ctypes.c_int32)], ctypes.c_int),
        "te_ensemble_create": ([P(Config), ctypes.c_int32, P(ctx)], ctypes.c_int),
        "te_ensemble_step": ([ctx, ctypes.c_int32], ctypes.c_int),
        "te_ensemble_sync": ([ctx], ctypes.c_int),
        "te_ensemble_size": ([ctx], ctypes.c_int32),
        "te_ensemble_member": ([ctx, ctypes.c_int32], ctx),
        "te_ensemble_points": ([ctx], u64),
        "te_ensemble_last_error": ([ctx], ctypes.c_char_p),
        "te_ensemble_destroy": ([ctx], None),
        "te_adjoint_create": ([ctx, ctypes.c_int32, P(ctx)], ctypes.c_int),
        "te_adjoint_program": ([ctx], ctx),
        "te_fwd_ckpt": ([ctx, ctypes.c_int32], ctypes.c_int),
        "te_bwd": ([ctx], ctypes.c_int),
        "te_adj_ops": ([ctx, ctypes.c_int32, ctypes.c_int32], ctypes.c_int),
        "te_adj_upload": ([ctx, u64, P(ctypes.c_float), u64], ctypes.c_int),
        "te_adj_download": ([ctx, u64, P(ctypes.c_float), u64], ctypes.c_int),
        "te_adj_zero": ([ctx], ctypes.c_int),
        "te_adj_last_ms": ([ctx, P(ctypes.c_float), P(ctypes.c_float)], ctypes.c_int),
        "te_adjoint_last_error": ([ctx], ctypes.c_char_p),
        "te_adjoint_destroy": ([ctx], None),
    }
    for name, (args, res) in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = res
    _lib = L
    return L


def kernel_names():
    """Names of the kernels libte.so implements."""
    L = lib()
    return [L.te_kernel_name(i).decode() for i in range(L.te_kernel_count())]


def selftest_math():
    """te_selftest_math: per function (log, exp, pow(cp/rd), pow(rd/cv),
    pow(1/3): the kernels' binary64 shortcuts against the double-float
    routines; recip: the recurrences' reciprocal against 1.0f / x) the
    binary32 inputs where the two differ, the first one, and the inputs
    where the shortcut defers."""
    mis, first, slow = (ctypes.c_uint64 * 6)(), (ctypes.c_uint32 * 6)(), (ctypes.c_uint64 * 6)()
    if lib().te_selftest_math(mis, first, slow) != TE_OK:
        raise TeError("te_selftest_math failed")
    names = ("log", "exp", "pow_cpdrd", "pow_rddcv", "pow_third", "recip")
    return {n: {"mismatch": int(mis[q]), "first": int(first[q]), "deferred": int(slow[q])} for q, n in enumerate(names)}


def config(**kw):
    """te_config with the library defaults (64x64x40 synthetic case) and overrides."""
    c = Config()
    lib().te_config_defaults(ctypes.byref(c))
    keep = []
    for k, v in kw.items():
        if k == "zf":
            arr = (ctypes.c_float * len(v))(*v)
            keep.append(arr)
            c.zf = ctypes.cast(arr, ctypes.POINTER(ctypes.c_float))
        elif k == "input_sounding":
            b = v.encode() if isinstance(v, str) else v
            keep.append(b)
            c.input_sounding = b
        elif k == "arena_fields":
            arr = (ctypes.c_char_p * len(v))(*[s.encode() for s in v])
            keep.append(arr)
            c.arena_fields = arr
            c.n_arena_fields = len(v)
        else:
            setattr(c, k, v)
    c._keep = keep
    return c


def tanh_faces(nz=120, dz_bottom=20.0, dz_top=300.0, scale=5000.0):
    """Face heights (m) of a column whose spacing grows from dz_bottom at the
    surface towards dz_top as dz(z) = dz_bottom + (dz_top - dz_bottom) tanh(z / scale):
    with the defaults 120 levels, 20 m at the surface, 297 m at 16 km, top
    24.7 km (the tornado-resolving grid's column).  For te_config.zf
    (VerticalGrid(faceHeights), web/core/grid.js:256)."""
    z = [0.0]
    for _ in range(nz):
        z.append(z[-1] + dz_bottom + (dz_top - dz_bottom) * float(np.tanh(z[-1] / scale)))
    return z


class _Ctx:
    def __init__(self):
        self.h = ctypes.c_void_p()

    def _check(self, rc, what):
        if rc != TE_OK:
            msg = lib().te_last_error(self.h)
            raise TeError("%s failed (%d): %s" % (what, rc, msg.decode() if msg else "?"))

    def close(self):
        if self.h:
            lib().te_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- memory
    @property
    def mem_words(self):
        return int(lib().te_memory_words(self.h))

    def upload(self, word_off, words):
        w = np.ascontiguousarray(words, dtype=np.uint32)
        self._check(lib().te_upload_words(self.h, word_off, w.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), w.size),
                    "te_upload_words")

    def read(self, word_off, nwords):
        out = np.empty(nwords, dtype=np.uint32)
        self._check(lib().te_download_words(self.h, word_off, out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)),
                                            nwords), "te_download_words")
        return out

    def region_hash(self, word_off, nwords):
        out = ctypes.c_uint32()
        self._check(lib().te_region_hash(self.h, word_off, nwords, ctypes.byref(out)), "te_region_hash")
        return out.value

    def fnv1a(self, word_off, nwords):
        out = ctypes.c_uint32()
        self._check(lib().te_fnv1a(self.h, word_off, nwords, ctypes.byref(out)), "te_fnv1a")
        return out.value

    # ---- execution
    def run(self, nsteps=1, first=0, last=-1):
        self._check(lib().te_run(self.h, nsteps, first, last), "te_run")

    def sync(self):
        self._check(lib().te_sync(self.h), "te_sync")

    def profile_ops(self, nsteps=1):
        """Per-op device milliseconds summed over nsteps eager steps (HIP events)."""
        n = int(lib().te_op_count(self.h))
        out = np.zeros(n, dtype=np.float32)
        self._check(lib().te_profile_ops(self.h, nsteps, out.ctypes.data_as(ctypes.POINTER(ctypes.c_float))),
                    "te_profile_ops")
        return out

    def last_run_ms(self):
        ms = ctypes.c_float()
        self._check(lib().te_last_run_ms(self.h, ctypes.byref(ms)), "te_last_run_ms")
        return ms.value

    def autotune(self, reps=5, force=False):
        """Chooses each generic dispatch's launch shape by timing candidates
        inside whole steps (te_autotune); the memory is left as it was.
        Returns the number of dispatches left on a tuned shape (force: keep
        the choices even when whole steps are not faster with them)."""
        self._check(lib().te_autotune(self.h, reps, int(force)), "te_autotune")
        return int(lib().te_tuned_ops(self.h))

    def autotune_ms(self):
        """(default, tuned) mean eager step ms of the tuner's final comparison."""
        a, b = ctypes.c_float(), ctypes.c_float()
        self._check(lib().te_autotune_ms(self.h, ctypes.byref(a), ctypes.byref(b)), "te_autotune_ms")
        return a.value, b.value

    def program(self):
        """(names, (nops, 64) uint32 argument slots) of the context's program."""
        n = int(lib().te_op_count(self.h))
        names, A = [], np.zeros((n, 64), dtype=np.uint32)
        nm = ctypes.c_char_p()
        for i in range(n):
            self._check(lib().te_op_get(self.h, i, ctypes.byref(nm), A[i].ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))),
                        "te_op_get")
            names.append(nm.value.decode())
        return names, A


class Replayer(_Ctx):
    """web/replay.js Replayer over the C ABI: plan + build + upload + encodeSteps + read."""

    def __init__(self, mem_words, names, args):
        super().__init__()
        n = len(names)
        cn = (ctypes.c_char_p * n)(*[s.encode() for s in names])
        A = np.ascontiguousarray(args, dtype=np.uint32)
        self._check(lib().te_program_create(mem_words, n, cn, A.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)),
                                            ctypes.byref(self.h)), "te_program_create")

    def encode_steps(self, n=1, last_op=-1, first_op=0):
        self.run(n, first_op, last_op)


class Run(_Ctx):
    """buildRun(configuration) (web/core/driver.js:23-75): program + arena + state."""

    def __init__(self, cfg=None, encode_only=False, **kw):
        super().__init__()
        self.cfg = cfg if cfg is not None else config(**kw)
        fn = lib().te_encode if encode_only else lib().te_create
        self._check(fn(ctypes.byref(self.cfg), ctypes.byref(self.h)), "te_encode" if encode_only else "te_create")
        g = (ctypes.c_int32 * 5)()
        self._check(lib().te_grid(self.h, g), "te_grid")
        self.nx, self.ny, self.nz, self.sJ, self.sK = list(g)
        self.field_words = int(lib().te_field_stride_words(self.h))
        self.fields = [lib().te_field_name(self.h, i).decode() for i in range(lib().te_field_count(self.h))]

    def init_state(self):
        self._check(lib().te_init_state(self.h), "te_init_state")

    def step(self, n=1):
        self._check(lib().te_step(self.h, n), "te_step")

    def set_time_step(self, dt, nsound):
        """TimeStep.setTimeStep (web/core/timestep.js:711-716): the step for a new dt and nsound."""
        self._check(lib().te_set_time_step(self.h, dt, nsound), "te_set_time_step")

    def adaptive_start(self):
        """adapt_dt = 1: the t = 0 stability maxima start the clock (driver.js:46-47)."""
        self._check(lib().te_adaptive_start(self.h), "te_adaptive_start")

    def adaptive_step(self, n=1):
        """n large steps, each with the dt and nsound the clock chose from the last one."""
        self._check(lib().te_adaptive_step(self.h, n), "te_adaptive_step")

    def clock(self):
        st = ClockState()
        self._check(lib().te_adaptive_clock(self.h, ctypes.byref(st)), "te_adaptive_clock")
        return st.as_dict()

    def field_offset(self, name):
        o = ctypes.c_uint64()
        self._check(lib().te_field_offset(self.h, name.encode(), ctypes.byref(o)), "te_field_offset(%s)" % name)
        return o.value

    def region_offset(self, name):
        o = ctypes.c_uint64()
        self._check(lib().te_region_offset(self.h, name.encode(), ctypes.byref(o)), "te_region_offset(%s)" % name)
        return o.value

    def field(self, name):
        """FieldArena.field(name): (nz+1, strideK, strideJ) float32 copy of the field."""
        n = self.sJ * self.sK * (self.nz + 1)
        w = self.read(self.field_offset(name), n)
        return w.view(np.float32).reshape(self.nz + 1, self.sK, self.sJ)

    def interior(self, name, stagger=(0, 0, 0)):
        f = self.field(name)
        return f[: self.nz + stagger[2], 3: 3 + self.ny + stagger[1], 3: 3 + self.nx + stagger[0]]


class Clock:
    """AdaptiveTimeStep (web/core/adaptive.js:15-109) through te_clock_*:
    freq maps 'tape' 'restart' 'stat' 'domainDiag' 'azimAvg' 'hifrq' to
    frequencies (s); absent kinds have no clock."""
    KINDS = ("tape", "restart", "stat", "domainDiag", "azimAvg", "hifrq")

    def __init__(self, dtl, min_dx, min_dz=0.0, two_dimensional=False, ks_check=True, freq=None):
        fr = (ctypes.c_float * 6)(*[float((freq or {}).get(k, float("nan"))) for k in self.KINDS])
        self.h = lib().te_clock_create(dtl, min_dx, min_dz, int(two_dimensional), int(ks_check), fr)
        if not self.h:
            raise TeError("te_clock_create failed")

    def start(self, cflmax, ksmax):
        lib().te_clock_start(self.h, cflmax, ksmax)

    def begin_step(self):
        lib().te_clock_begin_step(self.h)

    def end_step(self, cflmax, ksmax):
        lib().te_clock_end_step(self.h, cflmax, ksmax)

    def state(self):
        st = ClockState()
        lib().te_clock_get(self.h, ctypes.byref(st))
        d = st.as_dict()
        d["fired_kinds"] = [k for i, k in enumerate(self.KINDS) if d["fired"] >> i & 1]
        return d

    def fired(self, kind):
        return kind in self.state()["fired_kinds"]

    def close(self):
        if self.h:
            lib().te_clock_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Ensemble:
    """Members of uneven grids / perturbed soundings stepped together
    (te_ensemble_*): each member on its own stream, graph replays."""

    def __init__(self, cfgs):
        arr = (Config * len(cfgs))(*cfgs)
        self._cfgs = cfgs
        self.h = ctypes.c_void_p()
        rc = lib().te_ensemble_create(arr, len(cfgs), ctypes.byref(self.h))
        if rc != TE_OK:
            msg = lib().te_ensemble_last_error(self.h)
            self.close()
            raise TeError("te_ensemble_create failed (%d): %s" % (rc, msg.decode() if msg else "?"))

    def step(self, n=1):
        rc = lib().te_ensemble_step(self.h, n)
        if rc != TE_OK:
            raise TeError("te_ensemble_step failed (%d): %s" % (rc, lib().te_ensemble_last_error(self.h).decode()))

    def sync(self):
        if lib().te_ensemble_sync(self.h) != TE_OK:
            raise TeError("te_ensemble_sync failed")

    def autotune(self, reps=5):
        """te_autotune on every member (each keeps its own shapes)."""
        n = 0
        for m in range(len(self)):
            c = lib().te_ensemble_member(self.h, m)
            if lib().te_autotune(c, reps, 0) != TE_OK:
                raise TeError("te_autotune failed on member %d: %s" % (m, lib().te_last_error(c).decode()))
            n += int(lib().te_tuned_ops(c))
        return n

    @property
    def points(self):
        return int(lib().te_ensemble_points(self.h))

    def __len__(self):
        return int(lib().te_ensemble_size(self.h))

    def member_words(self, m, word_off, nwords):
        """Words of member m's memory (a borrowed context)."""
        c = lib().te_ensemble_member(self.h, m)
        out = np.empty(nwords, dtype=np.uint32)
        if lib().te_download_words(c, word_off, out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), nwords) != TE_OK:
            raise TeError("te_download_words failed")
        return out

    def close(self):
        if self.h:
            lib().te_ensemble_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Adjoint:
    """Adjoint of the large step over a config context (te_adjoint_*): a
    checkpointed forward (te_fwd_ckpt) and the reverse sweep (te_bwd) over the
    recorded steps.  The adjoint memory has M's word layout, one float per
    word: seed it with dJ/d(final state), read dJ/d(initial state)."""

    def __init__(self, run, every=10):
        self.run = run
        self.h = ctypes.c_void_p()
        rc = lib().te_adjoint_create(run.h, every, ctypes.byref(self.h))
        if rc != TE_OK:
            msg = lib().te_adjoint_last_error(self.h)
            self.close()
            raise TeError("te_adjoint_create failed (%d): %s" % (rc, msg.decode() if msg else "?"))
        # the unfused dispatch list on the adjoint's own memory (borrowed)
        self.program = _Borrowed(lib().te_adjoint_program(self.h))

    def _check(self, rc, what):
        if rc != TE_OK:
            msg = lib().te_adjoint_last_error(self.h)
            raise TeError("%s failed (%d): %s" % (what, rc, msg.decode() if msg else "?"))

    def forward(self, nsteps):
        """nsteps large steps of the context, checkpointed."""
        self._check(lib().te_fwd_ckpt(self.h, nsteps), "te_fwd_ckpt")

    def backward(self):
        self._check(lib().te_bwd(self.h), "te_bwd")

    def dispatch_adjoints(self, first, last):
        self._check(lib().te_adj_ops(self.h, first, last), "te_adj_ops")

    def upload(self, word_off, values):
        v = np.ascontiguousarray(values, dtype=np.float32)
        self._check(lib().te_adj_upload(self.h, word_off, v.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), v.size),
                    "te_adj_upload")

    def read(self, word_off, nwords):
        out = np.empty(nwords, dtype=np.float32)
        self._check(lib().te_adj_download(self.h, word_off, out.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), nwords),
                    "te_adj_download")
        return out

    def zero(self):
        self._check(lib().te_adj_zero(self.h), "te_adj_zero")

    def last_ms(self):
        f, b = ctypes.c_float(), ctypes.c_float()
        self._check(lib().te_adj_last_ms(self.h, ctypes.byref(f), ctypes.byref(b)), "te_adj_last_ms")
        return f.value, b.value

    def close(self):
        if self.h:
            lib().te_adjoint_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class _Borrowed(_Ctx):
    """A context owned by another object (not destroyed here)."""

    def __init__(self, h):
        self.h = ctypes.c_void_p(h)

    def close(self):
        self.h = ctypes.c_void_p()


class Box:
    """Word indices of a field (default w) over an (i, j, k) box of interior points:
    i, j, k are (start, stop) ranges."""

    def __init__(self, e, i, j, k, field="w"):
        o = e.field_offset(field)
        ii, jj, kk = np.meshgrid(np.arange(*i), np.arange(*j), np.arange(*k), indexing="ij")
        self.words = (o + (ii + 3) + e.sJ * ((jj + 3) + e.sK * kk)).ravel()


class Controls:
    """Initial-condition controls: an interior value of each named field at
    levels 0..nz-1.  A perturbation of one moves that point's word in the field
    and in its _a copy, and every periodic halo image of them (the t=0 state
    holds the field in both, halos filled periodically)."""

    def __init__(self, e, names):
        self.e = e
        nx, ny, nz, sJ, sK = e.nx, e.ny, e.nz, e.sJ, e.sK
        ii = np.arange(-3, nx + 4)
        jj = np.arange(-3, ny + 4)
        I, J = np.meshgrid(ii, jj, indexing="ij")
        src = ((I % nx) + nx * (J % ny)).ravel()          # interior point each image copies
        rel = ((I + 3) + sJ * (J + 3)).ravel()            # word of the image within a level
        self.maps = []
        for f in names:
            offs = [e.field_offset(f)] + ([e.field_offset(f + "_a")] if f + "_a" in e.fields else [])
            self.maps.append(offs)
        self.src, self.rel = src, rel
        self.per_field = nx * ny * nz
        self.size = self.per_field * len(names)

    def _iter(self):
        e = self.e
        for q, offs in enumerate(self.maps):
            for k in range(e.nz):
                base = k * e.sJ * e.sK
                cols = q * self.per_field + k * e.nx * e.ny + self.src
                for o in offs:
                    yield o + base + self.rel, cols

    def pull(self, Mb):
        """Gradient with respect to the controls from an adjoint memory."""
        g = np.zeros(self.size)
        for words, cols in self._iter():
            np.add.at(g, cols, Mb[words].astype(np.float64))
        return g

    def push(self, d, n):
        """Word perturbation of length n for a control direction d."""
        x = np.zeros(n)
        for words, cols in self._iter():
            x[words] += d[cols]
        return x
