"""ctypes mirror of include/mahyco_gpu.h.

The structures here match the C structures field for field; `Library` binds
every entry point of a shared library that implements the ABI under a given
symbol prefix, and `Simulation` owns one context plus numpy views of the
state.  Nothing in this module computes physics: every number comes from the
C-ABI library it wraps.
"""
import ctypes as C
import os

import numpy as np

MAX_ENV = 8
MAX_BC = 16

OK, STOP = 0, 1
E_ARG, E_DT_MIN, E_HIP, E_STATE, E_UNSUPPORTED, E_ALLOC = -1, -2, -3, -4, -5, -6
ERRORS = {E_ARG: "bad argument", E_DT_MIN: "dt below deltat-min", E_HIP: "HIP error",
          E_STATE: "call out of order", E_UNSUPPORTED: "unsupported option",
          E_ALLOC: "device allocation failed"}

BC_VX, BC_VY, BC_VZ = 0, 1, 2
SURF = {"XMIN": 0, "XMAX": 1, "YMIN": 2, "YMAX": 3, "ZMIN": 4, "ZMAX": 5}
EOS_PERFECT_GAS, EOS_STIFFENED_GAS = 0, 1
LC_RACINE_CUBIQUE_VOLUME, LC_FACES_OPPOSEES = 0, 1
PREC_FP64, PREC_FP32_BULK = 0, 1
LIM_MINMOD, LIM_SUPERBEE, LIM_VANLEER, LIM_LEARNED, LIM_LEARNED_ALL = 0, 1, 2, 64, 65
LIM_MINMODG, LIM_SUPERBEEG, LIM_VANLEERG, LIM_ARITHMETICG, LIM_ULTRABEEG = 3, 4, 5, 6, 7


class MeshDesc(C.Structure):
    _fields_ = [("dim", C.c_int32), ("n", C.c_int32 * 3),
                ("origin", C.c_double * 3), ("length", C.c_double * 3)]


class BC(C.Structure):
    _fields_ = [("surface", C.c_int32), ("type", C.c_int32), ("value", C.c_double)]


class EnvDesc(C.Structure):
    _fields_ = [("eos", C.c_int32), ("adiabatic_cst", C.c_double),
                ("limit_tension", C.c_double)]


class Options(C.Structure):
    _fields_ = [
        ("cfl", C.c_double), ("threshold", C.c_double), ("deltat_init", C.c_double),
        ("deltat_min", C.c_double), ("deltat_max", C.c_double), ("final_time", C.c_double),
        ("with_projection", C.c_int32), ("pseudo_centree", C.c_int32),
        ("schema_csts", C.c_int32), ("with_newton", C.c_int32), ("sans_lagrange", C.c_int32),
        ("longueur_caracteristique", C.c_int32),
        ("nb_bc", C.c_int32), ("bc", BC * MAX_BC),
        ("nb_env", C.c_int32), ("env", EnvDesc * MAX_ENV),
        ("remap_threshold", C.c_double), ("is_euler_scheme", C.c_int32),
        ("ordre_projection", C.c_int32), ("projection_limiteur_id", C.c_int32),
        ("projection_pente_borne", C.c_int32), ("conservation_energie_totale", C.c_int32),
        ("visc_quad_coeff", C.c_double), ("precision", C.c_int32),
        ("land_on_final_time", C.c_int32),
        ("projection_pente_borne_mixte", C.c_int32), ("projection_limiteur_pure_id", C.c_int32),
        ("projection_pente_borne_debar_fix", C.c_int32), ("reserved", C.c_int32 * 3),
    ]


_DP = C.POINTER(C.c_double)
_IP = C.POINTER(C.c_int32)
_FP = C.POINTER(C.c_float)

STATE_CELL = ["density", "internal_energy", "pressure", "sound_speed", "pseudo_viscosity",
              "cell_mass", "fracvol", "mass_fraction"]
STATE_ENV = [f + "_env" for f in STATE_CELL]


class State(C.Structure):
    _fields_ = ([(f, _DP) for f in STATE_CELL] + [("env_mask", _IP)] +
                [(f, _DP) for f in STATE_ENV] +
                [("velocity", _DP), ("node_mass", _DP), ("node_coord", _DP),
                 ("dual_grad_phi", _DP), ("caracteristic_length", _DP),
                 ("global_deltat", C.c_double), ("old_deltat", C.c_double),
                 ("global_time", C.c_double), ("sens_projection", C.c_int32),
                 ("global_iteration", C.c_int32)])


ENTRY_POINTS = ["begin_iteration", "save_values_at_n", "update_velocity",
                "apply_boundary_condition", "update_position", "compute_geometric_values",
                "update_density", "compute_artificial_viscosity",
                "update_energy_and_pressure", "remap"]


class MahycoError(RuntimeError):
    def __init__(self, code, msg=""):
        super().__init__(f"{ERRORS.get(code, code)} ({code}): {msg}")
        self.code = code


class Library:
    """Binds a shared library implementing the mahyco_gpu.h ABI under `prefix`."""

    def __init__(self, path, prefix="mhg_"):
        if not os.path.exists(path):
            raise FileNotFoundError(f"C-ABI library not built: {path}")
        self.path = path
        self.prefix = prefix
        self.lib = C.CDLL(path)
        p = prefix
        L = self.lib
        vp = C.c_void_p

        def fn(name, res, args):
            f = getattr(L, p + name)
            f.restype = res
            f.argtypes = args
            return f

        self.abi_version = fn("abi_version", C.c_int, [])
        self.default_options = fn("default_options", None, [C.POINTER(Options)])
        self.create = fn("create", C.c_int, [C.POINTER(MeshDesc), C.POINTER(Options), C.POINTER(vp)])
        self.destroy = fn("destroy", C.c_int, [vp])
        self.last_error = fn("last_error", C.c_char_p, [vp])
        self.sizes = fn("sizes", C.c_int, [vp, C.POINTER(C.c_int64), C.POINTER(C.c_int64)])
        self.init_sod = fn("init_sod", C.c_int, [vp, C.c_int32])
        self.init_triple_point = fn("init_triple_point", C.c_int, [vp, _DP, _DP])
        self.init_cells = fn("init_cells", C.c_int, [vp, _IP, _DP, _DP, _DP])
        self.upload_state = fn("upload_state", C.c_int, [vp, C.POINTER(State)])
        self.download_state = fn("download_state", C.c_int, [vp, C.POINTER(State)])
        for e in ENTRY_POINTS:
            setattr(self, e, fn(e, C.c_int, [vp]))
        self.compute_delta_t = fn("compute_delta_t", C.c_int, [vp, _DP, _IP])
        self.step = fn("step", C.c_int, [vp, _IP])
        self.run = fn("run", C.c_int, [vp, C.c_int32, _IP])
        self.conservation = fn("conservation", C.c_int, [vp, _DP])
        for g in ["remap_get_ordre_projection", "remap_has_projection_pente_borne",
                  "remap_has_conservation_energie_totale", "remap_is_euler"]:
            setattr(self, g, fn(g, C.c_int, [vp]))
        self.optional = {}
        for name, res, args in [
                ("run_async", C.c_int, [vp, C.c_int32]),
                ("synchronize", C.c_int, [vp]),
                ("kernel_times", C.c_int, [vp, C.c_int32, C.POINTER(C.c_char_p), _DP, _IP, C.c_int32]),
                ("set_threads", C.c_int, [C.c_int32]),
                ("autotune", C.c_int, [vp, C.c_int32, C.c_char_p, C.c_int32]),
                ("variant_count", C.c_int32, [C.c_char_p]),
                ("select_variant", C.c_int, [vp, C.c_char_p, C.c_int32]),
                ("batch_create", C.c_int, [C.c_int32, C.POINTER(MeshDesc), C.POINTER(Options), C.POINTER(vp)]),
                ("batch_destroy", C.c_int, [vp]),
                ("batch_last_error", C.c_char_p, [vp]),
                ("batch_problem", vp, [vp, C.c_int32]),
                ("batch_run", C.c_int, [vp, C.c_int32, _IP, C.POINTER(C.c_int64)]),
                ("batch_status", C.c_int, [vp, C.c_int32, _IP, _DP, _DP, _IP]),
                ("batch_autotune", C.c_int, [vp, C.c_int32, C.c_char_p, C.c_int32]),
                ("set_learned_alpha", C.c_int, [vp, _DP]),
                ("learned_alpha", C.c_int, [vp, _DP]),
                ("learned_features", C.c_int, [vp, _FP, _FP]),
                ("set_limiter_mlp", C.c_int, [vp, vp]),
                ("set_learned_interface_alpha", C.c_int, [vp, C.c_double]),
                ("mlp_create", C.c_int, [C.POINTER(vp)]),
                ("mlp_destroy", C.c_int, [vp]),
                ("mlp_last_error", C.c_char_p, [vp]),
                ("mlp_param_count", C.c_int32, []),
                ("mlp_set_params", C.c_int, [vp, _FP]),
                ("mlp_get_params", C.c_int, [vp, _FP]),
                ("mlp_forward", C.c_int, [vp, _FP, C.c_int64, _FP]),
                ("mlp_set_data", C.c_int, [vp, _FP, _FP, C.c_int64]),
                ("mlp_gradient", C.c_int, [vp, C.c_int64, C.c_int64, _FP, _FP]),
                ("mlp_train", C.c_int, [vp, C.c_int32, C.c_int32, C.c_float, C.c_uint64, _FP]),
                ("remap_prepare", C.c_int, [vp]),
                ("remap_appli", C.c_int, [vp, C.c_int32, C.c_int32, C.c_int32, C.c_int32]),
                ("remap_variables", C.c_int, [vp, C.c_int32, C.c_int32, C.c_int32, C.c_int32]),
                ("remap_resize_variables", C.c_int, [vp, C.c_int32, C.c_int32]),
                ("remap_synchronize_uremap", C.c_int, [vp]),
                ("eos_init", C.c_int, [vp, C.c_int32]),
                ("eos_apply", C.c_int, [vp, C.c_int32]),
                ("eos_apply_one_cell", C.c_int, [vp, C.c_int32, C.c_int64]),
                ("eos_get_adiabatic_cst", C.c_int, [vp, C.c_int32, _DP]),
                ("eos_get_tension_limit_cst", C.c_int, [vp, C.c_int32, _DP]),
                ("compute_pression_moyenne", C.c_int, [vp])]:
            if hasattr(L, p + name):
                self.optional[name] = fn(name, res, args)

    def options(self, **kw):
        o = Options()
        self.default_options(C.byref(o))
        for k, v in kw.items():
            setattr(o, k, v)
        return o


def set_envs(opt, envs):
    """envs: list of (eos, gamma, limit_tension)."""
    opt.nb_env = len(envs)
    for i, (eos, g, t) in enumerate(envs):
        opt.env[i].eos = eos
        opt.env[i].adiabatic_cst = g
        opt.env[i].limit_tension = t
    return opt


def set_bcs(opt, bcs):
    """bcs: list of (surface name, type, value)."""
    opt.nb_bc = len(bcs)
    for i, (s, t, v) in enumerate(bcs):
        opt.bc[i].surface = SURF[s]
        opt.bc[i].type = t
        opt.bc[i].value = v
    return opt


def reflecting_walls(dim):
    b = [("XMIN", BC_VX, 0.), ("XMAX", BC_VX, 0.), ("YMIN", BC_VY, 0.), ("YMAX", BC_VY, 0.)]
    if dim == 3:
        b += [("ZMIN", BC_VZ, 0.), ("ZMAX", BC_VZ, 0.)]
    return b


def mesh_desc(dim, n, origin, length):
    m = MeshDesc()
    m.dim = dim
    for d in range(3):
        m.n[d] = n[d] if d < len(n) else 1
        m.origin[d] = origin[d] if d < len(origin) else 0.
        m.length[d] = length[d] if d < len(length) else 1.
    return m


class Simulation:
    """One context of a C-ABI library (GPU product library or, in tests, the oracle)."""

    def __init__(self, library, mesh, options, ctx=None):
        """ctx: an existing context owned elsewhere (a problem of a Batch)."""
        self.L = library
        self.mesh = mesh
        self.opt = options
        self.dim = mesh.dim
        self.nenv = options.nb_env
        self.owned = ctx is None
        if ctx is None:
            ctx = C.c_void_p()
            r = self.L.create(C.byref(mesh), C.byref(options), C.byref(ctx))
            if r != OK:
                raise MahycoError(r, "create")
        self.ctx = ctx
        nc, nn = C.c_int64(), C.c_int64()
        self.L.sizes(self.ctx, C.byref(nc), C.byref(nn))
        self.nc, self.nn = nc.value, nn.value

    def close(self):
        if self.ctx and self.owned:
            self.L.destroy(self.ctx)
        self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, r, what):
        if r < 0:
            msg = self.L.last_error(self.ctx)
            raise MahycoError(r, f"{what}: {msg.decode() if msg else ''}")
        return r

    # -- init --
    def init_sod(self, cas_test=3):
        return self._check(self.L.init_sod(self.ctx, cas_test), "init_sod")

    def init_triple_point(self, rho=(1.0, 1.0, 0.125), p=(1.0, 0.1, 0.1)):
        r = np.ascontiguousarray(rho, dtype=np.float64)
        q = np.ascontiguousarray(p, dtype=np.float64)
        return self._check(self.L.init_triple_point(self.ctx, r.ctypes.data_as(_DP),
                                                    q.ctypes.data_as(_DP)), "init_triple_point")

    def init_cells(self, env_of_cell, density, pressure, velocity=None):
        e = np.ascontiguousarray(env_of_cell, dtype=np.int32).ravel()
        d = np.ascontiguousarray(density, dtype=np.float64).ravel()
        p = np.ascontiguousarray(pressure, dtype=np.float64).ravel()
        assert e.size == self.nc and d.size == self.nc and p.size == self.nc
        vp = None
        if velocity is not None:
            v = np.ascontiguousarray(velocity, dtype=np.float64).reshape(self.dim, self.nn)
            vp = v.ctypes.data_as(_DP)
        return self._check(self.L.init_cells(self.ctx, e.ctypes.data_as(_IP), d.ctypes.data_as(_DP),
                                             p.ctypes.data_as(_DP), vp), "init_cells")

    # -- state --
    def alloc_state(self):
        nc, nn, ne, dim = self.nc, self.nn, self.nenv, self.dim
        a = {f: np.zeros(nc) for f in STATE_CELL}
        a.update({f: np.zeros((ne, nc)) for f in STATE_ENV})
        a["env_mask"] = np.zeros(nc, dtype=np.int32)
        a["velocity"] = np.zeros((dim, nn))
        a["node_mass"] = np.zeros(nn)
        a["node_coord"] = np.zeros((dim, nn))
        a["dual_grad_phi"] = np.zeros((dim, nn))
        a["caracteristic_length"] = np.zeros(nc)
        return a

    @staticmethod
    def _fill_struct(arrays, scal=None):
        st = State()
        for name, _ in State._fields_:
            if name in arrays and arrays[name] is not None:
                arr = arrays[name]
                assert arr.flags.c_contiguous
                ptype = _IP if arr.dtype == np.int32 else _DP
                setattr(st, name, arr.ctypes.data_as(ptype))
        if scal:
            for k, v in scal.items():
                setattr(st, k, v)
        return st

    def download(self, fields=None):
        a = self.alloc_state()
        if fields is not None:
            a = {k: v for k, v in a.items() if k in fields}
        st = self._fill_struct(a)
        self._check(self.L.download_state(self.ctx, C.byref(st)), "download_state")
        for k in ["global_deltat", "old_deltat", "global_time", "sens_projection",
                  "global_iteration"]:
            a[k] = getattr(st, k)
        return a

    def upload(self, a):
        arrays = {k: np.ascontiguousarray(v) for k, v in a.items() if isinstance(v, np.ndarray)}
        scal = {k: a[k] for k in ["global_deltat", "old_deltat", "global_time",
                                  "sens_projection", "global_iteration"] if k in a}
        st = self._fill_struct(arrays, scal)
        return self._check(self.L.upload_state(self.ctx, C.byref(st)), "upload_state")

    # -- loop --
    def entry(self, name):
        return self._check(getattr(self.L, name)(self.ctx), name)

    def compute_delta_t(self):
        dt = C.c_double()
        stop = C.c_int32()
        self._check(self.L.compute_delta_t(self.ctx, C.byref(dt), C.byref(stop)), "compute_delta_t")
        return dt.value, bool(stop.value)

    # -- the remap and EOS services, called as MahycoModule calls them --
    def remap_args(self):
        """(dimension, with_dual_projection, nb_vars_to_project, nb_env) of
        IRemap::appliRemap for this context."""
        return (self.dim, 1, 3 * self.nenv + 5, self.nenv)

    def remap_prepare(self):
        return self._check(self.L.optional["remap_prepare"](self.ctx), "remap_prepare")

    def remap_appli(self, args=None):
        return self._check(self.L.optional["remap_appli"](self.ctx, *(args or self.remap_args())), "remap_appli")

    def remap_variables(self, args=None):
        return self._check(self.L.optional["remap_variables"](self.ctx, *(args or self.remap_args())),
                           "remap_variables")

    def eos_init(self, env):
        return self._check(self.L.optional["eos_init"](self.ctx, env), "eos_init")

    def eos_apply(self, env):
        return self._check(self.L.optional["eos_apply"](self.ctx, env), "eos_apply")

    def eos_apply_one_cell(self, env, cell):
        return self._check(self.L.optional["eos_apply_one_cell"](self.ctx, env, cell), "eos_apply_one_cell")

    def eos_constants(self, env):
        g, t = C.c_double(), C.c_double()
        self._check(self.L.optional["eos_get_adiabatic_cst"](self.ctx, env, C.byref(g)), "eos_get_adiabatic_cst")
        self._check(self.L.optional["eos_get_tension_limit_cst"](self.ctx, env, C.byref(t)),
                    "eos_get_tension_limit_cst")
        return g.value, t.value

    def compute_pression_moyenne(self):
        return self._check(self.L.optional["compute_pression_moyenne"](self.ctx), "compute_pression_moyenne")

    def remap_split(self):
        """MahycoModule::remap (PrepareRemap.cc:448-563) through the service
        calls: prepare, appliRemap, applyEOS per environment, pressure
        average."""
        self.remap_prepare()
        self.remap_appli()
        for k in range(self.nenv):
            self.eos_apply(k)
        self.compute_pression_moyenne()

    def step(self):
        stop = C.c_int32()
        r = self._check(self.L.step(self.ctx, C.byref(stop)), "step")
        return bool(stop.value) or r == STOP

    def run(self, max_steps):
        done = C.c_int32()
        self._check(self.L.run(self.ctx, max_steps, C.byref(done)), "run")
        return done.value

    def autotune(self, reps=10):
        f = self.L.optional.get("autotune")
        if f is None:
            return ""
        buf = C.create_string_buffer(4096)
        self._check(f(self.ctx, reps, buf, 4096), "autotune")
        return buf.value.decode()

    def variant_count(self, kernel="cell_sweep_y"):
        f = self.L.optional.get("variant_count")
        return 0 if f is None else int(f(kernel.encode()))

    def select_variant(self, variant, kernel="cell_sweep_y"):
        self._check(self.L.optional["select_variant"](self.ctx, kernel.encode(), variant),
                    "select_variant")

    def set_learned_alpha(self, alpha):
        """Fixed learned-limiter coefficients [dim][N_c] (MHG_LIM_LEARNED)."""
        a = np.ascontiguousarray(alpha, dtype=np.float64).reshape(self.dim, self.nc)
        self._check(self.L.optional["set_learned_alpha"](self.ctx, a.ctypes.data_as(_DP)), "set_learned_alpha")

    def learned_alpha(self):
        a = np.zeros((2, self.nc))
        self._check(self.L.optional["learned_alpha"](self.ctx, a.ctypes.data_as(_DP)), "learned_alpha")
        return a

    def learned_features(self):
        X = np.zeros((self.nc, 128), dtype=np.float32)
        S = np.zeros((self.nc, 5), dtype=np.float32)
        self._check(self.L.optional["learned_features"](self.ctx, X.ctypes.data_as(_FP), S.ctypes.data_as(_FP)),
                    "learned_features")
        return X, S

    def set_limiter_mlp(self, mlp):
        self._check(self.L.optional["set_limiter_mlp"](self.ctx, mlp.h), "set_limiter_mlp")

    def set_learned_interface_alpha(self, alpha):
        self._check(self.L.optional["set_learned_interface_alpha"](self.ctx, float(alpha)),
                    "set_learned_interface_alpha")

    def conservation(self):
        out = np.zeros(3)
        self._check(self.L.conservation(self.ctx, out.ctypes.data_as(_DP)), "conservation")
        return out


class Batch:
    """A batch of independent 2D problems (mhg_batch_*): problem(i) is a
    Simulation over the batch's context for problem i."""

    def __init__(self, library, meshes, options):
        self.L = library
        n = len(meshes)
        assert n == len(options) and n > 0
        self.meshes = (MeshDesc * n)(*meshes)
        self.opts = (Options * n)(*options)
        h = C.c_void_p()
        r = self.L.optional["batch_create"](n, self.meshes, self.opts, C.byref(h))
        if r != OK:
            raise MahycoError(r, "batch_create")
        self.h = h
        self.n = n
        self.problems = [Simulation(library, self.meshes[i], self.opts[i],
                                    ctx=C.c_void_p(self.L.optional["batch_problem"](h, i)))
                         for i in range(n)]

    def _check(self, r, what):
        if r < 0:
            msg = self.L.optional["batch_last_error"](self.h)
            raise MahycoError(r, f"{what}: {msg.decode() if msg else ''}")
        return r

    def problem(self, i):
        return self.problems[i]

    def run(self, max_steps=1 << 30):
        """Returns (problems still running, zone-cycles done in this call)."""
        nr = C.c_int32()
        zc = C.c_int64()
        self._check(self.L.optional["batch_run"](self.h, max_steps, C.byref(nr), C.byref(zc)), "batch_run")
        return nr.value, zc.value

    def status(self, i):
        it, st = C.c_int32(), C.c_int32()
        t, dt = C.c_double(), C.c_double()
        self._check(self.L.optional["batch_status"](self.h, i, C.byref(it), C.byref(t), C.byref(dt),
                                                    C.byref(st)), "batch_status")
        return {"iterations": it.value, "time": t.value, "dt": dt.value, "state": st.value}

    def autotune(self, reps=2):
        buf = C.create_string_buffer(8192)
        self._check(self.L.optional["batch_autotune"](self.h, reps, buf, 8192), "batch_autotune")
        return buf.value.decode()

    def close(self):
        if getattr(self, "h", None):
            for p in self.problems:
                p.ctx = None
            self.L.optional["batch_destroy"](self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _f32(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float32)
    return a if shape is None else a.reshape(shape)


class Mlp:
    """The learned limiter's network (mhg_mlp_*)."""

    def __init__(self, library, params=None):
        self.L = library
        self.f = library.optional
        h = C.c_void_p()
        r = self.f["mlp_create"](C.byref(h))
        if r != OK:
            raise MahycoError(r, "mlp_create")
        self.h = h
        self.nparam = int(self.f["mlp_param_count"]())
        if params is not None:
            self.set_params(params)

    def _check(self, r, what):
        if r < 0:
            msg = self.f["mlp_last_error"](self.h)
            raise MahycoError(r, f"{what}: {msg.decode() if msg else ''}")
        return r

    def set_params(self, p):
        p = _f32(p)
        assert p.size == self.nparam
        self._check(self.f["mlp_set_params"](self.h, p.ctypes.data_as(_FP)), "mlp_set_params")

    def params(self):
        p = np.zeros(self.nparam, dtype=np.float32)
        self._check(self.f["mlp_get_params"](self.h, p.ctypes.data_as(_FP)), "mlp_get_params")
        return p

    def forward(self, X):
        X = _f32(X)
        Y = np.zeros((X.shape[0], 2), dtype=np.float32)
        self._check(self.f["mlp_forward"](self.h, X.ctypes.data_as(_FP), X.shape[0], Y.ctypes.data_as(_FP)),
                    "mlp_forward")
        return Y

    def set_data(self, X, T):
        self._X, self._T = _f32(X), _f32(T)
        self._check(self.f["mlp_set_data"](self.h, self._X.ctypes.data_as(_FP), self._T.ctypes.data_as(_FP),
                                           self._X.shape[0]), "mlp_set_data")

    def gradient(self, row0, m):
        g = np.zeros(self.nparam, dtype=np.float32)
        loss = C.c_float()
        self._check(self.f["mlp_gradient"](self.h, row0, m, C.byref(loss), g.ctypes.data_as(_FP)), "mlp_gradient")
        return loss.value, g

    def train(self, epochs, batch, lr, seed=0):
        out = np.zeros(epochs, dtype=np.float32)
        self._check(self.f["mlp_train"](self.h, epochs, batch, lr, seed, out.ctypes.data_as(_FP)), "mlp_train")
        return out

    def close(self):
        if getattr(self, "h", None):
            self.f["mlp_destroy"](self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
