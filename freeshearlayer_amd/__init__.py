"""freeshearlayer_amd -- MI355X-native per-step update of the freeShearLayer
compressible LES solver.

The product is the native library ``lib/libfsl.so`` (hand-written HIP kernels
for gfx950 behind the C-ABI of ``include/fsl.h``, plus the ANSI-C host layer).
This module is a thin ctypes binding over that C-ABI used by the tests and by
``bench.py``.  There is no Python or CPU fallback: if the library is missing,
``load()`` raises.
"""
import ctypes as C
import os

import numpy as np

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(PKG_DIR, "lib")
LIB_PATH = os.path.join(LIB_DIR, "libfsl.so")
LAYER2_PATH = os.path.join(LIB_DIR, "libfsl_layer2.so")

FSL_OK = 0
# error codes of include/fsl.h
FSL_ERR_ARG, FSL_ERR_HIP, FSL_ERR_NOMEM, FSL_ERR_NONFINITE, FSL_ERR_IO, FSL_ERR_UNSUPPORTED = -1, -2, -3, -4, -5, -6
FSL_SGS_CONSTANT, FSL_SGS_DYNAMIC, FSL_SGS_LEARNED = 0, 1, 2
FSL_SGS_AVG_NONE, FSL_SGS_AVG_Z, FSL_SGS_AVG_XZ = 0, 1, 2
FSL_PREC_FP64, FSL_PREC_MIXED, FSL_PREC_FP64_FAST = 0, 1, 2
FSL_CFL_CONVECTIVE, FSL_CFL_ACOUSTIC = 0, 1
FSL_MLP_TANH, FSL_MLP_RELU = 0, 1
MLP_ACTIVATIONS = {"tanh": FSL_MLP_TANH, "relu": FSL_MLP_RELU}

_dp = C.POINTER(C.c_double)
_fp = C.POINTER(C.c_float)


class FslCase(C.Structure):
    """The 22 positional items of layer2.ini (include/fsl.h: fsl_case)."""
    _fields_ = [("LEN", C.c_uint), ("HIG", C.c_uint), ("DEP", C.c_uint),
                ("deltaX", C.c_double), ("deltaY", C.c_double), ("deltaZ", C.c_double),
                ("deltaT", C.c_double), ("Answer", C.c_uint), ("numstep", C.c_uint),
                ("mu_L", C.c_double), ("Pr_L", C.c_double), ("Pr_T", C.c_double),
                ("Cs", C.c_double), ("BL_HIG", C.c_uint), ("Ua", C.c_double), ("Va", C.c_double),
                ("Wa", C.c_double), ("Ns_min", C.c_int), ("Nst", C.c_int), ("f_step", C.c_uint),
                ("nStages", C.c_int), ("maxCoNum", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class FslConfig(C.Structure):
    _fields_ = [("sgs_model", C.c_int), ("precision", C.c_int), ("cfl_mode", C.c_int),
                ("cfl", C.c_double), ("z_quirk", C.c_int), ("fused", C.c_int),
                ("rho_in", C.c_double), ("p_in", C.c_double), ("p_out", C.c_double),
                ("device", C.c_int)]


class FslBlock(C.Structure):
    _fields_ = [("X", C.c_double), ("counter", C.c_int), ("k_min", C.c_uint), ("k_max", C.c_uint),
                ("Ud", C.c_double), ("Vd", C.c_double), ("Wd", C.c_double)]


class FslShearStats(C.Structure):
    _fields_ = [("delta_theta", C.c_double), ("delta_theta_out", C.c_double), ("delta_w", C.c_double),
                ("tke", C.c_double)]


class FslLearnedStatus(C.Structure):
    _fields_ = [("interval", C.c_int), ("tol", C.c_double), ("in_fallback", C.c_int),
                ("last_error", C.c_double), ("judged", C.c_longlong), ("fallbacks", C.c_longlong),
                ("learned_steps", C.c_longlong), ("dynamic_steps", C.c_longlong), ("scale", C.c_double),
                ("adapt", C.c_int), ("profile_tol", C.c_double), ("local_tol", C.c_double),
                ("last_profile_error", C.c_double), ("last_local_error", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


MLP_NPARAM = 4801


class FslShearCase(C.Structure):
    _fields_ = [("mach_c", C.c_double), ("ratio", C.c_double), ("delta_w0", C.c_double),
                ("reynolds", C.c_double), ("pert", C.c_double), ("rho0", C.c_double),
                ("p0", C.c_double), ("seed", C.c_ulonglong)]


_lib = None


def load():
    """Load lib/libfsl.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("freeshearlayer_amd: native library %s is missing; run "
                           "`python -m freeshearlayer_amd.build` (no fallback exists)" % LIB_PATH)
    L = C.CDLL(LIB_PATH)
    vp, i_, d_, u_, cc = C.c_void_p, C.c_int, C.c_double, C.c_uint, C.c_char_p
    ll = C.c_longlong
    sig = {
        "fsl_case_read_ini": (i_, [cc, C.POINTER(FslCase)]),
        "fsl_case_read_bin": (i_, [cc, C.POINTER(FslCase)]),
        "fsl_case_ini_to_bin": (i_, [cc, cc]),
        "fsl_case_validate": (i_, [C.POINTER(FslCase), C.c_char_p, C.c_size_t]),
        "fsl_random": (d_, [d_]),
        "fsl_minmod": (d_, [d_, d_]),
        "fsl_block_init": (None, [C.POINTER(FslBlock), d_]),
        "fsl_block_advance": (i_, [C.POINTER(FslBlock), C.POINTER(FslCase)]),
        "fsl_ic_reference": (None, [C.POINTER(FslCase), _dp]),
        "fsl_shear_defaults": (None, [C.POINTER(FslShearCase)]),
        "fsl_ic_shear_layer": (d_, [C.POINTER(FslCase), C.POINTER(FslShearCase), _dp, _dp]),
        "fsl_backup_write": (i_, [cc, C.POINTER(FslCase), _dp, d_, u_, i_]),
        "fsl_backup_read": (i_, [cc, C.POINTER(FslCase), _dp, _dp, C.POINTER(u_), i_, i_]),
        "fsl_config_defaults": (None, [C.POINTER(FslConfig)]),
        "fsl_create": (i_, [C.POINTER(FslCase), C.POINTER(FslConfig), C.POINTER(vp)]),
        "fsl_destroy": (None, [vp]),
        "fsl_last_error": (cc, [vp]),
        "fsl_ncell": (C.c_size_t, [vp]),
        "fsl_nface": (C.c_size_t, [vp, i_]),
        "fsl_device_bytes": (C.c_size_t, [vp]),
        "fsl_set_inflow": (i_, [vp, _dp]),
        "fsl_upload_state": (i_, [vp, i_, _dp]),
        "fsl_download_state": (i_, [vp, i_, _dp]),
        "fsl_download_face": (i_, [vp, i_, _dp]),
        "fsl_upload_face": (i_, [vp, i_, _dp]),
        "fsl_download_mu_sgs": (i_, [vp, _dp]),
        "fsl_upload_mu_sgs": (i_, [vp, _dp]),
        "fsl_set_dt": (i_, [vp, d_]),
        "fsl_get_dt": (d_, [vp]),
        "fsl_set_block": (i_, [vp, u_, u_, d_, d_, d_]),
        "fsl_get_cur": (i_, [vp]),
        "fsl_set_cur": (i_, [vp, i_]),
        "fsl_ghost_fill": (i_, [vp]),
        "fsl_reconstruct": (i_, [vp]),
        "fsl_iface_bc": (i_, [vp]),
        "fsl_fluxes": (i_, [vp]),
        "fsl_evolution": (i_, [vp, i_, i_]),
        "fsl_rk_update": (i_, [vp, i_, i_]),
        "fsl_check_co_num": (i_, [vp, _dp]),
        "fsl_dynamic_smagorinsky": (i_, [vp]),
        "fsl_set_sgs_average": (i_, [vp, i_]),
        "fsl_get_sgs_average": (i_, [vp]),
        "fsl_filter": (i_, [vp, _dp, _dp, _dp]),
        "fsl_stage": (i_, [vp, i_]),
        "fsl_step": (i_, [vp]),
        "fsl_run": (i_, [vp, i_]),
        "fsl_set_lcg": (i_, [vp, d_, i_]),
        "fsl_get_lcg": (i_, [vp, _dp, C.POINTER(i_), C.POINTER(u_), C.POINTER(u_), _dp, _dp, _dp]),
        "fsl_get_step": (u_, [vp]),
        "fsl_set_step": (i_, [vp, u_]),
        "fsl_get_block_state": (i_, [vp, C.POINTER(FslBlock)]),
        "fsl_set_block_state": (i_, [vp, C.POINTER(FslBlock)]),
        "fsl_get_case": (i_, [vp, C.POINTER(FslCase)]),
        "fsl_get_config": (i_, [vp, C.POINTER(FslConfig)]),
        "fsl_synchronize": (i_, [vp]),
        "fsl_conservation_sums": (i_, [vp, _dp]),
        "fsl_profile_enable": (i_, [vp, i_]),
        "fsl_profile_reset": (i_, [vp]),
        "fsl_profile_read": (i_, [vp, i_, _dp, C.POINTER(C.c_longlong)]),
        "fsl_kernel_name": (cc, [i_]),
        "fsl_kernel_count": (i_, []),
        "fsl_copy_kernel": (i_, [vp, C.c_size_t, i_]),
        "fsl_plane_profiles": (i_, [vp, _dp]),
        "fsl_shear_statistics": (i_, [C.POINTER(FslCase), _dp, d_, d_, d_, _dp, _dp,
                                      C.POINTER(FslShearStats)]),
        "fsl_mlp_init_params": (None, [C.c_ulonglong, _fp, _fp]),
        "fsl_mlp_file_write": (i_, [cc, _fp, _fp]),
        "fsl_mlp_file_read": (i_, [cc, _fp, _fp]),
        "fsl_mlp_file_write_act": (i_, [cc, _fp, _fp, i_]),
        "fsl_mlp_file_read_act": (i_, [cc, _fp, _fp, C.POINTER(i_)]),
        "fsl_mlp_set_activation": (i_, [vp, i_]),
        "fsl_mlp_get_activation": (i_, [vp]),
        "fsl_mlp_set_params": (i_, [vp, _fp, _fp]),
        "fsl_mlp_get_params": (i_, [vp, _fp, _fp]),
        "fsl_mlp_forward": (i_, [vp, _fp, ll, _fp]),
        "fsl_mlp_features": (i_, [vp, ll, ll, ll, _fp]),
        "fsl_mlp_update_thickness": (i_, [vp]),
        "fsl_mlp_get_thickness": (i_, [vp, _fp]),
        "fsl_mlp_set_thickness": (i_, [vp, _fp]),
        "fsl_learned_smagorinsky": (i_, [vp]),
        "fsl_mlp_reserve": (i_, [vp, ll]),
        "fsl_mlp_collect": (i_, [vp, ll, ll, C.POINTER(ll)]),
        "fsl_mlp_sample_count": (ll, [vp]),
        "fsl_mlp_clear_samples": (i_, [vp]),
        "fsl_mlp_samples": (i_, [vp, ll, ll, _fp, _fp]),
        "fsl_mlp_load_samples": (i_, [vp, _fp, _fp, ll]),
        "fsl_mlp_fit_norm": (i_, [vp, _fp]),
        "fsl_mlp_train": (i_, [vp, i_, ll, C.c_float, C.c_ulonglong, _fp]),
        "fsl_mlp_gradient": (i_, [vp, ll, _fp, _fp]),
        "fsl_mlp_set_loss_weight": (i_, [vp, C.c_float]),
        "fsl_learned_configure": (i_, [vp, i_, d_]),
        "fsl_learned_set_scale": (i_, [vp, d_, i_]),
        "fsl_learned_history": (i_, [vp, i_, _dp, _dp]),
        "fsl_learned_get_status": (i_, [vp, C.POINTER(FslLearnedStatus)]),
        "fsl_learned_judge": (i_, [vp, _dp, _dp, _dp]),
        "fsl_learned_judge_errors": (i_, [vp, _dp, _dp, _dp, _dp, _dp]),
        "fsl_learned_set_tolerances": (i_, [vp, d_, d_]),
        "fsl_learned_level_factors": (C.c_longlong, [vp, C.c_longlong, _fp]),
        "fsl_learned_history_errors": (i_, [vp, i_, _dp, _dp, _dp, _dp]),
        "fsl_mfma_probe": (i_, [vp, _fp, _fp, _fp]),
        "fsl_output_fields": (i_, [vp, i_, i_, _dp]),
        "fsl_conservation_sums_noalloc": (i_, [vp, u_, u_, u_, vp, vp]),
        "fsl_write_plt": (i_, [vp, cc]),
        "fsl_plt_write_fields": (i_, [cc, C.POINTER(FslCase), _dp]),
        "fsl_probe_values": (i_, [vp, _dp]),
        "fsl_write_probes": (i_, [vp, cc, i_]),
        "fsl_probes_append": (i_, [cc, u_, u_, _dp, i_]),
        "fsl_favre_average": (i_, [cc, u_, u_, u_, i_, _dp, C.POINTER(u_)]),
        "fsl_favre_write_profiles": (i_, [cc, u_, _dp]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def exported_symbols():
    """Names of the C-ABI functions this binding expects libfsl.so to export."""
    import re
    hdr = os.path.join(os.path.dirname(PKG_DIR), "include", "fsl.h")
    txt = open(hdr).read()
    return sorted(set(re.findall(r"\b(fsl_[a-z0-9_]+)\s*\(", txt)))


def _ptr(a):
    if a.dtype != np.float64 or not a.flags.c_contiguous:
        raise TypeError("expected a C-contiguous float64 array")
    return a.ctypes.data_as(_dp)


def _fptr(a):
    if a.dtype != np.float32 or not a.flags.c_contiguous:
        raise TypeError("expected a C-contiguous float32 array")
    return a.ctypes.data_as(_fp)


def _f32(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float32)
    return a.reshape(shape) if shape is not None else a


def case_reference(L, H, D, **over):
    """Case items of the reference layer2.ini with the mesh replaced."""
    c = FslCase(LEN=L, HIG=H, DEP=D, deltaX=1.6666667, deltaY=1.0, deltaZ=1.5, deltaT=0.0007,
                Answer=0, numstep=3000, mu_L=0.253358, Pr_L=0.72, Pr_T=0.8, Cs=0.01, BL_HIG=5,
                Ua=35.0, Va=35.0, Wa=35.0, Ns_min=10, Nst=50, f_step=50, nStages=3, maxCoNum=0.1)
    for k, v in over.items():
        setattr(c, k, v)
    return c


def config(**over):
    cfg = FslConfig()
    load().fsl_config_defaults(C.byref(cfg))
    for k, v in over.items():
        setattr(cfg, k, v)
    return cfg


class FslError(RuntimeError):
    pass


class Context:
    """Device context over the C-ABI (one HIP stream, device-resident state)."""

    def __init__(self, case, cfg=None):
        self.lib = load()
        self.case = case
        self.cfg = cfg if cfg is not None else config()
        h = C.c_void_p()
        rc = self.lib.fsl_create(C.byref(self.case), C.byref(self.cfg), C.byref(h))
        if rc != FSL_OK:
            raise FslError("fsl_create failed (%d)" % rc)
        self.h = h
        self.L, self.H, self.D = case.LEN, case.HIG, case.DEP
        self.cell_shape = (self.L + 2, self.H + 2, self.D + 2)

    def close(self):
        if getattr(self, "h", None):
            self.lib.fsl_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _chk(self, rc, what):
        if rc != FSL_OK:
            raise FslError("%s failed (%d): %s" % (what, rc, self.lib.fsl_last_error(self.h).decode()))
        return rc

    # --- transfers ---
    def upload_state(self, U, which=0):
        U = np.ascontiguousarray(U, dtype=np.float64).reshape((5,) + self.cell_shape)
        self._chk(self.lib.fsl_upload_state(self.h, which, _ptr(U)), "upload_state")

    def download_state(self, which=0):
        U = np.zeros((5,) + self.cell_shape)
        self._chk(self.lib.fsl_download_state(self.h, which, _ptr(U)), "download_state")
        return U

    def face_shape(self, fid):
        d = fid // 10
        return (self.L + (d == 0), self.H + (d == 1), self.D + (d == 2))

    def download_face(self, fid):
        out = np.zeros(self.face_shape(fid))
        self._chk(self.lib.fsl_download_face(self.h, fid, _ptr(out)), "download_face")
        return out

    def upload_face(self, fid, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        self._chk(self.lib.fsl_upload_face(self.h, fid, _ptr(a)), "upload_face")

    def download_mu(self):
        out = np.zeros(self.cell_shape)
        self._chk(self.lib.fsl_download_mu_sgs(self.h, _ptr(out)), "download_mu_sgs")
        return out

    def upload_mu(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        self._chk(self.lib.fsl_upload_mu_sgs(self.h, _ptr(a)), "upload_mu_sgs")

    def set_inflow(self, u_in):
        u = np.ascontiguousarray(u_in, dtype=np.float64)
        self._chk(self.lib.fsl_set_inflow(self.h, _ptr(u)), "set_inflow")

    # --- scalars ---
    def set_dt(self, dt):
        self._chk(self.lib.fsl_set_dt(self.h, dt), "set_dt")

    @property
    def dt(self):
        return self.lib.fsl_get_dt(self.h)

    def set_block(self, kmin, kmax, ud, vd, wd):
        self._chk(self.lib.fsl_set_block(self.h, kmin, kmax, ud, vd, wd), "set_block")

    def set_lcg(self, X, counter=0):
        self._chk(self.lib.fsl_set_lcg(self.h, X, counter), "set_lcg")

    def lcg(self):
        X = C.c_double(); cnt = C.c_int(); kmin = C.c_uint(); kmax = C.c_uint()
        ud = C.c_double(); vd = C.c_double(); wd = C.c_double()
        self.lib.fsl_get_lcg(self.h, C.byref(X), C.byref(cnt), C.byref(kmin), C.byref(kmax),
                             C.byref(ud), C.byref(vd), C.byref(wd))
        return dict(X=X.value, counter=cnt.value, k_min=kmin.value, k_max=kmax.value,
                    Ud=ud.value, Vd=vd.value, Wd=wd.value)

    @property
    def cur_is_p(self):
        return self.lib.fsl_get_cur(self.h)

    def set_cur(self, is_p):
        self._chk(self.lib.fsl_set_cur(self.h, int(is_p)), "set_cur")

    @property
    def step_count(self):
        return self.lib.fsl_get_step(self.h)

    # --- operations (reference entry points) ---
    def ghost_fill(self): self._chk(self.lib.fsl_ghost_fill(self.h), "ghost_fill")
    def reconstruct(self): self._chk(self.lib.fsl_reconstruct(self.h), "reconstruct")
    def iface_bc(self): self._chk(self.lib.fsl_iface_bc(self.h), "iface_bc")
    def fluxes(self): self._chk(self.lib.fsl_fluxes(self.h), "fluxes")
    def evolution(self, n, s): self._chk(self.lib.fsl_evolution(self.h, n, s), "evolution")
    def rk_update(self, scheme, s): self._chk(self.lib.fsl_rk_update(self.h, scheme, s), "rk_update")
    def dynamic_smagorinsky(self): self._chk(self.lib.fsl_dynamic_smagorinsky(self.h), "dynamic")

    def set_sgs_average(self, mode):
        """Averaging of the dynamic coefficient: FSL_SGS_AVG_NONE (pointwise, the
        reference's), FSL_SGS_AVG_Z (z lines) or FSL_SGS_AVG_XZ (x-z planes)."""
        self._chk(self.lib.fsl_set_sgs_average(self.h, int(mode)), "set_sgs_average")

    def sgs_average(self): return self.lib.fsl_get_sgs_average(self.h)
    def stage(self, s): self._chk(self.lib.fsl_stage(self.h, s), "stage")
    def step(self): self._chk(self.lib.fsl_step(self.h), "step")
    def run(self, n): self._chk(self.lib.fsl_run(self.h, n), "run")
    def synchronize(self): self._chk(self.lib.fsl_synchronize(self.h), "synchronize")

    def check_co_num(self):
        m = C.c_double()
        self._chk(self.lib.fsl_check_co_num(self.h, C.byref(m)), "check_co_num")
        return m.value

    def filter(self, a, h=(1. / 3., 1. / 3., 1. / 3.)):
        a = np.ascontiguousarray(a, dtype=np.float64).reshape(self.cell_shape)
        out = np.zeros(self.cell_shape)
        hh = np.array(h, dtype=np.float64)
        self._chk(self.lib.fsl_filter(self.h, _ptr(a), _ptr(out), _ptr(hh)), "filter")
        return out

    def conservation_sums(self):
        out = np.zeros(5)
        self._chk(self.lib.fsl_conservation_sums(self.h, _ptr(out)), "sums")
        return out

    # --- measurement ---
    def profile(self, on=True):
        self.lib.fsl_profile_enable(self.h, int(on))

    def profile_reset(self):
        self.lib.fsl_profile_reset(self.h)

    def profile_read(self):
        n = self.lib.fsl_kernel_count()
        ms = np.zeros(n)
        cnt = (C.c_longlong * n)()
        self.lib.fsl_profile_read(self.h, n, _ptr(ms), cnt)
        return {self.lib.fsl_kernel_name(i).decode(): (ms[i], cnt[i]) for i in range(n) if cnt[i]}

    def copy_kernel(self, n, reps):
        self._chk(self.lib.fsl_copy_kernel(self.h, n, reps), "copy_kernel")

    @property
    def device_bytes(self):
        return self.lib.fsl_device_bytes(self.h)

    # --- shear-layer statistics ---
    def plane_profiles(self):
        c = self.case
        out = np.zeros((7, c.LEN, c.HIG))
        self._chk(self.lib.fsl_plane_profiles(self.h, _ptr(out)), "plane_profiles")
        return out

    # --- learned closure ---
    def mlp_set_params(self, params, norm):
        p, n = _f32(params), _f32(norm)
        if p.size != MLP_NPARAM or n.size != 16:
            raise ValueError("expected %d parameters and 16 normalisation values" % MLP_NPARAM)
        self._chk(self.lib.fsl_mlp_set_params(self.h, _fptr(p), _fptr(n)), "mlp_set_params")

    def mlp_set_activation(self, activation):
        """Hidden activation of the network: FSL_MLP_TANH / FSL_MLP_RELU or 'tanh' / 'relu'."""
        act = MLP_ACTIVATIONS.get(activation, activation)
        self._chk(self.lib.fsl_mlp_set_activation(self.h, int(act)), "mlp_set_activation")

    def mlp_activation(self):
        return self.lib.fsl_mlp_get_activation(self.h)

    def mlp_load(self, path):
        """Parameters, normalisation and activation from a model file."""
        p, n, act = mlp_file_read_act(path)
        self.mlp_set_params(p, n)
        self.mlp_set_activation(act)

    def mlp_get_params(self):
        p = np.zeros(MLP_NPARAM, np.float32)
        n = np.zeros(16, np.float32)
        self._chk(self.lib.fsl_mlp_get_params(self.h, _fptr(p), _fptr(n)), "mlp_get_params")
        return p, n

    def mlp_forward(self, X):
        X = _f32(X)
        n = X.shape[0]
        y = np.zeros(n, np.float32)
        self._chk(self.lib.fsl_mlp_forward(self.h, _fptr(X.reshape(-1)), n, _fptr(y)), "mlp_forward")
        return y

    def mlp_features(self, stride=1, offset=0, n=None):
        N = self.case.LEN * self.case.HIG * self.case.DEP
        if n is None:
            n = (N - offset + stride - 1) // stride
        X = np.zeros((n, 8), np.float32)
        self._chk(self.lib.fsl_mlp_features(self.h, stride, offset, n, _fptr(X.reshape(-1))), "mlp_features")
        return X

    def mlp_update_thickness(self):
        self._chk(self.lib.fsl_mlp_update_thickness(self.h), "mlp_update_thickness")

    def mlp_thickness(self):
        dw = np.zeros(self.case.LEN + 2, np.float32)
        self._chk(self.lib.fsl_mlp_get_thickness(self.h, _fptr(dw)), "mlp_get_thickness")
        return dw

    def mlp_set_thickness(self, dw):
        dw = _f32(dw)
        self._chk(self.lib.fsl_mlp_set_thickness(self.h, _fptr(dw)), "mlp_set_thickness")

    def learned_smagorinsky(self):
        self._chk(self.lib.fsl_learned_smagorinsky(self.h), "learned_smagorinsky")

    def mlp_reserve(self, cap):
        self._chk(self.lib.fsl_mlp_reserve(self.h, cap), "mlp_reserve")

    def mlp_collect(self, stride, offset=0):
        n = C.c_longlong(0)
        self._chk(self.lib.fsl_mlp_collect(self.h, stride, offset, C.byref(n)), "mlp_collect")
        return n.value

    def mlp_sample_count(self):
        return self.lib.fsl_mlp_sample_count(self.h)

    def mlp_clear_samples(self):
        self._chk(self.lib.fsl_mlp_clear_samples(self.h), "mlp_clear_samples")

    def mlp_samples(self, first=0, n=None):
        if n is None:
            n = self.mlp_sample_count() - first
        X = np.zeros((n, 8), np.float32)
        t = np.zeros(n, np.float32)
        self._chk(self.lib.fsl_mlp_samples(self.h, first, n, _fptr(X.reshape(-1)), _fptr(t)), "mlp_samples")
        return X, t

    def mlp_load_samples(self, X, t):
        X = _f32(X)
        t = _f32(t)
        self._chk(self.lib.fsl_mlp_load_samples(self.h, _fptr(X.reshape(-1)), _fptr(t), t.size),
                  "mlp_load_samples")

    def mlp_fit_norm(self):
        n = np.zeros(16, np.float32)
        self._chk(self.lib.fsl_mlp_fit_norm(self.h, _fptr(n)), "mlp_fit_norm")
        return n

    def mlp_train(self, epochs, batch, lr, seed=0):
        loss = np.zeros(max(epochs, 1), np.float32)
        self._chk(self.lib.fsl_mlp_train(self.h, epochs, batch, lr, seed, _fptr(loss)), "mlp_train")
        return loss[:epochs]

    def mlp_set_loss_weight(self, power):
        self._chk(self.lib.fsl_mlp_set_loss_weight(self.h, power), "mlp_set_loss_weight")

    def mlp_gradient(self, batch):
        g = np.zeros(MLP_NPARAM, np.float32)
        loss = np.zeros(1, np.float32)
        self._chk(self.lib.fsl_mlp_gradient(self.h, batch, _fptr(g), _fptr(loss)), "mlp_gradient")
        return g, float(loss[0])

    def learned_configure(self, interval=50, tol=0.10):
        self._chk(self.lib.fsl_learned_configure(self.h, interval, tol), "learned_configure")

    def learned_set_scale(self, scale=1.0, adapt=2):
        """adapt: 0 / False keeps the scale, 1 / True one global level set by
        each judge, 2 a level per y row, 3 per x band and y row."""
        self._chk(self.lib.fsl_learned_set_scale(self.h, scale, int(adapt)), "learned_set_scale")

    def learned_level_factors(self):
        """Level factors f(i, j) of the network's scale, shape (LEN, HIG) (all 1 below adapt 2)."""
        n = self.lib.fsl_learned_level_factors(self.h, 0, None)
        out = np.zeros(n, np.float32)
        self.lib.fsl_learned_level_factors(self.h, n, _fptr(out))
        return out.reshape(self.case.LEN, self.case.HIG)

    def learned_history(self, errors=False):
        """Judge history: (errors, scales), or with errors=True (errors,
        profile errors, local errors, scales)."""
        n = self.lib.fsl_learned_history(self.h, 0, None, None)
        e, pe, le, sc = (np.zeros(max(n, 1)) for _ in range(4))
        self.lib.fsl_learned_history_errors(self.h, n, _ptr(e), _ptr(pe), _ptr(le), _ptr(sc))
        return (e[:n], pe[:n], le[:n], sc[:n]) if errors else (e[:n], sc[:n])

    def learned_set_tolerances(self, profile_tol, local_tol):
        self._chk(self.lib.fsl_learned_set_tolerances(self.h, profile_tol, local_tol), "learned_set_tolerances")

    def learned_status(self):
        s = FslLearnedStatus()
        self._chk(self.lib.fsl_learned_get_status(self.h, C.byref(s)), "learned_status")
        return s.as_dict()

    def learned_judge(self, errors=False):
        """One judge on the step state: (error, D_dynamic, D_learned), or with
        errors=True (error, profile error, local error, D_dynamic, D_learned)."""
        e, pe, le, dd, dl = (C.c_double() for _ in range(5))
        self._chk(self.lib.fsl_learned_judge_errors(self.h, C.byref(e), C.byref(pe), C.byref(le), C.byref(dd),
                                                    C.byref(dl)), "learned_judge")
        if errors:
            return e.value, pe.value, le.value, dd.value, dl.value
        return e.value, dd.value, dl.value

    def output_fields(self, k0=1, nk=None):
        """Output()'s nine computed columns (rho, u, v, w, p, T, |omega|, Q,
        mu_T/mu_L) of planes k0 .. k0+nk-1: array (nk, HIG, LEN, 9)."""
        case = self.case
        if nk is None:
            nk = case.DEP - k0 + 1
        out = np.zeros((nk, case.HIG, case.LEN, 9))
        self._chk(self.lib.fsl_output_fields(self.h, k0, nk, _ptr(out)), "output_fields")
        return out

    def write_plt(self, path=None):
        self._chk(self.lib.fsl_write_plt(self.h, path.encode() if path else None), "write_plt")

    def probe_values(self):
        out = np.zeros((6, self.case.HIG))
        self._chk(self.lib.fsl_probe_values(self.h, _ptr(out)), "probe_values")
        return out

    def write_probes(self, path=None, real_bytes=8):
        self._chk(self.lib.fsl_write_probes(self.h, path.encode() if path else None, real_bytes), "write_probes")

    def mfma_probe(self, A, B):
        A = _f32(A, (16, 16))
        B = _f32(B, (16, 16))
        D = np.zeros((16, 16), np.float32)
        self._chk(self.lib.fsl_mfma_probe(self.h, _fptr(A), _fptr(B), _fptr(D)), "mfma_probe")
        return D


def plt_write_fields(path, case, fields):
    """The Tecplot frame of Output() from host fields (LEN HIG DEP cells x 9)."""
    fields = np.ascontiguousarray(fields, dtype=np.float64)
    assert fields.size == 9 * case.LEN * case.HIG * case.DEP
    rc = load().fsl_plt_write_fields(path.encode(), C.byref(case), _ptr(fields))
    if rc:
        raise FslError("fsl_plt_write_fields (%d)" % rc)


def probes_append(path, step, vals, real_bytes=8):
    vals = np.ascontiguousarray(vals, dtype=np.float64)
    H = vals.size // 6
    rc = load().fsl_probes_append(path.encode(), step, H, _ptr(vals), real_bytes)
    if rc:
        raise FslError("fsl_probes_append (%d)" % rc)


def favre_average(path, HIG, step_start, step_end, real_bytes=4):
    """Favre-averaged u~ of the three probe stations: (array (3, HIG), records used)."""
    ru = np.zeros((3, HIG))
    n = C.c_uint()
    rc = load().fsl_favre_average(path.encode(), HIG, step_start, step_end, real_bytes, _ptr(ru), C.byref(n))
    if rc:
        raise FslError("fsl_favre_average (%d)" % rc)
    return ru, n.value


def favre_write_profiles(dirname, ru):
    ru = np.ascontiguousarray(ru, dtype=np.float64)
    rc = load().fsl_favre_write_profiles(dirname.encode(), ru.shape[1], _ptr(ru))
    if rc:
        raise FslError("fsl_favre_write_profiles (%d)" % rc)


def shear_statistics(case, prof, U1, U2, rho0=1.0):
    """delta_theta(x), delta_w(x) and integrated resolved TKE from plane profiles."""
    prof = np.ascontiguousarray(prof, dtype=np.float64)
    th = np.zeros(case.LEN)
    dw = np.zeros(case.LEN)
    st = FslShearStats()
    rc = load().fsl_shear_statistics(C.byref(case), _ptr(prof.reshape(-1)), U1, U2, rho0, _ptr(th), _ptr(dw),
                                     C.byref(st))
    if rc != 0:
        raise FslError("fsl_shear_statistics returned %d" % rc)
    return {"delta_theta": st.delta_theta, "delta_theta_out": st.delta_theta_out, "delta_w": st.delta_w,
            "tke": st.tke, "delta_theta_x": th, "delta_w_x": dw}


def mlp_init_params(seed=0):
    p = np.zeros(MLP_NPARAM, np.float32)
    n = np.zeros(16, np.float32)
    load().fsl_mlp_init_params(seed, _fptr(p), _fptr(n))
    return p, n


def mlp_file_write(path, params, norm):
    p, n = _f32(params), _f32(norm)
    rc = load().fsl_mlp_file_write(path.encode(), _fptr(p), _fptr(n))
    if rc != 0:
        raise FslError("fsl_mlp_file_write(%s) returned %d" % (path, rc))


def mlp_file_read(path):
    p = np.zeros(MLP_NPARAM, np.float32)
    n = np.zeros(16, np.float32)
    rc = load().fsl_mlp_file_read(path.encode(), _fptr(p), _fptr(n))
    if rc != 0:
        raise FslError("fsl_mlp_file_read(%s) returned %d" % (path, rc))
    return p, n


def mlp_file_write_act(path, params, norm, activation):
    p, n = _f32(params), _f32(norm)
    act = MLP_ACTIVATIONS.get(activation, activation)
    rc = load().fsl_mlp_file_write_act(path.encode(), _fptr(p), _fptr(n), int(act))
    if rc != 0:
        raise FslError("fsl_mlp_file_write_act(%s) returned %d" % (path, rc))


def mlp_file_read_act(path):
    """(params, norm, activation) of a model file of either format."""
    p = np.zeros(MLP_NPARAM, np.float32)
    n = np.zeros(16, np.float32)
    act = C.c_int(0)
    rc = load().fsl_mlp_file_read_act(path.encode(), _fptr(p), _fptr(n), C.byref(act))
    if rc != 0:
        raise FslError("fsl_mlp_file_read_act(%s) returned %d" % (path, rc))
    return p, n, act.value


def ic_reference(case):
    U = np.zeros((5, case.LEN + 2, case.HIG + 2, case.DEP + 2))
    load().fsl_ic_reference(C.byref(case), _ptr(U))
    return U


def shear_defaults():
    s = FslShearCase()
    load().fsl_shear_defaults(C.byref(s))
    return s


def ic_shear_layer(case, shear=None):
    """Synthetic shear-layer state and inflow profile; returns (U, u_in, mu_L)."""
    s = shear if shear is not None else shear_defaults()
    U = np.zeros((5, case.LEN + 2, case.HIG + 2, case.DEP + 2))
    u_in = np.zeros(case.HIG)
    mu = load().fsl_ic_shear_layer(C.byref(case), C.byref(s), _ptr(U), _ptr(u_in))
    return U, u_in, mu
