"""ctypes wrappers over the CPU checkers (test infrastructure only).

* ``Oracle``  -- the CPU restatement in oracle/fsl_oracle.c (oracle/liboracle.so)
* ``RefLib``  -- the compiled reference src/ path (oracle/_ref/libfslref_*.so),
                present only where oracle/Makefile could build it.

Both use the dense layouts of oracle/fsl_oracle.h: cell arrays (L+2, H+2, D+2)
and face arrays x (L+1, H, D), y (L, H+1, D), z (L, H, D+1), k fastest.
"""
import ctypes as C
import os
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")
REF_DIR = os.path.join(ROOT, "oracle", "_ref")

_dp = C.POINTER(C.c_double)


def cell_shape(L, H, D):
    return (L + 2, H + 2, D + 2)


def face_shapes(L, H, D):
    return [(L + 1, H, D), (L, H + 1, D), (L, H, D + 1)]


def face_shape(fid, L, H, D):
    return face_shapes(L, H, D)[fid // 10]


def _ptr(a):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_dp)


class Oracle:
    """Handle on one or_sim instance (reference defaults of layer2.ini)."""

    _lib = None

    @classmethod
    def lib(cls):
        if cls._lib is None:
            if not os.path.exists(ORACLE_SO):
                raise RuntimeError("oracle/liboracle.so is not built (run __graft_entry__.build())")
            lib = C.CDLL(ORACLE_SO)
            lib.or_sim_create.restype = C.c_void_p
            lib.or_sim_create.argtypes = [C.c_int] * 4
            lib.or_sim_destroy.argtypes = [C.c_void_p]
            lib.or_sim_set.argtypes = [C.c_void_p, C.c_char_p, C.c_double]
            lib.or_sim_set.restype = C.c_int
            lib.or_sim_get.argtypes = [C.c_void_p, C.c_char_p]
            lib.or_sim_get.restype = C.c_double
            for n in ("cell", "face"):
                f = getattr(lib, "or_sim_" + n)
                f.argtypes = [C.c_void_p, C.c_int]
                f.restype = _dp
            lib.or_sim_u_in.argtypes = [C.c_void_p]
            lib.or_sim_u_in.restype = _dp
            for n in ("derive", "ghost", "recon", "iface", "fluxes", "block_draw", "dynamic",
                      "step", "ic_reference"):
                getattr(lib, "or_sim_" + n).argtypes = [C.c_void_p]
            lib.or_sim_evolution.argtypes = [C.c_void_p, C.c_int, C.c_int]
            lib.or_sim_check_co_num.argtypes = [C.c_void_p]
            lib.or_sim_check_co_num.restype = C.c_double
            lib.or_sim_filter.argtypes = [C.c_void_p, _dp, _dp]
            lib.or_sim_output_fields.argtypes = [C.c_void_p, _dp]
            lib.or_sim_dynamic_terms.argtypes = [C.c_void_p, _dp, _dp, _dp]
            lib.or_sim_probe_values.argtypes = [C.c_void_p, _dp]
            lib.or_sim_cur_is_p.argtypes = [C.c_void_p]
            lib.or_sim_cur_is_p.restype = C.c_int
            lib.or_sim_set_cur.argtypes = [C.c_void_p, C.c_int]
            lib.or_random.argtypes = [C.c_double]
            lib.or_random.restype = C.c_double
            lib.or_minmod.argtypes = [C.c_double, C.c_double]
            lib.or_minmod.restype = C.c_double
            cls._lib = lib
        return cls._lib

    def __init__(self, L, H, D, dynamic=1):
        self.L, self.H, self.D = L, H, D
        self.h = self.lib().or_sim_create(L, H, D, int(dynamic))
        if not self.h:
            raise MemoryError("or_sim_create failed")

    def close(self):
        if self.h:
            self.lib().or_sim_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set(self, key, value):
        if self.lib().or_sim_set(self.h, key.encode(), float(value)) != 0:
            raise KeyError(key)

    def get(self, key):
        return self.lib().or_sim_get(self.h, key.encode())

    def derive(self):
        self.lib().or_sim_derive(self.h)

    def cell(self, cid):
        """numpy view of cell array cid (0..4 U, 5..9 Up, 10 mu_SGS)."""
        p = self.lib().or_sim_cell(self.h, cid)
        n = int(np.prod(cell_shape(self.L, self.H, self.D)))
        return np.ctypeslib.as_array(p, shape=(n,)).reshape(cell_shape(self.L, self.H, self.D))

    def face(self, fid):
        p = self.lib().or_sim_face(self.h, fid)
        shp = face_shape(fid, self.L, self.H, self.D)
        return np.ctypeslib.as_array(p, shape=(int(np.prod(shp)),)).reshape(shp)

    def u_in(self):
        return np.ctypeslib.as_array(self.lib().or_sim_u_in(self.h), shape=(self.H,))

    def state(self):
        return np.stack([self.cell(f).copy() for f in range(5)])

    def set_state(self, U):
        for f in range(5):
            self.cell(f)[...] = U[f]

    def ghost(self): self.lib().or_sim_ghost(self.h)
    def recon(self): self.lib().or_sim_recon(self.h)
    def iface(self): self.lib().or_sim_iface(self.h)
    def fluxes(self): self.lib().or_sim_fluxes(self.h)
    def evolution(self, n, s): self.lib().or_sim_evolution(self.h, n, s)
    def check_co_num(self): return self.lib().or_sim_check_co_num(self.h)
    def block_draw(self): self.lib().or_sim_block_draw(self.h)
    def dynamic(self): self.lib().or_sim_dynamic(self.h)

    def dynamic_terms(self):
        """The dynamic model's L:M, M:M and rho DD |S| per cell (interior set)."""
        shp = cell_shape(self.L, self.H, self.D)
        lm, mm, fac = (np.zeros(shp) for _ in range(3))
        self.lib().or_sim_dynamic_terms(self.h, _ptr(lm), _ptr(mm), _ptr(fac))
        return lm, mm, fac

    def step(self): self.lib().or_sim_step(self.h)
    def ic_reference(self): self.lib().or_sim_ic_reference(self.h)
    def cur_is_p(self): return self.lib().or_sim_cur_is_p(self.h)
    def set_cur(self, is_p): self.lib().or_sim_set_cur(self.h, int(is_p))

    def output_fields(self):
        """Output()'s nine computed columns per interior cell: (D, H, L, 9)."""
        out = np.zeros((self.D, self.H, self.L, 9))
        self.lib().or_sim_output_fields(self.h, _ptr(out))
        return out

    def probe_values(self):
        out = np.zeros((6, self.H))
        self.lib().or_sim_probe_values(self.h, _ptr(out))
        return out

    def filter(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        out = np.zeros_like(a)
        self.lib().or_sim_filter(self.h, _ptr(a), _ptr(out))
        return out


INI_TEMPLATE = """{L}          LEN
{H}          HIG
{D}          DEP
{dx}  deltaX
{dy}         deltaY
{dz}        deltaZ
{dt}     deltaT
{answer}         Answer
{numstep}       numstep
{mu_L}   mu_L
{Pr_L}       Pr_L
{Pr_T}        Pr_T
{Cs}       Cs
{BL_HIG}          BL_HIG
{Ua}        Ua
{Va}        Va
{Wa}        Wa
{Ns_min}         Nst_min
{Nst}         Nst
{f_step}    f_step
{nStages}          nStages
{maxCoNum}        maxCoNum
"""

REF_INI_DEFAULTS = dict(dx="1.6666667", dy="1.", dz="1.5", dt="0.0007", answer=0, numstep=3000,
                        mu_L="0.253358", Pr_L="0.72", Pr_T="0.8", Cs="0.01", BL_HIG=5, Ua="35.",
                        Va="35.", Wa="35.", Ns_min=10, Nst=50, f_step=100000, nStages=3,
                        maxCoNum="0.1")


def write_ini(path, L, H, D, **over):
    d = dict(REF_INI_DEFAULTS)
    d.update(over)
    with open(path, "w") as f:
        f.write(INI_TEMPLATE.format(L=L, H=H, D=D, **d))


def ref_available(variant="dyn"):
    return os.path.exists(os.path.join(REF_DIR, "libfslref_%s.so" % variant))


class RefLib:
    """The compiled reference.  One instance per variant ('dyn' or 'cs'); the
    reference keeps its state in process globals, so each variant's library is
    loaded once and re-initialised by ``initialize``."""

    _libs = {}

    def __init__(self, variant="dyn"):
        if variant not in self._libs:
            path = os.path.join(REF_DIR, "libfslref_%s.so" % variant)
            if not os.path.exists(path):
                raise RuntimeError("compiled reference not built: " + path)
            lib = C.CDLL(path)
            lib.fslref_initialize.argtypes = [C.c_char_p]
            lib.fslref_initialize.restype = C.c_int
            for n in ("get_cell", "set_cell", "get_face", "set_face"):
                getattr(lib, "fslref_" + n).argtypes = [C.c_int, _dp]
                getattr(lib, "fslref_" + n).restype = C.c_int
            lib.fslref_get_scalars.argtypes = [_dp]
            lib.fslref_set_time.argtypes = [C.c_double, C.c_double, C.c_int]
            lib.fslref_set_block.argtypes = [C.c_uint, C.c_uint, C.c_double, C.c_double, C.c_double]
            lib.fslref_set_nstages.argtypes = [C.c_int]
            lib.fslref_evolution.argtypes = [C.c_int, C.c_int]
            lib.fslref_random.argtypes = [C.c_double]
            lib.fslref_random.restype = C.c_double
            lib.fslref_minmod.argtypes = [C.c_double, C.c_double]
            lib.fslref_minmod.restype = C.c_double
            lib.fslref_filter.argtypes = [_dp, _dp]
            lib.fslref_set_cur.argtypes = [C.c_int]
            lib.fslref_cur_is_p.restype = C.c_int
            self._libs[variant] = lib
        self.lib = self._libs[variant]
        self.variant = variant

    def initialize(self, L, H, D, **ini):
        self.L, self.H, self.D = L, H, D
        with tempfile.TemporaryDirectory() as d:
            write_ini(os.path.join(d, "layer2.ini"), L, H, D, **ini)
            rc = self.lib.fslref_initialize(d.encode())
        if rc != 0:
            raise RuntimeError("fslref_initialize failed: %d" % rc)

    def cell(self, cid):
        out = np.zeros(cell_shape(self.L, self.H, self.D))
        self.lib.fslref_get_cell(cid, _ptr(out))
        return out

    def set_cell(self, cid, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        self.lib.fslref_set_cell(cid, _ptr(a))

    def face(self, fid):
        out = np.zeros(face_shape(fid, self.L, self.H, self.D))
        self.lib.fslref_get_face(fid, _ptr(out))
        return out

    def set_face(self, fid, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        self.lib.fslref_set_face(fid, _ptr(a))

    def state(self):
        return np.stack([self.cell(f) for f in range(5)])

    def set_state(self, U):
        for f in range(5):
            self.set_cell(f, U[f])

    def scalars(self):
        o = np.zeros(22)
        self.lib.fslref_get_scalars(_ptr(o))
        keys = ["deltaT", "deltaT_X", "deltaT_Y", "deltaT_Z", "k_min", "k_max", "Ud", "Vd", "Wd",
                "X", "DD", "CsDD", "lambda_L", "cp_Pr_T", "counter", "step", "nStages",
                "maxCoNum", "_2deltaX", "_4deltaX", "_deltaX", "mu_L"]
        return dict(zip(keys, o))

    def set_time(self, dt, X, counter):
        self.lib.fslref_set_time(dt, X, counter)

    def set_block(self, kmin, kmax, ud, vd, wd):
        self.lib.fslref_set_block(kmin, kmax, ud, vd, wd)

    def ghost(self): self.lib.fslref_ghost()
    def recon(self): self.lib.fslref_recon()
    def iface(self): self.lib.fslref_iface()
    def fluxes(self): self.lib.fslref_fluxes()
    def evolution(self, n, s): self.lib.fslref_evolution(n, s)
    def check_co_num(self): self.lib.fslref_check_co_num()
    def block_draw(self): self.lib.fslref_block_draw()
    def dynamic(self): self.lib.fslref_dynamic()
    def step(self): self.lib.fslref_step()
    def cur_is_p(self): return self.lib.fslref_cur_is_p()
    def set_cur(self, is_p): self.lib.fslref_set_cur(int(is_p))

    def filter(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        out = np.zeros_like(a)
        self.lib.fslref_filter(_ptr(a), _ptr(out))
        return out

    def random(self, x):
        return self.lib.fslref_random(x)

    def minmod(self, x, y):
        return self.lib.fslref_minmod(x, y)
