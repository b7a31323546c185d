"""The allocation-free conservation sums (fsl_conservation_sums_noalloc)
through the C-ABI, against the context entry point fsl_conservation_sums on
the same state: bitwise equal, on two calls in a row, the second on a
smaller mesh (the scratch array keeps the first call's partials).  The
device buffers of the test come from hipMalloc here; the entry point itself
allocates nothing."""
import ctypes as C

import numpy as np
import pytest

import freeshearlayer_amd as fsl
from fsl_test_util import random_state

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


def _hip():
    hip = C.CDLL("libamdhip64.so")
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipFree.argtypes = [C.c_void_p]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipDeviceSynchronize.argtypes = []
    hip.hipStreamCreate.argtypes = [C.POINTER(C.c_void_p)]
    hip.hipStreamSynchronize.argtypes = [C.c_void_p]
    hip.hipStreamDestroy.argtypes = [C.c_void_p]
    return hip


def _dev_copy(hip, a):
    p = C.c_void_p()
    assert hip.hipMalloc(C.byref(p), a.nbytes) == 0
    assert hip.hipMemcpy(p, a.ctypes.data_as(C.c_void_p), a.nbytes, H2D) == 0
    return p


def _context_sums(mesh, U):
    case = fsl.case_reference(*mesh)
    with fsl.Context(case, fsl.config(sgs_model=fsl.FSL_SGS_CONSTANT, fused=1)) as ctx:
        ctx.upload_state(U)
        return ctx.conservation_sums()


@pytest.mark.parametrize("use_stream", [False, True])
def test_noalloc_sums_equal_the_context_sums_on_two_calls(use_stream):
    hip = _hip()
    lib = fsl.load()
    stream = C.c_void_p()
    if use_stream:
        assert hip.hipStreamCreate(C.byref(stream)) == 0
    d_sums = C.c_void_p()
    assert hip.hipMalloc(C.byref(d_sums), 5 * 8) == 0
    try:
        for mesh, seed in (((64, 48, 32), 1), ((16, 12, 8), 2)):   # the second call is smaller
            U = random_state(*mesh, seed=seed)
            want = _context_sums(mesh, U)
            d_U = _dev_copy(hip, U)
            rc = lib.fsl_conservation_sums_noalloc(d_U, mesh[0], mesh[1], mesh[2], d_sums,
                                                   stream if use_stream else None)
            assert rc == 0
            if use_stream:
                assert hip.hipStreamSynchronize(stream) == 0
            else:
                assert hip.hipDeviceSynchronize() == 0
            got = np.zeros(5)
            assert hip.hipMemcpy(got.ctypes.data_as(C.c_void_p), d_sums, 40, D2H) == 0
            hip.hipFree(d_U)
            assert np.array_equal(got, np.asarray(want)), (mesh, got, want)
            # a closed form: the interior sums of the state itself
            ref = U[:, 1:-1, 1:-1, 1:-1].reshape(5, -1).sum(axis=1)
            assert np.allclose(got, ref, rtol=1e-12)
    finally:
        hip.hipFree(d_sums)
        if use_stream:
            hip.hipStreamDestroy(stream)


def test_noalloc_sums_reject_bad_arguments():
    lib = fsl.load()
    assert lib.fsl_conservation_sums_noalloc(None, 4, 4, 4, None, None) == fsl.FSL_ERR_ARG
    assert lib.fsl_conservation_sums_noalloc(C.c_void_p(8), 0, 4, 4, C.c_void_p(8), None) == fsl.FSL_ERR_ARG
