"""Size limits of fsl_create, checked before any device call (runs without a
GPU): the mixed-precision stage kernel and the learned closure's network
kernel address one (j, k) plane of a field with 32-bit byte offsets; beyond
that fsl_create returns FSL_ERR_UNSUPPORTED.  A mesh with an empty direction
is FSL_ERR_ARG."""
import pytest

import freeshearlayer_amd as fsl


def test_mixed_path_rejects_a_plane_beyond_32_bit_offsets():
    case = fsl.case_reference(4, 16384, 16384)     # (H + 2)(D + 2) * 8 bytes > 2^31 - 1
    with pytest.raises(fsl.FslError, match=r"\(%d\)" % fsl.FSL_ERR_UNSUPPORTED):
        fsl.Context(case, fsl.config(precision=fsl.FSL_PREC_MIXED, fused=1))


def test_learned_closure_rejects_a_plane_beyond_32_bit_offsets():
    case = fsl.case_reference(4, 16384, 16384)     # (H + 2)(D + 2) * 8 bytes > 2^31 - 1
    with pytest.raises(fsl.FslError, match=r"\(%d\)" % fsl.FSL_ERR_UNSUPPORTED):
        fsl.Context(case, fsl.config(sgs_model=fsl.FSL_SGS_LEARNED, fused=1))


@pytest.mark.parametrize("mesh", [(0, 8, 8), (8, 0, 8), (8, 8, 0), (8, -1, 8)])
def test_create_rejects_an_empty_direction(mesh):
    case = fsl.case_reference(*mesh)
    with pytest.raises(fsl.FslError, match=r"\(%d\)" % fsl.FSL_ERR_ARG):
        fsl.Context(case, fsl.config(fused=1))


def test_error_codes_match_the_header():
    import os
    import re
    hdr = open(os.path.join(os.path.dirname(__file__), "..", "include", "fsl.h")).read()
    codes = dict((m.group(1), int(m.group(2))) for m in re.finditer(r"#define (FSL_ERR_\w+)\s+(-?\d+)", hdr))
    assert codes and all(getattr(fsl, k) == v for k, v in codes.items()), codes


def test_fused_path_rejects_two_stages():
    """The fused stage kernels implement the three-stage TVD RK only: the
    reference's nStages = 2 (its two-stage update falling through into the
    three-stage one, src/evolution.c:15-27) is an unfused-path case, and
    fsl_create refuses the fused path for it before any device call."""
    case = fsl.case_reference(16, 12, 8, nStages=2)
    with pytest.raises(fsl.FslError, match=r"\(%d\)" % fsl.FSL_ERR_UNSUPPORTED):
        fsl.Context(case, fsl.config(fused=1))
    for n in (1, 4):   # not a stage count of the reference (src/evolution.c:15-27)
        with pytest.raises(fsl.FslError, match=r"\(%d\)" % fsl.FSL_ERR_ARG):
            fsl.Context(fsl.case_reference(16, 12, 8, nStages=n), fsl.config(fused=0))


def test_fast_fp64_needs_the_fused_path():
    """FSL_PREC_FP64_FAST is a build of the fused stage kernel: an unfused
    context with it is refused before any device call."""
    case = fsl.case_reference(16, 12, 8)
    with pytest.raises(fsl.FslError, match=r"\(%d\)" % fsl.FSL_ERR_UNSUPPORTED):
        fsl.Context(case, fsl.config(precision=fsl.FSL_PREC_FP64_FAST, fused=0))
    with pytest.raises(fsl.FslError, match=r"\(%d\)" % fsl.FSL_ERR_ARG):
        fsl.Context(case, fsl.config(precision=3, fused=1))


def test_precision_codes_match_the_header():
    import os
    import re
    hdr = open(os.path.join(os.path.dirname(__file__), "..", "include", "fsl.h")).read()
    codes = dict((m.group(1), int(m.group(2))) for m in re.finditer(r"#define (FSL_PREC_\w+)\s+(\d+)", hdr))
    assert len(codes) == 3 and all(getattr(fsl, k) == v for k, v in codes.items()), codes
