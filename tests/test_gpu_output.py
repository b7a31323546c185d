"""The output files from the device state, against the reference's own.

fsl_layer2 runs the reference main's case (16x12x8, 5 steps, f_step 3,
seed on stdin) and writes probes.dat every step, <step>.plt every f_step
steps and at the end; the files must equal, byte for byte, those the
reference's own main wrote for the same case (tests/golden/g7_*, written by
make_golden.py from the compiled reference): the frame computed on the device
(vorticity magnitude, Q, mu_T/mu_L in FP64) and formatted on the host, the
probe columns gathered on the device.  Larger meshes: the device columns
against the oracle's restatement, bit for bit."""
import gzip
import os
import subprocess

import numpy as np
import pytest

import freeshearlayer_amd as fsl
from fsl_test_util import assert_bitwise, random_state
from oracle_lib import Oracle, write_ini

pytestmark = pytest.mark.gpu

EXE = os.path.join(fsl.LIB_DIR, "fsl_layer2")
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
L, H, D = 16, 12, 8


def gold(name):
    return open(os.path.join(GOLD, name), "rb").read()


@pytest.mark.parametrize("real4", [False, True])
@pytest.mark.parametrize("fused", [False, True])
@pytest.mark.parametrize("variant", ["cs", "dyn"])
def test_executable_writes_the_reference_files(tmp_path, variant, fused, real4):
    wd = tmp_path / "run"
    wd.mkdir()
    write_ini(str(wd / "layer2.ini"), L, H, D, numstep=5, f_step=3, BL_HIG=5, nStages=3)
    args = [EXE, "-q"] + (["--const-cs"] if variant == "cs" else []) + (["--fused"] if fused else []) + \
           (["--real4"] if real4 else [])
    r = subprocess.run(args, cwd=str(wd), input="0.317391239817\n", stdout=subprocess.PIPE,
                       stderr=subprocess.PIPE, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert (wd / "3.plt").exists()                       # the frame of step % f_step == 0
    assert open(wd / "5.plt", "rb").read() == gzip.decompress(gold("g7_%s_5.plt.gz" % variant))
    want = gold("g7_%s_probes_f32.dat" % variant) if real4 else gold("g7_%s_probes.dat" % variant)
    assert open(wd / "probes.dat", "rb").read() == want
    if real4:   # the Favre post-processor over the run's own probes
        r = subprocess.run([os.path.join(fsl.LIB_DIR, "fsl_layer2_p")], cwd=str(wd), input="2\n4\n%d\n" % H,
                           stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, timeout=60)
        assert r.returncode == 0
        for q in (1, 2, 3):
            assert open(wd / ("probe%d.dat" % q), "rb").read() == gold("g7_%s_probe%d.dat" % (variant, q))


@pytest.mark.parametrize("precision", [fsl.FSL_PREC_FP64, fsl.FSL_PREC_MIXED])
@pytest.mark.parametrize("mesh", [(64, 48, 32), (37, 21, 19)])
def test_device_columns_match_the_oracle(mesh, precision):
    Ls, Hs, Ds = mesh
    U = random_state(Ls, Hs, Ds, seed=21)
    orc = Oracle(Ls, Hs, Ds, dynamic=1)
    orc.set_state(U)
    case = fsl.case_reference(Ls, Hs, Ds)
    with fsl.Context(case, fsl.config(precision=precision, fused=1)) as ctx:
        ctx.upload_state(U)
        got = ctx.output_fields()
        assert_bitwise(orc.output_fields(), got, "output columns")
        # in slabs of planes, as the frame writer streams them
        part = ctx.output_fields(k0=5, nk=7)
        assert_bitwise(got[4:11], part, "planes 5..11")
        assert_bitwise(orc.probe_values(), ctx.probe_values(), "probe columns")
        with pytest.raises(fsl.FslError):
            ctx.output_fields(k0=Ds, nk=2)               # beyond the last plane
    orc.close()


def test_frame_of_a_larger_mesh_streams_in_slabs(tmp_path):
    """A frame bigger than one slab of the writer: same bytes as the host
    writer applied to the oracle's columns."""
    Ls, Hs, Ds = 128, 96, 64          # 37 planes per slab: two slabs
    U = random_state(Ls, Hs, Ds, seed=4)
    orc = Oracle(Ls, Hs, Ds, dynamic=1)
    orc.set_state(U)
    case = fsl.case_reference(Ls, Hs, Ds)
    with fsl.Context(case, fsl.config(fused=1)) as ctx:
        ctx.upload_state(U)
        ctx.write_plt(str(tmp_path / "dev.plt"))
    fsl.plt_write_fields(str(tmp_path / "orc.plt"), case, orc.output_fields())
    assert open(tmp_path / "dev.plt", "rb").read() == open(tmp_path / "orc.plt", "rb").read()
    orc.close()
