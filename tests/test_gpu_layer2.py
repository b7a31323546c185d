"""The layer2-compatible executable on the GPU, run as the reference's
layer2 is: layer2.ini in the working directory, the LCG seed on stdin,
backup.dat written at the end.  The unfused (reference entry points) and
fused paths must write the same bytes, and match the oracle's state; the
learned closure runs from its shipped parameter file."""
import os
import subprocess

import numpy as np
import pytest

import freeshearlayer_amd as fsl
from oracle_lib import Oracle, write_ini

pytestmark = pytest.mark.gpu

EXE = os.path.join(fsl.LIB_DIR, "fsl_layer2")
MODEL = os.path.join(os.path.dirname(fsl.LIB_DIR), "data", "sgs_mlp.bin")
L, H, D, STEPS = 16, 12, 8, 4


def run(tmp_path, name, *args):
    wd = tmp_path / name
    wd.mkdir()
    write_ini(str(wd / "layer2.ini"), L, H, D, numstep=STEPS, f_step=1000, BL_HIG=5)
    r = subprocess.run([EXE, "-q", *args], cwd=str(wd), input="0.317391239817\n", stdout=subprocess.PIPE,
                       stderr=subprocess.PIPE, text=True, timeout=120)
    print(name, r.stdout[-300:], r.stderr[-300:])
    assert r.returncode == 0 and "Simulation finished!" in r.stdout
    return wd


def test_unfused_fused_and_oracle_agree(tmp_path):
    a = run(tmp_path, "unfused")
    b = run(tmp_path, "fused", "--fused")
    ba, bb = open(a / "backup.dat", "rb").read(), open(b / "backup.dat", "rb").read()
    assert len(ba) == L * H * 5 * D * 8 + 8 + 4
    assert ba == bb
    case = fsl.case_reference(L, H, D)
    U = np.zeros((5, L + 2, H + 2, D + 2))
    dt = fsl.C.c_double()
    step = fsl.C.c_uint()
    assert fsl.load().fsl_backup_read(str(a / "backup.dat").encode(), fsl.C.byref(case), fsl._ptr(U),
                                      fsl.C.byref(dt), fsl.C.byref(step), 0, 8) == 0
    assert step.value == STEPS
    orc = Oracle(L, H, D, dynamic=1)
    orc.set("BL_HIG", 5)
    orc.set("X", 0.317391239817)   # the seed on the executable's stdin
    orc.ic_reference()
    for _ in range(STEPS):
        orc.step()
    assert np.array_equal(orc.state()[:, 1:-1, 1:-1, 1:-1], U[:, 1:-1, 1:-1, 1:-1])
    assert dt.value == orc.get("dt")
    orc.close()


def test_learned_closure_from_the_command_line(tmp_path):
    wd = tmp_path / "learned"
    wd.mkdir()
    write_ini(str(wd / "layer2.ini"), L, H, D, numstep=STEPS, f_step=1000, BL_HIG=5)
    r = subprocess.run([EXE, "--fused", "--mixed", "--learned", MODEL], cwd=str(wd), input="0.317391239817\n",
                       stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, timeout=120)
    print(r.stdout[-400:], r.stderr[-300:])
    assert r.returncode == 0 and "Learned closure:" in r.stdout
    assert os.path.getsize(wd / "backup.dat") == L * H * 5 * D * 8 + 12
    r = subprocess.run([EXE, "--learned", str(wd / "missing.bin")], cwd=str(wd), input="0.3\n",
                       stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, timeout=120)
    assert r.returncode == 1 and "cannot load the learned closure" in r.stderr
