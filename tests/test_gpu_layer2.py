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


def _backup_step(path):
    raw = open(path, "rb").read()
    return int(np.frombuffer(raw[-4:], dtype=np.uint32)[0]), raw


def test_stopfile_ends_the_run_after_the_current_step(tmp_path):
    """src-par/mpi_layer2.c:169-191: a stopfile holding 's' stops the run
    after the step in progress; backup.dat is written for that step."""
    wd = tmp_path / "stop"
    wd.mkdir()
    write_ini(str(wd / "layer2.ini"), L, H, D, numstep=1000, f_step=1000, BL_HIG=5)
    (wd / "stopfile").write_text("s\n")
    r = subprocess.run([EXE, "-q"], cwd=str(wd), input="0.317391239817\n", stdout=subprocess.PIPE,
                       stderr=subprocess.PIPE, text=True, timeout=120)
    assert r.returncode == 0 and "Simulation finished!" in r.stdout
    step, _ = _backup_step(wd / "backup.dat")
    assert step == 1
    (wd / "stopfile").write_text("c\n")   # anything else does not stop it
    write_ini(str(wd / "layer2.ini"), L, H, D, numstep=3, f_step=1000, BL_HIG=5)
    r = subprocess.run([EXE, "-q"], cwd=str(wd), input="0.317391239817\n", stdout=subprocess.PIPE,
                       stderr=subprocess.PIPE, text=True, timeout=120)
    assert r.returncode == 0 and _backup_step(wd / "backup.dat")[0] == 3


def test_sigint_ends_the_run_after_the_current_step(tmp_path):
    """src/layer2.c:107-127: SIGINT clears the continue flag; the loop ends
    after the current step and the run still writes backup.dat."""
    import signal
    import time
    wd = tmp_path / "sigint"
    wd.mkdir()
    write_ini(str(wd / "layer2.ini"), L, H, D, numstep=10 ** 7, f_step=1000, BL_HIG=5)
    p = subprocess.Popen([EXE], cwd=str(wd), stdin=subprocess.PIPE, stdout=subprocess.PIPE,
                         stderr=subprocess.PIPE, text=True)
    try:
        p.stdin.write("0.317391239817\n")
        p.stdin.close()
        seen = 0
        t0 = time.time()
        while time.time() - t0 < 60:
            line = p.stdout.readline()
            if line.startswith("Timestep no.:"):
                seen += 1
                if seen == 3:
                    break
        assert seen == 3, "the run printed no step lines"
        p.send_signal(signal.SIGINT)
        p.wait(timeout=60)
        out, err = p.stdout.read(), p.stderr.read()
    finally:
        if p.poll() is None:
            p.kill()
            p.wait()
    assert p.returncode == 0 and "Simulation finished!" in out, err
    step, _ = _backup_step(wd / "backup.dat")
    assert 3 <= step < 10 ** 7


def test_failed_step_leaves_backup_dat_unchanged(tmp_path):
    """A restart (Answer = 1) from a state with a NaN cell fails its first
    CFL rescale (FSL_ERR_NONFINITE); the run exits 1 and backup.dat -- its
    own input -- is left byte for byte as it was."""
    wd = tmp_path / "nan"
    wd.mkdir()
    case = fsl.case_reference(L, H, D, BL_HIG=5)
    U = fsl.ic_reference(case)
    U[0, 5, 4, 3] = np.nan
    assert fsl.load().fsl_backup_write(str(wd / "backup.dat").encode(), fsl.C.byref(case), fsl._ptr(U),
                                       fsl.C.c_double(7e-4), 0, 8) == 0
    before = open(wd / "backup.dat", "rb").read()
    write_ini(str(wd / "layer2.ini"), L, H, D, numstep=STEPS, f_step=1000, BL_HIG=5, answer=1)
    r = subprocess.run([EXE, "-q"], cwd=str(wd), input="0.317391239817\n", stdout=subprocess.PIPE,
                       stderr=subprocess.PIPE, text=True, timeout=120)
    assert r.returncode == 1 and "step failed" in r.stderr and "not written" in r.stderr
    assert open(wd / "backup.dat", "rb").read() == before
