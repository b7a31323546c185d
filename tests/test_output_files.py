"""The reference's output files: the Tecplot frame of Output()
(src/output.c:16-143), the probes.dat records of Probes() (src/probes.c:13-43)
and the Favre-averaged profiles of the post-processor
(favre-average-src/layer2_p.c:27-152), against fixtures written by the
reference itself (tests/golden/make_golden.py: its own main on 16x12x8,
5 steps, and its post-processor on those probes), byte for byte.

On the CPU: the oracle's restatement of the Output() columns and of the
probe stations, formatted by the library's host writers, and the library's
Favre averaging (function and fsl_layer2_p executable).  Closed forms pin
the averaging window and the Favre mean of constant records."""
import gzip
import os
import subprocess

import numpy as np
import pytest

import freeshearlayer_amd as fsl
from oracle_lib import Oracle

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
L, H, D, STEPS = 16, 12, 8, 5
X0 = 0.317391239817


def gold(name):
    return open(os.path.join(GOLD, name), "rb").read()


def gold_plt(variant):
    return gzip.decompress(gold("g7_%s_5.plt.gz" % variant))


def oracle_run(variant):
    """The oracle through the reference main's 5 steps, probes after each."""
    orc = Oracle(L, H, D, dynamic=1 if variant == "dyn" else 0)
    orc.set("BL_HIG", 5)
    orc.set("nStages", 3)
    orc.set("X", X0)
    orc.ic_reference()
    probes = []
    for _ in range(STEPS):
        orc.step()
        probes.append(orc.probe_values())
    return orc, probes


@pytest.mark.parametrize("variant", ["cs", "dyn"])
def test_oracle_frame_formatted_by_the_host_writer_matches_the_reference(tmp_path, variant):
    orc, _ = oracle_run(variant)
    case = fsl.case_reference(L, H, D, BL_HIG=5)
    path = str(tmp_path / "5.plt")
    fsl.plt_write_fields(path, case, orc.output_fields())
    got = open(path, "rb").read()
    ref = gold_plt(variant)
    assert got.splitlines()[:15] == ref.splitlines()[:15]        # header
    assert len(got.splitlines()) == 15 + L * H * D
    assert got == ref
    orc.close()


@pytest.mark.parametrize("real_bytes", [8, 4])
@pytest.mark.parametrize("variant", ["cs", "dyn"])
def test_oracle_probes_appended_by_the_host_writer_match_the_reference(tmp_path, variant, real_bytes):
    orc, probes = oracle_run(variant)
    path = str(tmp_path / "probes.dat")
    for n, p in enumerate(probes, 1):
        fsl.probes_append(path, n, p, real_bytes)
    want = gold("g7_%s_probes.dat" % variant) if real_bytes == 8 else gold("g7_%s_probes_f32.dat" % variant)
    assert open(path, "rb").read() == want
    orc.close()


@pytest.mark.parametrize("variant", ["cs", "dyn"])
def test_favre_profiles_match_the_reference_post_processor(tmp_path, variant):
    probes = str(tmp_path / "probes.dat")
    open(probes, "wb").write(gold("g7_%s_probes_f32.dat" % variant))
    ru, n = fsl.favre_average(probes, H, 2, 4, real_bytes=4)
    assert n == 3
    fsl.favre_write_profiles(str(tmp_path), ru)
    for q in (1, 2, 3):
        assert open(tmp_path / ("probe%d.dat" % q), "rb").read() == gold("g7_%s_probe%d.dat" % (variant, q))


def test_favre_executable_reads_the_reference_prompts(tmp_path):
    open(tmp_path / "probes.dat", "wb").write(gold("g7_dyn_probes_f32.dat"))
    exe = os.path.join(fsl.LIB_DIR, "fsl_layer2_p")
    r = subprocess.run([exe], cwd=str(tmp_path), input="2\n4\n%d\n" % H, stdout=subprocess.PIPE,
                       stderr=subprocess.PIPE, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    assert "stepStart = " in r.stdout and "3 records averaged" in r.stdout
    for q in (1, 2, 3):
        assert open(tmp_path / ("probe%d.dat" % q), "rb").read() == gold("g7_dyn_probe%d.dat" % q)


@pytest.mark.parametrize("real_bytes", [4, 8])
def test_favre_mean_of_constant_records_and_the_step_window(tmp_path, real_bytes):
    """Inside [start, end] every record holds rho = 2, rho u = 6 + station,
    so u~ = 3 + station/2 exactly; records outside the window hold other
    values and must not count; the record past `end` ends the sums."""
    Hs = 7
    path = str(tmp_path / "probes.dat")
    for step in range(1, 9):
        inside = 3 <= step <= 5
        vals = np.zeros((6, Hs))
        for s in range(3):
            vals[2 * s] = 2.0 if inside else 1.0
            vals[2 * s + 1] = (6.0 + s) if inside else 100.0
        fsl.probes_append(path, step, vals, real_bytes)
    ru, n = fsl.favre_average(path, Hs, 3, 5, real_bytes=real_bytes)
    assert n == 3
    for s in range(3):
        assert np.all(ru[s] == 3.0 + s / 2.0)
    # the end of the file also ends the sums (the reference's loop would not)
    ru, n = fsl.favre_average(path, Hs, 3, 100, real_bytes=real_bytes)
    assert n == 6


def test_probe_stations_follow_the_reference_integer_arithmetic():
    """i = LEN/2, 3 LEN/4, LEN and k = DEP/2 + 1 with unsigned division, on a
    state whose cells hold their own indices."""
    Ls, Hs, Ds = 7, 5, 3
    orc = Oracle(Ls, Hs, Ds, dynamic=0)
    i, j, k = np.meshgrid(np.arange(Ls + 2), np.arange(Hs + 2), np.arange(Ds + 2), indexing="ij")
    U = np.stack([1000.0 * i + 10.0 * j + k, -(1000.0 * i + 10.0 * j + k), 0 * i, 0 * i, 0 * i]).astype(float)
    orc.set_state(U)
    pv = orc.probe_values()
    for s, ist in enumerate((Ls // 2, 3 * Ls // 4, Ls)):
        assert np.array_equal(pv[2 * s], 1000.0 * ist + 10.0 * np.arange(1, Hs + 1) + (Ds // 2 + 1))
        assert np.array_equal(pv[2 * s + 1], -pv[2 * s])
    orc.close()


def test_output_columns_of_a_linear_shear_flow():
    """u = a y, v = w = 0, uniform rho and p: |omega| = |S| = a, Q = 0,
    mu_T/mu_L = rho CsDD a / mu_L, p and T as given."""
    Ls, Hs, Ds = 6, 5, 4
    a = 3.0
    orc = Oracle(Ls, Hs, Ds, dynamic=0)
    orc.derive()
    y = (np.arange(Hs + 2) - 0.5) * orc.get("dy")
    rho, p = 1.2, 1.0e5
    shp = (Ls + 2, Hs + 2, Ds + 2)
    u = np.broadcast_to((a * y)[None, :, None], shp)
    U = np.stack([np.full(shp, rho), rho * u, np.zeros(shp), np.zeros(shp), p / 0.4 + 0.5 * rho * u * u])
    orc.set_state(np.ascontiguousarray(U))
    f = orc.output_fields()
    assert np.allclose(f[..., 6], a, rtol=1e-12) and np.allclose(f[..., 0], rho)
    assert np.allclose(f[..., 7], 0.0, atol=1e-9)
    assert np.allclose(f[..., 4], p, rtol=1e-12)
    assert np.allclose(f[..., 5], p / (287.10 * rho), rtol=1e-12)
    assert np.allclose(f[..., 8], rho * orc.get("CsDD") * a / orc.get("mu_L"), rtol=1e-12)
    orc.close()
