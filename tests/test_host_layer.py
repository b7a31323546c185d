"""CPU tests of the ANSI-C host layer and of the C-ABI library itself: it
loads, exports every function include/*.h declares, and its host helpers
(case input, LCG, disturbance block, initial conditions, backup I/O) match
the reference's data files and the oracle.  No GPU calls."""
import ctypes as C
import os
import re
import struct
import subprocess

import numpy as np
import pytest

import freeshearlayer_amd as fsl
from oracle_lib import Oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")
X0 = 0.317391239817


def declared(header):
    txt = open(os.path.join(ROOT, "include", header)).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    names = re.findall(r"^[A-Za-z_][\w\s\*]*?\b([A-Za-z_]\w*)\s*\(", txt, flags=re.M)
    return sorted(set(n for n in names if n not in ("if", "sizeof")))


def defined_symbols(path):
    out = subprocess.run(["nm", "-D", "--defined-only", path], stdout=subprocess.PIPE, text=True,
                         check=True).stdout
    return set(l.split()[-1] for l in out.splitlines() if l.strip())


def test_libfsl_loads_and_exports_fsl_h():
    fsl.load()
    have = defined_symbols(fsl.LIB_PATH)
    missing = [n for n in declared("fsl.h") if n not in have]
    assert not missing, "libfsl.so lacks %s" % missing


def test_libfsl_layer2_exports_reference_names():
    C.CDLL(fsl.LAYER2_PATH)
    have = defined_symbols(fsl.LAYER2_PATH)
    missing = [n for n in declared("fsl_layer2.h") if n not in have]
    assert not missing, "libfsl_layer2.so lacks %s" % missing
    for n in ("Reconstruction", "Fluxes", "Evolution", "BounCondInGhostCells", "BounCondOnInterfaces",
              "checkCoNum", "DynamicSmagorinsky", "filter_", "Random", "minmod", "Array3D", "free3D",
              "Evolution_threeStage_TVD_RK", "Evolution_twoStage_TVD_RK"):
        assert n in have


def test_read_reference_ini():
    c = fsl.FslCase()
    assert fsl.load().fsl_case_read_ini(os.path.join(GOLD, "case_layer2.ini").encode(), C.byref(c)) == 0
    assert (c.LEN, c.HIG, c.DEP) == (100, 80, 35)
    assert c.deltaX == 1.6666667 and c.deltaY == 1.0 and c.deltaZ == 1.5
    assert c.deltaT == 0.0007 and c.Answer == 0 and c.numstep == 3000
    assert c.mu_L == 0.253358 and c.Pr_L == 0.72 and c.Pr_T == 0.8 and c.Cs == 0.01
    assert c.BL_HIG == 5 and c.Ua == 35.0 and c.Va == 35.0 and c.Wa == 35.0
    assert (c.Ns_min, c.Nst, c.f_step, c.nStages) == (10, 50, 50, 2)
    assert c.maxCoNum == 0.1


def test_read_reference_bin_matches_float32_payload():
    raw = open(os.path.join(GOLD, "case_mpi_layer2.bin"), "rb").read()
    v = struct.unpack("22f", raw)
    c = fsl.FslCase()
    assert fsl.load().fsl_case_read_bin(os.path.join(GOLD, "case_mpi_layer2.bin").encode(), C.byref(c)) == 0
    assert (c.LEN, c.HIG, c.DEP) == (int(v[0]), int(v[1]), int(v[2])) == (50, 80, 35)
    assert c.deltaX == np.float32(1.6666667) and c.numstep == 6000 and c.nStages == 3
    assert c.maxCoNum == np.float32(0.1)


def test_ini_to_bin_reproduces_reference_bin(tmp_path):
    out = tmp_path / "x.bin"
    assert fsl.load().fsl_case_ini_to_bin(os.path.join(GOLD, "case_ini2bin.ini").encode(),
                                          str(out).encode()) == 0
    assert out.read_bytes() == open(os.path.join(GOLD, "case_ini2bin.bin"), "rb").read()


def test_validate_rejects_out_of_bounds_block():
    c = fsl.case_reference(10, 6, 4)
    msg = C.create_string_buffer(256)
    assert fsl.load().fsl_case_validate(C.byref(c), msg, 256) != 0
    assert b"BL_HIG" in msg.value
    c.BL_HIG = 3
    assert fsl.load().fsl_case_validate(C.byref(c), msg, 256) == 0


def test_lcg_and_blocks_match_golden():
    g = np.load(os.path.join(GOLD, "g4_lcg.npz"), allow_pickle=False)
    lib = fsl.load()
    seq = [X0]
    for _ in range(64):
        seq.append(lib.fsl_random(seq[-1]))
    assert np.array_equal(np.array(seq), g["random_seq"])
    for D in (32, 96, 256):
        case = fsl.case_reference(4, 10, D)
        b = fsl.FslBlock()
        lib.fsl_block_init(C.byref(b), X0)
        draws = []
        for _ in range(16):
            b.counter = 0
            assert lib.fsl_block_advance(C.byref(b), C.byref(case)) == 1
            draws.append([b.counter, b.k_min, b.k_max, b.Ud, b.Vd, b.Wd, b.X])
        assert np.array_equal(np.array(draws), g["blocks_D%d" % D])


def test_block_lifetime_counts_down():
    lib = fsl.load()
    case = fsl.case_reference(4, 10, 32)
    b = fsl.FslBlock()
    lib.fsl_block_init(C.byref(b), X0)
    drawn = [lib.fsl_block_advance(C.byref(b), C.byref(case)) for _ in range(100)]
    assert drawn[0] == 1 and drawn[1:42] == [0] * 41 and drawn[42] == 1


def test_minmod_host_matches_oracle():
    rng = np.random.default_rng(1)
    for x, y in rng.uniform(-2, 2, (300, 2)):
        assert fsl.load().fsl_minmod(x, y) == Oracle.lib().or_minmod(x, y)


def test_ic_reference_matches_golden():
    g = np.load(os.path.join(GOLD, "g1_ref_cs.npz"), allow_pickle=False)
    L, H, D = (int(v) for v in g["mesh"])
    assert np.array_equal(fsl.ic_reference(fsl.case_reference(L, H, D)), g["U0"])


def test_shear_layer_ic_properties():
    case = fsl.case_reference(16, 32, 8)
    U, u_in, mu = fsl.ic_shear_layer(case)
    c0 = np.sqrt(1.4e5)
    dU = 2 * 0.3 * c0
    # streams at 2 dU and dU far from the layer, monotone profile, inflow = mean profile
    assert abs(u_in[-1] - 2 * dU) / dU < 1e-3 and abs(u_in[0] - dU) / dU < 1e-3
    assert np.all(np.diff(u_in) > 0)
    rho = U[0, 1:-1, 1:-1, 1:-1]
    assert np.all(rho == 1.0)
    u = U[1, 1:-1, 1:-1, 1:-1] / rho
    v = U[2, 1:-1, 1:-1, 1:-1] / rho
    p = (U[4, 1:-1, 1:-1, 1:-1] - 0.5 * rho * (u * u + v * v + (U[3, 1:-1, 1:-1, 1:-1] / rho) ** 2)) * 0.4
    np.testing.assert_allclose(p, 1e5, rtol=1e-12)
    # perturbations only inside |y| < delta_w0 = 4 dy, bounded by 2 % dU
    y = (np.arange(1, 33) - 16 - 0.5)
    inside = np.abs(y) < 4
    dev = np.abs(u - u_in[None, :, None])
    assert np.all(dev[:, ~inside, :] == 0) and np.all(v[:, ~inside, :] == 0)
    assert dev[:, inside, :].max() <= 0.02 * dU and dev[:, inside, :].max() > 0.01 * dU
    assert mu == pytest.approx(dU * 4.0 / 500.0)
    # seeded: same seed, same field; another seed, another field
    U2, _, _ = fsl.ic_shear_layer(case)
    assert np.array_equal(U, U2)
    s = fsl.shear_defaults()
    s.seed = 7
    U3, _, _ = fsl.ic_shear_layer(case, s)
    assert not np.array_equal(U, U3)


@pytest.mark.parametrize("real_bytes", [8, 4])
def test_backup_round_trip_and_reference_layout(tmp_path, real_bytes):
    lib = fsl.load()
    case = fsl.case_reference(5, 4, 3)
    U = np.random.default_rng(0).uniform(1, 2, (5, 7, 6, 5))
    U[:, 0] = 0; U[:, -1] = 0; U[:, :, 0] = 0; U[:, :, -1] = 0; U[:, :, :, 0] = 0; U[:, :, :, -1] = 0
    path = str(tmp_path / "backup.dat").encode()
    assert lib.fsl_backup_write(path, C.byref(case), fsl._ptr(U), 1.25e-3, 77, real_bytes) == 0
    n = 5 * 4 * 3 * 5
    assert os.path.getsize(path) == n * real_bytes + real_bytes + 4
    back = np.zeros_like(U)
    dt = C.c_double(); step = C.c_uint()
    assert lib.fsl_backup_read(path, C.byref(case), fsl._ptr(back), C.byref(dt), C.byref(step), 0,
                               real_bytes) == 0
    if real_bytes == 8:
        assert np.array_equal(back, U) and dt.value == 1.25e-3
    else:
        assert np.array_equal(back, U.astype(np.float32).astype(np.float64))
    assert step.value == 77
    # reference reader: step comes from the first 4 bytes of the deltaT slot
    assert lib.fsl_backup_read(path, C.byref(case), fsl._ptr(back), C.byref(dt), C.byref(step), 1,
                               real_bytes) == 0
    slot = struct.pack("d", 1.25e-3) if real_bytes == 8 else struct.pack("f", np.float32(1.25e-3))
    assert step.value == struct.unpack("I", slot[:4])[0]
    # layout of src/finalize.c: for i, j: DEP values of U1..U5
    raw = np.frombuffer(open(path, "rb").read()[: n * real_bytes],
                        dtype=np.float64 if real_bytes == 8 else np.float32)
    assert raw[0] == np.asarray(U[0, 1, 1, 1], raw.dtype)
    assert raw[3] == np.asarray(U[1, 1, 1, 1], raw.dtype)


def test_layer2_executable_usage():
    exe = os.path.join(fsl.LIB_DIR, "fsl_layer2")
    r = subprocess.run([exe, "--bogus"], stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True)
    assert r.returncode == 2 and "usage" in r.stderr
