"""Host side of the learned closure and the shear-layer statistics (no GPU):
parameter init and file format, and fsl_shear_statistics against closed
forms.  For the Favre profile u~ = Um + dU/2 tanh(2 y / delta_w) with
uniform density, delta_theta = delta_w / 4 exactly in the continuum
(integral of dU^2/4 sech^2), and the central-difference vorticity thickness
tends to delta_w; an alternating fluctuation +-a along z gives a Favre TKE
of exactly 1/2 a^2 per unit mass."""
import os

import numpy as np
import pytest

import freeshearlayer_amd as fsl
import mlp_ref


def test_init_params_layout_and_scale():
    p, n = fsl.mlp_init_params(42)
    assert p.shape == (fsl.MLP_NPARAM,) and p.dtype == np.float32
    W1, b1, W2, b2, w3, b3 = mlp_ref.unpack(p)
    assert np.all(b1 == 0) and np.all(b2 == 0) and b3 == np.float32(0.1)
    # Glorot uniform bounds sqrt(6 / (fan_in + fan_out))
    assert np.abs(W1).max() <= np.sqrt(6 / 72) and np.abs(W1).max() > 0.8 * np.sqrt(6 / 72)
    assert np.abs(W2).max() <= np.sqrt(6 / 128) and abs(W2.mean()) < 0.01
    assert np.array_equal(n, np.concatenate([np.zeros(8), np.ones(8)]).astype(np.float32))
    q, _ = fsl.mlp_init_params(42)
    r, _ = fsl.mlp_init_params(43)
    assert np.array_equal(p, q) and not np.array_equal(p, r)


def test_param_file_round_trip_and_rejects(tmp_path):
    p, n = fsl.mlp_init_params(1)
    n[:8] = np.arange(8)
    path = str(tmp_path / "m.bin")
    fsl.mlp_file_write(path, p, n)
    assert os.path.getsize(path) == 8 + 12 + 64 + 4 * fsl.MLP_NPARAM
    q, m = fsl.mlp_file_read(path)
    assert np.array_equal(p, q) and np.array_equal(n, m)
    raw = bytearray(open(path, "rb").read())
    raw[0] = ord("X")
    bad = str(tmp_path / "bad.bin")
    open(bad, "wb").write(bytes(raw))
    with pytest.raises(fsl.FslError):
        fsl.mlp_file_read(bad)
    open(bad, "wb").write(open(path, "rb").read()[:100])
    with pytest.raises(fsl.FslError):
        fsl.mlp_file_read(bad)


def test_param_file_activation(tmp_path):
    """FSLMLP02 files carry the activation; the FSLMLP01 reader refuses a
    non-tanh one, the _act reader takes either format."""
    p, n = fsl.mlp_init_params(2)
    relu = str(tmp_path / "relu.bin")
    fsl.mlp_file_write_act(relu, p, n, "relu")
    assert os.path.getsize(relu) == 8 + 16 + 64 + 4 * fsl.MLP_NPARAM
    assert open(relu, "rb").read(8) == b"FSLMLP02"
    q, m, act = fsl.mlp_file_read_act(relu)
    assert act == fsl.FSL_MLP_RELU and np.array_equal(p, q) and np.array_equal(n, m)
    with pytest.raises(fsl.FslError):
        fsl.mlp_file_read(relu)
    tanh = str(tmp_path / "tanh.bin")
    fsl.mlp_file_write_act(tanh, p, n, "tanh")          # written in the first format
    assert open(tanh, "rb").read(8) == b"FSLMLP01"
    assert fsl.mlp_file_read_act(tanh)[2] == fsl.FSL_MLP_TANH
    q, m = fsl.mlp_file_read(tanh)
    assert np.array_equal(p, q)
    raw = bytearray(open(relu, "rb").read())
    raw[20:24] = (7).to_bytes(4, "little")                # unknown activation
    bad = str(tmp_path / "bad.bin")
    open(bad, "wb").write(bytes(raw))
    with pytest.raises(fsl.FslError):
        fsl.mlp_file_read_act(bad)
    with pytest.raises(fsl.FslError):
        fsl.mlp_file_write_act(bad, p, n, 5)


def tanh_state(L, H, D, dw_cells, a=0.0):
    case = fsl.case_reference(L, H, D)
    dy = case.deltaY
    U1, U2 = 200.0, 100.0
    y = (np.arange(1, H + 1) - H // 2 - 0.5) * dy
    ut = 0.5 * (U1 + U2) + 0.5 * (U1 - U2) * np.tanh(2 * y / (dw_cells * dy))
    U = np.zeros((5, L + 2, H + 2, D + 2))
    U[0] = 1.0
    sgn = np.where(np.arange(D + 2) % 2 == 0, 1.0, -1.0)
    U[1, 1:-1, 1:-1, 1:-1] = ut[None, :, None] + a * sgn[None, None, 1:-1]
    U[2, 1:-1, 1:-1, 1:-1] = a * sgn[None, None, 1:-1]
    return case, U, U1, U2


def test_shear_statistics_closed_forms():
    case, U, U1, U2 = tanh_state(6, 400, 4, dw_cells=40)
    prof = mlp_ref.plane_profiles(U)
    st = fsl.shear_statistics(case, prof, U1, U2)
    dw = 40 * case.deltaY
    assert abs(st["delta_theta"] - dw / 4) < 1e-6 * dw
    assert abs(st["delta_theta_out"] - dw / 4) < 1e-6 * dw
    assert abs(st["delta_w"] - dw) < 5e-3 * dw
    assert np.allclose(st["delta_theta_x"], dw / 4, rtol=1e-6)
    assert abs(st["tke"]) < 1e-6
    # alternating +-a on u and v along z (D even): Favre variance a^2 each
    a = 3.0
    case, U, U1, U2 = tanh_state(6, 50, 4, dw_cells=10, a=a)
    st = fsl.shear_statistics(case, mlp_ref.plane_profiles(U), U1, U2)
    vol = case.deltaX * case.deltaY * case.deltaZ
    expect = 0.5 * (2 * a * a) * vol * 6 * 50 * 4
    assert abs(st["tke"] - expect) < 1e-9 * expect


def test_shear_statistics_rejects_bad_input():
    case, U, U1, U2 = tanh_state(4, 20, 2, dw_cells=4)
    with pytest.raises(fsl.FslError):
        fsl.shear_statistics(case, mlp_ref.plane_profiles(U), U2, U1)   # dU <= 0
