"""The training and comparison drivers end to end on small meshes: the
training script writes a loadable parameter file and a report whose
candidates carry held-out and deployment judge errors; the study driver
writes comparable records for both paths."""
import json

import numpy as np
import pytest

import freeshearlayer_amd as fsl
from freeshearlayer_amd import study, train_sgs

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("act", ["tanh", "relu"])
def test_train_sgs_small(tmp_path, act):
    out, rep = str(tmp_path / "m.bin"), str(tmp_path / "r.json")
    rc = train_sgs.main(["--out", out, "--report", rep, "--mesh", "24x20x16", "--runs", "0.3:500:0.02,0.5:500:0.02",
                         "--samples", "60000", "--steps", "60", "--every", "10",
                         "--batch", "2048", "--schedule", "2:2e-3", "--weight-power", "0,3", "--seeds", "1",
                         "--eval", "0.4:500:0.02", "--eval-steps", "100", "--activation", act])
    assert rc == 0
    p, n, a = fsl.mlp_file_read_act(out)
    assert a == fsl.MLP_ACTIVATIONS[act]
    assert p.shape == (fsl.MLP_NPARAM,) and np.all(np.isfinite(p)) and np.all(n[8:] > 0)
    r = json.load(open(rep))
    assert len(r["candidates"]) == 2 and r["selected"]["weight_power"] in (0.0, 3.0)
    for c in r["candidates"]:
        assert len(c["runs"]) == 2 and all(np.isfinite(x["judge_error"]) for x in c["runs"])
        assert len(c["deploy"]) == 1 and len(c["deploy"][0]["judge_error"]) == 2   # judges at steps 1 and 51


@pytest.mark.parametrize("mode", ["mixed", "learned"])
def test_study_small(tmp_path, mode):
    out = str(tmp_path / "s.json")
    rc = study.main([mode, "--mesh", "24x20x16", "--mach", "0.4", "--steps", "40", "--every", "20", "--out", out])
    assert rc == 0
    d = json.load(open(out))
    assert d["reference"]["step"] == [0, 20, 40] and d["candidate"]["step"] == [0, 20, 40]
    assert d["relative_difference"]["delta_theta"][0] == 0.0   # same initial state
    assert all(np.isfinite(d["final_field_error"]))
    if mode == "learned":
        assert d["candidate"]["learned_status"]["judged"] == 1
