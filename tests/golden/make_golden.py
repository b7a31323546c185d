"""Writes the golden fixtures of tests/golden/ from the compiled reference.

The reference (nikola-m/freeShearLayer, serial src/) is built from its own
sources by oracle/Makefile into oracle/_ref/ with real = double and
zero-initialised Array3D rows; this script drives it through
oracle/ref_harness.c and stores inputs and outputs as compressed .npz data.
It runs only where /root/reference exists; the fixtures are committed and
travel with the repository.

    python tests/golden/make_golden.py

Fixtures
  g1_ref_cs.npz / g1_ref_dyn.npz   16x12x8, two-stream IC, seed X0: state after
                                   the stage-1 ghost fill, 30 face arrays after
                                   Reconstruction and after BounCondOnInterfaces,
                                   15 flux arrays, mu_SGS, state after each stage
  g2_rand_cs.npz / g2_rand_dyn.npz 12x10x6 seeded random state: the same
                                   intermediates plus 2 full steps
  g3_long_cs.npz / g3_long_dyn.npz 64x48x32 two-stream IC: per-step dt and
                                   block log for 100 steps, per-field norms and
                                   mid-z x-lines after steps 1, 10, 100
  g4_lcg.npz                       Random() sequence and block draws for
                                   DEP in {32, 96, 256}
  g5_filter.npz                    filter_ of a seeded field, 9x7x5
  g6_nstages2.npz                  16x12x8 with nStages = 2 (Evolution falls
                                   through), 4 steps
  g7_<v>_5.plt.gz, g7_<v>_probes.dat   (v = cs, dyn) the reference's own main
                                   (src/layer2.c) on 16x12x8, 5 steps, f_step 3,
                                   seed X0: its Output() frame after step 5 and
                                   the probes.dat its Probes() appended (8-byte
                                   reals: real = double)
  g7_<v>_probes_f32.dat            the same records in the 4-byte layout of the
                                   shipped build (converted here)
  g7_<v>_probe{1,2,3}.dat          the reference's Favre post-processor
                                   (favre-average-src/layer2_p.c) on that file,
                                   stepStart 2, stepEnd 4, HIG 12
  case_layer2.ini, case_mpi_layer2.bin, case_ini2bin.ini, case_ini2bin.bin,
  case_seed.txt                    the reference's case files (data)
"""
import ctypes
import gzip
import hashlib
import json
import os
import shutil
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from oracle_lib import REF_DIR, RefLib, write_ini  # noqa: E402
from fsl_test_util import random_state  # noqa: E402

REF_ROOT = "/root/reference"
X0 = 0.317391239817


def stage1_intermediates(ref, out, prefix=""):
    """Runs checkCoNum, the block draw and stage 1 entry point by entry point."""
    ref.check_co_num()
    ref.block_draw()
    s = ref.scalars()
    out[prefix + "dt"] = s["deltaT"]
    out[prefix + "block"] = np.array([s["k_min"], s["k_max"], s["Ud"], s["Vd"], s["Wd"], s["X"]])
    ref.ghost()
    out[prefix + "U_ghost"] = ref.state()
    ref.recon()
    out[prefix + "faces_recon"] = np.concatenate([ref.face(f).ravel() for f in range(30)])
    ref.iface()
    out[prefix + "faces_bc"] = np.concatenate([ref.face(f).ravel() for f in range(30)])
    ref.fluxes()
    out[prefix + "fluxes"] = np.concatenate(
        [ref.face(f).ravel() for f in (0, 1, 2, 3, 4, 10, 11, 12, 13, 14, 20, 21, 22, 23, 24)])
    out[prefix + "mu_sgs"] = ref.cell(10)
    ref.evolution(3, 1)
    out[prefix + "Up_stage1"] = np.stack([ref.cell(5 + f) for f in range(5)])
    for st in (2, 3):
        ref.ghost(); ref.recon(); ref.iface(); ref.fluxes(); ref.evolution(3, st)
        out[prefix + "U_stage%d" % st] = ref.state()
        out[prefix + "Up_stage%d" % st] = np.stack([ref.cell(5 + f) for f in range(5)])


def g1(variant):
    ref = RefLib(variant)
    L, H, D = 16, 12, 8
    ref.initialize(L, H, D)
    ref.set_time(0.0007, X0, 0)
    out = {"mesh": np.array([L, H, D]), "U0": ref.state()}
    stage1_intermediates(ref, out)
    return out


def g2(variant):
    ref = RefLib(variant)
    L, H, D = 12, 10, 6
    ref.initialize(L, H, D)
    ref.set_time(0.0007, X0, 0)
    U0 = random_state(L, H, D, seed=1234)
    ref.set_state(U0)
    out = {"mesh": np.array([L, H, D]), "U0": U0}
    stage1_intermediates(ref, out)
    for n in range(2):
        ref.step()
        out["U_step%d" % (n + 2)] = ref.state()
        out["dt_step%d" % (n + 2)] = ref.scalars()["deltaT"]
    return out


def g3(variant):
    ref = RefLib(variant)
    L, H, D = 64, 48, 32
    ref.initialize(L, H, D)
    ref.set_time(0.0007, X0, 0)
    dts, blocks = [], []
    out = {"mesh": np.array([L, H, D])}
    for n in range(1, 101):
        ref.step()
        s = ref.scalars()
        dts.append(s["deltaT"])
        blocks.append([s["k_min"], s["k_max"], s["Ud"], s["Vd"], s["Wd"], s["counter"]])
        if n in (1, 10, 100):
            U = ref.state()[:, 1:-1, 1:-1, 1:-1]
            out["norms_step%d" % n] = np.array(
                [[np.abs(U[f]).sum(), np.sqrt((U[f] ** 2).sum()), np.abs(U[f]).max()] for f in range(5)])
            k = D // 2 + 1
            out["xlines_step%d" % n] = ref.state()[:, :, [H // 4, H // 2, 3 * H // 4], k]
    out["dt"] = np.array(dts)
    out["blocks"] = np.array(blocks)
    return out


def g4():
    ref = RefLib("cs")
    seq = [X0]
    for _ in range(64):
        seq.append(ref.random(seq[-1]))
    out = {"random_seq": np.array(seq)}
    for D in (32, 96, 256):
        ref.initialize(4, 4, D)
        ref.set_time(0.0007, X0, 0)
        draws = []
        for _ in range(16):
            ref.set_time(0.0007, ref.scalars()["X"], 0)   # counter 0: draw every call
            ref.block_draw()
            s = ref.scalars()
            draws.append([s["counter"], s["k_min"], s["k_max"], s["Ud"], s["Vd"], s["Wd"], s["X"]])
        out["blocks_D%d" % D] = np.array(draws)
    return out


def g5():
    ref = RefLib("cs")
    L, H, D = 9, 7, 5
    ref.initialize(L, H, D)
    a = random_state(L, H, D, seed=77)[1]
    return {"mesh": np.array([L, H, D]), "input": a, "output": ref.filter(a)}


def g6():
    ref = RefLib("dyn")
    L, H, D = 16, 12, 8
    ref.initialize(L, H, D, nStages=2)
    ref.set_time(0.0007, X0, 0)
    out = {"mesh": np.array([L, H, D])}
    for n in range(1, 5):
        ref.step()
        out["U_step%d" % n] = ref.state()
        out["Up_step%d" % n] = np.stack([ref.cell(5 + f) for f in range(5)])
        out["dt_step%d" % n] = ref.scalars()["deltaT"]
    return out


G7 = dict(L=16, H=12, D=8, numstep=5, f_step=3, BL_HIG=5, nStages=3)


def run_reference_main(variant, wd):
    """The reference's main in this (fresh) process, working directory wd."""
    lib = ctypes.CDLL(os.path.join(REF_DIR, "libfslref_%s.so" % variant))
    lib.fslref_main.argtypes = [ctypes.c_char_p, ctypes.c_char_p]
    seed = os.path.join(wd, "seed.txt")
    with open(seed, "w") as f:
        f.write("%.12f\n" % X0)
    return lib.fslref_main(wd.encode(), seed.encode())


def probes_to_f32(raw, H):
    """probes.dat records (uint32 step + 6 H doubles) in the float layout."""
    rec = 4 + 6 * H * 8
    out = b""
    for o in range(0, len(raw), rec):
        out += raw[o:o + 4] + np.frombuffer(raw[o + 4:o + rec], dtype=np.float64).astype(np.float32).tobytes()
    return out


def g7(variant):
    """Output files of the reference's own main and post-processor."""
    files = {}
    L, H, D = G7["L"], G7["H"], G7["D"]
    with tempfile.TemporaryDirectory() as wd:
        write_ini(os.path.join(wd, "layer2.ini"), L, H, D, numstep=G7["numstep"], f_step=G7["f_step"],
                  BL_HIG=G7["BL_HIG"], nStages=G7["nStages"])
        r = subprocess.run([sys.executable, os.path.abspath(__file__), "--run-main", variant, wd],
                           stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=300)
        if r.returncode != 0:
            raise RuntimeError("reference main failed: " + r.stdout[-500:])
        assert os.path.exists(os.path.join(wd, "3.plt"))
        files["g7_%s_5.plt.gz" % variant] = gzip.compress(open(os.path.join(wd, "5.plt"), "rb").read(), mtime=0)
        raw = open(os.path.join(wd, "probes.dat"), "rb").read()
        assert len(raw) == G7["numstep"] * (4 + 6 * H * 8)
        files["g7_%s_probes.dat" % variant] = raw
        f32 = probes_to_f32(raw, H)
        files["g7_%s_probes_f32.dat" % variant] = f32
    with tempfile.TemporaryDirectory() as wd:
        open(os.path.join(wd, "probes.dat"), "wb").write(f32)
        # stepEnd below the last record: the post-processor ends its loop on the
        # first record past stepEnd (it has no end-of-file test)
        r = subprocess.run([os.path.join(REF_DIR, "layer2_p")], cwd=wd, input="2\n4\n%d\n" % H,
                           stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=60)
        if r.returncode != 0:
            raise RuntimeError("layer2_p failed: " + r.stdout[-500:])
        for q in (1, 2, 3):
            files["g7_%s_probe%d.dat" % (variant, q)] = open(os.path.join(wd, "probe%d.dat" % q), "rb").read()
    return files


def main():
    if len(sys.argv) == 4 and sys.argv[1] == "--run-main":
        return run_reference_main(sys.argv[2], sys.argv[3])
    if not os.path.isdir(REF_ROOT):
        print("reference not present; nothing to do")
        return 1
    gens = {
        "g1_ref_cs": lambda: g1("cs"), "g1_ref_dyn": lambda: g1("dyn"),
        "g2_rand_cs": lambda: g2("cs"), "g2_rand_dyn": lambda: g2("dyn"),
        "g3_long_cs": lambda: g3("cs"), "g3_long_dyn": lambda: g3("dyn"),
        "g4_lcg": g4, "g5_filter": g5, "g6_nstages2": g6,
    }
    only = sys.argv[1:]
    manifest_path = os.path.join(HERE, "MANIFEST.json")
    manifest = json.load(open(manifest_path)) if os.path.exists(manifest_path) else {}
    for name, fn in gens.items():
        if only and name not in only:
            continue
        data = fn()
        path = os.path.join(HERE, name + ".npz")
        np.savez_compressed(path, **data)
        manifest[name + ".npz"] = {"sha256": hashlib.sha256(open(path, "rb").read()).hexdigest(),
                                   "keys": sorted(data.keys())}
        print("wrote", path, os.path.getsize(path))
    for variant in ("cs", "dyn"):
        if only and "g7_" + variant not in only:
            continue
        for name, data in g7(variant).items():
            open(os.path.join(HERE, name), "wb").write(data)
            manifest[name] = {"sha256": hashlib.sha256(data).hexdigest(), "case": G7, "seed": X0}
            print("wrote", name, len(data))
    for src, dst in (("layer2.ini", "case_layer2.ini"), ("mpi_layer2.bin", "case_mpi_layer2.bin"),
                     ("src-par/ini2bin/mpi_layer2.ini", "case_ini2bin.ini"),
                     ("src-par/ini2bin/mpi_layer2.bin", "case_ini2bin.bin"), ("file", "case_seed.txt")):
        shutil.copyfile(os.path.join(REF_ROOT, src), os.path.join(HERE, dst))
        manifest[dst] = {"sha256": hashlib.sha256(open(os.path.join(HERE, dst), "rb").read()).hexdigest(),
                         "source": src}
    json.dump(manifest, open(manifest_path, "w"), indent=1, sort_keys=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
