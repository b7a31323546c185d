"""Link check of the drop-in recipe: the reference's own main
(src/layer2.c) with the three calls the integration guide adds, compiled
with its non-hot-path files (initialize.c, output.c, probes.c, finalize.c)
and linked against libfsl_layer2.so instead of the hot-path objects.  Only
where the reference sources exist (the build container); the patched copy
of layer2.c lives in the test's temporary directory.  Link only: running it
needs a GPU."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = "/root/reference/src"
LIB = os.path.join(ROOT, "freeshearlayer_amd", "lib")

ATTACH = """
    {   /* hand the case and the state to the GPU */
        fsl_case c = { LEN, HIG, DEP, deltaX, deltaY, deltaZ, deltaT, Answer, numstep,
                       mu_L, Pr_L, Pr_T, Cs, BL_HIG, Ua, Va, Wa, Ns_min, Nst, f_step,
                       nStages, maxCoNum };
        fsl_config cfg;
        fsl_config_defaults(&cfg);
        if (fsl_layer2_attach(&c, &cfg, U1, U2, U3, U4, U5, mu_SGS, &deltaT) != 0) exit(-1);
    }
"""


def patch_layer2(src):
    """The integration guide's three additions to src/layer2.c."""
    def once(s, old, new):
        assert s.count(old) == 1, old
        return s.replace(old, new)
    s = once(src, '#include "finalize.h"', '#include "finalize.h"\n#include "fsl_layer2.h"')
    s = once(s, "\tInitialize();", "\tInitialize();" + ATTACH)
    s = once(s, "\t\tfor( Stage = 1; Stage <= nStages; Stage++){",
             "\t\tfsl_layer2_set_block(k_min, k_max, Ud, Vd, Wd);\n"
             "\t\tfor( Stage = 1; Stage <= nStages; Stage++){")
    s = once(s, "\t\tProbes();", "\t\tfsl_layer2_sync_host();\n\t\tProbes();")
    s = once(s, "\tFinalize();", "\tFinalize();\n\tfsl_layer2_detach();")
    return s


@pytest.mark.skipif(not os.path.isdir(REF), reason="reference sources not present")
def test_reference_main_links_against_the_drop_in(tmp_path):
    src = open(os.path.join(REF, "layer2.c")).read()
    (tmp_path / "layer2.c").write_text(patch_layer2(src))
    exe = tmp_path / "layer2_gpu"
    cmd = ["gcc", "-O2", "-DTYPE_H", "-Dreal=double", "-I" + os.path.join(ROOT, "include"), "-I" + REF,
           str(tmp_path / "layer2.c")] + [os.path.join(REF, f) for f in
                                          ("initialize.c", "output.c", "probes.c", "finalize.c")] + \
          ["-L" + LIB, "-lfsl_layer2", "-lfsl", "-Wl,-rpath," + LIB, "-lm", "-o", str(exe)]
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=120)
    assert r.returncode == 0, r.stdout
    assert os.access(str(exe), os.X_OK)
    # every hot-path name the reference's main calls resolves in the drop-in library
    nm = subprocess.run(["nm", "-D", "--defined-only", os.path.join(LIB, "libfsl_layer2.so")],
                        stdout=subprocess.PIPE, text=True, check=True).stdout
    for name in ("BounCondInGhostCells", "Reconstruction", "BounCondOnInterfaces", "Fluxes", "Evolution",
                 "checkCoNum", "Random", "Array3D", "free3D", "fsl_layer2_attach", "fsl_layer2_set_block",
                 "fsl_layer2_sync_host"):
        assert (" T %s\n" % name) in nm, name
