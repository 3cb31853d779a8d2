"""pg3gpu's command line without a GPU: the reference's default output file names
(Config::DefaultFilename, config.hxx:172-264, and Scene::GetSceneName's acronym,
scene.hxx:1035-1273), the output path, help and argument errors.  The -opof / -opop modes print
the name and exit before any GPU work.

Expected names: file names that the reference's own test scripts (testing/*.sh) use for their
reference images, and names composed by the same rules for other settings."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PG3GPU = os.path.join(ROOT, "pg3render_amd", "pg3gpu")


def run(*args):
    assert os.path.exists(PG3GPU), "pg3gpu is not built (make -C pg3render_amd/csrc)"
    r = subprocess.run([PG3GPU, *args], capture_output=True, text=True, timeout=60)
    return r.returncode, r.stdout


@pytest.mark.parametrize("args,name", [
    # names used by the reference's test scripts
    ("-s 20 -em 4 -a dlsa -i 72000 -e hdr -ot Reference", "1s_e4_Spd_dlsa_72000s_Reference.hdr"),
    ("-s 22 -em 10 -a dmis -i 24000 -e hdr -ot Reference", "1s_e10_Spdg_dmis_24000s_Reference.hdr"),
    ("-s 20 -em 12 -a dlsa -i 84000 -e hdr -ot Reference", "1s_e12_Spd_dlsa_84000s_Reference.hdr"),
    ("-s 22 -em 11 -a dmis -i 156000 -e hdr -ot Reference", "1s_e11_Spdg_dmis_156000s_Reference.hdr"),
    # the reference's defaults (-a pt, -maxpl 0, -sb 4, -iic 16, -i 1, bmp)
    ("-s 2", "w2s_c_SpdWpd_pt_rr_splt4,1.0,1.0_iic16.0_1s.bmp"),
    ("", "w2s_p_SpdWpd_pt_rr_splt4,1.0,1.0_iic16.0_1s.bmp"),
    # path length, splitting, clipping, time, trail and extension fields
    ("-s 33 -maxpl 100 -sb 1 -iic 0 -t 180 -e hdr", "w1s_b_SmgdWpd_pt_pl1-100_splt1,1.0,1.0_180sec.hdr"),
    ("-s 18 -a ptn -minpl 2 -maxpl 8 -i 64", "w2s_c_SffcWpd_ptn_pl2-8_64s.bmp"),
    ("-s 12 -em 3 -sb 16 -sl 0.5 -slbr 2 -iic 4 -t 5 -ot a -ot b", "w2s_bpe3_SpdWpd_pt_rr_splt16,0.5,2.0_iic4.0_05sec_ab.bmp"),
    ("-s 28 -a dbs -e jpg", "vrdr_e0_Rffd_dbs_1s.bmp"),
])
def test_default_output_names(args, name):
    rc, out = run(*args.split(), "-opof")
    assert rc == 1 and out.endswith(name) and (out == name or "Warning" in out), out


def test_output_path_and_custom_name():
    assert run("-s", "2", "-od", "out", "-o", "img", "-opop") == (1, "out/img.bmp")
    assert run("-s", "2", "-od", "out", "-o", "img.hdr", "-opof") == (1, "img.hdr")
    rc, out = run("-s", "2", "-od", "dir", "-opop")
    assert rc == 1 and out == "dir/w2s_c_SpdWpd_pt_rr_splt4,1.0,1.0_iic16.0_1s.bmp"


def test_help_and_argument_errors():
    rc, out = run("-h")
    assert rc == 1 and "Usage:" in out and "-slbr" in out
    for bad in (["-sl", "1.5"], ["-slbr", "0"], ["-sb", "0"], ["-i", "0"], ["-j", "0"], ["-t", "-1"], ["-a", "xyz"],
                ["-bogus"], ["-s"]):
        rc, out = run(*bad, "-opof")
        assert rc == 1 and "Error" in out, (bad, out)
