"""The per-solve results line against the reference's format and its reader.

Golden lines: the fprintf sequence of File_output_gen / File_output_smv /
File_output (GAGA/gaga/functions.cu:34-80) applied by hand to exactly
representable values.  Reader: the whitespace textscan of
CGIHTcode/code/process_{gen,smv,dct}_data.m:51 and the column lists those
scripts keep (process_gen_data.m:66-86 and the smv/dct twins), so a line this
build writes lands every field in the column the reference's scripts read.
No GPU needed: the line is host code in the library.
"""
import datetime
import os
import re

import numpy as np
import pytest

NORMS = np.array([1.25e-07, 2.5e-07, 5.0e-07], np.float32)
TIMES = np.array([1.5, 0.1015625, 2.5], np.float32)
SUP = np.array([20, 0, 1004, 0], np.int32)

GOLDEN = {
    "gen": ("CGIHT_S_gen output: k 20, m 256, n 1024, vecDistribution 1 matrixEnsemble 1 "
            "errorl1 1.2500000e-07 errorl2 2.5000000e-07 errorlinf 5.0000000e-07 "
            "timeALG 1.5000000e+00 timeIter 1.0156250e-01 timeTotal 2.5000000e+00 "
            "iterations 15 converge_rate 5.0000000e-01 "
            "TruePos 20 FalsePos 0 TrueNeg 1004 FalseNeg 0 seed 7 \n"),
    "smv": ("NIHT_S_smv output: k 20, m 256, n 1024, vecDistribution 2 nonzeros_per_column 7 "
            "matrixEnsemble 3 "
            "errorl1 1.2500000e-07 errorl2 2.5000000e-07 errorlinf 5.0000000e-07 "
            "timeALG 1.5000000e+00 timeIter 1.0156250e-01 timeTotal 2.5000000e+00 "
            "iterations 15 converge_rate 5.0000000e-01 "
            "TruePos 20 FalsePos 0 TrueNeg 1004 FalseNeg 0 seed 7 "
            "band_percentage 0.00 noise_level 0.0000 \n"),
    "dct": ("HTP_S_dct output: k 20, m 256, n 1024, vecDistribution 0 "
            "errorl1 1.2500000e-07 errorl2 2.5000000e-07 errorlinf 5.0000000e-07 "
            "timeALG 1.5000000e+00 timeIter 1.0156250e-01 timeTotal 2.5000000e+00 "
            "iterations 15 converge_rate 5.0000000e-01 "
            "TruePos 20 FalsePos 0 TrueNeg 1004 FalseNeg 0 seed 7 \n"),
}
ALG = {"gen": "CGIHT", "smv": "NIHT", "dct": "HTP"}
VEC = {"gen": 1, "smv": 2, "dct": 0}
MAT = {"gen": 1, "smv": 3, "dct": 0}

# column lists of the processing scripts (1-based, as in the scripts' comments)
COLUMNS = {
    "gen": dict(k=4, m=7, n=10, vec=13, mat=15, l1=17, l2=19, linf=21, talg=23, titer=25,
                ttot=27, iters=29, rate=31, tp=33, fp=35, tn=37, fn=39, seed=41),
    "smv": dict(k=4, m=7, n=10, vec=13, p=15, mat=17, l1=19, l2=21, linf=23, talg=25, titer=27,
                ttot=29, iters=31, rate=33, tp=35, fp=37, tn=39, fn=41, seed=43),
    "dct": dict(k=4, m=7, n=10, vec=13, l1=15, l2=17, linf=19, talg=21, titer=23, ttot=25,
                iters=27, rate=29, tp=31, fp=33, tn=35, fn=37, seed=39),
}
# the scripts' textscan format: three "key value," groups, the vecDistribution pair, then
# "key value" pairs up to the seed (41 fields for gen, 43 for smv, 39 for dct)
PAIRS = {"gen": 14, "smv": 15, "dct": 13}


def textscan(line, ens):
    """One record of MATLAB textscan(fid, FORMAT) with whitespace delimiters: a
    %f field takes the longest numeric prefix of the next token and leaves the
    rest (the commas after k, m, n) as the next token."""
    toks = line.split()
    out, i = [], 0
    num = re.compile(r"[-+]?(\d+\.?\d*|\.\d+)([eE][-+]?\d+)?")
    fmt = "%s %s %s %f %s %s %f %s %s %f %s %s %f " + " ".join(["%s %f"] * PAIRS[ens])
    for spec in fmt.split():
        tok = toks[i]
        if spec == "%f":
            mt = num.match(tok)
            assert mt, (spec, tok)
            out.append(float(mt.group(0)))
            rest = tok[mt.end():]
            if rest:
                toks[i] = rest
            else:
                i += 1
        else:
            out.append(tok)
            i += 1
    return out


def _opts(ga, ens):
    o = ga._capi.default_options(ALG[ens], ens, 1024)
    o.seed = 7
    o.vec_distribution = VEC[ens]
    o.matrix_ensemble = MAT[ens]
    return o


@pytest.mark.parametrize("ens", ["gen", "smv", "dct"])
def test_golden_line(ens, lib):
    import gaga_amd as ga
    o = _opts(ga, ens)
    line = ga.results_line(ALG[ens], ens, 20, 256, 1024, 7, o, NORMS, TIMES, 15, SUP, 0.5)
    assert line == GOLDEN[ens]


@pytest.mark.parametrize("ens", ["gen", "smv", "dct"])
def test_reference_reader_columns(ens, lib):
    import gaga_amd as ga
    o = _opts(ga, ens)
    rng = np.random.default_rng(3)
    norms = rng.random(3).astype(np.float32) * 1e-3
    times = (rng.random(3) * 100).astype(np.float32)
    sup = np.array([17, 3, 1001, 3], np.int32)
    line = ga.results_line(ALG[ens], ens, 20, 256, 1024, 7, o, norms, times, 123, sup, 0.875)
    rec = textscan(line, ens)
    col = COLUMNS[ens]
    assert rec[0] == "%s_S_%s" % (ALG[ens], ens)  # strcmp(results_full{1}, alg_list{j})
    assert (rec[col["k"] - 1], rec[col["m"] - 1], rec[col["n"] - 1]) == (20, 256, 1024)
    assert rec[col["vec"] - 1] == VEC[ens]
    if "mat" in col:
        assert rec[col["mat"] - 1] == MAT[ens]
    if "p" in col:
        assert rec[col["p"] - 1] == 7
    for key, v in zip(("l1", "l2", "linf"), norms):
        assert abs(rec[col[key] - 1] - v) <= 1e-7 * abs(v)  # %0.7e keeps 8 digits
    for key, v in zip(("talg", "titer", "ttot"), times):
        assert abs(rec[col[key] - 1] - v) <= 1e-7 * abs(v)
    assert rec[col["iters"] - 1] == 123 and rec[col["rate"] - 1] == 0.875
    assert [rec[col[key] - 1] for key in ("tp", "fp", "tn", "fn")] == list(sup)
    assert rec[col["seed"] - 1] == 7


def test_append_file_name_and_lines(tmp_path, lib):
    import gaga_amd as ga
    o = _opts(ga, "gen")
    for _ in range(2):
        ga._capi.results_append(tmp_path, "CGIHT", "gen", 20, 256, 1024, 0, o, NORMS, TIMES, 15,
                                SUP, 0.5)
    files = os.listdir(tmp_path)
    assert len(files) == 1
    today = datetime.date.today()
    assert files[0] in ("gpu_data_gen%04d%02d%02d.txt" % (d.year, d.month, d.day)
                        for d in (today, today - datetime.timedelta(days=1)))
    assert open(tmp_path / files[0]).read() == GOLDEN["gen"] * 2


def test_bad_arguments(lib):
    import gaga_amd as ga
    o = _opts(ga, "gen")
    with pytest.raises(ga.GagaError):
        ga.results_line("NoSuchAlg", "gen", 20, 256, 1024, 0, o, NORMS, TIMES, 15, SUP, 0.5)
    with pytest.raises(ga.GagaError):
        ga.results_line("CGIHT", "xyz", 20, 256, 1024, 0, o, NORMS, TIMES, 15, SUP, 0.5)


def test_smv_ensemble_names(lib):
    """matrixEnsemble names and codes of gaga_smv.cu:387-395 (default 3, :241)."""
    import gaga_amd as ga
    for name, code in (("ones", 1), ("binary", 2), ("expander", 3)):
        o = ga._apply_options("CGIHT", "smv", 1024, ga.options(matrixEnsemble=name))
        assert o.matrix_ensemble == code
    assert ga._apply_options("CGIHT", "smv", 1024, None).matrix_ensemble == 3


# ----------------------------------------------------- timing variants' text --
# File_output_timings_gen / _smv / File_output_timings (NIHT) and the _HTP twins
# (functions.cu:82-212), applied by hand to exactly representable values.
T_ITER = np.array([0.125, 0.25], np.float32)
T_SUPP = np.array([0.0625, 0.03125], np.float32)
CG = np.array([7, 15], np.int32)
T_CG = np.array([0.5, 0.75], np.float32)
HEAD = ("k 20, m 256, n 1024, vecDistribution {vec} "
        "errorl1 1.2500000e-07 errorl2 2.5000000e-07 errorlinf 5.0000000e-07 "
        "timeALG 1.5000000e+00 timeIter 1.0156250e-01 timeTotal 2.5000000e+00 "
        "iterations 2 converge_rate 5.0000000e-01 "
        "TruePos 20 FalsePos 0 TrueNeg 1004 FalseNeg 0 seed 7 alpha 2.5000000e-01 ")
TIMING_GOLDEN = {
    ("NIHT", "gen"): "NIHT_S_gen output: " + HEAD.format(vec=1) + "supp_flag 3 matrixEnsemble 1 \n"
    "iteration 0 time_iteration 1.2500000e-01 time_supp_set 6.2500000e-02\n"
    "iteration 1 time_iteration 2.5000000e-01 time_supp_set 3.1250000e-02\n",
    ("HTP", "smv"): "HTP_S_smv output: " + HEAD.format(vec=2)
    + "supp_flag 3 NonzerosPerCol 7 matrixEnsemble 3 \n"
    "iteration 0 time_iteration 1.2500000e-01 time_supp_set 6.2500000e-02 cg_per_iteration 7 "
    "time_for_cg 5.0000000e-01\n"
    "iteration 1 time_iteration 2.5000000e-01 time_supp_set 3.1250000e-02 cg_per_iteration 15 "
    "time_for_cg 7.5000000e-01\n",
    ("NIHT", "dct"): "NIHT_S_dct output: " + HEAD.format(vec=0) + "supp_flag 3 \n"
    "iteration 0 time_iteration 1.2500000e-01 time_supp_set 6.2500000e-02\n"
    "iteration 1 time_iteration 2.5000000e-01 time_supp_set 3.1250000e-02\n",
}


@pytest.mark.parametrize("alg,ens", list(TIMING_GOLDEN))
def test_timing_text_golden(alg, ens, lib):
    import gaga_amd as ga
    o = ga._capi.default_options(alg, ens, 1024)
    o.vec_distribution = VEC[ens]
    o.matrix_ensemble = MAT[ens]
    o.seed = 7
    htp = alg == "HTP"
    text = ga._capi.results_timing_text(alg, ens, 20, 256, 1024, 7, o, NORMS, TIMES, 2, SUP, 0.5,
                                        0.25, 3, T_ITER, T_SUPP, CG if htp else None,
                                        T_CG if htp else None)
    assert text == TIMING_GOLDEN[(alg, ens)]


def test_timing_append_file(tmp_path, lib):
    import gaga_amd as ga
    o = ga._capi.default_options("NIHT", "gen", 1024)
    o.vec_distribution, o.matrix_ensemble, o.seed = 1, 1, 7
    for _ in range(2):
        ga._capi.results_timing_append(tmp_path, "NIHT", "gen", 20, 256, 1024, 0, o, NORMS, TIMES,
                                       2, SUP, 0.5, 0.25, 3, T_ITER, T_SUPP)
    name = "gpu_timings_data_gen%s.txt" % datetime.date.today().strftime("%Y%m%d")
    assert os.listdir(tmp_path) == [name]
    assert open(tmp_path / name).read() == TIMING_GOLDEN[("NIHT", "gen")] * 2


def test_timing_text_errors(lib):
    import gaga_amd as ga
    o = ga._capi.default_options("NIHT", "gen", 1024)
    with pytest.raises(ga.GagaError):  # only NIHT and HTP have timing variants
        ga._capi.results_timing_text("CGIHT", "gen", 20, 256, 1024, 0, o, NORMS, TIMES, 2, SUP,
                                     0.5, 0.25, 0, T_ITER, T_SUPP)
    with pytest.raises(ga.GagaError):  # HTP needs its CG arrays
        ga._capi.results_timing_text("HTP", "gen", 20, 256, 1024, 0, o, NORMS, TIMES, 2, SUP,
                                     0.5, 0.25, 0, T_ITER, T_SUPP)
