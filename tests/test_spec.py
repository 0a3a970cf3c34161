"""The problem specification parser (stfem_spec_normalize, CPU only): the
reference's parameter files (tests/json/*.json of nlsmrg/dealii-stfem, kept
as fixtures under tests/golden/json/) are accepted as they are, their
camelCase keys (Parameters::parse, tests/include/gmg.h:591-649) map onto this
library's keys, and anything unknown or malformed is an error."""
import json
import os

import pytest

import stfem_amd as st

HERE = os.path.dirname(os.path.abspath(__file__))
JSON = os.path.join(HERE, "golden", "json")


def norm(spec):
    return st.spec_normalize(spec)


def err(spec):
    with pytest.raises(st.StfemError) as e:
        st.spec_normalize(spec)
    return e.value.code, str(e.value)


@pytest.mark.parametrize("name", [f"tf0{i}" for i in range(1, 9)])
def test_reference_parameter_files(name):
    text = open(os.path.join(JSON, name + ".json")).read()
    ref = json.loads(text)
    n = norm(text)
    assert n["type"] == ref["timeType"]
    assert n["problem"] == ref["problemType"]
    assert int(n["nts"]) == int(ref["nTimestepsAtOnce"])
    assert int(n["r"]) == int(ref["feDegree"])
    assert int(n["p"]) == int(ref["feDegree"]) + 1  # FE_Q(fe_degree + 1), tp_01.cc:77
    assert int(n["refine"]) == int(ref["refinement"])
    assert int(n["space_time_mg"]) == 1 and int(n["time_before_space"]) == 0
    # Parameters::parse derivations (gmg.h:637-648)
    assert int(n["nts_min"]) == max(1, int(ref["nTimestepsAtOnce"]) // 2)
    lowest = 0 if ref["timeType"] == "DG" else 1
    assert int(n["r_min"]) == max(lowest, int(ref["feDegree"]) - 1)
    # the same spec written with this library's keys
    mine = dict(type=ref["timeType"], problem=ref["problemType"], nts=ref["nTimestepsAtOnce"],
                r=ref["feDegree"], refine=ref["refinement"], frequency=ref["frequency"])
    assert norm(mine) == n


@pytest.mark.parametrize("name", ["practical01", "practical02"])
def test_practical_parameter_files_are_reported_unsupported(name):
    """spaceTimeConvergenceTest=false selects the region coefficient of
    operators.h:280-353, which this library does not build: a clear
    STFEM_ERR_UNSUPPORTED, not a silently different problem."""
    code, msg = err(open(os.path.join(JSON, name + ".json")).read())
    assert code == -4 and "spaceTimeConvergenceTest" in msg


def test_camel_case_and_cli_keys():
    n = norm("timeType=CGP;problemType=heat;feDegree=2;feDegreeMin=1;nTimestepsAtOnce=4;"
             "nTimestepsAtOnceMin=2;refinement=3;smoothingSteps=2;estimateRelaxation=false;"
             "variable=false;extrapolate=false;endTime=0.5;subdivisions=2,3;"
             "hyperRectLowerLeft=-1,-1;hyperRectUpperRight=1,2;mgTimeBeforeSpace=true;"
             "spaceTimeMg=true;restrictIsTransposeProlongate=false;precondition_float=1;dim=2;"
             "doOutput=true;printTiming=true;nDegCycles=3;nRefCycles=4;functionalFile=x.txt;"
             "sourcePoint=0,0;smoothingDegree=5;coarseGridSmootherType=Smoother;"
             "coarseGridMaxiter=10;coarseGridAbstol=1e-20;coarseGridReltol=1e-4;"
             "distortGrid=0.0;distortCoeff=0;spaceTimeConvergenceTest=true")
    assert n["type"] == "CGP" and n["problem"] == "heat"
    assert (n["r"], n["r_min"], n["p"], n["nts"], n["nts_min"]) == ("2", "1", "3", "4", "2")
    assert n["refine"] == "3" and n["smoothing_steps"] == "2"
    assert n["estimate_relaxation"] == "0" and n["variable"] == "0" and n["extrapolate"] == "0"
    assert float(n["end_time"]) == 0.5 and n["subdiv"].startswith("2,3,")
    assert n["lo"].startswith("-1,-1,") and n["hi"].startswith("1,2,")
    assert n["time_before_space"] == "1" and n["restrict_is_transpose"] == "0"
    assert n["mg_float"] == "1" and n["print_timing"] == "1"


def test_round_trip():
    n = norm(dict(dim=3, refine=2, p=4, type=2, r=3, nts=1, tau=0.25, source=3, seed=7,
                  reltol=1e-8, basis=60, coarse_coef=3, use_graph=-1))
    assert norm(";".join(f"{k}={v}" for k, v in n.items())) == n


@pytest.mark.parametrize("spec,needle", [
    ("dim=2;smoothing_step=2", "unknown key 'smoothing_step'"),
    ("dim=2;vanka_invert_T=0", "unknown key 'vanka_invert_T'"),
    ("dim=2;refine=2x", "needs an integer"),
    ("dim=2;reltol=abc", "needs a finite number"),
    ("dim=2;extrapolate=maybe", "needs true/false"),
    ("dim=2;refine", "is not key=value"),
    ("refinement=2;refine=3", "given twice"),
    ("dim=2;type=XY", "DG or CGP"),
    ("dim=2;problem=fluid", "heat or wave"),
    ("dim=2;use_graph=2", "use_graph"),
    ("dim=2;source=4", "source"),
    ("dim=2;subdiv=1,1,1,1", "more than 3"),
    ("{\"refinement\": [1, 2]}", "string, number or boolean"),
    ("{\"refinement\": \"2\" \"x\": 1}", "expected ','"),
])
def test_rejected_specs(spec, needle):
    code, msg = err(spec)
    assert code == -1 and needle in msg, msg


@pytest.mark.parametrize("spec,needle", [
    ("coarseGridSmootherType=GMRES", "coarseGridSmootherType"),
    ("distortGrid=0.25", "distortGrid"),
    ("distortCoeff=0.5", "distortCoeff"),
    ("timeType=GCC", "GCC"),
])
def test_unsupported_specs(spec, needle):
    code, msg = err(spec)
    assert code == -4 and needle in msg, msg
