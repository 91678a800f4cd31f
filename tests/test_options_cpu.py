"""The options parser (mc3d_options_check, host only) on decks in the
reference's format: the Samtaney D4 deck of the reference's HIT test
(tests/golden/samtaney_d4.in, a copy of tests/hit/Samtaney_D4/MPI_3DCompact.in),
which as shipped lacks the [LES] section the parser requires (src/Options.cpp:6-20,
the section read at src/Options.hpp:419-426), its values once that section is
added, spacing and comment variants of the boost config-file syntax, and the
reference's abort paths (mandatory keys, unknown enum spellings, face
mismatches, bcValidation src/Options.cpp:178-268, grid size src/Options.hpp:241),
each of which must come back as MC3D_ECONFIG instead of aborting.
"""
import os

import pytest

from mpi_3dcompact_amd import hit
from mpi_3dcompact_amd._lib import Mc3dError
from mpi_3dcompact_amd.solver import options_check

HERE = os.path.dirname(os.path.abspath(__file__))
LES = "\n[LES]\nLESMODEL = NONE\nLESAVERAGING = NONE\n"


def deck():
    with open(os.path.join(HERE, "golden", "samtaney_d4.in")) as f:
        return f.read()


def econfig(text, match):
    with pytest.raises(Mc3dError, match=match) as e:
        options_check(text)
    assert "error -5" in str(e.value), str(e.value)


def test_shipped_deck_lacks_the_les_section():
    econfig(deck(), "LES.LESMODEL not specified")


def test_samtaney_deck_values():
    v = options_check(deck() + LES)
    want = {
        "RESTART.FROMRESTART": "FALSE", "RESTART.ONLYGRIDFROMRESTART": "FALSE",
        "RESTART.SPONGEAVGFROMRESTART": "FALSE",
        "TIMESTEPPING.TSTYPE": "CONST_CFL", "TIMESTEPPING.CFL": "0.5", "TIMESTEPPING.MAXTIME": "10",
        "TIMESTEPPING.MAXTIMESTEP": "750000", "TIMESTEPPING.FILTERSTEP": "1",
        "TIMESTEPPING.SUBSTEPFILTERING": "FALSE", "TIMESTEPPING.CHECKSTEP": "1", "TIMESTEPPING.DUMPSTEP": "25000",
        "STATS.INTERVAL": "1", "STATS.VELOCITYSTATS": "FALSE", "STATS.THERMOSTATS": "FALSE",
        "STATS.STATSAVGTYPE": "XI3_AVG",
        "BC.BCXTYPE": "PERIODIC_SOLVE", "BC.BCYTYPE": "PERIODIC_SOLVE", "BC.BCZTYPE": "PERIODIC_SOLVE",
        "BC.BCY0": "PERIODIC", "BC.PERIODICDISPX_X": "6.2831853", "BC.PERIODICDISPY_Y": "6.2831853",
        "BC.PERIODICDISPZ_Z": "6.2831853", "BC.PERIODICDISPX_Y": "0",
        # no face is a sponge: the [SPONGE] section is not read (src/Options.hpp:312-362)
        "SPONGE.KIND": "RECTILINEAR", "SPONGE.STRENGTH": "1",
        "DOMAIN.NX": "128", "DOMAIN.NY": "128", "DOMAIN.NZ": "128",
        "PENCILDECOMP.PROW": "0", "PENCILDECOMP.PCOL": "0",
        "SOLVER.RKTYPE": "LSLDDRK4", "SOLVER.ALPHAF": "0.495", "SOLVER.MU_REF": "0.00033",
        "SOLVER.USETIMING": "FALSE", "SOLVER.XFDTYPE": "PADE6", "SOLVER.YFDTYPE": "PADE6",
        "SOLVER.ZFDTYPE": "PADE6", "SOLVER.FILTERTYPE": "COMPACT10",
        # NONE is accepted as an alias of NOMODEL (the README's spelling)
        "LES.LESMODEL": "NOMODEL", "LES.LESAVERAGING": "NONE",
        # defaults of the keys the deck leaves out
        "LES.DSMTESTFILTER": "1", "LES.DSMUSETAUKK": "TRUE", "LES.DSMCOEFFRANGE": "FALSE",
    }
    for k, w in want.items():
        try:  # numbers are written with 17 significant digits
            assert float(v[k]) == float(w), (k, v[k], w)
        except ValueError:
            assert v[k] == w, (k, v[k], w)


def test_spacing_and_comment_variants():
    """Tabs around '=', trailing blanks, comments after values and on their
    own lines, blank lines and section headers with inner blanks parse like
    the plain deck."""
    plain = options_check(deck() + LES)
    lines = []
    for ln in (deck() + LES).splitlines():
        s = ln.strip()
        if "=" in s and not s.startswith("#"):
            k, _, val = s.partition("=")
            ln = f"\t{k.strip()}\t =  {val.strip()}   # trailing comment"
        elif s.startswith("["):
            ln = "[ " + s[1:-1] + " ]  "
        lines.append(ln)
        lines.append("")
        lines.append("   # a comment line")
    assert options_check("\n".join(lines)) == plain


def test_booleans_in_every_spelling():
    base = hit.options_text(32, rk="RK4")
    for word, want in (("TRUE", "TRUE"), ("true", "TRUE"), ("1", "TRUE"), ("yes", "TRUE"), ("on", "TRUE"),
                       ("FALSE", "FALSE"), ("0", "FALSE"), ("no", "FALSE"), ("off", "FALSE")):
        t = base.replace("SUBSTEPFILTERING = FALSE", f"SUBSTEPFILTERING = {word}")
        assert "SUBSTEPFILTERING" in base
        assert options_check(t)["TIMESTEPPING.SUBSTEPFILTERING"] == want, word
    econfig(base.replace("SUBSTEPFILTERING = FALSE", "SUBSTEPFILTERING = maybe"), "invalid boolean")


def _drop(text, key):
    out = [ln for ln in text.splitlines() if ln.split("=")[0].strip() != key]
    assert len(out) < len(text.splitlines()), key
    return "\n".join(out) + "\n"


@pytest.mark.parametrize("key,msg", [
    ("MU_REF", "SOLVER.MU_REF is mandatory"),
    ("RKTYPE", "SOLVER.RKTYPE not specified"),           # the shipped src/MPI_3DCompact.in lacks it (T15)
    ("XFDTYPE", "SOLVER.XFDTYPE not specified"),
    ("FILTERTYPE", "SOLVER.FILTERTYPE not specified"),
    ("TSTYPE", "TIMESTEPPING.TSTYPE not specified"),
    ("BCY1", "BC.BCY1 not specified"),
    ("PERIODICDISPZ_Y", "BC.PERIODICDISPZ_Y is mandatory"),
    ("FROMRESTART", "RESTART.FROMRESTART is mandatory"),
    ("NZ", "DOMAIN.NZ is mandatory"),
    ("LESAVERAGING", "LES.LESAVERAGING not specified"),
    ("STATSAVGTYPE", "STATS.STATSAVGTYPE not specified"),
])
def test_mandatory_keys(key, msg):
    econfig(_drop(deck() + LES, key), msg)


@pytest.mark.parametrize("key,bad", [("RKTYPE", "RK5"), ("XFDTYPE", "PADE8"), ("FILTERTYPE", "COMPACT6"),
                                     ("TSTYPE", "CONST_STEP"), ("BCXTYPE", "PERIODIC"), ("BCZ0", "WALL"),
                                     ("STATSAVGTYPE", "XI4_AVG")])
def test_unknown_enum_spellings(key, bad):
    text = "\n".join(f"{key} = {bad}" if ln.split("=")[0].strip() == key else ln
                     for ln in (deck() + LES).splitlines())
    econfig(text, f"unknown value for .*{key}: '{bad}'")


def test_les_model_spellings():
    for model in ("NOMODEL", "VREMAN", "DSM", "LEARNED"):
        assert options_check(deck() + LES.replace("NONE\nLESAV", f"{model}\nLESAV"))["LES.LESMODEL"] == model
    econfig(deck() + LES.replace("NONE\nLESAV", "SMAGORINSKY\nLESAV"), "unknown value for LES.LESMODEL")
    # enum spellings are case-sensitive, as the reference's strcmp
    econfig(deck() + LES.replace("NONE\nLESAV", "vreman\nLESAV"), "unknown value for LES.LESMODEL")


def test_validation_paths():
    d = deck() + LES
    # a periodic face against a non-periodic one (src/Options.hpp:276-303)
    econfig(d.replace("BCY1 = PERIODIC", "BCY1 = ADIABATIC_WALL"), "BC mismatch for Y0-Y1")
    # a periodic solve with wall faces (bcValidation)
    t = d.replace("BCZ0 = PERIODIC", "BCZ0 = ADIABATIC_WALL").replace("BCZ1 = PERIODIC", "BCZ1 = ADIABATIC_WALL")
    econfig(t, "periodic solve in Z needs periodic faces")
    # periodic faces under a Dirichlet solve
    econfig(d.replace("BCXTYPE = PERIODIC_SOLVE", "BCXTYPE = DIRICHLET_SOLVE"),
            "periodic face kinds in a Dirichlet solve of X")
    # CONST_DT needs DT
    econfig(d.replace("TSTYPE\t    = CONST_CFL", "TSTYPE = CONST_DT"), "TIMESTEPPING.DT is mandatory")
    assert options_check(d.replace("TSTYPE\t    = CONST_CFL", "TSTYPE = CONST_DT").replace(
        "#DT\t    = 0.001", "DT = 0.001"))["TIMESTEPPING.DT"] == "0.001"
    # grid sizes above 10 (src/Options.hpp:241)
    econfig(d.replace("NX = 128", "NX = 10"), "greater than 10")
    # restart with both flags
    econfig(d.replace("ONLYGRIDFROMRESTART = FALSE", "ONLYGRIDFROMRESTART = TRUE").replace(
        "FROMRESTART = FALSE\nONLY", "FROMRESTART = TRUE\nONLY"), "conflict")
    # numbers and syntax
    econfig(d.replace("CFL \t    = 0.5", "CFL = 0.5x"), "invalid number for TIMESTEPPING.CFL")
    econfig(d.replace("NY = 128", "NY = 12.5"), "invalid integer for DOMAIN.NY")
    econfig(d.replace("NX = 128", "NX = 128\nNX = 64"), "duplicate key DOMAIN.NX")
    econfig(d + "[LES\n", "malformed section header")
    econfig(d + "JUSTAWORD\n", "expected KEY = VALUE")


def test_sponge_section_is_read_with_a_sponge_face():
    d = deck() + LES
    t = (d.replace("BCYTYPE = PERIODIC_SOLVE", "BCYTYPE = DIRICHLET_SOLVE")
         .replace("BCY0 = PERIODIC ", "BCY0 = ADIABATIC_WALL").replace("BCY1 = PERIODIC", "BCY1 = SPONGE"))
    v = options_check(t)
    assert (v["SPONGE.KIND"], float(v["SPONGE.STRENGTH"]), float(v["SPONGE.RMIN"])) == ("CYLINDRICAL", 0.15, 15.0)
    econfig(t.replace("RMIN = 15", ""), "SPONGE.RMIN is mandatory")
    econfig(t.replace("CYLAXIS_ORIENT = 2", "CYLAXIS_ORIENT = 3"), "orientation must be 0, 1 or 2")


def test_options_text_of_the_hit_cases_parses():
    for kw in (dict(), dict(fd="PENTA10", filt="COMPACT8", rk="KENRK4", les="VREMAN", les_avg="LOCAL"),
               dict(les="DSM", les_avg="XI1_AVG", coeff_range=True), dict(rk="LSLDDRK4", ts_type="CONST_DT",
                                                                           dt=1e-3)):
        v = options_check(hit.options_text(64, **kw))
        assert v["DOMAIN.NX"] == "64"


def test_closure_features_option():
    """CLOSUREFEATURES (no counterpart in the reference): AVERAGED when absent,
    AVERAGED or POINT when given, anything else refused."""
    d = deck() + LES.replace("NONE\nLESAV", "LEARNED\nLESAV")
    assert options_check(d)["LES.CLOSUREFEATURES"] == "AVERAGED"
    for v in ("AVERAGED", "POINT"):
        assert options_check(d + f"CLOSUREFEATURES = {v}\n")["LES.CLOSUREFEATURES"] == v
    econfig(d + "CLOSUREFEATURES = FILTERED\n", "unknown value for LES.CLOSUREFEATURES")
