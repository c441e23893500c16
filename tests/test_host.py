"""Host side without a GPU: the C-ABI library loads and exports what
include/tpic_mi.h declares, and the tpic_mi_run driver's deck parser, run
admission and command line behave as the reference's
(src/config/simulation_parameters.cpp, src/config/physics.cpp,
src/app/main.cpp).

The parser cases restate the reference's own parser tests
(tests/config/simulation_parameters_tests.cpp: the base deck of
InputFixture::base, 15 lines for one species, and the error fragments each
case expects); --print-input exposes the parsed deck, every key with its
resolved value."""
import ctypes
import os
import re
import subprocess

import pytest

from tpic_amd import runtime as R

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
# TPIC_MI_RUN selects another build of the driver (the sanitized one).
RUN = os.environ.get("TPIC_MI_RUN") or os.path.join(ROOT, "tpic_amd", "lib", "tpic_mi_run")
HEADER = os.path.join(ROOT, "include", "tpic_mi.h")
DECKS = os.path.join(ROOT, "tests", "decks")


def header_functions():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(tpic_mi_\w+)\s*\(", text, re.M)))


def test_library_loads_and_exports_every_declared_function():
    lib = ctypes.CDLL(R.LIB_PATH)
    declared = header_functions()
    assert len(declared) >= 25
    for name in declared:
        assert hasattr(lib, name), name
    assert sorted(R.EXPORTS) == declared


@pytest.fixture(scope="module")
def driver():
    if not os.path.exists(RUN):
        pytest.fail("tpic_mi_run is not built: run __graft_entry__.build()")
    return RUN


def run(driver, *args):
    r = subprocess.run([driver, *args], capture_output=True, text=True, timeout=60)
    return r.returncode, r.stdout, r.stderr


def base(species=1):
    text = "dt = 0.125\nncycles = 12\nCase = Uniform\nns = %d\n" % species
    for key in ("rhoINIT", "qom", "npcelx", "npcely", "npcelz", "uth", "vth", "wth", "u0", "v0", "w0"):
        text += key + " =" + "".join(" %d" % (i + 1) for i in range(species)) + "\n"
    return text


class Deck:
    def __init__(self, driver, tmp_path):
        self.driver = driver
        self.path = str(tmp_path / "inputfile")

    def read(self, text):
        with open(self.path, "w") as f:
            f.write(text)
        rc, out, err = run(self.driver, "--print-input", self.path)
        if rc != 0:
            raise ValueError(err.strip())
        values, unknown = {}, []
        for line in out.splitlines():
            if line.startswith("unknown: "):
                unknown.append(line[len("unknown: "):])
            else:
                key, value = line.split(" = ", 1) if " = " in line else (line.rstrip(" ="), "")
                values[key] = value
        return values, unknown

    def with_(self, text):
        return self.read(base() + text)

    def error(self, text, fragment, whole=False):
        with pytest.raises(ValueError) as e:
            self.read(text if whole else base() + text)
        assert fragment in str(e.value), (fragment, str(e.value))


@pytest.fixture
def deck(driver, tmp_path):
    return Deck(driver, tmp_path)


def test_parser_defaults(deck):
    p, unknown = deck.with_("")
    assert unknown == []
    expect = {
        "dt": 0.125, "ncycles": 12, "th": 1.0, "Smooth": 1.0, "SmoothNiter": 6, "c": 1, "B0x": 0, "B0y": 0,
        "B0z": 0, "delta": 0.5, "HarrisB0x": 0, "ForceFreeB0x": 0, "ForceFreeB0z": 0, "pertGEM": 0,
        "pertHump": 0, "deltaxHump": 8, "deltayHump": 4, "kxHump": -1, "kyHump": -1, "Lx": 10, "Ly": 10,
        "Lz": 10, "nxc": 64, "nyc": 64, "nzc": 64, "XLEN": 1, "YLEN": 1, "ZLEN": 1, "FluxRopeB0": 0.1,
        "FluxRopeWidth": 1, "FluxRopeCenterX": 5, "FluxRopeCenterY": 5, "FluxRopePressureFloor": 9e-4,
        "FluxRopeElectronPressureFraction": 0.1, "GaussianExpansionCenterX": 5, "GaussianExpansionCenterY": 5,
        "GaussianExpansionCenterZ": 5, "GaussianExpansionSigmaX": 0, "GaussianExpansionSigmaY": 0,
        "GaussianExpansionSigmaZ": 0, "GaussianExpansionCutoff": 0, "WhistlerTurbulenceEpsilon": 0,
        "WhistlerTurbulenceSeed": 1, "WhistlerNoiseSeed": 12345, "WhistlerNoiseModeMax": 0,
        "WhistlerFieldTimeOffset": 0, "WhistlerParticleTimeOffset": 0, "TaylorGreenVortexModeX": 1,
        "TaylorGreenVortexModeY": 1, "TaylorGreenVortexModeZ": 1, "TaylorGreenVortexVelocityAmplitude": 0,
        "TaylorGreenVortexMagneticAmplitude": 0, "ns": 1, "nsTestPart": 0, "GMREStol": 1e-3,
        "FieldOutputCycle": 100, "ParticlesOutputCycle": 0, "TestParticlesOutputCycle": 0,
        "RestartOutputCycle": 0, "DiagnosticsOutputCycle": 100, "DiagnosticEnergyCycle": 1,
        "DiagnosticGaussCycle": 0,
    }
    for k, v in expect.items():
        assert float(p[k]) == v, k
    for k in ("currentFromAmpere", "spatiallyVaryingThermal", "QuietStart", "TaylorGreenVortexSeedMotionalE"):
        assert p[k] == "false", k
    for k in ("PERIODICX", "PERIODICY", "PERIODICZ", "PERIODICX_P", "PERIODICY_P", "PERIODICZ_P",
              "WhistlerUseExactColdRoot", "CallFinalize"):
        assert p[k] == "true", k
    assert p["ECSimSmoothing"] == "none" and p["ECSimSmoothingOperator"] == "face7"
    assert p["WhistlerSeedMode"] == "eigenmode"
    assert p["SaveDirName"] == "data" and p["RestartDirName"] == "data" and p["Case"] == "Uniform"
    for k in ("WriteMethod", "FieldWriteMethod", "ParticlesWriteMethod", "RestartWriteMethod", "SimulationName",
              "FieldOutputTag", "ParticlesOutputTag", "MomentsOutputTag"):
        assert p[k] == "", k
    for face in ("Xright", "Xleft", "Yright", "Yleft", "Zright", "Zleft"):
        assert p["bcEMface" + face] == "0" and p["bcPface" + face] == "1"


def test_parser_lexical_grammar(deck):
    p, _ = deck.with_("\n # comment\n dt \t= 2.5e-1 # last duplicate wins\n"
                      "SimulationName = \t a = b; c \t\nSaveDirName = \"a b\"\n"
                      "RestartDirName = \"before#after\"\n")
    assert float(p["dt"]) == 0.25
    assert p["SimulationName"] == "a = b; c"
    assert p["SaveDirName"] == '"a b"' and p["RestartDirName"] == '"before'
    assert deck.with_("SaveDirName =\nRestartDirName =\n")[0]["SaveDirName"] == ""
    # every rejection names the deck line; the base deck has 15 lines
    deck.error("dt = broken\n", "line 16: dt = 'broken' is not a number")
    deck.error("dt = 2.5e-1 seconds\n", "line 16: dt = '2.5e-1 seconds' is not a number")
    deck.error("ncycles = 17.5\n", "line 16: ncycles = '17.5' is not an integer")
    deck.error("ns = 1trailing\n", "line 16: ns = '1trailing' is not an integer")
    deck.error("ncycles =\n", "line 16: ncycles = '' is not an integer")
    deck.error("not a key value line\n", "line 16: expected 'key = value'")
    deck.error(" = value\n", "line 16: expected 'key = value'")
    assert deck.read(base(1))[1] == []
    assert deck.with_("DT = 1\n")[1] == ["DT (line 16)"]
    assert deck.with_("\n\nUnknown = 1\nAlsoUnknown = 2\n")[1] == ["Unknown (line 18)", "AlsoUnknown (line 19)"]
    deck.error("DT = 1\n", "missing required key dt", whole=True)
    deck.error("dt = 1\n", "missing required key ncycles", whole=True)
    deck.error("dt = 1\nncycles = 1\n", "missing required key ns", whole=True)
    deck.error("dt = 1\nncycles = 1\nns = 1\nrhoINIT = 1\nqom = -1\nnpcelx = 1\nnpcely = 1\n"
               "npcelz = 1\nuth = 0\nvth = 0\nwth = 0\nu0 = 0\nv0 = 0\nw0 = 0\n", "missing required key Case",
               whole=True)


REMOVED_KEYS = (
    "NonlinearECSIMNewtonMaxIter", "NonlinearECSIMGMRESRestart", "NonlinearECSIMGMRESMaxIter",
    "NonlinearECSIMParticleMaxIter", "NonlinearECSIMMaxPathSegmentsPerParticle",
    "NonlinearECSIMScalarReductionSlots", "NonlinearECSIMNewtonTol", "NonlinearECSIMGMRESTol",
    "NonlinearECSIMFDGMRESTol", "NonlinearECSIMParticleTol", "NonlinearECSIMJVPEpsilon",
    "NonlinearECSIMLineSearch", "NonlinearECSIMFrozenPathPreconditioner", "bcPHIfaceXright", "bcPHIfaceXleft",
    "bcPHIfaceYright", "bcPHIfaceYleft", "bcPHIfaceZright", "bcPHIfaceZleft", "PoissonCorrection",
    "PoissonCorrectionCycle", "CGtol", "NiterMover", "PushWithBatTime", "PushWithEatTime", "ImplSusceptTime",
    "ImplSusceptMode", "x_center", "y_center", "z_center", "L_square", "NpMaxNpRatio", "Vinj", "rhoINJECT",
    "pitch_angle", "energy", "ParaviewScriptPath", "HarrisB0", "TaylorGreenVortexAmplitude",
    "TestPartOutputCycle")


def test_parser_ignores_and_lists_keys_it_does_not_read(deck):
    for key in REMOVED_KEYS:
        assert deck.with_(key + " = 1\n")[1] == [key + " (line 16)"], key


CASES = ("Uniform", "GEM", "GEMnoPert", "GEMHarris", "DoubleGEM", "DoubleGEMnoPert", "plasmoidChain", "ForceFree",
         "Weibel", "Whistler", "TaylorGreenVortex", "TwoStream", "fluxRope", "GaussianExpansion",
         "WhistlerTurbulence")


def test_parser_cases(deck):
    for name in CASES:
        assert deck.with_("Case = %s\n" % name)[0]["Case"] == name
    for name in ("", "gem", "GEMDoubleHarris", "HarrisSheet", "DoubleHarris", "testGEM3D", "FluxRope", "fluxrope",
                 "flux_rope", "FluxRope3D", "fluxRope3D", "FluxRopeKink", "fluxRopeKink", "PlasmoidChain",
                 "plasmoid_chain", "MysteryCase"):
        deck.error("Case = %s\n" % name, "line 16: Case = '%s' is unknown (expected Uniform, GEM," % name)


def test_parser_output_methods(deck):
    p, _ = deck.with_("WriteMethod=hdf5\nFieldWriteMethod=vtk\nParticlesWriteMethod=adios2\n")
    assert (p["WriteMethod"], p["FieldWriteMethod"], p["ParticlesWriteMethod"]) == ("hdf5", "vtk", "adios2")
    p, _ = deck.with_("WriteMethod=hdf5\nFieldWriteMethod=\nParticlesWriteMethod=default\n")
    assert (p["WriteMethod"], p["FieldWriteMethod"], p["ParticlesWriteMethod"]) == ("hdf5", "", "default")
    for key in ("WriteMethod", "FieldWriteMethod", "ParticlesWriteMethod"):
        for value in ("unknown", "VTK", "HDF5", "Default", "bp4"):
            deck.error("%s = %s\n" % (key, value),
                       "TPIC input: line 16: %s = '%s' is unknown (expected default, pvtk, vtk, adios2, bp, bp5, "
                       "hdf5 or h5)" % (key, value))


def test_parser_booleans(deck):
    p, _ = deck.with_("PERIODICX = false\nPERIODICX_P = true\nPERIODICY = false\nPERIODICZ = false\n")
    assert (p["PERIODICX"], p["PERIODICX_P"], p["PERIODICY_P"], p["PERIODICZ_P"]) == ("false", "true", "false",
                                                                                      "false")
    assert deck.with_("WhistlerUseExactColdRoot = false\n")[0]["WhistlerUseExactColdRoot"] == "false"
    assert deck.with_("TaylorGreenVortexSeedMotionalE = true\n")[0]["TaylorGreenVortexSeedMotionalE"] == "true"
    p, _ = deck.with_("currentFromAmpere = 1\nspatiallyVaryingThermal = true\n")
    assert p["currentFromAmpere"] == "true" and p["spatiallyVaryingThermal"] == "true"
    assert deck.with_("QuietStart = true\n")[0]["QuietStart"] == "true"
    assert deck.with_("QuietStart = 1\n")[0]["QuietStart"] == "true"
    assert deck.with_("QuietStart = 0\n")[0]["QuietStart"] == "false"
    p, _ = deck.with_("PERIODICX = 0\nPERIODICY = 1\nPERIODICZ = 0\nPERIODICZ_P = 1\n")
    assert (p["PERIODICX"], p["PERIODICY"], p["PERIODICZ"], p["PERIODICX_P"], p["PERIODICZ_P"]) == (
        "false", "true", "false", "false", "true")
    for value in ("", "2", "01", "F", "no", "NONE", "TRUE", "False", "true trailing", "1 trailing"):
        deck.error("PERIODICX = %s\n" % value,
                   "line 16: PERIODICX = '%s' is not a boolean (0, 1, true or false)" % value)


def test_parser_dynamic_defaults(deck):
    p, _ = deck.with_("TaylorGreenVortexMode = 3\nTaylorGreenVortexModeY = 7\nFieldOutputCycle = 13\nLx = 18\n"
                      "Ly = 26\n")
    assert (p["TaylorGreenVortexModeX"], p["TaylorGreenVortexModeY"], p["TaylorGreenVortexModeZ"]) == ("3", "7", "3")
    assert p["DiagnosticsOutputCycle"] == "13"
    assert float(p["FluxRopeCenterX"]) == 9 and float(p["FluxRopeCenterY"]) == 13
    assert float(p["GaussianExpansionCenterX"]) == 9 and float(p["GaussianExpansionCenterY"]) == 13
    p, _ = deck.with_("HarrisB0x = 2\nTaylorGreenVortexMagneticAmplitude = 4\nTestParticlesOutputCycle = 6\n"
                      "FieldOutputCycle = 7\nDiagnosticsOutputCycle = 8\nDiagnosticEnergyCycle = 5\n"
                      "DiagnosticGaussCycle = 4\npertGEM = .5\npertHump = .25\ndeltaxHump = 2\ndeltayHump = 3\n"
                      "kxHump = 0\nkyHump = 1.5\n")
    assert [float(p[k]) for k in ("pertGEM", "pertHump", "deltaxHump", "deltayHump", "kxHump", "kyHump")] == [
        .5, .25, 2, 3, 0, 1.5]
    assert [float(p[k]) for k in ("HarrisB0x", "TaylorGreenVortexMagneticAmplitude", "TestParticlesOutputCycle",
                                  "DiagnosticsOutputCycle", "DiagnosticEnergyCycle", "DiagnosticGaussCycle")] == [
        2, 4, 6, 8, 5, 4]
    deck.error("TaylorGreenVortexMode = bad\nTaylorGreenVortexModeX = 1\n",
               "TaylorGreenVortexMode = 'bad' is not an integer")


def test_parser_enumerations(deck):
    assert deck.with_("ECSimSmoothing = energy_conserving\n")[0]["ECSimSmoothing"] == "energy_conserving"
    assert deck.with_("ECSimSmoothingOperator = binomial27\n")[0]["ECSimSmoothingOperator"] == "binomial27"
    for mode in ("eigenmode", "magnetic_only", "noise"):
        assert deck.with_("WhistlerSeedMode = %s\n" % mode)[0]["WhistlerSeedMode"] == mode
    for t in ("OFF", "False", "0", "EnergyConserving", "energy-conserving", "NONE"):
        deck.error("ECSimSmoothing = %s\n" % t,
                   "line 16: ECSimSmoothing = '%s' is unknown (expected none or energy_conserving)" % t)
    for t in ("SEVEN_POINT", "7point", "face6", "27point", "27_point", "directional", "Directional27", "Face7"):
        deck.error("ECSimSmoothingOperator = %s\n" % t,
                   "line 16: ECSimSmoothingOperator = '%s' is unknown (expected face7 or binomial27)" % t)
    for t in ("EIGEN_MODE", "eigen-mode", "magnetic-only", "magnetic", "B_ONLY", "b-only", "NOISE",
              "magnetic_noise", "magnetic-noise"):
        deck.error("WhistlerSeedMode = %s\n" % t,
                   "line 16: WhistlerSeedMode = '%s' is unknown (expected eigenmode, magnetic_only or noise)" % t)


LIST_KEYS = ("rhoINIT", "qom", "npcelx", "npcely", "npcelz", "uth", "vth", "wth", "u0", "v0", "w0")


def test_parser_lists_and_species(deck):
    p, _ = deck.read(base(2) + "rhoINIT = 0.1 0.2 0.3\nqom = -25 1 99\nnpcelx = 2 3\nnpcely = 4 5\nnpcelz = 6 7\n"
                     "uth = 0.6 0.7\nvth = 0.8 0.9\nwth = 1.1 1.2\nu0 = 1.3 1.4\nv0 = 1.5 1.6\nw0 = 1.7 1.8\n"
                     "CenterInitialVelocity = true\n")
    cols = {k: p[k].split() for k in LIST_KEYS}
    assert [float(v) for v in cols["qom"]] == [-25, 1]
    assert [float(cols[k][1]) for k in LIST_KEYS] == [0.2, 1, 3, 5, 7, 0.7, 0.9, 1.2, 1.4, 1.6, 1.8]
    assert p["CenterInitialVelocity"] == "true false"
    assert deck.with_("CenterInitialVelocity =\n")[0]["CenterInitialVelocity"] == "false"
    for key in LIST_KEYS:
        deck.error("%s =\n" % key, "line 16: %s lists 0 values, expected at least 1" % key)
        deck.error(base(2) + "%s = 1\n" % key, "%s lists 1 values, expected at least 2" % key, whole=True)
        deck.error(base(2) + "%s = 1 nope 2\n" % key, "'nope' is not a", whole=True)
    deck.error("qom = [-1]\n", "'[-1]' is not a number")
    deck.error("qom = -2suffix\n", "'-2suffix' is not a number")
    deck.error("uth = 3,\n", "'3,' is not a number")
    assert deck.with_("CenterInitialVelocity = 1\n")[0]["CenterInitialVelocity"] == "true"
    assert deck.with_("CenterInitialVelocity = 0\n")[0]["CenterInitialVelocity"] == "false"
    deck.error("CenterInitialVelocity = -3\n", "'-3' is not a boolean (0, 1, true or false)")


def test_parser_test_species_block(deck):
    extras = "nsTestPart = 2\nqom = -1 2 -3\nnpcelx = 1 2 3\nnpcely = 4 5 6\nnpcelz = 7 8 9\n"
    p, _ = deck.with_(extras)
    assert [float(v) for v in p["qom"].split()] == [-1, 2, -3]
    assert p["npcelx"].split()[2] == "3" and p["npcely"].split()[2] == "6" and p["npcelz"].split()[2] == "9"
    for k in ("rhoINIT", "uth", "vth", "wth", "u0", "v0", "w0"):
        assert [float(v) for v in p[k].split()[1:]] == [0, 0], k
    assert p["CenterInitialVelocity"].split()[1:] == ["false", "false"]
    p, _ = deck.with_(extras + "rhoINIT = 1 2 3\nuth = 4 5 6\nCenterInitialVelocity = false true false\n")
    assert float(p["rhoINIT"].split()[1]) == 2 and float(p["uth"].split()[2]) == 6
    assert p["CenterInitialVelocity"].split()[1] == "true"
    for key in ("qom", "npcelx", "npcely", "npcelz"):
        deck.error(extras + "%s = 1 2\n" % key, "%s lists 2 values, expected at least 3" % key)


def test_parser_species_counts(deck):
    for count in range(1, 11):
        p, _ = deck.read(base(count))
        assert p["ns"] == str(count) and float(p["qom"].split()[-1]) == count
    for value in ("0", "-1", "-2147483648", "11", "2147483647"):
        deck.error("ns = %s\n" % value, "supported real species range: 1-10")
    for value in ("broken", "", "2147483648", "-2147483649"):
        deck.error("ns = %s\n" % value, "is not an integer")
    for value in ("-1", "-2147483648"):
        deck.error("nsTestPart = %s\n" % value, "nsTestPart must be nonnegative")
    deck.error("nsTestPart = broken\n", "is not an integer")
    deck.error("nsTestPart = 2147483648\n", "is not an integer")
    deck.error("nsTestPart = 2147483647\n", "ns + nsTestPart exceeds the supported integer range")
    deck.error("nsTestPart = 2147483646\n", "qom lists 1 values, expected at least 2147483647")


# ---- admission and the command line ----------------------------------------

STARTUP = os.path.join(DECKS, "uniform_startup.inp")


def test_check_input_admits_the_uniform_startup_deck(driver):
    rc, out, err = run(driver, "--check-input", STARTUP)
    assert rc == 0, err
    assert "TPIC host input check: passed" in out
    assert "Particles: local=1008, global=1008" in out
    for s, ids in enumerate(("[0,336)", "[336,672)", "[672,1008)")):
        assert "species %d: count=336, IDs=%s" % (s, ids) in out
    assert "Unsupported" not in out


def write(tmp_path, text, name="deck.inp"):
    path = str(tmp_path / name)
    with open(path, "w") as f:
        f.write(text)
    return path


def small(extra=""):
    return base(1).replace("ncycles = 12", "ncycles = 2") + "nxc = 4\nnyc = 4\nnzc = 4\n" + extra


@pytest.mark.parametrize("extra,args,fragment", [
    ("Case = Whistler\n", (),
     "Unsupported: Case Whistler: this build initializes the Uniform, TwoStream, Weibel and Harris-sheet cases"),
    ("PERIODICX = false\nbcEMfaceXleft = 2\n", (),
     "Unsupported: bcEMfaceXleft=2: open electromagnetic boundaries are not implemented"),
    ("c = 2\n", (), "Unsupported: c != 1 is not supported"),
    ("nsTestPart = 1\nqom = 1 2\nnpcelx = 1 1\nnpcely = 1 1\nnpcelz = 1 1\n", (),
     "Unsupported: nsTestPart: test-particle initialization and evolution are not implemented"),
    ("XLEN = 2\n", (), "Unsupported: XLEN*YLEN*ZLEN = 2: this build runs one partition on one GPU"),
    ("PERIODICY_P = false\nbcPfaceYleft = 3\n", (), "Unsupported: bcPfaceYleft=3: boundary particle injection"),
    ("", ("--precision", "single-position"), "Unsupported: Precision pair single-position/single-position"),
    ("", ("--compute-precision", "single-particles-moments-fields"),
     "Unsupported: Precision pair double/single-particles-moments-fields"),
    ("", ("--order", "unsorted"), "Unsupported: --order unsorted"),
])
def test_check_input_reports_unsupported_requests(driver, tmp_path, extra, args, fragment):
    rc, out, err = run(driver, "--check-input", write(tmp_path, small(extra)), *args)
    assert rc == 2, (out, err)
    assert "TPIC host input check: unsupported requests" in out
    assert fragment in out, out


@pytest.mark.parametrize("extra,fragment", [
    ("th = 0.3\n", "TPIC physics: th must be in [0.5, 1] for the admitted linear-ECSIM scheme"),
    ("qom = 0\n", "TPIC physics: species 0: qom=0 has no defined neutral-species semantics"),
    ("GMREStol = 1\n", "TPIC physics: GMREStol must be less than 1"),
    ("dt = -1\n", "TPIC physics: dt must be finite and positive"),
    ("uth = -0.1\n", "TPIC physics: species 0: uth must be finite and nonnegative"),
    ("ncycles = -1\n", "TPIC physics: ncycles must be nonnegative"),
    ("nxc = 0\n", "TPIC domain: x global cell count must be positive"),
    ("Lz = 0\n", "TPIC domain: z physical length must be finite and positive"),
    ("PERIODICY = false\nnyc = 1\n", "TPIC domain: y requires at least 2 cells per rank"),
    ("npcelx = 0\n", "TPIC particles: PPC factors must be positive and product representable"),
    ("bcEMfaceXleft = 3\n", "TPIC physics: bcEMfaceXleft=3: unknown code (expected 0 perfect electric "
                            "conductor, 1 perfect magnetic conductor or 2 open)"),
    ("ECSimSmoothing = energy_conserving\nSmooth = 1.5\n",
     "TPIC physics: energy-conserving smoothing requires 0 <= Smooth <= 1"),
    ("ECSimSmoothing = energy_conserving\nSmoothNiter = -1\n",
     "TPIC physics: energy-conserving smoothing requires SmoothNiter >= 0"),
    ("ECSimSmoothing = energy_conserving\nSmooth = 0.5\nPERIODICX = false\nbcEMfaceXleft = 2\n",
     "TPIC physics: active energy-conserving smoothing cannot use an effective open EM face"),
])
def test_check_input_rejects_invalid_decks(driver, tmp_path, extra, fragment):
    rc, out, err = run(driver, "--check-input", write(tmp_path, small(extra)))
    assert rc == 1, (out, err)
    assert fragment in err, err


def test_check_input_warns_about_unknown_keys_and_unwritten_output(driver, tmp_path):
    rc, out, _ = run(driver, "--check-input", write(tmp_path, small("Mystery = 1\n")))
    assert rc == 0
    assert "Warning: Unknown input keys ignored: Mystery (line 19)" in out
    # the base deck keeps the default FieldOutputCycle = 100 and its default tags
    assert "Warning: Moment output (rho, J, pressure) is not built: field snapshots hold E and B only" in out


def test_command_line(driver, tmp_path):
    deck_path = write(tmp_path, small())
    rc, out, _ = run(driver, "--help")
    assert rc == 0 and out.startswith("Usage: tpic_mi_run") and "double/double" in out
    for args, fragment in [
        ((), "give exactly one of --check-input, --initialize or --run and one input file"),
        (("--check-input",), "give exactly one of"),
        (("--check-input", "--run", deck_path), "give exactly one of"),
        (("--check-input", deck_path, deck_path), "TPIC options: expected exactly one input file"),
        (("--check-input", deck_path, "--bogus"), "TPIC options: unknown option: --bogus"),
        (("--check-input", deck_path, "--order"), "TPIC options: --order requires a value"),
        (("--check-input", deck_path, "--order", "sorted"), "--order must be full, partial or unsorted"),
        (("--check-input", deck_path, "--mass-matrix", "mfma"), "--mass-matrix must be scalar or matrix-core"),
        (("--check-input", deck_path, "--mass-apply", "half"), "--mass-apply must be full or split"),
        (("--check-input", deck_path, "--precision", "half"), "TPIC precision: --precision must be double"),
        (("--check-input", deck_path, "--rank", "-1"), "--rank requires a nonnegative signed integer"),
        (("--check-input", deck_path, "--rank", "1"), "TPIC domain: rank must be in [0, communicator_size)"),
        (("--run", deck_path, "--rank", "0"), "--rank is only valid with --check-input"),
        (("--check-input", deck_path, "--restart"), "--restart is only valid with --run or --initialize"),
        (("--check-input", deck_path, "--mass-precision", "fp16"), "--mass-precision must be fp64, fp32 or bf16x3"),
        (("--check-input", str(tmp_path / "missing.inp")), "TPIC input: unable to open input file"),
    ]:
        rc, out, err = run(driver, *args)
        assert rc == 1, (args, out, err)
        assert fragment in err, (args, err)
    rc, out, _ = run(driver, "--check-input", deck_path, "--order", "full", "--mass-matrix", "matrix-core",
                     "--mass-precision", "bf16x3")
    assert rc == 0 and "particle order plan: full" in out and "moment operands: bf16x3" in out
    rc, out, err = run(driver, "--run", write(tmp_path, small("Case = ForceFree\n"), "ff.inp"))
    assert rc == 2 and "unsupported request: Case ForceFree" in err


def test_check_input_two_stream_and_weibel_recipes(driver, tmp_path):
    for deck, recipe in (("two_stream_1d.inp", "TwoStream"), ("weibel_2d.inp", "uniform")):
        rc, out, err = run(driver, "--check-input", os.path.join(DECKS, deck))
        assert rc == 0, err
        assert "particle recipe: %s" % recipe in out
    two = open(os.path.join(DECKS, "two_stream_1d.inp")).read()
    weibel = open(os.path.join(DECKS, "weibel_2d.inp")).read()
    for text, fragment in [
        (small("Case = TwoStream\n"), "TPIC recipe: TwoStream requires exactly two electron species and one ion species"),
        (two + "TwoStreamSeedModeX = 32\n", "TPIC recipe: TwoStreamSeedModeX must be below the x-grid Nyquist mode"),
        (two + "TwoStreamSeedDensityAmplitude = 1\n",
         "TPIC recipe: abs(TwoStreamSeedDensityAmplitude) must be less than one"),
        (two + "u0 = 0.2 0.1 0\n", "TPIC recipe: TwoStream electron u0 must be equal and opposite and nonzero"),
        (two + "rhoINIT = 0.5 0.5 0.9\n", "TPIC recipe: TwoStream ion rhoINIT must equal the sum of electron rhoINIT"),
        (two + "TwoStreamSeedModeX = 0\n", "TPIC recipe: TwoStreamSeedModeX must be positive for nonzero seed amplitude"),
        (weibel + "WeibelSeedModeX = 0\n", "TPIC recipe: WeibelSeedModeX must be positive for nonzero seed amplitude"),
        (weibel + "WeibelSeedModeX = -1\n", "TPIC recipe: WeibelSeedModeX must be non-negative"),
    ]:
        rc, out, err = run(driver, "--check-input", write(tmp_path, text))
        assert rc == 1 and fragment in err, (fragment, out, err)


def test_check_input_admits_pec_and_pmc_walls(driver):
    rc, out, err = run(driver, "--check-input", os.path.join(DECKS, "uniform_pmc_box.inp"))
    assert rc == 0, err
    assert "EM faces [Xright Xleft Yright Yleft Zright Zleft]: PMC PMC periodic periodic PEC PMC" in out
    assert "Particle faces: mirror mirror periodic periodic mirror mirror" in out


@pytest.mark.parametrize("extra,state", [
    ("ECSimSmoothing = energy_conserving\nSmooth = 0.5\nSmoothNiter = 4\n", "active"),
    ("ECSimSmoothing = energy_conserving\nECSimSmoothingOperator = binomial27\nSmooth = 0\n", "active"),
    ("ECSimSmoothing = energy_conserving\nSmooth = 1\n", "inactive"),
    ("ECSimSmoothing = energy_conserving\nSmooth = 0.5\nSmoothNiter = 0\n", "inactive"),
    ("ECSimSmoothing = none\nSmooth = 0.5\n", "inactive"),
])
def test_check_input_admits_energy_conserving_smoothing(driver, tmp_path, extra, state):
    """physics.cpp:65-71: smoothing is active when ECSimSmoothing is
    energy_conserving, SmoothNiter > 0 and Smooth < 1; this build runs it."""
    rc, out, err = run(driver, "--check-input", write(tmp_path, small(extra)))
    assert rc == 0, (out, err)
    assert "Smoothing: %s; GMRES restart=30" % state in out, out


def test_check_input_harris_recipes(driver, tmp_path):
    """The Harris-sheet recipe checks (harris_sheet.cpp:19-227)."""
    gem = open(os.path.join(DECKS, "gem_harris_2d.inp")).read()
    rc, out, err = run(driver, "--check-input", os.path.join(DECKS, "gem_harris_2d.inp"))
    assert rc == 0 and "particle recipe: GEM/Harris" in out, err
    rc, out, err = run(driver, "--check-input", os.path.join(DECKS, "double_gem_2d.inp"))
    assert rc == 0 and "particle recipe: DoubleGEM" in out, err
    for text, fragment in [
        (gem + "delta = 0\n", "TPIC recipe: Harris delta must be finite and positive"),
        (gem + "deltaxHump = -1\n", "TPIC recipe: GEMHarris deltaxHump must be finite and positive"),
        (gem + "kyHump = 0\n", "TPIC recipe: GEMHarris kyHump must be nonzero"),
        (gem + "u0 = 0.01 0 0 0\n", "TPIC recipe: GEMHarris currentFromAmpere requires u0 = v0 = 0"),
        (gem + "w0 = 0 0 0 0\n", "TPIC recipe: GEMHarris currentFromAmpere requires a nonzero sum of sign(qom)*w0"),
        (gem + "pertGEM = 0.1\n",
         "TPIC recipe: GEMHarris currentFromAmpere drift exceeds c where the sheet density vanishes"),
    ]:
        rc, out, err = run(driver, "--check-input", write(tmp_path, text))
        assert rc == 1 and fragment in err, (fragment, out, err)
    rc, out, err = run(driver, "--check-input", write(tmp_path, gem + "CenterInitialVelocity = 1 0 0 0\n"))
    assert rc == 2 and "Unsupported: CenterInitialVelocity of species 0 with a position-dependent Harris drift" in out


def restart_file(path, **fields):
    """A restart file holding only the magic, the version and a header (what
    the admission reads before any GPU work)."""
    h = R.CheckpointHeader()
    h.cycle, h.species = 2, 1
    h.global_cells[:] = [4, 4, 4]
    h.lengths[:] = [10.0, 10.0, 10.0]
    h.dt, h.c = 0.125, 1.0
    h.faces[:] = [-1] * 12
    h.storage_bytes[:] = [8] * 5
    h.q_over_m[0], h.global_count[0] = 1.0, 64
    for key, value in fields.items():
        if isinstance(value, list):
            getattr(h, key)[:] = value
        else:
            setattr(h, key, value)
    with open(path, "wb") as f:
        f.write(b"TPICMICK" + (1).to_bytes(4, "little") + bytes(h))


def test_restart_admission(driver, tmp_path):
    """--restart continues from the newest checkpoint_t<cycle>_n1.tpicmi of
    RestartDirName; a checkpoint whose mesh, domain, boundaries, dt, c or
    species differ from the deck's, or whose cycle is not before ncycles, is
    refused before any GPU work (admit_restart, run_plan.cpp:199-214)."""
    ckpt = tmp_path / "ckpt"
    deck = write(tmp_path, small("RestartDirName = %s\nRestartOutputCycle = 2\nncycles = 4\n" % ckpt))
    rc, _, err = run(driver, "--run", deck, "--restart")
    assert rc == 1 and "TPIC restart: no restart file in " + str(ckpt) in err
    ckpt.mkdir()
    rc, _, err = run(driver, "--check-input", deck, "--restart")
    assert rc == 1 and "--restart is only valid with --run or --initialize" in err
    header = R.checkpoint_header
    for fields, fragment in [
            (dict(global_cells=[4, 4, 8]), "Restart: the mesh (nxc, nyc, nzc) differs from the checkpoint's"),
            (dict(lengths=[10.0, 10.0, 12.0]), "Restart: the domain (Lx, Ly, Lz) differs from the checkpoint's"),
            (dict(faces=[-1] * 6 + [0] * 6), "Restart: a field or particle boundary differs from the checkpoint's"),
            (dict(dt=0.25), "Restart: dt differs from the checkpoint's"),
            (dict(c=2.0), "Restart: c differs from the checkpoint's"),
            (dict(species=2), "Restart: the species (ns, qom) differs from the checkpoint's"),
            (dict(q_over_m=[-1.0] + [0.0] * 9), "Restart: the species (ns, qom) differs from the checkpoint's"),
            (dict(global_count=[65] + [0] * 9), "Restart: the particle count of species 0 differs"),
            (dict(cycle=4), "Restart: ncycles = 4 is not after the checkpoint cycle 4")]:
        path = ckpt / "checkpoint_t3_n1.tpicmi"
        restart_file(path, **fields)
        assert header(str(path)).dt == fields.get("dt", 0.125)
        rc, _, err = run(driver, "--run", deck, "--restart")
        assert rc == 2 and fragment in err, (fields, err)
        # an older file is not the one read
        restart_file(ckpt / "checkpoint_t1_n1.tpicmi")
        rc, _, err = run(driver, "--run", deck, "--restart")
        assert rc == 2 and fragment in err, (fields, err)
        os.remove(path)
    (ckpt / "checkpoint_t9_n1.tpicmi").write_bytes(b"not a checkpoint")
    rc, _, err = run(driver, "--run", deck, "--restart")
    assert rc == 1 and "not a checkpoint of this build" in err
    with pytest.raises(R.TpicError, match="not a checkpoint of this build"):
        header(str(ckpt / "checkpoint_t9_n1.tpicmi"))


def test_load_deck_errors(tmp_path):
    """tpic_mi_load_deck parses and admits before any GPU work: a broken deck
    is TPIC_MI_INVALID_ARGUMENT with the parser's message, a deck this build
    does not run TPIC_MI_UNSUPPORTED with every unsupported request; no
    context is returned."""
    L = R.load_library()
    h, run = ctypes.c_void_p(), R.DeckRun()
    for text, status, fragment in [
            (small("dt = fast\n"), R.INVALID_ARGUMENT, "dt"),
            (small("ns = 0\n"), R.INVALID_ARGUMENT, "ns"),
            (small("Case = Whistler\n"), R.UNSUPPORTED,
             "TPIC run plan: unsupported requests\nCase Whistler: this build initializes"),
    ]:
        path = write(tmp_path, text)
        rc = L.tpic_mi_load_deck(os.fsencode(path), None, None, ctypes.byref(h), ctypes.byref(run))
        assert rc == status and h.value is None, (text, rc)
        assert fragment in R.last_error(), R.last_error()
    rc = L.tpic_mi_load_deck(os.fsencode(str(tmp_path / "missing.inp")), None, None, ctypes.byref(h), None)
    assert rc == R.INVALID_ARGUMENT and "missing.inp" in R.last_error()
    opts = R.DeckOptions()
    opts.mass_kernel = 3
    rc = L.tpic_mi_load_deck(os.fsencode(path), ctypes.byref(opts), None, ctypes.byref(h), None)
    assert rc == R.INVALID_ARGUMENT and "mass_kernel" in R.last_error()
    opts.mass_kernel, opts.reserved[2] = 0, 1
    rc = L.tpic_mi_load_deck(os.fsencode(path), ctypes.byref(opts), None, ctypes.byref(h), None)
    assert rc == R.INVALID_ARGUMENT and "reserved" in R.last_error()
    with pytest.raises(R.TpicError, match="unsupported requests"):
        R.Simulation.from_deck(path)


def test_check_input_output_plan(driver, tmp_path):
    """Output requests: the fields and diagnostics this driver writes pass;
    what it does not write is a warning (the reference's text for a build
    without a checkpoint backend); checkpoints are this build's raw restart
    files; another field engine is refused."""
    rc, out, _ = run(driver, "--check-input", write(tmp_path, small(
        "FieldOutputCycle = 5\nParticlesOutputCycle = 5\nRestartOutputCycle = 10\nDiagnosticGaussCycle = 2\n")))
    assert rc == 0
    assert "Warning: Moment output (rho, J, pressure) is not built: field snapshots hold E and B only" in out
    assert "ParticlesOutputCycle" not in out  # particle snapshots are written
    assert ("Note: RestartOutputCycle: checkpoints are raw files data/checkpoint_t<cycle>.tpicmi (this build "
            "has no ADIOS2 or HDF5 engine)") in out
    assert "Warning: DiagnosticGaussCycle: Gauss-law residual samples are not built" in out
    rc, out, _ = run(driver, "--check-input", write(tmp_path, small("FieldOutputCycle = 5\nWriteMethod = hdf5\n")))
    assert rc == 2 and "Unsupported: FieldWriteMethod hdf5: this build writes fields with pvtk only" in out
    rc, out, _ = run(driver, "--check-input", write(tmp_path, small("ParticlesOutputCycle = 5\n"
                                                                    "ParticlesWriteMethod = adios2\n")))
    assert rc == 2 and "Unsupported: ParticlesWriteMethod adios2: this build writes particles with pvtk only" in out
    rc, out, err = run(driver, "--check-input", write(tmp_path, small("FieldOutputTag = B+rho\n")))
    assert rc == 1 and "TPIC output: FieldOutputTag = 'B+rho': 'rho' is unknown (expected B, E or Gauss)" in err
    rc, out, _ = run(driver, "--check-input", write(tmp_path, small()), "--restart")
    assert rc == 1  # --restart is only valid with --run or --initialize



def test_static_field_energy_rejects_arguments_before_any_device_work():
    """tpic_mi_field_energy_static checks its arguments on the host: a grid
    beyond its 1 MiB scratch (256^3 cells) and a null pointer are refused
    without a device call."""
    dummy = ctypes.c_void_p(16)
    rc = R.field_energy_static((512, 256, 256), (1.0, 1.0, 1.0), dummy, dummy, dummy)
    assert rc == R.INVALID_ARGUMENT
    assert "exceeds the scratch limit of 16777216" in R.last_error()
    rc = R.field_energy_static((8, 8, 8), (1.0, 1.0, 1.0), dummy, None, dummy)
    assert rc == R.INVALID_ARGUMENT
    assert "null field or output pointer" in R.last_error()
