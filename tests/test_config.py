"""Configuration files: the libconfig-syntax reader of the host layer
(impec2d_amd/host/impec_config.cpp) against the reference's shipped
configurations (copied as fixtures under tests/golden/configs) and its
mapping of keys and options (SimParams.cpp:45-463), through the driver's
--print-params (no device needed)."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
RUN = os.path.join(ROOT, "impec2d_amd", "impec2d_run")
CFG = os.path.join(ROOT, "tests", "golden", "configs")


def _params(*args):
    r = subprocess.run([RUN] + list(args) + ["--print-params"], capture_output=True, text=True, timeout=60)
    return r.returncode, r.stdout, r.stderr


def _parse(out):
    d = {}
    for line in out.splitlines():
        k, _, v = line.partition(" ")
        d[k] = v
    return d


@pytest.fixture(scope="module", autouse=True)
def _driver():
    if not os.path.exists(RUN):
        pytest.fail("impec2d_run is not built")


# values read by hand from the three production configurations
EXPECT = {
    "parallel.cfg": dict(N1=141, N2=141, stype="parallel", x_ub=1.41421356237, nu=2 / 3.0, mu_res=41.0,
                         mu_inj=1.0, diffusion=0, well_profile=1, eps=1e-14, log_skip_steps=30),
    "parallel_dispersion.cfg": dict(N1=141, N2=141, stype="parallel", x_ub=1.41421356237, nu=2 / 3.0,
                                    mu_res=41.0, mu_inj=1.0, diffusion=1, alpha_l=1.0, alpha_t=1.0,
                                    well_profile=1, para_eps=1e-14),
    "diagonal_dispersion.cfg": dict(N1=100, N2=50, stype="diagonal", x_ub=1.0, nu=2 / 3.0, mu_res=41.0,
                                    diffusion=1, well_profile=1, tfinal=0.4, cfl=0.5),
}


@pytest.mark.parametrize("name", sorted(EXPECT))
def test_shipped_configurations_read(name):
    rc, out, err = _params("--config", os.path.join(CFG, name))
    assert rc == 0, err
    d = _parse(out)
    for k, v in EXPECT[name].items():
        if isinstance(v, str):
            assert d[k] == v, (k, d[k], v)
        else:
            assert float(d[k]) == pytest.approx(v, rel=1e-15, abs=0), (k, d[k], v)
    assert d["avgproc"] == "harmonic" and d["dim_splitting"] == "4"       # "parallel" splitting
    assert float(d["dx"]) == float(d["x_ub"]) / int(d["N1"])               # SimParams.cpp: dx = (x_ub - x_lb) / N1


# the test sections of the two test configurations (SimParams.cpp:252-316)
TESTS = {
    "elliptic_test.cfg": dict(test_component="elliptic", testno="4", init_x_cells="20", init_y_cells="30",
                              levels="6", hyp_test="1", hyp_norm="1"),
    "hyperbolic_test.cfg": dict(test_component="hyperbolic", testno="1", init_x_cells="50", init_y_cells="50",
                                levels="4", hyp_test="0", hyp_norm="0"),
}


@pytest.mark.parametrize("name", sorted(TESTS))
def test_test_configurations_read(name):
    rc, out, err = _params("--config", os.path.join(CFG, name))
    assert rc == 0, err
    d = _parse(out)
    assert d["is_test"] == "1"
    for k, v in TESTS[name].items():
        assert d[k] == v, (k, d[k], v)


def _test_text():
    with open(os.path.join(CFG, "hyperbolic_test.cfg")) as f:
        return f.read()


@pytest.mark.parametrize("key,value,message", [
    ("component", '"parabolic"', "source type is not valid"),
    ("profile", '"square"', "Test profile is not valid"),
    ("hyp_norm", "3", "Norm is not supported"),
    ("levels", None, "section <<test>>"),
])
def test_invalid_test_settings_reported(tmp_path, key, value, message):
    rc, out, err = _run_text(tmp_path, _set(_test_text(), key, value))
    assert rc == 2 and message in err, err


def _base_text():
    with open(os.path.join(CFG, "diagonal_dispersion.cfg")) as f:
        return f.read()


def _run_text(tmp_path, text, *extra):
    p = tmp_path / "c.cfg"
    p.write_text(text)
    return _params("--config", str(p), *extra)


def test_syntax_features(tmp_path):
    """Comments of all three kinds, ':' and '=', concatenated strings, hex
    integers, 64-bit suffixes, integer values for floating settings, arrays
    and lists of output times and steps."""
    t = _base_text()
    t = _set(t, "x_cells", "0x40 /* hex */") + "\n# trailing comment\n// another\n"
    t = _set(t, "y_cells", "48L")
    t = _set(t, "log_time_list", "[0.1, 0.25, 1]")
    t = _set(t, "log_counter_list", "(3, 5)")
    t = _set(t, "results_dir", '"/tm" "p/out"')
    t = _set(t, "mu_res", "41")
    rc, out, err = _run_text(tmp_path, t)
    assert rc == 0, err
    d = _parse(out)
    assert d["N1"] == "64" and d["N2"] == "48"
    assert [float(x) for x in d["times"].split()] == [0.1, 0.25, 1.0]
    assert d["counters"].split() == ["3", "5"]
    assert d["results_dir"] == "/tmp/out"
    assert float(d["mu_res"]) == 41.0


def _set(text, key, value):
    """The setting `key = ...;` of the text replaced by `key = value;`
    (removed when value is None)."""
    import re
    pat = re.compile(r"(^|\s)" + re.escape(key) + r"\s*[=:][^;]*;", re.M)
    assert pat.search(text), key
    return pat.sub(lambda m: m.group(1) + ("" if value is None else "%s = %s;" % (key, value)), text, count=1)


@pytest.mark.parametrize("key,value,message", [
    ("x_cells", None, "section <<general>>"),
    ("mu_inj", None, "<<elliptic>>"),
    ("average_procedure", '"arithmetic"', "Averaging procedure is unknown"),
    ("dimensional_splitting", '"diagonal"', "Dimenstional Splitting"),
    ("source_type", '"horizontal"', "horizontal"),
    ("well_profile", '"square"', "Unknown profile"),
    ("ell_method", '"cg"', "requires a Multigrid-Solver"),
    ("x_cells", '"100"', "section <<general>>"),
    ("bound_type", '"dirichlet"', "Boundary type"),
    ("syncmode", '"lockstep"', "protocoll is not implemented"),
    ("max_mass_frac_error", None, "max_mass_frac_error"),
])
def test_invalid_settings_reported(tmp_path, key, value, message):
    rc, out, err = _run_text(tmp_path, _set(_base_text(), key, value))
    assert rc == 2 and message in err, err


@pytest.mark.parametrize("text,message", [
    ('application: { general: { x_cells = "abc; }; };', "unterminated string"),
    ('application: { general: { x_cells = 1; };', "'}' expected"),
    ('application: { general: { x_cells 1; }; };', "'=' or ':' expected"),
    ('application: { general: { a = [1, "x"]; }; };', "different types"),
    ('/* open', "unterminated comment"),
])
def test_syntax_errors_name_the_line(tmp_path, text, message):
    rc, out, err = _run_text(tmp_path, text)
    assert rc == 2 and message in err and "line " in err, err


def test_command_line_overrides_the_file(tmp_path):
    rc, out, err = _params("--config", os.path.join(CFG, "parallel.cfg"), "--N", "64", "--results-dir", "/x")
    assert rc == 0, err
    d = _parse(out)
    assert d["N1"] == "64" and d["results_dir"] == "/x" and d["stype"] == "parallel"
