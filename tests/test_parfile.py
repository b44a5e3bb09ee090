"""The reference's parameter files (readParamFile, BSSN_GR/src/grUtils.cpp:37-52)
read into the right-hand side's parameters: the reference's own JSON and TOML
files under tests/golden/pars/ (copied from BSSN_GR/pars/), the JSON and TOML
forms of one run agreeing, and the optional/required-key rules of the
reference's readers. CPU only (the library is loaded for its defaults)."""
import json
import os

import pytest

from dgr import parfile

PARS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "pars")


def test_json_tinytest():
    f = parfile.read_param_file(os.path.join(PARS, "q1.tinytest.par.json"))
    p = f.params
    assert list(p.lambda_) == [1.0, 1.0, 1.0, 1.0]
    assert list(p.lambda_f) == [1.0, 0.0]
    assert p.ko_sigma == 0.4 and p.chi_floor == 1e-4
    # keys absent from the file keep the reference defaults (parameters.cpp:36-42, 133-136)
    assert (p.eta_central, p.eta_outer, p.eta_width) == (2.0, 0.25, 40.0)
    assert p.cako_enabled == 0 and (p.eps_cako_gauge, p.eps_cako_other) == (0.99, 0.3)
    assert f.bh == ((4.0, 0.0, 0.0), (-4.0, 0.0, 0.0))
    assert f.constraint_r == (5.0, 5.0)
    assert f.cfl == 0.25 and f.time == (0.0, 20.0)


def test_toml_and_json_forms_agree_on_the_rhs_parameters():
    j = parfile.read_param_file(os.path.join(PARS, "q1.tinytest.par.json")).params
    t = parfile.read_param_file(os.path.join(PARS, "q1.tinytest.par.toml")).params
    for name, _ in j._fields_:
        a, b = getattr(j, name), getattr(t, name)
        assert (list(a) if hasattr(a, "__len__") else a) == (list(b) if hasattr(b, "__len__") else b), name


def test_toml_q2_constraint_radii_and_cako_flag():
    f = parfile.read_param_file(os.path.join(PARS, "q2.par.toml"))
    assert f.constraint_r == (1.2, 2.4)
    assert f.params.cako_enabled == 0  # BSSN_KO_SIGMA_SCALE_BY_CONFORMAL = false
    assert f.cfl == 0.25 and f.time == (0.0, 600.0)


def _variant(tmp_path, **changes):
    d = json.load(open(os.path.join(PARS, "q1.tinytest.par.json")))
    for k, v in changes.items():
        if v is None:
            d.pop(k)
        else:
            d[k] = v
    path = tmp_path / "v.par.json"
    path.write_text(json.dumps(d))
    return str(path)


def test_cako_and_eta_keys(tmp_path):
    f = parfile.read_param_file(_variant(tmp_path, BSSN_KO_SIGMA_SCALE_BY_CONFORMAL=True,
                                         BSSN_EPSILON_CAKO_GAUGE=0.5, RIT_ETA_CENTRAL=1.5,
                                         RIT_ETA_WIDTH=20.0))
    assert f.params.cako_enabled == 1 and f.params.eps_cako_gauge == 0.5
    assert f.params.eta_central == 1.5 and f.params.eta_width == 20.0
    # the post-merger-only flag turns the scaling off before the merger (grUtils.cpp:186-191)
    f = parfile.read_param_file(_variant(tmp_path, BSSN_KO_SIGMA_SCALE_BY_CONFORMAL=True,
                                         BSSN_KO_SIGMA_SCALE_BY_CONFORMAL_POST_MERGER_ONLY=True))
    assert f.params.cako_enabled == 0


def test_lambda_is_cast_to_unsigned_like_the_json_reader(tmp_path):
    lam = {"BSSN_LAMBDA_1": 1.9, "BSSN_LAMBDA_2": 0, "BSSN_LAMBDA_3": 2, "BSSN_LAMBDA_4": 1}
    f = parfile.read_param_file(_variant(tmp_path, BSSN_LAMBDA=lam))
    assert list(f.params.lambda_) == [1.0, 0.0, 2.0, 1.0]


@pytest.mark.parametrize("key", ["KO_DISS_SIGMA", "CHI_FLOOR", "BSSN_LAMBDA", "BSSN_LAMBDA_F"])
def test_missing_required_key_raises(tmp_path, key):
    with pytest.raises(KeyError):
        parfile.read_param_file(_variant(tmp_path, **{key: None}))


def test_rk_type_is_read():
    """BSSN_RK_TYPE (grUtils.cpp:100): the tinytest files ask for RK4 (1)."""
    for name in ("q1.tinytest.par.json", "q1.tinytest.par.toml"):
        assert parfile.read_param_file(os.path.join(PARS, name)).rk_type == 1


def test_eta_function_constants_are_read():
    """BSSN_ETA_R0 and BSSN_ETA_POWER (JSON object of _1/_2, TOML array;
    grUtils.cpp:205-207, parameters.cpp:329-331): 1.31 and (2, 2) in the
    reference's files; the eta function itself stays off (a build option
    there)."""
    for name in ("q1.tinytest.par.json", "q1.tinytest.par.toml", "q2.par.toml"):
        p = parfile.read_param_file(os.path.join(PARS, name)).params
        assert (p.eta_r0, p.eta_power[0], p.eta_power[1], p.eta_func) == (1.31, 2.0, 2.0, 0)


def test_rk45_keys_are_read():
    """BSSN_RK45_TIME_STEP_SIZE and BSSN_RK45_DESIRED_TOL (parameters.cpp:
    220-223) as the reference's files hold them: 0.01 and 1e-3."""
    for name in ("q1.tinytest.par.json", "q1.tinytest.par.toml", "q2.par.toml"):
        assert parfile.read_param_file(os.path.join(PARS, name)).rk45 == (0.01, 1e-3)


def test_xi_is_read_and_cast_to_unsigned(tmp_path):
    """BSSN_XI (grUtils.cpp:161-163 as an object of _0.._2, parameters.cpp:
    315-317 as an array; both cast to unsigned int): (0, 0, 0) in the
    reference's files; other values pass through as the integers the casts
    give; the Rochester gauge itself stays off (a build option there)."""
    for name in ("q1.tinytest.par.json", "q1.tinytest.par.toml", "q2.par.toml"):
        p = parfile.read_param_file(os.path.join(PARS, name)).params
        assert (list(p.xi), p.rochester, p.ssl) == ([0.0, 0.0, 0.0], 0, 0)
    with open(os.path.join(PARS, "q1.tinytest.par.json")) as fh:
        d = json.load(fh)
    d["BSSN_XI"] = {"BSSN_XI_0": 1, "BSSN_XI_1": 2, "BSSN_XI_2": -1}
    f = tmp_path / "xi.par.json"
    f.write_text(json.dumps(d))
    assert list(parfile.read_param_file(str(f)).params.xi) == [1.0, 2.0, 4294967295.0]
