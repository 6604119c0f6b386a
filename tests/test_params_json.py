"""gpe_params.json validation through the host C ABI (gpes_validate_params_json),
mirroring the reference's FileManager tests (tests/filemanager/file_manager_test.cpp:13-335):
required keys throw with a message naming the key, optional fftw_n_threads
defaults, edd vs edd_start/edd_stop, strategy-dependent keys, potential keys.
The valid base configuration is the reference test fixture's values
(file_manager_test.cpp:17-36)."""
import json

import pytest

from gpesolver_amd import hip

VALID = {
    "n_atoms": 1000, "m": 4.0, "dx": 0.1, "dy": 0.1, "dz": 0.1, "nx": 64, "ny": 64, "nz": 64,
    "omega_x": 1.0e9, "omega_y": 1.0e9, "omega_z": 1.0e9, "calc_strategy": "IT", "init_strategy": "GAUSS",
    "pote_strategy": "HARMONIC", "iter_imag": 1000, "iter_real": 5000, "edd": 0.0, "fftw_n_threads": 4,
}


def check(cfg):
    return hip.validate_params_json(json.dumps(cfg))


def variant(erase=(), **set_):
    c = dict(VALID)
    for k in erase:
        c.pop(k, None)
    c.update(set_)
    return c


def test_complete_config_success():
    assert check(VALID) is None


@pytest.mark.parametrize("key", ["dx", "dy", "dz", "nx", "ny", "nz"])
def test_load_box_missing_key_throws(key):
    err = check(variant(erase=[key]))
    assert err is not None and key in err


def test_empty_json_throws():
    assert check({}) is not None


@pytest.mark.parametrize("key", ["calc_strategy", "iter_imag", "iter_real", "n_atoms", "m"])
def test_load_simulation_missing_key_throws(key):
    err = check(variant(erase=[key]))
    assert err is not None and key in err


def test_optional_fftw_threads_uses_default():
    assert check(variant(erase=["fftw_n_threads"])) is None


def test_fftw_threads_must_be_positive():
    assert "FFTW" in check(variant(fftw_n_threads=0))


def test_erase_edd_with_ramp_no_throw():
    assert check(variant(erase=["edd"], calc_strategy="RT", edd_start=1.5, edd_stop=1.5)) is None


def test_erase_edd_throws_for_it():
    assert "edd" in check(variant(erase=["edd"]))


def test_erase_edd_fs_needs_ramp_keys():
    assert "edd_start" in check(variant(erase=["edd"], calc_strategy="FS"))


def test_single_gauss_success():
    assert check(variant(erase=["initial_maximas"], init_strategy="GAUSS")) is None


def test_multiple_gauss_missing_maximas_throws():
    assert "initial_maximas" in check(variant(erase=["initial_maximas"], init_strategy="MULTIPLE_GAUSS"))


def test_multiple_gauss_with_maximas_success():
    assert check(variant(init_strategy="MULTIPLE_GAUSS", initial_maximas=5)) is None


def test_multiple_gauss_nonpositive_maximas_throws():
    assert "positive" in check(variant(init_strategy="MULTIPLE_GAUSS", initial_maximas=0))


def test_setup_gauss_lowercase_is_unknown_strategy():
    # the reference's "missing droplets" test uses lowercase "setup_gauss" and
    # passes through the unknown-strategy throw (complex_param.hpp:93)
    assert "Unknown initialization strategy" in check(variant(init_strategy="setup_gauss"))


@pytest.mark.parametrize("key", ["bec_droplets_x", "bec_droplets_y", "bec_droplets_z"])
def test_setup_gauss_missing_droplets_throws(key):
    cfg = variant(init_strategy="SETUP_GAUSS", bec_droplets_x=1.0, bec_droplets_y=1.0, bec_droplets_z=1.0)
    cfg.pop(key)
    assert key in check(cfg)


def test_setup_gauss_with_droplets_success():
    assert check(variant(init_strategy="SETUP_GAUSS", bec_droplets_x=1.0, bec_droplets_y=1.0,
                         bec_droplets_z=1.0)) is None


@pytest.mark.parametrize("strategy", ["BINARY_FILE", "TEXT_FILE"])
def test_file_init_needs_load_filename(strategy):
    assert "load_filename" in check(variant(init_strategy=strategy))
    assert check(variant(init_strategy=strategy, load_filename="test")) is None


def test_wrong_init_strategy_throws():
    assert "Unknown initialization strategy" in check(variant(init_strategy="WRONG_NAME"))


def test_unknown_calc_strategy_is_ignored():
    # complex_param.hpp:38-47: unknown strings keep the default (FS)
    assert check(variant(calc_strategy="XYZ", edd_start=1.5, edd_stop=1.4)) is None


@pytest.mark.parametrize("pot", ["MEXICAN", "MEXICAN_FREE", "MEXICAN_ASYMETRIC", "MEXICAN_ASYMETRIC_FREE"])
def test_mexican_family_needs_dd(pot):
    assert "dd" in check(variant(erase=["dd"], pote_strategy=pot))
    assert check(variant(pote_strategy=pot, dd=0.5)) is None


@pytest.mark.parametrize("pot", ["HARMONIC", "CYLINDRICAL", "FREE"])
def test_other_potentials_do_not_need_dd(pot):
    assert check(variant(pote_strategy=pot)) is None


@pytest.mark.parametrize("key", ["omega_x", "omega_y", "omega_z"])
def test_missing_omega_throws(key):
    assert key in check(variant(erase=[key]))


def test_missing_pote_strategy_throws():
    assert "pote_strategy" in check(variant(erase=["pote_strategy"]))


def test_wrong_pote_strategy_throws():
    assert "potential" in check(variant(pote_strategy="IAMNOTREAL"))


def test_error_message_names_the_key():
    err = check({"other_key": 42, "init_strategy": "GAUSS", "calc_strategy": "IT", "iter_imag": 1,
                 "iter_real": 1, "n_atoms": 1, "m": 1, "edd": 1.0})
    assert err == "Missing required parameter: dx"


@pytest.mark.parametrize("field,value", [("nx", 0), ("dx", -1.0), ("n_atoms", -5), ("m", 0.0)])
def test_check_params_positivity(field, value):
    # file_manager.cpp:216-228
    assert "positive" in check(variant(**{field: value}))


def test_malformed_json_is_an_error_not_a_crash():
    assert "JSON" in hip.validate_params_json('{"nx": 64,, }')
    assert "JSON" in hip.validate_params_json('{"nx": ')


def test_precision_extension():
    assert check(variant(precision="fp32")) is None
    assert "precision" in check(variant(precision="fp16"))
