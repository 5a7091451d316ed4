"""The C ABI library and the tb3 driver without a GPU: the library loads,
exports every entry point include/tb3.h declares, and its error paths return
status codes instead of crashing (no compute call is made here)."""
import ctypes
import os
import re
import subprocess

import pytest

import threebodymdalgorithms_amd as tb3
from threebodymdalgorithms_amd import _abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "tb3.h")


def declared_symbols():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:tb3_status|const char \*)\s*(tb3_\w+)\s*\(", text, re.M)))


def has_gpu():
    return tb3.device_count() > 0


def test_library_exports_every_declared_symbol():
    lib = _abi.load()
    names = declared_symbols()
    assert len(names) >= 15
    for name in names:
        assert hasattr(lib, name), name
        assert name in _abi.SIGNATURES, f"{name} missing from the ctypes table"
    # and the ctypes table declares nothing the header lacks
    assert set(_abi.SIGNATURES) == set(names)


def test_exported_dynamic_symbols():
    out = subprocess.run(["nm", "-D", "--defined-only", _abi.LIB_PATH], capture_output=True, text=True, check=True)
    exported = set(re.findall(r"\bT (tb3_\w+)", out.stdout))
    assert set(declared_symbols()) <= exported


def test_version_string():
    v = _abi.load().tb3_version().decode()
    assert v.startswith("tb3 ") and "gfx950" in v


def test_null_context_is_einval():
    lib = _abi.load()
    assert lib.tb3_destroy(None) == _abi.EINVAL
    assert lib.tb3_configure(None, 0, 1.0, 0.0, None, 0) == _abi.EINVAL
    assert lib.tb3_set_particles(None, 0, None, None, None, None, None) == _abi.EINVAL
    assert lib.tb3_step(None, 0.001, None, None, None) == _abi.EINVAL
    assert lib.tb3_compute_forces(None, None) == _abi.EINVAL
    assert lib.tb3_energy(None, None, None) == _abi.EINVAL
    assert lib.tb3_device_count(None) == _abi.EINVAL
    assert lib.tb3_create(0, None) == _abi.EINVAL
    assert lib.tb3_get_tuning(None, None, None, None) == _abi.EINVAL
    assert lib.tb3_force_scale(None, None) == _abi.EINVAL
    assert lib.tb3_set_tuning(None, 1.0, 8) == _abi.EINVAL
    assert lib.tb3_set_cell_slots(None, -1) == _abi.EINVAL
    assert lib.tb3_get_cell_slots(None, None, None) == _abi.EINVAL


@pytest.mark.skipif(has_gpu(), reason="checks the no-device error path")
def test_create_without_device_reports_enodev():
    lib = _abi.load()
    ctx = ctypes.c_void_p()
    st = lib.tb3_create(0, ctypes.byref(ctx))
    assert st == _abi.ENODEV
    assert not ctx.value
    assert lib.tb3_last_error(None).decode()
    with pytest.raises(tb3.TB3Error) as e:
        tb3.Engine(0)
    assert e.value.status == _abi.ENODEV


def run_cli(*args):
    return subprocess.run([_abi.CLI_PATH, *args], capture_output=True, text=True, timeout=60)


def test_cli_missing_csv_fails_cleanly(tmp_path):
    r = run_cli("-a", "p3bca", "-csv", str(tmp_path / "nope.csv"))
    assert r.returncode == 1
    assert "cannot open" in r.stderr


def test_cli_bad_algorithm_prints_help():
    r = run_cli("-a", "fancy")
    assert r.returncode == 1
    assert "Help:" in r.stdout


def test_cli_flag_without_value_prints_help():
    r = run_cli("-i")
    assert r.returncode == 1 and "Help:" in r.stdout


def test_cli_unparsable_number_prints_help():
    r = run_cli("-d", "abc")
    assert r.returncode == 1 and "Help:" in r.stdout


def test_cli_unknown_flags_warn_and_continue(tmp_path):
    src = tmp_path / "in.csv"
    src.write_text("ID, pX, pY, pZ, vX, vY, vZ, aX, aY, aZ, m\n0, 1, 2, 3, 0, 0, 0, 0, 0, 0, 1\n")
    r = run_cli("-zzz", "stray", "-decomp", "naive", "-op", "x.json", "-csv", str(src))
    assert "CLI parser: unrecognized option: zzz" in r.stdout
    assert "CLI parser: unrecognized option: stray" in r.stdout
    if not has_gpu():
        assert r.returncode == 1 and "tb3_create" in r.stderr


def test_cli_malformed_csv(tmp_path):
    src = tmp_path / "bad.csv"
    src.write_text("ID, pX\n0, 1\n")
    r = run_cli("-csv", str(src))
    assert r.returncode == 1 and "expected at least 11 columns" in r.stderr


def test_csv_round_trip(tmp_path):
    import numpy as np
    ids = np.array([3, 1, 2], np.int32)
    pos = np.array([[1.5, 2.25, -3.0], [0.1, 0.2, 0.3], [1e-5, 123456789.0, 7.0]])
    z = np.zeros((3, 3))
    path = tmp_path / "p.csv"
    tb3.write_csv(path, ids, pos, z, z, np.ones(3))
    lines = path.read_text().splitlines()
    assert lines[0] == "ID, pX, pY, pZ, vX, vY, vZ, aX, aY, aZ, m"
    assert lines[1].startswith("1, 0.1, 0.2, 0.3")
    assert lines[3].startswith("3, 1.5, 2.25, -3")
    assert "1.23457e+08" in lines[2] and "1e-05" in lines[2]
    rid, rpos, *_ = tb3.read_csv(path)
    assert list(rid) == [1, 2, 3]
