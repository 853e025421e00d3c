"""Restart files (ufc_rt_restart_save / ufc_rt_restart_load), CPU only:
host-only form (no context).

The layout is checked three ways: a byte-level parse here that follows
RT_Out_RestartSave's record sequence (L5_RT/Output/RT_Out_Restart.f90:
239-277), a Fortran program that reads the file with RT_Out_RestartRestore's
READ sequence (tests/fortran/restart_restore.f90, built with flang), and the
library's own reader.  Errors follow RT_Out_RestartRestore's codes.
"""
import os
import shutil
import struct
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


def _records(b):
    """Splits a Fortran unformatted sequential file into record payloads."""
    out, p = [], 0
    while p < len(b):
        (n,) = struct.unpack_from("<i", b, p)
        out.append(b[p + 4:p + 4 + n])
        (m,) = struct.unpack_from("<i", b, p + 4 + n)
        assert m == n
        p += 8 + n
    return out


def _write(tmp_path, n=11, seed=0):
    from ufc_hip import restart as R
    u = np.random.default_rng(seed).standard_normal(n)
    step = R.StepState(time=0.625, increment=5, lambda_=0.625, residual_norm=3.25e-7,
                       iterations=4, converged=True)
    path = str(tmp_path / "job.rst")
    R.save(path, u, step)
    return path, u, step


def test_record_sequence_of_the_reference_writer(tmp_path):
    path, u, step = _write(tmp_path)
    recs = _records(open(path, "rb").read())
    assert [len(r) for r in recs] == [12, 88, 88, 88, 20, 16, 4]
    assert struct.unpack("<3i", recs[0]) == (1, 20260204, 11)
    assert np.array_equal(np.frombuffer(recs[1], "<f8"), u)
    assert not np.frombuffer(recs[2], "<f8").any() and not np.frombuffer(recs[3], "<f8").any()
    assert struct.unpack("<did", recs[4]) == (0.625, 5, 0.625)
    assert struct.unpack("<dii", recs[5]) == (3.25e-7, 4, 1)
    assert struct.unpack("<i", recs[6]) == (11,)
    assert not os.path.exists(path + ".tmp")


def test_round_trip(tmp_path):
    from ufc_hip import restart as R
    path, u, step = _write(tmp_path)
    u2, s2 = R.load(path, n_dof=u.size)
    assert np.array_equal(u2, u) and s2 == step


def test_fortran_reader_with_the_reference_read_sequence(tmp_path):
    fc = shutil.which("flang") or shutil.which("gfortran")
    if fc is None:
        pytest.skip("no Fortran compiler")
    exe = str(tmp_path / "restart_restore")
    subprocess.run([fc, "-O0", os.path.join(HERE, "fortran", "restart_restore.f90"), "-o", exe],
                   check=True, cwd=str(tmp_path), capture_output=True)
    path, u, step = _write(tmp_path, n=9, seed=4)
    r = subprocess.run([exe, path], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    kv = [ln.split() for ln in r.stdout.splitlines()]
    head = {k: v for k, v in kv if k != "u"}
    assert head["version"] == "1" and head["stamp"] == "20260204" and head["n"] == "9"
    assert float(head["time"]) == 0.625 and head["increment"] == "5"
    assert float(head["lambda"]) == 0.625 and float(head["residual_norm"]) == 3.25e-7
    assert head["iterations"] == "4" and head["converged"] == "1" and head["checksum"] == "9"
    assert float(head["max_abs_va"]) == 0.0
    assert np.array_equal(np.array([float(v) for k, v in kv if k == "u"]), u)


def _patched(tmp_path, offset, fmt, value, name="bad.rst"):
    path, u, _ = _write(tmp_path)
    b = bytearray(open(path, "rb").read())
    struct.pack_into(fmt, b, offset, value)
    bad = str(tmp_path / name)
    open(bad, "wb").write(bytes(b))
    return bad, u.size


@pytest.mark.parametrize("case,offset,fmt,value,code,what", [
    ("version", 4, "<i", 2, 1, "Incompatible file version: 2"),
    ("size", 12, "<i", 0, 1, "Invalid data size"),
    ("checksum", 4 + 12 + 4 + 3 * (8 + 88) + (8 + 20) + (8 + 16) + 4, "<i", 12, 2,
     "Checksum mismatch"),
    ("marker", 4 + 12, "<i", 13, 2, "bad record marker"),
])
def test_damaged_files(tmp_path, case, offset, fmt, value, code, what):
    from ufc_hip import restart as R, UfcError
    bad, n = _patched(tmp_path, offset, fmt, value)
    with pytest.raises(UfcError) as e:
        R.load(bad, n_dof=n)
    assert e.value.code == code and what in e.value.message, str(e.value)


def test_truncated_and_missing_files(tmp_path):
    from ufc_hip import restart as R, UfcError
    path, u, _ = _write(tmp_path)
    b = open(path, "rb").read()
    cut = str(tmp_path / "cut.rst")
    open(cut, "wb").write(b[:-6])
    with pytest.raises(UfcError) as e:
        R.load(cut, n_dof=u.size)
    assert e.value.code == 2 and "Read error" in e.value.message
    with pytest.raises(UfcError) as e:
        R.load(str(tmp_path / "none.rst"), n_dof=u.size)
    assert e.value.code == 1 and "File not found" in e.value.message
    with pytest.raises(UfcError) as e:
        R.load(path, n_dof=u.size + 1)
    assert e.value.code == 1 and "holds 11 DOFs" in e.value.message


def test_save_errors(tmp_path):
    from ufc_hip import restart as R, UfcError
    with pytest.raises(UfcError) as e:
        R.save("", np.zeros(3))
    assert e.value.code == 1 and "Empty filename" in e.value.message
    with pytest.raises(UfcError) as e:
        R.save(str(tmp_path / "no_dir" / "x.rst"), np.zeros(3))
    assert e.value.code == 2 and "Cannot open file" in e.value.message
    with pytest.raises(UfcError) as e:
        R.save(str(tmp_path / "s.rst"), np.zeros(3), with_state=True)
    assert e.value.code == 1 and "needs a context" in e.value.message


def test_save_replaces_the_previous_checkpoint(tmp_path):
    from ufc_hip import restart as R
    path = str(tmp_path / "job.rst")
    R.save(path, np.ones(4), R.StepState(increment=1))
    R.save(path, np.full(4, 2.0), R.StepState(increment=2))
    u, s = R.load(path, n_dof=4)
    assert np.all(u == 2.0) and s.increment == 2
