"""The reference's file protocol (src/main.cu:403-602) on the committed fixtures of tests/golden/.

CPU: the oracle reproduces the committed INFO_MX_OUT rows; RunControllerEngine() reports each input
error the reference reports (src/main.cu:295-300, 450-462, src/config.cu) with return code 1 and the
message appended to error.log, before touching the GPU.
GPU: RunControllerEngine() in a scratch working directory writes INFO_MX_OUT.bin bit-identical to the
committed rows (FP64 chain), and the stage dumps of dump_diagnostics."""
import ctypes
import os

import numpy as np
import pytest

import fixtures
from controllerengine_amd import _abi
from controllerengine_amd.metrics import compare
from oracle import oracle


@pytest.mark.parametrize("name", fixtures.NAMES)
def test_oracle_reproduces_fixture(name):
    p = fixtures.problem(name)
    want, _ = fixtures.expected(name)
    d = p.derived()
    rows = []
    for r in range(int(p.config["n_rounds"])):
        g, info = p.round_rows(r, d.n_threads)
        rows.append(oracle.evaluate(p, rnd=r, gains=g, info=info)[0])
    got = np.concatenate(rows).ravel()
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


def run_in(path):
    cwd = os.getcwd()
    os.chdir(path)
    try:
        return _abi.lib().RunControllerEngine()
    finally:
        os.chdir(cwd)


def errlog(path):
    with open(os.path.join(path, "error.log")) as f:
        return f.read()


def _short_cuda(files):
    files["CUDA_MX.bin"] = files["CUDA_MX.bin"][:-4]


def _bad_key(files):
    files["ControllerEngine.config"] += b"colour = blue\n"


def _short_a(files):
    files["DISCRETE_A.bin"] = files["DISCRETE_A.bin"][:-8]


def _bad_digit(files):
    dc = np.frombuffer(files["DC_INFO.bin"], dtype=np.int32).copy()
    dc[np.flatnonzero(dc)[0]] = 3 * 1296 + 0 * 36 + 30          # MUL state * BASE_PARAMS[29]
    files["DC_INFO.bin"] = dc.tobytes()


@pytest.mark.parametrize("name,drop,edit,msg", [
    ("c1_batch", ["ControllerEngine.config"], None, "ControllerEngine error: Cannot open ControllerEngine.config"),
    ("c1_batch", ["STATESPACE_B.bin"], None, "ControllerEngine error: Cannot open STATESPACE_B.bin"),
    ("c1_batch", [], _bad_key, "ControllerEngine error: ControllerEngine.config: unknown key `colour`"),
    ("c1_batch", [], _short_cuda, "ControllerEngine error: CUDA_MX.bin is 30716 bytes, expected 30720"),
    ("c2_serial", [], _short_a, "ControllerEngine error: DISCRETE_A.bin is 120 bytes, expected 128 (2 x 64)"),
    ("c1_batch", [], _bad_digit, "instruction digit 30 needs BASE_PARAMS[29] but n_params is"),
    ("c1_batch", ["INFO_MX.bin"], None, "ControllerEngine error: Cannot open INFO_MX.bin"),
])
def test_run_reports_input_errors(tmp_path, name, drop, edit, msg):
    d = fixtures.write(name, str(tmp_path), drop=drop, edit=edit)
    with open(os.path.join(d, "error.log"), "w") as f:
        f.write("earlier line\n")
    assert run_in(d) == 1
    log = errlog(d)
    assert log.startswith("earlier line\n")          # appended, not truncated
    assert msg in log, log
    assert not os.path.exists(os.path.join(d, "INFO_MX_OUT.bin"))


def test_exports_match_header():
    import re
    hdr = open(os.path.join(fixtures.HERE, "..", "include", "ControllerEngine.h")).read()
    declared = set(re.findall(r"CE_EXPORT\s+[\w\s\*]+?\b(\w+)\s*\(", hdr))
    assert declared == set(_abi.EXPORTS)
    lib = _abi.lib()
    for name in declared:
        assert hasattr(lib, name), name
    # every symbol the library needs resolves at load time (no missing internal definitions)
    ctypes.CDLL(_abi.LIB_PATH, mode=os.RTLD_NOW | os.RTLD_LOCAL)


# ------------------------------------------------------------------------------------------ GPU
@pytest.mark.gpu
@pytest.mark.parametrize("name", fixtures.NAMES)
def test_run_matches_fixture_fp64(tmp_path, name, monkeypatch, engine):
    monkeypatch.setenv("CE_PRECISION", "fp64")
    d = fixtures.write(name, str(tmp_path))
    print("RunControllerEngine fp64", name, flush=True)
    assert run_in(d) == 0, errlog(d) if os.path.exists(os.path.join(d, "error.log")) else ""
    got = np.fromfile(os.path.join(d, "INFO_MX_OUT.bin"), dtype=np.float32)
    want, _ = fixtures.expected(name)
    assert got.size == want.size
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    assert open(os.path.join(d, "params.log")).read().startswith("BASE_PARAMS_HOST[0] = ")


@pytest.mark.gpu
@pytest.mark.parametrize("name", fixtures.NAMES)
def test_run_mixed_matches_margins(tmp_path, name, monkeypatch, engine):
    monkeypatch.setenv("CE_PRECISION", "mixed")
    d = fixtures.write(name, str(tmp_path))
    print("RunControllerEngine mixed", name, flush=True)
    assert run_in(d) == 0
    p = fixtures.problem(name)
    dv = p.derived()
    iw = dv.info_mx_width
    got = np.fromfile(os.path.join(d, "INFO_MX_OUT.bin"), dtype=np.float32).reshape(-1, iw)
    want = fixtures.expected(name)[0].reshape(-1, iw)
    r = compare(p, want.astype(np.float64), got.astype(np.float64))
    assert r["max_dgm_db"] <= 0.01 and r["max_dpm_deg"] <= 0.01 and r["max_db"] <= 1e-4, r
    assert r["passfail_mismatch"] == 0 and r["stable_mismatch"] == 0, r


@pytest.mark.gpu
def test_run_diagnostic_dumps(tmp_path, monkeypatch, engine):
    monkeypatch.setenv("CE_PRECISION", "fp64")

    def dump_on(files):
        files["ControllerEngine.config"] = files["ControllerEngine.config"].replace(
            b"dump_diagnostics = false", b"dump_diagnostics = true")
    d = fixtures.write("c1_batch", str(tmp_path), edit=dump_on)
    print("RunControllerEngine diagnostics c1_batch", flush=True)
    assert run_in(d) == 0
    p = fixtures.problem("c1_batch")
    dv = p.derived()
    n, nt, T = dv.n_threads, dv.total_dim, int(p.config["time"])
    _, _, dg = oracle.evaluate(p, diag=True)
    sizes = {"YPLANT_OUT.bin": T, "LINA_MX_OUT.bin": nt * nt, "EIGEN_MX_OUT.bin": 2 * nt,
             "ATA_MX_OUT.bin": 4 * nt * nt, "ATB_MX_OUT.bin": 2 * nt}
    keys = {"YPLANT_OUT.bin": "y", "LINA_MX_OUT.bin": "lina", "EIGEN_MX_OUT.bin": "eig", "ATA_MX_OUT.bin": "ata",
            "ATB_MX_OUT.bin": "atb"}
    for f, per in sizes.items():
        a = np.fromfile(os.path.join(d, f), dtype=np.float32)
        assert a.size == n * per, f
        ref = np.asarray(dg[keys[f]], dtype=np.float32).ravel()
        assert np.array_equal(a.view(np.uint32), ref.view(np.uint32)), f
