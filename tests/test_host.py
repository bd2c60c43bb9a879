"""The C++ LINPACK host (hpcc_fpga_amd/host/Linpack_hip): command line, test
mode, error exits (CPU) and a full run with the JSON dump (GPU).

The stdout layout is checked with the reference's own result parser regex
(scripts/evaluation/parse_raw_to_csv.py:61-69, restated here), and the JSON
keys with the reference's host-integration test
(LINPACK/tests/test_kernel_functionality_and_host_integration.cpp:100-126).
"""
import json
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "hpcc_fpga_amd", "host", "Linpack_hip")
EXE_DP = os.path.join(ROOT, "hpcc_fpga_amd", "host", "Linpack_hip_dp")

NUM = r"((\d|\.|\+|-|e)+|nan)"
LINPACK_REGEX = (
    r"Version:\s+(?P<version>.+)\n"
    r"(.*\n)+Matrix\sSize\s+(?P<size>\d+)"
    r"(.*\n)+Device\s+(?P<device>.+)\n"
    r"(.*\n)+\s+norm\.\sresidual\s+res\.\serror\s+mach\.\seps\n"
    rf"\s+(?P<error>{NUM})\s+(?P<resid>{NUM})\s+(?P<epsilon>(\d|\.|\+|-|e)+)\n"
    r"(.*\n)+\sMethod\s+best\s+mean\s+GFLOPS\s+\n"
    r"\stotal\s+(?P<total_best_time>(\d|\.|\+|-|e)+)\s.+\s+(?P<total_avg_time>(\d|\.|\+|-|e)+)\s.+\s+(?P<total_gflops>(\d|\.|\+|-|e)+)\s.+\s+\n"
    r"\sGEFA\s+(?P<lu_best_time>(\d|\.|\+|-|e)+)\s.+\s+(?P<lu_avg_time>(\d|\.|\+|-|e)+)\s.+\s+(?P<lu_gflops>(\d|\.|\+|-|e)+)\s.+\s+\n"
    r"\sGESL\s+(?P<sl_best_time>(\d|\.|\+|-|e)+)\s.+\s+(?P<sl_avg_time>(\d|\.|\+|-|e)+)\s.+\s+(?P<sl_gflops>(\d|\.|\+|-|e)+)\s.+\s+\n"
)


def run(*args, timeout=120):
    if not os.path.exists(EXE):
        pytest.skip("host executable not built")
    return subprocess.run([EXE, *args], capture_output=True, text=True, timeout=timeout)


def test_help():
    r = run("-h")
    assert r.returncode == 0
    assert "-m arg" in r.stdout and "--dump-json arg" in r.stdout


def test_test_mode_parses_without_device():
    r = run("-f", "HIP", "--test", "-m", "8", "-b", "7")
    assert r.returncode == 0, r.stderr
    assert "SUCCESSFULLY parsed input parameters!" in r.stdout
    assert re.search(r"Matrix Size\s+1024", r.stdout)
    assert re.search(r"Block Size\s+128", r.stdout)
    assert re.search(r"Communication Type\s+HIP", r.stdout)


@pytest.mark.parametrize("args,msg", [
    (["-m", "4"], "Kernel file must be given with option -f"),
    (["-f", "HIP", "-p", "2", "--test"], "MPI size not dividable by P=2"),
    (["-f", "HIP", "--comm-type", "NOPE", "--test"], "could not be converted from string"),
    (["-f", "xyz", "--test"], "could not be autodetected"),
    (["-f", "HIP", "--bogus", "--test"], "does not exist"),
])
def test_setup_errors_exit_1(args, msg):
    r = run(*args)
    assert r.returncode == 1
    assert msg in r.stderr


def test_unhandled_backend_exits_1():
    """Backends other than HIP and CPU throw in executeKernel, as the
    reference's default case does (linpack_benchmark.hpp:319)."""
    r = run("-f", "PCIE", "-m", "2", "-b", "5", "-n", "1")
    assert r.returncode == 1
    assert "No calculate method implemented for communication type PCIE" in r.stderr


@pytest.mark.parametrize("uniform", [False, True])
def test_cpu_backend_is_the_reference_cpu_path(tmp_path, uniform):
    """-f CPU (cpu_only, communication_types.hpp:58,88) runs the reference's
    own gefa_ref_nopvt/gesl_ref_nopvt (--uniform: gefa_ref/gesl_ref) inside the
    host, with no device: its residual max|x - 1| equals, bit for bit, that of
    the oracle's restatement of the same loops on the same ref-generator
    matrix (mt19937(0), linpack_benchmark.hpp:211-297)."""
    import numpy as np

    from oracle import oracle as O

    n = 160
    out = tmp_path / "c.json"
    args = ["-f", "CPU", "-m", "5", "-b", "5", "-n", "2", "--dump-json", str(out)]
    r = run(*(args + (["--uniform"] if uniform else [])))
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    assert re.search(r"Communication Type\s+CPU", r.stdout)
    assert "Validation: SUCCESS!" in r.stdout
    m = re.search(LINPACK_REGEX, r.stdout)
    assert m and int(m.group("size")) == n
    d = json.loads(out.read_text())
    assert len(d["timings"]["gefa"]) == 2 and d["validated"] is True
    A, b, _, normb = O.generate_ref(n, seed=0, diag_dominant=not uniform)
    if uniform:
        F, ipvt = O.gefa_pvt(A)
        x = O.gesl_pvt(F, b, ipvt)
    else:
        x = O.gesl_nopvt(O.gefa_nopvt(A), b)
    resid, residn = O.residn_ref(x, normb)
    assert d["errors"]["residual"] == resid
    assert d["errors"]["residual_norm"] == pytest.approx(residn, rel=1e-12)


@pytest.mark.gpu
@pytest.mark.parametrize("gen", ["ref", "sized"])
def test_full_run_and_json(tmp_path, gen):
    out = tmp_path / "r.json"
    print(f"host run generator={gen}", flush=True)
    r = run("-f", "HIP", "-m", "8", "-b", "7", "-n", "3", "--generator", gen, "--dump-json",
            str(out))
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    assert "Validation: SUCCESS!" in r.stdout
    m = re.search(LINPACK_REGEX, r.stdout)
    assert m, r.stdout
    assert int(m.group("size")) == 1024
    assert float(m.group("error")) < 1.0
    d = json.loads(out.read_text())
    for key in ("gefa", "gesl"):
        assert len(d["timings"][key]) == 3
    for key in ("gflops", "gflops_lu", "gflops_sl", "t_mean", "t_min", "tlu_mean", "tlu_min",
                "tsl_mean", "tsl_min"):
        assert key in d["results"]
    assert d["validated"] is True
    assert d["errors"]["residual_norm"] < 1
    assert d["errors"]["hpl_scaled_residual"] < 1
    assert d["settings"]["Matrix Size"] == 1024
    assert d["settings"]["FPGA Torus"] == {"P": 1, "Q": 1}


@pytest.mark.gpu
@pytest.mark.parametrize("gen", ["ref", "sized"])
def test_uniform_runs_with_pivoting(tmp_path, gen):
    """--uniform: partial pivoting on the device (gefa_ref's path), validated by
    the reference's residn check (test_host_reference_implementations.cpp:82-89)."""
    out = tmp_path / "u.json"
    print(f"host --uniform generator={gen}", flush=True)
    r = run("-f", "HIP", "-m", "8", "-b", "7", "-n", "2", "--uniform", "--generator", gen,
            "--dump-json", str(out))
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    assert "Validation: SUCCESS!" in r.stdout
    assert re.search(r"Diagonally Dominant\s+No", r.stdout)
    d = json.loads(out.read_text())
    assert d["validated"] is True
    assert d["errors"]["residual_norm"] < 1
    assert d["errors"]["hpl_scaled_residual"] < 1


def test_uniform_rejects_large_blocks():
    r = run("-f", "HIP", "-m", "2", "-b", "9", "--uniform", "--test")
    assert r.returncode == 1
    assert "partial pivoting (--uniform) needs a block size in [32, 256]" in r.stderr + r.stdout


def test_double_build_parses():
    """The DATA_TYPE=double build (Linpack_hip_dp) reports cl_double and
    rejects what it cannot run."""
    r = subprocess.run([EXE_DP, "-f", "HIP", "--test", "-m", "16", "-b", "6"], capture_output=True,
                       text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    assert re.search(r"Data Type\s+cl_double", r.stdout)
    r = subprocess.run([EXE_DP, "-f", "HIP", "--test", "-m", "3", "-b", "5"], capture_output=True,
                       text=True, timeout=60)
    assert r.returncode == 1
    assert "multiple of 64" in r.stderr + r.stdout
    r = subprocess.run([EXE_DP, "-f", "HIP", "--test", "-m", "16", "-b", "6", "--uniform"],
                       capture_output=True, text=True, timeout=60)
    assert r.returncode == 1
    assert "without pivoting" in r.stderr + r.stdout


@pytest.mark.gpu
@pytest.mark.parametrize("gen", ["ref", "sized"])
def test_double_build_full_run(tmp_path, gen):
    out = tmp_path / "dp.json"
    print(f"host DATA_TYPE=double generator={gen}", flush=True)
    r = subprocess.run([EXE_DP, "-f", "HIP", "-m", "16", "-b", "6", "-n", "2", "--generator", gen,
                        "--dump-json", str(out)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    assert "Validation: SUCCESS!" in r.stdout
    d = json.loads(out.read_text())
    assert d["validated"] is True
    assert d["errors"]["epsilon"] < 1e-15  # double epsilon
    assert d["errors"]["residual_norm"] < 1
    assert d["errors"]["hpl_scaled_residual"] < 1
