"""The reference's own sample programs (sample/C++ and sample/C), compiled
unchanged against libp3dfft3_hip.so by tests/samples/Makefile, run on one
rank with their usual input files.  Each program checks its own known
answer and prints "Results are correct" / "Results are incorrect"; the
reference's test scripts grep for the same words (extra/TESTING_README).

These programs pass host arrays, so this exercises the staged host path of
the library as an unmodified P3DFFT++ program would."""
import os
import re
import subprocess
import tempfile

import pytest

pytestmark = pytest.mark.gpu

BIN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "samples", "bin")

STD3D = {"stdin": "64 64 64 1 2\n"}
CASES = [
    ("cxx_test3D_r2c", STD3D),
    ("cxx_test3D_r2c_single", STD3D),
    ("cxx_test3D_c2c", STD3D),
    ("cxx_test3D_c2c_inplace", STD3D),
    ("cxx_test_deriv", {"stdin": "64 48 40 1 1 1\n"}),
    ("cxx_test_deriv", {"stdin": "32 24 20 1 1 2\n"}),
    ("cxx_test_deriv", {"stdin": "32 24 20 1 1 3\n"}),
    ("cxx_test1D_cos", {"trans.in": "33 8 6 0 2\n0 1 2\n1 0 2\n"}),
    ("cxx_test1D_cos_complex", {"trans.in": "33 8 6 0 2\n0 1 2\n2 0 1\n"}),
    ("cxx_test1D_sin", {"trans.in": "31 8 6 0 2\n0 1 2\n0 1 2\n"}),
    ("cxx_test_transplan", {"trans.in": "64 8 6 0 2\n0 1 2\n1 2 0\n"}),
    ("cxx_test_transplan", {"trans.in": "8 48 6 1 2\n2 1 0\n0 2 1\n"}),
    ("c_test3D_r2c", STD3D),
    ("c_test3D_r2c_single", STD3D),
    ("c_test3D_c2c", STD3D),
    ("c_test3D_c2c_inplace", STD3D),
    ("c_test_deriv", {"stdin": "64 48 40 1 1 2\n"}),
    ("c_test1D_cos", {"trans.in": "33 8 6 0 2\n0 1 2\n0 1 2\n"}),
    ("c_test1D_cos_complex", {"trans.in": "33 8 6 0 2\n0 1 2\n1 0 2\n"}),
    ("c_test1D_r2c", {"trans.in": "64 8 6 0 2\n0 1 2\n0 1 2\n"}),
    ("c_test2D+empty", {"stdin": "64 64 64 1 2\n"}),
    ("c_test3D_r2c_memord", {"memord3d": "32 32 32 1 2\n2 1 0\n0 2 1\n"}),
    ("c_test3D_r2c_memord", {"memord3d": "48 48 48 1 2\n1 2 0\n2 0 1\n"}),
]


def run_sample(prog, files):
    exe = os.path.join(BIN, prog)
    if not os.path.exists(exe):
        pytest.skip("samples not built (tests/samples/Makefile needs the reference sources)")
    with tempfile.TemporaryDirectory() as d:
        for name, text in files.items():
            with open(os.path.join(d, name), "w") as f:
                f.write(text)
        print("case:", prog, files, flush=True)
        r = subprocess.run([exe], cwd=d, capture_output=True, text=True, timeout=120)
    return r, r.stdout + r.stderr


@pytest.mark.parametrize("prog,files", CASES, ids=[f"{p}-{i}" for i, (p, _) in enumerate(CASES)])
def test_reference_sample(prog, files):
    r, out = run_sample(prog, files)
    assert r.returncode == 0, out[-3000:]
    assert "Results are correct" in out, out[-3000:]
    assert "incorrect" not in out, out[-3000:]


TIMER_LINES = ["Reorder_trans", "Reorder_out", "Reorder_in", "Trans_exec", "Packsend", "Packsend_trans",
               "Unpackrecv", "Alltoall"]


def timer_report(out):
    """{phase: (avg, min, max)} from the TIMERS report lines."""
    rep = {}
    for line in out.splitlines():
        m = re.match(r"^(\w+)\s*\(([-\d.e+]+) ([-\d.e+]+) ([-\d.e+]+)\)", line.strip())
        if m:
            rep[m.group(1)] = tuple(float(m.group(i)) for i in (2, 3, 4))
    return rep


@pytest.mark.parametrize("prog", ["cxx_timers_test3D_c2c", "cxx_timers_test3D_c2c_inplace"])
def test_sample_timers_report(prog):
    """The samples built with -DTIMERS call timers.init() before their timing
    loop and timers.print(MPI_COMM_WORLD) after it (sample/C++/test3D_c2c.C:
    223-224, 263-264): the report lists the reference's phases, the GPU time
    of the fused reorder + transform passes lands in Reorder_trans /
    Trans_exec, and the exchange phases of a 1x1 grid stay 0."""
    r, out = run_sample(prog, {"stdin": "64 64 64 1 3\n"})
    assert r.returncode == 0, out[-3000:]
    assert "Results are correct" in out, out[-3000:]
    assert "TIMERS (avg/min/max)" in out, out[-3000:]
    rep = timer_report(out)
    for k in TIMER_LINES:
        assert k in rep, (k, out[-2000:])
    busy = rep["Reorder_trans"][0] + rep["Trans_exec"][0]
    assert busy > 0.0, out[-2000:]
    # 3 repetitions of a forward and a backward 64^3 C2C: well under a second
    assert busy < 1.0
    for k in ("Packsend", "Packsend_trans", "Unpackrecv", "Alltoall"):
        assert rep[k] == (0.0, 0.0, 0.0)


def test_bench_c2c_timers():
    """sample/C++/bench_c2c.C built with -DTIMERS: its proc-grid sweep on one
    rank, the transform and the TIMERS report."""
    r, out = run_sample("cxx_timers_bench_c2c", {"stdin": "64 64 64 1 2\n"})
    assert r.returncode == 0, out[-3000:]
    assert "incorrect" not in out, out[-3000:]
    assert "TIMERS (avg/min/max)" in out, out[-3000:]
    assert timer_report(out)["Reorder_trans"][0] + timer_report(out)["Trans_exec"][0] > 0.0
