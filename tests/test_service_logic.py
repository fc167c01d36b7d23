"""The Poisson service's scheduling (p3dfft.3_amd/csrc/host/serve.cpp) on the
CPU against a mock HIP runtime whose device thread executes the stream's
commands asynchronously (tests/svc): for several in-flight limits, pool
sizes and device speeds, every request completes, each request's passes
form the chain rhs -> ... -> solution in device order, no other pass
writes a buffer between its writer and its reader, and no more than four
large-grid requests (dedicated launches) hold working memory at once."""
import os
import subprocess

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "svc")


def test_service_scheduling_with_mock_device():
    subprocess.check_call(["make", "-C", HERE, "-s"])
    r = subprocess.run([os.path.join(HERE, "svc_test")], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.count("ok:") == 24
