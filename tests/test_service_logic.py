"""The Poisson service's scheduling (p3dfft.3_amd/csrc/host/serve.cpp) on the
CPU against a mock HIP runtime whose streams execute their commands
asynchronously and concurrently, one thread per stream, and whose frees and
pinned allocations synchronize the whole device (tests/svc): for one and
four lanes, several in-flight limits, pool sizes and device speeds, every
request completes, each request's passes form the chain rhs -> ... ->
solution in device order, no other pass writes a buffer between its writer
and its reader, and no more than max_inflight requests hold working memory
at once."""
import os
import subprocess

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "svc")


def test_service_scheduling_with_mock_device():
    subprocess.check_call(["make", "-C", HERE, "-s"])
    r = subprocess.run([os.path.join(HERE, "svc_test")], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.count("ok:") == 48
