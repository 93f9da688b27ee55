"""Host-side logic that needs no GPU: the autotuner's choice rule and the
bench's reading of the committed traffic profile."""
import importlib.util
import os

from starsmasher_amd.autotune import _pick

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_pick_keeps_incumbent_within_margin():
    table = [(64, 1, 0, 100.0), (256, 1, 0, 100.5), (128, 1, 0, 99.0)]
    assert _pick(table, (256, 1, 0), 0.02)[:3] == (256, 1, 0)     # 100.5 <= 99 * 1.02
    assert _pick(table, (256, 1, 0), 0.01)[:3] == (128, 1, 0)     # 100.5 >  99 * 1.01
    assert _pick(table, (64, 2, 0), 0.02)[:3] == (128, 1, 0)      # incumbent not in table
    assert _pick(table, (256, 1), 0.02)[:2] == (256, 1)           # batch rows: (wg, ipl)


def _bench():
    spec = importlib.util.spec_from_file_location("bench_mod", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_profiled_traffic_reads_committed_profile():
    b = _bench()
    t = b.profiled_traffic(256, 1)
    assert t is not None and t > 1e9
    assert b.profiled_traffic(64, 4) is None


def test_bench_refuses_several_gpus():
    import subprocess
    import sys
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and r.stdout == ""


def test_torch_tensors_refused_across_hip_runtimes():
    """torch bundles its own HIP runtime.  When the library is loaded first
    it binds ROCm's runtime, and importing torch then maps a second one: a
    torch tensor's pointer (and torch's stream) would reach a runtime that
    does not know them, so the device API refuses torch tensors with
    TypeError instead of enqueueing work on them.  Run in a fresh process so
    the load order is known."""
    import subprocess
    import sys
    code = (
        "import sys; sys.path.insert(0, %r)\n"
        "from starsmasher_amd import _lib, gravity\n"
        "_lib.load()\n"
        "import torch\n"
        "rts = gravity.hip_runtimes()\n"
        "assert len(rts) == 2, rts\n"
        "try:\n"
        "    gravity._dev_ptr(torch.zeros(8, dtype=torch.float64), 'x', 8)\n"
        "except TypeError as e:\n"
        "    assert 'HIP runtimes' in str(e), e\n"
        "else:\n"
        "    raise SystemExit('torch tensor accepted across runtimes')\n"
        "assert gravity._dev_ptr(12345, 'x') == 12345\n"
        "print('refused')\n" % ROOT)
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    assert r.stdout.strip() == "refused"
