"""bench.py's command-line contract that needs no GPU: more than one GPU is
refused with exit status 2 and a message, before anything touches a
device."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _run(*args):
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args], capture_output=True, text=True,
                          timeout=120, cwd=ROOT)


@pytest.mark.parametrize("args, words", [(("--gpus", "2"), "one MI355X"), (("--gpus", "8", "--tune"), "one MI355X")])
def test_refused_before_any_device_work(args, words):
    r = _run(*args)
    assert r.returncode == 2, (r.returncode, r.stderr)
    assert words in r.stderr
    assert r.stdout == ""
