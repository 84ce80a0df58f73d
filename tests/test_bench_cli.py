"""bench.py's command-line contract that needs no GPU: more than one GPU is
refused with one JSON line and exit status 2, before anything touches a device."""
import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _run(args, env_extra=None):
    env = dict(os.environ)
    env.update(env_extra or {})
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py")] + args,
                          capture_output=True, text=True, timeout=120, env=env, cwd=ROOT)


def test_more_than_one_gpu_is_refused():
    r = _run(["--gpus", "2", "--steps", "1", "--warmup", "0"])
    assert r.returncode == 2, r.stderr
    lines = [l for l in r.stdout.splitlines() if l.strip()]
    assert len(lines) == 1
    msg = json.loads(lines[0])
    assert "one GPU" in msg["error"]


def test_world_size_above_one_is_refused():
    r = _run(["--steps", "1", "--warmup", "0"], {"WORLD_SIZE": "8"})
    assert r.returncode == 2, r.stderr
    assert "error" in json.loads(r.stdout.strip().splitlines()[-1])


def test_help_lists_the_contract_flags():
    r = _run(["--help"])
    assert r.returncode == 0
    for flag in ("--gpus", "--steps", "--warmup", "--mask-stride"):
        assert flag in r.stdout
