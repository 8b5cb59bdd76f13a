"""bench.py's command line on the CPU: more than one GPU is refused with a message (the build
measures one GPU), and --help lists the legs' options."""
import os
import subprocess
import sys

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _run(*args):
    return subprocess.run([sys.executable, os.path.join(REPO, "bench.py"), *args], capture_output=True, text=True,
                          timeout=300, cwd=REPO)


def test_more_than_one_gpu_is_refused():
    r = _run("--gpus", "2")
    assert r.returncode != 0
    assert "one GPU" in (r.stdout + r.stderr) or "1 GPU" in (r.stdout + r.stderr)


def test_help_lists_every_leg():
    r = _run("--help")
    assert r.returncode == 0
    for opt in ("--steps", "--warmup", "--ts-N", "--batch", "--bf16-steps", "--ps-N", "--baryon-N", "--grid-hi",
                "--only", "--autotune"):
        assert opt in r.stdout, opt
