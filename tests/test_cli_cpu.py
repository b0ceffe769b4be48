"""The epirust-gpu command line on paths that end before any device work: flags as
engine-app parses them (engine-app/src/main.rs:60-87), the config read error
(main.rs:100-105 `Config::read(...).expect("Failed to read config file")`) and a
device-less machine failing loudly."""
import json
import os
import subprocess

import pytest

import epirust_amd as E


def run(args, cwd):
    env = dict(os.environ)
    env.pop("EPI_OUTPUT_DIR", None)
    return subprocess.run([E.EXE_PATH] + args, cwd=str(cwd), env=env, stdout=subprocess.PIPE,
                          stderr=subprocess.PIPE, text=True, timeout=60)


def test_help_and_unknown_flag(tmp_path):
    p = run(["--help"], tmp_path)
    assert p.returncode == 0 and "usage: epirust-gpu" in p.stderr
    p = run(["--frobnicate"], tmp_path)
    assert p.returncode == 2 and "unexpected argument '--frobnicate'" in p.stderr
    p = run(["-c"], tmp_path)  # a flag without its value
    assert p.returncode == 2


def test_only_standalone_mode(tmp_path):
    for m in ("kafka", "mpi"):
        p = run(["-m", m], tmp_path)
        assert p.returncode == 2 and "standalone only" in p.stderr
    # clap's ValueEnum names (kebab case of the variants, main.rs:41-46), case-sensitive
    p = run(["--mode", "STANDALONE"], tmp_path)
    assert p.returncode == 2 and "invalid value 'STANDALONE'" in p.stderr
    p = run(["--mode", "standalone", "-c", str(tmp_path / "missing.json")], tmp_path)
    assert p.returncode == 1 and p.stdout.startswith("Standalone")


def test_missing_and_invalid_config(tmp_path):
    p = run(["-c", str(tmp_path / "missing.json")], tmp_path)
    assert p.returncode == 1 and "Failed to read config file" in p.stderr
    bad = tmp_path / "bad.json"
    bad.write_text('{"hours": 10}')
    p = run(["--config=" + str(bad)], tmp_path)
    assert p.returncode == 1 and "missing field" in p.stderr
    # default config path is config/default.json relative to the working directory
    p = run([], tmp_path)
    assert p.returncode == 1 and "Failed to read config file" in p.stderr


def test_device_less_run_fails_and_writes_nothing(cfg1, tmp_path):
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    cfg = tmp_path / "c.json"
    cfg.write_text(json.dumps(cfg1))
    p = run(["-c", str(cfg), "-o", str(tmp_path / "o")], tmp_path)
    assert p.returncode == 1 and ("gfx950" in p.stderr or "HIP" in p.stderr)
    assert not (tmp_path / "o").exists()
