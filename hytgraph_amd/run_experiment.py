"""Experiment runner: a YAML configuration in, one JSON result out.

Counterpart of the original project's experiments/run_experiment.py:46-88
(same command line and the same additions to the record: git_commit, a
`runner` block with Python version, platform and UTC timestamp, and the
experiment configuration). The original hands the configuration to a
placeholder executable; here the executable is hyt_experiment (built next to
the libraries), which runs the workload on the GPU and records measurements.

    python -m hytgraph_amd.run_experiment --config exp.yaml --output out/result.json

Scalar entries of the YAML file are passed on as --config key=value
(booleans lower-cased); `experiment_id` and `algorithm` become their own
flags. The process exits with the executable's status (1 = the experiment
failed; its record, with status "failure" and the error, is still written).
"""
from __future__ import annotations

import argparse
import json
import platform
import subprocess
import sys
from datetime import datetime, timezone
from pathlib import Path

import yaml

DEFAULT_EXECUTABLE = Path(__file__).resolve().parent / "lib" / "hyt_experiment"


def parse_args(argv=None) -> argparse.Namespace:
    ap = argparse.ArgumentParser(description="Run one hytgraph_amd experiment")
    ap.add_argument("--config", required=True, type=Path)
    ap.add_argument("--output", required=True, type=Path)
    ap.add_argument("--executable", type=Path, default=DEFAULT_EXECUTABLE)
    return ap.parse_args(argv)


def git_commit() -> str:
    try:
        done = subprocess.run(["git", "rev-parse", "HEAD"], check=True, capture_output=True, text=True,
                              cwd=Path(__file__).resolve().parent)
        return done.stdout.strip()
    except (OSError, subprocess.CalledProcessError):
        return "uncommitted"


def command(executable: Path, output: Path, config: dict) -> list[str]:
    cmd = [str(executable), "--output", str(output),
           "--experiment-id", str(config.get("experiment_id", "experiment")),
           "--algorithm", str(config.get("algorithm", "sssp"))]
    for key, value in sorted(config.items()):
        if key in ("experiment_id", "algorithm"):
            continue
        if isinstance(value, (str, int, float, bool)):
            cmd += ["--config", f"{key}={str(value).lower() if isinstance(value, bool) else value}"]
    return cmd


def main(argv=None) -> int:
    args = parse_args(argv)
    if not args.executable.exists():
        raise SystemExit(f"{args.executable} is missing: build it with `python -m hytgraph_amd.build`")
    with args.config.open("r", encoding="utf-8") as fh:
        config = yaml.safe_load(fh) or {}
    if not isinstance(config, dict):
        raise SystemExit(f"{args.config}: the configuration must be a mapping")
    args.output.parent.mkdir(parents=True, exist_ok=True)
    rc = subprocess.run(command(args.executable, args.output, config)).returncode
    if not args.output.exists():
        raise SystemExit(f"{args.executable} wrote no result (exit status {rc})")
    result = json.loads(args.output.read_text(encoding="utf-8"))
    result["git_commit"] = git_commit()
    result["runner"] = {
        "name": "hytgraph_amd/run_experiment.py",
        "python_version": platform.python_version(),
        "platform": platform.platform(),
        "timestamp_utc": datetime.now(timezone.utc).isoformat(),
    }
    result["experiment_config"] = config
    args.output.write_text(json.dumps(result, indent=2, sort_keys=True) + "\n", encoding="utf-8")
    print(f"Experiment result: {args.output}")
    return rc


if __name__ == "__main__":
    sys.exit(main())
