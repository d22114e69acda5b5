"""Build and run the full-multiply benchmark front end for the 32-bit, hybrid
and/or 64-bit pipelines, appending every pipeline's rows to one CSV (the
reference's scripts/run_gpu_bench.py interface: pipelines, --no-build, --csv,
and the benchmark's own arguments after the options).

    python -m birdwing_amd.run_gpu_bench --limb-bits all 16-24 --iters 20
    python -m birdwing_amd.run_gpu_bench --limb-bits 64bit --csv out.csv -- 20 22

The rows carry the column set of bench/gpu_full_multiply_benchmark.cpp, so the
reference's plotting scripts read the CSV unchanged.
"""
import argparse
import os
import subprocess
import sys

APPS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "apps")
PIPELINES = ("32", "hybrid", "64bit")
ALIASES = {"64": "hybrid", "64native": "64bit"}     # older names the reference still accepts


def binary(pipeline):
    return os.path.join(APPS, "bin", "bench_full_multiply_%s" % pipeline)


def build(pipelines):
    targets = ["bin/bench_full_multiply_%s" % p for p in pipelines]
    subprocess.run(["make", "-C", APPS] + targets, check=True)


def selected(choice):
    if choice == "both":
        return ["32", "hybrid"]
    if choice == "all":
        return list(PIPELINES)
    return [ALIASES.get(choice, choice)]


def main(argv=None):
    ap = argparse.ArgumentParser(description="Run the full-multiply benchmark for 32-bit, hybrid and/or 64-bit builds.")
    ap.add_argument("--limb-bits", default="32", choices=PIPELINES + tuple(ALIASES) + ("both", "all"),
                    help="pipeline(s): 32 (default), hybrid, 64bit, both (=32+hybrid), all")
    ap.add_argument("--no-build", action="store_true", help="skip make; use the binaries already built")
    ap.add_argument("--csv", default="gpu_multiply_bench.csv", help="output CSV path")
    ap.add_argument("bench_args", nargs=argparse.REMAINDER, help="arguments for the benchmark (e.g. 16-24 --iters 20)")
    args = ap.parse_args(argv)
    extra = args.bench_args[1:] if args.bench_args[:1] == ["--"] else args.bench_args
    pipes = selected(args.limb_bits)
    if not args.no_build:
        build(pipes)
    for i, p in enumerate(pipes):
        cmd = [binary(p), "--csv", args.csv] + (["--append"] if i > 0 else []) + extra
        print("+", " ".join(cmd), flush=True)
        rc = subprocess.run(cmd).returncode
        if rc:
            return rc
    return 0


if __name__ == "__main__":
    sys.exit(main())
