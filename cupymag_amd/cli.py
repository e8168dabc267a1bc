"""``cupymag <config.yaml>`` command line (cupymag/cli.py:24-67): same
arguments, same exit codes (0 ok, 1 on a missing file or any error)."""
import argparse
import os
import sys
from pathlib import Path

from . import __version__


def main(argv=None):
    parser = argparse.ArgumentParser(
        description="CuPyMag on MI355X: GPU finite-element micromagnetics with magnetostriction",
        formatter_class=argparse.RawDescriptionHelpFormatter)
    parser.add_argument("config_file", nargs="?", default="config.yaml",
                        help="Path to YAML configuration file (default: config.yaml)")
    parser.add_argument("--version", action="version", version=f"CuPyMag v{__version__} (cupymag_amd)")
    parser.add_argument("--max-steps", type=int, default=None, help="cap on the total number of LLG steps")
    parser.add_argument("--engine", choices=("auto", "fp64", "fast64", "mixed"), default="auto",
                        help="step engine: auto follows simulation.precision (DP: fast64, FP64 CG on SELL operators; "
                             "SP: mixed, FP32 CG in FP64 refinement, multigrid where the mesh has a hierarchy); "
                             "fp64: CSR operators, the reference's sequence of solves")
    args = parser.parse_args(argv)
    config_path = Path(args.config_file).resolve()
    if not config_path.exists():
        print(f"Error: Configuration file '{config_path}' not found.")
        print("Please create a config.yaml file or copy one from the examples/ directory.")
        return 1
    try:
        os.environ["CUPYMAG_CONFIG_PATH"] = str(config_path)
        from .core.Micromagnetics import main as run_simulation
        run_simulation(config_path, max_steps=args.max_steps, engine=args.engine)
        return 0
    except KeyboardInterrupt:
        print("\nSimulation interrupted by user.")
        return 1
    except Exception as e:  # the reference maps every error to exit code 1
        print(f"Error: {e}")
        return 1


if __name__ == "__main__":
    sys.exit(main())
