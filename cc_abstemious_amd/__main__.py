"""Command line: ``python -m cc_abstemious_amd input.json`` runs the analysis
and writes the field results at the first frequency (src/main.rs:1-9)."""
import sys

from .analysis import Analysis


def main(argv=None) -> int:
    argv = sys.argv[1:] if argv is None else argv
    if len(argv) != 1:
        print("usage: python -m cc_abstemious_amd <input.json>", file=sys.stderr)
        return 2
    analysis = Analysis()
    analysis.input_from_file(argv[0])
    analysis.run()
    err = analysis.write_results_at_frequency(0)
    if err is not None:
        print(f"could not write results: {err}", file=sys.stderr)
        return 1
    return 0


if __name__ == "__main__":
    sys.exit(main())
