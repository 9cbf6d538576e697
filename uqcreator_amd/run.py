"""Run an input file end to end on the GPU (uqk_run_config) and print the run
report as one JSON line.

    python -m uqcreator_amd.run INPUT.toml [--max-steps N] [--precision fp64|mixed]
                                            [--mean-var OUT.csv] [--output PREFIX]

--max-steps and --precision override the file's [problem] maxSteps and
[moment_system] precision (a copy of the file with those keys replaced is run).
--output writes PREFIX.vtk (mean and variance fields), the run log PREFIX and
the restart files PREFIX_moments / PREFIX_duals; an input with [general]
restartFile = "PREFIX" continues from them.
"""
import argparse
import json
import os
import re
import sys
import tempfile

from . import native


def _override(text, table, key, value):
    """Sets key = value in [table] of a TOML text (appends the key if absent)."""
    parts = re.split(r"(?m)^(\[[^\]]+\])\s*$", text)
    out, found = [parts[0]], False
    for head, body in zip(parts[1::2], parts[2::2]):
        if head.strip("[] ") == table:
            pat = re.compile(r"(?m)^\s*%s\s*=.*$" % re.escape(key))
            if pat.search(body):
                body = pat.sub("%s = %s" % (key, value), body)
            else:
                body = "\n%s = %s" % (key, value) + body
            found = True
        out += [head, body]
    if not found:
        out.append("\n[%s]\n%s = %s\n" % (table, key, value))
    return "".join(out)


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("input")
    ap.add_argument("--max-steps", type=int)
    ap.add_argument("--precision", choices=["fp64", "mixed"])
    ap.add_argument("--mean-var")
    ap.add_argument("--output")
    args = ap.parse_args(argv)
    path = args.input
    tmp = None
    if args.max_steps is not None or args.precision:
        text = open(path).read()
        if args.max_steps is not None:
            text = _override(text, "problem", "maxSteps", str(args.max_steps))
        if args.precision:
            text = _override(text, "moment_system", "precision", '"%s"' % args.precision)
        # beside the input, so that paths in it (SU2MeshFile, restartFile) resolve alike
        fd, tmp = tempfile.mkstemp(suffix=".toml", dir=os.path.dirname(os.path.abspath(path)))
        with os.fdopen(fd, "w") as f:
            f.write(text)
        path = tmp
    try:
        rep = native.run_config(path, args.mean_var, args.output)
    finally:
        if tmp:
            os.unlink(tmp)
    rep["input"] = args.input
    print(json.dumps(rep), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
